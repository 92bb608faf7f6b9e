"""Quadratures on P2 (quadratic Lagrange) meshes, mirroring Inti.jl's
`Quadrature` (src/quadrature.jl:65-162): nodes are element-major, node i of
element e is qtags[i, e]; weights are sqrt(det(J^T J)) * w_ref
(src/utils.jl:90-100) and normals are the normalised (t2, -t1) in 2D and
t1 x t2 in 3D (src/utils.jl:121-135), here with analytic Jacobians instead of
automatic differentiation.

Reference rules (src/quadrature.jl:208-227): Gauss-Legendre on lines,
Vioreanu-Rokhlin on triangles (only the order-4, 6-point rule is tabulated
here: TRIANGLE_VR4N6, src/quad_rules_tables_triangle.jl:262-269).

Mesh generators (the Julia side builds these with meshgen / gmsh; here they
are written directly): circles and ellipses of P2 line elements, and
geodesic icosahedral / octahedral spheres of P2 triangles with an optional
radial map r(theta, phi).
"""
import math

import numpy as np

from . import _lib


# ----------------------------------------------------------- reference rules
def _legendre_and_derivative(n, t):
    """P_n(t) and P_n'(t) by the three-term recurrence."""
    p0 = np.ones_like(t)
    p1 = t.copy()
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * t * p1 - (k - 1) * p0) / k
    return p1, n * (t * p1 - p0) / (t * t - 1)


def gauss_legendre(n):
    """n-point Gauss-Legendre rule on [0, 1] (GaussLegendre{N},
    src/reference_integration.jl:139-158): ascending nodes, weights sum 1.
    Roots of P_n from numpy, polished by Newton steps in extended precision;
    weights 2 / ((1 - t^2) P_n'(t)^2)."""
    if n < 1:
        raise ValueError("n >= 1 required")
    t, _ = np.polynomial.legendre.leggauss(n)
    tl = t.astype(np.longdouble)
    for _ in range(2):
        p, dp = _legendre_and_derivative(n, tl)
        tl = tl - p / dp
    _, dp = _legendre_and_derivative(n, tl)
    wl = 2 / ((1 - tl * tl) * dp * dp)
    return np.asarray((tl + 1) / 2, dtype=np.float64), np.asarray(wl / 2, dtype=np.float64)


def gauss_legendre_order(order):
    """GaussLegendre(; order) -> ceil((order + 1)/2) nodes
    (src/reference_integration.jl:152)."""
    return gauss_legendre(int(math.ceil((order + 1) / 2)))


# TRIANGLE_VR4N6 (src/quad_rules_tables_triangle.jl:262-269): nodes on the
# reference triangle {u, v >= 0, u + v <= 1}, weights summing to 1/2.
VR4N6_NODES = np.array([
    [0.09157621350977069, 0.8168475729804586],
    [0.09157621350977072, 0.09157621350977081],
    [0.8168475729804584, 0.09157621350977081],
    [0.44594849091596483, 0.10810301816807029],
    [0.4459484909159649, 0.445948490915965],
    [0.10810301816807016, 0.445948490915965],
])
VR4N6_WEIGHTS = np.array([0.0549758718276609] * 3 + [0.1116907948390057] * 3)

# TRIANGLE_VR2N3 and TRIANGLE_VR5N10 (src/quad_rules_tables_triangle.jl:250-255, 277-288)
VR2N3_NODES = np.array([
    [0.16666666666666663, 0.6666666666666667],
    [0.16666666666666663, 0.16666666666666674],
    [0.6666666666666666, 0.16666666666666674],
])
VR2N3_WEIGHTS = np.array([0.1666666666666667] * 3)
VR5N10_NODES = np.array([
    [0.31397939767698096, 0.07109443567989363],
    [0.6149261666431254, 0.3139793976769811],
    [0.07109443567989354, 0.6149261666431255],
    [0.6149261666431254, 0.07109443567989363],
    [0.31397939767698096, 0.6149261666431255],
    [0.07109443567989351, 0.3139793976769811],
    [0.07111299453344089, 0.8577740109331182],
    [0.07111299453344097, 0.07111299453344105],
    [0.857774010933118, 0.07111299453344105],
    [0.33333333333333326, 0.3333333333333334],
])
VR5N10_WEIGHTS = np.array([0.051683140693245144] * 6 + [0.030139490203570284] * 3 + [0.09948268522981828])

# TRIANGLE_VR_ORDER_TO_NPTS (src/quad_rules_tables.jl:66-78), the rules tabulated here
TRIANGLE_VR_RULES = {2: (VR2N3_NODES, VR2N3_WEIGHTS), 4: (VR4N6_NODES, VR4N6_WEIGHTS),
                     5: (VR5N10_NODES, VR5N10_WEIGHTS)}
# TRIANGLE_VR_IORDER_TO_QORDER (src/quad_rules_tables.jl:83-95), restricted to them
TRIANGLE_VR_IORDER_TO_QORDER = {1: 2, 2: 4, 3: 5}


# TETRAHEDRON_VR2N4, VR3N10, VR5N20 (src/quad_rules_tables_tetrahedron.jl:222-385):
# nodes on the reference tetrahedron {u, v, w >= 0, u + v + w <= 1}, weights
# summing to 1/6; interpolation degrees 1, 2, 3 (src/quad_rules_tables.jl:145-157).
TET_VR2N4_NODES = np.array([
    [0.1381966011250106, 0.13819660112501048, 0.13819660112501042],
    [0.5854101966249685, 0.13819660112501053, 0.13819660112501053],
    [0.13819660112501048, 0.5854101966249685, 0.13819660112501053],
    [0.13819660112501053, 0.13819660112501053, 0.5854101966249683],
])
TET_VR2N4_WEIGHTS = np.array([0.04166666666666665, 0.041666666666666685, 0.041666666666666664,
                              0.041666666666666685])
TET_VR3N10_NODES = np.array([
    [0.07714091970968884, 0.07714091970968884, 0.07714091970968884],
    [0.40597452087424457, 0.09402547912575537, 0.09402547912575543],
    [0.7685772408709335, 0.07714091970968878, 0.07714091970968884],
    [0.09402547912575543, 0.4059745208742447, 0.09402547912575532],
    [0.4059745208742447, 0.40597452087424457, 0.09402547912575543],
    [0.07714091970968884, 0.7685772408709336, 0.07714091970968884],
    [0.09402547912575543, 0.09402547912575537, 0.4059745208742445],
    [0.4059745208742446, 0.09402547912575537, 0.4059745208742445],
    [0.09402547912575543, 0.4059745208742447, 0.4059745208742445],
    [0.07714091970968884, 0.07714091970968884, 0.7685772408709335],
])
TET_VR3N10_WEIGHTS = np.array([0.008403127078534032, 0.022175693058755084, 0.008403127078534035,
                               0.022175693058755084, 0.022175693058755084, 0.008403127078534032,
                               0.022175693058755098, 0.022175693058755098, 0.022175693058755098,
                               0.008403127078534035])
TET_VR5N20_NODES = np.array([
    [0.06994279207826776, 0.0699427920782677, 0.06994279207826776],
    [0.2928644237211868, 0.06164442753123828, 0.06164442753123833],
    [0.5838467212163365, 0.06164442753123828, 0.06164442753123833],
    [0.7901716237651969, 0.0699427920782677, 0.0699427920782677],
    [0.06164442753123828, 0.2928644237211868, 0.06164442753123833],
    [0.30974425517318926, 0.30974425517318926, 0.0707672344804322],
    [0.5838467212163364, 0.2928644237211869, 0.06164442753123833],
    [0.06164442753123833, 0.5838467212163365, 0.06164442753123828],
    [0.2928644237211869, 0.5838467212163365, 0.06164442753123833],
    [0.0699427920782677, 0.7901716237651969, 0.0699427920782677],
    [0.06164442753123828, 0.06164442753123828, 0.2928644237211868],
    [0.30974425517318926, 0.0707672344804322, 0.3097442551731892],
    [0.5838467212163365, 0.06164442753123833, 0.2928644237211868],
    [0.07076723448043226, 0.3097442551731893, 0.30974425517318915],
    [0.30974425517318926, 0.30974425517318926, 0.30974425517318926],
    [0.06164442753123828, 0.5838467212163365, 0.2928644237211868],
    [0.06164442753123828, 0.06164442753123833, 0.5838467212163364],
    [0.2928644237211868, 0.06164442753123833, 0.5838467212163365],
    [0.06164442753123828, 0.2928644237211869, 0.5838467212163365],
    [0.0699427920782677, 0.06994279207826765, 0.7901716237651969],
])
TET_VR5N20_WEIGHTS = np.array([0.004559181826208037, 0.006804733769335409, 0.0068047337693354125,
                               0.004559181826208037, 0.006804733769335409, 0.016693283532452394,
                               0.006804733769335416, 0.006804733769335412, 0.006804733769335412,
                               0.004559181826208037, 0.006804733769335412, 0.016693283532452394,
                               0.006804733769335415, 0.016693283532452394, 0.016693283532452394,
                               0.0068047337693354125, 0.006804733769335416, 0.006804733769335416,
                               0.006804733769335416, 0.004559181826208035])
# TETRAHEDRON_VR_ORDER_TO_NPTS (src/quad_rules_tables.jl:128-140), the rules tabulated here
TETRAHEDRON_VR_RULES = {2: (TET_VR2N4_NODES, TET_VR2N4_WEIGHTS), 3: (TET_VR3N10_NODES, TET_VR3N10_WEIGHTS),
                        5: (TET_VR5N20_NODES, TET_VR5N20_WEIGHTS)}


def vioreanu_rokhlin_tetrahedron(order):
    if order not in TETRAHEDRON_VR_RULES:
        raise ValueError(
            f"VioreanuRokhlin quadrature of order {order} not available for ReferenceTetrahedron "
            f"(available: {sorted(TETRAHEDRON_VR_RULES)})")
    x, w = TETRAHEDRON_VR_RULES[order]
    return x.copy(), w.copy()


def triangle_vr_interpolation_order_to_quadrature_order(i):
    """Triangle_VR_interpolation_order_to_quadrature_order
    (src/reference_integration.jl:320-322)."""
    return TRIANGLE_VR_IORDER_TO_QORDER[i]


def vioreanu_rokhlin_triangle(order):
    if order not in TRIANGLE_VR_RULES:
        raise ValueError(
            f"VioreanuRokhlin quadrature of order {order} not available for ReferenceTriangle "
            f"(available: {sorted(TRIANGLE_VR_RULES)})")
    x, w = TRIANGLE_VR_RULES[order]
    return x.copy(), w.copy()


# ------------------------------------------------------------- P2 elements
def p2_line_eval(vals, s):
    """LagrangeLine{3} with reference nodes 0, 1/2, 1
    (src/reference_interpolation.jl:374-396). vals: (ne, 3, dim); s: (nq,).
    Returns x (ne, nq, dim) and dx/ds (ne, nq, dim)."""
    s = np.asarray(s)[None, :, None]
    l0 = 2 * (s - 0.5) * (s - 1)
    l1 = -4 * s * (s - 1)
    l2 = 2 * s * (s - 0.5)
    d0 = 4 * s - 3
    d1 = -8 * s + 4
    d2 = 4 * s - 1
    v0, v1, v2 = (vals[:, i, None, :] for i in range(3))
    return l0 * v0 + l1 * v1 + l2 * v2, d0 * v0 + d1 * v1 + d2 * v2


def p2_triangle_eval(vals, uv):
    """LagrangeTriangle{6}; reference nodes in the order of
    src/reference_interpolation.jl:499-509: (0,0), (1/2,0), (1,0), (0,1/2),
    (1/2,1/2), (0,1). vals: (ne, 6, 3); uv: (nq, 2).
    Returns x, dx/du, dx/dv, each (ne, nq, 3)."""
    u = uv[None, :, 0, None]
    v = uv[None, :, 1, None]
    w = 1 - u - v
    B = [w * (2 * w - 1), 4 * u * w, u * (2 * u - 1), 4 * v * w, 4 * u * v, v * (2 * v - 1)]
    Bu = [-(4 * w - 1), 4 * (w - u), 4 * u - 1, -4 * v, 4 * v, 0 * u]
    Bv = [-(4 * w - 1), -4 * u, 0 * u, 4 * (w - v), 4 * u, 4 * v - 1]
    x = sum(B[i] * vals[:, i, None, :] for i in range(6))
    xu = sum(Bu[i] * vals[:, i, None, :] for i in range(6))
    xv = sum(Bv[i] * vals[:, i, None, :] for i in range(6))
    return x, xu, xv


# --------------------------------------------------------------- Quadrature
class Quadrature:
    """SoA quadrature: coords (dim, N), normals (dim, N), weights (N,),
    qtags (nq, ne) element-major (node i of element e is qtags[i, e]).

    Several element types (Quadrature.etype2qtags, src/quadrature.jl:60-120):
    pass qtags as a list with one (nq_t, ne_t) array per type; etype_qtags
    is that list (one entry for a single type), and qtags / nq / ne are
    defined only for a single type."""

    def __init__(self, coords, normals, weights, qtags, elements=None, name="", qorder=None):
        self.qorder = qorder
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.normals = None if normals is None else np.ascontiguousarray(normals, dtype=np.float64)
        self.weights = np.ascontiguousarray(weights, dtype=np.float64)
        if isinstance(qtags, (list, tuple)):
            self.etype_qtags = [np.ascontiguousarray(q, dtype=np.int64) for q in qtags]
        else:
            self.etype_qtags = [np.ascontiguousarray(qtags, dtype=np.int64)]
        if not self.etype_qtags or any(q.ndim != 2 for q in self.etype_qtags):
            raise ValueError("qtags: one (nq, ne) array per element type")
        self.elements = elements
        self.name = name
        self._dev = {}

    @property
    def qtags(self):
        if len(self.etype_qtags) != 1:
            raise ValueError(f"{len(self.etype_qtags)} element types: use etype_qtags")
        return self.etype_qtags[0]

    @property
    def dim(self):
        return self.coords.shape[0]

    def __len__(self):
        return self.coords.shape[1]

    @property
    def nq(self):
        return self.qtags.shape[0]

    @property
    def ne(self):
        return self.qtags.shape[1]

    def device(self, ctx):
        """Device copy on ctx (uploaded once, cached)."""
        key = id(ctx)
        if key not in self._dev:
            import ctypes as C
            h = C.c_void_p()
            ctx.check(ctx.lib.inti_quad_upload(
                ctx.handle, self.dim, len(self), _lib.dptr(self.coords),
                _lib.dptr(self.normals), _lib.dptr(self.weights), C.byref(h)))
            self._dev[key] = (ctx, h)
        return self._dev[key][1]

    def __del__(self):
        if _lib.exiting():
            return
        for ctx, h in getattr(self, "_dev", {}).values():
            try:
                ctx.lib.inti_quad_destroy(h)
            except Exception:
                pass

    def etype_flat(self):
        """(nq per type int32, ne per type int64, tags): the column-major
        (nq_t, ne_t) blocks of every type one after another."""
        nqs = np.array([q.shape[0] for q in self.etype_qtags], dtype=np.int32)
        nes = np.array([q.shape[1] for q in self.etype_qtags], dtype=np.int64)
        tags = np.concatenate([q.T.reshape(-1) for q in self.etype_qtags]).astype(np.int64)
        return nqs, nes, np.ascontiguousarray(tags)

    def subset(self, idx):
        """Point set of the nodes idx (no element structure), e.g. target[idx]."""
        idx = np.asarray(idx, dtype=np.int64)
        return Quadrature(self.coords[:, idx], None if self.normals is None else self.normals[:, idx],
                          self.weights[idx], np.zeros((1, 0), dtype=np.int64), name=self.name + "[sub]")


def join_quadratures(*quads, name="joined"):
    """One quadrature whose element types are the given quadratures' (a mesh
    with several element types, each with its own rule, as
    Quadrature(msh, etype2qrule) builds, src/quadrature.jl:103-120); node tags
    are offset by the nodes that precede them."""
    off, tags = 0, []
    for Q in quads:
        tags += [q + off for q in Q.etype_qtags]
        off += len(Q)
    nrm = None if any(Q.normals is None for Q in quads) else np.concatenate([Q.normals for Q in quads], axis=1)
    return Quadrature(np.concatenate([Q.coords for Q in quads], axis=1), nrm,
                      np.concatenate([Q.weights for Q in quads]), tags, name=name)


def points(coords):
    """A target point set without normals or weights (e.g. volume targets)."""
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    return Quadrature(coords, None, np.ones(coords.shape[1]), np.zeros((1, 0), dtype=np.int64),
                      name="points")


def _line_quadrature(vals, qorder, orientation=1.0):
    x_ref, w_ref = gauss_legendre_order(qorder)
    x, dx = p2_line_eval(vals, x_ref)
    ne, nq, _ = x.shape
    mu = np.sqrt(np.sum(dx * dx, axis=2))
    w = mu * w_ref[None, :]
    n = orientation * np.stack([dx[..., 1], -dx[..., 0]], axis=2) / mu[..., None]
    coords = x.reshape(ne * nq, 2).T
    normals = n.reshape(ne * nq, 2).T
    qtags = np.arange(ne * nq, dtype=np.int64).reshape(ne, nq).T
    return coords, normals, w.reshape(-1), qtags


def _triangle_quadrature(vals, qorder):
    uv, w_ref = vioreanu_rokhlin_triangle(qorder)
    x, xu, xv = p2_triangle_eval(vals, uv)
    ne, nq, _ = x.shape
    cr = np.cross(xu, xv)
    mu = np.linalg.norm(cr, axis=2)
    w = mu * w_ref[None, :]
    n = cr / mu[..., None]
    coords = x.reshape(ne * nq, 3).T
    normals = n.reshape(ne * nq, 3).T
    qtags = np.arange(ne * nq, dtype=np.int64).reshape(ne, nq).T
    return coords, normals, w.reshape(-1), qtags


def ellipse_quadrature(a, b, nel, qorder=8, center=(0.0, 0.0), rotation=0.0):
    """Closed ellipse (a cos t, b sin t), counter-clockwise, nel P2 line
    elements with nodes at parameters t_e, (t_e + t_{e+1})/2, t_{e+1}."""
    e = np.arange(nel)
    t0 = 2 * np.pi * e / nel
    t1 = 2 * np.pi * (e + 1) / nel
    tm = (t0 + t1) / 2
    ts = np.stack([t0, tm, t1], axis=1)
    c, s = math.cos(rotation), math.sin(rotation)
    px, py = a * np.cos(ts), b * np.sin(ts)
    vals = np.stack([c * px - s * py + center[0], s * px + c * py + center[1]], axis=2)
    coords, normals, w, qtags = _line_quadrature(vals, qorder)
    return Quadrature(coords, normals, w, qtags, elements=vals,
                      name=f"ellipse(a={a},b={b},nel={nel},q={qorder})",
                      qorder=2 * len(gauss_legendre_order(qorder)[0]) - 1)


def circle_quadrature(nel, qorder=8, radius=1.0, center=(0.0, 0.0)):
    """Unit circle of nel P2 line elements (all nodes on the circle)."""
    q = ellipse_quadrature(radius, radius, nel, qorder, center)
    q.name = f"circle(nel={nel},q={qorder})"
    return q


_ICO_V = None


def _icosahedron():
    p = (1 + 5 ** 0.5) / 2
    V = np.array([[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0], [0, -1, p], [0, 1, p],
                  [0, -1, -p], [0, 1, -p], [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]],
                 dtype=np.float64)
    V /= np.linalg.norm(V, axis=1)[:, None]
    F = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                  [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                  [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                  [8, 6, 7], [9, 8, 1]])
    return V, F


def _octahedron():
    V = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]],
                 dtype=np.float64)
    F = np.array([[0, 2, 4], [2, 1, 4], [1, 3, 4], [3, 0, 4],
                  [2, 0, 5], [1, 2, 5], [3, 1, 5], [0, 3, 5]])
    return V, F


def _geodesic_p2(base, freq):
    """P2 node sets (ne, 6, 3) of a geodesic sphere: every base face split
    into freq^2 flat triangles whose 6 P2 nodes (vertices and edge midpoints)
    are projected to the unit sphere; faces oriented outward."""
    V, F = _icosahedron() if base == "ico" else _octahedron()
    tri = []
    for (ia, ib, ic) in F:
        A, B, Cc = V[ia], V[ib], V[ic]
        if np.dot(np.cross(B - A, Cc - A), A + B + Cc) < 0:
            B, Cc = Cc, B
        P = lambda i, j: A + (i / freq) * (B - A) + (j / freq) * (Cc - A)  # noqa: E731
        for i in range(freq):
            for j in range(freq - i):
                tri.append((P(i, j), P(i + 1, j), P(i, j + 1)))
                if i + j + 1 < freq:
                    tri.append((P(i + 1, j), P(i + 1, j + 1), P(i, j + 1)))
    T = np.array(tri)  # (ne, 3, 3)
    a, b, c = T[:, 0], T[:, 1], T[:, 2]
    nodes = np.stack([a, (a + b) / 2, b, (a + c) / 2, (b + c) / 2, c], axis=1)
    return nodes / np.linalg.norm(nodes, axis=2)[..., None]


def sphere_quadrature(freq, qorder=4, base="ico", radius=1.0, radial=None, center=(0, 0, 0)):
    """Geodesic sphere of P2 triangles: 20*freq^2 ('ico') or 8*freq^2
    ('octa') elements. radial(theta, phi) -> r maps the unit-sphere nodes
    (theta polar angle from +z, phi azimuth); default r = radius."""
    unit = _geodesic_p2(base, freq)
    if radial is None:
        r = radius
    else:
        th = np.arccos(np.clip(unit[..., 2], -1, 1))
        ph = np.arctan2(unit[..., 1], unit[..., 0])
        r = radial(th, ph)[..., None]
    vals = unit * r + np.asarray(center, dtype=np.float64)
    coords, normals, w, qtags = _triangle_quadrature(vals, qorder)
    # outward orientation check (star-shaped surfaces): flip elements whose
    # normal points towards the centre (build_orientation!, src/mesh.jl:166-185)
    c = np.asarray(center, dtype=np.float64)[:, None]
    nq = qtags.shape[0]
    s = np.sign(np.sum((coords - c) * normals, axis=0)).reshape(-1, nq).sum(axis=1)
    if np.any(s < 0):
        flip = np.repeat(np.where(s < 0, -1.0, 1.0), nq)
        normals = normals * flip[None, :]
    return Quadrature(coords, normals, w, qtags, elements=vals,
                      name=f"sphere({base},freq={freq},q={qorder})", qorder=qorder)


def _triangle_quadrature_2d(vals, qorder):
    """Planar P2 triangles: weights |det J| w_ref, no normals."""
    uv, w_ref = vioreanu_rokhlin_triangle(qorder)
    x, xu, xv = p2_triangle_eval(vals, uv)
    ne, nq, _ = x.shape
    det = xu[..., 0] * xv[..., 1] - xu[..., 1] * xv[..., 0]
    w = np.abs(det) * w_ref[None, :]
    coords = x.reshape(ne * nq, 2).T
    qtags = np.arange(ne * nq, dtype=np.int64).reshape(ne, nq).T
    return coords, w.reshape(-1), qtags


def _zip_rings(ia, ta, ib, tb):
    """Triangulate the band between two closed rings of point indices with
    angles ta (inner) and tb (outer), both increasing, counter-clockwise."""
    tris = []
    na, nb = len(ia), len(ib)
    if na == 1:  # a fan around the centre
        return [(ia[0], ib[j], ib[(j + 1) % nb]) for j in range(nb)]
    a = b = 0
    while a < na or b < nb:
        ta_next = ta[(a + 1) % na] + (2 * np.pi if a + 1 >= na else 0.0)
        tb_next = tb[(b + 1) % nb] + (2 * np.pi if b + 1 >= nb else 0.0)
        if b < nb and (a >= na or tb_next <= ta_next):
            tris.append((ia[a % na], ib[b % nb], ib[(b + 1) % nb]))
            b += 1
        else:
            tris.append((ia[a % na], ib[b % nb], ia[(a + 1) % na]))
            a += 1
    return tris


def _p2_disk(V, T, tpar, a, b, qorder, boundary_qorder, name):
    """Volume and boundary quadratures of a triangulated elliptical disk.
    V: vertices (nv, 2); T: triangles (ne, 3); tpar: the ellipse parameter t
    of each boundary vertex (NaN for interior ones). Edges between two
    consecutive boundary vertices become curved P2 edges (midpoint at the mean
    parameter, on the ellipse); all other edges are straight."""
    def xy(t):
        return np.array([a * np.cos(t), b * np.sin(t)])

    def tmid(tu, tv):
        if abs(tv - tu) > np.pi:
            tv += 2 * np.pi if tv < tu else -2 * np.pi
        return 0.5 * (tu + tv)

    vals = np.zeros((len(T), 6, 2))
    for e, (i0, i1, i2) in enumerate(T):
        p0, p1, p2 = V[i0], V[i1], V[i2]
        if (p1[0] - p0[0]) * (p2[1] - p0[1]) - (p1[1] - p0[1]) * (p2[0] - p0[0]) < 0:
            i1, i2 = i2, i1  # counter-clockwise in the (x, y) plane

        def mid(u, v):
            if not (np.isnan(tpar[u]) or np.isnan(tpar[v])):
                return xy(tmid(tpar[u], tpar[v]))
            return 0.5 * (V[u] + V[v])

        vals[e] = [V[i0], mid(i0, i1), V[i1], mid(i0, i2), mid(i1, i2), V[i2]]
    coords, w, qtags = _triangle_quadrature_2d(vals, qorder)
    vol = Quadrature(coords, None, w, qtags, elements=vals, name=name, qorder=qorder)
    # boundary: the same curved edges as P2 line elements, counter-clockwise
    bidx = np.nonzero(~np.isnan(tpar))[0]
    tb = np.sort(tpar[bidx])
    bvals = np.zeros((len(tb), 3, 2))
    for j in range(len(tb)):
        t0, t1 = tb[j], tb[j + 1] if j + 1 < len(tb) else tb[0] + 2 * np.pi
        bvals[j] = [xy(t0), xy(0.5 * (t0 + t1)), xy(t1)]
    vol.boundary_elements = bvals
    bq = qorder if boundary_qorder is None else boundary_qorder
    return vol, curve_quadrature(bvals, bq, name + "-boundary")


def curve_quadrature(elements, qorder, name="curve"):
    """Gauss-Legendre quadrature (order qorder) on P2 line elements
    (ne, 3, 2), nodes in element order; normals (t2, -t1)/|t|."""
    bc, bn, bw, bqt = _line_quadrature(elements, qorder)
    return Quadrature(bc, bn, bw, bqt, elements=elements, name=name,
                      qorder=2 * len(gauss_legendre_order(qorder)[0]) - 1)


def area_quadrature(elements, qorder, name="area"):
    """Vioreanu-Rokhlin quadrature (order qorder) on planar P2 triangles
    (ne, 6, 2), nodes in element order (a 2D volume quadrature)."""
    coords, w, qtags = _triangle_quadrature_2d(elements, qorder)
    return Quadrature(coords, None, w, qtags, elements=elements, name=name, qorder=qorder)


def surface_quadrature(elements, qorder, name="surface", center=None):
    """Vioreanu-Rokhlin quadrature (order qorder) on P2 triangles (ne, 6, 3),
    nodes in element order; normals along x_u x x_v, turned away from center
    element by element when center is given (star-shaped surfaces)."""
    coords, normals, w, qtags = _triangle_quadrature(elements, qorder)
    if center is not None:
        c = np.asarray(center, dtype=np.float64)[:, None]
        nq = qtags.shape[0]
        s = np.sign(np.sum((coords - c) * normals, axis=0)).reshape(-1, nq).sum(axis=1)
        normals = normals * np.repeat(np.where(s < 0, -1.0, 1.0), nq)[None, :]
    return Quadrature(coords, normals, w, qtags, elements=elements, name=name, qorder=qorder)


def boundary_quadrature(volume, qorder):
    """Quadrature of order qorder on the outer boundary of a meshed 2D volume
    (the boundary the reference's volume_potential builds from the parent mesh
    when none is given, api.jl:314-322)."""
    el = getattr(volume, "boundary_elements", None)
    if el is None:
        raise ValueError("the volume quadrature carries no boundary mesh")
    return curve_quadrature(el, qorder, volume.name + "-boundary")


def disk_quadratures(nrings, a=1.0, b=1.0, qorder=5, boundary_qorder=None):
    """(volume, boundary) quadratures of the elliptical disk x^2/a^2 + y^2/b^2 <= 1:
    P2 triangles on concentric rings (ring i has 6 i vertices), interior edges
    straight, the outer edges curved (midpoints on the ellipse); the boundary
    quadrature uses the same outer edges as P2 line elements, so both discretise
    the same curve. qorder: Vioreanu-Rokhlin order of the volume rule;
    boundary_qorder: Gauss-Legendre order of the boundary rule (default qorder)."""
    pts = [(0.0, 0.0)]
    rings, angles = [[0]], [np.array([0.0])]
    for i in range(1, nrings + 1):
        t = 2 * np.pi * np.arange(6 * i) / (6 * i)
        idx = list(range(len(pts), len(pts) + 6 * i))
        pts.extend((i / nrings * np.cos(tt), i / nrings * np.sin(tt)) for tt in t)
        rings.append(idx)
        angles.append(t)
    P = np.array(pts)
    rho = np.hypot(P[:, 0], P[:, 1])
    theta = np.arctan2(P[:, 1], P[:, 0])
    tris = []
    for i in range(1, nrings + 1):
        tris += _zip_rings(rings[i - 1], angles[i - 1], rings[i], angles[i])
    V = np.stack([a * rho * np.cos(theta), b * rho * np.sin(theta)], axis=1)
    tpar = np.full(len(P), np.nan)
    tpar[rings[-1]] = np.mod(angles[-1], 2 * np.pi)
    return _p2_disk(V, np.array(tris), tpar, a, b, qorder, boundary_qorder,
                    f"disk(a={a},b={b},rings={nrings},q={qorder})")


def ellipse_disk_quadratures(meshsize, a=1.0, b=1.0, qorder=5, boundary_qorder=None):
    """(volume, boundary) quadratures of the elliptical disk with a quasi-uniform
    mesh of size about `meshsize` (what gmsh's addDisk + MeshSizeMax/Min gives
    the reference's examples): boundary vertices equally spaced in arc length,
    interior vertices on a hexagonal lattice kept at least 0.6 meshsize from
    the boundary, Delaunay triangulation, P2 with curved boundary edges."""
    from scipy.spatial import Delaunay, cKDTree
    h = float(meshsize)
    ts = np.linspace(0.0, 2 * np.pi, 20001)
    arc = np.concatenate([[0.0], np.cumsum(np.hypot(np.diff(a * np.cos(ts)), np.diff(b * np.sin(ts))))])
    nb = max(6, int(round(arc[-1] / h)))
    tb = np.interp(np.arange(nb) * arc[-1] / nb, arc, ts)
    B = np.stack([a * np.cos(tb), b * np.sin(tb)], axis=1)
    dense = cKDTree(np.stack([a * np.cos(ts), b * np.sin(ts)], axis=1))
    ny = int(np.ceil(b / (h * np.sqrt(3) / 2))) + 1
    nx = int(np.ceil(a / h)) + 1
    I = []
    for j in range(-ny, ny + 1):
        y = j * h * np.sqrt(3) / 2
        for i in range(-nx, nx + 1):
            x = (i + 0.5 * (j % 2)) * h
            if (x / a) ** 2 + (y / b) ** 2 < 1.0:
                I.append((x, y))
    I = np.array(I)
    I = I[dense.query(I)[0] > 0.6 * h]
    V = np.concatenate([B, I], axis=0)
    T = Delaunay(V).simplices
    # drop zero-area triangles (collinear boundary triples cannot occur on a
    # strictly convex curve, but guard anyway)
    p0, p1, p2 = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    area = 0.5 * np.abs((p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p1[:, 1] - p0[:, 1]) * (p2[:, 0] - p0[:, 0]))
    T = T[area > 1e-12 * h * h]
    tpar = np.full(len(V), np.nan)
    tpar[:nb] = tb
    # every boundary edge must join consecutive boundary vertices
    for tri in T:
        on = [v for v in tri if v < nb]
        if len(on) == 3:
            raise RuntimeError("triangle with three boundary vertices")
        if len(on) == 2 and (on[0] - on[1]) % nb not in (1, nb - 1):
            raise RuntimeError("boundary chord in the triangulation")
    return _p2_disk(V, T, tpar, a, b, qorder, boundary_qorder,
                    f"ellipse-disk(a={a},b={b},h={h},q={qorder})")


def box_quadratures(ncell, half=0.5, qorder=3, boundary_qorder=4):
    """(volume, boundary) quadratures of the cube [-half, half]^3: ncell^3
    cells, each cut into six tetrahedra around its main diagonal (every
    tetrahedron positively oriented), with the Vioreanu-Rokhlin tetrahedron
    rule of order qorder (3: ten nodes, interpolation degree 2; 5: twenty
    nodes, degree 3); the boundary is the six faces cut into two flat
    triangles per cell face, Vioreanu-Rokhlin triangle rule of order
    boundary_qorder, outward normals. Straight elements are exact here."""
    uvw, wref = vioreanu_rokhlin_tetrahedron(qorder)
    g = np.linspace(-half, half, ncell + 1)
    corner = np.array([[i, j, k] for k in (0, 1) for j in (0, 1) for i in (0, 1)], dtype=float)
    # six tetrahedra sharing the diagonal (0,0,0)-(1,1,1): paths along the axes
    paths = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    tets = []
    for k in range(ncell):
        for j in range(ncell):
            for i in range(ncell):
                o = np.array([g[i], g[j], g[k]])
                h = g[1] - g[0]
                for pth in paths:
                    v = [np.zeros(3)]
                    for ax in pth:
                        nv = v[-1].copy()
                        nv[ax] += 1.0
                        v.append(nv)
                    V = o + h * np.array(v)
                    J = np.stack([V[1] - V[0], V[2] - V[0], V[3] - V[0]], axis=1)
                    if np.linalg.det(J) < 0:
                        V[[1, 2]] = V[[2, 1]]
                    tets.append(V)
    T = np.array(tets)                                               # (ne, 4, 3)
    J = np.stack([T[:, 1] - T[:, 0], T[:, 2] - T[:, 0], T[:, 3] - T[:, 0]], axis=2)  # (ne, 3, 3)
    x = T[:, None, 0, :] + np.einsum("eij,qj->eqi", J, uvw)          # (ne, nq, 3)
    det = np.linalg.det(J)
    w = det[:, None] * wref[None, :]
    ne, nq, _ = x.shape
    vol = Quadrature(x.reshape(ne * nq, 3).T, None, w.reshape(-1),
                     np.arange(ne * nq, dtype=np.int64).reshape(ne, nq).T, name=f"box(n={ncell},q={qorder})",
                     qorder=qorder)
    vol.tetrahedra = T
    # boundary: per face (axis, side), cells (a, b) -> two triangles, outward
    uv, wt = vioreanu_rokhlin_triangle(boundary_qorder)
    coords, normals, weights = [], [], []
    for ax in range(3):
        for side in (-1.0, 1.0):
            a1, a2 = [d for d in range(3) if d != ax]
            for b in range(ncell):
                for a in range(ncell):
                    P = np.zeros((4, 3))
                    for c, (da, db) in enumerate(((0, 0), (1, 0), (1, 1), (0, 1))):
                        P[c, ax] = side * half
                        P[c, a1] = g[a + da]
                        P[c, a2] = g[b + db]
                    nrm = np.zeros(3)
                    nrm[ax] = side
                    for tri in ((0, 1, 2), (0, 2, 3)):
                        A, B, Cc = P[list(tri)]
                        if np.dot(np.cross(B - A, Cc - A), nrm) < 0:
                            B, Cc = Cc, B
                        area2 = np.linalg.norm(np.cross(B - A, Cc - A))
                        pts = A[None, :] + uv[:, 0:1] * (B - A)[None, :] + uv[:, 1:2] * (Cc - A)[None, :]
                        coords.append(pts)
                        normals.append(np.repeat(nrm[None, :], len(wt), axis=0))
                        weights.append(area2 * wt)
    C = np.concatenate(coords)
    nb = len(C)
    bq = len(wt)
    bnd = Quadrature(C.T, np.concatenate(normals).T, np.concatenate(weights),
                     np.arange(nb, dtype=np.int64).reshape(nb // bq, bq).T, name=f"box-boundary(n={ncell})",
                     qorder=boundary_qorder)
    return vol, bnd
