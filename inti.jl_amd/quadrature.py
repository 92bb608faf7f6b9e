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

# TRIANGLE_VR_ORDER_TO_NPTS (src/quad_rules_tables.jl:66-78), tabulated here
TRIANGLE_VR_RULES = {4: (VR4N6_NODES, VR4N6_WEIGHTS)}


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
    qtags (nq, ne) element-major (node i of element e is qtags[i, e])."""

    def __init__(self, coords, normals, weights, qtags, elements=None, name=""):
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.normals = None if normals is None else np.ascontiguousarray(normals, dtype=np.float64)
        self.weights = np.ascontiguousarray(weights, dtype=np.float64)
        self.qtags = np.ascontiguousarray(qtags, dtype=np.int64)
        self.elements = elements
        self.name = name
        self._dev = {}

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

    def subset(self, idx):
        """Point set of the nodes idx (no element structure), e.g. target[idx]."""
        idx = np.asarray(idx, dtype=np.int64)
        return Quadrature(self.coords[:, idx], None if self.normals is None else self.normals[:, idx],
                          self.weights[idx], np.zeros((1, 0), dtype=np.int64), name=self.name + "[sub]")


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
                      name=f"ellipse(a={a},b={b},nel={nel},q={qorder})")


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
                      name=f"sphere({base},freq={freq},q={qorder})")
