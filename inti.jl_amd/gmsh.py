"""Gmsh mesh ingest (MSH 4.1, ASCII and binary): what `Inti.import_mesh`
(ext/IntiGmshExt.jl:49-221) gives the GPU path, without the gmsh library.

`read_msh(path)` returns the nodes, the entities (with their physical names
and signed bounding entities) and the element blocks, node indices 0-based
and permuted to the reference's node order of each element type
(ext/IntiGmshExt.jl:190-213: Gmsh lists a P2 triangle as its three vertices
then the edge midpoints 01, 12, 20; the reference orders the nodes of a
simplex by `Iterators.product` with coordinate sum <= order, first index
fastest: (0,0), (1/2,0), (1,0), (0,1/2), (1/2,1/2), (0,1)). Quadratures
are then built on the P2 (or P1, lifted to P2) elements of chosen entities:
`boundary_quadrature` for the surface of a volume (each bounding surface
oriented against the volume's tetrahedra, normals outward), `entity_quadrature`
for the elements of given entities, `write_msh` for ASCII output.
"""
import struct

import numpy as np

# element type -> (name, dimension, number of nodes)
ELEMENT_TYPES = {
    15: ("point", 0, 1),
    1: ("line2", 1, 2),
    8: ("line3", 1, 3),
    2: ("tri3", 2, 3),
    9: ("tri6", 2, 6),
    3: ("quad4", 2, 4),
    4: ("tet4", 3, 4),
    11: ("tet10", 3, 10),
}

# Gmsh node order -> reference node order (index into the Gmsh list)
NODE_PERMUTATION = {
    8: [0, 2, 1],            # line: v0, v1, mid -> v0, mid, v1
    9: [0, 3, 1, 5, 4, 2],   # triangle: v0 v1 v2 m01 m12 m20 -> v0 m01 v1 m20 m12 v2
}


class GmshMesh:
    def __init__(self):
        self.nodes = None            # (N, 3)
        self.node_tags = None        # (N,) Gmsh tags
        self.physical_names = {}     # (dim, tag) -> name
        self.entities = {}           # (dim, tag) -> {"physical": [...], "bounding": [signed tags]}
        self.blocks = []             # dicts: dim, entity, type, tags (ne,), nodes (ne, nn) 0-based

    def element_types(self):
        return sorted({ELEMENT_TYPES[b["type"]][0] for b in self.blocks})

    def elements(self, dim, entities=None, type_names=None):
        """(type, nodes (ne, nn)) blocks of dimension dim (optionally of the
        given entity tags / element type names)."""
        out = []
        for b in self.blocks:
            if b["dim"] != dim:
                continue
            if entities is not None and b["entity"] not in entities:
                continue
            name = ELEMENT_TYPES[b["type"]][0]
            if type_names is not None and name not in type_names:
                continue
            out.append((name, b["nodes"]))
        return out

    def entity_tags(self, dim):
        return sorted(t for (d, t) in self.entities if d == dim)


class _Reader:
    def __init__(self, data):
        self.d = data
        self.p = 0

    def line(self):
        q = self.d.index(b"\n", self.p)
        s = self.d[self.p:q].decode("ascii", "replace").strip()
        self.p = q + 1
        return s

    def expect(self, tag):
        while True:
            s = self.line()
            if s:
                break
        if s != tag:
            raise ValueError(f"msh: expected {tag}, found {s[:40]!r}")

    def skip_to(self, tag):
        while True:
            if self.p >= len(self.d):
                raise ValueError(f"msh: {tag} not found")
            if self.line() == tag:
                return

    def unpack(self, fmt):
        n = struct.calcsize(fmt)
        v = struct.unpack_from(fmt, self.d, self.p)
        self.p += n
        return v

    def array(self, dtype, count):
        a = np.frombuffer(self.d, dtype=dtype, count=count, offset=self.p)
        self.p += a.nbytes
        return a


class _Tokens:
    """Whitespace tokens of an ASCII section."""

    def __init__(self, text):
        self.t = text.split()
        self.i = 0

    def next(self, cast=int):
        v = cast(self.t[self.i])
        self.i += 1
        return v

    def many(self, n, cast=int):
        v = [cast(x) for x in self.t[self.i:self.i + n]]
        self.i += n
        return v


def _section_text(r, end):
    q = r.d.index(end.encode(), r.p)
    text = r.d[r.p:q].decode("ascii")
    r.p = q
    r.line()
    return text


def read_msh(path):
    """Read a Gmsh MSH 4.1 file (ASCII or binary, 8-byte size_t)."""
    with open(path, "rb") as f:
        data = f.read()
    r = _Reader(data)
    r.expect("$MeshFormat")
    head = r.line().split()
    if head[0] != "4.1":
        raise ValueError(f"msh: version {head[0]} not supported (4.1 only)")
    binary = head[1] == "1"
    if binary:
        if head[2] != "8":
            raise ValueError("msh: binary files need 8-byte size_t")
        one, = r.unpack("<i")
        if one != 1:
            raise ValueError("msh: big-endian binary files are not supported")
        r.line()
    r.expect("$EndMeshFormat")
    mesh = GmshMesh()
    sz = "<Q"
    while r.p < len(data):
        s = r.line()
        if not s:
            continue
        if s == "$PhysicalNames":
            tk = _Tokens(_section_text(r, "$EndPhysicalNames"))
            for _ in range(tk.next()):
                d, t = tk.next(), tk.next()
                name = tk.t[tk.i]
                # a quoted name may contain spaces
                while not name.endswith('"'):
                    tk.i += 1
                    name += " " + tk.t[tk.i]
                tk.i += 1
                mesh.physical_names[(d, t)] = name.strip('"')
        elif s == "$Entities":
            if binary:
                counts = [r.unpack(sz)[0] for _ in range(4)]
                for dim in range(4):
                    for _ in range(counts[dim]):
                        tag, = r.unpack("<i")
                        r.unpack("<3d" if dim == 0 else "<6d")
                        nphys, = r.unpack(sz)
                        phys = list(r.unpack(f"<{nphys}i")) if nphys else []
                        bnd = []
                        if dim > 0:
                            nb, = r.unpack(sz)
                            bnd = list(r.unpack(f"<{nb}i")) if nb else []
                        mesh.entities[(dim, tag)] = {"physical": phys, "bounding": bnd}
                r.skip_to("$EndEntities")
            else:
                tk = _Tokens(_section_text(r, "$EndEntities"))
                counts = tk.many(4)
                for dim in range(4):
                    for _ in range(counts[dim]):
                        tag = tk.next()
                        tk.many(3 if dim == 0 else 6, float)
                        phys = tk.many(tk.next())
                        bnd = tk.many(tk.next()) if dim > 0 else []
                        mesh.entities[(dim, tag)] = {"physical": phys, "bounding": bnd}
        elif s == "$Nodes":
            tags, xyz = [], []
            if binary:
                nblk, nn, _, _ = r.unpack("<4Q")
                for _ in range(nblk):
                    edim, etag, par = r.unpack("<3i")
                    cnt, = r.unpack(sz)
                    tags.append(r.array("<u8", cnt).astype(np.int64))
                    ncoord = 3 + (edim if par else 0)
                    xyz.append(r.array("<f8", cnt * ncoord).reshape(cnt, ncoord)[:, :3])
                r.skip_to("$EndNodes")
            else:
                tk = _Tokens(_section_text(r, "$EndNodes"))
                nblk, nn = tk.next(), tk.next()
                tk.many(2)
                for _ in range(nblk):
                    edim, etag, par, cnt = tk.many(4)
                    tags.append(np.array(tk.many(cnt), dtype=np.int64))
                    ncoord = 3 + (edim if par else 0)
                    xyz.append(np.array(tk.many(cnt * ncoord, float)).reshape(cnt, ncoord)[:, :3])
            mesh.node_tags = np.concatenate(tags) if tags else np.zeros(0, dtype=np.int64)
            mesh.nodes = np.concatenate(xyz) if xyz else np.zeros((0, 3))
            if len(mesh.node_tags) != nn:
                raise ValueError("msh: node count mismatch")
        elif s == "$Elements":
            if mesh.nodes is None:
                raise ValueError("msh: $Elements before $Nodes")
            index = np.full(int(mesh.node_tags.max()) + 1 if len(mesh.node_tags) else 1, -1, dtype=np.int64)
            index[mesh.node_tags] = np.arange(len(mesh.node_tags))
            if binary:
                nblk, _, _, _ = r.unpack("<4Q")
                for _ in range(nblk):
                    edim, etag, etype = r.unpack("<3i")
                    cnt, = r.unpack(sz)
                    if etype not in ELEMENT_TYPES:
                        raise ValueError(f"msh: element type {etype} not supported")
                    nn = ELEMENT_TYPES[etype][2]
                    rec = r.array("<u8", cnt * (nn + 1)).reshape(cnt, nn + 1).astype(np.int64)
                    mesh.blocks.append(_block(edim, etag, etype, rec, index))
                r.skip_to("$EndElements")
            else:
                tk = _Tokens(_section_text(r, "$EndElements"))
                nblk = tk.next()
                tk.many(3)
                for _ in range(nblk):
                    edim, etag, etype, cnt = tk.many(4)
                    if etype not in ELEMENT_TYPES:
                        raise ValueError(f"msh: element type {etype} not supported")
                    nn = ELEMENT_TYPES[etype][2]
                    rec = np.array(tk.many(cnt * (nn + 1)), dtype=np.int64).reshape(cnt, nn + 1)
                    mesh.blocks.append(_block(edim, etag, etype, rec, index))
        elif s.startswith("$") and not s.startswith("$End"):
            r.skip_to("$End" + s[1:])
    if mesh.nodes is None:
        raise ValueError("msh: no $Nodes section")
    return mesh


def _block(edim, etag, etype, rec, index):
    nodes = index[rec[:, 1:]]
    if (nodes < 0).any():
        raise ValueError("msh: element refers to an unknown node")
    perm = NODE_PERMUTATION.get(etype)
    if perm is not None:
        nodes = nodes[:, perm]
    return {"dim": edim, "entity": etag, "type": etype, "tags": rec[:, 0], "nodes": nodes}


def _p2_triangles(mesh, nodes3, name):
    """P2 node values (ne, 6, 3) in the reference order from tri6 (already
    permuted) or tri3 (midpoints of straight edges)."""
    X = mesh.nodes
    if name == "tri6":
        return X[nodes3]
    v0, v1, v2 = X[nodes3[:, 0]], X[nodes3[:, 1]], X[nodes3[:, 2]]
    return np.stack([v0, 0.5 * (v0 + v1), v1, 0.5 * (v0 + v2), 0.5 * (v1 + v2), v2], axis=1)


def _p2_lines(mesh, nodes, name):
    X = mesh.nodes
    if name == "line3":
        return X[nodes]
    v0, v1 = X[nodes[:, 0]], X[nodes[:, 1]]
    return np.stack([v0, 0.5 * (v0 + v1), v1], axis=1)


def _surface_orientation(mesh, volume, surface, sign_in_file):
    """+1 if the triangles of `surface` (as stored) have normals pointing out
    of `volume`, -1 if into it: a vote over the triangles that are faces of
    the volume's tetrahedra (the outward side is away from the tetrahedron's
    fourth vertex). The signs of the file's bounding lists are used only when
    the volume has no tetrahedra: they do not always match the element
    orientation (on docs/assets/piece.msh, 4 of the 10 faces of volume 1)."""
    X = mesh.nodes
    tets = [n[:, :4] for name, n in mesh.elements(3, entities={volume}) if name in ("tet4", "tet10")]
    if not tets:
        return 1 if sign_in_file > 0 else -1
    opp = {}
    for t in np.concatenate(tets):
        for f, o in (((0, 1, 2), 3), ((0, 1, 3), 2), ((0, 2, 3), 1), ((1, 2, 3), 0)):
            opp[tuple(sorted(t[list(f)]))] = t[o]
    vote = 0.0
    for name, tri in mesh.elements(2, entities={surface}, type_names={"tri3", "tri6"}):
        v = tri[:, [0, 2, 5]] if name == "tri6" else tri  # vertices (reference order for tri6)
        for row in v:
            o = opp.get(tuple(sorted(row)))
            if o is None:
                continue
            nrm = np.cross(X[row[1]] - X[row[0]], X[row[2]] - X[row[0]])
            vote += np.sign(np.dot(nrm, X[row].mean(axis=0) - X[o]))
    if vote == 0.0:
        return 1 if sign_in_file > 0 else -1
    return 1 if vote > 0 else -1


def boundary_quadrature(mesh, volumes=None, qorder=4):
    """Quadrature on the boundary of 3D entities `volumes` (all if None): the
    triangles of their bounding surfaces, surfaces shared by two of the
    volumes left out, each surface oriented so that its normals point out of
    its volume (Inti.boundary of a Domain, then Quadrature(msh[Γ]; qorder))."""
    from .quadrature import Quadrature, _triangle_quadrature
    vols = mesh.entity_tags(3) if volumes is None else list(volumes)
    owner = {}
    for v in vols:
        for s in mesh.entities[(3, v)]["bounding"]:
            t = abs(s)
            owner[t] = None if t in owner else (v, s)
    vals = []
    for t, o in sorted(owner.items()):
        if o is None:
            continue
        sign = _surface_orientation(mesh, o[0], t, o[1])
        for name, nodes in mesh.elements(2, entities={t}, type_names={"tri3", "tri6"}):
            v = _p2_triangles(mesh, nodes, name)
            if sign < 0:  # reverse the orientation: swap the second and third vertices
                v = v[:, [0, 3, 5, 1, 4, 2]]
            vals.append(v)
    if not vals:
        raise ValueError("no boundary triangles found")
    vals = np.concatenate(vals)
    coords, normals, w, qtags = _triangle_quadrature(vals, qorder)
    return Quadrature(coords, normals, w, qtags, elements=vals, name="gmsh-boundary", qorder=qorder)


def entity_quadrature(mesh, dim, entities=None, qorder=None, ambient=None):
    """Quadrature on the elements of dimension `dim` of the given entities:
    dim 1 (lines, 2D meshes: Gauss-Legendre, normals (t2, -t1)), dim 2 in a
    3D mesh (surface, Vioreanu-Rokhlin, normals from the element orientation)
    or dim 2 in a 2D mesh (planar volume, Vioreanu-Rokhlin, no normals)."""
    from .quadrature import (Quadrature, _line_quadrature, _triangle_quadrature,
                             _triangle_quadrature_2d)
    amb = ambient or (2 if np.allclose(mesh.nodes[:, 2], 0.0) else 3)
    blocks = mesh.elements(dim, entities=entities)
    if dim == 1:
        vals = np.concatenate([_p2_lines(mesh, n, nm) for nm, n in blocks if nm in ("line2", "line3")])
        q = qorder or 8
        coords, normals, w, qtags = _line_quadrature(vals[..., :2], q)
        return Quadrature(coords, normals, w, qtags, elements=vals[..., :2], name="gmsh-lines", qorder=q)
    vals = np.concatenate([_p2_triangles(mesh, n, nm) for nm, n in blocks if nm in ("tri3", "tri6")])
    q = qorder or 4
    if amb == 2:
        coords, w, qtags = _triangle_quadrature_2d(vals[..., :2], q)
        return Quadrature(coords, None, w, qtags, elements=vals[..., :2], name="gmsh-volume", qorder=q)
    coords, normals, w, qtags = _triangle_quadrature(vals, q)
    return Quadrature(coords, normals, w, qtags, elements=vals, name="gmsh-surface", qorder=q)


def tetra_volume(mesh, volumes=None):
    """Sum of the (straight) tetrahedron volumes of the given 3D entities."""
    X = mesh.nodes
    total = 0.0
    for name, nodes in mesh.elements(3, entities=None if volumes is None else set(volumes)):
        a, b, c, d = (X[nodes[:, i]] for i in range(4))
        total += np.abs(np.einsum("ij,ij->i", b - a, np.cross(c - a, d - a))).sum() / 6.0
    return total


def write_msh(path, nodes, blocks, entities=None, physical_names=None):
    """Write an ASCII MSH 4.1 file. nodes (N, 3); blocks: (dim, entity_tag,
    gmsh_type, elements (ne, nn) 0-based node indices in the reference order,
    permuted back to Gmsh's); entities: {(dim, tag): {"physical": [...],
    "bounding": [...]}} (defaults: one untagged entity per block)."""
    inverse = {t: np.argsort(p) for t, p in NODE_PERMUTATION.items()}
    ents = dict(entities or {})
    for dim, tag, _, _ in blocks:
        ents.setdefault((dim, tag), {"physical": [], "bounding": []})
    out = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat"]
    if physical_names:
        out.append("$PhysicalNames")
        out.append(str(len(physical_names)))
        for (d, t), name in sorted(physical_names.items()):
            out.append(f'{d} {t} "{name}"')
        out.append("$EndPhysicalNames")
    counts = [sum(1 for (d, _) in ents if d == k) for k in range(4)]
    out.append("$Entities")
    out.append(" ".join(map(str, counts)))
    for dim in range(4):
        for (d, t), e in sorted(ents.items()):
            if d != dim:
                continue
            box = "0 0 0" if dim == 0 else "0 0 0 0 0 0"
            rec = [str(t), box, str(len(e["physical"]))] + [str(x) for x in e["physical"]]
            if dim > 0:
                rec += [str(len(e["bounding"]))] + [str(x) for x in e["bounding"]]
            out.append(" ".join(rec))
    out.append("$EndEntities")
    n = len(nodes)
    out.append("$Nodes")
    out.append(f"1 {n} 1 {n}")
    out.append(f"0 1 0 {n}")
    out.extend(str(i + 1) for i in range(n))
    out.extend(f"{x:.17g} {y:.17g} {z:.17g}" for x, y, z in nodes)
    out.append("$EndNodes")
    ne = sum(len(b[3]) for b in blocks)
    out.append("$Elements")
    out.append(f"{len(blocks)} {ne} 1 {ne}")
    tag = 1
    for dim, etag, etype, els in blocks:
        out.append(f"{dim} {etag} {etype} {len(els)}")
        inv = inverse.get(etype)
        for row in np.asarray(els):
            r = row[inv] if inv is not None else row
            out.append(" ".join([str(tag)] + [str(int(v) + 1) for v in r]))
            tag += 1
    out.append("$EndElements")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
