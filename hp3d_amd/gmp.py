"""GMP geometry files for hexahedral meshes: reader and writer.

The file format is the one hp3D reads in GMP/base/input_DEFAULT.F90:98-644
(points, 'Seglin' curves, 'BilQua' rectangles, 'Linear' hexas; no surfaces,
triangles, prisms, tetrahedra or pyramids), e.g.
problems/POISSON/GALERKIN/geometries/hexa27_orient_0.

Reading restates the connectivity step of GMP/base/connect.F90:265-380 that
hp3gen (meshgen/hp3gen.F90:160-199) turns into element orientation codes:

* hexa edge j runs from vertex BRICK_EDGE_TO_VERT(1, j) to (2, j); its curve
  is the curve through both points; nori_e(j) = 1 when the curve starts at
  the edge's second vertex (the negative EdgeNo of connect.F90:293-301);
* hexa face j has edges BRICK_FACE_TO_EDGE(1:4, j); its rectangle is the one
  bounded by their curves; with k1, k2 the positions (1..4, QUADR_EDGE_TO_VERT
  order) of the curves of the face's first and fourth edges among the
  rectangle's edges, the face orientation is the code of k1*10 + k2 in
  {14: 0, 43: 1, 32: 2, 21: 3, 41: 4, 12: 5, 23: 6, 34: 7} (connect.F90:354-378).

The global edge and face nodes are the curves and rectangles; their orders
follow the minimum rule as for every HexMesh.  Writing emits a HexMesh with
one curve per edge node (from its lower to its higher vertex number) and one
rectangle per face node with vertices (origin, s end, opposite, t end), so
that reading the file back reproduces the mesh's orientation codes.
"""
import numpy as np

from .mesh import FACE_VERTS, HexMesh, _orders

# element_data.F90:67-70, 113-116 (0-based)
BRICK_EDGE_TO_VERT = np.array([(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7),
                               (0, 4), (1, 5), (2, 6), (3, 7)])
BRICK_FACE_TO_EDGE = np.array([(0, 1, 2, 3), (4, 5, 6, 7), (0, 9, 4, 8), (1, 10, 5, 9),
                               (2, 10, 6, 11), (3, 11, 7, 8)])
QUADR_EDGE_TO_VERT = np.array([(0, 1), (1, 2), (3, 2), (0, 3)])
FACE_NICK_CODE = {14: 0, 43: 1, 32: 2, 21: 3, 41: 4, 12: 5, 23: 6, 34: 7}


def _tokens(path):
    """Whitespace tokens of the file with Fortran 'd' exponents made readable."""
    with open(path) as f:
        return f.read().split()


def read_gmp(path, porder=2, coef=None):
    """HexMesh of the Linear hexas of a GMP file (vertex numbers as in the
    file, 0-based), with hp3gen's orientation codes and minimum-rule orders."""
    tok = _tokens(path)
    pos = [0]

    def nxt():
        t = tok[pos[0]]
        pos[0] += 1
        return t

    def num():
        return int(nxt())

    def skip_label():          # "NAME" after a count, e.g. "64 NRPOINT"
        pos[0] += 1

    ndim, mandim = num(), num()
    pos[0] += 1                # NDIM,MANDIM label
    if ndim != 3 or mandim != 3:
        raise ValueError("only 3D GMP files")
    nsurf = num(); skip_label()
    if nsurf:
        raise ValueError("surfaces are not supported")
    num(); skip_label()        # NRDOMAIN
    num(); skip_label()        # BDOM_FLAG
    npt = num(); skip_label()
    verts = np.zeros((npt, 3))
    for i in range(npt):
        typ = nxt()
        if typ != "Regular":
            raise ValueError("point type %s" % typ)
        pos[0] += 2            # "point", index
        verts[i] = [float(nxt().replace("d", "e").replace("D", "e")) for _ in range(3)]
    ncur = num(); skip_label()
    curves = np.zeros((ncur, 2), np.int64)
    for i in range(ncur):
        typ = nxt()
        if typ != "Seglin":
            raise ValueError("curve type %s" % typ)
        pos[0] += 2
        curves[i] = [num() - 1, num() - 1]
    if num() != 0:
        raise ValueError("triangles are not supported")
    skip_label()
    nrec = num(); skip_label()
    rects = np.zeros((nrec, 4), np.int64)
    for i in range(nrec):
        typ = nxt()
        if typ not in ("BilQua", "TraQua"):
            raise ValueError("rectangle type %s" % typ)
        pos[0] += 2
        rects[i] = [num() - 1 for _ in range(4)]
    if num() != 0:
        raise ValueError("prisms are not supported")
    skip_label()
    nhex = num(); skip_label()
    elems = np.zeros((nhex, 8), np.int64)
    for i in range(nhex):
        typ = nxt()
        if typ != "Linear":
            raise ValueError("hexa type %s" % typ)
        pos[0] += 2
        num()                  # domain
        elems[i] = [num() - 1 for _ in range(8)]
    return _from_blocks(verts, curves, rects, elems, porder, coef)


def _from_blocks(verts, curves, rects, elems, porder, coef):
    nel = elems.shape[0]
    ckey = {}
    for c, (a, b) in enumerate(curves):
        ckey[(min(a, b), max(a, b))] = c
    # rectangle edges as curve numbers (QUADR_EDGE_TO_VERT order)
    redge = np.zeros((rects.shape[0], 4), np.int64)
    rkey = {}
    for r, rv in enumerate(rects):
        for k, (i, j) in enumerate(QUADR_EDGE_TO_VERT):
            a, b = rv[i], rv[j]
            redge[r, k] = ckey[(min(a, b), max(a, b))]
        rkey[tuple(sorted(redge[r]))] = r
    elem_edges = np.zeros((nel, 12), np.int64)
    nori_e = np.zeros((nel, 12), np.int32)
    elem_faces = np.zeros((nel, 6), np.int64)
    nori_f = np.zeros((nel, 6), np.int32)
    for e in range(nel):
        v = elems[e]
        for j, (i1, i2) in enumerate(BRICK_EDGE_TO_VERT):
            a, b = v[i1], v[i2]
            c = ckey[(min(a, b), max(a, b))]
            elem_edges[e, j] = c
            nori_e[e, j] = 1 if curves[c, 0] == b else 0
        for j in range(6):
            fe = elem_edges[e, BRICK_FACE_TO_EDGE[j]]
            r = rkey[tuple(sorted(fe))]
            k1 = int(np.nonzero(redge[r] == fe[0])[0][0]) + 1
            k2 = int(np.nonzero(redge[r] == fe[3])[0][0]) + 1
            elem_faces[e, j] = r
            nori_f[e, j] = FACE_NICK_CODE[k1 * 10 + k2]
    p = np.asarray(porder)
    if p.ndim == 0:
        p = np.full((nel, 3), int(p))
    elif p.ndim == 1:
        p = np.broadcast_to(p, (nel, 3)).copy()
    c = np.ones(nel) if coef is None else np.broadcast_to(np.asarray(coef, float), (nel,)).copy()
    lo = np.minimum(verts.min(0), 0.0)
    box = tuple(verts.max(0) - lo)
    m = HexMesh(verts=verts, elems=elems, porder=p.astype(np.int64), coef=c, shape=(nel, 1, 1),
                box=box)
    m.elem_edges, m.nori_e = elem_edges, nori_e
    m.elem_faces, m.nori_f = elem_faces, nori_f
    m.edges = curves.copy()
    # face-node corners (origin, s end, opposite, t end) are the rectangle's vertices
    m.faces = rects.copy()
    _orders(m)
    return m


def write_gmp(m, path):
    """Write a HexMesh as a GMP file of points, Seglin curves, BilQua
    rectangles and Linear hexas (all in domain 1)."""
    lines = ["3   3     NDIM,MANDIM", "", "0         NRSURFS", "", "1         NRDOMAIN", "",
             "0         BDOM_FLAG", "", "%d        NRPOINT" % m.verts.shape[0], ""]
    for i, x in enumerate(m.verts):
        lines += ["Regular             point %d" % (i + 1),
                  " ".join("%.17e" % c for c in x), ""]
    lines += ["", "%d          NRCURVE" % m.edges.shape[0], ""]
    for i, (a, b) in enumerate(m.edges):
        lines += ["Seglin              curve %d" % (i + 1), "%d %d" % (a + 1, b + 1), ""]
    lines += ["0             NRTRIAN", "", "%d          NRRECTA" % m.faces.shape[0], ""]
    for i, f in enumerate(m.faces):
        lines += ["BilQua              rectangle %d" % (i + 1), " ".join(str(v + 1) for v in f), ""]
    lines += ["0     NRPRISM", "", "%d          NRHEXAS" % m.nel, ""]
    for i, v in enumerate(m.elems):
        lines += ["Linear              hexa %d" % (i + 1), "1    " + " ".join(str(x + 1) for x in v), ""]
    lines += ["0     NRTETRA", "", "0     NRPYRAM", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))


__all__ = ["read_gmp", "write_gmp", "FACE_VERTS"]
