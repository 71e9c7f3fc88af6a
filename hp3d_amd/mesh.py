"""Synthetic hexahedral meshes with hp3D's element codes.

A structured nx x ny x nz grid of trilinear hexahedra ('Linear' GMP blocks),
with per-element order codes, edge/face orientation codes and global node
numbering in hp3D's conventions:

* element vertex order v1..v8 = (000, 100, 110, 010, 001, 101, 111, 011)
  (element_data / BlendHexaV, src/element/shape_1/BlendProject.F90:197-253);
* local edges (v1v2, v2v3, v4v3, v1v4, v5v6, v6v7, v8v7, v5v8, v1v5, v2v6,
  v3v7, v4v8) (BlendProject.F90:270-273);
* local faces (1234, 5678, 1265, 2376, 4378, 1485) with local axes from the
  first vertex to the second (s) and to the fourth (t) (BlendProject.F90:380-383);
* norder(19): 12 edge orders, 6 face orders h*10+v in face-node axes,
  interior x*100+y*10+z (find_order, src/datstrs/find_order.F90:5-58);
* nori_e in {0,1}, nori_f in {0..7} with OrientQuad's meaning
  (src/element/shape_1/Orient.F90:39-116).

Global edge nodes run from their lower to their higher global vertex number;
a global face node has its origin at its lowest-numbered vertex and its s axis
towards the lower-numbered of the two neighbouring vertices.  Edge and face
orders follow hp3D's minimum rule (src/adapt/enforce_min_rule.F90): the order
of a shared node is the minimum, per direction, of the orders its elements
imply.  A seeded permutation of the global vertex numbers gives every
orientation code a chance to occur (hp3gen numbers nodes from the geometry
file; any consistent numbering is admissible).
"""
from dataclasses import dataclass, field

import numpy as np

MAXP = 6

EDGE_VERTS = np.array([(0, 1), (1, 2), (3, 2), (0, 3), (4, 5), (5, 6), (7, 6), (4, 7),
                       (0, 4), (1, 5), (2, 6), (3, 7)])
EDGE_DIR = np.array([0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2])
FACE_VERTS = np.array([(0, 1, 2, 3), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (3, 2, 6, 7),
                       (0, 3, 7, 4)])
FACE_AXES = np.array([(0, 1), (0, 1), (0, 2), (1, 2), (0, 2), (1, 2)])
VERT_XYZ = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1),
                     (1, 1, 1), (0, 1, 1)])

# OrientQuad (Orient.F90:39-116): local pair feeding the face-node s and t
# axes (0 = local s, 1 = local t) and whether each is reversed.
FACE_ORI = np.array([
    # ps, pt, rs, rt
    (0, 1, 0, 0), (1, 0, 0, 1), (0, 1, 1, 1), (1, 0, 1, 0),
    (1, 0, 0, 0), (0, 1, 1, 0), (1, 0, 1, 1), (0, 1, 0, 1)])


def _face_ori_table():
    """code[m, sdir]: orientation code whose face-node origin is local corner m
    (0..3 in the order (0,0),(1,0),(1,1),(0,1)) and whose s axis goes to the
    next corner (sdir=0) or to the previous one (sdir=1)."""
    corners = np.array([(0, 0), (1, 0), (1, 1), (0, 1)])
    table = -np.ones((4, 2), dtype=np.int32)
    for code, (ps, pt, rs, rt) in enumerate(FACE_ORI):
        gs = np.where(rs, 1 - corners[:, ps], corners[:, ps])
        gt = np.where(rt, 1 - corners[:, pt], corners[:, pt])
        m = int(np.nonzero((gs == 0) & (gt == 0))[0][0])
        s_to = int(np.nonzero((gs == 1) & (gt == 0))[0][0])
        sdir = 0 if s_to == (m + 1) % 4 else 1
        assert table[m, sdir] < 0
        table[m, sdir] = code
    assert (table >= 0).all()
    return table


FACE_CODE = _face_ori_table()
SWAPPED = np.array([0, 1, 0, 1, 1, 0, 1, 0])  # NFAXES(3,:)


@dataclass
class HexMesh:
    verts: np.ndarray            # (nv, 3)
    elems: np.ndarray            # (nel, 8) global vertex ids
    porder: np.ndarray           # (nel, 3) interior orders
    coef: np.ndarray             # (nel,) kappa_e or eps_e
    shape: tuple                 # (nx, ny, nz)
    box: tuple = (1.0, 1.0, 1.0)
    norder: np.ndarray = field(default=None)   # (nel, 19) int32
    nori_e: np.ndarray = field(default=None)   # (nel, 12) int32
    nori_f: np.ndarray = field(default=None)   # (nel, 6) int32
    elem_edges: np.ndarray = field(default=None)  # (nel, 12) edge node ids
    elem_faces: np.ndarray = field(default=None)  # (nel, 6) face node ids
    edges: np.ndarray = field(default=None)    # (ned, 2) global vertex ids (low, high)
    faces: np.ndarray = field(default=None)    # (nfa, 4) face-node corners (origin, s, opp, t)
    edge_order: np.ndarray = field(default=None)
    face_order: np.ndarray = field(default=None)  # (nfa, 2) along face-node s, t

    @property
    def nel(self):
        return self.elems.shape[0]

    def xnod(self):
        """Geometry dofs of the trilinear map: (nel, 8, 3) vertex coordinates."""
        return self.verts[self.elems]

    def boundary_vertices(self):
        eps = 1e-12
        v = self.verts
        b = np.zeros(v.shape[0], bool)
        for d in range(3):
            b |= (np.abs(v[:, d]) < eps) | (np.abs(v[:, d] - self.box[d]) < eps)
        return b


def _topology(m: HexMesh):
    el = m.elems
    nel = el.shape[0]
    # edges
    ev = el[:, EDGE_VERTS]                       # (nel, 12, 2)
    lo = ev.min(-1)
    hi = ev.max(-1)
    key = lo.astype(np.int64) * (m.verts.shape[0] + 1) + hi
    ukey, einv = np.unique(key.ravel(), return_inverse=True)
    m.elem_edges = einv.reshape(nel, 12).astype(np.int64)
    m.edges = np.stack([ukey // (m.verts.shape[0] + 1), ukey % (m.verts.shape[0] + 1)], 1)
    m.nori_e = (ev[..., 0] > ev[..., 1]).astype(np.int32)
    # faces
    fv = el[:, FACE_VERTS]                       # (nel, 6, 4) local corners (0,0),(1,0),(1,1),(0,1)
    srt = np.sort(fv, axis=-1).astype(np.int64)
    nv1 = m.verts.shape[0] + 1
    fkey = ((srt[..., 0] * nv1 + srt[..., 1]) * nv1 + srt[..., 2]) * nv1 + srt[..., 3]
    ufk, finv = np.unique(fkey.ravel(), return_inverse=True)
    m.elem_faces = finv.reshape(nel, 6).astype(np.int64)
    mpos = np.argmin(fv, axis=-1)                # local corner of the face-node origin
    nxt = np.take_along_axis(fv, ((mpos + 1) % 4)[..., None], -1)[..., 0]
    prv = np.take_along_axis(fv, ((mpos + 3) % 4)[..., None], -1)[..., 0]
    sdir = (prv < nxt).astype(np.int64)          # s axis towards the lower neighbour
    m.nori_f = FACE_CODE[mpos, sdir].astype(np.int32)
    # face-node corners (origin, s-end, opposite, t-end) in global ids
    org = np.take_along_axis(fv, mpos[..., None], -1)[..., 0]
    s_end = np.where(sdir == 0, nxt, prv)
    t_end = np.where(sdir == 0, prv, nxt)
    opp = np.take_along_axis(fv, ((mpos + 2) % 4)[..., None], -1)[..., 0]
    faces = np.zeros((ufk.size, 4), np.int64)
    faces[m.elem_faces.ravel()] = np.stack([org, s_end, opp, t_end], -1).reshape(-1, 4)
    m.faces = faces


def _orders(m: HexMesh):
    """Minimum rule for edge and face nodes, then the norder(19) codes."""
    nel = m.nel
    p = m.porder
    ned = m.edges.shape[0]
    nfa = m.faces.shape[0]
    eo = np.full(ned, 99, np.int64)
    np.minimum.at(eo, m.elem_edges.ravel(), p[:, EDGE_DIR].ravel())
    # element direction of each face-node axis: local pair ps/pt -> element axis
    ori = FACE_ORI[m.nori_f]                                   # (nel, 6, 4)
    ax = np.broadcast_to(FACE_AXES, (nel, 6, 2))
    ds = np.take_along_axis(ax, ori[..., 0:1], -1)[..., 0]
    dt = np.take_along_axis(ax, ori[..., 1:2], -1)[..., 0]
    ps = np.take_along_axis(p[:, None, :].repeat(6, 1), ds[..., None], -1)[..., 0]
    pt = np.take_along_axis(p[:, None, :].repeat(6, 1), dt[..., None], -1)[..., 0]
    fo = np.full((nfa, 2), 99, np.int64)
    np.minimum.at(fo[:, 0], m.elem_faces.ravel(), ps.ravel())
    np.minimum.at(fo[:, 1], m.elem_faces.ravel(), pt.ravel())
    m.edge_order = eo
    m.face_order = fo
    no = np.zeros((nel, 19), np.int32)
    no[:, :12] = eo[m.elem_edges]
    fe = fo[m.elem_faces]
    no[:, 12:18] = fe[..., 0] * 10 + fe[..., 1]
    no[:, 18] = p[:, 0] * 100 + p[:, 1] * 10 + p[:, 2]
    m.norder = no


def structured(shape, box=(1.0, 1.0, 1.0), porder=2, perturb=0.0, mesh_seed=None,
               perm_seed=None, coef=None):
    """Structured hexahedral mesh.

    porder  int, (3,) or (nel, 3) interior orders
    perturb interior vertices move by U(-perturb*h, perturb*h) per coordinate
    coef    None (ones), scalar, or (nel,) array
    """
    nx, ny, nz = shape
    xs = [np.linspace(0.0, box[d], n + 1) for d, n in enumerate(shape)]
    gx, gy, gz = np.meshgrid(*xs, indexing="ij")
    verts = np.stack([gx.ravel(order="F"), gy.ravel(order="F"), gz.ravel(order="F")], 1)
    if perturb:
        rng = np.random.default_rng(mesh_seed)
        h = np.array([box[d] / shape[d] for d in range(3)])
        ii = np.stack(np.meshgrid(*[np.arange(n + 1) for n in shape], indexing="ij"), -1)
        ii = ii.reshape(-1, 3, order="F")
        interior = np.all((ii > 0) & (ii < np.array(shape)), axis=1)
        delta = rng.uniform(-perturb, perturb, size=verts.shape) * h
        verts = verts + delta * interior[:, None]
    vid = lambda i, j, k: i + (nx + 1) * (j + (ny + 1) * k)
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ex, ey, ez = ex.ravel(order="F"), ey.ravel(order="F"), ez.ravel(order="F")
    elems = np.stack([vid(ex + a, ey + b, ez + c) for a, b, c in VERT_XYZ], 1).astype(np.int64)
    nel = elems.shape[0]
    if perm_seed is not None:
        perm = np.random.default_rng(perm_seed).permutation(verts.shape[0])
        inv = np.empty_like(perm)
        inv[perm] = np.arange(perm.size)
        verts = verts[perm]
        elems = inv[elems]
    p = np.asarray(porder)
    if p.ndim == 0:
        p = np.full((nel, 3), int(p))
    elif p.ndim == 1:
        p = np.broadcast_to(p, (nel, 3)).copy()
    if (p < 1).any() or (p > MAXP).any():
        raise ValueError("orders must lie in [1, %d]" % MAXP)
    if coef is None:
        c = np.ones(nel)
    else:
        c = np.broadcast_to(np.asarray(coef, dtype=np.float64), (nel,)).copy()
    m = HexMesh(verts=verts, elems=elems, porder=p.astype(np.int64), coef=c,
                shape=tuple(shape), box=tuple(box))
    _topology(m)
    _orders(m)
    return m


# ------------------------------------------------------------------ configs

def config_mesh(name):
    """The synthetic meshes the workloads are defined on (seeds fixed here).

    cfg0  8^3 unit cube, regular grid, p = 3, kappa = 1
    cfg1  32^3 unit cube, +-0.15h interior perturbation, p = 4, kappa = exp(N(0,1))
    cfg2  20x20x50 in 1x1x2.5, +-0.1h, (px,py,pz) drawn from {1..6}: weighted to
          2-4, with 5-6 concentrated in a seeded ball of radius 0.3; lognormal kappa
    cfg3  16^3, +-0.1h, H(curl) p = 5, eps_e ~ U[1,4]
    cfg4  24^3, +-0.1h, p = 3, kappa = exp(theta), theta ~ N(0, 0.5^2)
    """
    if name == "cfg0":
        return structured((8, 8, 8), porder=3)
    if name == "cfg1":
        kappa = np.exp(np.random.default_rng(2001).normal(0.0, 1.0, 32 ** 3))
        return structured((32, 32, 32), porder=4, perturb=0.15, mesh_seed=1001,
                          perm_seed=5001, coef=kappa)
    if name == "cfg2":
        return mixed_order_mesh()
    if name == "cfg3":
        eps = np.random.default_rng(2003).uniform(1.0, 4.0, 16 ** 3)
        return structured((16, 16, 16), porder=5, perturb=0.1, mesh_seed=1003,
                          perm_seed=5003, coef=eps)
    if name == "cfg4":
        theta = np.random.default_rng(2004).normal(0.0, 0.5, 24 ** 3)
        return structured((24, 24, 24), porder=3, perturb=0.1, mesh_seed=1004,
                          perm_seed=5004, coef=np.exp(theta))
    raise KeyError(name)


# order distributions of the mixed-order mesh: P(p = 1..6) outside / inside the ball
P_OUT = np.array([0.10, 0.30, 0.30, 0.24, 0.04, 0.02])
P_IN = np.array([0.00, 0.00, 0.05, 0.15, 0.40, 0.40])


def mixed_order_mesh(shape=(20, 20, 50), box=(1.0, 1.0, 2.5)):
    nel = shape[0] * shape[1] * shape[2]
    centre = np.random.default_rng(3002).uniform([0.3, 0.3, 0.3], [0.7, 0.7, 2.2])
    rng = np.random.default_rng(2002)
    # element centroids of the unperturbed grid, x fastest
    h = np.array(box) / np.array(shape)
    ex, ey, ez = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    cen = (np.stack([ex.ravel(order="F"), ey.ravel(order="F"), ez.ravel(order="F")], 1) + 0.5) * h
    inside = np.linalg.norm(cen - centre, axis=1) < 0.3
    pin = rng.choice(np.arange(1, 7), size=(nel, 3), p=P_IN)
    pout = rng.choice(np.arange(1, 7), size=(nel, 3), p=P_OUT)
    porder = np.where(inside[:, None], pin, pout)
    kappa = np.exp(np.random.default_rng(4002).normal(0.0, 1.0, nel))
    return structured(shape, box=box, porder=porder, perturb=0.1, mesh_seed=1002,
                      perm_seed=5002, coef=kappa)
