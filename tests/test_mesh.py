"""Host-side mesh logic: orientation codes, minimum rule, dof counts.

The decisive check is trace conformity: on a face shared by two elements the
global face and edge dofs, evaluated from either side with the oracle's
shape functions (pinned bitwise against the reference), give the same trace.
That holds only if the orientation codes the mesh assigns mean what
OrientE / OrientQuad (src/element/shape_1/Orient.F90) mean.
"""
import numpy as np
import pytest

import oracle
from hp3d_amd import mesh
from hp3d_amd.elem import H1, HCURL, ndof, nquad

# element faces that meet: (face of A, face of B, direction of the step A -> B)
PAIRS = [(3, 5, 0), (4, 2, 1), (1, 0, 2)]   # x: face4|face6, y: face5|face3, z: face2|face1
FACE_NORMAL = [2, 2, 1, 0, 1, 0]
FACE_VAL = [0, 1, 0, 1, 1, 0]


def dof_offsets(space, no):
    """Start of the dofs of each edge / face in hp3D's element order."""
    pe, fh, fv = no[:12], no[12:18] // 10, no[12:18] % 10
    off = 8 if space == H1 else 0
    eo, fo = [], []
    for e in range(12):
        eo.append(off)
        off += pe[e] - 1 if space == H1 else pe[e]
    for f in range(6):
        fo.append(off)
        off += (fh[f] - 1) * (fv[f] - 1) if space == H1 else fh[f] * (fv[f] - 1) + (fh[f] - 1) * fv[f]
    return eo, fo


def face_points(fa, rng, k=5):
    pts = rng.uniform(0.05, 0.95, (k, 3))
    pts[:, FACE_NORMAL[fa]] = FACE_VAL[fa]
    return pts


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("space", [H1, HCURL])
def test_trace_conformity_across_faces(seed, space):
    rng = np.random.default_rng(seed)
    shape = (3, 3, 3)
    porder = rng.integers(1, 5, size=(27, 3))
    m = mesh.structured(shape, porder=porder, perturb=0.1, mesh_seed=seed, perm_seed=100 + seed)
    nx, ny, nz = shape
    eid = lambda i, j, k: i + nx * (j + ny * k)
    shp = oracle.shape3dh if space == H1 else oracle.shape3de
    checked = 0
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                a = eid(i, j, k)
                for fa, fb, d in PAIRS:
                    ijk = [i, j, k]
                    ijk[d] += 1
                    if ijk[d] >= shape[d]:
                        continue
                    b = eid(*ijk)
                    assert m.elem_faces[a, fa] == m.elem_faces[b, fb]
                    pa = face_points(fa, rng)
                    pb = pa.copy()
                    pb[:, FACE_NORMAL[fb]] = FACE_VAL[fb]
                    _, foa = dof_offsets(space, m.norder[a])
                    _, fob = dof_offsets(space, m.norder[b])
                    assert m.norder[a, 12 + fa] == m.norder[b, 12 + fb]
                    h, v = m.norder[a, 12 + fa] // 10, m.norder[a, 12 + fa] % 10
                    nfd = (h - 1) * (v - 1) if space == H1 else h * (v - 1) + (h - 1) * v
                    tang = [c for c in range(3) if c != FACE_NORMAL[fa]]
                    if nfd == 0:
                        continue
                    for xa, xb in zip(pa, pb):
                        sa = shp(xa, m.norder[a], m.nori_e[a], m.nori_f[a])[0]
                        sb = shp(xb, m.norder[b], m.nori_e[b], m.nori_f[b])[0]
                        ua = sa[foa[fa]:foa[fa] + nfd]
                        ub = sb[fob[fb]:fob[fb] + nfd]
                        if space == HCURL:
                            ua, ub = ua[:, tang], ub[:, tang]
                        assert np.max(np.abs(ua - ub)) < 1e-13
                        checked += 1
    assert checked > 0


def test_edge_conformity_all_orientations():
    m = mesh.structured((2, 2, 1), porder=4, perturb=0.0, perm_seed=9)
    rng = np.random.default_rng(0)
    # elements 0 and 1 share the x = 1 face of element 0; its edges 10, 11 (v2v6, v3v7)
    # are edges 9, 12 (v1v5, v4v8) of element 1
    for ea, eb in ((9, 8), (10, 11)):
        assert m.elem_edges[0, ea] == m.elem_edges[1, eb]
        _eo_a, _ = dof_offsets(H1, m.norder[0])
        _eo_b, _ = dof_offsets(H1, m.norder[1])
        if m.norder[0, ea] < 2:
            continue
        for t in rng.uniform(0.05, 0.95, 5):
            xa = np.array([1.0, 0.0 if ea == 9 else 1.0, t])
            xb = np.array([0.0, xa[1], t])
            sa = oracle.shape3dh(xa, m.norder[0], m.nori_e[0], m.nori_f[0])[0]
            sb = oracle.shape3dh(xb, m.norder[1], m.nori_e[1], m.nori_f[1])[0]
            na = m.norder[0, ea] - 1
            assert np.max(np.abs(sa[_eo_a[ea]:_eo_a[ea] + na] - sb[_eo_b[eb]:_eo_b[eb] + na])) < 1e-14


def test_orientation_codes_cover_all_values():
    m = mesh.config_mesh("cfg1")
    assert set(np.unique(m.nori_f)) == set(range(8))
    assert set(np.unique(m.nori_e)) == {0, 1}
    # natural numbering gives the trivial codes
    m0 = mesh.config_mesh("cfg0")
    assert not m0.nori_e.any() and not m0.nori_f.any()


def test_minimum_rule():
    rng = np.random.default_rng(3)
    porder = rng.integers(1, 7, size=(4 * 4 * 4, 3))
    m = mesh.structured((4, 4, 4), porder=porder, perm_seed=1)
    # every edge order is the minimum over its elements of their order along it
    for e in range(m.nel):
        for le in range(12):
            g = m.elem_edges[e, le]
            els, locs = np.nonzero(m.elem_edges == g)
            d = mesh.EDGE_DIR[le]
            assert m.norder[e, le] == min(porder[x, mesh.EDGE_DIR[l]] for x, l in zip(els, locs))
            assert m.norder[e, le] <= porder[e, d]


def test_counts_match_oracle_and_configs():
    m = mesh.config_mesh("cfg2")
    n, nb = ndof(H1, m.norder)
    ne_, nbe = ndof(HCURL, m.norder)
    for e in range(0, m.nel, 37):
        assert n[e] == oracle.ndof_h1(m.norder[e]) and nb[e] == oracle.nbubble_h1(m.norder[e][18])
        assert ne_[e] == oracle.ndof_hcurl(m.norder[e])
        assert nbe[e] == oracle.nbubble_hcurl(m.norder[e][18])
        xi, _w = oracle.set_3d_int(m.norder[e], m.nori_f[e])
        assert xi.shape[0] == np.prod(nquad(m.norder[e], m.nori_f[e])[0])
    assert m.nel == 20000 and n.min() >= 8 and n.max() <= 343
    c1 = mesh.config_mesh("cfg1")
    assert c1.nel == 32768 and set(ndof(H1, c1.norder)[0]) == {125}
    c3 = mesh.config_mesh("cfg3")
    assert set(ndof(HCURL, c3.norder)[0]) == {540} and set(ndof(HCURL, c3.norder)[1]) == {240}
    # interior vertices move by at most 0.15 h, boundary vertices stay on the boundary
    v = c1.verts
    assert np.all(v >= -1e-15) and np.all(v <= 1 + 1e-15)
