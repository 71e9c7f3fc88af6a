"""Solution-level checks on the CPU: the oracle's element systems with the
reference's own numbering, Dirichlet data and lift reproduce what the
reference's tests assert, as they state it (test/poly_pois.F90: u = xyz at
p = 1, two uniform refinements of the unit cube, H1 error <= 1e-15;
test/conv_pois.F90: u = sin(pi x) sinh(pi y) z^2 at p = 2, three
refinements, rate in [1.95, 2.05]) and the boundary-vanishing analogues; the
product's host logic (first-visit numbering, NEXTRACT layout, Dirichlet
interpolation) against independent restatements; and the CPU adjoint
gradient against central finite differences.  These pin the FP64 CPU
reference that the GPU solution tests compare against."""
import numpy as np
import pytest

import oracle
from hp3d_amd import dirichlet, mesh
from hp3d_amd.mesh import FACE_AXES, FACE_VERTS, VERT_XYZ
from hp3d_amd.solve import DofMap

import solution_ref as sr


def unit_cube(nref, porder, seed=None):
    """The single-hexahedron unit cube of files/mesh/hexa_orient_0 after nref
    global h-refinements (global_href): a 2^nref grid."""
    n = 2 ** nref
    return mesh.structured((n, n, n), porder=porder, perm_seed=seed)


def test_poly_pois_as_stated():
    # test/poly_pois.F90:60-101: order 111, Dirichlet data u = xyz on all faces,
    # f = 0, two refinements; sqrt(sum errorH) <= 1e-15 after each
    for nref in (1, 2):
        m = unit_cube(nref, 1, seed=nref)
        ud = dirichlet.dirichlet_values(m, DofMap(m, 0), oracle.LOAD_XYZ)
        r = sr.solve(m, 0, oracle.LOAD_ZERO, exact=oracle.LOAD_XYZ, ud=ud)
        assert np.sqrt(r["error2"]) <= 1e-15, (nref, np.sqrt(r["error2"]))


def test_conv_pois_as_stated():
    # test/conv_pois.F90:35-75: order 222, u = sin(pi x) sinh(pi y) z^2 with its
    # Dirichlet data on all faces, three refinements, rate in [1.95, 2.05]
    errs = []
    for nref in (1, 2, 3):
        m = unit_cube(nref, 2, seed=nref)
        ud = dirichlet.dirichlet_values(m, DofMap(m, 0), oracle.LOAD_CONV)
        errs.append(np.sqrt(sr.solve(m, 0, oracle.LOAD_CONV, ud=ud)["error2"]))
    for i in (1, 2):
        rate = np.log(errs[i - 1] / errs[i]) / np.log(2.0)
        assert 1.95 <= rate <= 2.05, (errs, rate)


@pytest.mark.parametrize("porder,tol", [(1, 1e-14), (2, 1e-13), (3, 1e-12), ("mixed", 1e-12)])
def test_dirichlet_interpolation_reproduces_polynomials(porder, tol):
    """update_Ddof's projection-based interpolation reproduces the exact
    solution on the boundary when it lies in the trace space: the boundary
    values, summed with the oracle's shape functions at random boundary
    points, equal u (all orientation codes present)."""
    polys = {1: lambda x: (x[:, 0] * x[:, 1] * x[:, 2],
                           np.stack([x[:, 1] * x[:, 2], x[:, 0] * x[:, 2], x[:, 0] * x[:, 1]], 1)),
             2: lambda x: (x[:, 0] ** 2 * x[:, 1] + x[:, 2] ** 2,
                           np.stack([2 * x[:, 0] * x[:, 1], x[:, 0] ** 2, 2 * x[:, 2]], 1))}
    cubic = lambda x: (x[:, 0] ** 3 * x[:, 1] ** 2 - x[:, 2] ** 3 * x[:, 1],
                       np.stack([3 * x[:, 0] ** 2 * x[:, 1] ** 2,
                                 2 * x[:, 0] ** 3 * x[:, 1] - x[:, 2] ** 3,
                                 -3 * x[:, 2] ** 2 * x[:, 1]], 1))
    if porder == "mixed":
        po = np.random.default_rng(1).integers(3, 5, size=(8, 3))
        m, u_fn = mesh.structured((2, 2, 2), porder=po, perm_seed=7), cubic
    else:
        m = mesh.structured((2, 2, 2), porder=porder, perm_seed=3 + porder)
        u_fn = polys.get(porder, cubic)
    dm = DofMap(m, 0)
    ud = dirichlet.dirichlet_values(m, dm, u_fn)
    rng = np.random.default_rng(0)
    xn = m.xnod()
    worst = 0.0
    for e in range(m.nel):
        for lf in range(6):
            if not dm.fixed[2][m.elem_faces[e, lf]]:
                continue
            ni = dm.ni[e]
            for _ in range(4):
                xi = VERT_XYZ[FACE_VERTS[lf][0]].astype(float)
                xi[list(FACE_AXES[lf])] = rng.uniform(0, 1, 2)
                sv = np.asarray(oracle.shape3dh(xi, m.norder[e], m.nori_e[e], m.nori_f[e])[0])
                x = sv[:8] @ xn[e]
                worst = max(worst, abs(sv[:ni] @ ud[e, :ni] - u_fn(x[None])[0][0]))
    assert worst < tol, worst


def test_first_visit_numbering_matches_nfirst_dof():
    """DofMap(numbering='first_visit') equals the test's own restatement of
    mumps_sc's NFIRST_DOF loop, H1 and H(curl), mixed orders."""
    po = np.random.default_rng(2).integers(1, 5, size=(3 * 3 * 2, 3))
    m = mesh.structured((3, 3, 2), porder=po, perm_seed=9)
    for space in (0, 1):
        dm = DofMap(m, space, numbering="first_visit")
        ref = sr.Numbering(m, space)
        assert dm.ndof == ref.ndof
        for e in range(m.nel):
            assert np.array_equal(dm.gid[e, :dm.ni[e]], ref.gid[e]), (space, e)


def test_element_system_layout_and_triplet_assembly():
    """DofMap.element_system gives celem_systemI's layout (nodes in reverse
    element order, Dirichlet dofs dropped, loads lifted) and LCON: the
    triplets of all elements sum to the same global system as the direct
    assembly by global numbers, with the Dirichlet lift (conv_pois data)."""
    import scipy.sparse as sp

    m = unit_cube(2, 2, seed=3)
    dm = DofMap(m, 0, numbering="first_visit")
    ud = dirichlet.dirichlet_values(m, dm, oracle.LOAD_CONV)
    ref, _ = sr.element_data(m, 0, oracle.LOAD_CONV)
    a_ref, b_ref = sr.assemble(m, 0, ref, sr.Numbering(m, 0), ud)
    rows, cols, vals = [], [], []
    rhs = np.zeros(dm.ndof)
    for e in range(m.nel):
        n, nb = oracle.ndof_h1(m.norder[e]), oracle.nbubble_h1(m.norder[e][18])
        aii, bi, _, _ = sr.unpack(ref, e, n, nb)
        zastif, zbload, lcon = dm.element_system(e, aii, bi, ud[e])
        k = lcon.size
        # the reverse node order: the first entries belong to the last face
        free = np.nonzero(dm.gid[e, :n - nb] >= 0)[0]
        assert set(lcon - 1) == set(dm.gid[e, free])
        rows.append(np.repeat(lcon - 1, k))
        cols.append(np.tile(lcon - 1, k))
        vals.append(zastif)
        np.add.at(rhs, lcon - 1, zbload)
    a = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(dm.ndof, dm.ndof)).tocsr()
    assert abs(a - a_ref).max() <= 1e-14 * abs(a_ref).max()
    assert np.abs(rhs - b_ref).max() <= 1e-14 * np.abs(b_ref).max()


def test_element_system_reverse_node_order():
    """NEXTRACT walks the element's nodes from the last to the first: on an
    interior element of a p = 2 mesh the compressed order is the 6 faces
    (5..0), the 12 edges (11..0), then the 8 vertices (7..0)."""
    m = mesh.structured((3, 3, 3), porder=2)
    dm = DofMap(m, 0)
    e = 13                                    # the centre element: no Dirichlet dof
    ni = int(dm.ni[e])
    aii = np.arange(ni * ni, dtype=float).reshape(ni, ni)
    _, zb, lcon = dm.element_system(e, aii, np.arange(ni, dtype=float))
    expect = list(range(20 + 5, 19, -1)) + list(range(8 + 11, 7, -1)) + list(range(7, -1, -1))
    assert list(zb.astype(int)) == expect
    assert list(lcon - 1) == list(dm.gid[e, expect])


def test_poly_exact_solution():
    # u = x(1-x)y(1-y)z(1-z) lies in Q2 and vanishes on the boundary: exact for
    # p >= 2 on affine elements, whatever the orientations (poly_pois analogue)
    for porder, seed in ((2, 3), ([[2, 3, 2]] * 27, 4), (3, 5)):
        m = mesh.structured((3, 3, 3), porder=porder, perm_seed=seed)
        r = sr.solve(m, 0, oracle.LOAD_POLY2)
        assert np.sqrt(r["error2"]) < 1e-14 * max(1.0, np.sqrt(r["norm2"]))


@pytest.mark.parametrize("space,p,load,lo,hi", [(0, 2, oracle.LOAD_SINSINSIN, 1.95, 2.05),
                                                 (0, 3, oracle.LOAD_SINSINSIN, 2.95, 3.05),
                                                 (1, 2, oracle.LOAD_CURL_SIN, 1.95, 2.05)])
def test_convergence_rate(space, p, load, lo, hi):
    # conv_pois analogue: rate of the H1 (H(curl)) error between the last two meshes
    errs = []
    for nx in (2, 4, 8):
        m = mesh.structured((nx,) * 3, porder=p, perm_seed=nx)
        errs.append(np.sqrt(sr.solve(m, space, load)["error2"]))
    rate = np.log(errs[-2] / errs[-1]) / np.log(2.0)
    assert lo <= rate <= hi, (errs, rate)


def test_cpu_adjoint_against_finite_differences():
    m = mesh.structured((3, 3, 2), porder=2, perturb=0.1, mesh_seed=1, perm_seed=2)
    rng = np.random.default_rng(0)
    th, tt = rng.normal(0, 0.5, m.nel), rng.normal(0, 0.5, m.nel)
    J, g = sr.adjoint(m, th, tt, oracle.LOAD_SINSINSIN)
    h = 1e-6
    for e in (0, 7, 17):
        tp, tm = th.copy(), th.copy()
        tp[e] += h
        tm[e] -= h
        fd = (sr.adjoint(m, tp, tt, oracle.LOAD_SINSINSIN)[0] -
              sr.adjoint(m, tm, tt, oracle.LOAD_SINSINSIN)[0]) / (2 * h)
        assert abs(fd - g[e]) <= 1e-6 * np.abs(g).max(), (e, fd, g[e])
