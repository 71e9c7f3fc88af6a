"""Solution-level checks on the CPU: the global numbering (hp3d_amd.solve.DofMap)
with the oracle's element systems reproduces what the reference's own tests
assert (test/poly_pois.F90: exact polynomial solution; test/conv_pois.F90:
H1 error rate p under uniform h-refinement), and the CPU adjoint gradient
agrees with central finite differences.  These pin the FP64 CPU reference
that the GPU solution tests compare against."""
import numpy as np
import pytest

import oracle
from hp3d_amd import mesh

import solution_ref as sr


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
