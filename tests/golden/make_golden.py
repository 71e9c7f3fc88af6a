"""Writes tests/golden/shape_ref.npz from the reference's own routines.

Inputs are seeded random points, anisotropic orders and all orientation
codes; outputs are what the reference Fortran (shape3DHHexa, shape3DEHexa,
PolyLegendre, PolyILegendre, the Gauss tables after init_gauss_quadrature,
geom) returns for them, through oracle/_ref (built by oracle/Makefile from the
reference sources).  Run from the repository root:

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def cases(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(n):
        no = np.zeros(19, np.int32)
        no[:12] = rng.integers(1, 7, 12)
        no[12:18] = rng.integers(1, 7, 6) * 10 + rng.integers(1, 7, 6)
        p = rng.integers(1, 7, 3)
        no[18] = p[0] * 100 + p[1] * 10 + p[2]
        ne = rng.integers(0, 2, 12).astype(np.int32)
        nf = (np.arange(6) + t) % 8
        xi = rng.random(3)
        out.append((xi, no, ne, nf.astype(np.int32)))
    return out


def main():
    if oracle.ref() is None:
        raise SystemExit("oracle/_ref is not built (needs the reference tree)")
    data = {}
    cs = cases(48, 20261202)
    data["xi"] = np.array([c[0] for c in cs])
    data["norder"] = np.array([c[1] for c in cs])
    data["nori_e"] = np.array([c[2] for c in cs])
    data["nori_f"] = np.array([c[3] for c in cs])
    h_s, h_g, e_s, e_c = [], [], [], []
    for xi, no, ne, nf in cs:
        s, g = oracle.ref_shape3dh(xi, no, ne, nf)
        h_s.append(np.pad(s, (0, oracle.MAXBRICKH - s.size)))
        h_g.append(np.pad(g, ((0, oracle.MAXBRICKH - g.shape[0]), (0, 0))))
        s, c = oracle.ref_shape3de(xi, no, ne, nf)
        e_s.append(np.pad(s, ((0, oracle.MAXBRICKE - s.shape[0]), (0, 0))))
        e_c.append(np.pad(c, ((0, oracle.MAXBRICKE - c.shape[0]), (0, 0))))
    data["h1_shap"], data["h1_grad"] = np.array(h_s), np.array(h_g)
    data["hc_shap"], data["hc_curl"] = np.array(e_s), np.array(e_c)
    xs = np.linspace(0.0, 1.0, 7)
    data["leg_x"] = xs
    data["leg_p"] = np.array([oracle.ref_poly_legendre(x, 0.75 + x, 6) for x in xs])
    data["ileg_l"] = np.array([oracle.ref_poly_ilegendre(x, 6)[0] for x in xs])
    gx, gw = np.zeros((10, 10)), np.zeros((10, 10))
    for n in range(1, 11):
        x, w = oracle.ref_gauss1d(n)
        gx[n - 1, :n], gw[n - 1, :n] = x, w
    data["gauss_x"], data["gauss_w"] = gx, gw
    rng = np.random.default_rng(7)
    jac = rng.random((16, 3, 3)) + np.eye(3)
    jac[3] = -jac[3]  # a negative Jacobian
    data["geom_in"] = jac
    res = [oracle.ref_geom(a) for a in jac]
    data["geom_inv"] = np.array([r[0] for r in res])
    data["geom_det"] = np.array([r[1] for r in res])
    data["geom_flag"] = np.array([r[2] for r in res])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "shape_ref.npz")
    np.savez_compressed(out, **data)
    print("wrote", out)


if __name__ == "__main__":
    main()
