"""The Fortran ISO_C_BINDING binding (hp3d_amd/fortran/hp3d_gpu_mod.F90) and a
Fortran host program calling it the way an hp3D elem routine would:
ALOC with the caller's leading dimension, codes from find_order/find_orient."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from hp3d_amd import build as hb
from hp3d_amd import mesh

pytestmark = pytest.mark.skipif(not os.path.exists(hb.FDRIVER),
                                reason="Fortran driver not built (amdflang absent)")


def driver_input(mode, space, ld, no, ne, nf, xnod, coef, load):
    lines = ["%d" % mode, "%d %d" % (space, ld),
             " ".join(str(int(v)) for v in list(no) + list(ne) + list(nf)),
             "%d" % xnod.shape[1],
             " ".join("%.17e" % v for v in np.asarray(xnod, float).T.reshape(-1)),
             "%.17e %d" % (coef, load)]
    return "\n".join(lines) + "\n"


def run_driver(text):
    return subprocess.run([hb.FDRIVER], input=text, capture_output=True, text=True, timeout=120)


def random_codes(rng):
    no = np.zeros(19, np.int32)
    no[:12] = rng.integers(1, 7, 12)
    no[12:18] = rng.integers(1, 7, 6) * 10 + rng.integers(1, 7, 6)
    p = rng.integers(1, 7, 3)
    no[18] = p[0] * 100 + p[1] * 10 + p[2]
    return no, rng.integers(0, 2, 12), rng.integers(0, 8, 6)


def test_fortran_counts_and_error_stop():
    rng = np.random.default_rng(4)
    xn = np.zeros((3, 8))
    for _ in range(10):
        no, ne, nf = random_codes(rng)
        for space, fn, fb in ((0, oracle.ndof_h1, oracle.nbubble_h1),
                              (1, oracle.ndof_hcurl, oracle.nbubble_hcurl)):
            r = run_driver(driver_input(1, space, 1, no, ne, nf, xn, 1.0, 1))
            assert r.returncode == 0, r.stderr
            n, nb = map(int, r.stdout.split()[1:3])
            assert n == fn(no) and nb == fb(no[18])
    no[0] = 7  # above MAXP: the binding stops like checkorder
    r = run_driver(driver_input(1, 0, 1, no, ne, nf, xn, 1.0, 1))
    assert r.returncode != 0 and "failed (-1)" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("space", [0, 1])
def test_fortran_elem_and_celem_on_gpu(gpu, space):
    m = mesh.structured((2, 2, 2), porder=[[3, 2, 4]] * 8 if space == 0 else 2, perturb=0.1,
                        mesh_seed=5, perm_seed=6)
    e = 5
    no, ne, nf = m.norder[e], m.nori_e[e], m.nori_f[e]
    xn = m.xnod()[e].T
    coef, load = 1.7, (oracle.LOAD_SINSINSIN if space == 0 else oracle.LOAD_CURL_SIN)
    if space == 0:
        kr, br, _ = oracle.elem_h1(no, ne, nf, xn, kappa=coef, load=load)
        nb = oracle.nbubble_h1(no[18])
    else:
        kr, br, _ = oracle.elem_hcurl(no, ne, nf, xn, eps=coef, load=load)
        nb = oracle.nbubble_hcurl(no[18])
    n = kr.shape[0]
    ld = n + 17  # MAXDOFS of a mixed-order mesh: larger than this element's N
    r = run_driver(driver_input(2, space, ld, no, ne, nf, xn, coef, load))
    assert r.returncode == 0, r.stdout + r.stderr
    tok = r.stdout.split()
    assert tok[:3] == ["counts", str(n), str(nb)] and tok[3] == "rc"
    vals = np.array(tok[5:], float)
    k = vals[:n * n].reshape(n, n).T
    b = vals[n * n:]
    assert np.linalg.norm(k - kr) <= 1e-12 * np.linalg.norm(kr)
    assert np.linalg.norm(b - br) <= 1e-12 * np.linalg.norm(br)
    r = run_driver(driver_input(3, space, ld, no, ne, nf, xn, coef, load))
    assert r.returncode == 0, r.stdout + r.stderr
    vals = np.array(r.stdout.split()[5:], float)
    ni = n - nb
    aii_r, bi_r, as_r, bs_r = oracle.stc_fwd(kr, br, nb, herm=True)
    o = 0
    parts = []
    for size in (ni * ni, ni, nb * ni, nb):
        parts.append(vals[o:o + size])
        o += size
    assert o == vals.size
    aii = parts[0].reshape(ni, ni).T
    asc = parts[2].reshape(ni, nb).T
    for got, ref in ((aii, aii_r), (parts[1], bi_r), (asc, as_r), (parts[3], bs_r)):
        assert np.linalg.norm(got - ref) <= 1e-11 * np.linalg.norm(ref)


def batch_input(mode, space, m, load):
    xn = m.xnod()
    lines = ["%d" % mode, "%d 0 %d 8 %d" % (space, m.nel, load)]
    for e in range(m.nel):
        lines.append(" ".join(str(int(v)) for v in list(m.norder[e]) + list(m.nori_e[e])
                              + list(m.nori_f[e])))
        lines.append(" ".join("%.17e" % v for v in xn[e].reshape(-1)))
        lines.append("%.17e" % m.coef[e])
    return "\n".join(lines) + "\n"


@pytest.mark.gpu
@pytest.mark.parametrize("space", [0, 1])
@pytest.mark.parametrize("mode", [4, 5])
def test_fortran_batched_celem_on_gpu(gpu, space, mode):
    """A Fortran element loop over a batch: codes, geometry and offsets packed
    on the host, staged with hipMalloc/hipMemcpy, condensed by
    hp3d_gpu_celem_batch (mode 4) or hp3d_gpu_celem_batch_noalloc (mode 5);
    mixed orders up to p = 5 in H1 and p = 3 in H(curl), so that both the
    fused and the two-pass condensation serve the batch."""
    rng = np.random.default_rng(20 + space)
    po = rng.integers(1, 6 if space == 0 else 4, size=(4, 3))
    po[0] = 5 if space == 0 else 3
    m = mesh.structured((2, 2, 1), porder=po, perturb=0.1, mesh_seed=3, perm_seed=4,
                        coef=np.exp(rng.normal(0, 0.5, 4)))
    load = oracle.LOAD_SINSINSIN if space == 0 else oracle.LOAD_CURL_SIN
    r = run_driver(batch_input(mode, space, m, load))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    tok = r.stdout.split()
    assert tok[:2] == ["batch", str(m.nel)]
    assert [int(t) for t in tok[2:2 + m.nel]] == [0] * m.nel
    vals = np.array(tok[2 + m.nel:], float)
    cnt = oracle.ndof_h1 if space == 0 else oracle.ndof_hcurl
    bub = oracle.nbubble_h1 if space == 0 else oracle.nbubble_hcurl
    ns = [int(cnt(m.norder[e])) for e in range(m.nel)]
    nbs = [int(bub(m.norder[e][18])) for e in range(m.nel)]
    sc = sum((n - nb) ** 2 for n, nb in zip(ns, nbs))
    sv = sum(ns)
    ss = sum(nb * (n - nb) for n, nb in zip(ns, nbs))
    assert vals.size == sc + 2 * sv + ss
    aii_all, bi_all = vals[:sc], vals[sc:sc + sv]
    bs_all, as_all = vals[sc + sv:sc + 2 * sv], vals[sc + 2 * sv:]
    xn = m.xnod()
    oc = ov = os_ = 0
    for e in range(m.nel):
        n, nb = ns[e], nbs[e]
        ni = n - nb
        no, ne, nf = m.norder[e], m.nori_e[e], m.nori_f[e]
        if space == 0:
            k, b, _ = oracle.elem_h1(no, ne, nf, xn[e].T, kappa=m.coef[e], load=load)
        else:
            k, b, _ = oracle.elem_hcurl(no, ne, nf, xn[e].T, eps=m.coef[e], load=load)
        aii_r, bi_r, as_r, bs_r = oracle.stc_fwd(k, b, nb, herm=True)
        aii = aii_all[oc:oc + ni * ni].reshape(ni, ni).T
        asc = as_all[os_:os_ + nb * ni].reshape(ni, nb).T
        got = ((aii, aii_r), (bi_all[ov:ov + ni], bi_r), (asc, as_r), (bs_all[ov:ov + nb], bs_r))
        for g, ref in got:
            if ref.size:
                assert np.linalg.norm(g - ref) <= 1e-11 * np.linalg.norm(ref), (e, n, nb)
        oc += ni * ni
        ov += n
        os_ += nb * ni
