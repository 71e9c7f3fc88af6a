"""Parity of the HIP element path (through the C ABI) against the oracle.

Tolerances (relative Frobenius, per element):
- FP64: K, b, Aii and Bi 1e-12.  ASchur = Abb^-1 Abi and BSchur = Abb^-1 Bb
  are solutions of the bubble system, whose forward error is bounded by the
  condition number: both the GPU and the oracle sit within a small multiple
  of nb cond(Abb) eps of the exact value, so they are held to
  max(1e-12, nb cond(Abb) eps) with cond(Abb) computed here per element.
- FP32 contraction (FP64 geometry and condensation): K, Aii and Bi within
  1e-6.  ASchur and BSchur solve the bubble system whose matrix carries the
  FP32 contraction's rounding: to first order their relative error is at
  most cond(Abb) times the relative perturbation of Abb and Abi, itself a
  small multiple of the FP32 unit roundoff u32 = 2^-24: with relative
  perturbations e_bi of Abi and e_bb of Abb, the relative error of
  Abb^-1 Abi is at most cond(Abb) (e_bi + sqrt(nb) e_bb).  They are held to
  C32 cond(Abb) u32 per element (cond computed here; C32 = 32 covers K
  perturbations of up to 4 u32, as measured, with sqrt(nb) up to 7).
"""
import numpy as np
import pytest

import oracle
from hp3d_amd import mesh

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-6
EPS = np.finfo(np.float64).eps
U32 = 2.0 ** -24
# ASchur / BSchur in the FP32 form: error <= C32 cond(Abb) u32 per element
C32 = 32.0


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def run_batch(m, space, precision, want_stc=True, herm=True):
    import torch

    import hp3d_amd

    batch = hp3d_amd.ElementBatch.from_mesh(space, m)
    out = batch.run(precision, want_stc=want_stc)
    torch.cuda.synchronize()
    batch.check_status()
    load = oracle.LOAD_SINSINSIN if space == 0 else oracle.LOAD_CURL_SIN
    ldk = int(batch.n.max())
    ref = oracle.element_loop(space, m.norder, m.nori_e, m.nori_f, m.xnod(), m.coef, load,
                              ldk=ldk, want_stc=want_stc, herm=herm)
    return batch, out, ref


def cond_bb(ref, e, n, nb):
    """2-norm condition number of element e's bubble block (the oracle's K)."""
    kr = ref["k"][e].T[:n, :n]
    return np.linalg.cond(kr[n - nb:, n - nb:])


def compare(batch, out, ref, tol_k, tol_c, want_stc=True, schur32=False, sample=None):
    """Worst per-element errors of every output against the oracle.

    tol_k bounds K, tol_c bounds Aii and Bi (the load b is always FP64 and
    held to 1e-12).  ASchur and BSchur are held to max(tol_c, nb cond(Abb)
    eps) in FP64 and, with schur32, to C32 cond(Abb) u32.  `sample` picks the
    elements (all by default; ref then holds those elements in that order)."""
    worst = {"k": 0.0, "b": 0.0, "aii": 0.0, "bi": 0.0, "aschur": 0.0, "bschur": 0.0}
    schur_ratio = 0.0   # worst error / bound of ASchur and BSchur
    c_eff = 0.0         # worst error / (cond(Abb) u), the constant the bound states
    elems = range(batch.nel) if sample is None else sample
    for i, e in enumerate(elems):
        n, nb, ni = int(batch.n[e]), int(batch.nb[e]), int(batch.ni[e])
        worst["k"] = max(worst["k"], rel(batch.k_of(out, e), ref["k"][i].T[:n, :n]))
        worst["b"] = max(worst["b"], rel(batch.vec_of(out, "b", e), ref["b"][i][:n]))
        if want_stc and nb > 0:
            worst["aii"] = max(worst["aii"], rel(batch.aii_of(out, e), ref["aii"][i].T[:ni, :ni]))
            worst["bi"] = max(worst["bi"], rel(batch.vec_of(out, "bi", e, ni), ref["bi"][i][:ni]))
            asr = ref["aschur"][i].reshape(-1)[: nb * ni].reshape(ni, nb).T
            ea = rel(batch.aschur_of(out, e), asr)
            eb = rel(batch.vec_of(out, "bschur", e, nb), ref["bschur"][i][:nb])
            worst["aschur"] = max(worst["aschur"], ea)
            worst["bschur"] = max(worst["bschur"], eb)
            kappa = cond_bb(ref, i, n, nb)
            unit = U32 if schur32 else EPS
            bound = C32 * kappa * U32 if schur32 else max(tol_c, nb * kappa * EPS)
            schur_ratio = max(schur_ratio, ea / bound, eb / bound)
            c_eff = max(c_eff, ea / (kappa * unit), eb / (kappa * unit))
    print({k: "%.2e" % v for k, v in worst.items()},
          "schur error / bound %.3f, error / (cond u) %.3f" % (schur_ratio, c_eff))
    assert worst["k"] < tol_k and worst["b"] < TOL64, worst
    if want_stc:
        for k in ("aii", "bi"):
            assert worst[k] < tol_c, (k, worst[k])
        assert schur_ratio < 1.0, schur_ratio
    return worst


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_h1_uniform_fp64(gpu, p):
    m = mesh.structured((3, 2, 2), porder=p, perturb=0.15, mesh_seed=p, perm_seed=10 + p,
                        coef=np.exp(np.random.default_rng(p).normal(size=12)))
    batch, out, ref = run_batch(m, 0, 0)
    compare(batch, out, ref, TOL64, TOL64)


def test_h1_mixed_orders_fp64(gpu):
    rng = np.random.default_rng(5)
    porder = rng.integers(1, 7, size=(4 * 3 * 3, 3))
    m = mesh.structured((4, 3, 3), porder=porder, perturb=0.1, mesh_seed=3, perm_seed=4,
                        coef=np.exp(rng.normal(size=36)))
    batch, out, ref = run_batch(m, 0, 0)
    compare(batch, out, ref, TOL64, TOL64)


def test_h1_large_p6_fp64(gpu):
    # p = 6: N = 343 > 128, several K blocks per element
    m = mesh.structured((2, 1, 1), porder=6, perturb=0.1, mesh_seed=1, perm_seed=2)
    batch, out, ref = run_batch(m, 0, 0)
    compare(batch, out, ref, TOL64, TOL64)


def test_h1_fp32(gpu):
    m = mesh.structured((3, 3, 2), porder=4, perturb=0.15, mesh_seed=9, perm_seed=9,
                        coef=np.exp(np.random.default_rng(9).normal(size=18)))
    batch, out, ref = run_batch(m, 0, 1)
    w = compare(batch, out, ref, TOL32, TOL32, schur32=True)
    print("fp32 worst", w)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_hcurl_fp64(gpu, p):
    m = mesh.structured((2, 2, 1), porder=p, perturb=0.1, mesh_seed=20 + p, perm_seed=30 + p,
                        coef=np.random.default_rng(p).uniform(1, 4, 4))
    batch, out, ref = run_batch(m, 1, 0)
    compare(batch, out, ref, TOL64, TOL64)


@pytest.mark.parametrize("precision", [0, 1])
def test_fused_matches_two_pass(gpu, precision):
    """The single-launch form (condensation on the register-resident K) and the
    two-pass form (element kernel, then the blocked condensation from HBM)
    agree.  In FP32 the fused form condenses the carried sum h + c in FP64 and
    the two-pass form the K rounded once to FP32, so they agree to FP32
    rounding; in FP64 to a few ulps."""
    import torch

    import hp3d_amd

    m = mesh.structured((3, 2, 2), porder=4, perturb=0.15, mesh_seed=2, perm_seed=5,
                        coef=np.exp(np.random.default_rng(0).normal(size=12)))
    batch = hp3d_amd.ElementBatch.from_mesh(0, m)
    o1 = batch.run(precision, fused=True)
    o2 = batch.run(precision, fused=False)
    o3 = batch.run(precision, want_k=False, fused=True)  # K not written at all
    torch.cuda.synchronize()
    batch.check_status()
    assert "k" not in o3
    t1, t2 = (1e-13, 1e-12) if precision == 0 else (1e-6, 1e-6)
    for e in range(batch.nel):
        for o in (o1, o3):
            assert rel(batch.aii_of(o, e), batch.aii_of(o2, e)) < t1
            assert rel(batch.aschur_of(o, e), batch.aschur_of(o2, e)) < t2
            nb = int(batch.nb[e])
            assert rel(batch.vec_of(o, "bschur", e, nb), batch.vec_of(o2, "bschur", e, nb)) < t2
            ni = int(batch.ni[e])
            assert rel(batch.vec_of(o, "bi", e, ni), batch.vec_of(o2, "bi", e, ni)) < t1
        assert rel(batch.k_of(o1, e), batch.k_of(o2, e)) == 0.0


@pytest.mark.parametrize("space,p", [(0, 3), (0, 5), (1, 3)])
def test_dropin_elem_and_stc(gpu, space, p):
    """The per-element entries (host arrays): elem, stc_fwd and celem, whose
    batch of one runs the fused kernel (H1 p = 3) or the element pass and the
    strip condensation (H1 p = 5, H(curl) p = 3)."""
    import hp3d_amd

    m = mesh.structured((2, 2, 2), porder=p, perturb=0.1, mesh_seed=1, perm_seed=3,
                        coef=np.exp(np.random.default_rng(p).normal(0, 0.5, 8)))
    load = oracle.LOAD_SINSINSIN if space == 0 else oracle.LOAD_CURL_SIN
    for e in range(m.nel):
        args = (m.norder[e], m.nori_e[e], m.nori_f[e])
        c = m.coef[e]
        k, b, flag = hp3d_amd.elem(space, *args, m.xnod()[e], coef=c, load=load)
        if space == 0:
            kr, br, _ = oracle.elem_h1(*args, m.xnod()[e].T, kappa=c, load=load)
            nb = oracle.nbubble_h1(m.norder[e][18])
        else:
            kr, br, _ = oracle.elem_hcurl(*args, m.xnod()[e].T, eps=c, load=load)
            nb = oracle.nbubble_hcurl(m.norder[e][18])
        assert flag == 0
        assert rel(k, kr) < TOL64 and rel(b, br) < TOL64
        aii, bi, asc, bsc = hp3d_amd.stc_fwd(kr, br, nb)
        ar, bir, ascr, bscr = oracle.stc_fwd(kr, br, nb, herm=False)
        assert rel(aii, ar) < TOL64 and rel(bi, bir) < TOL64
        assert rel(asc, ascr) < TOL64 and rel(bsc, bscr) < TOL64
        aii2, bi2, asc2, bsc2, fl2 = hp3d_amd.celem(space, *args, m.xnod()[e], coef=c, load=load)
        assert fl2 == 0
        assert rel(aii2, ar) < TOL64 and rel(bi2, bir) < TOL64
        assert rel(asc2, ascr) < TOL64 and rel(bsc2, bscr) < TOL64


def test_order_bins_mixed_mesh_and_graph_replay(gpu):
    """Mixed anisotropic orders with the minimum rule: the three order bins
    (64-row fused, 128-row fused, two-pass) against the oracle, then the captured graph replayed three
    times with eager work in between, each replay equal to the eager result."""
    import torch

    import hp3d_amd

    rng = np.random.default_rng(21)
    porder = rng.choice(np.arange(1, 7), size=(5 * 4 * 3, 3), p=mesh.P_OUT)
    porder[:6] = 6  # a few large elements
    m = mesh.structured((5, 4, 3), porder=porder, perturb=0.1, mesh_seed=8, perm_seed=9,
                        coef=np.exp(rng.normal(size=60)))
    batch = hp3d_amd.ElementBatch.from_mesh(0, m)
    assert batch.bin_paths == ["fused", "fused", "two_pass"]
    assert batch.bins[0][2] <= 64 and batch.bins[1][2] > 64
    # bi / bschur use length-N slots of which ni / nb entries are written:
    # start from zeros so the untouched tails compare equal across replays
    out = {k: v.zero_() for k, v in batch.alloc_outputs(0).items()}
    batch.run(0, out=out)
    torch.cuda.synchronize()
    batch.check_status()
    ref = oracle.element_loop(0, m.norder, m.nori_e, m.nori_f, m.xnod(), m.coef,
                              oracle.LOAD_SINSINSIN, ldk=int(batch.n.max()), herm=True)
    compare(batch, out, ref, TOL64, TOL64)
    eager = {k: v.clone() for k, v in out.items()}
    g = batch.capture(out, 0)
    for rep in range(3):
        for v in out.values():
            v.zero_()
        g.replay()
        torch.cuda.synchronize()
        for key in ("k", "b", "aii", "bi", "aschur", "bschur"):
            assert torch.equal(out[key], eager[key]), (rep, key)
        batch.run(1)  # eager work between replays (different precision, own buffers)
        torch.cuda.synchronize()


@pytest.mark.parametrize("space,p", [(0, 4), (1, 2)])
def test_bf16_trial(gpu, space, p):
    """BF16 operands, FP32 accumulation: K within bfloat16 accuracy of the
    FP64 K (and clearly not FP32-exact), condensation still well defined."""
    import torch

    import hp3d_amd

    m = mesh.structured((2, 2, 2), porder=p, perturb=0.1, mesh_seed=3, perm_seed=4)
    batch = hp3d_amd.ElementBatch.from_mesh(space, m)
    o16 = batch.run(hp3d_amd.BF16)
    o64 = batch.run(hp3d_amd.FP64)
    torch.cuda.synchronize()
    batch.check_status()
    errs = [rel(batch.k_of(o16, e), batch.k_of(o64, e)) for e in range(batch.nel)]
    aerr = [rel(batch.aii_of(o16, e), batch.aii_of(o64, e)) for e in range(batch.nel)]
    print("bf16 K error max %.2e median %.2e, Aii max %.2e" % (max(errs), np.median(errs), max(aerr)))
    assert 1e-5 < max(errs) < 2e-2
    assert max(aerr) < 5e-2


def test_bin_rules_and_autotune(gpu):
    """Every rule set the autotuner may pick computes the same element and
    condensed matrices (FP64, to rounding); autotune returns one of them."""
    import torch

    import hp3d_amd
    from hp3d_amd import tune

    rng = np.random.default_rng(8)
    porder = rng.choice(np.arange(1, 7), size=(4 * 4 * 3, 3), p=mesh.P_OUT)
    m = mesh.structured((4, 4, 3), porder=porder, perturb=0.1, mesh_seed=2, perm_seed=3,
                        coef=np.exp(rng.normal(size=48)))
    ref_batch = hp3d_amd.ElementBatch.from_mesh(0, m)
    ref = ref_batch.run(0)
    for rules in tune.candidate_rules()[::3]:
        b = hp3d_amd.ElementBatch.from_mesh(0, m, rules=rules)
        o = b.run(0)
        torch.cuda.synchronize()
        b.check_status()
        for e in range(m.nel):
            assert rel(b.k_of(o, e), ref_batch.k_of(ref, e)) < 1e-13
            assert rel(b.aii_of(o, e), ref_batch.aii_of(ref, e)) < 1e-12
            assert rel(b.aschur_of(o, e), ref_batch.aschur_of(ref, e)) < 1e-11
    best, info = tune.autotune(0, m, hp3d_amd.FP32, reps=2)
    assert any(tuple(c["rules"]) == best for c in info["candidates"])
    # the winner is the fastest of the confirmation stage, whose entries are
    # the three fastest of the first pass
    confirmed = [c for c in info["candidates"] if "ms_confirm" in c]
    first = sorted(c["ms"] for c in info["candidates"])
    assert 1 <= len(confirmed) <= 3 and all(c["ms"] <= first[len(confirmed) - 1] for c in confirmed)
    win = min(confirmed, key=lambda c: c["ms_confirm"])
    assert tuple(win["rules"]) == best and info["best_ms"] == win["ms_confirm"]


@pytest.mark.parametrize("p", [4, 5, 6])
def test_hcurl_large_fp64(gpu, p):
    """H(curl) up to MAXP (N = 882, nb = 450): block-tiled element kernel and
    the large-panel blocked condensation, against the oracle."""
    m = mesh.structured((2, 1, 1), porder=p, perturb=0.1, mesh_seed=40 + p, perm_seed=50 + p,
                        coef=np.array([1.5, 3.0]))
    batch, out, ref = run_batch(m, 1, 0)
    compare(batch, out, ref, TOL64, TOL64)


def test_hcurl_fp32_carry(gpu):
    """FP32 contraction with the TwoSum carry on a long k-sum (H(curl) p = 4,
    k = 6 x 125 = 750 columns): K and the condensed Aii and Bi within 1e-6 of
    the oracle; ASchur and BSchur within the stated H(curl) FP32 bound."""
    m = mesh.structured((2, 2, 1), porder=4, perturb=0.1, mesh_seed=7, perm_seed=8,
                        coef=np.random.default_rng(4).uniform(1, 4, 4))
    batch, out, ref = run_batch(m, 1, 1)
    w = compare(batch, out, ref, TOL32, TOL32, schur32=True)
    assert w["k"] < 5e-7


@pytest.mark.parametrize("precision", [0, 1])
def test_persistent_single_block_kernel(gpu, precision):
    """The persistent grid gives the same bits as one workgroup per element
    (64-row and 128-row bins)."""
    import torch

    import hp3d_amd

    rng = np.random.default_rng(12)
    porder = rng.choice(np.arange(1, 5), size=(6 * 5 * 4, 3), p=[0.1, 0.4, 0.3, 0.2])
    m = mesh.structured((6, 5, 4), porder=porder, perturb=0.1, mesh_seed=5, perm_seed=6,
                        coef=np.exp(rng.normal(size=120)))
    rules = ({"max_ndof": 64, "max_nbubble": 16, "path": "fused"},
             {"max_ndof": 128, "max_nbubble": 32, "path": "fused"}, {"path": "two_pass"})
    plain = hp3d_amd.ElementBatch.from_mesh(0, m, rules=rules)
    pers = hp3d_amd.ElementBatch.from_mesh(0, m, rules=tuple(
        dict(r, persistent=True) if r["path"] == "fused" else r for r in rules))
    assert pers.bin_persistent[:2] == [True, True] and plain.bins == pers.bins
    o1 = plain.run(precision)
    o2 = pers.run(precision)
    torch.cuda.synchronize()
    plain.check_status()
    pers.check_status()
    for key in ("k", "b", "aii", "aschur"):
        assert torch.equal(o1[key], o2[key]), key


@pytest.mark.parametrize("space,pmax", [(0, 3), (1, 2)])
def test_small_fused_kernel(gpu, space, pmax):
    """The 64-row single-block kernel (N <= 64, nb <= 16) against the oracle in
    FP64 and against the 128-row kernel in FP64 and FP32."""
    import torch

    import hp3d_amd

    rng = np.random.default_rng(31 + space)
    porder = rng.integers(1, pmax + 1, size=(4 * 4 * 3, 3))
    m = mesh.structured((4, 4, 3), porder=porder, perturb=0.1, mesh_seed=4, perm_seed=5,
                        coef=np.exp(rng.normal(size=48)))
    small = hp3d_amd.ElementBatch.from_mesh(space, m)
    assert small.bin_paths == ["fused"] and small.bins[0][2] <= 64 and small.bins[0][3] <= 16
    big = hp3d_amd.ElementBatch.from_mesh(space, m, rules=(
        {"max_ndof": 128, "max_nbubble": 32, "path": "fused"}, {"path": "two_pass"}))
    ref = oracle.element_loop(space, m.norder, m.nori_e, m.nori_f, m.xnod(), m.coef,
                              oracle.LOAD_SINSINSIN if space == 0 else oracle.LOAD_CURL_SIN,
                              ldk=int(small.n.max()), herm=True)
    for prec, tol in ((0, 1e-13), (1, 1e-6)):
        o1, o2 = small.run(prec), big.run(prec)
        torch.cuda.synchronize()
        small.check_status()
        big.check_status()
        if prec == 0:
            compare(small, o1, ref, TOL64, TOL64)
        for e in range(m.nel):
            ni = int(small.ni[e])
            assert rel(small.k_of(o1, e), big.k_of(o2, e)) < tol
            assert rel(small.vec_of(o1, "b", e), big.vec_of(o2, "b", e)) < 1e-13
            if int(small.nb[e]) == 0:
                continue
            assert rel(small.aii_of(o1, e), big.aii_of(o2, e)) < tol
            assert rel(small.vec_of(o1, "bi", e, ni), big.vec_of(o2, "bi", e, ni)) < tol
            assert rel(small.aschur_of(o1, e), big.aschur_of(o2, e)) < (tol if prec == 0 else 1e-3)


@pytest.mark.parametrize("porder", [2, 5])
def test_status_codes_on_device(gpu, porder):
    """Failure paths of the fused (p = 2) and two-pass (p = 5) bins, forced from
    the host data (every kernel runs to its end):
    - an inverted element warns (status 1, geom3D.F90:93-110) when only K is
      asked for; its K is NaN, as the reference's sqrt(rjac*wa) makes it
      (elem_opt.F90:101,122), so its condensation fails the SPD check;
    - kappa_e = 0 makes the bubble block singular (HP3D_E_FACTOR,
      stc.F90:360-396);
    - an order code above MAXP given straight to the C ABI reports HP3D_E_ORDER
      for that element only."""
    import ctypes

    import torch

    import hp3d_amd
    from hp3d_amd import _lib

    m = mesh.structured((2, 2, 2), porder=porder, perturb=0.0, mesh_seed=1, perm_seed=2)
    xnod = m.xnod().copy()
    xnod[3, :, 0] *= -1.0          # element 3 mirrored in x: det J < 0
    coef = np.ones(m.nel)
    coef[5] = 0.0                  # element 5: K = 0, Abb singular
    batch = hp3d_amd.ElementBatch(0, m.norder, m.nori_e, m.nori_f, xnod, coef)
    assert batch.bin_paths == (["fused"] if porder == 2 else ["two_pass"])
    ok = [e for e in range(m.nel) if e not in (3, 5)]
    ref = oracle.element_loop(0, m.norder[ok], m.nori_e[ok], m.nori_f[ok], xnod[ok], coef[ok],
                              oracle.LOAD_SINSINSIN, ldk=int(batch.n.max()), herm=True)
    # K only: the warning, and the other elements exact
    out = batch.run(hp3d_amd.FP64, want_stc=False)
    torch.cuda.synchronize()
    st = batch.status.cpu().numpy()[batch.iperm]
    assert st[3] == _lib.NEG_JACOBIAN and all(st[e] == 0 for e in range(m.nel) if e != 3), st
    assert np.isnan(batch.k_of(out, 3)).any()
    for i, e in enumerate(ok):
        n = int(batch.n[e])
        assert rel(batch.k_of(out, e), ref["k"][i].T[:n, :n]) < TOL64
    # with condensation both forced failures are fatal, the rest condensed
    out = batch.run(hp3d_amd.FP64)
    torch.cuda.synchronize()
    st = batch.status.cpu().numpy()[batch.iperm]
    assert st[3] == _lib.E_FACTOR and st[5] == _lib.E_FACTOR and all(st[e] == 0 for e in ok), st
    with pytest.raises(hp3d_amd.Hp3dError):
        batch.check_status()
    for i, e in enumerate(ok):
        ni = int(batch.ni[e])
        assert rel(batch.aii_of(out, e), ref["aii"][i].T[:ni, :ni]) < TOL64

    # an order code above MAXP, through the C ABI: only that element fails
    lib = _lib.load()
    no = np.ascontiguousarray(m.norder, dtype=np.int32).copy()
    no[2, 0] = 7
    n, nb = [np.asarray(v) for v in hp3d_amd.ndof(0, m.norder)]
    dev = torch.device("cuda")
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(dev)
    off_v = np.concatenate([[0], np.cumsum(n)[:-1]])
    off_k = np.concatenate([[0], np.cumsum(n * n)[:-1]])
    off_c = np.concatenate([[0], np.cumsum((n - nb) ** 2)[:-1]])
    off_s = np.concatenate([[0], np.cumsum(nb * (n - nb))[:-1]])
    d_no, d_oe, d_of = t(no, torch.int32), t(m.nori_e, torch.int32), t(m.nori_f, torch.int32)
    d_x, d_c = t(m.xnod(), torch.float64), t(np.ones(m.nel), torch.float64)
    d_ok, d_ov, d_oc, d_os = (t(a, torch.int64) for a in (off_k, off_v, off_c, off_s))
    k = torch.zeros(int((n * n).sum()), dtype=torch.float64, device=dev)
    b, bi, bs = (torch.zeros(int(n.sum()), dtype=torch.float64, device=dev) for _ in range(3))
    aii = torch.zeros(int(((n - nb) ** 2).sum()), dtype=torch.float64, device=dev)
    asc = torch.zeros(max(int((nb * (n - nb)).sum()), 1), dtype=torch.float64, device=dev)
    mnd, mnb = int(n.max()), int(nb.max())
    ws = torch.zeros(int(lib.hp3d_gpu_celem_workspace(0, 0, m.nel, mnd, mnb)) // 8 + 1,
                     dtype=torch.float64, device=dev)
    status = torch.full((m.nel,), 99, dtype=torch.int32, device=dev)
    p = lambda x: ctypes.c_void_p(x.data_ptr())
    rc = lib.hp3d_gpu_celem_batch(0, 0, m.nel, p(d_no), p(d_oe), p(d_of), p(d_x), m.xnod().shape[1],
                                  p(d_c), 1, p(d_ok), 0, p(k), p(d_ov), p(b), p(d_oc), p(aii), p(bi),
                                  p(d_os), p(asc), p(bs), p(ws), p(status), mnd, mnb,
                                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    s = status.cpu().numpy()
    assert s[2] == _lib.E_ORDER and all(s[e] == 0 for e in range(m.nel) if e != 2), s


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_concurrent_bins_match_sequential(gpu, precision):
    """The bins on forked streams give the same bits as one stream, in every
    precision (the BF16 trial included: its fused bins use no workspace, so
    only its two-pass bins share it, on one stream)."""
    import torch

    import hp3d_amd

    rng = np.random.default_rng(41)
    porder = rng.choice(np.arange(1, 7), size=(5 * 4 * 3, 3), p=mesh.P_OUT)
    porder[:6] = 6
    m = mesh.structured((5, 4, 3), porder=porder, perturb=0.1, mesh_seed=8, perm_seed=9,
                        coef=np.exp(rng.normal(size=60)))
    batch = hp3d_amd.ElementBatch.from_mesh(0, m)
    assert len(batch.bins) == 3
    o1 = {k: v.zero_() for k, v in batch.alloc_outputs(precision).items()}
    o2 = {k: v.zero_() for k, v in batch.alloc_outputs(precision).items()}
    batch.celem(o1, precision, concurrent=True)
    torch.cuda.synchronize()
    batch.check_status()
    batch.celem(o2, precision, concurrent=False)
    torch.cuda.synchronize()
    batch.check_status()
    for key in o1:
        assert torch.equal(o1[key], o2[key]), key


def test_per_element_entries_from_threads(gpu):
    """The per-element entries called from 8 host threads at once (an OpenMP
    element loop): each thread runs on its own context, and every result is
    bit-identical to the same call made alone.  The threaded loop's rate
    against the one-thread loop is printed."""
    import time
    from concurrent.futures import ThreadPoolExecutor

    import hp3d_amd as he

    rng = np.random.default_rng(12)
    po = rng.integers(1, 5, size=(48, 3))
    m = mesh.structured((4, 4, 3), porder=po, perturb=0.1, mesh_seed=1, perm_seed=2,
                        coef=np.exp(rng.normal(0, 0.5, 48)))
    xn = m.xnod()

    def one(e):
        if e % 2:
            return he.elem(0, m.norder[e], m.nori_e[e], m.nori_f[e], xn[e], coef=m.coef[e])
        return he.celem(0, m.norder[e], m.nori_e[e], m.nori_f[e], xn[e], coef=m.coef[e])

    alone = [one(e) for e in range(m.nel)]
    with ThreadPoolExecutor(8) as ex:
        first = list(ex.map(one, range(m.nel)))       # makes the 8 contexts
        second = list(ex.map(one, range(m.nel)))
        # rates over 10 passes of the mesh
        reps = list(range(m.nel)) * 10
        t0 = time.perf_counter()
        for e in reps:
            one(e)
        t1 = time.perf_counter()
        list(ex.map(one, reps))
        t2 = time.perf_counter()
    for e in range(m.nel):
        for a, b, c in zip(alone[e], first[e], second[e]):
            assert np.array_equal(np.asarray(a), np.asarray(b)), e
            assert np.array_equal(np.asarray(a), np.asarray(c)), e
    print("per-element loop: 1 thread %.0f el/s, 8 threads %.0f el/s"
          % (len(reps) / (t1 - t0), len(reps) / (t2 - t1)))


def test_finalize_and_reinit_with_thread_contexts(gpu):
    """hp3d_gpu_finalize frees every thread's context; after hp3d_gpu_init
    again, threads that held one get a fresh context on their next call and
    compute the same bits."""
    from concurrent.futures import ThreadPoolExecutor

    import hp3d_amd as he
    from hp3d_amd import _lib

    m = mesh.structured((2, 2, 2), porder=[[2, 3, 2]] * 8, perturb=0.1, mesh_seed=3, perm_seed=1)
    xn = m.xnod()
    one = lambda e: he.celem(0, m.norder[e], m.nori_e[e], m.nori_f[e], xn[e], coef=1.0 + e)
    lib = _lib.load()
    with ThreadPoolExecutor(4) as ex:
        before = list(ex.map(one, range(m.nel)))
        assert lib.hp3d_gpu_finalize() == 0
        assert lib.hp3d_gpu_init(0, 0) == 0
        after = list(ex.map(one, range(m.nel)))
    for b, a in zip(before, after):
        for x, y in zip(b, a):
            assert np.array_equal(np.asarray(x), np.asarray(y))
