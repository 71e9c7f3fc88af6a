"""The workloads' own forms on the GPU, through the C ABI, against the oracle:

- the H(curl) p = 5 precision workload in FP32 (and FP64) on a few elements;
- seeded 256-element samples of the full 32^3 p = 4 mesh and of the
  20 x 20 x 50 mixed-order mesh, FP64 and FP32, every output of each sampled
  element against the oracle's element loop;
- the adjoint gradient on a p = 3 mesh: central differences (step 1e-6) on
  50 seeded elements and the CPU FP64 reference gradient, with the FP64 and
  the FP32 forward forms;
- an element above the caller's declared max_ndof (HP3D_E_ARG);
- isoparametric geometry (ngeo = N geometry dofs, nodcor.F90:19-200 and the
  sum over all of them in geom3D.F90:77-83), staged (N <= 27) and unstaged.
"""
import ctypes

import numpy as np
import pytest

import oracle
from hp3d_amd import mesh

import solution_ref as sr
from test_gpu_elem import TOL32, TOL64, compare, rel, run_batch

pytestmark = pytest.mark.gpu


# ------------------------------------------------ H(curl) p = 5 (540 dofs)

def test_hcurl_p5_fp32_and_fp64(gpu):
    """Four perturbed H(curl) p = 5 elements with eps_e in [1, 4]: the FP32
    contraction (FP64 geometry and condensation) against the FP64 oracle, K
    within 1e-6, Aii and Bi within 1e-6, ASchur and BSchur within the stated
    H(curl) FP32 bound; the FP64 form at the FP64 tolerances."""
    m = mesh.structured((2, 2, 1), porder=5, perturb=0.1, mesh_seed=1003, perm_seed=5003,
                        coef=np.random.default_rng(2003).uniform(1.0, 4.0, 4))
    batch, out, ref = run_batch(m, 1, 1)
    assert int(batch.n.max()) == 540 and int(batch.nb.max()) == 240
    w = compare(batch, out, ref, TOL32, TOL32, schur32=True)
    print("H(curl) p=5 FP32", w)
    batch, out, ref = run_batch(m, 1, 0)
    compare(batch, out, ref, TOL64, TOL64)


# ------------------------------------------------- config-mesh samples

@pytest.mark.parametrize("name", ["cfg1", "cfg2"])
def test_config_mesh_sample(gpu, name):
    """The whole workload mesh through the batched path (order bins, fused and
    two-pass forms), then 256 seeded elements against the oracle: FP64 at the
    FP64 tolerances, FP32 with K, Aii and Bi within 1e-6."""
    import torch

    import hp3d_amd

    m = mesh.config_mesh(name)
    batch = hp3d_amd.ElementBatch.from_mesh(0, m)
    sample = np.sort(np.random.default_rng(7001).choice(m.nel, 256, replace=False))
    ldk = int(batch.n[sample].max())
    ref = oracle.element_loop(0, m.norder[sample], m.nori_e[sample], m.nori_f[sample],
                              m.xnod()[sample], m.coef[sample], oracle.LOAD_SINSINSIN, ldk=ldk,
                              herm=True)
    for prec in (hp3d_amd.FP64, hp3d_amd.FP32):
        out = batch.run(prec)
        torch.cuda.synchronize()
        batch.check_status()
        if prec == hp3d_amd.FP64:
            compare(batch, out, ref, TOL64, TOL64, sample=sample)
        else:
            w = compare(batch, out, ref, TOL32, TOL32, sample=sample, schur32=True)
            print(name, "FP32", w)
        del out
        torch.cuda.empty_cache()


# ------------------------------------------------------- adjoint, p = 3

@pytest.fixture(scope="module")
def adjoint_case():
    """A 5 x 4 x 4 mesh at p = 3 with +-0.1h perturbation, theta ~ N(0, 0.5^2)
    and a second seeded theta for the target, central differences (step 1e-6)
    of the GPU FP64 forward on 50 seeded elements, and the CPU FP64 gradient."""
    from hp3d_amd import adjoint

    m = mesh.structured((5, 4, 4), porder=3, perturb=0.1, mesh_seed=1004, perm_seed=5004)
    th = np.random.default_rng(2004).normal(0.0, 0.5, m.nel)
    tt = np.random.default_rng(3004).normal(0.0, 0.5, m.nel)
    picks = np.random.default_rng(4004).choice(m.nel, 50, replace=False)
    p64 = adjoint.AdjointProblem(m, tt, precision=0)
    h = 1e-6
    fd = {}
    for e in picks:
        tp, tm = th.copy(), th.copy()
        tp[e] += h
        tm[e] -= h
        fd[int(e)] = (p64.objective(tp)[0] - p64.objective(tm)[0]) / (2 * h)
    Jc, gc = sr.adjoint(m, th, tt, oracle.LOAD_SINSINSIN)
    return m, th, tt, fd, Jc, gc


@pytest.mark.parametrize("precision", [0, 1])
def test_adjoint_p3_fd50(gpu, adjoint_case, precision):
    """dJ/dtheta through torch.autograd on the GPU against the CPU FP64
    gradient (1e-10 FP64, 1e-5 FP32) and against central differences of the
    FP64 forward on 50 elements (1e-6 FP64, 1e-5 FP32, relative to max |g|)."""
    import torch

    from hp3d_amd import adjoint

    m, th, tt, fd, Jc, gc = adjoint_case
    prob = adjoint.AdjointProblem(m, tt, precision=precision)
    theta = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    J = adjoint.objective(theta, prob)
    J.backward()
    g = theta.grad.numpy()
    tol = 1e-10 if precision == 0 else 1e-5
    gmax = np.abs(gc).max()
    print("J %.3e vs %.3e, grad vs CPU %.2e" % (float(J.detach()), Jc, np.abs(g - gc).max() / gmax))
    assert abs(float(J.detach()) - Jc) < tol * Jc
    assert np.max(np.abs(g - gc)) < tol * gmax
    fd_tol = 1e-6 if precision == 0 else 1e-5
    worst = max(abs(v - g[e]) for e, v in fd.items()) / np.abs(g).max()
    print("FD (50 elements) worst %.2e" % worst)
    assert worst <= fd_tol


def test_adjoint_two_forwards_then_backward(gpu):
    """Two objectives before a backward: each gradient is its own (the state
    restores the shared element buffers a later forward overwrote)."""
    import torch

    from hp3d_amd import adjoint

    m = mesh.structured((3, 3, 2), porder=2, perturb=0.1, mesh_seed=1, perm_seed=2)
    rng = np.random.default_rng(5)
    t1, t2, tt = (rng.normal(0, 0.5, m.nel) for _ in range(3))
    prob = adjoint.AdjointProblem(m, tt)
    a = torch.tensor(t1, requires_grad=True)
    b = torch.tensor(t2, requires_grad=True)
    Ja = adjoint.objective(a, prob)
    Jb = adjoint.objective(b, prob)
    Ja.backward()
    Jb.backward()
    for th, got in ((t1, a.grad.numpy()), (t2, b.grad.numpy())):
        _, gref = sr.adjoint(m, th, tt, oracle.LOAD_SINSINSIN)
        assert np.max(np.abs(got - gref)) < 1e-10 * np.abs(gref).max()


# ------------------------------------------- declared bounds (E_ARG)

def test_element_above_declared_max_ndof(gpu):
    """celem_batch over p = 2 elements and two p = 4 elements with max_ndof
    declared as 64: the p = 4 elements (N = 71 and 83 under the minimum rule)
    report HP3D_E_ARG, and the p = 2 elements are condensed exactly.  The
    two-pass form (two p = 6 elements, N = 203 and 235, declared max_ndof 200)
    reports HP3D_E_ARG the same way."""
    import torch

    import hp3d_amd
    from hp3d_amd import _lib

    lib = _lib.load()
    dev = torch.device("cuda")
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(dev)
    p = lambda x: ctypes.c_void_p(x.data_ptr())
    for porder, big, declared in ((2, 4, 64), (2, 6, 200)):
        po = np.full((3 * 2 * 2, 3), porder)
        po[[1, 6]] = big
        m = mesh.structured((3, 2, 2), porder=po, perturb=0.1, mesh_seed=3, perm_seed=4)
        n, nb = [np.asarray(v) for v in hp3d_amd.ndof(0, m.norder)]
        small = np.nonzero(n <= declared)[0]
        off = lambda c: np.concatenate([[0], np.cumsum(c)[:-1]])
        off_k, off_v, off_c, off_s = off(n * n), off(n), off((n - nb) ** 2), off(nb * (n - nb))
        k = torch.zeros(int((n * n).sum()), dtype=torch.float64, device=dev)
        b, bi, bs = (torch.zeros(int(n.sum()), dtype=torch.float64, device=dev) for _ in range(3))
        aii = torch.zeros(int(((n - nb) ** 2).sum()), dtype=torch.float64, device=dev)
        asc = torch.zeros(max(int((nb * (n - nb)).sum()), 1), dtype=torch.float64, device=dev)
        mnb = int(nb[small].max())
        ws = torch.zeros(int(lib.hp3d_gpu_celem_workspace(0, 0, m.nel, declared, mnb)) // 8 + 1,
                         dtype=torch.float64, device=dev)
        status = torch.full((m.nel,), 99, dtype=torch.int32, device=dev)
        d = [t(a, torch.int32) for a in (m.norder, m.nori_e, m.nori_f)]
        d_x, d_c = t(m.xnod(), torch.float64), t(np.ones(m.nel), torch.float64)
        o = [t(a, torch.int64) for a in (off_k, off_v, off_c, off_s)]
        print("case p=%d/%d max_ndof=%d" % (porder, big, declared), flush=True)
        rc = lib.hp3d_gpu_celem_batch(0, 0, m.nel, p(d[0]), p(d[1]), p(d[2]), p(d_x), 8, p(d_c), 1,
                                      p(o[0]), 0, p(k), p(o[1]), p(b), p(o[2]), p(aii), p(bi),
                                      p(o[3]), p(asc), p(bs), p(ws), p(status), declared, mnb,
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert rc == 0
        torch.cuda.synchronize()
        s = status.cpu().numpy()
        assert all(s[e] == _lib.E_ARG for e in (1, 6)), s
        assert all(s[e] == 0 for e in small), s
        ref = oracle.element_loop(0, m.norder[small], m.nori_e[small], m.nori_f[small],
                                  m.xnod()[small], np.ones(small.size), oracle.LOAD_SINSINSIN,
                                  ldk=int(n[small].max()), herm=True)
        aii_h = aii.cpu().numpy()
        for i, e in enumerate(small):
            ni = int(n[e] - nb[e])
            got = aii_h[off_c[e]:off_c[e] + ni * ni].reshape(ni, ni).T
            assert rel(got, ref["aii"][i].T[:ni, :ni]) < TOL64


# ------------------------------------------------ isoparametric geometry

def iso_mesh(porder, space, seed):
    """A 2 x 2 x 1 mesh whose geometry is isoparametric: xnod holds all N_H1
    geometry dofs, the vertex dofs at the (perturbed) vertices and seeded
    edge, face and bubble dofs of 2% of h, i.e. curved elements."""
    m = mesh.structured((2, 2, 1), porder=porder, perturb=0.1, mesh_seed=seed, perm_seed=seed + 1,
                        coef=np.random.default_rng(seed).uniform(1.0, 3.0, 4))
    nh = oracle.ndof_h1(m.norder[0])
    x = np.zeros((m.nel, nh, 3))
    x[:, :8] = m.xnod()
    x[:, 8:] = np.random.default_rng(seed + 2).uniform(-0.01, 0.01, (m.nel, nh - 8, 3))
    return m, x


@pytest.mark.parametrize("space,porder", [(0, 2), (0, 3), (1, 2)])
def test_isoparametric_geometry(gpu, space, porder):
    """ngeo = N geometry dofs (27 for p = 2: staged in LDS; 64 for p = 3:
    read from HBM) in the fused element + condensation kernel, FP64, against
    the oracle with the same geometry dofs."""
    import torch

    import hp3d_amd

    m, x = iso_mesh(porder, space, 11 + porder)
    load = oracle.LOAD_SINSINSIN if space == 0 else oracle.LOAD_CURL_SIN
    batch = hp3d_amd.ElementBatch(space, m.norder, m.nori_e, m.nori_f, x, m.coef, load=load)
    out = batch.run(hp3d_amd.FP64)
    torch.cuda.synchronize()
    assert (batch.check_status() == 0).all()
    ref = oracle.element_loop(space, m.norder, m.nori_e, m.nori_f, x, m.coef, load,
                              ldk=int(batch.n.max()), herm=True)
    compare(batch, out, ref, TOL64, TOL64)
    # the curved geometry matters: the trilinear map gives a different K
    ref8 = oracle.element_loop(space, m.norder, m.nori_e, m.nori_f, m.xnod(), m.coef, load,
                               ldk=int(batch.n.max()), herm=True)
    n = int(batch.n[0])
    assert rel(ref8["k"][0].T[:n, :n], ref["k"][0].T[:n, :n]) > 1e-4
