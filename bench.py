"""Throughput of the hp3D element-computation hot path on one MI355X.

One step = element stiffness + load of every element of the workload's mesh
(elem_kernel) followed by static condensation of all bubble dofs
(stc_kernel), inputs resident in HBM, outputs left in HBM.  The default
workload is the 32x32x32 perturbed hexahedral mesh at uniform p = 4 (32,768
elements, 125 dofs and 125 Gauss points each, lognormal kappa_e).

value      element matrices/s of the batched FP32-contraction form (FP64
           geometry and bubble factorisation), whole job
forms      the same for the batched FP64 form and for the plain FP64 drop-in
           (one C-ABI call per element from the host, as the reference's
           OpenMP loop would make it)
roofline   dominant kernel (elem_kernel): algorithmic flops k N (N+1) per
           element over its average launch time (HIP events on its stream)
cpu_baseline  the oracle's OpenMP element loop (elem + LU condensation) on a
           bounded sample of the same mesh, on this host's cores

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cfg1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("element matrices/s + condensed DOF/s: plain FP64 vs batched FP32 at "
          "≤1e-6 rel err")
PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "hbm": 8.0e12}

WORKLOADS = {
    "cfg1": "Poisson H1, 32x32x32 hexahedra on the unit cube, uniform p=4 (125 dofs, 125 Gauss "
            "points), interior vertices perturbed +-0.15h, lognormal kappa_e",
    "cfg2": "Poisson H1, 20x20x50 hexahedra in 1x1x2.5, anisotropic (px,py,pz) in 1..6 with the "
            "minimum rule, +-0.1h perturbation, lognormal kappa_e",
    "cfg3": "H(curl) curl-curl + mass, 16x16x16 hexahedra, uniform p=5 (540 dofs, 216 Gauss "
            "points), +-0.1h perturbation, eps_e ~ U[1,4]",
    "cfg4": "Poisson H1, 24x24x24 hexahedra, uniform p=3, +-0.1h perturbation, kappa_e = exp(theta)",
}


def algorithmic_flops(batch):
    """k N (N+1) per element (SYRK count, k = 3Q for H1, 2 x 3Q for H(curl))."""
    from hp3d_amd.elem import nquad

    q = np.prod(nquad(batch.norder_h, batch.nori_f_h), axis=1)
    k = 3 * q * (1 if batch.space == 0 else 2)
    n = batch._n
    fk = (k * n * (n + 1)).astype(np.float64)
    nb, ni = batch._nb.astype(np.float64), batch._ni.astype(np.float64)
    fc = nb ** 3 / 3 + 2 * nb ** 2 * (ni + 1) + 2 * ni ** 2 * nb + 2 * ni * nb
    return float(fk.sum()), float(fc.sum())


def rel_err_per_element(batch, a, b, key_sizes):
    """Relative Frobenius error of packed per-element blocks a vs b (device)."""
    import torch

    lengths = torch.as_tensor(np.asarray(key_sizes), device=a.device)
    n = int(lengths.sum())
    a, b = a[:n].double(), b[:n].double()
    num = torch.segment_reduce((a - b) ** 2, "sum", lengths=lengths)
    den = torch.segment_reduce(b * b, "sum", lengths=lengths)
    r = torch.sqrt(num / torch.clamp(den, min=1e-300)).cpu().numpy()
    return float(r.max()), float(np.median(r))


def time_steps(batch, out, precision, steps, warmup):
    """Wall time per step and the mean duration of the element launch (HIP
    events recorded on the stream the kernels run on)."""
    import torch

    for _ in range(warmup):
        batch.celem(out, precision)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    t0 = time.perf_counter()
    for s in range(steps):
        ev[s][0].record()
        batch.celem(out, precision)
        ev[s][1].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elem_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    return (t1 - t0) / steps, elem_ms


def plain_dropin_rate(m, nsample):
    """Per-element C-ABI calls (host arrays in and out), as the reference loop would make."""
    import hp3d_amd

    xn = m.xnod()
    t0 = time.perf_counter()
    for e in range(nsample):
        hp3d_amd.celem(0, m.norder[e], m.nori_e[e], m.nori_f[e], xn[e], coef=m.coef[e])
    return nsample / (time.perf_counter() - t0)


def cpu_baseline(m, batch, out32, out64, seconds=15.0):
    """Oracle element loop (OpenMP, dynamic schedule) on the first elements of the mesh."""
    import oracle

    threads = min(os.cpu_count() or 1, 16)
    space = batch.space
    load = oracle.LOAD_SINSINSIN if space == 0 else oracle.LOAD_CURL_SIN
    ldk = int(batch.max_ndof)
    # calibrate the sample so the timed run is ~seconds of work
    probe = min(m.nel, 4 * threads)
    t0 = time.perf_counter()
    oracle.element_loop(space, m.norder[:probe], m.nori_e[:probe], m.nori_f[:probe],
                        m.xnod()[:probe], m.coef[:probe], load, ldk=ldk, nthreads=threads)
    dt = max(time.perf_counter() - t0, 1e-3)
    ns = int(min(m.nel, max(probe, probe * seconds / dt)))
    t0 = time.perf_counter()
    ref = oracle.element_loop(space, m.norder[:ns], m.nori_e[:ns], m.nori_f[:ns],
                              m.xnod()[:ns], m.coef[:ns], load, ldk=ldk, nthreads=threads)
    dt = time.perf_counter() - t0
    # FP32 and FP64 GPU outputs against the oracle on the sample
    err = {"k32": [], "aii32": [], "k64": [], "aii64": []}
    for e in range(0, ns, max(1, ns // 200)):
        n, ni = int(batch.n[e]), int(batch.ni[e])
        kr = ref["k"][e].T[:n, :n]
        ar = ref["aii"][e].T[:ni, :ni]
        for tag, o in (("32", out32), ("64", out64)):
            err["k" + tag].append(np.linalg.norm(batch.k_of(o, e) - kr) / np.linalg.norm(kr))
            err["aii" + tag].append(np.linalg.norm(batch.aii_of(o, e) - ar) / np.linalg.norm(ar))
    # the same loop on one core, on a smaller sample
    n1 = int(min(ns, max(1, ns / threads / 3)))
    t0 = time.perf_counter()
    oracle.element_loop(space, m.norder[:n1], m.nori_e[:n1], m.nori_f[:n1], m.xnod()[:n1],
                        m.coef[:n1], load, ldk=ldk, nthreads=1)
    dt1 = time.perf_counter() - t0
    res = {"value": ns / dt, "unit": "element matrices/s", "cores": threads, "kind": "port",
           "sample": "first %d of %d elements, elem_opt + stc_fwd_gen (LU) per element, "
                     "OpenMP dynamic schedule, %.1f s" % (ns, m.nel, dt),
           "single_core": {"value": n1 / dt1, "sample": "first %d elements, %.1f s" % (n1, dt1)}}
    errors = {k: {"max": float(np.max(v)), "median": float(np.median(v))} for k, v in err.items()}
    return res, errors


def adjoint_leg(m, nfd):
    """dJ/dtheta_e for all elements of the 24^3 p = 3 mesh (J = 1/2 ||u - u_t||^2,
    u_t the solution for a second seeded theta): forward (element kernels with
    condensation, device assembly, PCG, back substitution, L2 field kernel) and
    gradient (adjoint condensation and solve, field kernel) timed per form; the
    gradient kernel checked against -lambda_e^T K_e u_e with the oracle's K_e on
    a sample of elements, and against central differences (step 1e-6) of the
    FP64 forward on nfd seeded elements."""
    import torch

    import hp3d_amd
    import oracle
    from hp3d_amd import adjoint
    from hp3d_amd.solve import field

    theta = np.log(m.coef)
    theta_t = np.random.default_rng(3004).normal(0.0, 0.5, m.nel)
    res = {}
    g64 = None
    for prec, name in ((hp3d_amd.FP64, "fp64"), (hp3d_amd.FP32, "fp32")):
        prob = adjoint.AdjointProblem(m, theta_t, precision=prec, solver="pcg", rtol=1e-13)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J, st = prob.objective(theta)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        its_fwd = st["solver"].iterations
        g, lam = prob.gradient(st)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        # the gradient kernel alone (HIP events on its stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            field(prob.batch, hp3d_amd.FIELD_ADJOINT, st["u"], v=lam)
        e1.record()
        torch.cuda.synchronize()
        kms = e0.elapsed_time(e1) / 5
        b = prob.batch
        u, lv = st["u"].cpu().numpy(), lam.cpu().numpy()
        xn = m.xnod()
        sample = np.random.default_rng(4004).choice(m.nel, 256, replace=False)
        err = 0.0
        for e in sample:
            i = int(b.iperm[e])
            o, n = int(b.off_v_h[i]), int(b._n[i])
            k, _, _ = oracle.elem_h1(m.norder[e], m.nori_e[e], m.nori_f[e], xn[e].T,
                                     kappa=float(np.exp(theta[e])))
            err = max(err, abs(g[e] + lv[o:o + n] @ k @ u[o:o + n]))
        entry = {"J": J, "forward_s": t1 - t0, "gradient_s": t2 - t1,
                 "pcg_iterations": [its_fwd, st["solver"].iterations],
                 "gradient_kernel_ms": kms, "gradient_kernel_elements_per_s": m.nel / (kms * 1e-3),
                 "grad_vs_oracle_K_max_rel": err / np.abs(g).max(),
                 "grad_sample": "256 seeded elements"}
        if prec == hp3d_amd.FP64:
            g64 = g
            fd_err = []
            x0 = st["x"]
            prob.rtol = 1e-14   # J(theta +- h) must resolve differences of ~h * g
            for e in np.random.default_rng(4004).choice(m.nel, nfd, replace=False):
                tp, tm = theta.copy(), theta.copy()
                tp[e] += 1e-6
                tm[e] -= 1e-6
                fd = (prob.objective(tp, x0=x0)[0] - prob.objective(tm, x0=x0)[0]) / 2e-6
                fd_err.append(abs(fd - g[e]) / np.abs(g).max())
            entry["fd_elements"] = nfd
            entry["fd_max_rel"] = max(fd_err) if fd_err else None
        else:
            entry["grad_vs_fp64_max_rel"] = float(np.abs(g - g64).max() / np.abs(g64).max())
        res[name] = entry
        del prob
        torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cfg1", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--dropin-sample", type=int, default=256)
    ap.add_argument("--kernel-only", action="store_true",
                    help="time only the FP32 element+condensation launches (for counter runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run one device-to-device copy of MIB MiB and exit (counter scale)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (default bin rules)")
    ap.add_argument("--fd", type=int, default=8,
                    help="cfg4: elements checked by central differences (step 1e-6)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the element path runs on a single GPU and has no "
              "cross-GPU exchange (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}))
        return

    import hp3d_amd
    from hp3d_amd import mesh

    hp3d_amd.init()
    m = mesh.config_mesh(args.workload)
    space = hp3d_amd.HCURL if args.workload == "cfg3" else hp3d_amd.H1
    rules, tuning = None, None
    if not args.no_tune and not args.kernel_only:
        from hp3d_amd import tune

        rules, tuning = tune.autotune(space, m, hp3d_amd.FP32, reps=3)
    batch = hp3d_amd.ElementBatch.from_mesh(space, m, rules=rules)
    out32 = batch.alloc_outputs(hp3d_amd.FP32)
    if args.kernel_only:
        s32, e32 = time_steps(batch, out32, hp3d_amd.FP32, args.steps, args.warmup)
        batch.check_status()
        print(json.dumps({"workload": args.workload, "ms_per_step": s32 * 1e3,
                          "kernel_ms": e32, "bins": [list(b) for b in batch.bins]}))
        return
    out64 = batch.alloc_outputs(hp3d_amd.FP64)

    # batched FP64 (reference form on the GPU) and FP32 (headline form)
    s64, e64 = time_steps(batch, out64, hp3d_amd.FP64, max(2, args.steps // 2), 1)
    batch.check_status()
    s32, e32 = time_steps(batch, out32, hp3d_amd.FP32, args.steps, args.warmup)
    batch.check_status()

    # the same FP32 step as one replay of a captured graph of the per-bin launches
    g = batch.capture(out32, hp3d_amd.FP32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.replay()
    torch.cuda.synchronize()
    sg = (time.perf_counter() - t0) / args.steps

    # end to end: the FP32 step and the copy of K, the load and the condensed
    # system into pinned host memory (what a host caller's ALOC/CLOC receive)
    host = {k: torch.empty(v.shape, dtype=v.dtype, pin_memory=True) for k, v in out32.items()}
    d2h_bytes = sum(v.numel() * v.element_size() for v in out32.values())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.celem(out32, hp3d_amd.FP32)
        for k, v in out32.items():
            host[k].copy_(v, non_blocking=True)
    torch.cuda.synchronize()
    se2e = (time.perf_counter() - t0) / args.steps
    del host

    # BF16 trial: bfloat16 operands, FP32 accumulation, block-tiled element
    # kernel then the blocked condensation
    out16 = batch.alloc_outputs(hp3d_amd.BF16)
    s16, e16 = time_steps(batch, out16, hp3d_amd.BF16, max(2, args.steps // 2), 1)
    batch.check_status()
    k16max, k16med = rel_err_per_element(batch, out16["k"], out64["k"], batch._n * batch._n)
    a16max, _a16med = rel_err_per_element(batch, out16["aii"], out64["aii"], batch._ni * batch._ni)
    del out16

    fk, fc = algorithmic_flops(batch)
    nel = batch.nel
    nb_total = float(batch.nb.sum())
    kmax, kmed = rel_err_per_element(batch, out32["k"], out64["k"], batch._n * batch._n)
    amax, amed = rel_err_per_element(batch, out32["aii"], out64["aii"], batch._ni * batch._ni)

    dropin = plain_dropin_rate(m, min(args.dropin_sample, nel)) if space == 0 else None
    value = nel / s32
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "element matrices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "%s: %s" % (args.workload, WORKLOADS[args.workload]),
                   "elements": nel, "ndof_per_element": int(batch.max_ndof),
                   "form": "batched FP32 MFMA contraction, FP64 geometry and Cholesky condensation"},
        "condensed_dof_per_s": nb_total / s32,
        "interface_rows_per_s": float(batch.ni.sum()) / s32,
        "forms": {
            "batched_fp32": {"element_matrices_per_s": value, "ms_per_step": s32 * 1e3,
                             "elem_kernel_ms": e32,
                             "bins": [list(b) + [p] for b, p in zip(batch.bins, batch.bin_paths)]},
            "batched_fp32_graph": {"element_matrices_per_s": nel / sg, "ms_per_step": sg * 1e3},
            "batched_fp32_end_to_end": {"element_matrices_per_s": nel / se2e, "ms_per_step": se2e * 1e3,
                                        "d2h_bytes": d2h_bytes,
                                        "note": "step + D2H of K, load, Aii, Bi, ASchur, BSchur "
                                                "into pinned host memory"},
            "batched_fp64": {"element_matrices_per_s": nel / s64, "ms_per_step": s64 * 1e3,
                             "elem_kernel_ms": e64,
                             "condensed_dof_per_s": nb_total / s64},
            "plain_fp64_dropin": ({"element_matrices_per_s": dropin,
                                   "note": "hp3d_gpu_celem per element, host arrays, %d elements"
                                           % min(args.dropin_sample, nel)} if dropin else None),
            "fp32_over_dropin": (value / dropin) if dropin else None,
            "bf16_trial": {"element_matrices_per_s": nel / s16, "ms_per_step": s16 * 1e3,
                           "k_err_max": k16max, "k_err_median": k16med, "aii_err_max": a16max,
                           "note": "bfloat16 operands, FP32 accumulation (two-pass path)"},
        },
        "fp32_vs_fp64_error": {"k_max": kmax, "k_median": kmed, "aii_max": amax,
                               "aii_median": amed},
        "tuning": tuning,
    }
    parts = []
    for (lo, cnt, mnd, mnb), path in zip(batch.bins, batch.bin_paths):
        if path == "fused":
            r = 64 if mnd <= 64 and mnb <= 16 else 128
            parts.append("elem_single_kernel<float, %d rows> (element + condensation, %d elements)"
                         % (r, cnt))
        else:
            ek = ("elem_single_kernel<float>" if mnd <= 128 else
                  "geometry_kernel + elem_kernel<float>")
            parts.append("%s + stc_blocked_kernel<float> (%d elements)" % (ek, cnt))
    kname = " ; ".join(parts)
    achieved = (fk + fc) / (e32 * 1e-3)
    line["roofline"] = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK["fp32"] / 1e12,
                        "unit": "TFLOP/s", "frac": achieved / PEAK["fp32"], "traffic": None,
                        "kernel": kname, "flops_per_launch": fk + fc,
                        "flops_k": fk, "flops_condensation": fc}
    tfile = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.workload)
    if os.path.exists(tfile):
        with open(tfile) as f:
            line["roofline"]["traffic"] = json.load(f).get("elem_kernel_fp32_bytes")
    line["roofline_fp64"] = {"achieved": (fk + fc) / (e64 * 1e-3) / 1e12,
                             "peak": PEAK["fp64"] / 1e12, "unit": "TFLOP/s",
                             "frac": (fk + fc) / (e64 * 1e-3) / PEAK["fp64"]}
    if args.workload == "cfg4":
        line["adjoint"] = adjoint_leg(m, args.fd)
    if not args.no_cpu:
        cb, errs = cpu_baseline(m, batch, out32, out64)
        line["cpu_baseline"] = cb
        line["error_vs_oracle_sample"] = errs
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))


if __name__ == "__main__":
    main()
