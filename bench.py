"""Throughput of the hp3D element-computation hot path on one MI355X.

One step = element stiffness + load of every element of the workload's mesh
followed by static condensation of all bubble dofs, inputs resident in HBM,
outputs left in HBM.  The headline workload is the 32x32x32 perturbed
hexahedral mesh at uniform p = 4 (32,768 elements, 125 dofs and 125 Gauss
points each, lognormal kappa_e); `per_config` carries the same measurement for
the mixed-order mesh (20,000 elements, orders 1..6), the H(curl) p = 5 mesh
(4,096 elements) and the 24^3 p = 3 mesh (13,824 elements).

value      element matrices/s of the batched FP32-contraction form (FP64
           geometry and bubble factorisation), whole job
forms      the same for the batched FP64 form and for the plain FP64 drop-in
           (one C-ABI call per element from the host, as the reference's
           OpenMP loop would make it, from one thread and from a pool of
           host threads)
roofline   the step's launches: algorithmic flops (k N (N+1) per element plus
           the condensation count) over their mean duration (HIP events on
           the launch stream)
cpu_baseline  the oracle's OpenMP element loop in the reference's own form
           (DSYRK, DGETRF/DGETRS/DGEMM from the host OpenBLAS) on a bounded
           sample of the same mesh, on this host's cores

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload cfg1]
"""
import argparse
import json
import os
import sys
import threading
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


def plain_dropin_rate(m, space, nsample, threads=1):
    """Per-element C-ABI calls (host arrays in and out), as the reference loop
    would make them: from one thread, or from `threads` host threads at once
    (the OpenMP loop; each thread has its own context, made before timing)."""
    from concurrent.futures import ThreadPoolExecutor

    import hp3d_amd

    xn = m.xnod()
    load = hp3d_amd.LOAD_SINSINSIN if space == 0 else hp3d_amd.LOAD_CURL_SIN
    one = lambda e: hp3d_amd.celem(space, m.norder[e], m.nori_e[e], m.nori_f[e], xn[e],
                                   coef=m.coef[e], load=load)
    if threads == 1:
        one(0)
        t0 = time.perf_counter()
        for e in range(nsample):
            one(e)
        return nsample / (time.perf_counter() - t0)
    with ThreadPoolExecutor(threads) as ex:
        # one call per worker first, held at a barrier so each worker makes its context
        bar = threading.Barrier(threads)
        list(ex.map(lambda e: (one(e), bar.wait(timeout=300)), range(threads)))
        t0 = time.perf_counter()
        list(ex.map(one, range(nsample)))
        return nsample / (time.perf_counter() - t0)


def host_cpu():
    """Model name and the CPUs this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def cpu_threads():
    """OpenMP threads of the CPU baseline: every CPU of this process's
    affinity set, at most the host share OMP_NUM_THREADS grants (16 per GPU on
    the measurement hosts)."""
    _, ncpu = host_cpu()
    cap = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(ncpu, cap))


def cpu_baseline(m, batch, out32, out64, seconds=15.0, single_core=True):
    """The reference-shaped element loop (elem_opt with DSYRK, stc_fwd_gen
    with DGETRF/DGETRS/DGEMM, OpenMP dynamic schedule: mumps_sc.F90:224-342)
    on the first elements of the mesh, sized to about `seconds` of work; the
    GPU outputs are compared with it on up to 200 of the sampled elements."""
    import oracle

    threads = cpu_threads()
    space = batch.space
    load = oracle.LOAD_SINSINSIN if space == 0 else oracle.LOAD_CURL_SIN
    ldk = int(batch.max_ndof)
    blas = oracle.host_blas()
    loop = lambda sl, nt: oracle.element_loop_blas(
        space, m.norder[sl], m.nori_e[sl], m.nori_f[sl], m.xnod()[sl], m.coef[sl], load,
        ldk=ldk, nthreads=nt)
    # calibrate the sample so the timed run is ~seconds of work: a first call
    # warms the BLAS and OpenMP thread pools, a second one gives the rate
    probe = min(m.nel, max(64, 8 * threads))
    loop(slice(0, min(m.nel, 2 * threads)), threads)
    t0 = time.perf_counter()
    loop(slice(0, probe), threads)
    dt = max(time.perf_counter() - t0, 1e-3)
    ns = int(min(m.nel, max(probe, probe * seconds / dt)))
    t0 = time.perf_counter()
    ref = loop(slice(0, ns), threads)
    dt = time.perf_counter() - t0
    # FP32 and FP64 GPU outputs against the oracle on the sample
    err = {"k32": [], "aii32": [], "k64": [], "aii64": []}
    for e in range(0, ns, max(1, ns // 200)):
        n, ni = int(batch.n[e]), int(batch.ni[e])
        kr = ref["k"][e].T[:n, :n]
        ar = ref["aii"][e].T[:ni, :ni]
        for tag, o in (("32", out32), ("64", out64)):
            err["k" + tag].append(np.linalg.norm(batch.k_of(o, e) - kr) / np.linalg.norm(kr))
            if ni < n:
                err["aii" + tag].append(np.linalg.norm(batch.aii_of(o, e) - ar) / np.linalg.norm(ar))
    del ref
    model, ncpu = host_cpu()
    res = {"value": ns / dt, "unit": "element matrices/s", "cores": threads, "kind": "port",
           "sample": "first %d of %d elements, elem_opt (DSYRK) + stc_fwd_gen (DGETRF/DGETRS/"
                     "DGEMM) per element, OpenMP dynamic schedule, %.1f s" % (ns, m.nel, dt),
           "blas": blas, "cpu_model": model, "cpus_visible": ncpu}
    if single_core:
        # the same loop on one core, on a sample of about a third of `seconds`
        n1 = int(min(ns, max(1, ns / dt / threads * seconds / 3)))
        t0 = time.perf_counter()
        loop(slice(0, n1), 1)
        dt1 = time.perf_counter() - t0
        res["single_core"] = {"value": n1 / dt1, "sample": "first %d elements, %.1f s" % (n1, dt1)}
    errors = {k: {"max": float(np.max(v)), "median": float(np.median(v))} for k, v in err.items() if v}
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


def kernel_names(batch):
    parts = []
    for (lo, cnt, mnd, mnb), path in zip(batch.bins, batch.bin_paths):
        if path == "fused":
            r = 64 if mnd <= 64 and mnb <= 16 else 128
            parts.append("elem_single_kernel<float, %d rows> (element + condensation, %d elements)"
                         % (r, cnt))
        else:
            ek = ("elem_single_kernel<float>" if mnd <= 128 else
                  "geometry_kernel + elem_kernel<float>")
            parts.append("%s + stc_strip_kernel<float> (%d elements)" % (ek, cnt))
    return " ; ".join(parts)


def traffic_of(workload):
    """HBM bytes per step from the committed counter summary of this workload
    (FETCH_SIZE and WRITE_SIZE passes, scaled by a plain copy's reading)."""
    tfile = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    if os.path.exists(tfile):
        with open(tfile) as f:
            return json.load(f).get("elem_kernel_fp32_bytes")
    return None


def measure(workload, args, headline):
    """Batched FP32 and FP64 steps of one workload (autotuned order bins), the
    plain per-element drop-in, FP32-vs-FP64 errors, roofline and (unless
    --no-cpu) the CPU baseline; the headline workload adds graph replay, the
    end-to-end form and the BF16 trial."""
    import torch

    import hp3d_amd
    from hp3d_amd import mesh

    m = mesh.config_mesh(workload)
    space = hp3d_amd.HCURL if workload == "cfg3" else hp3d_amd.H1
    rules, tuning = None, None
    if not args.no_tune:
        from hp3d_amd import tune

        rules, tuning = tune.autotune(space, m, hp3d_amd.FP32, reps=3)
    batch = hp3d_amd.ElementBatch.from_mesh(space, m, rules=rules)
    steps = args.steps
    out32 = batch.alloc_outputs(hp3d_amd.FP32)
    out64 = batch.alloc_outputs(hp3d_amd.FP64)

    # batched FP64 (reference form on the GPU) and FP32 (headline form)
    s64, e64 = time_steps(batch, out64, hp3d_amd.FP64, max(2, steps // 2), 1)
    batch.check_status()
    s32, e32 = time_steps(batch, out32, hp3d_amd.FP32, steps, args.warmup)
    batch.check_status()
    nel = batch.nel
    nb_total = float(batch.nb.sum())
    value = nel / s32
    fk, fc = algorithmic_flops(batch)
    kmax, kmed = rel_err_per_element(batch, out32["k"], out64["k"], batch._n * batch._n)
    amax, amed = rel_err_per_element(batch, out32["aii"], out64["aii"], batch._ni * batch._ni)
    nd = min(args.dropin_sample if space == 0 else 16, nel)
    dropin = plain_dropin_rate(m, space, nd)
    dropin_t = plain_dropin_rate(m, space, 2 * nd, threads=cpu_threads())
    forms = {
        "batched_fp32": {"element_matrices_per_s": value, "ms_per_step": s32 * 1e3,
                         "elem_kernel_ms": e32, "condensed_dof_per_s": nb_total / s32,
                         "bins": [list(b) + [p] for b, p in zip(batch.bins, batch.bin_paths)]},
        "batched_fp64": {"element_matrices_per_s": nel / s64, "ms_per_step": s64 * 1e3,
                         "elem_kernel_ms": e64, "condensed_dof_per_s": nb_total / s64},
        "plain_fp64_dropin": {"element_matrices_per_s": dropin,
                              "note": "hp3d_gpu_celem per element, host arrays, %d elements" % nd},
        "plain_fp64_dropin_threads": {"element_matrices_per_s": dropin_t, "threads": cpu_threads(),
                                      "note": "the same calls from an OpenMP-like pool of host "
                                              "threads, %d elements" % (2 * nd)},
        "fp32_over_dropin": value / dropin,
        "fp32_over_fp64_batched": s64 / s32,
    }
    if headline:
        # the same FP32 step as one replay of a captured graph of the per-bin launches
        g = batch.capture(out32, hp3d_amd.FP32)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.replay()
        torch.cuda.synchronize()
        sg = (time.perf_counter() - t0) / steps
        forms["batched_fp32_graph"] = {"element_matrices_per_s": nel / sg, "ms_per_step": sg * 1e3}
        del g
        # end to end: the FP32 step and the copy of K, the load and the
        # condensed system into pinned host memory (what a host caller's
        # ALOC/CLOC receive)
        host = {k: torch.empty(v.shape, dtype=v.dtype, pin_memory=True) for k, v in out32.items()}
        d2h_bytes = sum(v.numel() * v.element_size() for v in out32.values())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            batch.celem(out32, hp3d_amd.FP32)
            for k, v in out32.items():
                host[k].copy_(v, non_blocking=True)
        torch.cuda.synchronize()
        se2e = (time.perf_counter() - t0) / steps
        del host
        forms["batched_fp32_end_to_end"] = {
            "element_matrices_per_s": nel / se2e, "ms_per_step": se2e * 1e3, "d2h_bytes": d2h_bytes,
            "note": "step + D2H of K, load, Aii, Bi, ASchur, BSchur into pinned host memory"}
        # BF16 trial: bfloat16 operands, FP32 accumulation, on the same bins
        # (fused single-block bins, block-tiled element kernel plus strip
        # condensation for the larger elements)
        out16 = batch.alloc_outputs(hp3d_amd.BF16)
        s16, _e16 = time_steps(batch, out16, hp3d_amd.BF16, max(2, steps // 2), 1)
        batch.check_status()
        k16max, k16med = rel_err_per_element(batch, out16["k"], out64["k"], batch._n * batch._n)
        a16max, _ = rel_err_per_element(batch, out16["aii"], out64["aii"], batch._ni * batch._ni)
        del out16
        forms["bf16_trial"] = {"element_matrices_per_s": nel / s16, "ms_per_step": s16 * 1e3,
                               "k_err_max": k16max, "k_err_median": k16med, "aii_err_max": a16max,
                               "note": "bfloat16 operands, FP32 accumulation, the FP32 form's bins"}
    achieved = (fk + fc) / (e32 * 1e-3)
    entry = {
        "workload": "%s: %s" % (workload, WORKLOADS[workload]),
        "elements": nel,
        "ndof_per_element_max": int(batch.max_ndof),
        "element_matrices_per_s": value,
        "ms_per_step": s32 * 1e3,
        "condensed_dof_per_s": nb_total / s32,
        "interface_rows_per_s": float(batch.ni.sum()) / s32,
        "forms": forms,
        "fp32_vs_fp64_error": {"k_max": kmax, "k_median": kmed, "aii_max": amax,
                               "aii_median": amed},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK["fp32"] / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK["fp32"],
                     "traffic": traffic_of(workload), "kernel": kernel_names(batch),
                     "flops_per_launch": fk + fc, "flops_k": fk, "flops_condensation": fc},
        "roofline_fp64": {"achieved": (fk + fc) / (e64 * 1e-3) / 1e12, "peak": PEAK["fp64"] / 1e12,
                          "unit": "TFLOP/s", "frac": (fk + fc) / (e64 * 1e-3) / PEAK["fp64"]},
        "tuning": tuning,
    }
    if workload == "cfg2":
        entry["target"] = {
            "statement": ">= 10x the plain FP64 drop-in's element matrices/s at <= 1e-6 "
                         "relative element-matrix error",
            "speedup_over_dropin": value / dropin, "k_err_max": kmax,
            "met": bool(value / dropin >= 10.0 and kmax <= 1e-6)}
    if not args.no_cpu:
        cb, errs = cpu_baseline(m, batch, out32, out64, seconds=15.0 if headline else 5.0,
                                single_core=headline)
        entry["cpu_baseline"] = cb
        entry["error_vs_oracle_sample"] = errs
    else:
        entry["cpu_baseline"] = None
    extra = (m, batch)
    del out32, out64
    torch.cuda.empty_cache()
    return entry, extra


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="cfg1", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-per-config", action="store_true",
                    help="measure only --workload (no per_config block)")
    ap.add_argument("--dropin-sample", type=int, default=256)
    ap.add_argument("--kernel-only", action="store_true",
                    help="time only the FP32 element+condensation launches (for counter runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run one device-to-device copy of MIB MiB and exit (counter scale)")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (default bin rules)")
    ap.add_argument("--fd", type=int, default=50,
                    help="--workload cfg4: elements checked by central differences (step 1e-6)")
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
    if args.kernel_only:
        m = mesh.config_mesh(args.workload)
        space = hp3d_amd.HCURL if args.workload == "cfg3" else hp3d_amd.H1
        batch = hp3d_amd.ElementBatch.from_mesh(space, m)
        out32 = batch.alloc_outputs(hp3d_amd.FP32)
        s32, e32 = time_steps(batch, out32, hp3d_amd.FP32, args.steps, args.warmup)
        batch.check_status()
        print(json.dumps({"workload": args.workload, "ms_per_step": s32 * 1e3,
                          "kernel_ms": e32, "bins": [list(b) for b in batch.bins]}))
        return

    t_start = time.perf_counter()
    head, (m, _batch) = measure(args.workload, args, headline=True)
    print("# %s measured (%.0f s)" % (args.workload, time.perf_counter() - t_start),
          file=sys.stderr, flush=True)
    line = {
        "metric": METRIC,
        "value": head["element_matrices_per_s"],
        "unit": "element matrices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": head["workload"], "elements": head["elements"],
                   "ndof_per_element": head["ndof_per_element_max"],
                   "form": "batched FP32 MFMA contraction, FP64 geometry and Cholesky condensation"},
        "condensed_dof_per_s": head["condensed_dof_per_s"],
        "interface_rows_per_s": head["interface_rows_per_s"],
        "forms": head["forms"],
        "fp32_vs_fp64_error": head["fp32_vs_fp64_error"],
        "tuning": head["tuning"],
        "roofline": head["roofline"],
        "roofline_fp64": head["roofline_fp64"],
        "cpu_baseline": head["cpu_baseline"],
    }
    if "error_vs_oracle_sample" in head:
        line["error_vs_oracle_sample"] = head["error_vs_oracle_sample"]
    if "target" in head:
        line["target"] = head["target"]
    if args.workload == "cfg4":
        line["adjoint"] = adjoint_leg(m, args.fd)
    del _batch
    if not args.no_per_config:
        per = {}
        for w in sorted(WORKLOADS):
            if w == args.workload:
                continue
            t0 = time.perf_counter()
            entry, _ = measure(w, args, headline=False)
            entry.pop("tuning", None)
            per[w] = entry
            print("# %s measured (%.0f s)" % (w, time.perf_counter() - t0), file=sys.stderr,
                  flush=True)
        line["per_config"] = per
    print(json.dumps(line))


if __name__ == "__main__":
    main()
