"""hp3d_gpu_celem_batch_noalloc (no workspace argument, the library's static
device scratch) against hp3d_gpu_celem_batch with a caller workspace, through
the C ABI: every output bit-identical, on two calls in a row, the second with
a smaller input.  The first call's two-pass bin (H(curl) p = 6, about 1 MB
of workspace per element) needs more than one chunk of the scratch."""
import ctypes

import numpy as np
import pytest

from hp3d_amd import mesh

pytestmark = pytest.mark.gpu


def _celem_all_bins(lib, batch, out, precision, noalloc):
    """One C-ABI call per order bin: the allocation-free entry, or the
    existing entry with the batch's workspace."""
    import torch

    from hp3d_amd import _lib

    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    tp = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
    for lo, cnt, mnd, mnb in batch.bins:
        at = batch._ptrs(lo)
        head = (batch.space, precision, cnt, at(batch.norder, 19), at(batch.nori_e, 12),
                at(batch.nori_f, 6), at(batch.xnod, 3 * batch.ngeo), batch.ngeo, at(batch.coef, 1),
                batch.load, at(batch.off_k, 1), 0, tp(out["k"]), at(batch.off_v, 1), tp(out["b"]),
                at(batch.off_c, 1), tp(out["aii"]), tp(out["bi"]), at(batch.off_s, 1),
                tp(out["aschur"]), tp(out["bschur"]))
        if noalloc:
            rc = lib.hp3d_gpu_celem_batch_noalloc(*head, at(batch.status, 1), mnd, mnb, st)
        else:
            rc = lib.hp3d_gpu_celem_batch(*head, tp(batch.workspace), at(batch.status, 1), mnd,
                                          mnb, st)
        _lib.check(rc, "celem_batch%s" % ("_noalloc" if noalloc else ""))


def _zeroed(batch, precision):
    out = batch.alloc_outputs(precision)
    for v in out.values():
        v.zero_()
    return out


def _mesh(dims, porder, seed):
    nel = int(np.prod(dims))
    return mesh.structured(dims, porder=porder, perturb=0.1, mesh_seed=seed, perm_seed=seed + 1,
                           coef=np.random.default_rng(seed).uniform(1.0, 4.0, nel))


def test_noalloc_matches_workspace_entry(gpu):
    import torch

    import hp3d_amd
    from hp3d_amd import _lib

    lib = _lib.load()
    # the first case's two-pass bin needs more than one chunk of the scratch
    per = lib.hp3d_gpu_celem_workspace(1, 1, 1, 882, 450)
    assert 0 < per < _lib.SCRATCH_BYTES and 150 > _lib.SCRATCH_BYTES // per
    # then a mixed-order H1 mesh (fused bins and a two-pass bin) in FP64,
    # smaller than the first call
    cases = [(_mesh((5, 5, 6), 6, 71), 1, hp3d_amd.FP32)]
    rng = np.random.default_rng(72)
    po = rng.choice(np.arange(1, 7), size=(3 * 3 * 2, 3), p=mesh.P_OUT)
    po[:3] = 6
    cases.append((mesh.structured((3, 3, 2), porder=po, perturb=0.1, mesh_seed=73, perm_seed=74,
                                  coef=np.exp(rng.normal(size=18))), 0, hp3d_amd.FP64))
    cases.append((_mesh((2, 2, 1), 5, 75), 1, hp3d_amd.FP32))
    for i, (m, space, prec) in enumerate(cases):
        print("case %d: space %d, %d elements, precision %d" % (i, space, m.nel, prec), flush=True)
        batch = hp3d_amd.ElementBatch.from_mesh(space, m)
        assert "two_pass" in batch.bin_paths
        ref = _zeroed(batch, prec)
        _celem_all_bins(lib, batch, ref, prec, noalloc=False)
        torch.cuda.synchronize()
        st_ref = batch.status.clone()
        got = _zeroed(batch, prec)
        batch.status.fill_(99)
        _celem_all_bins(lib, batch, got, prec, noalloc=True)
        torch.cuda.synchronize()
        assert torch.equal(batch.status, st_ref) and int(st_ref.abs().max()) == 0
        for key in ("k", "b", "aii", "bi", "aschur", "bschur"):
            assert torch.equal(got[key], ref[key]), (i, key)
