"""Host-side mirror of hp3D's element interface over libhp3d_gpu.

Per-element drop-ins (host arrays, one C-ABI call each), mirroring
  elem / elem_opt        problems/POISSON/GALERKIN/elem.F90:20-55, elem_opt.F90:22-141
                         problems/MAXWELL/GALERKIN/elem_opt.F90:22-158 (real form)
  stc_fwd_wrapper        src/modules/stc.F90:182-311
  celem_systemI up to    src/constrs/celem_systemI.F90:523-534 (elem + stc_fwd)
and the batched form of the mumps_sc Step 2 element loop
(src/solver/omp_mumps/mumps_sc.F90:224-342) over device tensors.

Matrices are column-major per element, as ALOC is: element e's K occupies
kflat[off_k[e] : off_k[e] + N_e*N_e] with leading dimension N_e.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (BF16, FP32, FP64, H1, HCURL, LOAD_CURL_SIN, LOAD_ONE, LOAD_SINSINSIN,  # noqa: F401
                   LOAD_ZERO, LOAD_CONV, Hp3dError, check)

MAXP = 6


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _tptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


# ------------------------------------------------------------------ counts

def decode_orders(norder):
    """(edge orders (..,12), face h (..,6), face v (..,6), interior (..,3))."""
    no = np.asarray(norder)
    pe = no[..., :12]
    fh = no[..., 12:18] // 10
    fv = no[..., 12:18] % 10
    b = np.stack([no[..., 18] // 100, (no[..., 18] // 10) % 10, no[..., 18] % 10], -1)
    return pe, fh, fv, b


def ndof(space, norder):
    """(N, nb) per element: celndof / ndof_nod (element/util/celndof.F90:24,
    element_data.F90:845-855) for hexahedra, nb = bubbles of the middle node."""
    pe, fh, fv, b = decode_orders(norder)
    if space == H1:
        nb = (b[..., 0] - 1) * (b[..., 1] - 1) * (b[..., 2] - 1)
        n = 8 + (pe - 1).sum(-1) + ((fh - 1) * (fv - 1)).sum(-1) + nb
    else:
        nb = (b[..., 0] * (b[..., 1] - 1) * (b[..., 2] - 1) + (b[..., 0] - 1) * b[..., 1] *
              (b[..., 2] - 1) + (b[..., 0] - 1) * (b[..., 1] - 1) * b[..., 2])
        n = pe.sum(-1) + (fh * (fv - 1) + (fh - 1) * fv).sum(-1) + nb
    return n, nb


def check_orders(norder):
    """checkorder (element/shape_1/CheckOrder.F90:139-152): orders in [1, MAXP]."""
    pe, fh, fv, b = decode_orders(norder)
    allp = np.concatenate([pe.reshape(-1), fh.reshape(-1), fv.reshape(-1), b.reshape(-1)])
    if allp.size and (allp.min() < 1 or allp.max() > MAXP):
        raise Hp3dError("order outside [1, %d]" % MAXP)


def nquad(norder, nori_f):
    """Gauss points per element direction (set_3D_int brick branch with
    find_order_loc, src/element/quadrature/set_3D_int.F90:203-244)."""
    pe, fh, fv, b = decode_orders(norder)
    nf = np.atleast_2d(nori_f)
    pe, fh, fv, b = [np.atleast_2d(x) for x in (pe, fh, fv, b)]
    m = b.copy()
    edir = [0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2]
    for e in range(12):
        m[:, edir[e]] = np.maximum(m[:, edir[e]], pe[:, e])
    axes = [(0, 1), (0, 1), (0, 2), (1, 2), (0, 2), (1, 2)]
    swapped = np.array([0, 1, 0, 1, 1, 0, 1, 0])[nf]
    for f in range(6):
        o0 = np.where(swapped[:, f] == 1, fv[:, f], fh[:, f])
        o1 = np.where(swapped[:, f] == 1, fh[:, f], fv[:, f])
        m[:, axes[f][0]] = np.maximum(m[:, axes[f][0]], o0)
        m[:, axes[f][1]] = np.maximum(m[:, axes[f][1]], o1)
    return np.minimum(m, MAXP) + 1


# ---------------------------------------------------- per-element drop-ins

def _codes(norder, nori_e, nori_f):
    no = np.ascontiguousarray(np.asarray(norder, dtype=np.int32).reshape(19))
    oe = np.ascontiguousarray(np.asarray(nori_e, dtype=np.int32).reshape(12))
    of = np.ascontiguousarray(np.asarray(nori_f, dtype=np.int32).reshape(6))
    return no, oe, of


def elem(space, norder, nori_e, nori_f, xnod, coef=1.0, load=LOAD_SINSINSIN, ld=None):
    """elem_opt for one element on the GPU.  xnod: (ngeo, 3) geometry dofs.

    Returns (K (N,N), b (N,), flag) with flag 1 when a negative Jacobian was seen."""
    lib = _lib.init()
    no, oe, of = _codes(norder, nori_e, nori_f)
    n, _ = ndof(space, no)
    n = int(n)
    ld = n if ld is None else int(ld)
    x = np.ascontiguousarray(xnod, dtype=np.float64)
    aloc = np.zeros((ld, n), order="F")  # ALOC(1,1)%array with leading dimension ld
    bloc = np.zeros(n)
    rc = lib.hp3d_gpu_elem(space, _ptr(no), _ptr(oe), _ptr(of), _ptr(x), x.shape[0],
                           float(coef), int(load), _ptr(aloc), ld, _ptr(bloc))
    flag = check(rc, "hp3d_gpu_elem")
    return np.array(aloc[:n, :n]), bloc, flag


def stc_fwd(k, b, nb):
    """stc_fwd_wrapper for one element: condense the last nb dofs.

    Returns (Aii (ni,ni), Bi (ni,), ASchur (nb,ni), BSchur (nb,))."""
    lib = _lib.init()
    n = k.shape[0]
    ni = n - nb
    a = np.asfortranarray(np.array(k, dtype=np.float64))
    bb = np.array(b, dtype=np.float64)
    asch = np.zeros((nb, ni), order="F")
    bsch = np.zeros(nb)
    rc = lib.hp3d_gpu_stc_fwd(ni, nb, _ptr(a), n, _ptr(bb), _ptr(asch), _ptr(bsch))
    check(rc, "hp3d_gpu_stc_fwd")
    return np.array(a[:ni, :ni]), bb[:ni].copy(), np.array(asch), bsch


def celem(space, norder, nori_e, nori_f, xnod, coef=1.0, load=LOAD_SINSINSIN):
    """elem + stc_fwd for one element: (Aii, Bi, ASchur, BSchur, flag)."""
    lib = _lib.init()
    no, oe, of = _codes(norder, nori_e, nori_f)
    n, nb = ndof(space, no)
    n, nb = int(n), int(nb)
    ni = n - nb
    x = np.ascontiguousarray(xnod, dtype=np.float64)
    a = np.zeros((n, n), order="F")
    bl = np.zeros(n)
    asch = np.zeros((max(nb, 1), ni), order="F")
    bsch = np.zeros(max(nb, 1))
    rc = lib.hp3d_gpu_celem(space, _ptr(no), _ptr(oe), _ptr(of), _ptr(x), x.shape[0],
                            float(coef), int(load), _ptr(a), n, _ptr(bl), _ptr(asch), _ptr(bsch))
    flag = check(rc, "hp3d_gpu_celem")
    if nb == 0:
        return np.array(a), bl, np.zeros((0, n)), np.zeros(0), flag
    return np.array(a[:ni, :ni]), bl[:ni].copy(), np.array(asch), bsch, flag


# ------------------------------------------------------------ batched form

class ElementBatch:
    """Element data resident in device memory, ready for the batched kernels.

    norder (nel,19), nori_e (nel,12), nori_f (nel,6) int; xnod (nel,ngeo,3);
    coef (nel,).  Elements are grouped into order bins (the replacement of the
    OpenMP element loop by one launch per bin): bin 0 holds the elements whose
    matrix fits one 128 x 128 block with at most 32 bubbles -- K stays in
    registers and condensation runs in the same kernel -- and bin 1 the rest
    (block-tiled element kernel, then the condensation kernel); `rules`
    changes the bin limits and paths (hp3d_amd.tune picks them by timing).
    Internally the elements are stored bin by bin; every accessor takes the
    caller's element index.  Outputs are packed: element i's N_i x N_i matrix starts at
    off_k[i] (internal order) with leading dimension N_i.
    """

    FUSED_MAX_NDOF = 128
    FUSED_MAX_NBUBBLE = 32
    # bin rules: an element goes to the first rule whose limits it meets; a
    # "fused" bin runs the single-block kernel (K in registers, condensation in
    # the same launch; needs N <= 128, nb <= 32; a bin with N <= 64 and
    # nb <= 16 gets its 64-row form), as a persistent grid when the rule says
    # "persistent": True; a "two_pass" bin the block-tiled element kernel and
    # then the blocked condensation
    SMALL_MAX_NDOF = 64
    SMALL_MAX_NBUBBLE = 16
    DEFAULT_RULES = ({"max_ndof": 64, "max_nbubble": 16, "path": "fused"},
                     {"max_ndof": 128, "max_nbubble": 32, "path": "fused"}, {"path": "two_pass"})

    def __init__(self, space, norder, nori_e, nori_f, xnod, coef, load=LOAD_SINSINSIN,
                 device=None, binned=True, rules=None):
        import torch

        _lib.init()
        self.space = space
        self.load = int(load)
        norder = np.ascontiguousarray(norder, dtype=np.int32)
        check_orders(norder)
        self.nel = norder.shape[0]
        n, nb = [np.asarray(v, dtype=np.int64) for v in ndof(space, norder)]
        rules = list(self.DEFAULT_RULES if rules is None else rules)
        if not binned:
            rules = [{"path": "fused"}] if ((n <= self.FUSED_MAX_NDOF) &
                                           (nb <= self.FUSED_MAX_NBUBBLE)).all() else [{"path": "two_pass"}]
        which = np.full(self.nel, -1, np.int64)
        for ri, r in enumerate(rules):
            ok = (which < 0) & (n <= r.get("max_ndof", 10 ** 9)) & (nb <= r.get("max_nbubble", 10 ** 9))
            if r["path"] == "fused":
                ok &= (n <= self.FUSED_MAX_NDOF) & (nb <= self.FUSED_MAX_NBUBBLE)
            which[ok] = ri
        if (which < 0).any():
            raise ValueError("bin rules leave elements unassigned (the last rule should be open)")
        self.rules = rules
        # bin by bin; within a bin the largest elements first, so the last
        # workgroups of a launch are the cheapest ones
        self.perm = np.lexsort((-nb, -n, which))
        self.iperm = np.empty_like(self.perm)
        self.iperm[self.perm] = np.arange(self.nel)
        P = self.perm
        # caller order (public) and internal (bin) order
        self.n, self.nb, self.ni = n, nb, n - nb
        self._n, self._nb = n[P], nb[P]
        self._ni = self._n - self._nb
        self.max_ndof = int(n.max()) if self.nel else 1
        self.max_nbubble = int(nb.max()) if self.nel else 0
        self.norder_h = norder[P]
        self.bins, self.bin_paths, self.bin_persistent = [], [], []
        ws = which[P]
        for ri, r in enumerate(rules):
            idx = np.nonzero(ws == ri)[0]
            if idx.size:
                lo, hi = int(idx[0]), int(idx[-1]) + 1
                self.bins.append((lo, hi - lo, int(self._n[lo:hi].max()), int(self._nb[lo:hi].max())))
                self.bin_paths.append(r["path"])
                self.bin_persistent.append(bool(r.get("persistent", False)))
        dev = torch.device("cuda") if device is None else torch.device(device)
        self.device = dev
        t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(dev)
        self.norder = t(self.norder_h, torch.int32)
        self.nori_e = t(np.asarray(nori_e)[P], torch.int32)
        self.nori_f = t(np.asarray(nori_f)[P], torch.int32)
        self.nori_f_h = np.asarray(nori_f)[P]
        xnod = np.ascontiguousarray(np.asarray(xnod, dtype=np.float64)[P])
        self.ngeo = xnod.shape[1]
        self.xnod = t(xnod, torch.float64)
        self.coef = t(np.asarray(coef, dtype=np.float64)[P], torch.float64)
        # packed offsets (internal order)
        n2 = self._n * self._n
        self.off_k_h = np.concatenate([[0], np.cumsum(n2)[:-1]]).astype(np.int64)
        self.off_v_h = np.concatenate([[0], np.cumsum(self._n)[:-1]]).astype(np.int64)
        ni2 = self._ni * self._ni
        self.off_c_h = np.concatenate([[0], np.cumsum(ni2)[:-1]]).astype(np.int64)
        nbi = self._nb * self._ni
        self.off_s_h = np.concatenate([[0], np.cumsum(nbi)[:-1]]).astype(np.int64)
        self.size_k = int(n2.sum())
        self.size_v = int(self._n.sum())
        self.size_c = int(ni2.sum())
        self.size_s = int(nbi.sum())
        self.off_k = t(self.off_k_h, torch.int64)
        self.off_v = t(self.off_v_h, torch.int64)
        self.off_c = t(self.off_c_h, torch.int64)
        self.off_s = t(self.off_s_h, torch.int64)
        self.status = torch.zeros(self.nel, dtype=torch.int32, device=dev)
        self.workspace = torch.empty(self._workspace_bytes(FP32) // 8 + 1, dtype=torch.float64, device=dev)

    def _workspace_bytes(self, precision):
        """Workspace of the per-bin calls: the geometry of the block-tiled
        element pass and the condensation slabs (one after the other)."""
        lib = _lib.load()
        ws = 8
        for _, cnt, mnd, mnb in self.bins:
            ws = max(ws, int(lib.hp3d_gpu_stc_workspace(cnt, mnd, mnb)),
                     int(lib.hp3d_gpu_elem_workspace(self.space, precision, cnt, mnd)),
                     int(lib.hp3d_gpu_celem_workspace(self.space, precision, cnt, mnd, mnb)))
        return ws

    def _ensure_workspace(self, precision):
        """Grow the workspace for `precision`; never called during a capture."""
        import torch

        need = self._workspace_bytes(precision)
        if self.workspace.numel() * 8 < need:
            self.workspace = torch.empty(need // 8 + 1, dtype=torch.float64, device=self.device)

    def _ptrs(self, start):
        """Device pointers of the per-element arrays from internal element `start`."""
        def at(tensor, per):
            if tensor is None:
                return None
            return ctypes.c_void_p(tensor.data_ptr() + start * per * tensor.element_size())
        return at

    @classmethod
    def from_mesh(cls, space, mesh, load=None, device=None, rules=None):
        if load is None:
            load = LOAD_SINSINSIN if space == H1 else LOAD_CURL_SIN
        return cls(space, mesh.norder, mesh.nori_e, mesh.nori_f, mesh.xnod(), mesh.coef,
                   load=load, device=device, rules=rules)

    def alloc_outputs(self, precision=FP64, want_k=True, want_stc=True):
        import torch

        dt = torch.float64 if precision == FP64 else torch.float32
        out = {}
        dev = self.device
        # two-pass bins condense from K in HBM
        if want_k or (want_stc and "two_pass" in self.bin_paths):
            out["k"] = torch.empty(self.size_k, dtype=dt, device=dev)
            out["b"] = torch.empty(self.size_v, dtype=torch.float64, device=dev)
        if want_stc:
            out["aii"] = torch.empty(max(self.size_c, 1), dtype=torch.float64, device=dev)
            out["bi"] = torch.empty(self.size_v, dtype=torch.float64, device=dev)
            out["aschur"] = torch.empty(max(self.size_s, 1), dtype=torch.float64, device=dev)
            out["bschur"] = torch.empty(self.size_v, dtype=torch.float64, device=dev)
        return out

    def _element_bin(self, lib, out, precision, st, lo, cnt, mnd):
        at = self._ptrs(lo)
        rc = lib.hp3d_gpu_elem_batch(
            self.space, precision, cnt, at(self.norder, 19), at(self.nori_e, 12),
            at(self.nori_f, 6), at(self.xnod, 3 * self.ngeo), self.ngeo, at(self.coef, 1),
            self.load, at(self.off_k, 1), 0, _tptr(out["k"]), at(self.off_v, 1),
            _tptr(out["b"]), _tptr(self.workspace), at(self.status, 1), mnd, ctypes.c_void_p(st))
        check(rc, "hp3d_gpu_elem_batch")

    def _condense_bin(self, lib, out, precision, st, lo, cnt, mnd, mnb):
        at = self._ptrs(lo)
        rc = lib.hp3d_gpu_stc_batch(
            self.space, precision, cnt, at(self.norder, 19), at(self.off_k, 1), 0,
            _tptr(out["k"]), at(self.off_v, 1), _tptr(out["b"]), at(self.off_c, 1), 0,
            _tptr(out["aii"]), _tptr(out["bi"]), at(self.off_s, 1), _tptr(out["aschur"]),
            _tptr(out["bschur"]), _tptr(self.workspace), at(self.status, 1), mnd, mnb,
            ctypes.c_void_p(st))
        check(rc, "hp3d_gpu_stc_batch")

    def element_matrices(self, out, precision=FP64, stream=None):
        """Launch the element kernel(s) into out['k'], out['b'] (no sync)."""
        import torch

        self._ensure_workspace(precision)
        lib = _lib.load()
        st = torch.cuda.current_stream().cuda_stream if stream is None else stream
        for lo, cnt, mnd, _mnb in self.bins:
            self._element_bin(lib, out, precision, st, lo, cnt, mnd)

    def condense(self, out, precision=FP64, stream=None):
        """Launch static condensation of out['k'], out['b'] (no sync)."""
        import torch

        lib = _lib.load()
        st = torch.cuda.current_stream().cuda_stream if stream is None else stream
        for lo, cnt, mnd, mnb in self.bins:
            self._condense_bin(lib, out, precision, st, lo, cnt, mnd, mnb)

    def celem(self, out, precision=FP64, stream=None, concurrent=True):
        """Element matrices and condensation: per bin one launch (fused bins)
        or two (two-pass bins), no sync.  out['k'] and out['b'] may be absent
        when every bin is fused.  With concurrent=True (and stream left to the
        current torch stream) the fused bins and the two-pass bins run on two
        side streams forked from and joined back to the current stream, so the
        tail of one launch overlaps the other group's work; the bins write
        disjoint outputs and only the two-pass bins use the workspace."""
        import torch

        self._ensure_workspace(precision)
        lib = _lib.load()
        # the fused bins in one group, the two-pass bins (sharing the workspace)
        # in the other; a stream per fused bin measured slower (cfg2 2.49 ms
        # against 2.31 ms): the fused kernels pack worse beside each other
        groups = [g for g in ([i for i, p in enumerate(self.bin_paths) if p == "fused"],
                              [i for i, p in enumerate(self.bin_paths) if p == "two_pass"]) if g]
        if concurrent and stream is None and len(groups) > 1:
            main = torch.cuda.current_stream(self.device)
            side = getattr(self, "_side", [])
            while len(side) < len(groups):
                side.append(torch.cuda.Stream(self.device))
            self._side = side
            for sd, group in zip(side, groups):
                sd.wait_stream(main)
                for i in group:
                    self._bin_celem(lib, out, precision, sd.cuda_stream, i)
            for sd in side[:len(groups)]:
                main.wait_stream(sd)
            return
        st = torch.cuda.current_stream().cuda_stream if stream is None else stream
        for i in range(len(self.bins)):
            self._bin_celem(lib, out, precision, st, i)

    def _bin_celem(self, lib, out, precision, st, i):
        (lo, cnt, mnd, mnb), path, pers = self.bins[i], self.bin_paths[i], self.bin_persistent[i]
        if path == "two_pass":
            self._element_bin(lib, out, precision, st, lo, cnt, mnd)
            self._condense_bin(lib, out, precision, st, lo, cnt, mnd, mnb)
            return
        if pers:
            check(lib.hp3d_gpu_set_option(_lib.OPT_PERSISTENT, 1), "hp3d_gpu_set_option")
        at = self._ptrs(lo)
        rc = lib.hp3d_gpu_celem_batch(
            self.space, precision, cnt, at(self.norder, 19), at(self.nori_e, 12),
            at(self.nori_f, 6), at(self.xnod, 3 * self.ngeo), self.ngeo, at(self.coef, 1),
            self.load, at(self.off_k, 1), 0, _tptr(out.get("k")), at(self.off_v, 1),
            _tptr(out.get("b")), at(self.off_c, 1), _tptr(out["aii"]), _tptr(out["bi"]),
            at(self.off_s, 1), _tptr(out["aschur"]), _tptr(out["bschur"]),
            _tptr(self.workspace), at(self.status, 1), mnd, mnb, ctypes.c_void_p(st))
        if pers:
            lib.hp3d_gpu_set_option(_lib.OPT_PERSISTENT, 0)
        check(rc, "hp3d_gpu_celem_batch")

    def capture(self, out, precision=FP64):
        """The per-bin launches of celem() captured in a HIP graph (replay with
        graph.replay()); inputs and outputs stay where they are (the eager
        run before the capture sizes the workspace)."""
        import torch

        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.celem(out, precision)
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.celem(out, precision)
        return g

    def run(self, precision=FP64, want_k=True, want_stc=True, out=None, fused=True):
        """Element matrices and (optionally) condensation; returns the outputs.
        fused=False runs every bin as two passes."""
        if out is None:
            out = self.alloc_outputs(precision, want_k or not fused, want_stc)
        if want_stc and fused:
            self.celem(out, precision)
            return out
        self.element_matrices(out, precision)
        if want_stc:
            self.condense(out, precision)
        return out

    def check_status(self):
        """Per-element status in the caller's order; raises on a fatal code."""
        s = self.status.cpu().numpy()[self.iperm]
        if (s < 0).any():
            e = int(np.nonzero(s < 0)[0][0])
            check(int(s[e]), "element %d" % e)
        return s

    # ---------------------------------------------------------- host views
    def k_of(self, out, e, key="k"):
        e = int(self.iperm[e])
        n = int(self._n[e])
        o = int(self.off_k_h[e])
        return out[key][o:o + n * n].double().cpu().numpy().reshape(n, n).T

    def aii_of(self, out, e):
        e = int(self.iperm[e])
        ni = int(self._ni[e])
        o = int(self.off_c_h[e])
        return out["aii"][o:o + ni * ni].cpu().numpy().reshape(ni, ni).T

    def aschur_of(self, out, e):
        e = int(self.iperm[e])
        ni, nb = int(self._ni[e]), int(self._nb[e])
        o = int(self.off_s_h[e])
        return out["aschur"][o:o + nb * ni].cpu().numpy().reshape(ni, nb).T

    def vec_of(self, out, key, e, length=None):
        e = int(self.iperm[e])
        o = int(self.off_v_h[e])
        n = int(self._n[e]) if length is None else int(length)
        return out[key][o:o + n].cpu().numpy()
