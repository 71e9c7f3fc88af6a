"""Test oracle for the hp3D hexahedral element path.

TEST INFRASTRUCTURE ONLY.  ``oracle.lib`` is the C restatement in
``oracle/hp3d_oracle.c`` (the checker and the timed CPU baseline);
``oracle.ref()`` is the reference's own Fortran shape-function / quadrature /
geom subset compiled by ``oracle/Makefile`` into ``oracle/_ref`` (present only
where it was built).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this package.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "liboracle.so")
_REF = os.path.join(HERE, "_ref", "libhp3d_ref.so")
REF_PATH = _REF  # tests check for it without loading the library

MAXP = 6
MAXBRICKH = 343
MAXBRICKE = 882
MAX_NINT3 = 392

LOAD_ZERO, LOAD_SINSINSIN, LOAD_CONV, LOAD_ONE, LOAD_CURL_SIN = 0, 1, 2, 3, 10
LOAD_POLY2 = 4
LOAD_XYZ = 5

_i = ctypes.c_int
_d = ctypes.c_double
_p = ctypes.c_void_p


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def build():
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = ctypes.CDLL(_LIB)
    sig = {
        "orc_poly_legendre": (None, [_d, _d, _i, _p]),
        "orc_poly_ilegendre": (None, [_d, _d, _i, _i, _p, _p, _p]),
        "orc_gauss1d": (None, [_i, _p, _p]),
        "orc_set_3d_int": (None, [_p, _p, _p, _p, _p]),
        "orc_find_order_loc": (None, [_p, _p, _p]),
        "orc_ndof_h1": (_i, [_p]),
        "orc_ndof_hcurl": (_i, [_p]),
        "orc_nbubble_h1": (_i, [_i]),
        "orc_nbubble_hcurl": (_i, [_i]),
        "orc_checkorder": (_i, [_p]),
        "orc_shape3dh_hexa": (_i, [_p, _p, _p, _p, _p, _p, _p]),
        "orc_shape3de_hexa": (_i, [_p, _p, _p, _p, _p, _p, _p]),
        "orc_geom": (_i, [_p, _p, _p]),
        "orc_geom3d": (_i, [_p, _i, _p, _p, _p, _p, _p, _p]),
        "orc_elem_h1": (_i, [_p, _p, _p, _p, _i, _d, _i, _p, _i, _p]),
        "orc_elem_hcurl": (_i, [_p, _p, _p, _p, _i, _d, _d, _i, _p, _i, _p]),
        "orc_stc_fwd_gen": (_i, [_i, _i, _p, _i, _p, _p, _p]),
        "orc_stc_fwd_herm": (_i, [_i, _i, _p, _i, _p, _p, _p]),
        "orc_stc_bwd": (None, [_i, _i, _p, _p, _p, _p]),
        "orc_set_3dint": (None, [_p, _i, _p, _p, _p]),
        "orc_exact_h1": (None, [_i, _p, _p, _p]),
        "orc_exact_hcurl": (None, [_i, _p, _p, _p]),
        "orc_element_error": (_i, [_i, _p, _p, _p, _p, _i, _p, _i, _p, _p]),
        "orc_elem_mass": (_i, [_i, _p, _p, _p, _p, _i, _p, _i]),
        "orc_element_loop": (_i, [_i, _i, _p, _p, _p, _p, _i, _p, _i, _i,
                                  _p, _p, _p, _p, _p, _p, _i, _i]),
        "orc_set_blas": (_i, [_p, _p, _p, _p]),
        "orc_element_loop_blas": (_i, [_i, _i, _p, _p, _p, _p, _i, _p, _i, _i,
                                       _p, _p, _p, _p, _p, _p, _i]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()
_ref_lib = None


def ref():
    """The reference's own routines (oracle/_ref), or None when not built."""
    global _ref_lib
    if _ref_lib is None and os.path.exists(_REF):
        r = ctypes.CDLL(_REF)
        r.ref_shape3dh_hexa.argtypes = [_p] * 7
        r.ref_shape3de_hexa.argtypes = [_p] * 7
        r.ref_poly_legendre.argtypes = [_d, _d, _i, _p]
        r.ref_poly_ilegendre.argtypes = [_d, _d, _i, _p, _p]
        r.ref_gauss1d.argtypes = [_i, _p, _p]
        r.ref_geom.argtypes = [_p, _p, _p, _p]
        for f in (r.ref_shape3dh_hexa, r.ref_shape3de_hexa, r.ref_poly_legendre,
                  r.ref_poly_ilegendre, r.ref_gauss1d, r.ref_geom):
            f.restype = None
        _ref_lib = r
    return _ref_lib


def _i32(a, n):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))
    assert a.size == n
    return a


# ------------------------------------------------------------------ wrappers

def poly_legendre(x, t, nord):
    p = np.zeros(nord + 1)
    lib.orc_poly_legendre(x, t, nord, _ptr(p))
    return p


def poly_ilegendre(x, nord):
    """L[2..nord] and P[1..nord-1] with t = 1 (returned zero-padded from index 0)."""
    l_ = np.zeros(nord + 1)
    p = np.zeros(nord + 1)
    r = np.zeros(nord + 1)
    lib.orc_poly_ilegendre(x, 1.0, nord, 1, _ptr(l_), _ptr(p), _ptr(r))
    return l_, p


def gauss1d(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib.orc_gauss1d(n, _ptr(x), _ptr(w))
    return x, w


def set_3d_int(norder, nori_f):
    no = _i32(norder, 19)
    nf = _i32(nori_f, 6)
    nint = np.zeros(1, np.int32)
    xi = np.zeros((MAX_NINT3, 3))
    wa = np.zeros(MAX_NINT3)
    lib.orc_set_3d_int(_ptr(no), _ptr(nf), _ptr(nint), _ptr(xi), _ptr(wa))
    n = int(nint[0])
    return xi[:n].copy(), wa[:n].copy()


def find_order_loc(norder, nori_f):
    out = np.zeros(19, np.int32)
    lib.orc_find_order_loc(_ptr(_i32(norder, 19)), _ptr(_i32(nori_f, 6)), _ptr(out))
    return out


def ndof_h1(norder):
    return lib.orc_ndof_h1(_ptr(_i32(norder, 19)))


def ndof_hcurl(norder):
    return lib.orc_ndof_hcurl(_ptr(_i32(norder, 19)))


def nbubble_h1(nord_mdl):
    return lib.orc_nbubble_h1(int(nord_mdl))


def nbubble_hcurl(nord_mdl):
    return lib.orc_nbubble_hcurl(int(nord_mdl))


def shape3dh(xi, norder, nori_e, nori_f):
    """(shap[N], grad[N,3]) of the oracle's shape3DHHexa."""
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    shap = np.zeros(MAXBRICKH)
    grad = np.zeros((MAXBRICKH, 3))
    n = np.zeros(1, np.int32)
    st = lib.orc_shape3dh_hexa(_ptr(xi), _ptr(_i32(norder, 19)), _ptr(_i32(nori_e, 12)),
                               _ptr(_i32(nori_f, 6)), _ptr(n), _ptr(shap), _ptr(grad))
    if st:
        raise ValueError("shape3dh: order outside [1, MAXP]")
    k = int(n[0])
    return shap[:k].copy(), grad[:k].copy()


def shape3de(xi, norder, nori_e, nori_f):
    """(shap[N,3], curl[N,3]) of the oracle's shape3DEHexa."""
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    shap = np.zeros((MAXBRICKE, 3))
    curl = np.zeros((MAXBRICKE, 3))
    n = np.zeros(1, np.int32)
    st = lib.orc_shape3de_hexa(_ptr(xi), _ptr(_i32(norder, 19)), _ptr(_i32(nori_e, 12)),
                               _ptr(_i32(nori_f, 6)), _ptr(n), _ptr(shap), _ptr(curl))
    if st:
        raise ValueError("shape3de: order outside [1, MAXP]")
    k = int(n[0])
    return shap[:k].copy(), curl[:k].copy()


def ref_shape3dh(xi, norder, nori_e, nori_f):
    r = ref()
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    shap = np.zeros(MAXBRICKH)
    grad = np.zeros((MAXBRICKH, 3))
    n = np.zeros(1, np.int32)
    r.ref_shape3dh_hexa(_ptr(xi), _ptr(_i32(norder, 19)), _ptr(_i32(nori_e, 12)),
                        _ptr(_i32(nori_f, 6)), _ptr(n), _ptr(shap), _ptr(grad))
    k = int(n[0])
    return shap[:k].copy(), grad[:k].copy()


def ref_shape3de(xi, norder, nori_e, nori_f):
    r = ref()
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    shap = np.zeros((MAXBRICKE, 3))
    curl = np.zeros((MAXBRICKE, 3))
    n = np.zeros(1, np.int32)
    r.ref_shape3de_hexa(_ptr(xi), _ptr(_i32(norder, 19)), _ptr(_i32(nori_e, 12)),
                        _ptr(_i32(nori_f, 6)), _ptr(n), _ptr(shap), _ptr(curl))
    k = int(n[0])
    return shap[:k].copy(), curl[:k].copy()


def ref_poly_legendre(x, t, nord):
    p = np.zeros(nord + 1)
    ref().ref_poly_legendre(x, t, nord, _ptr(p))
    return p


def ref_poly_ilegendre(x, nord):
    l_ = np.zeros(nord + 1)
    p = np.zeros(nord + 1)
    lbuf = np.zeros(max(nord - 1, 1))
    pbuf = np.zeros(max(nord - 1, 1))
    ref().ref_poly_ilegendre(x, 1.0, nord, _ptr(lbuf), _ptr(pbuf))
    l_[2:] = lbuf[: nord - 1]
    p[1:nord] = pbuf[: nord - 1]
    return l_, p


def ref_gauss1d(n):
    x = np.zeros(n)
    w = np.zeros(n)
    ref().ref_gauss1d(n, _ptr(x), _ptr(w))
    return x, w


def geom(dxdxi):
    """dxdxi[3,3] row = physical, col = master.  Returns (dxidx, rjac, iflag)."""
    a = np.asfortranarray(dxdxi, dtype=np.float64)
    b = np.zeros((3, 3), order="F")
    r = np.zeros(1)
    fl = lib.orc_geom(_ptr(a), _ptr(b), _ptr(r))
    return b, float(r[0]), fl


def ref_geom(dxdxi):
    a = np.asfortranarray(dxdxi, dtype=np.float64)
    b = np.zeros((3, 3), order="F")
    r = np.zeros(1)
    fl = np.zeros(1, np.int32)
    ref().ref_geom(_ptr(a), _ptr(b), _ptr(r), _ptr(fl))
    return b, float(r[0]), int(fl[0])


def elem_h1(norder, nori_e, nori_f, xnod, kappa=1.0, load=LOAD_SINSINSIN):
    """(K[N,N], b[N], iflag) of the oracle's elem_opt (Poisson)."""
    xnod = np.asfortranarray(xnod, dtype=np.float64)  # 3 x ngeo
    n = ndof_h1(norder)
    k = np.zeros((n, n), order="F")
    b = np.zeros(n)
    st = lib.orc_elem_h1(_ptr(_i32(norder, 19)), _ptr(_i32(nori_e, 12)), _ptr(_i32(nori_f, 6)),
                         _ptr(xnod), xnod.shape[1], float(kappa), load, _ptr(k), n, _ptr(b))
    if st < 0:
        raise ValueError("elem_h1 failed: %d" % st)
    return np.array(k), b, st


def elem_hcurl(norder, nori_e, nori_f, xnod, mu=1.0, eps=1.0, load=LOAD_CURL_SIN):
    xnod = np.asfortranarray(xnod, dtype=np.float64)
    n = ndof_hcurl(norder)
    k = np.zeros((n, n), order="F")
    b = np.zeros(n)
    st = lib.orc_elem_hcurl(_ptr(_i32(norder, 19)), _ptr(_i32(nori_e, 12)),
                            _ptr(_i32(nori_f, 6)), _ptr(xnod), xnod.shape[1], float(mu),
                            float(eps), load, _ptr(k), n, _ptr(b))
    if st < 0:
        raise ValueError("elem_hcurl failed: %d" % st)
    return np.array(k), b, st


def stc_fwd(k, b, nb, herm=False):
    """Condense the last nb dofs.  Returns (Aii, Bi, ASchur[nb,ni], BSchur[nb])."""
    n = k.shape[0]
    ni = n - nb
    a = np.asfortranarray(k, dtype=np.float64).copy(order="F")
    bb = np.array(b, dtype=np.float64)
    asch = np.zeros((nb, ni), order="F")
    bsch = np.zeros(nb)
    f = lib.orc_stc_fwd_herm if herm else lib.orc_stc_fwd_gen
    st = f(ni, nb, _ptr(a), n, _ptr(bb), _ptr(asch), _ptr(bsch))
    if st:
        raise ValueError("stc_fwd failed: %d" % st)
    return np.array(a[:ni, :ni]), bb[:ni].copy(), np.array(asch), bsch


def stc_bwd(aschur, bschur, xi):
    nb, ni = aschur.shape
    a = np.asfortranarray(aschur)
    xb = np.zeros(nb)
    lib.orc_stc_bwd(ni, nb, _ptr(a), _ptr(np.ascontiguousarray(bschur)),
                    _ptr(np.ascontiguousarray(xi, dtype=np.float64)), _ptr(xb))
    return xb


def element_loop(space, norder, nori_e, nori_f, xnod, coef, load, ldk, want_k=True,
                 want_stc=True, herm=False, nthreads=0):
    """Batched loop.  norder [nel,19], nori_e [nel,12], nori_f [nel,6], xnod [nel,ngeo,3].

    Returns dict of arrays shaped [nel, ldk, ldk] (column-major per element, i.e. the
    [e, j, i] entry is row i, column j) and [nel, ldk]."""
    norder = np.ascontiguousarray(norder, dtype=np.int32)
    nel = norder.shape[0]
    nori_e = np.ascontiguousarray(nori_e, dtype=np.int32)
    nori_f = np.ascontiguousarray(nori_f, dtype=np.int32)
    xnod = np.ascontiguousarray(xnod, dtype=np.float64)
    ngeo = xnod.shape[1]
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    out = {}
    z2 = lambda: np.zeros((nel, ldk, ldk))
    z1 = lambda: np.zeros((nel, ldk))
    if want_k:
        out["k"], out["b"] = z2(), z1()
    if want_stc:
        out["aii"], out["bi"], out["aschur"], out["bschur"] = z2(), z1(), z2(), z1()
    g = out.get
    st = lib.orc_element_loop(space, nel, _ptr(norder), _ptr(nori_e), _ptr(nori_f), _ptr(xnod),
                              ngeo, _ptr(coef), load, ldk, _ptr(g("k")), _ptr(g("b")),
                              _ptr(g("aii")), _ptr(g("bi")), _ptr(g("aschur")),
                              _ptr(g("bschur")), 1 if herm else 0, nthreads)
    out["status"] = st
    return out


_blas_name = None


def host_blas():
    """The host BLAS/LAPACK the reference-shaped loop calls: scipy's bundled
    OpenBLAS (32-bit integers, Fortran ABI, symbols prefixed scipy_), loaded
    once and handed to orc_set_blas.  Returns a description of the library."""
    global _blas_name
    if _blas_name is not None:
        return _blas_name
    import glob

    import scipy

    libdir = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
    cands = sorted(glob.glob(os.path.join(libdir, "libscipy_openblas*.so")))
    if not cands:
        raise RuntimeError("no host BLAS found under %s" % libdir)
    blas = ctypes.CDLL(cands[0])
    addr = lambda n: ctypes.cast(getattr(blas, n), ctypes.c_void_p)
    blas.scipy_openblas_set_num_threads.argtypes = [_i]
    blas.scipy_openblas_set_num_threads(1)   # one BLAS thread per OpenMP thread
    rc = lib.orc_set_blas(addr("scipy_dsyrk_"), addr("scipy_dgemm_"), addr("scipy_dgetrf_"),
                          addr("scipy_dgetrs_"))
    if rc != 0:
        raise RuntimeError("orc_set_blas failed")
    cfg = getattr(blas, "scipy_openblas_get_config", None)
    desc = os.path.basename(cands[0])
    if cfg is not None:
        cfg.restype = ctypes.c_char_p
        desc += " (" + cfg().decode().strip() + ")"
    _blas_name = desc
    return desc


def element_loop_blas(space, norder, nori_e, nori_f, xnod, coef, load, ldk, nthreads=0,
                      want_k=True, want_stc=True):
    """The element loop as the reference builds it: per element elem_opt with
    DSYRK and stc_fwd_gen with DGETRF/DGETRS/DGEMM from the host BLAS, OpenMP
    dynamic schedule over elements (mumps_sc.F90:224-342).  Outputs as
    element_loop."""
    host_blas()
    norder = np.ascontiguousarray(norder, dtype=np.int32)
    nel = norder.shape[0]
    nori_e = np.ascontiguousarray(nori_e, dtype=np.int32)
    nori_f = np.ascontiguousarray(nori_f, dtype=np.int32)
    xnod = np.ascontiguousarray(xnod, dtype=np.float64)
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    out = {}
    if want_k:
        out["k"], out["b"] = np.zeros((nel, ldk, ldk)), np.zeros((nel, ldk))
    if want_stc:
        out["aii"], out["bi"] = np.zeros((nel, ldk, ldk)), np.zeros((nel, ldk))
        out["aschur"], out["bschur"] = np.zeros((nel, ldk, ldk)), np.zeros((nel, ldk))
    g = out.get
    st = lib.orc_element_loop_blas(space, nel, _ptr(norder), _ptr(nori_e), _ptr(nori_f),
                                   _ptr(xnod), xnod.shape[1], _ptr(coef), load, ldk, _ptr(g("k")),
                                   _ptr(g("b")), _ptr(g("aii")), _ptr(g("bi")),
                                   _ptr(g("aschur")), _ptr(g("bschur")), nthreads)
    out["status"] = st
    return out


def set_3dint(norder, integration=2):
    """set_3Dint (no orientation swap) with INTEGRATION = integration."""
    xi = np.zeros((MAX_NINT3 + 200, 3))
    w = np.zeros(MAX_NINT3 + 200)
    n = ctypes.c_int()
    lib.orc_set_3dint(_ptr(_i32(norder, 19)), int(integration), ctypes.byref(n), _ptr(xi), _ptr(w))
    return xi[:n.value].copy(), w[:n.value].copy()


def exact_h1(sel, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    u = np.zeros(1)
    du = np.zeros(3)
    lib.orc_exact_h1(int(sel), _ptr(x), _ptr(u), _ptr(du))
    return u[0], du


def exact_hcurl(sel, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    e = np.zeros(3)
    c = np.zeros(3)
    lib.orc_exact_hcurl(int(sel), _ptr(x), _ptr(e), _ptr(c))
    return e, c


def element_error(space, norder, nori_e, nori_f, xnod, dof, sel):
    """element_error: (squared error, squared exact norm) of one element.
    xnod is 3 x ngeo; dof the element's full dof vector (element order)."""
    xnod = np.asfortranarray(xnod, dtype=np.float64)
    dof = np.ascontiguousarray(dof, dtype=np.float64)
    err = np.zeros(1)
    nrm = np.zeros(1)
    st = lib.orc_element_error(int(space), _ptr(_i32(norder, 19)), _ptr(_i32(nori_e, 12)),
                               _ptr(_i32(nori_f, 6)), _ptr(xnod), xnod.shape[1], _ptr(dof),
                               int(sel), _ptr(err), _ptr(nrm))
    if st < 0:
        raise ValueError("element_error failed: %d" % st)
    return err[0], nrm[0]


def elem_mass(space, norder, nori_e, nori_f, xnod):
    xnod = np.asfortranarray(xnod, dtype=np.float64)
    n = ndof_h1(norder) if space == 0 else ndof_hcurl(norder)
    m = np.zeros((n, n), order="F")
    st = lib.orc_elem_mass(int(space), _ptr(_i32(norder, 19)), _ptr(_i32(nori_e, 12)),
                           _ptr(_i32(nori_f, 6)), _ptr(xnod), xnod.shape[1], _ptr(m), n)
    if st < 0:
        raise ValueError("elem_mass failed: %d" % st)
    return np.array(m)
