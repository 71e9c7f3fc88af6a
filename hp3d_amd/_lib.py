"""ctypes binding of libhp3d_gpu.so (the C ABI declared in include/hp3d_gpu.h).

The library is loaded after torch so that both share one HIP runtime
(torch's libamdhip64.so.7 is already resident when the library's NEEDED entry
is resolved).  There is no fallback: if the library is missing or no GPU is
visible, calls raise.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhp3d_gpu.so")

H1, HCURL = 0, 1
FP64, FP32, BF16 = 0, 1, 2
LOAD_ZERO, LOAD_SINSINSIN, LOAD_CONV, LOAD_ONE, LOAD_CURL_SIN = 0, 1, 2, 3, 10
LOAD_POLY2 = 4
LOAD_XYZ = 5
FIELD_ERROR, FIELD_ADJOINT, FIELD_L2 = 0, 1, 2
OPT_PERSISTENT = 1
SCRATCH_BYTES = 128 << 20   # HP3D_SCRATCH_BYTES

OK = 0
NEG_JACOBIAN = 1
E_ORDER, E_NOIFACE, E_FACTOR, E_ARG, E_HIP, E_INIT = -1, -3, -4, -5, -6, -7
ERRORS = {E_ORDER: "order outside [1, MAXP]", E_NOIFACE: "no interface dof",
          E_FACTOR: "bubble block is not positive definite", E_ARG: "bad argument",
          E_HIP: "HIP runtime error", E_INIT: "library not initialised / no device"}

_i, _d, _p, _sz = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t

SIGNATURES = {
    "hp3d_gpu_init": (_i, [_i, _i]),
    "hp3d_gpu_finalize": (_i, []),
    "hp3d_gpu_last_error": (ctypes.c_char_p, []),
    "hp3d_gpu_set_option": (_i, [_i, _i]),
    "hp3d_gpu_ndof": (_i, [_i, _p, _p, _p]),
    "hp3d_gpu_elem": (_i, [_i, _p, _p, _p, _p, _i, _d, _i, _p, _i, _p]),
    "hp3d_gpu_stc_fwd": (_i, [_i, _i, _p, _i, _p, _p, _p]),
    "hp3d_gpu_celem": (_i, [_i, _p, _p, _p, _p, _i, _d, _i, _p, _i, _p, _p, _p]),
    "hp3d_gpu_elem_batch": (_i, [_i, _i, _i, _p, _p, _p, _p, _i, _p, _i, _p, _i, _p, _p, _p,
                                 _p, _p, _i, _p]),
    "hp3d_gpu_elem_workspace": (_sz, [_i, _i, _i, _i]),
    "hp3d_gpu_celem_workspace": (_sz, [_i, _i, _i, _i, _i]),
    "hp3d_gpu_stc_batch": (_i, [_i, _i, _i, _p, _p, _i, _p, _p, _p, _p, _i, _p, _p, _p, _p,
                                _p, _p, _p, _i, _i, _p]),
    "hp3d_gpu_stc_workspace": (_sz, [_i, _i, _i]),
    "hp3d_gpu_celem_batch": (_i, [_i, _i, _i, _p, _p, _p, _p, _i, _p, _i, _p, _i, _p, _p, _p,
                                  _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _p]),
    "hp3d_gpu_celem_batch_noalloc": (_i, [_i, _i, _i, _p, _p, _p, _p, _i, _p, _i, _p, _i, _p,
                                          _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _p]),
    "hp3d_gpu_stc_bwd_batch": (_i, [_i, _i, _p, _p, _p, _p, _p, _p, _p]),
    "hp3d_gpu_field_batch": (_i, [_i, _i, _i, _p, _p, _p, _p, _i, _p, _p, _p, _p, _i, _p, _p,
                                  _p, _p]),
    "hp3d_gpu_dirichlet_lift": (_i, [_i, _i, _p, _p, _p, _p, _p, _p, _p]),
    "hp3d_gpu_shape": (_i, [_i, _p, _p, _p, _i, _p, _p, _p]),
    "hp3d_gpu_gather_sum": (_i, [ctypes.c_longlong, _p, _p, _p, _p, _p]),
    "hp3d_gpu_csr_spmv": (_i, [ctypes.c_longlong, _p, _p, _p, _p, _p, _p]),
}


class Hp3dError(RuntimeError):
    pass


_lib = None


def load():
    """Load libhp3d_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  -- one HIP runtime for torch and the library

    if not os.path.exists(LIB_PATH):
        raise Hp3dError("%s is not built; run __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc, what):
    """Raise on fatal codes; return rc (0, or 1 for a negative Jacobian)."""
    if rc < 0:
        msg = ERRORS.get(rc, "error %d" % rc)
        if rc in (E_HIP, E_INIT):
            msg += ": " + (load().hp3d_gpu_last_error() or b"").decode()
        raise Hp3dError("%s: %s" % (what, msg))
    return rc


_initialised = False


def init(device=0, max_ndof=0):
    """hp3d_gpu_init on `device` (idempotent)."""
    global _initialised
    import torch

    if not torch.cuda.is_available():
        raise Hp3dError("no GPU visible: the HIP path has no CPU fallback")
    lib = load()
    if not _initialised:
        torch.cuda.set_device(device)
        check(lib.hp3d_gpu_init(device, max_ndof), "hp3d_gpu_init")
        _initialised = True
    return lib
