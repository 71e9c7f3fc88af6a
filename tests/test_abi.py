"""The C ABI library loads, exports what include/hp3d_gpu.h declares, and its
host-side entry points behave as documented (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hp3d_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hp3d_gpu_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from hp3d_amd import _lib

    assert os.path.exists(_lib.LIB_PATH), "build() first"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 12
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"


def test_ndof_entry_matches_oracle_and_rejects_bad_orders():
    import oracle
    from hp3d_amd import _lib

    lib = _lib.load()
    rng = np.random.default_rng(0)
    n, nb = ctypes.c_int(), ctypes.c_int()
    for _ in range(50):
        no = np.zeros(19, np.int32)
        no[:12] = rng.integers(1, 7, 12)
        no[12:18] = rng.integers(1, 7, 6) * 10 + rng.integers(1, 7, 6)
        p = rng.integers(1, 7, 3)
        no[18] = p[0] * 100 + p[1] * 10 + p[2]
        ptr = no.ctypes.data_as(ctypes.c_void_p)
        assert lib.hp3d_gpu_ndof(0, ptr, ctypes.byref(n), ctypes.byref(nb)) == 0
        assert n.value == oracle.ndof_h1(no) and nb.value == oracle.nbubble_h1(no[18])
        assert lib.hp3d_gpu_ndof(1, ptr, ctypes.byref(n), ctypes.byref(nb)) == 0
        assert n.value == oracle.ndof_hcurl(no) and nb.value == oracle.nbubble_hcurl(no[18])
    no[3] = 7  # order above MAXP: checkorder stops the reference
    assert lib.hp3d_gpu_ndof(0, no.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n),
                             ctypes.byref(nb)) == _lib.E_ORDER
    assert lib.hp3d_gpu_ndof(5, no.ctypes.data_as(ctypes.c_void_p), None, None) == _lib.E_ARG


def test_entries_before_init_and_bad_arguments():
    from hp3d_amd import _lib

    lib = _lib.load()
    no = np.array([2] * 12 + [22] * 6 + [222], np.int32)
    z = np.zeros(18, np.int32)
    x = np.zeros((8, 3))
    a = np.zeros((27, 27))
    b = np.zeros(27)
    p = lambda arr: arr.ctypes.data_as(ctypes.c_void_p)
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        # without a device the per-element entries refuse to run
        assert lib.hp3d_gpu_elem(0, p(no), p(z), p(z[12:]), p(x), 8, 1.0, 1, p(a), 27, p(b)) == _lib.E_INIT
        assert lib.hp3d_gpu_init(0, 0) == _lib.E_INIT
    # argument checks happen before any device work
    assert lib.hp3d_gpu_elem_batch(7, 0, 1, None, None, None, None, 8, None, 0, None, 0, None,
                                   None, None, None, None, 1, None) == _lib.E_ARG
    assert lib.hp3d_gpu_elem_batch(0, 0, 1, None, None, None, None, 8, None, 0, None, 0, None,
                                   None, None, None, None, 1, None) == _lib.E_ARG
    assert lib.hp3d_gpu_stc_batch(0, 3, 1, None, None, 0, None, None, None, None, 0, None, None,
                                  None, None, None, None, None, 8, 1, None) == _lib.E_ARG
    assert lib.hp3d_gpu_elem_batch(0, 0, 0, None, None, None, None, 8, None, 0, None, 0, None,
                                   None, None, None, None, 1, None) == 0  # empty batch
    assert lib.hp3d_gpu_stc_workspace(10, 125, 0) == 0
    # slab per element, in 16 x 16 tiles over nbk = ceil(nb / 16) block rows: the
    # lower factor L and its transformed copy (nbk (nbk + 1) / 2 tiles each) and
    # the inverted diagonal blocks and their transposes (nbk each); independent
    # of N
    for nb in (27, 240):
        nbk = -(-nb // 16)
        assert lib.hp3d_gpu_stc_workspace(10, 540, nb) == 10 * (nbk * (nbk + 1) + 2 * nbk) * 256 * 8
    assert lib.hp3d_gpu_stc_workspace(10, 125, 27) == lib.hp3d_gpu_stc_workspace(10, 540, 27)
    # geometry slabs: one per element of the block-tiled path (343 points x 10 or
    # 21 doubles + a flag), none for the single-block kernel
    assert lib.hp3d_gpu_elem_workspace(0, 1, 10, 128) == 0
    assert lib.hp3d_gpu_elem_workspace(0, 2, 10, 27) == 0  # BF16 trial: single-block too
    assert lib.hp3d_gpu_elem_workspace(0, 2, 10, 200) == 10 * (343 * 10 + 2) * 8
    assert lib.hp3d_gpu_elem_workspace(1, 0, 10, 540) == 10 * (343 * 21 + 2) * 8
    assert lib.hp3d_gpu_celem_workspace(0, 1, 10, 125, 27) == 0
    assert lib.hp3d_gpu_celem_workspace(1, 1, 10, 540, 240) == max(
        lib.hp3d_gpu_elem_workspace(1, 1, 10, 540), lib.hp3d_gpu_stc_workspace(10, 540, 240))


def test_python_mirror_counts():
    from hp3d_amd.elem import check_orders, ndof, nquad
    from hp3d_amd import Hp3dError

    no = np.array([[4] * 12 + [44] * 6 + [444]])
    assert ndof(0, no)[0][0] == 125 and ndof(0, no)[1][0] == 27
    assert ndof(1, no)[0][0] == 300 and ndof(1, no)[1][0] == 108
    assert (nquad(no, np.zeros((1, 6), int)) == 5).all()
    bad = no.copy()
    bad[0, 18] = 171
    with pytest.raises(Hp3dError):
        check_orders(bad)
