"""CPU-side checks of the shipped library: it loads, exports every symbol the
header declares, and its host-side set-up equals the checker's."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cwhip
from cwhip import lib as cl
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "cwhip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cwhip_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    L = cwhip.load()
    names = header_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(cl.EXPORTS)


def test_header_limits_equal_python_constants():
    txt = open(os.path.join(ROOT, "include", "cwhip.h")).read()
    defs = {name: int(v) << int(sh or 0) for name, v, sh in
            re.findall(r"#define\s+(CWHIP_[A-Z_]+)\s+\(?([0-9]+)(?:ll)?(?:\s*<<\s*([0-9]+))?", txt)}
    assert defs["CWHIP_MAX_LEN"] == cl.MAX_LEN == 32767
    assert defs["CWHIP_NOALLOC_SCRATCH_BYTES"] == cl.NOALLOC_SCRATCH_BYTES
    assert defs["CWHIP_NOALLOC_MAX_PAIRS"] == cl.NOALLOC_MAX_PAIRS
    # (row << 16 | column) keys stay non-negative up to the limit
    assert (cl.MAX_LEN << 16 | cl.MAX_LEN) < 2 ** 31


def test_calibration_library_loads():
    """The VALU calibration kernels bench.py --valu-probe times live in their
    own library, and without a device the call reports an error, not a
    number."""
    lib = C.CDLL(os.path.join(ROOT, "cuda-clustalw_amd", "libcwcal.so"))
    lib.cwcal_valu.argtypes = [C.c_int, C.c_int, C.c_int] + [C.POINTER(C.c_double)] * 3
    ms, instrs, ops = C.c_double(), C.c_double(), C.c_double()
    assert lib.cwcal_valu(99, 10, 4, C.byref(ms), C.byref(instrs), C.byref(ops)) == -1
    assert lib.cwcal_valu(0, 10, 6, C.byref(ms), C.byref(instrs), C.byref(ops)) == -1
    assert lib.cwcal_valu(0, 10, 4, C.byref(ms), C.byref(instrs), C.byref(ops)) < 0


def test_binding_return_types():
    """Entry points that return something other than an int status keep their
    ctypes return type (a float read back as int reads 0)."""
    import ctypes as C
    L = cl.load()
    assert L.cwhip_distance.restype is C.c_float
    assert L.cwhip_fast_last_ms.restype is C.c_float
    assert L.cwhip_range_npairs.restype is C.c_int64
    assert L.cwhip_fast_pairwise_range.restype is C.c_int


def test_params_equal_checker():
    p = cwhip.pair_params(max_len=6000)
    m = np.ctypeslib.as_array(p.matrix).reshape(32, 32)
    sc = O.Scoring()
    assert (m == sc.mat).all()
    assert p.gap_extend == 1
    assert (p._keep[1:] == sc.gap_table(6000)[1:]).all()
    d = cwhip.pair_params(dna=True, max_len=100)
    sd = O.Scoring(dna=True)
    assert (np.ctypeslib.as_array(d.matrix).reshape(32, 32) == sd.mat).all()
    assert d.gap_extend == 666 and d._keep[50] == 3000


def test_distance_matches_checker():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        l1, l2 = (int(x) for x in rng.integers(0, 3000, 2))
        c = int(rng.integers(0, max(1, min(l1, l2)) + 1))
        a = np.float32(cwhip.distance(c, l1, l2))
        b = np.float32(O.lib().cwo_distance(c, l1, l2))
        assert a.view(np.uint32) == b.view(np.uint32)


@pytest.mark.parametrize("rng_", [(0, 10, 0, 10), (3, 7, 0, 10), (0, 10, 5, 8), (2, 2, 0, 10),
                                  (0, 100, 0, 100), (9, 20, 8, 20)])
def test_range_npairs(rng_):
    n = 10
    L = cwhip.load()
    assert L.cwhip_range_npairs(n, *rng_) == len(O.range_pairs(n, *rng_))


def test_no_gpu_context_fails_loudly():
    # in a container without a GPU the context must fail, not fall back
    try:
        import torch
        has = torch.cuda.is_available()
    except Exception:
        has = None
    if has:
        pytest.skip("a GPU is present")
    with pytest.raises(cwhip.CwhipError):
        cwhip.Context(0)


def test_server_needs_a_device():
    """No CPU fallback: the request loop cannot start without a device."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(cwhip.CwhipError):
        cwhip.Server(0, slots=2, hbm_budget=1 << 30)
