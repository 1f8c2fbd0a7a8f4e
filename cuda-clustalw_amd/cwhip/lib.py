"""ctypes binding of libcwhip.so (the C ABI in include/cwhip.h).

This module only marshals arrays; every alignment runs in the gfx950 kernels.
Loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libcwhip.so")

KERNEL_PLAIN = 0
KERNEL_BATCHED = 1

EXPORTS = [
    "cwhip_create", "cwhip_destroy", "cwhip_last_error", "cwhip_pair_params_default",
    "cwhip_load_sequences", "cwhip_range_npairs", "cwhip_pairwise_range",
    "cwhip_pairwise_pairs", "cwhip_distance", "cwhip_stage_pairs", "cwhip_run_staged",
    "cwhip_fetch_counts", "cwhip_fetch_info", "cwhip_set_tuning",
]


class PairParams(C.Structure):
    _fields_ = [("matrix", C.c_int32 * 1024), ("gap_extend", C.c_int32),
                ("gap_pos1", C.c_int32), ("gap_pos2", C.c_int32), ("max_len", C.c_int32),
                ("gap_open_by_len", C.POINTER(C.c_int32))]


class CwhipError(RuntimeError):
    pass


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CwhipError("libcwhip.so not built (run __graft_entry__.build() or make -C cuda-clustalw_amd)")
    L = C.CDLL(LIB_PATH)
    P = np.ctypeslib.ndpointer
    vp = C.c_void_p
    L.cwhip_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.cwhip_destroy.argtypes = [vp]
    L.cwhip_destroy.restype = None
    L.cwhip_last_error.argtypes = [vp]
    L.cwhip_last_error.restype = C.c_char_p
    L.cwhip_pair_params_default.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, C.c_int,
                                            C.POINTER(PairParams), P(np.int32, flags="C")]
    L.cwhip_load_sequences.argtypes = [vp, P(np.uint8, flags="C"), P(np.int64, flags="C"),
                                       P(np.int32, flags="C"), C.c_int]
    L.cwhip_range_npairs.argtypes = [C.c_int] * 5
    L.cwhip_range_npairs.restype = C.c_int64
    L.cwhip_pairwise_range.argtypes = [vp, C.POINTER(PairParams), C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.c_int, vp, vp, vp]
    L.cwhip_pairwise_pairs.argtypes = [vp, C.POINTER(PairParams), P(np.int32, flags="C"),
                                       C.c_int64, C.c_int, vp, vp]
    L.cwhip_distance.argtypes = [C.c_int32, C.c_int32, C.c_int32]
    L.cwhip_distance.restype = C.c_float
    L.cwhip_stage_pairs.argtypes = [vp, C.POINTER(PairParams), P(np.int32, flags="C"), C.c_int64]
    L.cwhip_run_staged.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.cwhip_fetch_counts.argtypes = [vp, P(np.int32, flags="C"), C.c_int64]
    L.cwhip_fetch_info.argtypes = [vp, P(np.int32, flags="C"), C.c_int64]
    L.cwhip_set_tuning.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    for name in EXPORTS:
        if name not in ("cwhip_destroy", "cwhip_last_error", "cwhip_range_npairs", "cwhip_distance"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def pair_params(dna: bool = False, pw_gap_open: float | None = None,
                pw_gap_extend: float | None = None, transition_weight: float = 0.5,
                max_len: int = 20000):
    """The reference's default pairwise scoring (Gonnet 250 or IUB)."""
    L = load()
    go = (15.0 if dna else 10.0) if pw_gap_open is None else pw_gap_open
    ge = (6.66 if dna else 0.1) if pw_gap_extend is None else pw_gap_extend
    p = PairParams()
    buf = np.zeros(max_len + 1, dtype=np.int32)
    rc = L.cwhip_pair_params_default(int(dna), go, ge, transition_weight, max_len, C.byref(p), buf)
    if rc:
        raise CwhipError("cwhip_pair_params_default failed: %d" % rc)
    p._keep = buf            # the struct points into buf
    return p


class Context:
    """One device context: sequences resident in HBM plus per-wave scratch."""

    def __init__(self, device: int = 0):
        self._L = load()
        h = C.c_void_p()
        rc = self._L.cwhip_create(device, C.byref(h))
        if rc:
            raise CwhipError("cwhip_create(%d) failed: %d" % (device, rc))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.cwhip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int, what: str):
        if rc:
            msg = self._L.cwhip_last_error(self._h)
            raise CwhipError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def load_sequences(self, codes, offs, lens):
        self._codes = np.ascontiguousarray(codes, dtype=np.uint8)
        self._offs = np.ascontiguousarray(offs, dtype=np.int64)
        self._lens = np.ascontiguousarray(lens, dtype=np.int32)
        self._chk(self._L.cwhip_load_sequences(self._h, self._codes, self._offs, self._lens,
                                               len(self._lens)), "cwhip_load_sequences")

    def set_tuning(self, rmax=0, waves_per_group=0, groups_per_cu=0):
        self._chk(self._L.cwhip_set_tuning(self._h, rmax, waves_per_group, groups_per_cu),
                  "cwhip_set_tuning")

    def pairwise_pairs(self, params, pairs, kernel=KERNEL_BATCHED):
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        n = len(pairs)
        cnt = np.zeros(max(n, 1), np.int32)
        dist = np.zeros(max(n, 1), np.float32)
        self._chk(self._L.cwhip_pairwise_pairs(self._h, C.byref(params), pairs, n, kernel,
                                               cnt.ctypes.data, dist.ctypes.data),
                  "cwhip_pairwise_pairs")
        return cnt[:n], dist[:n]

    def pairwise_range(self, params, i_start, i_end, j_start, j_end, kernel=KERNEL_BATCHED):
        n = self._L.cwhip_range_npairs(len(self._lens), i_start, i_end, j_start, j_end)
        n = max(int(n), 0)
        pr = np.zeros(2 * max(n, 1), np.int32)
        cnt = np.zeros(max(n, 1), np.int32)
        dist = np.zeros(max(n, 1), np.float32)
        self._chk(self._L.cwhip_pairwise_range(self._h, C.byref(params), i_start, i_end, j_start,
                                               j_end, kernel, pr.ctypes.data, cnt.ctypes.data,
                                               dist.ctypes.data), "cwhip_pairwise_range")
        return pr[:2 * n].reshape(-1, 2), cnt[:n], dist[:n]

    # --- device-resident timing path -------------------------------------
    def stage(self, params, pairs):
        self._staged = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        self._chk(self._L.cwhip_stage_pairs(self._h, C.byref(params), self._staged,
                                            len(self._staged)), "cwhip_stage_pairs")

    def run_staged(self, kernel=KERNEL_BATCHED, reps=1):
        ms, kms = C.c_float(), C.c_float()
        self._chk(self._L.cwhip_run_staged(self._h, kernel, reps, C.byref(ms), C.byref(kms)),
                  "cwhip_run_staged")
        return ms.value, kms.value

    def fetch_counts(self):
        n = len(self._staged)
        out = np.zeros(max(n, 1), np.int32)
        self._chk(self._L.cwhip_fetch_counts(self._h, out, n), "cwhip_fetch_counts")
        return out[:n]

    def fetch_info(self):
        n = len(self._staged)
        out = np.zeros(8 * max(n, 1), np.int32)
        self._chk(self._L.cwhip_fetch_info(self._h, out, n), "cwhip_fetch_info")
        return out[:8 * n].reshape(-1, 8)


def distance(count: int, len1: int, len2: int) -> float:
    return float(load().cwhip_distance(count, len1, len2))
