"""ctypes front end of the CPU restatement (oracle/cw_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker, never as the thing
measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libcworacle.so")
NUMRES = 32


class PairResult(C.Structure):
    _fields_ = [("max_score", C.c_int), ("se1", C.c_int), ("se2", C.c_int),
                ("sb1", C.c_int), ("sb2", C.c_int), ("mm_score", C.c_int),
                ("count", C.c_int), ("n_displ", C.c_int), ("count_leaves", C.c_int),
                ("cells_fwd", C.c_int64), ("cells_rev", C.c_int64), ("cells_mm", C.c_int64),
                ("mm_nodes", C.c_int64)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P = np.ctypeslib.ndpointer
        L.cwo_pairwise_matrix.argtypes = [C.c_int, C.c_float, P(np.int32, flags="C"),
                                          C.POINTER(C.c_int), C.POINTER(C.c_int),
                                          C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.cwo_gap_params.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                     C.c_float, C.c_float, C.c_int, C.c_int,
                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.cwo_distance.restype = C.c_float
        L.cwo_distance.argtypes = [C.c_int, C.c_int, C.c_int]
        L.cwo_align_pair.argtypes = [P(np.int32, flags="C"), C.c_int, P(np.int32, flags="C"),
                                     C.c_int, P(np.int32, flags="C"), C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_void_p, C.POINTER(PairResult)]
        L.cwo_count_pairs.restype = C.c_int64
        L.cwo_count_pairs.argtypes = [C.c_int] * 5
        L.cwo_nj_dnd.restype = C.c_void_p
        L.cwo_nj_dnd.argtypes = [P(np.float64, flags="C"), C.c_int, C.POINTER(C.c_char_p)]
        L.cwo_free.argtypes = [C.c_void_p]
        L.cwo_pairwise_list.argtypes = [P(np.uint8, flags="C"), P(np.int64, flags="C"),
                                        P(np.int32, flags="C"), P(np.int32, flags="C"),
                                        C.c_int64, P(np.int32, flags="C"), P(np.int32, flags="C"),
                                        C.c_int, C.c_int, C.c_int, C.c_int,
                                        P(np.int32, flags="C"), P(np.float32, flags="C"),
                                        C.c_void_p]
        _lib = L
    return _lib


class Scoring:
    """Pairwise scoring as the reference sets it up (protein: Gonnet 250,
    pw gap open 10.0 / extend 0.1; DNA: IUB, 15.0 / 6.66)."""

    def __init__(self, dna: bool = False, pw_go: float = None, pw_ge: float = None,
                 transition_weight: float = 0.5):
        L = lib()
        self.dna = dna
        self.mat = np.zeros((NUMRES, NUMRES), dtype=np.int32)
        a, s = C.c_int(), C.c_int()
        gs, es = C.c_float(), C.c_float()
        L.cwo_pairwise_matrix(1 if dna else 0, transition_weight, self.mat, C.byref(a),
                              C.byref(s), C.byref(gs), C.byref(es))
        self.mat_avg, self.int_scale = a.value, s.value
        self.go_scale, self.ge_scale = gs.value, es.value
        self.pw_go = (15.0 if dna else 10.0) if pw_go is None else pw_go
        self.pw_ge = (6.66 if dna else 0.1) if pw_ge is None else pw_ge

    def gaps(self, n: int, m: int):
        go, ge = C.c_int(), C.c_int()
        lib().cwo_gap_params(int(self.dna), self.mat_avg, self.int_scale, self.go_scale,
                             self.ge_scale, self.pw_go, self.pw_ge, n, m, C.byref(go), C.byref(ge))
        return go.value, ge.value

    def gap_table(self, maxlen: int) -> np.ndarray:
        t = np.zeros(maxlen + 1, dtype=np.int32)
        for L in range(1, maxlen + 1):
            t[L] = self.gaps(L, L)[0]
        return t


def align_pair(s1: np.ndarray, s2: np.ndarray, sc: Scoring, want_displ: bool = False):
    """One pair through forward/reverse/diff/trace; s1, s2 are 0-based code arrays."""
    n, m = len(s1), len(s2)
    a = np.zeros(n + 1, np.int32); a[1:] = s1
    b = np.zeros(m + 1, np.int32); b[1:] = s2
    go, ge = sc.gaps(n, m)
    r = PairResult()
    displ = np.zeros(2 * (n + m) + 2, np.int32) if want_displ else None
    lib().cwo_align_pair(a, n, b, m, sc.mat, go, ge, 30, 31,
                         displ.ctypes.data if want_displ else None, C.byref(r))
    if want_displ:
        return r, displ[1:r.n_displ + 1].copy()
    return r


def pairwise_pairs(codes, offs, lens, pairs, sc: Scoring, nthreads: int = 8, want_results=False):
    """Counts and float distances for an explicit (npairs, 2) pair list."""
    pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
    npairs = len(pairs)
    maxlen = int(lens.max()) if len(lens) else 1
    cnt = np.zeros(npairs, np.int32)
    dist = np.zeros(npairs, np.float32)
    res = (PairResult * max(npairs, 1))() if want_results else None
    lib().cwo_pairwise_list(np.ascontiguousarray(codes, np.uint8),
                            np.ascontiguousarray(offs, np.int64),
                            np.ascontiguousarray(lens, np.int32),
                            pairs, npairs, sc.mat, sc.gap_table(max(maxlen, 1)),
                            sc.gaps(1, 1)[1], 30, 31, nthreads, cnt, dist,
                            C.cast(res, C.c_void_p) if want_results else None)
    if want_results:
        return cnt, dist, res
    return cnt, dist


def range_pairs(nseq: int, i_start: int, i_end: int, j_start: int, j_end: int) -> np.ndarray:
    """Pairs in the order FullPairwiseAlign.cpp:97,110 visits them."""
    out = []
    for si in range(max(0, i_start), min(nseq, i_end)):
        for sj in range(max(si + 1, j_start + 1), min(nseq, j_end)):
            out.append((si, sj))
    return np.array(out, dtype=np.int32).reshape(-1, 2)


def nj_dnd(dist_full: np.ndarray, names) -> str:
    """Guide-tree text (.dnd) of NJTree + ClusterTreeOutput for an (n+1)x(n+1)
    1-based distance matrix (row/column 0 unused)."""
    n = len(names)
    d = np.ascontiguousarray(dist_full, dtype=np.float64).copy()
    assert d.shape == (n + 1, n + 1)
    arr = (C.c_char_p * (n + 1))(*([b""] + [nm.encode() for nm in names]))
    p = lib().cwo_nj_dnd(d, n, arr)
    txt = C.string_at(p).decode()
    lib().cwo_free(p)
    return txt


def distance_matrix(n: int, pairs: np.ndarray, dist: np.ndarray) -> np.ndarray:
    """(n+1)x(n+1) symmetric matrix from per-pair float distances (SetAt twice)."""
    D = np.zeros((n + 1, n + 1), dtype=np.float64)
    D[pairs[:, 0] + 1, pairs[:, 1] + 1] = dist.astype(np.float64)
    D[pairs[:, 1] + 1, pairs[:, 0] + 1] = dist.astype(np.float64)
    return D
