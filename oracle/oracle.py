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


class MsaParams(C.Structure):
    _fields_ = [("dna", C.c_int), ("gap_open", C.c_float), ("gap_extend", C.c_float),
                ("gap_dist", C.c_int), ("no_var_penalties", C.c_int),
                ("no_hyd_penalties", C.c_int), ("no_pref_penalties", C.c_int),
                ("use_end_gaps", C.c_int), ("end_gap_penalties", C.c_int),
                ("transition_weight", C.c_float), ("neg_matrix", C.c_int),
                ("no_weights", C.c_int), ("divergence_cutoff", C.c_int),
                ("hyd_residues", C.c_char_p), ("distance_tree", C.c_int),
                ("profile_matrix", C.c_void_p), ("profile_matrix_user", C.c_void_p)]


PROFILE_MATRIX_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_int),
                                C.POINTER(C.c_float), C.POINTER(C.c_int))


def set_profile_matrix(params, fn):
    """Use fn(pcid, min_len) -> (maxRes, int32[32,32], scale, mat_avg) as the
    params' matrix series (a stand-in for SubMatrix::getProfileAlignMatrix)."""
    def thunk(user, pcid, min_len, mat, scale, avg):
        r, m, sc, a = fn(pcid, min_len)
        m = np.ascontiguousarray(m, np.int32).reshape(-1)
        for k in range(32 * 32):
            mat[k] = int(m[k])
        scale[0] = sc
        avg[0] = a
        return int(r)
    cb = PROFILE_MATRIX_FN(thunk)
    params._matrix_cb = cb
    params.profile_matrix = C.cast(cb, C.c_void_p)
    params.profile_matrix_user = None
    return params


class DpStats(C.Structure):
    _fields_ = [("fwd_cells", C.c_int64), ("rev_cells", C.c_int64),
                ("leaf_cells", C.c_int64), ("nodes", C.c_int64)]


class StepInfo(C.Structure):
    _fields_ = [("step", C.c_int), ("score", C.c_int), ("prf_len1", C.c_int),
                ("prf_len2", C.c_int), ("nseqs1", C.c_int), ("nseqs2", C.c_int),
                ("aln_len", C.c_int), ("n_displ", C.c_int), ("pcid", C.c_double),
                ("gapcoef", C.c_int), ("lencoef", C.c_int), ("mat_avg", C.c_int),
                ("stats", DpStats)]


PRF_STRIDE = 35


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
        L.cwo_msa_params_default.argtypes = [C.c_int, C.POINTER(MsaParams)]
        L.cwo_guide_from_dnd.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p),
                                         P(np.int32, flags="C"), P(np.float64, flags="C"),
                                         C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.c_int)]
        L.cwo_prog_diff.argtypes = [P(np.int32, flags="C"), C.c_int, P(np.int32, flags="C"),
                                    C.c_int, C.c_int, P(np.int32, flags="C"),
                                    C.POINTER(C.c_int), C.POINTER(DpStats)]
        L.cwo_prog_diff_trace.argtypes = [P(np.int32, flags="C"), C.c_int, P(np.int32, flags="C"),
                                          C.c_int, C.c_int, P(np.int32, flags="C"),
                                          C.POINTER(C.c_int), C.POINTER(DpStats),
                                          P(np.int32, flags="C"), C.c_int, C.POINTER(C.c_int)]
        L.cwo_step_profiles.argtypes = [C.POINTER(MsaParams), C.c_int, C.c_void_p,
                                        P(np.int32, flags="C"), P(np.float64, flags="C"),
                                        P(np.int32, flags="C"), P(np.int32, flags="C"),
                                        P(np.int32, flags="C"),
                                        C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.c_int),
                                        C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.c_int)]
        L.cwo_msa.argtypes = [C.POINTER(MsaParams), C.c_int, C.c_void_p,
                              P(np.int32, flags="C"), P(np.float64, flags="C"),
                              P(np.int32, flags="C"), P(np.int32, flags="C"), C.c_int,
                              P(np.int32, flags="C"), C.c_void_p, C.POINTER(C.c_int)]
        L.cwo_profile_matrix.argtypes = [C.c_int, C.c_double, C.c_int, C.c_float, C.c_int,
                                         C.c_int, P(np.int32, flags="C"),
                                         C.POINTER(C.c_float), C.POINTER(C.c_int)]
        L.cwo_pairwise_list.argtypes = [P(np.uint8, flags="C"), P(np.int64, flags="C"),
                                        P(np.int32, flags="C"), P(np.int32, flags="C"),
                                        C.c_int64, P(np.int32, flags="C"), P(np.int32, flags="C"),
                                        C.c_int, C.c_int, C.c_int, C.c_int,
                                        P(np.int32, flags="C"), P(np.float32, flags="C"),
                                        C.c_void_p]
        L.cwo_fast_pairwise.restype = C.c_long
        L.cwo_fast_pairwise.argtypes = [C.c_int, C.c_void_p, P(np.int32, flags="C")] + \
            [C.c_int] * 11 + [C.c_void_p, C.c_void_p]
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


# ---------------------------------------------------------------------------
# Progressive alignment (cw_guide.c, cw_msa.c)

def msa_params(dna: bool = False, **kw) -> MsaParams:
    p = MsaParams()
    lib().cwo_msa_params_default(1 if dna else 0, C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def guide_from_dnd(dnd: str, names):
    """(weights int32[n], sim float64[(n+1),(n+1)], steps int32[nsteps, n+1]) from
    the .dnd text as TreeInterface::getWeightsAndStepsUseOldGuideTreeNJ derives them."""
    n = len(names)
    w = np.zeros(max(n, 1), np.int32)
    sim = np.zeros((n + 1) * (n + 1), np.float64)
    arr = (C.c_char_p * max(n, 1))(*[nm.encode() for nm in names])
    sp = C.POINTER(C.c_int)()
    ns = C.c_int()
    st = lib().cwo_guide_from_dnd(dnd.encode(), n, arr, w, sim, C.byref(sp), C.byref(ns))
    if st != 1:
        raise ValueError("guide tree not usable (code %d)" % st)
    steps = np.ctypeslib.as_array(sp, shape=(ns.value * (n + 1),)).copy().reshape(ns.value, n + 1)
    lib().cwo_free(C.cast(sp, C.c_void_p))
    return w[:n].copy(), sim.reshape(n + 1, n + 1), steps


class _SeqSet:
    """malloc-free holder of the sequences as an int** the C side may replace."""

    def __init__(self, seqs):
        libc = C.CDLL(None)
        self.malloc = libc.malloc
        self.malloc.restype = C.c_void_p
        self.malloc.argtypes = [C.c_size_t]
        self.free = libc.free
        self.free.argtypes = [C.c_void_p]
        n = len(seqs)
        self.n = n
        self.ptrs = (C.c_void_p * max(n, 1))()
        self.lens = np.zeros(max(n, 1), np.int32)
        for i, s in enumerate(seqs):
            a = np.ascontiguousarray(s, dtype=np.int32)
            p = self.malloc(max(1, a.nbytes))
            C.memmove(p, a.ctypes.data, a.nbytes)
            self.ptrs[i] = p
            self.lens[i] = len(a)

    def rows(self):
        out = []
        for i in range(self.n):
            L = int(self.lens[i])
            buf = (C.c_int * max(L, 1)).from_address(self.ptrs[i])
            out.append(np.ctypeslib.as_array(buf)[:L].copy())
        return out

    def close(self):
        for i in range(self.n):
            if self.ptrs[i]:
                self.free(self.ptrs[i])
                self.ptrs[i] = None


def msa(seqs, sim, weights, steps, params: MsaParams):
    """multiSeqAlign: (aligned rows as int32 arrays, 1-based output order, step infos)."""
    n = len(seqs)
    ss = _SeqSet(seqs)
    try:
        order = np.zeros(max(n, 1), np.int32)
        info = (StepInfo * (len(steps) + n + 1))()
        ninfo = C.c_int()
        r = lib().cwo_msa(C.byref(params), n, C.cast(ss.ptrs, C.c_void_p), ss.lens,
                          np.ascontiguousarray(sim, np.float64).reshape(-1),
                          np.ascontiguousarray(weights, np.int32).copy(),
                          np.ascontiguousarray(steps, np.int32).reshape(-1), len(steps),
                          order, C.cast(info, C.c_void_p), C.byref(ninfo))
        if r < 0:
            raise RuntimeError("multiSeqAlign failed")
        return ss.rows(), order[:n].copy(), list(info[:ninfo.value])
    finally:
        ss.close()


def step_profiles(seqs, sim, weights, group, aligned, params: MsaParams):
    """The two int32 profiles ((len+2) x 35) profileAlign builds for one step, or None."""
    n = len(seqs)
    ss = _SeqSet(seqs)
    try:
        p1, p2 = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
        l1, l2 = C.c_int(), C.c_int()
        r = lib().cwo_step_profiles(C.byref(params), n, C.cast(ss.ptrs, C.c_void_p), ss.lens,
                                    np.ascontiguousarray(sim, np.float64).reshape(-1),
                                    np.ascontiguousarray(weights, np.int32),
                                    np.ascontiguousarray(group, np.int32),
                                    np.ascontiguousarray(aligned, np.int32),
                                    C.byref(p1), C.byref(l1), C.byref(p2), C.byref(l2))
        if r == 0:
            return None
        if r < 0:
            raise RuntimeError("the profile matrix failed")
        a = np.ctypeslib.as_array(p1, shape=((l1.value + 2) * PRF_STRIDE,)).copy()
        b = np.ctypeslib.as_array(p2, shape=((l2.value + 2) * PRF_STRIDE,)).copy()
        lib().cwo_free(C.cast(p1, C.c_void_p))
        lib().cwo_free(C.cast(p2, C.c_void_p))
        return a.reshape(-1, PRF_STRIDE), l1.value, b.reshape(-1, PRF_STRIDE), l2.value
    finally:
        ss.close()


def prog_diff(prf1, len1, prf2, len2, end_gap_penalties=0):
    """(score, edit script int32[], DpStats) of progDiff on two profiles."""
    a = np.ascontiguousarray(prf1, np.int32).reshape(-1)
    b = np.ascontiguousarray(prf2, np.int32).reshape(-1)
    displ = np.zeros(len1 + len2 + 4, np.int32)
    nd = C.c_int()
    st = DpStats()
    score = lib().cwo_prog_diff(a, len1, b, len2, end_gap_penalties, displ, C.byref(nd),
                                C.byref(st))
    return score, displ[:nd.value].copy(), st


def profile_matrix(dna: bool, pcid: float, min_len: int, neg_matrix=0, distance_tree=1):
    m = np.zeros((NUMRES, NUMRES), np.int32)
    sc = C.c_float()
    avg = C.c_int()
    r = lib().cwo_profile_matrix(1 if dna else 0, pcid, min_len, 0.5, neg_matrix, distance_tree,
                                 m, C.byref(sc), C.byref(avg))
    return r, m, sc.value, avg.value


def prog_diff_trace(prf1, len1, prf2, len2, end_gap_penalties=0):
    """progDiff's calls in depth-first order: int32[ncalls, 8] records
    (A, B, M, N, go1 | go2 << 1, type, midj, value)."""
    a = np.ascontiguousarray(prf1, np.int32).reshape(-1)
    b = np.ascontiguousarray(prf2, np.int32).reshape(-1)
    displ = np.zeros(len1 + len2 + 4, np.int32)
    nd = C.c_int()
    st = DpStats()
    cap = 4 * (len1 + len2) + 16
    tr = np.zeros(8 * cap, np.int32)
    nt = C.c_int()
    lib().cwo_prog_diff_trace(a, len1, b, len2, end_gap_penalties, displ, C.byref(nd),
                              C.byref(st), tr, cap, C.byref(nt))
    return tr[:8 * min(nt.value, cap)].reshape(-1, 8)


def fast_params(dna: bool = False, **kw):
    """The reference's k-tuple defaults (UserParameters.cpp:86-94): protein
    ktup 1, window 5, window gap 3, 5 top diagonals; DNA 2, 4, 5, 4; percent."""
    p = dict(ktup=2, window=4, wind_gap=5, signif=4) if dna else \
        dict(ktup=1, window=5, wind_gap=3, signif=5)
    p["percent"] = 1
    p.update(kw)
    return p


def fast_pairwise(seqs, dna=False, ranges=None, max_aln=None, **kw):
    """k-tuple distances (FastPairwiseAlign) over the range (iStart, iEnd,
    jStart, jEnd), default the full matrix: (pairs int32 [k,2], dist float64 [k])."""
    n = len(seqs)
    fp = fast_params(dna, **kw)
    iS, iE, jS, jE = ranges if ranges is not None else (0, n, 0, n)
    lens = np.array([len(s) for s in seqs], np.int32)
    if max_aln is None:
        max_aln = 2 * int(lens.max(initial=0))
    cap = max(1, sum(max(0, jE - max(jS + 2, i + 1) + 1) for i in range(iS + 1, iE + 1)))
    pairs = np.zeros((cap, 2), np.int32)
    dist = np.zeros(cap, np.float64)
    ss = _SeqSet(seqs)
    try:
        k = lib().cwo_fast_pairwise(n, C.cast(ss.ptrs, C.c_void_p), ss.lens, 1 if dna else 0,
                                    fp["ktup"], fp["window"], fp["wind_gap"], fp["signif"],
                                    fp["percent"], max_aln, iS, iE, jS, jE,
                                    pairs.ctypes.data, dist.ctypes.data)
    finally:
        ss.close()
    if k < 0:
        raise ValueError("bad range")
    return pairs[:k].copy(), dist[:k].copy()
