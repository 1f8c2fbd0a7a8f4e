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
    "cwhip_fetch_counts", "cwhip_fetch_info", "cwhip_set_tuning", "cwhip_set_msa_tuning",
    "cwhip_get_tune_table",
    "cwhip_set_tune_table", "cwhip_autotune", "cwhip_set_graph",
    "cwhip_profile_dp", "cwhip_msa_params_default", "cwhip_msa", "cwhip_msa_rows",
    "cwhip_msa_stats", "cwhip_profile_align", "cwhip_step_profiles", "cwhip_fast_params_default",
    "cwhip_fast_pairwise_range", "cwhip_fast_last_ms", "cwhip_nj_tree", "cwhip_server_create",
    "cwhip_server_destroy", "cwhip_server_submit", "cwhip_server_wait", "cwhip_server_distances",
    "cwhip_server_text", "cwhip_server_times", "cwhip_server_release", "cwhip_server_stats", "cwhip_guide_from_tree", "cwhip_clustal_text",
    "cwhip_pairwise_noalloc", "cwhip_nj_tree_from_run", "cwhip_nj_tree_on_device",
]

PRF_W = 35


class PrfPair(C.Structure):
    _fields_ = [("prf1", C.c_void_p), ("prf2", C.c_void_p), ("len1", C.c_int32),
                ("len2", C.c_int32), ("displ", C.c_void_p), ("n_displ", C.c_int32),
                ("score", C.c_int32)]


class MsaParams(C.Structure):
    _fields_ = [("dna", C.c_int32), ("gap_open", C.c_float), ("gap_extend", C.c_float),
                ("gap_dist", C.c_int32), ("no_var_penalties", C.c_int32),
                ("no_hyd_penalties", C.c_int32), ("no_pref_penalties", C.c_int32),
                ("use_end_gaps", C.c_int32), ("end_gap_penalties", C.c_int32),
                ("transition_weight", C.c_float), ("neg_matrix", C.c_int32),
                ("no_weights", C.c_int32), ("divergence_cutoff", C.c_int32),
                ("hyd_residues", C.c_char_p), ("distance_tree", C.c_int32),
                ("profile_matrix", C.c_void_p), ("profile_matrix_user", C.c_void_p)]


# cwhip_profile_matrix_fn: (user, pcid, min_len, int32 matrix[32*32], float* scale, int32* mat_avg) -> maxRes
PROFILE_MATRIX_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_int32, C.POINTER(C.c_int32),
                                C.POINTER(C.c_float), C.POINTER(C.c_int32))


def set_profile_matrix(params, fn):
    """Install a Python profile-matrix function fn(pcid, min_len) -> (maxRes,
    int32[32,32] matrix, scale, mat_avg) as params' cwhip_profile_matrix_fn;
    the ctypes thunk is kept on params."""
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


class FastParams(C.Structure):
    _fields_ = [("dna", C.c_int32), ("ktup", C.c_int32), ("window", C.c_int32),
                ("wind_gap", C.c_int32), ("signif", C.c_int32), ("percent", C.c_int32),
                ("max_aln", C.c_int32)]


class PairParams(C.Structure):
    _fields_ = [("matrix", C.c_int32 * 1024), ("gap_extend", C.c_int32),
                ("gap_pos1", C.c_int32), ("gap_pos2", C.c_int32), ("max_len", C.c_int32),
                ("gap_open_by_len", C.POINTER(C.c_int32))]


class PairDevParams(C.Structure):
    """cwhip_pair_dev_params: the scoring in device memory."""
    _fields_ = [("matrix", C.c_void_p), ("gap_open_by_len", C.c_void_p), ("gap_extend", C.c_int32),
                ("gap_pos1", C.c_int32), ("gap_pos2", C.c_int32), ("max_len", C.c_int32)]


NOALLOC_SCRATCH_BYTES = 128 << 20
NOALLOC_MAX_PAIRS = 1 << 22
MAX_LEN = 32767                 # CWHIP_MAX_LEN: longest sequence the pairwise kernels take

TUNE_CLASSES = 6
TUNE_CLASS_EDGES = (128, 256, 512, 1024, 2048)   # tuning class = first edge >= the longer sequence


class TuneTable(C.Structure):
    _fields_ = [("rmax", C.c_int32 * TUNE_CLASSES), ("colcap", C.c_int32 * TUNE_CLASSES)]

    def as_dict(self):
        return {"rmax": list(self.rmax), "colcap": list(self.colcap)}


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
    L.cwhip_set_msa_tuning.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.cwhip_get_tune_table.argtypes = [vp, C.POINTER(TuneTable)]
    L.cwhip_set_tune_table.argtypes = [vp, C.POINTER(TuneTable)]
    L.cwhip_autotune.argtypes = [vp, C.POINTER(PairParams), P(np.int32, flags="C"), C.c_int64,
                                 C.POINTER(TuneTable), C.POINTER(C.c_float)]
    L.cwhip_set_graph.argtypes = [vp, C.c_int]
    L.cwhip_profile_dp.argtypes = [vp, vp, C.c_int, C.c_int]
    L.cwhip_msa_params_default.argtypes = [C.c_int, C.POINTER(MsaParams)]
    L.cwhip_msa.argtypes = [vp, C.POINTER(MsaParams), C.c_int, P(np.uint8, flags="C"),
                            P(np.int64, flags="C"), P(np.int32, flags="C"),
                            P(np.float64, flags="C"), P(np.int32, flags="C"),
                            P(np.int32, flags="C"), C.c_int, P(np.int32, flags="C"),
                            C.POINTER(C.c_int32)]
    L.cwhip_msa_rows.argtypes = [vp, P(np.uint8, flags="C"), C.c_int64]
    L.cwhip_profile_align.argtypes = [vp, C.POINTER(MsaParams), C.c_int, P(np.uint8, flags="C"),
                                      P(np.int64, flags="C"), P(np.int32, flags="C"),
                                      P(np.float64, flags="C"), P(np.int32, flags="C"),
                                      P(np.int32, flags="C"), P(np.int32, flags="C"), vp,
                                      C.c_int64, P(np.int32, flags="C")]
    L.cwhip_step_profiles.argtypes = [vp, C.POINTER(MsaParams), C.c_int, P(np.uint8, flags="C"),
                                      P(np.int64, flags="C"), P(np.int32, flags="C"),
                                      P(np.float64, flags="C"), P(np.int32, flags="C"),
                                      P(np.int32, flags="C"), P(np.int32, flags="C"), vp, C.c_int64,
                                      vp, C.c_int64, P(np.int32, flags="C")]
    L.cwhip_msa_stats.argtypes = [vp, P(np.float64, flags="C")]
    L.cwhip_fast_params_default.argtypes = [C.c_int, C.POINTER(FastParams)]
    L.cwhip_fast_pairwise_range.argtypes = [vp, C.POINTER(FastParams), C.c_int, C.c_int, C.c_int,
                                            C.c_int, vp, vp, C.POINTER(C.c_int64)]
    L.cwhip_fast_last_ms.restype = C.c_float
    L.cwhip_fast_last_ms.argtypes = [vp]
    L.cwhip_nj_tree.argtypes = [C.c_int, P(np.float64, flags="C"), vp, vp, C.c_int64,
                                C.POINTER(C.c_int64)]
    L.cwhip_nj_tree_from_run.argtypes = [vp, vp, vp, C.c_int64, C.POINTER(C.c_int64)]
    L.cwhip_nj_tree_on_device.argtypes = [vp, C.c_int, P(np.float64, flags="C"), vp, vp, C.c_int64,
                                          C.POINTER(C.c_int64)]
    L.cwhip_guide_from_tree.argtypes = [C.c_int, C.c_char_p, vp, P(np.int32, flags="C"),
                                        P(np.float64, flags="C"), P(np.int32, flags="C"),
                                        C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.cwhip_clustal_text.argtypes = [C.c_int, vp, P(np.uint8, flags="C"), C.c_int64,
                                     P(np.int32, flags="C"), P(np.int32, flags="C"), C.c_int,
                                     vp, C.c_int64, C.POINTER(C.c_int64)]
    L.cwhip_server_create.argtypes = [C.c_int, C.c_int, C.c_int64, C.POINTER(vp)]
    L.cwhip_server_destroy.argtypes = [vp]
    L.cwhip_server_destroy.restype = None
    L.cwhip_server_submit.argtypes = [vp, C.c_int, C.c_int, C.c_int, P(np.uint8, flags="C"),
                                      P(np.int64, flags="C"), P(np.int32, flags="C"), vp,
                                      C.POINTER(C.c_int64)]
    L.cwhip_server_wait.argtypes = [vp, C.c_int64, C.c_double]
    L.cwhip_server_distances.argtypes = [vp, C.c_int64, P(np.float32, flags="C")]
    L.cwhip_server_text.argtypes = [vp, C.c_int64, C.c_int, vp, C.c_int64, C.POINTER(C.c_int64)]
    L.cwhip_server_times.argtypes = [vp, C.c_int64, P(np.float64, flags="C")]
    L.cwhip_server_release.argtypes = [vp, C.c_int64]
    L.cwhip_server_stats.argtypes = [vp, P(np.float64, flags="C")]
    L.cwhip_pairwise_noalloc.argtypes = [C.POINTER(PairDevParams), vp, vp, vp, C.c_int, vp,
                                         C.c_int64, vp, vp, vp]
    for name in EXPORTS:
        if name not in ("cwhip_destroy", "cwhip_last_error", "cwhip_range_npairs", "cwhip_distance",
                        "cwhip_server_destroy", "cwhip_fast_last_ms"):
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

    def set_msa_tuning(self, wide_rows=-1, wide_levels=-1, wide_pairs=-1, wait_limit=-1):
        """Schedule of the profile alignments' Myers-Miller levels (-1 keeps a value)."""
        self._chk(self._L.cwhip_set_msa_tuning(self._h, wide_rows, wide_levels, wide_pairs, wait_limit),
                  "cwhip_set_msa_tuning")

    def tune_table(self):
        t = TuneTable()
        self._chk(self._L.cwhip_get_tune_table(self._h, C.byref(t)), "cwhip_get_tune_table")
        return t

    def set_tune_table(self, rmax, colcap):
        t = TuneTable()
        for c in range(TUNE_CLASSES):
            t.rmax[c] = rmax[c] if hasattr(rmax, "__len__") else rmax
            t.colcap[c] = colcap[c] if hasattr(colcap, "__len__") else colcap
        self._chk(self._L.cwhip_set_tune_table(self._h, C.byref(t)), "cwhip_set_tune_table")

    def autotune(self, params, pairs):
        """Time the batched kernel's candidate shapes per tuning class on a
        sample of `pairs`; returns (TuneTable, device ms spent)."""
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        t, ms = TuneTable(), C.c_float()
        self._chk(self._L.cwhip_autotune(self._h, C.byref(params), pairs, len(pairs), C.byref(t),
                                         C.byref(ms)), "cwhip_autotune")
        return t, ms.value

    def set_graph(self, enable):
        self._chk(self._L.cwhip_set_graph(self._h, 1 if enable else 0), "cwhip_set_graph")

    def pairwise_pairs(self, params, pairs, kernel=KERNEL_BATCHED):
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        n = len(pairs)
        cnt = np.zeros(max(n, 1), np.int32)
        dist = np.zeros(max(n, 1), np.float32)
        self._chk(self._L.cwhip_pairwise_pairs(self._h, C.byref(params), pairs, n, kernel,
                                               cnt.ctypes.data, dist.ctypes.data),
                  "cwhip_pairwise_pairs")
        return cnt[:n], dist[:n]

    def nj_tree_from_run(self, names) -> str:
        """.dnd text of the NJ guide tree built on the device from the last
        pairwise run over all pairs of the loaded sequences."""
        return _tree_text(lambda a, b, c, r: self._L.cwhip_nj_tree_from_run(self._h, a, b, c, r), names,
                          "cwhip_nj_tree_from_run", self._h)

    def nj_tree_on_device(self, dist_full, names) -> str:
        """nj_tree's text, the joins run on the device."""
        d = np.ascontiguousarray(dist_full, dtype=np.float64).reshape(-1)
        return _tree_text(lambda a, b, c, r: self._L.cwhip_nj_tree_on_device(self._h, len(names), d, a, b, c, r),
                          names, "cwhip_nj_tree_on_device", self._h)

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

    def fast_pairwise_range(self, i_start, i_end, j_start, j_end, dna=False, **kw):
        """k-tuple distances (FastPairwiseAlign) of the loaded sequences over the
        reference's 1-based range rules: (pairs int32 [k, 2], dist float64 [k])."""
        fp = fast_params(dna, **kw)
        n = len(self._lens)
        cap = max(1, sum(max(0, j_end - max(j_start + 2, i + 1) + 1)
                         for i in range(max(i_start, 0) + 1, min(i_end, n) + 1)))
        pr = np.zeros(2 * cap, np.int32)
        dist = np.zeros(cap, np.float64)
        k = C.c_int64()
        self._chk(self._L.cwhip_fast_pairwise_range(self._h, C.byref(fp), i_start, i_end, j_start,
                                                    j_end, pr.ctypes.data, dist.ctypes.data,
                                                    C.byref(k)), "cwhip_fast_pairwise_range")
        return pr[:2 * k.value].reshape(-1, 2), dist[:k.value]

    def fast_last_ms(self):
        return float(self._L.cwhip_fast_last_ms(self._h))

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


    # --- progressive alignment --------------------------------------------
    def profile_dp(self, problems, end_gap_penalties=0):
        """progDiff on a batch of (prf1, len1, prf2, len2) profile pairs:
        [(score, edit script int32[])]."""
        keep = []
        arr = (PrfPair * max(len(problems), 1))()
        for k, (p1, l1, p2, l2) in enumerate(problems):
            a = np.ascontiguousarray(p1, dtype=np.int32).reshape(-1)
            b = np.ascontiguousarray(p2, dtype=np.int32).reshape(-1)
            if a.size != (l1 + 2) * PRF_W or b.size != (l2 + 2) * PRF_W:
                raise ValueError("profile shape does not match its length")
            d = np.zeros(l1 + l2 + 4, np.int32)
            keep += [a, b, d]
            arr[k].prf1 = a.ctypes.data
            arr[k].prf2 = b.ctypes.data
            arr[k].len1 = l1
            arr[k].len2 = l2
            arr[k].displ = d.ctypes.data
        self._chk(self._L.cwhip_profile_dp(self._h, C.cast(arr, C.c_void_p), len(problems),
                                           int(end_gap_penalties)), "cwhip_profile_dp")
        out = []
        for k in range(len(problems)):
            d = keep[3 * k + 2]
            out.append((int(arr[k].score), d[:arr[k].n_displ].copy()))
        return out

    def msa(self, seqs, sim, weights, steps, params=None, dna=False):
        """multiSeqAlign: (aligned rows as uint8 arrays, 1-based output order, stats)."""
        if params is None:
            params = msa_params(dna)
        n = len(seqs)
        codes = np.concatenate([np.asarray(s, np.uint8) for s in seqs]) if n else np.zeros(0, np.uint8)
        lens = np.array([len(s) for s in seqs], np.int32)
        offs = np.zeros(n, np.int64)
        if n > 1:
            offs[1:] = np.cumsum(lens[:-1])
        order = np.zeros(max(n, 1), np.int32)
        alen = C.c_int32()
        steps = np.ascontiguousarray(steps, np.int32).reshape(-1)
        nsteps = steps.size // (n + 1) if n else 0
        self._chk(self._L.cwhip_msa(self._h, C.byref(params), n, np.ascontiguousarray(codes),
                                    offs, lens, np.ascontiguousarray(sim, np.float64).reshape(-1),
                                    np.ascontiguousarray(weights, np.int32), steps, nsteps, order,
                                    C.byref(alen)), "cwhip_msa")
        L = alen.value
        rows = np.zeros((n, max(L, 1)), np.uint8)
        self._chk(self._L.cwhip_msa_rows(self._h, rows, rows.shape[1]), "cwhip_msa_rows")
        st = np.zeros(6, np.float64)
        self._chk(self._L.cwhip_msa_stats(self._h, st), "cwhip_msa_stats")
        stats = dict(calls=int(st[0]), batches=int(st[1]), cells=float(st[2]),
                     host_ms=float(st[3]), device_ms=float(st[4]), prf_entries=float(st[5]))
        return [rows[i, :L].copy() for i in range(n)], order[:n].copy(), stats


def _names(names):
    arr = (C.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
    return arr


def _tree_text(call, names, what, ctx=None) -> str:
    """Runs call(names_array, buf, cap, len_ref) with room for every name and
    two branch lengths per node (the tree is built once); a second call only if
    the text did not fit."""
    n = len(names)
    arr = _names(names)
    ln = C.c_int64()
    cap = sum(len(x.encode()) for x in names) + 80 * (n + 2) + 64
    buf = C.create_string_buffer(cap)
    rc = call(C.cast(arr, C.c_void_p), C.cast(buf, C.c_void_p), cap, C.byref(ln))
    if rc and ln.value + 1 > cap:
        buf = C.create_string_buffer(ln.value + 1)
        rc = call(C.cast(arr, C.c_void_p), C.cast(buf, C.c_void_p), ln.value + 1, C.byref(ln))
    if rc:
        msg = load().cwhip_last_error(ctx).decode() if ctx else ""
        raise CwhipError("%s failed: %d %s" % (what, rc, msg))
    return buf.value.decode()


def nj_tree(dist_full, names) -> str:
    """.dnd text of the NJ guide tree of an (n+1)x(n+1) 1-based distance matrix."""
    L = load()
    d = np.ascontiguousarray(dist_full, dtype=np.float64).reshape(-1)
    return _tree_text(lambda a, b, c, r: L.cwhip_nj_tree(len(names), d, a, b, c, r), names, "cwhip_nj_tree")


def guide_from_tree(dnd: str, names, with_distance_tree: bool = False):
    """(weights int32[n], sim float64[n+1, n+1], steps int32[nsteps, n+1]), and
    UserParameters::distanceTree as readTree leaves it when with_distance_tree."""
    L = load()
    n = len(names)
    w = np.zeros(max(n, 1), np.int32)
    sim = np.zeros((n + 1) * (n + 1), np.float64)
    steps = np.zeros(max(n - 1, 1) * (n + 1), np.int32)
    ns = C.c_int32()
    arr = _names(names)
    dt = C.c_int32()
    rc = L.cwhip_guide_from_tree(n, dnd.encode(), C.cast(arr, C.c_void_p), w, sim, steps,
                                 C.byref(ns), C.byref(dt))
    if rc:
        raise CwhipError("cwhip_guide_from_tree failed: %d" % rc)
    out = (w[:n].copy(), sim.reshape(n + 1, n + 1), steps[:ns.value * (n + 1)].reshape(-1, n + 1))
    return out + (dt.value,) if with_distance_tree else out


def clustal_text(names, rows, order, dna=False) -> str:
    """CLUSTAL (.aln) text of aligned rows (uint8 code arrays) in 1-based order."""
    L = load()
    n = len(names)
    lens = np.array([len(r) for r in rows], np.int32)
    stride = max(int(lens.max()) if n else 1, 1)
    buf = np.zeros((n, stride), np.uint8)
    for i, r in enumerate(rows):
        buf[i, :len(r)] = r
    order = np.ascontiguousarray(order, np.int32)
    arr = _names(names)
    ln = C.c_int64()
    # room for the text in one call: lines of at most 30 + 6 name columns and 60
    # residues, at least 50 residues per block
    blocks = -(-stride // 50)
    cap = 64 + blocks * (1 + (n + 1) * 98)
    out = C.create_string_buffer(cap)
    rc = L.cwhip_clustal_text(n, C.cast(arr, C.c_void_p), buf, stride, lens, order, int(dna),
                              C.cast(out, C.c_void_p), cap, C.byref(ln))
    if rc and ln.value + 1 > cap:
        out = C.create_string_buffer(ln.value + 1)
        rc = L.cwhip_clustal_text(n, C.cast(arr, C.c_void_p), buf, stride, lens, order, int(dna),
                                  C.cast(out, C.c_void_p), ln.value + 1, C.byref(ln))
    if rc:
        raise CwhipError("cwhip_clustal_text failed: %d" % rc)
    return out.raw[:ln.value].decode()


def profile_align(ctx, seqs, sim, weights, group, aligned, params):
    """One profileAlign step: (info[4], group after the call, new rows by sequence
    index for the members, None for the others)."""
    n = len(seqs)
    lens = np.array([len(s) for s in seqs], np.int32)
    offs = np.zeros(n, np.int64)
    if n > 1:
        offs[1:] = np.cumsum(lens[:-1])
    codes = np.concatenate([np.asarray(s, np.uint8) for s in seqs]) if n else np.zeros(0, np.uint8)
    stride = int(lens.sum()) + 2
    rows = np.zeros((n, stride), np.uint8)
    g = np.ascontiguousarray(group, np.int32).copy()
    info = np.zeros(4, np.int32)
    ctx._chk(ctx._L.cwhip_profile_align(ctx._h, C.byref(params), n, codes, offs, lens,
                                        np.ascontiguousarray(sim, np.float64).reshape(-1),
                                        np.ascontiguousarray(weights, np.int32), g,
                                        np.ascontiguousarray(aligned, np.int32),
                                        rows.ctypes.data, stride, info), "cwhip_profile_align")
    out = [rows[i, :info[1]].copy() if g[i + 1] in (1, 2) and info[1] > 0 else None for i in range(n)]
    return info, g, out


def step_profiles(ctx, seqs, sim, weights, group, aligned, params):
    """The two int32 profiles ((len+2) x PRF_W) the device builds for one
    profileAlign step, as (prf1, len1, prf2, len2), or None when a group is empty."""
    n = len(seqs)
    lens = np.array([len(s) for s in seqs], np.int32)
    offs = np.zeros(n, np.int64)
    if n > 1:
        offs[1:] = np.cumsum(lens[:-1])
    codes = np.concatenate([np.asarray(s, np.uint8) for s in seqs]) if n else np.zeros(0, np.uint8)
    cap = (int(lens.max(initial=0)) + 2) * PRF_W
    p1 = np.zeros(cap, np.int32)
    p2 = np.zeros(cap, np.int32)
    l12 = np.zeros(2, np.int32)
    ctx._chk(ctx._L.cwhip_step_profiles(ctx._h, C.byref(params), n, codes, offs, lens,
                                        np.ascontiguousarray(sim, np.float64).reshape(-1),
                                        np.ascontiguousarray(weights, np.int32),
                                        np.ascontiguousarray(group, np.int32),
                                        np.ascontiguousarray(aligned, np.int32),
                                        p1.ctypes.data, cap, p2.ctypes.data, cap, l12),
             "cwhip_step_profiles")
    l1, l2 = int(l12[0]), int(l12[1])
    if l1 < 0:
        return None
    return (p1[:(l1 + 2) * PRF_W].reshape(-1, PRF_W), l1, p2[:(l2 + 2) * PRF_W].reshape(-1, PRF_W), l2)


JOB_DISTANCES = 0
JOB_ALIGNMENT = 1


class Server:
    """The library's request loop: jobs are admitted by device footprint into
    `slots` concurrent device contexts (see include/cwhip.h)."""

    def __init__(self, device: int = 0, slots: int = 4, hbm_budget: int = 64 << 30):
        self._L = load()
        h = C.c_void_p()
        rc = self._L.cwhip_server_create(device, slots, hbm_budget, C.byref(h))
        if rc:
            raise CwhipError("cwhip_server_create failed: %d" % rc)
        self._h = h
        self._n = {}

    def close(self):
        if getattr(self, "_h", None):
            self._L.cwhip_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, kind, seqs, names=None, dna=False) -> int:
        n = len(seqs)
        lens = np.array([len(x) for x in seqs], np.int32)
        offs = np.zeros(n, np.int64)
        if n > 1:
            offs[1:] = np.cumsum(lens[:-1])
        codes = np.ascontiguousarray(np.concatenate([np.asarray(x, np.uint8) for x in seqs]))
        arr = _names(names) if names is not None else None
        job = C.c_int64()
        rc = self._L.cwhip_server_submit(self._h, kind, int(dna), n, codes, offs, lens,
                                         C.cast(arr, C.c_void_p) if arr is not None else None,
                                         C.byref(job))
        if rc:
            raise CwhipError("cwhip_server_submit failed: %d" % rc)
        self._n[job.value] = n
        return job.value

    def wait(self, job: int, timeout: float = -1.0) -> bool:
        rc = self._L.cwhip_server_wait(self._h, job, timeout)
        if rc == -7:
            return False
        if rc:
            try:
                why = self.text(job, 2)
            except CwhipError:
                why = ""
            raise CwhipError("job %d failed: %d %s" % (job, rc, why))
        return True

    def distances(self, job: int) -> np.ndarray:
        n = self._n[job]
        d = np.zeros(max(n * (n - 1) // 2, 1), np.float32)
        rc = self._L.cwhip_server_distances(self._h, job, d)
        if rc:
            raise CwhipError("cwhip_server_distances failed: %d" % rc)
        return d[:n * (n - 1) // 2]

    def text(self, job: int, which: int) -> str:
        ln = C.c_int64()
        self._L.cwhip_server_text(self._h, job, which, None, 0, C.byref(ln))
        buf = C.create_string_buffer(ln.value + 1)
        rc = self._L.cwhip_server_text(self._h, job, which, C.cast(buf, C.c_void_p), ln.value + 1,
                                       C.byref(ln))
        if rc:
            raise CwhipError("cwhip_server_text failed: %d" % rc)
        return buf.value.decode()

    def times(self, job: int):
        t = np.zeros(4, np.float64)
        rc = self._L.cwhip_server_times(self._h, job, t)
        if rc:
            raise CwhipError("cwhip_server_times failed: %d" % rc)
        return tuple(float(x) for x in t)

    def release(self, job: int):
        self._L.cwhip_server_release(self._h, job)
        self._n.pop(job, None)

    def stats(self) -> dict:
        st = np.zeros(6, np.float64)
        self._L.cwhip_server_stats(self._h, st)
        return dict(done=int(st[0]), cells=float(st[1]), peak_bytes=int(st[2]),
                    max_running=int(st[3]), budget=int(st[4]), slots=int(st[5]))


def msa_params(dna: bool = False, **kw) -> MsaParams:
    """The reference's multiple-alignment defaults (UserParameters.cpp:33-175)."""
    p = MsaParams()
    rc = load().cwhip_msa_params_default(1 if dna else 0, C.byref(p))
    if rc:
        raise CwhipError("cwhip_msa_params_default failed: %d" % rc)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def fast_params(dna: bool = False, **kw) -> FastParams:
    """The reference's k-tuple defaults (UserParameters.cpp:86-94)."""
    p = FastParams()
    rc = load().cwhip_fast_params_default(1 if dna else 0, C.byref(p))
    if rc:
        raise CwhipError("cwhip_fast_params_default failed: %d" % rc)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def distance(count: int, len1: int, len2: int) -> float:
    return float(load().cwhip_distance(count, len1, len2))
