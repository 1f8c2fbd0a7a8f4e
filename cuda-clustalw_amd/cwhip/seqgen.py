"""Seeded synthetic protein families for the pairwise and profile workloads.

Every workload is generated from one integer seed:

* a root sequence drawn from background amino-acid frequencies,
* a random binary tree (Yule splitting of a random leaf) with branch lengths
  ~ Exp(mean 0.1),
* along each edge: per-site substitution with probability 1 - exp(-b), the new
  residue drawn from the background; and indels that start at each site with
  probability ``indel_rate * b`` (insertion or deletion with equal odds),
  geometric lengths with mean 3.

Randomness comes only from the raw 64-bit output of numpy's PCG64 bit
generator (a stream numpy keeps stable across versions); uniforms, exponentials,
geometrics, normals and categorical draws are computed here from those words
by their textbook inverse-CDF / Box-Muller formulas, so a seed names the same
sequences on every machine with this image.

Residue codes follow ClustalW's alphabet "ABCDEFGHIKLMNOPQRSTUVWXYZ-" (A=0 ..
Z=24, '-'=31), the code assignment of UserParameters.cpp:132 and
Sequence.cpp:63-83 (CPU/clustalw-2.0.11/src).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

ALPHABET = "ABCDEFGHIKLMNOPQRSTUVWXYZ"
GAP_CODE = 31

# Background frequencies of the 20 standard amino acids (Robinson & Robinson,
# PNAS 1991), renormalised below.
_BACKGROUND = {
    "A": 0.07805, "R": 0.05129, "N": 0.04487, "D": 0.05364, "C": 0.01925,
    "Q": 0.04264, "E": 0.06295, "G": 0.07377, "H": 0.02199, "I": 0.05142,
    "L": 0.09019, "K": 0.05744, "M": 0.02243, "F": 0.03856, "P": 0.05203,
    "S": 0.07120, "T": 0.05841, "W": 0.01330, "Y": 0.03216, "V": 0.06441,
}
_BG_CODES = np.array([ALPHABET.index(a) for a in sorted(_BACKGROUND)], dtype=np.uint8)
_BG_P = np.array([_BACKGROUND[a] for a in sorted(_BACKGROUND)], dtype=np.float64)
_BG_CDF = np.cumsum(_BG_P / _BG_P.sum())
_BG_CDF[-1] = 1.0


class Rng:
    """Draws derived from PCG64's raw 64-bit stream only."""

    def __init__(self, seed: int):
        self._bg = np.random.PCG64(seed)

    def raw(self, n: int) -> np.ndarray:
        return np.asarray(self._bg.random_raw(n), dtype=np.uint64)

    def uniform(self, n: int) -> np.ndarray:
        """U[0,1) with 53-bit resolution."""
        return (self.raw(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def uniform1(self) -> float:
        return float(self.uniform(1)[0])

    def exponential(self, mean: float, n: int) -> np.ndarray:
        return -mean * np.log1p(-self.uniform(n))

    def geometric(self, mean: float, n: int) -> np.ndarray:
        """Support 1,2,...; P(k) = (1-p)^(k-1) p with p = 1/mean."""
        p = 1.0 / mean
        u = self.uniform(n)
        return 1 + np.floor(np.log1p(-u) / math.log1p(-p)).astype(np.int64)

    def normal(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.concatenate([r * np.cos(2.0 * math.pi * u2), r * np.sin(2.0 * math.pi * u2)])
        return z[:n]

    def residues(self, n: int) -> np.ndarray:
        idx = np.searchsorted(_BG_CDF, self.uniform(n), side="right")
        return _BG_CODES[np.minimum(idx, len(_BG_CODES) - 1)]

    def integers(self, lo: int, hi: int, n: int) -> np.ndarray:
        """Uniform integers in [lo, hi)."""
        return lo + np.floor(self.uniform(n) * (hi - lo)).astype(np.int64)


@dataclass
class Family:
    seqs: list          # list of np.uint8 code arrays
    names: list         # list of str

    def packed(self):
        """(codes uint8[total], offsets int64[n], lengths int32[n])."""
        lens = np.array([len(s) for s in self.seqs], dtype=np.int32)
        offs = np.zeros(len(self.seqs), dtype=np.int64)
        if len(self.seqs) > 1:
            offs[1:] = np.cumsum(lens[:-1], dtype=np.int64)
        codes = np.concatenate(self.seqs).astype(np.uint8) if self.seqs else np.zeros(0, np.uint8)
        return codes, offs, lens

    def fasta(self) -> str:
        out = []
        for nm, s in zip(self.names, self.seqs):
            txt = "".join("-" if c == GAP_CODE else ALPHABET[c] for c in s)
            out.append(">" + nm)
            out.extend(txt[k:k + 60] for k in range(0, len(txt), 60))
        return "\n".join(out) + "\n"


def _evolve(rng: Rng, parent: np.ndarray, b: float, indel_rate: float,
            indel_mean: float) -> np.ndarray:
    L = len(parent)
    child = parent.copy()
    if L:
        sub = rng.uniform(L) < (1.0 - math.exp(-b))
        k = int(sub.sum())
        if k:
            child[sub] = rng.residues(k)
        starts = np.nonzero(rng.uniform(L) < indel_rate * b)[0]
    else:
        starts = np.zeros(0, np.int64)
    if len(starts) == 0:
        return child
    kinds = rng.uniform(len(starts)) < 0.5           # True: insertion
    lens = rng.geometric(indel_mean, len(starts))
    pieces, pos = [], 0
    for st, ins, ln in zip(starts.tolist(), kinds.tolist(), lens.tolist()):
        if st < pos:                                  # inside a previous deletion
            continue
        if ins:
            pieces.append(child[pos:st + 1])
            pieces.append(rng.residues(int(ln)))
            pos = st + 1
        else:
            pieces.append(child[pos:st])
            pos = min(L, st + int(ln))
    pieces.append(child[pos:])
    return np.concatenate(pieces).astype(np.uint8)


def family(seed: int, n: int, root_len: int, branch_mean: float = 0.1,
           indel_rate: float = 0.02, indel_mean: float = 3.0,
           target_lens=None, name_prefix: str = "seq") -> Family:
    """n leaf sequences evolved along a random binary tree from one root.

    If ``target_lens`` is given (one length per leaf), leaf s is cut to a window
    of that length at a random offset (sequences shorter than the target are
    kept whole), which gives related sequences of very uneven length.
    """
    rng = Rng(seed)
    root = rng.residues(root_len)
    if n <= 0:
        return Family([], [])
    # Yule tree: nodes[k] = (parent, branch length); leaves listed in `leaves`
    parents, blen = [-1], [0.0]
    leaves = [0]
    while len(leaves) < n:
        pick = int(rng.integers(0, len(leaves), 1)[0])
        node = leaves[pick]
        kids = []
        for _ in range(2):
            parents.append(node)
            blen.append(float(rng.exponential(branch_mean, 1)[0]))
            kids.append(len(parents) - 1)
        leaves[pick:pick + 1] = kids
    seqs = [None] * len(parents)
    seqs[0] = root
    for v in range(1, len(parents)):                 # parents precede children
        seqs[v] = _evolve(rng, seqs[parents[v]], blen[v], indel_rate, indel_mean)
    out = [seqs[v] for v in leaves]
    if target_lens is not None:
        tl = np.asarray(target_lens, dtype=np.int64)
        cut = []
        for s, t in zip(out, tl.tolist()):
            if len(s) > t:
                off = int(rng.integers(0, len(s) - t + 1, 1)[0])
                s = s[off:off + t]
            cut.append(s)
        out = cut
    width = max(4, len(str(n)))
    names = ["%s%0*d" % (name_prefix, width, k + 1) for k in range(n)]
    return Family([np.ascontiguousarray(s, dtype=np.uint8) for s in out], names)


def lognormal_lengths(seed: int, n: int, mu: float, sigma: float, lo: int, hi: int) -> np.ndarray:
    rng = Rng(seed)
    z = rng.normal(n)
    return np.clip(np.round(np.exp(mu + sigma * z)), lo, hi).astype(np.int64)


# --- the named workloads -------------------------------------------------

def serving_request(seed: int, index: int, n_range=(10, 500), len_cap: int = 2000):
    """One request of the serving stream: (kind, Family); kind 1 (distances +
    progressive alignment) with probability 0.2, else 0 (distances only).
    N ~ Uniform[n_range], lengths lognormal(ln 300, 0.7) clipped to [30, len_cap],
    the sequences evolved along a seeded tree and cut to those lengths."""
    rng = Rng(seed * 1000003 + index)
    n = int(rng.integers(n_range[0], n_range[1] + 1, 1)[0])
    kind = 1 if float(rng.uniform(1)[0]) < 0.2 else 0
    tl = lognormal_lengths(seed * 7919 + index, n, math.log(300.0), 0.7, 30, len_cap)
    fam = family(seed * 31 + index, n, int(tl.max()), target_lens=tl, name_prefix="r%d_" % index)
    return kind, fam


def workload(name: str, seed: int | None = None) -> Family:
    """Named synthetic workloads.

    small97   : 100 seqs, root 97        (4,950 pairs)
    large1523 : 1000 seqs, root 1523     (499,500 pairs)
    uneven    : 2000 seqs, lengths lognormal(ln 400, 0.8) clipped [50, 3000]
    prog1002  : 1000 seqs, root 1002     (progressive-alignment workload)
    mid498    : 1000 seqs, root 498      (k-tuple distances at ~500 residues)
    """
    if name == "small97":
        return family(20260101 if seed is None else seed, 100, 97)
    if name == "large1523":
        return family(20260102 if seed is None else seed, 1000, 1523)
    if name == "uneven":
        s = 20260103 if seed is None else seed
        tl = lognormal_lengths(s + 7, 2000, math.log(400.0), 0.8, 50, 3000)
        return family(s, 2000, 3000, target_lens=tl)
    if name == "prog1002":
        return family(20260104 if seed is None else seed, 1000, 1002)
    if name == "mid498":
        return family(20260106 if seed is None else seed, 1000, 498)
    raise KeyError(name)
