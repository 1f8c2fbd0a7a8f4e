// cw_fast_kernels.hip -- gfx950 kernels of ClustalW's k-tuple distances
// (pairwise/FastPairwiseAlign.cpp, the -QUICKTREE path).
//
// The reference's pairAlign is a sequential walk per pair: k-tuple matches
// counted per diagonal, a quicksort of the diagonal counts whose tie order
// decides which diagonals are "top", then a fragment list kept in score order
// that every match on a flagged diagonal walks.  Pairs are independent.
// cw_fast_wave (below) runs one pair per wave with its arrays in LDS; this
// first kernel, cw_fast_pairs, is for pairs whose arrays do not fit there: one
// lane runs one pair from start to end and the grid's lanes take pairs from a
// counter.  A lane's arrays (positions per k-tuple, diagonals, fragments) lie
// contiguously in HBM, so its sequential scans (clearing, the partition scans
// of the sort) stay within cache lines; the data-dependent walks are
// latency-bound.  All arithmetic is integer except the final double distance,
// which is IEEE on both sides.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cw_fast_kernels.h"

namespace {

// FastPairwiseAlign.cpp:147-168: protein codes A, C, G, T, U to DNA 0..3;
// every other code breaks a k-tuple
__device__ __forceinline__ int dna_of(int c)
{
    switch (c) {
    case 0: return 0;
    case 2: return 1;
    case 6: return 2;
    case 18: case 19: return 3;
    default: return -1;
    }
}

struct Lane {
    int32_t* sc;     // this lane's rows
    __device__ __forceinline__ int32_t& at(int64_t arr, int64_t k) const { return sc[arr + k]; }
};

// residue code of 1-based position p of a sequence (0-based packed codes)
__device__ __forceinline__ int res_at(const FastArgs& a, const uint8_t* s, int p)
{
    const int c = s[p - 1];
    if (a.dna) return dna_of(c);
    return c;
}

// makePPtrs / makeNPtrs (:404-521)
__device__ void make_ptrs(const FastArgs& a, const Lane& ln, int64_t tptr, int64_t pl, const uint8_t* s,
                          int length)
{
    const int mr = a.dna ? 3 : 24;
    for (int i = 1; i <= a.limit; ++i) ln.at(pl, i) = 0;
    for (int i = 1; i <= length; ++i) ln.at(tptr, i) = 0;
    for (int i = 1; i <= length - a.ktup + 1; ++i) {
        int code = 0, pot = 1;
        bool bad = false;
        for (int j = 1; j <= a.ktup; ++j) {
            const int r = res_at(a, s, i + j - 1);
            if (r < 0 || r > mr) {
                bad = true;
                break;
            }
            code += r * pot;
            pot *= mr + 1;
        }
        if (bad) continue;
        ++code;
        const int prev = ln.at(pl, code);
        if (prev != 0) ln.at(tptr, i) = prev;
        ln.at(pl, code) = i;
    }
}

// desQuickSort (:584-659) over elements 1 .. n-1 (element n stays in place).
// Only positions >= keep are read afterwards (the top SIGNIF diagonals), and a
// partition's elements never leave it, so a partition wholly below keep is left
// unsorted: the values at positions >= keep are exactly the full sort's.
__device__ void des_quick_sort(const FastArgs& a, const Lane& ln, int n, int keep)
{
    int lst[48], ust[48];
    int p = 1;
    lst[1] = 1;
    ust[1] = n - 1;
    while (p > 0) {
        if (lst[p] >= ust[p] || ust[p] < keep) {
            p--;
            continue;
        }
        int i = lst[p] - 1, j = ust[p];
        const int piv = ln.at(a.o_displ, j);
        while (i < j) {
            for (i = i + 1; ln.at(a.o_displ, i) < piv; i++)
                ;
            for (j = j - 1; j > i; j--)
                if (ln.at(a.o_displ, j) <= piv) break;
            if (i < j) {
                int t = ln.at(a.o_displ, i);
                ln.at(a.o_displ, i) = ln.at(a.o_displ, j);
                ln.at(a.o_displ, j) = t;
                t = ln.at(a.o_diag, i);
                ln.at(a.o_diag, i) = ln.at(a.o_diag, j);
                ln.at(a.o_diag, j) = t;
            }
        }
        j = ust[p];
        int t = ln.at(a.o_displ, i);
        ln.at(a.o_displ, i) = ln.at(a.o_displ, j);
        ln.at(a.o_displ, j) = t;
        t = ln.at(a.o_diag, i);
        ln.at(a.o_diag, i) = ln.at(a.o_diag, j);
        ln.at(a.o_diag, j) = t;
        if (i - lst[p] < ust[p] - i) {
            lst[p + 1] = lst[p];
            ust[p + 1] = i - 1;
            lst[p] = i + 1;
        } else {
            lst[p + 1] = i + 1;
            ust[p + 1] = ust[p];
            ust[p] = i - 1;
        }
        p = p + 1;
    }
}

// pairAlign (:171-402) followed by the score of :87-100; returns the score
// (accum[0][maxSoFar], 0 when no fragment) through *score_out.
__device__ void pair_align(const FastArgs& a, const Lane& ln, const uint8_t* s1, int l1, int l2, int* score_out)
{
    const int mr = a.dna ? 3 : 24;
    const int tl1 = l1 + l2 - 1;
    for (int i = 1; i <= tl1; ++i) {
        ln.at(a.o_slopes, i) = 0;
        ln.at(a.o_displ, i) = 0;
        ln.at(a.o_diag, i) = i;
    }
    for (int c = 1; c <= a.limit; ++c)                  // k-tuple matches per diagonal
        for (int vn1 = ln.at(a.o_zzc, c); vn1; vn1 = ln.at(a.o_zza, vn1))
            for (int vn2 = ln.at(a.o_zzd, c); vn2; vn2 = ln.at(a.o_zzb, vn2)) ++ln.at(a.o_displ, vn1 - vn2 + l2);
    int j = tl1 - a.signif + 1;
    if (j < 1) j = 1;
    des_quick_sort(a, ln, tl1, j);
    for (int i = tl1; i >= j; i--)                      // WINDOW around the top SIGNIF diagonals
        if (ln.at(a.o_displ, i) > 0) {
            const int pos = ln.at(a.o_diag, i);
            const int l = 1 > pos - a.window ? 1 : pos - a.window;
            const int m = tl1 < pos + a.window ? tl1 : pos + a.window;
            for (int q = l; q <= m; q++) ln.at(a.o_slopes, q) = 1;
        }
    for (int i = 1; i <= tl1; ++i) ln.at(a.o_displ, i) = 0;
    int curr = 0, maxsofar = 0;
    // accum rows 0 (score), 1 and 2 (start in sequence 1 and 2) and 4 (next in
    // score order); row 3 (the predecessor) is written but never read by the
    // reference and is left out
    const int64_t A0 = a.o_acc, A1 = A0 + a.acc_len, A2 = A1 + a.acc_len, A4 = A2 + a.acc_len;
    ln.at(A0, 0) = 0;        // the score putFrag's walk compares with at the end of the list
    for (int i = 1; i <= l1 - a.ktup + 1; ++i) {
        int encrypt = 0, pot = 1;
        bool bad = false;
        for (int q = 1; q <= a.ktup; ++q) {
            const int r = res_at(a, s1, i + q - 1);
            if (r < 0 || r > mr) {
                bad = true;
                break;
            }
            encrypt += r * pot;
            pot *= mr + 1;
        }
        if (bad) continue;
        ++encrypt;
        for (int vn2 = ln.at(a.o_zzd, encrypt); vn2; vn2 = ln.at(a.o_zzb, vn2)) {
            const int osptr = i - vn2 + l2;
            if (ln.at(a.o_slopes, osptr) != 1) continue;
            int fs = a.ktup, next = maxsofar;
            for (;;) {                                  // A-loop (:299-393)
                if (!next) {
                    ++curr;
                    if (curr >= 2 * a.max_aln) {        // "(Partial alignment)": stop here
                        *score_out = maxsofar ? ln.at(A0, maxsofar) : 0;
                        return;
                    }
                    ln.at(a.o_displ, osptr) = curr;
                    // putFrag (:523-559): into the list kept in descending score order
                    ln.at(A0, curr) = fs;
                    ln.at(A1, curr) = i;
                    ln.at(A2, curr) = vn2;
                    if (!maxsofar) {
                        maxsofar = 1;
                        ln.at(A4, curr) = 0;
                    } else if (fs >= ln.at(A0, maxsofar)) {
                        ln.at(A4, curr) = maxsofar;
                        maxsofar = curr;
                    } else {
                        int nx = maxsofar, end;
                        for (;;) {
                            end = nx;
                            nx = ln.at(A4, nx);
                            if (fs >= ln.at(A0, nx)) break;
                        }
                        ln.at(A4, curr) = nx;
                        ln.at(A4, end) = curr;
                    }
                    break;
                }
                const int t1 = ln.at(A1, next), t2 = ln.at(A2, next);
                bool rel;                               // fragRelPos (:561-582)
                if (i - vn2 == t1 - t2) rel = t1 < i;
                else rel = t1 + a.ktup - 1 < i && t2 + a.ktup - 1 < vn2;
                if (!rel) {
                    next = ln.at(A4, next);
                    continue;
                }
                if (i - vn2 == t1 - t2) {
                    if (i > t1 + (a.ktup - 1)) fs = ln.at(A0, next) + a.ktup;
                    else fs = ln.at(A0, next) + (i - t1);
                } else {
                    const int d = ln.at(a.o_displ, osptr);
                    int subt1;
                    if (d == 0) subt1 = a.ktup;
                    else if (i > ln.at(A1, d) + (a.ktup - 1)) subt1 = ln.at(A0, d) + a.ktup;
                    else subt1 = ln.at(A0, d) + (i - ln.at(A1, d));
                    const int subt2 = ln.at(A0, next) - a.wind_gap + a.ktup;
                    fs = subt2 > subt1 ? subt2 : subt1;
                }
                next = 0;
            }
        }
    }
    *score_out = maxsofar ? ln.at(A0, maxsofar) : 0;
}

}  // namespace

extern "C" __global__ void __launch_bounds__(256) cw_fast_pairs(FastArgs a)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.lanes) return;
    const Lane ln{a.scratch + g * a.lane_ints};
    for (;;) {
        const int64_t k = atomicAdd(a.next, 1u);
        if (k >= a.npairs) break;
        const int si = a.pairs[2 * k], sj = a.pairs[2 * k + 1];
        const int l1 = a.lens[si], l2 = a.lens[sj];
        const uint8_t* s1 = a.codes + a.offs[si];
        const uint8_t* s2 = a.codes + a.offs[sj];
        make_ptrs(a, ln, a.o_zza, a.o_zzc, s1, l1);
        make_ptrs(a, ln, a.o_zzb, a.o_zzd, s2, l2);
        int sc = 0;                                     // a fragment scores >= ktup, so 0 = none
        pair_align(a, ln, s1, l1, l2, &sc);
        double score = 0.0;                             // FastPairwiseAlign.cpp:87-100
        if (sc != 0) {
            score = (double)sc;
            if (a.percent) score = (score / (double)(l1 < l2 ? l1 : l2)) * 100.0;
        }
        a.out_dist[k] = (100.0 - score) / 100.0;
    }
}

// ---- one wave per pair, arrays in LDS ------------------------------------------
// pairAlign with a pair's arrays in LDS (16-bit where values fit).  The
// lane-parallel parts: k-tuple codes, clearing, the diagonal counts (LDS
// atomics), the list of flagged diagonals, and for each position of sequence 1
// the check of every flagged diagonal for a match.  The two position lists are
// built by two lanes at once.  The sequential parts (the sort, one match after
// another) are run by every lane alike, so the wave never diverges there; the
// sort's scans and the searches of the fragment list use all 64 lanes and a
// ballot.  Every lane reads what it needs into registers before a wave barrier
// and writes after it, so no lane reads a location another lane of the same
// step writes.  For a position i of sequence 1 the reference
// visits the positions vn2 of sequence 2 with the same k-tuple in descending
// order and skips unflagged diagonals; taking the flagged diagonals in
// ascending order (vn2 = i + l2 - diagonal, descending) and keeping those whose
// k-tuple matches visits the same (i, vn2) in the same order.
// The reference keeps the fragments in a list linked in score order; here
// that order is implicit (see the fragment loop), so a fragment is its score
// and its start: 6 bytes.  Values every lane reads alike are taken into scalar
// registers, so the sequential parts branch on the scalar unit.
struct WaveLds {
    uint32_t* fp;                   // fragments: start in sequence 1 << 16 | start in sequence 2
    uint16_t* fsc;                  // fragments: score
    int16_t* displ;                 // [D] diagonal counts (two per 32-bit word while counting), then fragments
    int16_t *kt1, *kt2;             // k-tuple code + 1 per position, 0: none (positions 1..L)
    int16_t* diag;                  // [D] diagonal index, then the flagged diagonals in order
    // one region, first:
    int16_t *tp1, *tp2;             // previous position with the same k-tuple
    int16_t *pl1, *pl2;             // last position of each code
    uint8_t* slopes;                // [D]
    // then, while the fragments are built:
    uint32_t* bk;                   // per score: newest fragment | oldest fragment << 16 (0: none)
};

__device__ __forceinline__ WaveLds wave_lds(const FastArgs& a, unsigned char* base)
{
    WaveLds w;
    w.fp = (uint32_t*)base;
    w.fsc = (uint16_t*)(w.fp + a.acc_len);
    w.displ = (int16_t*)(w.fsc + a.acc_len);   // acc_len is even
    w.kt1 = w.displ + a.wD;
    w.kt2 = w.kt1 + (a.wL + 2);
    w.diag = w.kt2 + (a.wL + 2);
    w.tp1 = w.diag + a.wD;          // 4-byte aligned: wD and wL + 2 + wL + 2 are even
    w.tp2 = w.tp1 + (a.wL + 2);
    w.pl1 = w.tp2 + (a.wL + 2);
    w.pl2 = w.pl1 + (a.limit + 2);
    w.slopes = (uint8_t*)(w.pl2 + (a.limit + 2));
    w.bk = (uint32_t*)w.tp1;
    return w;
}

__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); __builtin_amdgcn_wave_barrier(); __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront"); }

__device__ __forceinline__ int fr_score(uint32_t v) { return (int)(v >> 16); }
// lane l receives src of lane l - 1, lane 0 receives old (DPP wave_shr:1)
__device__ __forceinline__ int wave_shr1(int old, int src) { return __builtin_amdgcn_update_dpp(old, src, 0x138, 0xf, 0xf, false); }
// a value every lane holds alike, in a scalar register
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

extern "C" __global__ void __launch_bounds__(256) cw_fast_wave(FastArgs a)
{
    // one or more waves per workgroup, each with its own lds_wave bytes
    extern __shared__ __align__(16) unsigned char lds[];
    const WaveLds w = wave_lds(a, lds + (size_t)(threadIdx.x >> 6) * (size_t)a.lds_wave);
    const int lane = threadIdx.x & 63;
    const int mr = a.dna ? 3 : 24;
    // a second pass takes the pairs a first one with fewer fragment slots listed
    const int64_t nwork = a.plist ? (int64_t)*a.plist_n : a.npairs;
    for (;;) {
        const unsigned got = atomicAdd(a.next, lane == 0 ? 1u : 0u);
        const int64_t kk = (int64_t)uni(got);        // lane 0's count: the work item of this wave
        if (kk >= nwork) break;
        const int64_t k = a.plist ? (int64_t)a.plist[kk] : kk;
        const int si = a.pairs[2 * k], sj = a.pairs[2 * k + 1];
        const int l1 = a.lens[si], l2 = a.lens[sj];
        const uint8_t* s1 = a.codes + a.offs[si];
        const uint8_t* s2 = a.codes + a.offs[sj];
        // k-tuple codes (+1; 0 where a k-tuple does not fit or has a bad residue)
        for (int h = 0; h < 2; ++h) {
            const int L = h ? l2 : l1;
            const uint8_t* s = h ? s2 : s1;
            int16_t* kt = h ? w.kt2 : w.kt1;
            for (int p = 1 + lane; p <= L; p += 64) {
                int code = 0, pot = 1;
                bool bad = p > L - a.ktup + 1;
                for (int j = 0; j < a.ktup && !bad; ++j) {
                    const int r = a.dna ? dna_of(s[p - 1 + j]) : (int)s[p - 1 + j];
                    if (r < 0 || r > mr) bad = true;
                    code += r * pot;
                    pot *= mr + 1;
                }
                kt[p] = bad ? 0 : (int16_t)(code + 1);
            }
        }
        for (int c = 1 + lane; c <= a.limit; c += 64) {
            w.pl1[c] = 0;
            w.pl2[c] = 0;
        }
        const int tl1 = l1 + l2 - 1;
        for (int d = lane; d <= tl1 + 1; d += 64) {     // whole words of the counts
            w.slopes[d] = 0;
            w.displ[d] = 0;
            w.diag[d] = (int16_t)d;
        }
        wave_sync();
        if ((lane & 31) == 0) {   // makePPtrs / makeNPtrs: lane 0 sequence 1's lists, lane 32 sequence 2's
            const int h = lane >> 5;
            const int L = h ? l2 : l1;
            const int16_t* kt = h ? w.kt2 : w.kt1;
            int16_t* tp = h ? w.tp2 : w.tp1;
            int16_t* pl = h ? w.pl2 : w.pl1;
            for (int i = 1; i <= L; ++i) {
                const int c = kt[i];
                int16_t prev = 0;
                if (c) {
                    prev = pl[c];
                    pl[c] = (int16_t)i;
                }
                tp[i] = prev;
            }
        }
        wave_sync();
        // k-tuple matches per diagonal: 16-bit counts, two to a 32-bit word
        // (a count stays below 2^15, so an add never carries into its neighbour)
        uint32_t* words = (uint32_t*)w.displ;
        for (int v1 = 1 + lane; v1 <= l1; v1 += 64) {
            const int c = w.kt1[v1];
            if (!c) continue;
            for (int v2 = w.pl2[c]; v2; v2 = w.tp2[v2]) {
                const int d = v1 - v2 + l2;
                atomicAdd(&words[d >> 1], 1u << (16 * (d & 1)));
            }
        }
        wave_sync();
        int jtop = tl1 - a.signif + 1;
        if (jtop < 1) jtop = 1;
        // desQuickSort over 1 .. tl1-1, partitions wholly below jtop left unsorted
        // (every lane runs it alike)
        {
            int lst[48], ust[48];
            int p = 1;
            lst[1] = 1;
            ust[1] = tl1 - 1;
            while (p > 0) {
                if (lst[p] >= ust[p] || ust[p] < jtop) {
                    p--;
                    continue;
                }
                int i = lst[p] - 1, j = ust[p];
                const int top_p = ust[p];
                const int piv = w.displ[j];
                while (i < j) {
                    // the two scans, 64 positions at a time: i to the first
                    // position after i with a count >= piv (the pivot at top_p
                    // stops it), j to the last one in (i, j) with a count <= piv,
                    // else to i
                    for (int b = i + 1;; b += 64) {
                        const int q = b + lane;
                        const uint64_t hb = __ballot(q <= top_p && w.displ[q] >= piv);
                        if (hb) {
                            i = b + __builtin_ctzll(hb);
                            break;
                        }
                    }
                    int jn = i;
                    for (int e = j - 1; e > i; e -= 64) {
                        const int q = e - lane;
                        const uint64_t hb = __ballot(q > i && w.displ[q] <= piv);
                        if (hb) {
                            jn = e - __builtin_ctzll(hb);
                            break;
                        }
                    }
                    j = jn;
                    if (i < j) {
                        const int16_t ti = w.displ[i], tj = w.displ[j];
                        const int16_t ui = w.diag[i], uj = w.diag[j];
                        wave_sync();
                        w.displ[i] = tj;
                        w.diag[i] = uj;
                        w.displ[j] = ti;
                        w.diag[j] = ui;
                        wave_sync();
                    }
                }
                j = ust[p];
                const int16_t ti = w.displ[i], tj = w.displ[j];
                const int16_t ui = w.diag[i], uj = w.diag[j];
                wave_sync();
                w.displ[i] = tj;
                w.diag[i] = uj;
                w.displ[j] = ti;
                w.diag[j] = ui;
                wave_sync();
                if (i - lst[p] < ust[p] - i) {
                    lst[p + 1] = lst[p];
                    ust[p + 1] = i - 1;
                    lst[p] = i + 1;
                } else {
                    lst[p + 1] = i + 1;
                    ust[p + 1] = ust[p];
                    ust[p] = i - 1;
                }
                p = p + 1;
            }
        }
        // WINDOW around the top SIGNIF diagonals
        for (int i = tl1; i >= jtop; i--) {
            const int cnt = w.displ[i];
            const int pos = w.diag[i];
            wave_sync();
            if (cnt > 0) {
                const int l = 1 > pos - a.window ? 1 : pos - a.window;
                const int m = tl1 < pos + a.window ? tl1 : pos + a.window;
                for (int q = l + lane; q <= m; q += 64) w.slopes[q] = 1;
            }
            wave_sync();
        }
        // clear the diagonals' fragment slots; list the flagged diagonals in order
        int nf = 0;
        for (int b0 = 1; b0 <= tl1; b0 += 64) {
            const int d = b0 + lane;
            const bool f = d <= tl1 && w.slopes[d] == 1;
            const uint64_t bal = __ballot(f);
            const int before = __popcll(bal & ((1ull << lane) - 1ull));
            if (d <= tl1) w.displ[d] = 0;
            wave_sync();
            if (f) w.diag[nf + before] = (int16_t)d;
            nf += __popcll(bal);
            wave_sync();
        }
        // the position lists and flags are done with: their region holds the
        // score buckets from here on (scores 1 .. ktup l1)
        const int smax = a.ktup * l1;
        for (int q = lane; q <= smax; q += 64) w.bk[q] = 0u;
        int curr = 0, maxsofar = 0, topscore = 0;
        uint32_t top = 0, tx = 0, ty = 0;   // the list's first 64 fragments, one per lane
        // the first 64 flagged diagonals one per lane, with the latest fragment on
        // each (its number, and score << 16 | start in sequence 1)
        const int dg = lane < nf ? (int)w.diag[lane] : 0;
        uint32_t fl_d = 0, fl_rec = 0;
        int k1 = 0;                         // k-tuple codes of 64 positions of sequence 1
        bool stop = false, ovf = false;
        wave_sync();
        for (int i = 1; i <= l1 - a.ktup + 1 && !stop; ++i) {
            if (((i - 1) & 63) == 0) k1 = i + lane <= l1 ? (int)w.kt1[i + lane] : 0;
            const int c = __builtin_amdgcn_readlane(k1, (i - 1) & 63);
            if (!c) continue;
            for (int f0 = 0; f0 < nf && !stop; f0 += 64) {
                const int fi = f0 + lane;
                bool m = false;
                if (fi < nf) {
                    const int v2c = i + l2 - (f0 == 0 ? dg : (int)w.diag[fi]);
                    m = v2c >= 1 && v2c <= l2 && w.kt2[v2c] == c;
                }
                uint64_t bal = __ballot(m);
                while (bal && !stop) {
                    const int b = __builtin_ctzll(bal);
                    bal &= bal - 1;
                    const int fb = f0 + b;          // the diagonal's place in the flagged list
                    const int osptr = f0 == 0 ? __builtin_amdgcn_readlane(dg, b)
                                              : (int)uni((uint32_t)(int)w.diag[fb]);
                    const int vn2 = i + l2 - osptr;
                    // A-loop (:299-393): the first fragment, in list order, that
                    // (i, vn2) can extend (fragRelPos, :561-582).  Lane l holds the
                    // list's l-th fragment (top: its number, 0 past the list's end;
                    // tx, ty: its record), so the first 64 are tested at once; the
                    // rest of the list is searched by all lanes over the records.
                    int fs = a.ktup, nx = 0;
                    uint32_t nrx = 0, nry = 0;
                    {
                        const int t1v = (int)(ty >> 16), t2v = (int)(ty & 0xffffu);
                        const bool relv = top != 0u && (i - vn2 == t1v - t2v
                                                            ? t1v < i
                                                            : (t1v + a.ktup - 1 < i && t2v + a.ktup - 1 < vn2));
                        const uint64_t rb = __ballot(relv);
                        if (rb) {
                            const int l = __builtin_ctzll(rb);
                            nx = __builtin_amdgcn_readlane((int)top, l);
                            nrx = (uint32_t)__builtin_amdgcn_readlane((int)tx, l);
                            nry = (uint32_t)__builtin_amdgcn_readlane((int)ty, l);
                        } else if (curr > 64) {
                            // past the 64th: list order is that of score << 16 |
                            // number, descending, so the first fragment after the
                            // 64th that (i, vn2) can extend has the largest such key
                            // below the 64th's among all fragments
                            const uint32_t key63 = ((uint32_t)__builtin_amdgcn_readlane((int)tx, 63) & 0xffff0000u) |
                                                   (uint32_t)__builtin_amdgcn_readlane((int)top, 63);
                            uint32_t best = 0;
                            for (int f = 1 + lane; f <= curr; f += 64) {
                                const uint32_t r = w.fp[f];
                                const uint32_t key = ((uint32_t)w.fsc[f] << 16) | (uint32_t)f;
                                const int t1 = (int)(r >> 16), t2 = (int)(r & 0xffffu);
                                const bool rel = i - vn2 == t1 - t2 ? t1 < i
                                                                    : (t1 + a.ktup - 1 < i && t2 + a.ktup - 1 < vn2);
                                if (rel && key < key63 && key > best) best = key;
                            }
                            for (int o = 32; o; o >>= 1) {
                                const uint32_t q = (uint32_t)__shfl_xor((int)best, o, 64);
                                best = q > best ? q : best;
                            }
                            best = uni(best);
                            if (best) {
                                nx = (int)(best & 0xffffu);
                                nrx = best & 0xffff0000u;
                                nry = uni(w.fp[nx]);
                            }
                        }
                    }
                    if (nx) {
                        const int t1 = (int)(nry >> 16), t2 = (int)(nry & 0xffffu);
                        const int sn = fr_score(nrx);
                        if (i - vn2 == t1 - t2) {
                            fs = i > t1 + (a.ktup - 1) ? sn + a.ktup : sn + (i - t1);
                        } else {
                            // the latest fragment on this diagonal (displ[osptr])
                            int d;
                            uint32_t drec = 0;
                            if (fb < 64) {
                                d = __builtin_amdgcn_readlane((int)fl_d, fb);
                                drec = (uint32_t)__builtin_amdgcn_readlane((int)fl_rec, fb);
                            } else {
                                d = (int)uni((uint32_t)(int)w.displ[osptr]);
                                if (d) drec = (uni((uint32_t)w.fsc[d]) << 16) | (uni(w.fp[d]) >> 16);
                            }
                            int subt1;
                            if (d == 0) {
                                subt1 = a.ktup;
                            } else {
                                const int d1 = (int)(drec & 0xffffu), ds = (int)(drec >> 16);
                                subt1 = i > d1 + (a.ktup - 1) ? ds + a.ktup : ds + (i - d1);
                            }
                            const int subt2 = sn - a.wind_gap + a.ktup;
                            fs = subt2 > subt1 ? subt2 : subt1;
                        }
                    }
                    ++curr;
                    if (curr >= 2 * a.max_aln) {
                        stop = true;                    // "(Partial alignment)"
                        break;
                    }
                    if (curr >= a.acc_len) {
                        stop = ovf = true;              // no slot left: the pair goes to the second pass
                        break;
                    }
                    // putFrag (:523-559): into the list kept in descending score
                    // order.  That list is ordered by score, descending, and among
                    // equal scores by fragment number, descending (a new fragment
                    // goes before the first one with a score <= its own), so the
                    // order is kept without links: the search above reads it off
                    // the keys, and the first 64 places live in the lanes.  The new
                    // fragment follows the oldest fragment of the lowest score above
                    // fs (end; none: it heads the list), which the score buckets
                    // give: lanes look at 64 scores from fs up.
                    int end = 0;
                    uint32_t bv = 0;                    // oldest fragment of score fs (0: none)
                    if (maxsofar && fs < topscore) {
                        for (int up = fs;; up += 64) {  // ends at topscore at the latest
                            const int sq = up + lane;
                            const uint32_t v = sq <= topscore ? w.bk[sq] : 0u;
                            if (up == fs) bv = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
                            const uint64_t sb = __ballot(v != 0u && sq > fs);
                            if (sb) {
                                end = __builtin_amdgcn_readlane((int)v, __builtin_ctzll(sb));
                                break;
                            }
                        }
                    } else if (maxsofar && fs == topscore) {
                        bv = uni(w.bk[fs]);
                    }
                    // list position of the new fragment where it is among the first 64
                    int pos = 0;
                    if (end) {
                        const uint64_t eb = __ballot(top == (uint32_t)end);
                        pos = eb ? __builtin_ctzll(eb) + 1 : 65;
                    }
                    wave_sync();
                    const uint32_t nrec_y = ((uint32_t)i << 16) | (uint32_t)vn2;
                    w.displ[osptr] = (int16_t)curr;
                    w.fp[curr] = nrec_y;
                    w.fsc[curr] = (uint16_t)fs;
                    if (!bv) w.bk[fs] = (uint32_t)curr;
                    wave_sync();
                    // the same in the lanes: the new fragment at lane pos, the ones
                    // after it one lane further on
                    const uint32_t st = (uint32_t)wave_shr1(0, (int)top);
                    const uint32_t sx = (uint32_t)wave_shr1(0, (int)tx);
                    const uint32_t sy = (uint32_t)wave_shr1(0, (int)ty);
                    if (pos < 64) {
                        if (lane > pos) {
                            top = st;
                            tx = sx;
                            ty = sy;
                        } else if (lane == pos) {
                            top = (uint32_t)curr;
                            tx = (uint32_t)fs << 16;
                            ty = nrec_y;
                        }
                    }
                    if (lane == fb) {
                        fl_d = (uint32_t)curr;
                        fl_rec = ((uint32_t)fs << 16) | (uint32_t)i;
                    }
                    if (!maxsofar) maxsofar = 1;
                    else if (!end) maxsofar = curr;
                    if (!end) topscore = fs;
                }
            }
        }
        const int sc = maxsofar ? topscore : 0;     // the list head's score
        if (lane == 0 && ovf) {
            const unsigned o = atomicAdd(a.ovf_n, 1u);
            a.ovf_list[o] = (int32_t)k;
        } else if (lane == 0) {
            double score = 0.0;                         // FastPairwiseAlign.cpp:87-100
            if (sc != 0) {
                score = (double)sc;
                if (a.percent) score = (score / (double)(l1 < l2 ? l1 : l2)) * 100.0;
            }
            a.out_dist[k] = (100.0 - score) / 100.0;
        }
        wave_sync();
    }
}

hipError_t cwf_launch_wave(int blocks, int waves, size_t lds_bytes, hipStream_t st, const FastArgs& a)
{
    hipError_t e = hipFuncSetAttribute((const void*)cw_fast_wave, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds_bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(cw_fast_wave, dim3(blocks), dim3(64 * waves), lds_bytes, st, a);
    return hipGetLastError();
}

hipError_t cwf_launch(int blocks, hipStream_t st, const FastArgs& a)
{
    hipLaunchKernelGGL(cw_fast_pairs, dim3(blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}
