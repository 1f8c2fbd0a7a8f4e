// cw_pairwise_kernels.hip -- gfx950 kernels of ClustalW's full-DP pairwise stage.
//
// One wavefront (64 lanes) aligns one pair end to end:
//   forward local Gotoh pass   (FullPairwiseAlign.cpp:280-341)  -> maxScore, (se1,se2)
//   reverse pass                (FullPairwiseAlign.cpp:343-416)  -> (sb1,sb2)
//   Myers-Miller recursion      (FullPairwiseAlign.cpp:418-612)  -> identity count
// (paths relative to CPU/clustalw-2.0.11/src of coldfunction/CUDA-clustalW).
//
// DP layout: every pass is a "sweep" over a rows x cols rectangle.  Lanes own
// rows (R consecutive rows per lane, a stripe of up to 64*R rows) and the wave
// walks the columns as an anti-diagonal wavefront: at step t lane k works on
// column t-k.  The H/E values of a lane's last row move to the next lane with
// one DPP wave_shr:1 per value; the lane-0 input (the stripe's top boundary
// row) and the lane-63 output (the stripe's bottom row) stream through 64-wide
// register chunks rotated with DPP wave_shl:1 and read/written to memory one
// coalesced chunk per 64 steps.  The substitution matrix sits in LDS, sequence
// residues in per-wave LDS buffers.  All arithmetic is int32 and every
// comparison keeps the reference's order, so results are bit-identical to the
// CPU code:
//   * forward max: first cell in row-major order (smallest i, then j);
//   * reverse start: first cell in the reverse scan order with H >= maxScore;
//   * diff midpoint: the two scans of :570-593 as parallel reductions;
//   * the identity count is summed over the M==1 leaves of the recursion, which
//     is order-free, so the recursion runs as a depth-first stack in LDS.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#include "cw_pairwise_kernels.h"

namespace cwk {

constexpr int WAVE = 64;
constexpr int BIG = 0x7fffffff;

enum { MODE_FWD = 0, MODE_REV = 1, MODE_GLB = 2 };

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// lane l receives src[l-1]; lane 0 receives old
__device__ __forceinline__ int shr1(int old, int src) {
    return __builtin_amdgcn_update_dpp(old, src, 0x138, 0xf, 0xf, false);
}
// lane l receives src[l+1]; lane 63 receives old
__device__ __forceinline__ int shl1(int old, int src) {
    return __builtin_amdgcn_update_dpp(old, src, 0x130, 0xf, 0xf, false);
}
// lane l receives src[l+1]; lane 63 receives 0 (no copy of an old value)
__device__ __forceinline__ int shl1z(int src) {
    return __builtin_amdgcn_mov_dpp(src, 0x130, 0xf, 0xf, true);
}

// Wave reductions with DPP (all 64 lanes active): within quads, half-rows and
// rows of 16, then row_bcast:15 / :31 carry the rows into row 3, so lane 63
// holds the result, which is read into a scalar register.  Six VALU steps,
// no LDS traffic.
template <typename Op>
__device__ __forceinline__ int wave_reduce(int x, Op op)
{
    x = op(x, __builtin_amdgcn_update_dpp(x, x, 0xB1, 0xf, 0xf, false));   // quad_perm [1,0,3,2]
    x = op(x, __builtin_amdgcn_update_dpp(x, x, 0x4E, 0xf, 0xf, false));   // quad_perm [2,3,0,1]
    x = op(x, __builtin_amdgcn_update_dpp(x, x, 0x141, 0xf, 0xf, false));  // row_half_mirror
    x = op(x, __builtin_amdgcn_update_dpp(x, x, 0x140, 0xf, 0xf, false));  // row_mirror
    x = op(x, __builtin_amdgcn_update_dpp(x, x, 0x142, 0xa, 0xf, false));  // row_bcast:15 -> rows 1, 3
    x = op(x, __builtin_amdgcn_update_dpp(x, x, 0x143, 0xc, 0xf, false));  // row_bcast:31 -> rows 2, 3
    return __builtin_amdgcn_readlane(x, 63);
}
__device__ __forceinline__ int wave_max(int v) {
    return wave_reduce(v, [](int a, int b) { return a > b ? a : b; });
}
__device__ __forceinline__ int wave_min(int v) {
    return wave_reduce(v, [](int a, int b) { return a < b ? a : b; });
}

// make LDS/global writes of this wave visible to its later reads
__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
}

// LDS address space: keeps sequence and matrix reads on ds_read_* rather
// than flat loads.
#define CW_LDS __attribute__((address_space(3)))
typedef CW_LDS const uint8_t lcu8;
typedef CW_LDS uint8_t lu8;
// Global address space for the boundary rows in HBM: global (not flat)
// loads and stores, which an LDS wait (lgkmcnt) does not wait for.
#define CW_GLB __attribute__((address_space(1)))
typedef CW_GLB int gint;
typedef CW_GLB const int gcint;

struct Sweep {
    int nrows, ncols;
    lcu8* rs; int rb, rd;            // code of row r          : rs[rb + rd*r]
    lcu8* cs; int cb, cd;            // 4 * code of column c   : cs[cb + cd*c]
    int th0, ths, teo;               // top row   : H[0][c] = th0 + c*ths, E = H + teo (c>=1)
    int lh0, lhs, lfo;               // left col  : H[r][0] = lh0 + r*lhs, F = H + lfo (r>=1)
    int go, ge, goe;                 // H[0][0] = 0 in every pass
    int thr;                         // MODE_REV threshold (maxScore)
    int kb;                          // packed tracking: column bits (ncols < 2^kb)
    lcu8* tab;                       // LDS substitution matrix, int32 [32][CW_TAB_STRIDE], byte-addressed
};

struct Track {
    int best;   // MODE_FWD: best value; MODE_REV: 1 when found
    int key;    // (row << 16) | col of the tracked cell
};

__device__ __forceinline__ int left_h(const Sweep& S, int r) { return r == 0 ? 0 : S.lh0 + r * S.lhs; }

__device__ __forceinline__ int score_at(lcu8* tab, int byte_off)
{
    return *(CW_LDS const int*)(tab + byte_off);
}

template <bool V> struct Tag { static constexpr bool value = V; };

// One stripe: rows row0+1 .. row0+P.  R rows per lane.  DIRECT: R==1 and P<64,
// lanes >= P idle and lane P-1 writes its row straight to out (only ever the
// single stripe of a sweep, so its top boundary is the formula row).
//
// PACK (MODE_FWD, when the pair's scores fit beside S.kb column bits): a
// row's tracked cell is one int, max over (h << kb) | (2^kb-1 - c), i.e. the
// largest h and among equal h the smallest column -- one shift-or and one max
// per cell instead of a compare and two selects.  PACK also clamps E and F at
// 0: with ge >= 0 that gives max(E, 0) and max(F, 0) at every cell (induction
// over E' = max(E - ge, H - go - ge, 0)) and leaves every H unchanged, so
// H = max3(diag + s, E, F) needs no clamp.
//
// PACK with MODE_GLB: the stripe carries every H, E and F of cell (i, j) as
// X + (i + j) * ge (i, j the sweep's own row and column).  In that frame both
// gap recurrences lose their extension term, E' = max(E~, H~ - go) and
// F' = max(F~, H~ - go) share one H~ - go, and the diagonal carries + 2 ge,
// kept in the stored H; the top input row and the bottom output row convert
// once per 64-column chunk.  One add per cell fewer, values identical.
template <int R, int MODE, bool DIRECT, bool PACK>
__device__ __forceinline__ void stripe(const Sweep& S, int row0, int P, int first,
                                       const int* inH, const int* inE, int* outH,
                                       int* outE, int want_out, Track& T)
{
    const int lane = lane_id();
    // wave-uniform values in scalar registers, so the step loop's control
    // (chunk turns, output flushes, loop bounds) runs on the scalar unit
    const int ncols = uni(S.ncols);
    P = uni(P);
    first = uni(first);
    want_out = uni(want_out);
    const int k1 = DIRECT ? P : P - WAVE * (R - 1);    // lanes < k1 hold R rows
    const bool dummy = (!DIRECT) && (R > 1) && (lane >= k1);
    const bool lane_on = DIRECT ? (lane < P) : true;
    const int d = dummy ? 1 : 0;
    const int firstrow = row0 + 1 + (lane < k1 ? lane * R : k1 * R + (lane - k1) * (R - 1));

    constexpr bool SH = PACK && MODE == MODE_GLB;        // shifted frame
    const int ge = uni(S.ge), goe = uni(S.goe), thr = uni(S.thr);
    const int go = uni(S.go), ge2 = 2 * ge;
    int sb[R], hL[R], hLg[R], f[R], tj[R], tv[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int row = firstrow + r - d;
        int rr = row < 1 ? 1 : (row > S.nrows ? S.nrows : row);
        sb[r] = (int)S.rs[S.rb + S.rd * rr] * CW_TAB_ROW;
        hL[r] = left_h(S, row);
        f[r] = hL[r] + S.lfo;
        hLg[r] = hL[r] - S.goe;
        if (SH) {
            const int sh = row * ge;                      // column 0
            f[r] += sh;
            hLg[r] = hL[r] + sh - go;
            hL[r] += sh + ge2;
        }
        tv[r] = 0;
        tj[r] = BIG;
    }
    int dprev = left_h(S, firstrow - 1);
    if (SH) dprev += (firstrow - 1) * ge + ge2;
    const int kb = uni(S.kb), cmask = (1 << kb) - 1;
    // 0 in a scalar register the compiler cannot see through: with a literal
    // 0 it proves E, F >= 0, turns the H maximum into a signed and an unsigned
    // max and loses the single v_max3
    int zero;
    asm("s_mov_b32 %0, 0" : "=s"(zero));
    lcu8* const tab = S.tab;
    lcu8* const cs = S.cs;
    const int cd = uni(S.cd);

    // lane-0 input chunks: lane l of chunk ci holds column 64*ci + 1 + l
    auto chunk = [&](int ci, int& h, int& e) {
        int col = 64 * ci + 1 + lane;
        if (first) {
            h = S.th0 + col * S.ths;
            e = h + S.teo;
        } else if (col <= ncols) {
            h = ((gcint*)inH)[col];
            e = ((gcint*)inE)[col];
        } else {
            h = 0;
            e = 0;
        }
        if (SH) {
            const int sh = (row0 + col) * ge;
            h += sh;
            e += sh;
        }
    };
    int cinH, cinE, nxtH = 0, nxtE = 0;
    chunk(0, cinH, cinE);
    const int nchunks = (ncols + 63) >> 6;
    if (nchunks > 1) chunk(1, nxtH, nxtE);

    int outHv = 0, outEv = 0, coutH = 0, coutE = 0;
    int caddr = S.cb + cd * (1 - lane);
    int b4 = cs[caddr];                                 // 4*code of this lane's column at t=1

    // masked: some lane may be off its valid columns; flush: this step may end
    // an output chunk (every 64th output column, and the last one)
    auto step = [&](int t, auto masked, auto flush) {
        const int nH = shl1z(cinH);                     // chunk rotation for the next step (lane 63's
        const int nE = shl1z(cinE);                     // element is past the chunk's last use)
        const int inHv = shr1(cinH, outHv);             // lane 0: chunk element; others: lane above
        const int inEv = shr1(cinE, outEv);
        cinH = nH;
        cinE = nE;
        const int bcur = b4;
        caddr += cd;
        b4 = cs[caddr];                                 // prefetch next step's column code
        const int c = t - lane;
        const int ck = cmask - c;                       // packed column key: smaller c ranks higher
        bool act = true;
        if constexpr (decltype(masked)::value) act = lane_on && c >= 1 && c <= ncols;
        if (act) {
            int hup = inHv, eup = inEv, dg = dprev;
            int hupg = SH ? inHv - go : inHv - goe;
            dprev = SH ? inHv + ge2 : inHv;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int s = score_at(tab, sb[r] + bcur);
                int e, h;
                if (MODE == MODE_FWD && PACK) {
                    e = max(max(eup - ge, hupg), zero);
                    f[r] = max(max(f[r] - ge, hLg[r]), zero);
                    h = max(max(dg + s, e), f[r]);
                } else if (SH) {
                    e = max(eup, hupg);
                    f[r] = max(f[r], hLg[r]);
                    h = max(max(dg + s, e), f[r]);
                } else {
                    e = max(eup - ge, hupg);
                    f[r] = max(f[r] - ge, hLg[r]);
                    h = max(max(dg + s, e), f[r]);
                    if (MODE == MODE_FWD) h = max(h, 0);
                }
                if (r == 0 && R > 1 && !DIRECT) {
                    if (dummy) { h = hup; e = eup; }
                }
                dg = hL[r];
                hL[r] = SH ? h + ge2 : h;
                hLg[r] = SH ? h - go : h - goe;
                hup = h;
                hupg = hLg[r];
                eup = e;
                if (PACK && MODE == MODE_FWD) {
                    tv[r] = max(tv[r], (int)(((unsigned)h << kb) | (unsigned)ck));
                } else if (MODE == MODE_FWD) {
                    if (h > tv[r]) { tv[r] = h; tj[r] = c; }
                }
            }
            if (MODE == MODE_REV) {
                // no reverse score exceeds maxScore, so a hit is h == thr, which
                // few cells reach: one max over the lane's rows per step, and
                // the columns are recorded only in a step where some row hits
                int mrow = hL[0];
#pragma unroll
                for (int r = 1; r < R; ++r) mrow = max(mrow, hL[r]);
                if (mrow >= thr) {
#pragma unroll
                    for (int r = 0; r < R; ++r) tj[r] = min(tj[r], hL[r] >= thr ? c : BIG);
                }
            }
            outHv = hup;
            outEv = eup;
            if (DIRECT && want_out && lane == P - 1) {
                const int sh = SH ? (row0 + P + c) * ge : 0;
                ((gint*)outH)[c] = hup - sh;
                ((gint*)outE)[c] = eup - sh;
            }
        }
        if (!DIRECT && want_out) {
            coutH = shl1(outHv, coutH);
            coutE = shl1(outEv, coutE);
            const int oc = t - 63;
            if (decltype(flush)::value && oc >= 1 && (((oc & 63) == 0) || oc == ncols)) {
                const int col = oc - 63 + lane;
                const int cstart = ((oc - 1) & ~63) + 1;
                if (col >= cstart) {
                    const int sh = SH ? (row0 + P + col) * ge : 0;
                    ((gint*)outH)[col] = coutH - sh;
                    ((gint*)outE)[col] = coutE - sh;
                }
            }
        }
    };

    // Steps run in blocks of 64: block b covers t = 64b+1 .. 64b+64, the
    // lane-0 input chunk turns at its start.  A block in which every lane is on
    // a valid column (t in [64, ncols]) runs without exec masking; a full one
    // flushes its output chunk only after its step 64b+63 (output column 64b).
    const int tend = DIRECT ? ncols + P - 1 : ncols + 63;
    for (int b = 0; 64 * b + 1 <= tend; ++b) {
        const int t0 = 64 * b + 1;
        if (b > 0) {
            cinH = nxtH;
            cinE = nxtE;
            if (b + 1 < nchunks) chunk(b + 1, nxtH, nxtE);
        }
        const int t1 = t0 + 63 < tend ? t0 + 63 : tend;
        if (DIRECT || t0 < 64) {
#pragma unroll 1
            for (int t = t0; t <= t1; ++t) step(t, Tag<true>{}, Tag<true>{});
        } else if (t1 > ncols) {       // the steps up to ncols still have every lane on
            int t = t0;
#pragma unroll 1
            for (; t <= ncols; ++t) step(t, Tag<false>{}, Tag<true>{});
#pragma unroll 1
            for (; t <= t1; ++t) step(t, Tag<true>{}, Tag<true>{});
        } else if (t1 == t0 + 63) {
#pragma unroll 1
            for (int k = 0; k < 62; k += 2) {
                step(t0 + k, Tag<false>{}, Tag<false>{});
                step(t0 + k + 1, Tag<false>{}, Tag<false>{});
            }
            step(t0 + 62, Tag<false>{}, Tag<true>{});
            step(t0 + 63, Tag<false>{}, Tag<false>{});
        } else {
#pragma unroll 1
            for (int t = t0; t <= t1; ++t) step(t, Tag<false>{}, Tag<true>{});
        }
    }

    // per-stripe reduction of the tracked cell
    if (MODE == MODE_FWD) {
        int v = 0, key = BIG;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = firstrow + r - d;
            const bool valid = lane_on && !(r == 0 && dummy);
            const int tvr = PACK ? (tv[r] >> kb) : tv[r];
            const int tjr = PACK ? cmask - (tv[r] & cmask) : tj[r];
            if (valid && tvr > v) { v = tvr; key = (row << 16) | tjr; }
        }
        const int V = wave_max(v);
        const int K = wave_min(v == V ? key : BIG);
        if (V > T.best) { T.best = V; T.key = K; }
    } else if (MODE == MODE_REV) {
        int key = BIG;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = firstrow + r - d;
            const bool valid = lane_on && !(r == 0 && dummy);
            if (valid && tj[r] != BIG) key = min(key, (row << 16) | tj[r]);
        }
        const int K = wave_min(key);
        if (K != BIG) { T.best = 1; T.key = K; }
    }
}

// Multi-row stripes are called out of line: each is register-allocated on its
// own (no spills at <= 128 VGPRs) and the call costs nothing next to a stripe
// of >= 128 rows.  Single-row stripes, which dominate the deep Myers-Miller
// nodes, stay inline.
template <int R, int MODE, bool PACK>
__device__ __attribute__((noinline)) void stripe_call(const Sweep S, int row0, int P, int first,
                                                      const int* inH, const int* inE, int* outH,
                                                      int* outE, int want_out, Track* T)
{
    stripe<R, MODE, false, PACK>(S, row0, P, first, inH, inE, outH, outE, want_out, *T);
}

template <int MODE, bool PACK>
__device__ void run_stripe(int R, bool direct, const Sweep& S, int row0, int P, int first,
                           const int* inH, const int* inE, int* outH, int* outE, int want_out,
                           Track& T)
{
    if (direct) { stripe<1, MODE, true, PACK>(S, row0, P, first, inH, inE, outH, outE, want_out, T); return; }
    switch (R) {
    case 1: stripe<1, MODE, false, PACK>(S, row0, P, first, inH, inE, outH, outE, want_out, T); break;
    case 2: stripe_call<2, MODE, PACK>(S, row0, P, first, inH, inE, outH, outE, want_out, &T); break;
    case 3: stripe_call<3, MODE, PACK>(S, row0, P, first, inH, inE, outH, outE, want_out, &T); break;
    case 4: stripe_call<4, MODE, PACK>(S, row0, P, first, inH, inE, outH, outE, want_out, &T); break;
    case 5: stripe_call<5, MODE, PACK>(S, row0, P, first, inH, inE, outH, outE, want_out, &T); break;
    case 6: stripe_call<6, MODE, PACK>(S, row0, P, first, inH, inE, outH, outE, want_out, &T); break;
    case 7: stripe_call<7, MODE, PACK>(S, row0, P, first, inH, inE, outH, outE, want_out, &T); break;
    default: stripe_call<8, MODE, PACK>(S, row0, P, first, inH, inE, outH, outE, want_out, &T); break;
    }
}

// Full sweep over S.nrows rows in balanced stripes.  Intermediate stripe rows
// ping-pong through g[0..3] (H0,E0,H1,E1); the last stripe of a MODE_GLB sweep
// writes the final row to finH/finE (indexed by column 1..ncols).
struct Bufs { int *h0, *e0, *h1, *e1; };   // ping-pong stripe rows in global scratch

template <int MODE, bool PACK = false>
__device__ void sweep(const Sweep& S, int rmax, const Bufs& g, int* finH, int* finE, Track& T)
{
    const int nrows = S.nrows;
    const int per = WAVE * rmax;
    const int nst = uni((nrows + per - 1) / per);
    const int base = nrows / nst, rem = nrows % nst;
    int row0 = 0;
    for (int s = 0; s < nst; ++s) {
        const int P = uni(base + (s < rem ? 1 : 0));
        const int R = uni((P + WAVE - 1) / WAVE);
        const bool last = (s == nst - 1);
        const bool direct = (R == 1 && P < WAVE);
        const int* inH = (s & 1) ? g.h1 : g.h0;
        const int* inE = (s & 1) ? g.e1 : g.e0;
        int* oH = last ? finH : ((s & 1) ? g.h0 : g.h1);
        int* oE = last ? finE : ((s & 1) ? g.e0 : g.e1);
        const int want = (!last || MODE == MODE_GLB) ? 1 : 0;
        run_stripe<MODE, PACK>(R, direct, S, row0, P, s == 0, inH, inE, oH, oE, want, T);
        wave_fence();
        if (MODE == MODE_REV && T.best) break;
        row0 += P;
    }
}

struct Node { int A, B, M, N, tb, te; };

// Final rows of both halves of a Myers-Miller node (:482-568) with lanes owning
// columns: lanes 0..31 sweep the forward half (rows 1..midi from corner (A,B),
// end penalty tb), lanes 32..63 the reverse half (rows M..midi+1 from corner
// (A+M,B+N), end penalty te) at the same time, C columns per lane (N <= 32*C).
// Lane l of a half works on its row t-l at step t; the (H,F) of its last
// column moves to the next lane by DPP.  Writes HH/DD[j] to fH/fE[j] and
// RR/SS[j] to rH/rE[N-j] (j = 1..N and 0..N-1), the layout of the row path.
template <int C, bool SH>
__device__ __forceinline__ void mm_halves_cols(lcu8* tab, lcu8* s1, lcu8* s2, int A, int B, int M,
                                               int N, int tb, int te, int go, int ge, int* fH,
                                               int* fE, int* rH, int* rE)
{
    const int lane = lane_id();
    const int grp = lane >> 5;                 // 0: forward half, 1: reverse half
    const int L = lane & 31;
    const int midi = M / 2;
    const int rows = grp ? (M - midi) : midi;
    const int t0 = grp ? te : tb;
    const int goe = go + ge;
    const int rb = grp ? (A + M + 1) : A;      // code of half-row i: s1[rb + rd*i]
    const int rd = grp ? -1 : 1;
    const int nl = (N + C - 1) / C;            // lanes in use per half
    int b4[C], hup[C], hupg[C], e[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int jj = L * C + c + 1;          // half-local column
        const int jc = jj <= N ? jj : N;
        b4[c] = s2[grp ? (B + N + 1 - jc) : (B + jc)];
        hup[c] = -t0 - jj * ge;                // top row H[0][jj]
        e[c] = hup[c] - go;
        hupg[c] = hup[c] - goe;
        if (SH) {                              // cell (i, jj) carried as X + (i + jj) ge
            hupg[c] = -t0 - go;
            e[c] = -t0 - go;
            hup[c] = -t0 + 2 * ge;             // stored with the diagonal's + 2 ge
        }
    }
    int dleft = (L == 0) ? 0 : (-t0 - (L * C) * ge);   // H[0][L*C]; corner is 0
    if (SH) dleft = (L == 0 ? 0 : -t0) + 2 * ge;
    int hl_out = 0, fl_out = 0;
    int raddr = rb + rd * (1 - L);
    const int T = (midi > M - midi ? midi : M - midi) + nl - 1;
    for (int t = 1; t <= T; ++t) {
        int hl = shr1(0, hl_out);
        int fl = shr1(0, fl_out);
        const int i = t - L;
        if (L == 0) {                          // left column H[i][0], F[i][0]
            hl = SH ? -t0 : -t0 - i * ge;
            fl = hl - go;
        }
        const int arow = (int)s1[raddr] * CW_TAB_ROW;
        raddr += rd;
        if (i >= 1 && i <= rows && L < nl) {
            int d = dleft;
            int hc = hl, fc = fl;
            int hcg = SH ? hl - go : 0;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int sc = score_at(tab, arow + b4[c]);
                if (SH) {
                    const int en = max(e[c], hupg[c]);
                    const int fn = max(fc, hcg);
                    const int h = max(max(d + sc, en), fn);
                    d = hup[c];
                    hup[c] = h + 2 * ge;
                    hupg[c] = h - go;
                    hcg = hupg[c];
                    e[c] = en;
                    hc = h;
                    fc = fn;
                } else {
                    const int en = max(e[c] - ge, hupg[c]);
                    const int fn = max(fc - ge, hc - goe);
                    const int h = max(max(d + sc, en), fn);
                    d = hup[c];
                    hup[c] = h;
                    hupg[c] = h - goe;
                    e[c] = en;
                    hc = h;
                    fc = fn;
                }
            }
            dleft = SH ? hl + 2 * ge : hl;
            hl_out = hc;
            fl_out = fc;
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int jj = L * C + c + 1;
        if (jj <= N) {
            const int sh = SH ? (rows + jj) * ge : 0;
            const int hv = SH ? hup[c] - 2 * ge - sh : hup[c];
            if (grp == 0) { fH[jj] = hv; fE[jj] = e[c] - sh; }
            else { rH[jj] = hv; rE[jj] = e[c] - sh; }
        }
    }
}

// Both halves of a Myers-Miller node on the row path in one sweep, for nodes
// whose halves have at most 256 rows: lanes 0..31 sweep the forward half (rows
// 1..midi from corner (A,B), end penalty tb), lanes 32..63 the reverse half
// (rows M..midi+1 from corner (A+M,B+N), end penalty te), each half a single
// stripe of 32 lanes with R rows per lane (the row path's recurrence and
// dummy first register), so a node costs one ramp of 31 steps instead of two
// sweeps with 63-step ramps, at twice the rows per lane.  With R == 1 a half of
// P < 32 rows uses lanes 0..P-1 of its 32.  Lanes 0 and 32 take their half's
// top row from its formula; the last lane of each half writes the half's final
// row (H and E at each column, indexed by the half's own column) to fH/fE and
// rH/rE, the layout the two row-path sweeps leave.
template <int R, bool SH>
__device__ __attribute__((noinline)) void dual_halves(lcu8* tab, lcu8* s1, lcu8* s2, int A, int B, int M,
                                                      int N, int tb, int te, int go, int ge, int* fH,
                                                      int* fE, int* rH, int* rE)
{
    A = uni(A); B = uni(B); M = uni(M); N = uni(N);
    tb = uni(tb); te = uni(te); go = uni(go); ge = uni(ge);
    const int lane = lane_id();
    const int grp = lane >> 5, L = lane & 31;
    const int midi = M / 2;
    const int P = grp ? M - midi : midi;
    const int t0v = grp ? te : tb;
    const int rb = grp ? A + M + 1 : A, rd = grp ? -1 : 1;    // code of half-row i: s1[rb + rd*i]
    const int cb = grp ? B + N + 1 : B, cd = grp ? -1 : 1;    // 4*code of half-column j: s2[cb + cd*j]
    const int goe = go + ge;
    const int k1 = P - 32 * (R - 1);                           // lanes < k1 hold R rows
    const bool dummy = (R > 1) && (L >= k1);
    // R == 1: lanes 0..P-1 hold a row each, the others idle
    const int used = (R > 1) ? 32 : uni(M - midi);             // lanes of the taller half
    const bool lane_on = (R > 1) || (L < P);
    const int lastlane = (R > 1) ? 31 : P - 1;
    const int d = dummy ? 1 : 0;
    const int firstrow = 1 + (L < k1 ? L * R : k1 * R + (L - k1) * (R - 1));
    // SH: H, E and F of cell (i, j) carried as X + (i + j) * ge, as the
    // row-path stripes do (the gap recurrences lose their extension term)
    const int ge2 = 2 * ge;
    int sb[R], hL[R], hLg[R], f[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int row = firstrow + r - d;
        const int rr = row < 1 ? 1 : (row > P ? P : row);
        sb[r] = (int)s1[rb + rd * rr] * CW_TAB_ROW;
        hL[r] = row == 0 ? 0 : -t0v - row * ge;                // left column H[row][0]
        f[r] = hL[r] - go;
        hLg[r] = hL[r] - goe;
        if (SH) {
            const int sh = row * ge;
            f[r] += sh;
            hLg[r] = hL[r] + sh - go;
            hL[r] += sh + ge2;
        }
    }
    int dprev = (firstrow == 1) ? 0 : -t0v - (firstrow - 1) * ge;
    if (SH) dprev += (firstrow - 1) * ge + ge2;
    int outHv = 0, outEv = 0;
    int caddr = cb + cd * (1 - L);
    int b4 = s2[caddr];
    gint* const oH = (gint*)(grp ? rH : fH);
    gint* const oE = (gint*)(grp ? rE : fE);
    const bool top = (L == 0), bottom = (L == lastlane);

    auto step = [&](int t, auto masked) {
        // lanes 0 and 32: their half's top row H[0][t] = -t0 - t*ge, E = H - go
        const int topH = SH ? -t0v : -t0v - t * ge;
        int inHv = shr1(topH, outHv);
        int inEv = shr1(topH - go, outEv);
        if (top) {
            inHv = topH;
            inEv = topH - go;
        }
        const int bcur = b4;
        caddr += cd;
        b4 = s2[caddr];
        const int c = t - L;
        bool act = true;
        if constexpr (decltype(masked)::value) act = lane_on && c >= 1 && c <= N;
        if (act) {
            int hup = inHv, eup = inEv, dg = dprev;
            int hupg = SH ? inHv - go : inHv - goe;
            dprev = SH ? inHv + ge2 : inHv;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int sc = score_at(tab, sb[r] + bcur);
                int e = SH ? max(eup, hupg) : max(eup - ge, hupg);
                f[r] = SH ? max(f[r], hLg[r]) : max(f[r] - ge, hLg[r]);
                int h = max(max(dg + sc, e), f[r]);
                if (r == 0 && R > 1) {
                    if (dummy) { h = hup; e = eup; }
                }
                dg = hL[r];
                hL[r] = SH ? h + ge2 : h;
                hLg[r] = SH ? h - go : h - goe;
                hup = h;
                hupg = hLg[r];
                eup = e;
            }
            outHv = hup;
            outEv = eup;
            if (bottom) {
                const int sh = SH ? (P + c) * ge : 0;             // the bottom lane's row is P
                oH[c] = hup - sh;
                oE[c] = eup - sh;
            }
        }
    };
    // every lane on a valid column for t in [32, N] (when all 32 lanes hold rows)
    const int tend = N + used - 1;
    int t = 1;
    if (used == 32) {
        const int up_end = N < 31 ? tend : 31;
#pragma unroll 1
        for (; t <= up_end; ++t) step(t, Tag<true>{});
#pragma unroll 1
        for (; t + 1 <= N; t += 2) {
            step(t, Tag<false>{});
            step(t + 1, Tag<false>{});
        }
    }
#pragma unroll 1
    for (; t <= tend; ++t) step(t, Tag<true>{});
}

template <int C, bool SH>
__device__ __attribute__((noinline)) void halves_cols_call(lcu8* tab, lcu8* s1, lcu8* s2, int A, int B,
                                                           int M, int N, int tb, int te, int go, int ge,
                                                           int* fH, int* fE, int* rH, int* rE)
{
    mm_halves_cols<C, SH>(tab, s1, s2, uni(A), uni(B), uni(M), uni(N), uni(tb), uni(te), uni(go), uni(ge), fH,
                      fE, rH, rE);
}

// ---- small column-path nodes, two or four per sweep ---------------------------
// A node with few columns leaves most of a 32-lane half idle in
// mm_halves_cols (a ramp of nl - 1 steps around a few rows, one or two columns
// per lane).  The depth-first walk queues such nodes, G at a time (G = 4 for
// N <= 63, G = 2 for 64 <= N <= 127), and one sweep serves all G: lanes
// (64/G)k .. (64/G)(k+1)-1 take queued node k, the first half of them its
// forward half and the second half its reverse half, C columns per lane
// (N <= 32C/G), the recurrence of mm_halves_cols.  The final rows of node k go
// to F/R + 2 * (256/G) * k (H at [j], E at [256/G + j]).
constexpr int kQ4Cols = 63, kQ4Rows = 96;
constexpr int kQ2Cols = 127, kQ2Rows = 192;

__device__ __forceinline__ bool pw_q4(int M, int N, int colcap)
{
    return M >= 2 && M <= kQ4Rows && N >= 1 && N <= kQ4Cols && N <= colcap;
}
__device__ __forceinline__ bool pw_q2(int M, int N, int colcap)
{
    return M >= 2 && M <= kQ2Rows && N > kQ4Cols && N <= kQ2Cols && N <= colcap;
}

template <int G, int C, bool SH>
__device__ __attribute__((noinline)) void colsG_sweep(lcu8* tab, lcu8* s1, lcu8* s2, const int* q, int cnt,
                                                      int go, int ge, int* F, int* R)
{
    constexpr int LPH = 32 / G;                // lanes per half
    constexpr int STRIDE = 256 / G;
    const int lane = lane_id();
    const int k = lane / (2 * LPH), grp = (lane / LPH) & 1, L = lane % LPH;
    const bool on = k < cnt;
    const int* e = q + 6 * (on ? k : 0);
    const int A = e[0], B = e[1], M = on ? e[2] : 0, N = on ? e[3] : 0;
    const int midi = M / 2;
    const int rows = grp ? (M - midi) : midi;
    const int t0 = grp ? e[5] : e[4];
    const int goe = go + ge;
    const int rb = grp ? (A + M + 1) : A;      // code of half-row i: s1[rb + rd*i]
    const int rd = grp ? -1 : 1;
    const int nl = (N + C - 1) / C;            // lanes in use in this half
    int b4[C], hup[C], hupg[C], ev[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int jj = L * C + c + 1;          // half-local column
        const int jc = jj <= N ? jj : (N > 0 ? N : 1);
        b4[c] = s2[grp ? (B + N + 1 - jc) : (B + jc)];
        hup[c] = -t0 - jj * ge;                // top row H[0][jj]
        ev[c] = hup[c] - go;
        hupg[c] = hup[c] - goe;
        if (SH) {                              // cell (i, jj) carried as X + (i + jj) ge
            hupg[c] = -t0 - go;
            ev[c] = -t0 - go;
            hup[c] = -t0 + 2 * ge;             // stored with the diagonal's + 2 ge
        }
    }
    int dleft = (L == 0) ? 0 : (-t0 - (L * C) * ge);   // H[0][L*C]; corner is 0
    if (SH) dleft = (L == 0 ? 0 : -t0) + 2 * ge;
    int hl_out = 0, fl_out = 0;
    const int T = wave_max(rows > 0 && L < nl ? rows + nl - 1 : 0);
    for (int t = 1; t <= T; ++t) {
        int hl = shr1(0, hl_out);
        int fl = shr1(0, fl_out);
        const int i = t - L;
        if (L == 0) {                          // left column H[i][0], F[i][0]
            hl = SH ? -t0 : -t0 - i * ge;
            fl = hl - go;
        }
        if (i >= 1 && i <= rows && L < nl) {
            const int arow = (int)s1[rb + rd * i] * CW_TAB_ROW;
            int d = dleft;
            int hc = hl, fc = fl;
            int hcg = SH ? hl - go : 0;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int sc = score_at(tab, arow + b4[c]);
                if (SH) {
                    const int en = max(ev[c], hupg[c]);
                    const int fn = max(fc, hcg);
                    const int h = max(max(d + sc, en), fn);
                    d = hup[c];
                    hup[c] = h + 2 * ge;
                    hupg[c] = h - go;
                    hcg = hupg[c];
                    ev[c] = en;
                    hc = h;
                    fc = fn;
                } else {
                    const int en = max(ev[c] - ge, hupg[c]);
                    const int fn = max(fc - ge, hc - goe);
                    const int h = max(max(d + sc, en), fn);
                    d = hup[c];
                    hup[c] = h;
                    hupg[c] = h - goe;
                    ev[c] = en;
                    hc = h;
                    fc = fn;
                }
            }
            dleft = SH ? hl + 2 * ge : hl;
            hl_out = hc;
            fl_out = fc;
        }
    }
    int* oH = (grp ? R : F) + 2 * STRIDE * k;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int jj = L * C + c + 1;
        if (on && jj <= N) {
            const int sh = SH ? (rows + jj) * ge : 0;
            oH[jj] = SH ? hup[c] - 2 * ge - sh : hup[c];
            oH[STRIDE + jj] = ev[c] - sh;
        }
    }
}

// Reductions within each group of 64/G lanes; every lane of a group holds the
// result (rows of 16 by DPP, then the two rows of a 32-lane group by a swizzle).
template <int G, typename Op>
__device__ __forceinline__ int group_reduce(int x, Op op)
{
    x = op(x, __builtin_amdgcn_update_dpp(x, x, 0xB1, 0xf, 0xf, false));   // quad_perm [1,0,3,2]
    x = op(x, __builtin_amdgcn_update_dpp(x, x, 0x4E, 0xf, 0xf, false));   // quad_perm [2,3,0,1]
    x = op(x, __builtin_amdgcn_update_dpp(x, x, 0x141, 0xf, 0xf, false));  // row_half_mirror
    x = op(x, __builtin_amdgcn_update_dpp(x, x, 0x140, 0xf, 0xf, false));  // row_mirror
    if (G == 2) x = op(x, __shfl_xor(x, 16));
    return x;
}

// The midpoints of the queued nodes (:570-593), node k in lanes
// (64/G)k .. (64/G)(k+1)-1: returns type | midj << 2 in every lane of the group.
template <int G>
__device__ __attribute__((noinline)) int colsG_midpoints(const int* q, int cnt, int go, int ge,
                                                         const int* F, const int* R)
{
    constexpr int W = 64 / G;                  // lanes per node
    constexpr int STRIDE = 256 / G;
    const int lane = lane_id();
    const int k = lane / W, lw = lane % W;
    const bool on = k < cnt;
    const int* e = q + 6 * (on ? k : 0);
    const int M = on ? e[2] : 2, N = on ? e[3] : 1, tb = e[4], te = e[5];
    const int midi = M / 2;
    const int* fH = F + 2 * STRIDE * k;
    const int* fE = fH + STRIDE;
    const int* rH = R + 2 * STRIDE * k;
    const int* rE = rH + STRIDE;
    const int HH0 = -tb - midi * ge;
    const int RRN = -te - (M - midi) * ge;
    int vmax = -BIG, jfirst = BIG, jc = -1, wmax = -BIG, jw = -1;
#pragma unroll
    for (int it = 0; it < STRIDE / W; ++it) {
        const int j = lw + W * it;
        if (on && j <= N) {
            const int HHj = (j == 0) ? HH0 : fH[j];
            const int DDj = (j == 0) ? HH0 : fE[j];
            const int RRj = (j == N) ? RRN : rH[N - j];
            const int SSj = (j == N) ? RRN : rE[N - j];
            const int v = HHj + RRj;
            const int w = DDj + SSj + go;
            if (v > vmax) { vmax = v; jfirst = j; jc = -1; }
            if (v == vmax && HHj != DDj && RRj == SSj) jc = j;
            if (w >= wmax) { wmax = w; jw = j; }
        }
    }
    auto mx = [](int a, int b) { return a > b ? a : b; };
    auto mn = [](int a, int b) { return a < b ? a : b; };
    const int Mv = group_reduce<G>(vmax, mx);
    const int J0 = group_reduce<G>(vmax == Mv ? jfirst : BIG, mn);
    const int JC = group_reduce<G>(vmax == Mv ? jc : -1, mx);
    const int Wv = group_reduce<G>(wmax, mx);
    const int JW = group_reduce<G>(wmax == Wv ? jw : -1, mx);
    int type, midj;
    if (Wv > Mv) { type = 2; midj = JW; }
    else { type = 1; midj = J0 > JC ? J0 : JC; }
    return type | (midj << 2);
}

// ---- tiny subtrees, one lane each ---------------------------------------------
// A node with 2 <= M <= kTinyRows and N <= kTinyCols, and every M == 1 leaf, is
// not worth a wave: the depth-first walk queues such nodes (two ints each) and
// every lane then solves one of them with its whole subtree as scalar code:
// the columns unrolled in registers (the forward half's H/E and the reverse
// half's H/E kept at the same column index j), the midpoint as the
// reference's own scan over j = 0..N, the children on a small register stack.
// The end penalties of every node are 0 or the gap-open penalty, so a node
// packs into {A | B << 16, M | N << 5 | tb? << 30 | te? << 31}.

constexpr int kTinyRows = 16;
constexpr int kTinyCols = 16;
constexpr int kTinyStack = 6;

__device__ __forceinline__ bool pw_tiny(int M, int N)
{
    return M == 1 || (M >= 2 && M <= kTinyRows && N >= 1 && N <= kTinyCols);
}

__device__ __attribute__((noinline)) void pw_tiny_subtree(lcu8* tab, lcu8* s1, lcu8* s2, int go, int ge, int gap1,
                                                int gap2, int w0, int w1, int& count, int& cells,
                                                int& nodes, int& err)
{
    constexpr int NT = kTinyCols;
    const int goe = go + ge;
    int st0[kTinyStack], st1[kTinyStack];
#pragma unroll
    for (int k = 0; k < kTinyStack; ++k) st0[k] = st1[k] = 0;
    st0[0] = w0;
    st1[0] = w1;
    int sp = 1;
    while (sp > 0) {
        --sp;
        int e0 = 0, e1 = 0;
#pragma unroll
        for (int k = 0; k < kTinyStack; ++k)
            if (k == sp) { e0 = st0[k]; e1 = st1[k]; }
        const int A = e0 & 0xffff, B = (e0 >> 16) & 0xffff;
        const int M = e1 & 31, N = (e1 >> 5) & 0x1ffffff;
        const int tb = ((e1 >> 30) & 1) ? go : 0, te = ((e1 >> 31) & 1) ? go : 0;
        if (N <= 0 || M <= 0) continue;
        if (M == 1) {                                     // :435-478
            cells += N;
            const int tegN = te + ge * N, tbgN = tb + ge * N;
            int midh0 = -(tb + ge) - tegN;
            const int alt = -(te + ge) - tbgN;
            if (alt > midh0) midh0 = alt;
            const int a = s1[A + 1];
            int bv = -BIG, bj = 0;
            for (int j = 1; j <= N; ++j) {
                const int k1 = N - j, k2 = j - 1;
                const int v = score_at(tab, a * CW_TAB_ROW + s2[B + j]) - (k1 <= 0 ? 0 : te + ge * k1) -
                              (k2 <= 0 ? 0 : tb + ge * k2);
                if (v > bv) { bv = v; bj = j; }
            }
            if (bv > midh0) {
                const int r2 = s2[B + bj] >> 2;
                if (a != gap1 && r2 != gap2 && a == r2) ++count;
            }
            continue;
        }
        ++nodes;
        cells += M * N;
        const int midi = M / 2;
        // forward half: rows 1..midi from (A,B); H/E[k] hold column k
        int H[NT + 1], E[NT + 1];
#pragma unroll
        for (int k = 0; k <= NT; ++k) {
            H[k] = -tb - k * ge;
            E[k] = H[k] - go;
        }
        H[0] = 0;
        for (int i = 1; i <= midi; ++i) {
            const int arow = (int)s1[A + i] * CW_TAB_ROW;
            int d = H[0];
            int hc = -tb - i * ge, fc = hc - go;
            H[0] = hc;
#pragma unroll
            for (int k = 1; k <= NT; ++k) {
                if (k <= N) {
                    const int sc = score_at(tab, arow + s2[B + k]);
                    const int en = max(E[k] - ge, H[k] - goe);
                    const int fn = max(fc - ge, hc - goe);
                    const int h = max(max(d + sc, en), fn);
                    d = H[k];
                    H[k] = h;
                    E[k] = en;
                    hc = h;
                    fc = fn;
                }
            }
        }
        // reverse half: rows M..midi+1 from (A+M+1,B+N+1), columns N..1;
        // RH/RE[k] hold column k (reversed column N+1-k)
        int RH[NT + 2], RE[NT + 2];
#pragma unroll
        for (int k = 0; k <= NT + 1; ++k) {
            RH[k] = -te - (N + 1 - k) * ge;
            RE[k] = RH[k] - go;
        }
        int rprev = 0;                                   // H of the reverse left column, row r-1
        for (int r = 1; r <= M - midi; ++r) {
            const int arow = (int)s1[A + M + 1 - r] * CW_TAB_ROW;
            int d = rprev;
            int hc = -te - r * ge, fc = hc - go;
            rprev = hc;
#pragma unroll
            for (int k = NT; k >= 1; --k) {
                if (k <= N) {
                    const int sc = score_at(tab, arow + s2[B + k]);
                    const int en = max(RE[k] - ge, RH[k] - goe);
                    const int fn = max(fc - ge, hc - goe);
                    const int h = max(max(d + sc, en), fn);
                    d = RH[k];
                    RH[k] = h;
                    RE[k] = en;
                    hc = h;
                    fc = fn;
                }
            }
        }
        // midpoint (:570-593) in the reference's scan order
        const int HH0 = -tb - midi * ge;
        const int RRN = -te - (M - midi) * ge;
        int vmax = -BIG, jfirst = 0, jc = -1, wmax = -BIG, jw = -1;
#pragma unroll
        for (int j = 0; j <= NT; ++j) {
            if (j <= N) {
                const int HHj = (j == 0) ? HH0 : H[j];
                const int DDj = (j == 0) ? HH0 : E[j];
                const int RRj = (j == N) ? RRN : RH[j + 1];
                const int SSj = (j == N) ? RRN : RE[j + 1];
                const int v = HHj + RRj;
                const int w = DDj + SSj + go;
                if (v > vmax) { vmax = v; jfirst = j; jc = -1; }
                if (v == vmax && HHj != DDj && RRj == SSj) jc = j;
                if (w >= wmax) { wmax = w; jw = j; }
            }
        }
        int type, midj;
        if (wmax > vmax) { type = 2; midj = jw; }
        else { type = 1; midj = jfirst > jc ? jfirst : jc; }
        int c0a, c0b, c1a, c1b;                          // children, packed
        const int tbf = (e1 >> 30) & 1, tef = (e1 >> 31) & 1;
        if (type == 1) {
            c0a = A | (B << 16);
            c0b = midi | (midj << 5) | (tbf << 30) | (1 << 31);
            c1a = (A + midi) | ((B + midj) << 16);
            c1b = (M - midi) | ((N - midj) << 5) | (1 << 30) | (tef << 31);
        } else {
            c0a = A | (B << 16);
            c0b = (midi - 1) | (midj << 5) | (tbf << 30);
            c1a = (A + midi + 1) | ((B + midj) << 16);
            c1b = (M - midi - 1) | ((N - midj) << 5) | (tef << 31);
        }
        if (sp + 2 > kTinyStack) { err = 1; break; }
#pragma unroll
        for (int k = 0; k < kTinyStack; ++k) {
            if (k == sp) { st0[k] = c1a; st1[k] = c1b; }
            if (k == sp + 1) { st0[k] = c0a; st1[k] = c0b; }
        }
        sp += 2;
    }
}

__device__ __forceinline__ int wave_sum(int v) {
    return wave_reduce(v, [](int a, int b) { return a + b; });
}

// FullPairwiseAlign::diff (:418-612) restated as a depth-first walk; returns the
// identity count summed over the M==1 leaves (what tracePath counts).
__device__ int mm_count(const Sweep& base, lcu8* s1, lcu8* s2, int go, int ge,
                        int gap1, int gap2, int sb1, int sb2, int se1, int se2, int rmax,
                        const Bufs& g, int* g2H, int* g2E, int* g3H, int* g3E,
                        int* smallF, int* smallR, int small_cap, int colcap, int* stack,
                        int stack_cap, int* tq, int* status, int* nodes_out, int* cells_out)
{
    const int lane = lane_id();
    int count = 0, sp = 0, nodes = 0, cells = 0;
    // per-lane totals of the tiny subtrees, summed over the wave at the end
    int lcount = 0, lcells = 0, lnodes = 0, lerr = 0, nq = 0;
    auto drain = [&]() {
        if (lane < nq)
            pw_tiny_subtree(base.tab, s1, s2, go, ge, gap1, gap2, tq[2 * lane], tq[2 * lane + 1], lcount,
                            lcells, lnodes, lerr);
        nq = 0;
        wave_fence();
    };
    // the top six stack entries hold the queues of small column-path nodes
    // (four entries for N <= 63, two for 64 <= N <= 127)
    // queued nodes always run in the shifted frame: (i + j) * ge < 320 * 2^16
    const bool small_ge = ge > -65536 && ge < 65536;
    int* const q4 = stack + 6 * (stack_cap - 6);
    int* const q2 = stack + 6 * (stack_cap - 2);
    const int cap = stack_cap - 6;
    int n4 = 0, n2 = 0;
    bool bad = false;
    auto flushG = [&](auto gtag) {
        constexpr int G = decltype(gtag)::value;
        int* const q = G == 4 ? q4 : q2;
        const int cnt = G == 4 ? n4 : n2;
        int mN = 1;
        for (int kk = 0; kk < cnt; ++kk) mN = max(mN, uni(q[6 * kk + 3]));
        constexpr int LPH = 32 / G;
        switch ((mN + LPH - 1) / LPH) {                  // columns per lane: N <= C * lanes per half
        case 1: colsG_sweep<G, 1, true>(base.tab, s1, s2, q, cnt, go, ge, smallF, smallR); break;
        case 2: colsG_sweep<G, 2, true>(base.tab, s1, s2, q, cnt, go, ge, smallF, smallR); break;
        case 3: colsG_sweep<G, 3, true>(base.tab, s1, s2, q, cnt, go, ge, smallF, smallR); break;
        case 4: colsG_sweep<G, 4, true>(base.tab, s1, s2, q, cnt, go, ge, smallF, smallR); break;
        case 5: colsG_sweep<G, 5, true>(base.tab, s1, s2, q, cnt, go, ge, smallF, smallR); break;
        case 6: colsG_sweep<G, 6, true>(base.tab, s1, s2, q, cnt, go, ge, smallF, smallR); break;
        case 7: colsG_sweep<G, 7, true>(base.tab, s1, s2, q, cnt, go, ge, smallF, smallR); break;
        default: colsG_sweep<G, 8, true>(base.tab, s1, s2, q, cnt, go, ge, smallF, smallR); break;
        }
        wave_fence();
        const int tm = colsG_midpoints<G>(q, cnt, go, ge, smallF, smallR);
        for (int kk = 0; kk < cnt; ++kk) {
            const int v = __builtin_amdgcn_readlane(tm, (64 / G) * kk);
            const int type = v & 3, midj = v >> 2;
            const int* e = q + 6 * kk;
            const int A = uni(e[0]), B = uni(e[1]), M = uni(e[2]), N = uni(e[3]);
            const int tb = uni(e[4]), te = uni(e[5]);
            const int midi = M / 2;
            if (sp + 2 > cap) { bad = true; break; }
            int* c0 = stack + 6 * sp;                    // every lane stores the same values
            int* c1 = stack + 6 * (sp + 1);
            if (type == 1) {
                c0[0] = A; c0[1] = B; c0[2] = midi; c0[3] = midj; c0[4] = tb; c0[5] = go;
                c1[0] = A + midi; c1[1] = B + midj; c1[2] = M - midi; c1[3] = N - midj; c1[4] = go; c1[5] = te;
            } else {
                c0[0] = A; c0[1] = B; c0[2] = midi - 1; c0[3] = midj; c0[4] = tb; c0[5] = 0;
                c1[0] = A + midi + 1; c1[1] = B + midj; c1[2] = M - midi - 1; c1[3] = N - midj; c1[4] = 0; c1[5] = te;
            }
            sp += 2;
        }
        if (G == 4) n4 = 0; else n2 = 0;
        wave_fence();
    };
    using G4 = std::integral_constant<int, 4>;
    using G2 = std::integral_constant<int, 2>;
    {
        int* e = stack;                                  // every lane stores the same values
        e[0] = sb1 - 1; e[1] = sb2 - 1; e[2] = se1 - sb1 + 1; e[3] = se2 - sb2 + 1; e[4] = 0; e[5] = 0;
    }
    sp = 1;
    wave_fence();
    while (sp > 0 || n4 > 0 || n2 > 0) {
        if (sp == 0) {                                   // only queued nodes left
            if (n4 > 0) flushG(G4{});
            else flushG(G2{});
            if (bad) { atomicOr(status, 2); break; }
            continue;
        }
        --sp;
        const int* e = stack + 6 * sp;
        const int A = uni(e[0]), B = uni(e[1]), M = uni(e[2]), N = uni(e[3]);
        const int tb = uni(e[4]), te = uni(e[5]);
        wave_fence();
        if (N <= 0 || M <= 0) continue;                 // only del()/add(): no identities
        if (pw_tiny(M, N) && A < 0x10000 && B < 0x10000) {
            tq[2 * nq] = A | (B << 16);                  // same value from every lane
            tq[2 * nq + 1] = M | (N << 5) | ((tb != 0) << 30) | ((te != 0) << 31);
            nq += 1;
            wave_fence();
            if (nq == CW_TINYQ) drain();
            continue;
        }
        if (M == 1) {                                     // :435-478
            cells += N;
            const int tegN = te + ge * N;                 // N >= 1
            const int tbgN = tb + ge * N;
            int midh0 = -(tb + ge) - tegN;
            const int alt = -(te + ge) - tbgN;
            if (alt > midh0) midh0 = alt;
            const int a = s1[A + 1];
            int bv = -BIG, bj = BIG;
            for (int j = 1 + lane; j <= N; j += WAVE) {
                const int k1 = N - j, k2 = j - 1;
                const int v = score_at(base.tab, a * CW_TAB_ROW + s2[B + j]) - (k1 <= 0 ? 0 : te + ge * k1) -
                              (k2 <= 0 ? 0 : tb + ge * k2);
                if (v > bv) { bv = v; bj = j; }
            }
            const int V = wave_max(bv);
            const int J = wave_min(bv == V ? bj : BIG);
            if (V > midh0) {
                const int r2 = s2[B + J] >> 2;
                if (a != gap1 && r2 != gap2 && a == r2) ++count;
            }
            continue;
        }
        ++nodes;
        cells += M * N;
        if (pw_q4(M, N, colcap) && small_ge) {
            int* qe = q4 + 6 * n4;                       // same value from every lane
            qe[0] = A; qe[1] = B; qe[2] = M; qe[3] = N; qe[4] = tb; qe[5] = te;
            n4 += 1;
            wave_fence();
            if (n4 == 4) {
                flushG(G4{});
                if (bad) { atomicOr(status, 2); break; }
            }
            continue;
        }
        if (pw_q2(M, N, colcap) && small_ge) {
            int* qe = q2 + 6 * n2;                       // same value from every lane
            qe[0] = A; qe[1] = B; qe[2] = M; qe[3] = N; qe[4] = tb; qe[5] = te;
            n2 += 1;
            wave_fence();
            if (n2 == 2) {
                flushG(G2{});
                if (bad) { atomicOr(status, 2); break; }
            }
            continue;
        }
        const int midi = M / 2;
        int *fH, *fE, *rH, *rE;          // HH/DD[j] at [j]; RR/SS[j] at [N-j]
        // shifted frame while (i + j) * ge stays small beside the scores
        const bool shf = (int64_t)(M + N + 2) * (ge < 0 ? -(int64_t)ge : (int64_t)ge) < (1 << 26);
        if (N <= colcap) {
            fH = smallF; fE = smallF + (small_cap + 1);
            rH = smallR; rE = smallR + (small_cap + 1);
            // C = ceil(N / 32) columns per lane: every lane of a half in use
            switch ((N + 31) >> 5) {
            case 1: if (shf) halves_cols_call<1, true>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE);
                    else halves_cols_call<1, false>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE); break;
            case 2: if (shf) halves_cols_call<2, true>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE);
                    else halves_cols_call<2, false>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE); break;
            case 3: if (shf) halves_cols_call<3, true>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE);
                    else halves_cols_call<3, false>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE); break;
            case 4: if (shf) halves_cols_call<4, true>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE);
                    else halves_cols_call<4, false>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE); break;
            case 5: if (shf) halves_cols_call<5, true>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE);
                    else halves_cols_call<5, false>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE); break;
            case 6: if (shf) halves_cols_call<6, true>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE);
                    else halves_cols_call<6, false>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE); break;
            case 7: if (shf) halves_cols_call<7, true>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE);
                    else halves_cols_call<7, false>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE); break;
            default: if (shf) halves_cols_call<8, true>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE);
                     else halves_cols_call<8, false>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE); break;
            }
        } else if (M - midi <= 32 * CW_RMAX) {
            fH = g2H; fE = g2E;
            rH = g3H; rE = g3E;
            switch ((M - midi + 31) >> 5) {               // rows per lane of the taller half
            case 1: if (shf) dual_halves<1, true>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE);
                    else dual_halves<1, false>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE); break;
            case 2: if (shf) dual_halves<2, true>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE);
                    else dual_halves<2, false>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE); break;
            case 3: if (shf) dual_halves<3, true>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE);
                    else dual_halves<3, false>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE); break;
            case 4: if (shf) dual_halves<4, true>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE);
                    else dual_halves<4, false>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE); break;
            case 5: if (shf) dual_halves<5, true>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE);
                    else dual_halves<5, false>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE); break;
            case 6: if (shf) dual_halves<6, true>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE);
                    else dual_halves<6, false>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE); break;
            case 7: if (shf) dual_halves<7, true>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE);
                    else dual_halves<7, false>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE); break;
            default: if (shf) dual_halves<8, true>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE);
                     else dual_halves<8, false>(base.tab, s1, s2, A, B, M, N, tb, te, go, ge, fH, fE, rH, rE); break;
            }
        } else {
            Track T{0, 0};
            Sweep F = base;
            F.nrows = midi; F.ncols = N;
            F.rs = s1; F.rb = A; F.rd = 1;
            F.cs = s2; F.cb = B; F.cd = 1;
            F.th0 = -tb; F.ths = -ge; F.teo = -go;
            F.lh0 = -tb; F.lhs = -ge; F.lfo = -go;
            fH = g2H; fE = g2E;
            if (shf) sweep<MODE_GLB, true>(F, rmax, g, fH, fE, T);
            else sweep<MODE_GLB, false>(F, rmax, g, fH, fE, T);
            Sweep Rv = base;
            Rv.nrows = M - midi; Rv.ncols = N;
            Rv.rs = s1; Rv.rb = A + M + 1; Rv.rd = -1;
            Rv.cs = s2; Rv.cb = B + N + 1; Rv.cd = -1;
            Rv.th0 = -te; Rv.ths = -ge; Rv.teo = -go;
            Rv.lh0 = -te; Rv.lhs = -ge; Rv.lfo = -go;
            rH = g3H; rE = g3E;
            if (shf) sweep<MODE_GLB, true>(Rv, rmax, g, rH, rE, T);
            else sweep<MODE_GLB, false>(Rv, rmax, g, rH, rE, T);
        }
        wave_fence();
        // midpoint (:570-593)
        const int HH0 = -tb - midi * ge;
        const int RRN = -te - (M - midi) * ge;
        int vmax = -BIG, jfirst = BIG, jc = -1, wmax = -BIG, jw = -1;
        for (int j = lane; j <= N; j += WAVE) {
            const int HHj = (j == 0) ? HH0 : fH[j];
            const int DDj = (j == 0) ? HH0 : fE[j];
            const int RRj = (j == N) ? RRN : rH[N - j];
            const int SSj = (j == N) ? RRN : rE[N - j];
            const int v = HHj + RRj;
            const int w = DDj + SSj + go;
            if (v > vmax) { vmax = v; jfirst = j; jc = -1; }
            if (v == vmax && HHj != DDj && RRj == SSj) jc = j;
            if (w >= wmax) { wmax = w; jw = j; }
        }
        const int Mv = wave_max(vmax);
        const int J0 = wave_min(vmax == Mv ? jfirst : BIG);
        const int JC = wave_max(vmax == Mv ? jc : -1);
        const int Wv = wave_max(wmax);
        const int JW = wave_max(wmax == Wv ? jw : -1);
        int type, midj;
        if (Wv > Mv) { type = 2; midj = JW; }
        else { type = 1; midj = J0 > JC ? J0 : JC; }
        midj = uni(midj);
        type = uni(type);
        wave_fence();
        if (sp + 2 > cap) { atomicOr(status, 2); break; }
        {
            int* c0 = stack + 6 * sp;
            int* c1 = stack + 6 * (sp + 1);
            if (type == 1) {
                c0[0] = A; c0[1] = B; c0[2] = midi; c0[3] = midj; c0[4] = tb; c0[5] = go;
                c1[0] = A + midi; c1[1] = B + midj; c1[2] = M - midi; c1[3] = N - midj; c1[4] = go; c1[5] = te;
            } else {
                c0[0] = A; c0[1] = B; c0[2] = midi - 1; c0[3] = midj; c0[4] = tb; c0[5] = 0;
                c1[0] = A + midi + 1; c1[1] = B + midj; c1[2] = M - midi - 1; c1[3] = N - midj; c1[4] = 0; c1[5] = te;
            }
        }
        sp += 2;
        wave_fence();
    }
    drain();
    count += wave_sum(lcount);
    nodes += wave_sum(lnodes);
    cells += wave_sum(lcells);
    if (wave_sum(lerr) != 0) atomicOr(status, 2);
    *nodes_out = nodes;
    *cells_out = cells;
    return count;
}

// dst[-pad .. n+pad]: residue k (1-based) at dst[k] as scale*code; padding = 0
__device__ void load_seq(lu8* dst, const uint8_t* src, int n, int pad, int scale)
{
    const int lane = lane_id();
    for (int k = lane - pad; k <= n + pad; k += WAVE) {
        int v = 0;
        if (k >= 1 && k <= n) v = src[k - 1] * scale;
        dst[k] = (uint8_t)v;
    }
}

// Out of line, and every store and atomic below is made by all 64 lanes (the
// same value from each): a lane-subset block just before a return lets the
// structurizer give lanes 1..63 a loop of their own in the caller.  In the
// persistent loop that compiled to lanes 1..63 taking work item 0 and entering
// the next pair without lane 0 after an empty sequence (the DP then reads
// Myers-Miller stack entries that only lane 0 would have written).
// Bits that hold 1..v (v >= 1).
__device__ __forceinline__ int bits_for(int v) { return 32 - __builtin_clz((unsigned)v); }

// Upper bound of every forward score of the pair: an alignment pairs each
// residue of a sequence at most once and (with non-negative gap penalties)
// gaps only subtract, so the score is at most the sum over seq1 of the
// largest entry of each residue's matrix row, and likewise over seq2's
// columns.  rcmax: lane c < 32 holds max(row c), lane 32 + c max(column c),
// clamped to [0, 32767] (the caller refuses packing when a clamp bit), so the
// per-lane sums (<= 512 residues each) and their total stay below 2^31.
__device__ int score_bound(lcu8* s1, lcu8* s2, int n, int m, int rcmax)
{
    const int lane = lane_id();
    int a1 = 0, a2 = 0;
    for (int b = 0; b < n || b < m; b += WAVE) {        // uniform trip count
        const int k = b + 1 + lane;
        const int c1 = k <= n ? (int)s1[k] : 0;
        const int c2 = k <= m ? ((int)s2[k] >> 2) : 0;
        const int r1 = __shfl(rcmax, c1);
        const int r2 = __shfl(rcmax, 32 + c2);
        a1 += k <= n ? r1 : 0;
        a2 += k <= m ? r2 : 0;
    }
    const int S1 = wave_sum(a1), S2 = wave_sum(a2);
    return S1 < S2 ? S1 : S2;
}

// rcmax: row and column maxima of the matrix (see score_bound), rcfit: none
// of them was clamped.
__device__ __attribute__((noinline)) void do_pair(const DevArgs& A, int64_t p, int slot, lcu8* tab,
                                                  lu8* wlds, int rcmax, int rcfit)
{
    const int si = uni(A.pairs[2 * p]), sj = uni(A.pairs[2 * p + 1]);
    const bool in_set = (unsigned)si < (unsigned)A.nseq && (unsigned)sj < (unsigned)A.nseq;
    const int n = in_set ? uni(A.lens[si]) : -1, m = in_set ? uni(A.lens[sj]) : -1;
    if (n < 0 || m < 0 || n > A.maxlen || m > A.maxlen) {   // outside what the buffers cover
        A.out_count[p] = -2;
        if (A.out_info)
            for (int k = 0; k < 8; ++k) A.out_info[8 * p + k] = 0;
        atomicOr(A.status, CW_ST_BADPAIR);
        return;
    }
    if (n == 0 || m == 0) {                         // FullPairwiseAlign.cpp:113-118
        A.out_count[p] = -1;
        if (A.out_info)
            for (int k = 0; k < 8; ++k) A.out_info[8 * p + k] = 0;
        return;
    }
    const int pad = CW_SEQ_PAD;
    lu8* s1 = wlds + pad;                  // codes of seq1
    lu8* s2 = wlds + A.lcap + pad;         // 4 * codes of seq2 (byte offsets into a matrix row)
    int* smallF = (int*)(uint8_t*)(wlds + 2 * A.lcap);
    int* smallR = smallF + 2 * (CW_SMALL_COLS + 1);
    int* stack = smallR + 2 * (CW_SMALL_COLS + 1);
    int* tq = stack + 6 * CW_STACK_CAP;
    load_seq(s1, A.codes + A.offs[si], n, pad, 1);
    load_seq(s2, A.codes + A.offs[sj], m, pad, 4);
    wave_fence();

    int* scr = A.scratch + (int64_t)slot * A.scratch_stride;
    const int rc = A.rowcap;
    const Bufs g{scr, scr + rc, scr + 2 * rc, scr + 3 * rc};
    int* g2H = scr + 4 * rc;
    int* g2E = scr + 5 * rc;
    int* g3H = scr + 6 * rc;
    int* g3E = scr + 7 * rc;

    const int go = A.go_by_len[n < m ? n : m];
    const int ge = A.ge;
    const int cls = cw_class_of(n, m);
    const int rmax = uni(A.rmax_cls[cls]);
    const int colcap = uni(A.colcap_cls[cls]);
    Sweep S;
    S.tab = tab;
    S.go = go; S.ge = ge; S.goe = go + ge;
    S.thr = 0;
    S.kb = 0;

    // forward local pass (:280-341)
    Sweep F = S;
    F.nrows = n; F.ncols = m;
    F.rs = s1; F.rb = 0; F.rd = 1;
    F.cs = s2; F.cb = 0; F.cd = 1;
    F.th0 = 0; F.ths = 0; F.teo = -go;
    F.lh0 = 0; F.lhs = 0; F.lfo = -go;
    // packed tracking when every forward score (at most score_bound with
    // non-negative gap penalties) fits above the column bits
    F.kb = bits_for(m);
    const bool packF = go >= 0 && ge >= 0 && rcfit &&
        uni(score_bound(s1, s2, n, m, rcmax)) < (1 << (31 - F.kb));
    Track T{0, 0};
    if (packF) sweep<MODE_FWD, true>(F, rmax, g, nullptr, nullptr, T);
    else sweep<MODE_FWD, false>(F, rmax, g, nullptr, nullptr, T);
    const int maxScore = uni(T.best);
    int se1 = 0, se2 = 0, sb1 = 1, sb2 = 1, status = 0;
    if (maxScore > 0) {
        se1 = uni(T.key >> 16);
        se2 = uni(T.key & 0xffff);
        // reverse pass (:343-416) in reversed coordinates
        Sweep Rv = S;
        Rv.nrows = se1; Rv.ncols = se2;
        Rv.rs = s1; Rv.rb = se1 + 1; Rv.rd = -1;
        Rv.cs = s2; Rv.cb = se2 + 1; Rv.cd = -1;
        Rv.th0 = -1; Rv.ths = 0; Rv.teo = 0;
        Rv.lh0 = -1; Rv.lhs = 0; Rv.lfo = 0;
        Rv.thr = maxScore;
        Track U{0, 0};
        sweep<MODE_REV, false>(Rv, rmax, g, nullptr, nullptr, U);
        if (U.best) {
            const int r = uni(U.key >> 16), c = uni(U.key & 0xffff);
            sb1 = se1 - r + 1;
            sb2 = se2 - c + 1;
        } else {
            status |= 1;   // cannot happen when maxScore > 0; reported, never guessed
        }
    }
    int nodes = 0, mmcells = 0;
    int count = 0;
    if (se1 - sb1 + 1 > 0 && se2 - sb2 + 1 > 0 && maxScore > 0) {
        count = mm_count(S, s1, s2, go, ge, A.gap1, A.gap2, sb1, sb2, se1, se2, rmax, g, g2H,
                         g2E, g3H, g3E, smallF, smallR, CW_SMALL_COLS, colcap, stack, CW_STACK_CAP, tq,
                         A.status, &nodes, &mmcells);
    }
    A.out_count[p] = count;
    if (A.out_info) {
        int* info = A.out_info + 8 * p;
        info[0] = maxScore; info[1] = se1; info[2] = se2; info[3] = sb1; info[4] = sb2;
        info[5] = status; info[6] = nodes; info[7] = mmcells;
    }
    if (status) atomicOr(A.status, status);
}

__device__ void load_table(const DevArgs& A, CW_LDS int* tab)
{
    for (int k = threadIdx.x; k < 32 * 32; k += blockDim.x) tab[(k >> 5) * CW_TAB_STRIDE + (k & 31)] = A.matrix[k];
    __syncthreads();
}

// Row maxima (lanes 0..31) and column maxima (lanes 32..63) of the matrix,
// clamped to [0, 32767]; every lane of the wave active.  *fit: no clamp from above.
__device__ int table_rcmax(lcu8* tab, int* fit)
{
    const int lane = lane_id();
    const int c = lane & 31;
    const CW_LDS int* t = (const CW_LDS int*)tab;
    int v = (int)0x80000000;
    for (int k = 0; k < 32; ++k) v = max(v, lane < 32 ? t[c * CW_TAB_STRIDE + k] : t[k * CW_TAB_STRIDE + c]);
    *fit = uni(wave_max(v)) <= 32767;
    return v < 0 ? 0 : (v > 32767 ? 32767 : v);
}

} // namespace cwk

using namespace cwk;

extern "C" __global__ void __launch_bounds__(CW_MAX_THREADS, 4) cw_pairwise_plain(DevArgs A)
{
    extern __shared__ __align__(16) uint8_t lds_raw[];
    lu8* lds = (lu8*)lds_raw;
    load_table(A, (CW_LDS int*)lds);
    const int wave = threadIdx.x / WAVE;
    lcu8* tab = lds;
    lu8* wlds = lds + CW_TAB_BYTES + wave * A.wave_lds;
    const int64_t slot = (int64_t)blockIdx.x * (blockDim.x / WAVE) + wave;
    const int64_t p = A.pair_base + slot;
    if (p >= A.pair_end) return;
    const int64_t q = A.order ? A.order[p] : p;
    int rcfit;
    const int rcmax = table_rcmax(tab, &rcfit);
    do_pair(A, q, (int)slot, tab, wlds, rcmax, rcfit);
}

extern "C" __global__ void __launch_bounds__(CW_MAX_THREADS, 4) cw_pairwise_persistent(DevArgs A)
{
    extern __shared__ __align__(16) uint8_t lds_raw[];
    lu8* lds = (lu8*)lds_raw;
    load_table(A, (CW_LDS int*)lds);
    const int wave = threadIdx.x / WAVE;
    lcu8* tab = lds;
    lu8* wlds = lds + CW_TAB_BYTES + wave * A.wave_lds;
    const int slot = blockIdx.x * (blockDim.x / WAVE) + wave;
    int rcfit;
    const int rcmax = table_rcmax(tab, &rcfit);
    // this launch's slice of the binned order (the binning kernels left the
    // long part's size and the total after the cursors) and its own counter
    int qb = 0, qe = (int)A.pair_end;
    unsigned int* qc = A.queue;
    if (A.qpart) {
        const int nl = uni(A.bins[2 * CW_QBINS]), tot = uni(A.bins[2 * CW_QBINS + 1]);
        qb = A.qpart == 1 ? 0 : nl;
        qe = A.qpart == 1 ? nl : tot;
        qc = A.queue + (A.qpart - 1);
    }
    for (;;) {
        // one atomic per pop (lane 0's); every lane reaches the broadcast, where
        // lane 0 is the first active lane
        unsigned got = 0;
        if (lane_id() == 0) got = atomicAdd(qc, 1u);
        const int item = qb + uni((int)got);
        if (item >= qe) break;
        const int64_t q = A.order ? A.order[item] : item;
        do_pair(A, q, slot, tab, wlds, rcmax, rcfit);
    }
}

// ---- host-side launchers (used by cwhip_api.cpp) -------------------------
hipError_t cw_launch_plain(int blocks, int threads, size_t lds, hipStream_t st, const DevArgs& a)
{
    hipLaunchKernelGGL(cw_pairwise_plain, dim3(blocks), dim3(threads), lds, st, a);
    return hipGetLastError();
}

hipError_t cw_launch_persistent(int blocks, int threads, size_t lds, hipStream_t st, const DevArgs& a)
{
    hipLaunchKernelGGL(cw_pairwise_persistent, dim3(blocks), dim3(threads), lds, st, a);
    return hipGetLastError();
}

extern "C" __global__ void cw_reset_counter(unsigned int* q, int32_t* status)
{
    if (threadIdx.x == 0) { *q = 0u; *status = 0; }
}

hipError_t cw_launch_reset(hipStream_t st, unsigned int* q, int32_t* status)
{
    hipLaunchKernelGGL(cw_reset_counter, dim3(1), dim3(64), 0, st, q, status);
    return hipGetLastError();
}

// ---- length-binned work queue of the batched kernel --------------------------
// A counting sort of the pairs by cw_bin_of (4 bins per octave of n*m), largest
// bins first, built on the device in three launches: count, scan, scatter.
// Pairs whose longer sequence exceeds DevArgs::lsplit take bins of their own
// above all others, so they form the front of the order (the long part) and
// can run in a launch of their own with per-wave sequence buffers that hold
// them, the rest in a launch whose smaller buffers leave room for more waves.
// Within a bin the order is whatever the atomics give; every pair's result is
// independent of the order, so the output does not depend on it.

constexpr int kBinItems = 1024;       // pairs per workgroup in count/scatter

extern "C" __global__ void __launch_bounds__(256) cw_bin_reset(DevArgs A)
{
    for (int k = threadIdx.x; k < CW_BINS_INTS; k += blockDim.x) A.bins[k] = 0;
    if (threadIdx.x == 0) { A.queue[0] = 0u; A.queue[1] = 0u; *A.status = 0; }
}

__device__ __forceinline__ int pair_bin(const DevArgs& A, int64_t p)
{
    const int si = A.pairs[2 * p], sj = A.pairs[2 * p + 1];
    if ((unsigned)si >= (unsigned)A.nseq || (unsigned)sj >= (unsigned)A.nseq) return 0;  // refused by do_pair
    const int n = A.lens[si], m = A.lens[sj];
    const int b = cw_bin_of(n < 0 ? 0 : n > 0xffff ? 0xffff : n, m < 0 ? 0 : m > 0xffff ? 0xffff : m);
    return (n > A.lsplit || m > A.lsplit) ? b + CW_NBINS : b;
}

extern "C" __global__ void __launch_bounds__(256) cw_bin_count(DevArgs A)
{
    __shared__ int cnt[CW_QBINS];
    for (int k = threadIdx.x; k < CW_QBINS; k += blockDim.x) cnt[k] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kBinItems;
    for (int k = threadIdx.x; k < kBinItems; k += blockDim.x) {
        const int64_t p = base + k;
        if (p < A.pair_end) atomicAdd(&cnt[pair_bin(A, p)], 1);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < CW_QBINS; k += blockDim.x)
        if (cnt[k]) atomicAdd(&A.bins[k], cnt[k]);
}

// cursors: bin CW_QBINS-1 starts at 0, each smaller bin after the larger ones;
// then the long part's size and the total
extern "C" __global__ void __launch_bounds__(64) cw_bin_scan(DevArgs A)
{
    if (threadIdx.x == 0) {
        int run = 0;
        for (int b = CW_QBINS - 1; b >= 0; --b) {
            if (b == CW_NBINS - 1) A.bins[2 * CW_QBINS] = run;
            A.bins[CW_QBINS + b] = run;
            run += A.bins[b];
        }
        A.bins[2 * CW_QBINS + 1] = run;
    }
}

extern "C" __global__ void __launch_bounds__(256) cw_bin_scatter(DevArgs A)
{
    __shared__ int cnt[CW_QBINS], at[CW_QBINS];
    for (int k = threadIdx.x; k < CW_QBINS; k += blockDim.x) cnt[k] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kBinItems;
    constexpr int per = kBinItems / 256;
    int rank[per], bin[per];
#pragma unroll
    for (int u = 0; u < per; ++u) {
        const int64_t p = base + threadIdx.x + 256 * u;
        bin[u] = -1;
        if (p < A.pair_end) {
            bin[u] = pair_bin(A, p);
            rank[u] = atomicAdd(&cnt[bin[u]], 1);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < CW_QBINS; k += blockDim.x)
        at[k] = cnt[k] ? atomicAdd(&A.bins[CW_QBINS + k], cnt[k]) : 0;
    __syncthreads();
#pragma unroll
    for (int u = 0; u < per; ++u)
        if (bin[u] >= 0) A.order_out[at[bin[u]] + rank[u]] = (int32_t)(base + threadIdx.x + 256 * u);
}

// The same counting sort by one workgroup for a batch of at most
// kBinSmallPairs pairs (counts and cursors in LDS): one launch instead of four.
constexpr int kBinSmallPairs = 65536;

extern "C" __global__ void __launch_bounds__(1024) cw_bin_small(DevArgs A)
{
    __shared__ int cnt[CW_QBINS], at[CW_QBINS];
    for (int k = threadIdx.x; k < CW_QBINS; k += blockDim.x) cnt[k] = 0;
    if (threadIdx.x == 0) { A.queue[0] = 0u; A.queue[1] = 0u; *A.status = 0; }
    __syncthreads();
    for (int64_t p = threadIdx.x; p < A.pair_end; p += blockDim.x) atomicAdd(&cnt[pair_bin(A, p)], 1);
    __syncthreads();
    if (threadIdx.x == 0) {                 // bin CW_QBINS-1 first
        int run = 0;
        for (int b = CW_QBINS - 1; b >= 0; --b) {
            if (b == CW_NBINS - 1) A.bins[2 * CW_QBINS] = run;
            at[b] = run;
            run += cnt[b];
        }
        A.bins[2 * CW_QBINS + 1] = run;
    }
    __syncthreads();
    for (int64_t p = threadIdx.x; p < A.pair_end; p += blockDim.x)
        A.order_out[atomicAdd(&at[pair_bin(A, p)], 1)] = (int32_t)p;
}

// The batched run's stages on one stream: reset, count, scan, scatter (or
// cw_bin_small for a small batch), then the persistent kernel over the binned
// order (captured into one hipGraph by the host).  With `s` (a.lsplit below the
// longest sequence), the persistent kernel runs twice: over the long part
// with a's shape, then over the short part with s's (smaller sequence buffers,
// more waves per CU).
static hipError_t launch_bins(hipStream_t st, const DevArgs& a)
{
    const int nb = (int)((a.pair_end + kBinItems - 1) / kBinItems);
    if (a.pair_end <= kBinSmallPairs) {
        hipLaunchKernelGGL(cw_bin_small, dim3(1), dim3(1024), 0, st, a);
    } else {
        hipLaunchKernelGGL(cw_bin_reset, dim3(1), dim3(256), 0, st, a);
        hipLaunchKernelGGL(cw_bin_count, dim3(nb), dim3(256), 0, st, a);
        hipLaunchKernelGGL(cw_bin_scan, dim3(1), dim3(64), 0, st, a);
        hipLaunchKernelGGL(cw_bin_scatter, dim3(nb), dim3(256), 0, st, a);
    }
    return hipGetLastError();
}

hipError_t cw_launch_batched(int blocks, int threads, size_t lds, hipStream_t st, const DevArgs& a,
                             const CwSplitShape* s)
{
    hipError_t e = launch_bins(st, a);
    if (e != hipSuccess) return e;
    DevArgs b = a;
    b.order = a.order_out;
    b.qpart = s ? 1 : 0;
    hipLaunchKernelGGL(cw_pairwise_persistent, dim3(blocks), dim3(threads), lds, st, b);
    e = hipGetLastError();
    if (e != hipSuccess || !s) return e;
    DevArgs c = s->args;
    c.order = a.order_out;
    c.qpart = 2;
    hipLaunchKernelGGL(cw_pairwise_persistent, dim3(s->blocks), dim3(s->threads), s->lds, st, c);
    return hipGetLastError();
}

hipError_t cw_occupancy(int threads, size_t lds, int* plain_blocks, int* pers_blocks)
{
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(plain_blocks, cw_pairwise_plain, threads, lds);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(pers_blocks, cw_pairwise_persistent, threads, lds);
}
