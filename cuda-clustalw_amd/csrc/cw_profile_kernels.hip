// cw_profile_kernels.hip -- gfx950 kernels of ClustalW's profile-profile
// Myers-Miller alignment (MyersMillerProfileAlign::progDiff,
// multipleAlign/MyersMillerProfileAlign.cpp:521-811 of CPU/clustalw-2.0.11/src).
//
// Two kernels per batch of profile pairs:
//   cw_prf_score  S[J][I] = prfScore(I, J) (:1002-1016): a 27-term integer dot
//                 product, C truncation by 10; 64x64 tiles, profile-2 rows in LDS.
//   cw_prf_mm     one workgroup per pair runs the Myers-Miller recursion level
//                 by level: the nodes of a level are independent and are spread
//                 over the waves; every node gets its two half sweeps, the
//                 midpoint reductions, and writes (type, midj) plus its children
//                 to a node table.  The host turns the table into the edit
//                 script in the reference's depth-first order (progAdd/progDel
//                 merge runs, so the order of emission matters).
//
// Gap model.  With g/l the GAPCOL/LENCOL of a profile position and m = 0 at
// the profile ends (endGapPenalties off) else all ones:
//   open1(I,J) = m1[I] & (g1[I]+g2[J])   ext1(I,J) = m1[I] & l2[J]
//   open2(I,J) = m2[J] & (g1[I]+g2[J])   ext2(I,J) = m2[J] & l1[I]
// Forward half (:636-680), cell (i,j) at (I,J) = (A+i, B+j):
//   F = max(H[i][j-1] - open1 - ext1, F[i][j-1] - ext1)
//   E = max(H[i-1][j] - open2 - ext2, E[i-1][j] - ext2)
//   H = max(H[i-1][j-1] + S(I,J), F, E)
// Reverse half (:684-747) in reversed coordinates r = M-i, c = N-j: the
// reference carries the opening penalty of the previous column/row along
// (f - h - g + t, SS - h - g + open2(A+i+2,.)); since max(a-x, b-x) = max(a,b)-x
// exactly in int32 this is the Gotoh recurrence
//   G = max(H[r][c-1], G[r][c-1]) - ext1(I-1,J),  F = G - open1(I-1,J)
//   K = max(H[r-1][c], K[r-1][c]) - ext2(I,J-1),  E = K - open2(I,J-1)
//   H = max(H[r-1][c-1] + S(I,J), F, E)          (I,J) = (A+M+1-r, B+N+1-c)
// and SS[j] + gS[j] of the type-2 midpoint test is K of the last row.  In both
// halves the F/E (G/K) boundary values are never selected (the first interior
// cell's two candidates are equal), so they are -inf here; the H boundaries
// are the reference's rows and columns.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cw_profile_kernels.h"

namespace cwp {

constexpr int WAVE = 64;
constexpr int BIG = 0x7fffffff;
constexpr int NEG = -(1 << 28);

enum { MODE_FWD = 0, MODE_REV = 1 };

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
__device__ __forceinline__ int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, WAVE));
    return v;
}
__device__ __forceinline__ int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, WAVE));
    return v;
}
__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
}
__device__ __forceinline__ int max3(int a, int b, int c) { return max(max(a, b), c); }

// LDS address space: staged reads compile to ds_read_* rather than flat loads
#define CWP_LDS __attribute__((address_space(3)))
typedef CWP_LDS int lint;

// One profile pair as the device sees it.
struct Prob {
    int L1, L2, ldS;
    const int4* row;   // {g1, l1, m1, g1+l1}[I], I = 0..L1+1
    const int4* col;   // {g2, l2, m2, g2+l2}[J], J = 0..L2+1
    const int* pl1;    // pl1[x] = sum_{y<x} l1[y]
    const int* pl2;
    const int* S;      // S[J*ldS + I]
};

__device__ __forceinline__ int open1(const Prob& p, int I, int J) { return p.row[I].z & (p.row[I].x + p.col[J].x); }
__device__ __forceinline__ int ext1(const Prob& p, int I, int J) { return p.row[I].z & p.col[J].y; }
__device__ __forceinline__ int open2(const Prob& p, int I, int J) { return p.col[J].z & (p.row[I].x + p.col[J].x); }
__device__ __forceinline__ int ext2(const Prob& p, int I, int J) { return p.col[J].z & p.row[I].y; }

// gapPenalty1 (:1166-1189) / gapPenalty2 (:1239-1262) with prefix sums for the
// LENCOL runs.
__device__ __forceinline__ int gap_pen1(const Prob& p, int i, int j, int k) {
    if (k <= 0 || p.row[i].z == 0) return 0;
    const int g = p.row[i].x + p.col[j].x;
    const int e = j + k < p.L2 ? j + k : p.L2;
    return g + (j < p.L2 ? p.pl2[e] - p.pl2[j] : 0);
}
__device__ __forceinline__ int gap_pen2(const Prob& p, int i, int j, int k) {
    if (k <= 0 || p.col[j].z == 0) return 0;
    const int g = p.row[i].x + p.col[j].x;
    const int e = i + k < p.L1 ? i + k : p.L1;
    return g + (i < p.L1 ? p.pl1[e] - p.pl1[i] : 0);
}

// One half sweep over nrows x ncols.  Local row i (1..nrows) scores at row
// I = ib + id*i, local column c at J = jb + jd*c.
struct PSweep {
    Prob p;
    int nrows, ncols;
    int ib, id, jb, jd;
    // left column H(i,0), i >= 1:
    //   forward: lt0 + i*ltl
    //   reverse: (go2 ? -open2(I(i), jedge) : 0) - (m2[jedge] & (pl1[ib] - pl1[I(i)]))
    int lt0, ltl;
    int go2, jedge;
    lint* stage;       // this wave's LDS buffer (CWP_STAGE_INTS ints)
    int jrec0;         // column of staged record 0 (B): staged record x is column jrec0 + x
};

template <int MODE>
__device__ __forceinline__ int left_h(const PSweep& W, int i) {
    if (i <= 0) return 0;
    if (MODE == MODE_FWD) return W.lt0 + i * W.ltl;
    const int I = W.ib + W.id * i;
    const int4 ce = W.p.col[W.jedge];
    const int g = W.go2 ? -(ce.z & (W.p.row[I].x + ce.x)) : 0;
    return g - (ce.z & (W.p.pl1[W.ib] - W.p.pl1[I]));
}

template <bool V> struct Tag { static constexpr bool value = V; };

// One stripe: local rows row0+1 .. row0+P, R rows per lane (lanes >= k1 hold
// R-1 rows behind a pass-through row).  The top boundary row comes from inH
// (inE, or -inf on the first stripe); lane 63's rows leave through outH/outE.
// DIRECT: R == 1, P < 64, lane P-1 writes its row straight to out.
template <int R, int MODE, bool DIRECT, bool STAGED = false>
__device__ __forceinline__ void pstripe(const PSweep& W, int row0, int P, int first,
                                        const int* inH, const int* inE, int* outH, int* outE)
{
    const int lane = lane_id();
    const int ncols = W.ncols;
    const int k1 = DIRECT ? P : P - WAVE * (R - 1);
    const bool dummy = (!DIRECT) && (R > 1) && (lane >= k1);
    const bool lane_on = DIRECT ? (lane < P) : true;
    const int d = dummy ? 1 : 0;
    const int firstrow = row0 + 1 + (lane < k1 ? lane * R : k1 * R + (lane - k1) * (R - 1));
    const Prob& p = W.p;
    const int ldS = p.ldS;

    int gF[R], mF[R], gE[R], lE[R], hL[R], f[R], sI[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = firstrow + r - d;
        const int ii = i < 1 ? 1 : (i > W.nrows ? W.nrows : i);
        const int I = W.ib + W.id * ii;
        const int4 rf = p.row[MODE == MODE_FWD ? I : I - 1];
        const int4 re = p.row[I];
        gF[r] = rf.x;
        mF[r] = rf.z;
        gE[r] = (MODE == MODE_FWD) ? re.w : re.x;
        lE[r] = re.y;
        hL[r] = left_h<MODE>(W, i < W.nrows ? i : W.nrows);   // idle lanes: any in-range row
        f[r] = NEG;
        sI[r] = I;
    }
    int dprev = left_h<MODE>(W, firstrow - 1 < W.nrows ? firstrow - 1 : W.nrows);

    // lane-0 input chunks: lane l of chunk ci holds column 64*ci + 1 + l
    auto chunk = [&](int ci, int& h, int& e) {
        const int col = 64 * ci + 1 + lane;
        if (col <= ncols) {
            h = inH[col];
            e = first ? NEG : inE[col];
        } else {
            h = 0;
            e = NEG;
        }
    };
    int cinH, cinE, nxtH = 0, nxtE = NEG;
    chunk(0, cinH, cinE);
    const int nchunks = (ncols + 63) >> 6;
    if (nchunks > 1) chunk(1, nxtH, nxtE);

    int outHv = 0, outEv = NEG, coutH = 0, coutE = 0;

    // column records and scores of this lane's column, one step ahead
    auto colJ = [&](int c) {
        const int cc = c < 1 ? 1 : (c > ncols ? ncols : c);
        return W.jb + W.jd * cc;
    };
    int4 nrecF, nrecE;
    int nsv[R];
    auto prefetch = [&](int c) {
        const int J = colJ(c);
        if constexpr (STAGED) {
            // single stripe, one row per lane: records at x = J - jrec0, scores
            // column-major over (local column, local row)
            const lint* lrec = W.stage;
            const int x = J - W.jrec0;
            nrecF = make_int4(lrec[4 * x], lrec[4 * x + 1], lrec[4 * x + 2], lrec[4 * x + 3]);
            if (MODE == MODE_FWD) nrecE = nrecF;
            else nrecE = make_int4(lrec[4 * x - 4], lrec[4 * x - 3], lrec[4 * x - 2], lrec[4 * x - 1]);
            const int cc = c < 1 ? 1 : (c > ncols ? ncols : c);
            const int rr = firstrow < 1 ? 1 : (firstrow > W.nrows ? W.nrows : firstrow);
            nsv[0] = W.stage[4 * (ncols + 2) + (cc - 1) * W.nrows + (rr - 1)];
        } else {
            nrecF = p.col[J];
            nrecE = (MODE == MODE_FWD) ? nrecF : p.col[J - 1];
            const int* Sc = p.S + (int64_t)J * ldS;
#pragma unroll
            for (int r = 0; r < R; ++r) nsv[r] = Sc[sI[r]];
        }
    };
    prefetch(1 - lane);

    auto step = [&](int t, auto masked) {
        if (((t - 1) & 63) == 0 && t > 1) {
            const int ci = (t - 1) >> 6;
            cinH = nxtH;
            cinE = nxtE;
            if (ci + 1 < nchunks) chunk(ci + 1, nxtH, nxtE);
        }
        const int nH = shl1(cinH, cinH);                // chunk rotation for the next step
        const int nE = shl1(cinE, cinE);
        const int inHv = shr1(cinH, outHv);             // lane 0: chunk element; others: lane above
        const int inEv = shr1(cinE, outEv);
        cinH = nH;
        cinE = nE;
        const int4 cF = nrecF, cE = nrecE;
        int s[R];
#pragma unroll
        for (int r = 0; r < R; ++r) s[r] = nsv[r];
        prefetch(t + 1 - lane);
        const int c = t - lane;
        bool act = true;
        if constexpr (decltype(masked)::value) act = lane_on && c >= 1 && c <= ncols;
        if (act) {
            int hup = inHv, eup = inEv, dg = dprev;
            dprev = inHv;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                int h, e;
                if (MODE == MODE_FWD) {
                    const int fn = max(hL[r] - ((gF[r] + cF.w) & mF[r]), f[r] - (cF.y & mF[r]));
                    const int en = max(hup - ((gE[r] + cE.x) & cE.z), eup - (lE[r] & cE.z));
                    h = max3(dg + s[r], fn, en);
                    f[r] = fn;
                    e = en;
                } else {
                    const int g = max(hL[r], f[r]) - (cF.y & mF[r]);
                    const int k = max(hup, eup) - (lE[r] & cE.z);
                    h = max3(dg + s[r], g - ((gF[r] + cF.x) & mF[r]), k - ((gE[r] + cE.x) & cE.z));
                    f[r] = g;
                    e = k;
                }
                if (r == 0 && R > 1 && !DIRECT) {
                    if (dummy) { h = hup; e = eup; }
                }
                dg = hL[r];
                hL[r] = h;
                hup = h;
                eup = e;
            }
            outHv = hup;
            outEv = eup;
            if (DIRECT && lane == P - 1) {
                outH[c] = hup;
                outE[c] = eup;
            }
        }
        if (!DIRECT) {
            coutH = shl1(outHv, coutH);
            coutE = shl1(outEv, coutE);
            const int oc = t - 63;
            if (oc >= 1 && (((oc & 63) == 0) || oc == ncols)) {
                const int col = oc - 63 + lane;
                const int cstart = ((oc - 1) & ~63) + 1;
                if (col >= cstart) {
                    outH[col] = coutH;
                    outE[col] = coutE;
                }
            }
        }
    };

    if (DIRECT) {
        const int tend = ncols + P - 1;
        for (int t = 1; t <= tend; ++t) step(t, Tag<true>{});
    } else {
        const int tend = ncols + 63;
        const int up_end = ncols < 63 ? ncols + 63 : 63;
        int t = 1;
        for (; t <= up_end; ++t) step(t, Tag<true>{});
        for (; t <= ncols; ++t) step(t, Tag<false>{});
        for (; t <= tend; ++t) step(t, Tag<true>{});
    }
}

template <int R, int MODE>
__device__ __attribute__((noinline)) void pstripe_call(const PSweep W, int row0, int P, int first,
                                                       const int* inH, const int* inE, int* outH,
                                                       int* outE)
{
    pstripe<R, MODE, false>(W, row0, P, first, inH, inE, outH, outE);
}

template <int MODE>
__device__ void run_pstripe(int R, bool direct, const PSweep& W, int row0, int P, int first,
                            const int* inH, const int* inE, int* outH, int* outE)
{
    if (direct) { pstripe<1, MODE, true>(W, row0, P, first, inH, inE, outH, outE); return; }
    if (R == 0) { pstripe<1, MODE, true, true>(W, row0, P, first, inH, inE, outH, outE); return; }
    switch (R) {
    case 1: pstripe<1, MODE, false>(W, row0, P, first, inH, inE, outH, outE); break;
    case 2: pstripe_call<2, MODE>(W, row0, P, first, inH, inE, outH, outE); break;
    case 3: pstripe_call<3, MODE>(W, row0, P, first, inH, inE, outH, outE); break;
    case 4: pstripe_call<4, MODE>(W, row0, P, first, inH, inE, outH, outE); break;
    case 5: pstripe_call<5, MODE>(W, row0, P, first, inH, inE, outH, outE); break;
    case 6: pstripe_call<6, MODE>(W, row0, P, first, inH, inE, outH, outE); break;
    case 7: pstripe_call<7, MODE>(W, row0, P, first, inH, inE, outH, outE); break;
    default: pstripe_call<8, MODE>(W, row0, P, first, inH, inE, outH, outE); break;
    }
}

struct Bufs { int *h0, *e0, *h1, *e1; };

// Whole half sweep in balanced stripes; the top boundary row must already be
// in g.h0[1..ncols].  The final row goes to finH/finE[1..ncols].
// Loads a single-stripe node's column records (columns jrec0 .. jrec0+N+1) and
// its nrows x N score block into this wave's LDS buffer.
template <int MODE>
__device__ void stage_node(const PSweep& W)
{
    const int lane = lane_id();
    const Prob& p = W.p;
    const int N = W.ncols, P = W.nrows;
    lint* st = W.stage;
    for (int x = lane; x < N + 2; x += WAVE) {
        const int4 r = p.col[W.jrec0 + x];
        st[4 * x + 0] = r.x;
        st[4 * x + 1] = r.y;
        st[4 * x + 2] = r.z;
        st[4 * x + 3] = r.w;
    }
    lint* lS = st + 4 * (N + 2);
    const int per = WAVE / P;                 // columns per pass (P < 64)
    const int cc = lane / P, rr = lane - cc * P;
    if (cc < per) {
        const int I = W.ib + W.id * (rr + 1);
        for (int c0 = 0; c0 < N; c0 += per) {
            const int c = c0 + cc;
            if (c < N) {
                const int J = W.jb + W.jd * (c + 1);
                lS[c * P + rr] = p.S[(int64_t)J * p.ldS + I];
            }
        }
    }
    wave_fence();
}

template <int MODE>
__device__ void psweep(const PSweep& W, int rmax, const Bufs& g, int* finH, int* finE)
{
    const int nrows = W.nrows;
    if (nrows < WAVE && nrows * W.ncols + 4 * (W.ncols + 2) <= CWP_STAGE_INTS) {
        // one stripe, one row per lane, from LDS
        stage_node<MODE>(W);
        run_pstripe<MODE>(0, false, W, 0, nrows, 1, g.h0, g.e0, finH, finE);
        wave_fence();
        return;
    }
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
        run_pstripe<MODE>(R, direct, W, row0, P, s == 0, inH, inE, oH, oE);
        wave_fence();
        row0 += P;
    }
}

// The top boundary row H(0,c), c = 1..N, of a half sweep (:636-641, :684-693).
template <int MODE>
__device__ void fill_top(const Prob& p, int A, int B, int M, int N, int* h0)
{
    const int lane = lane_id();
    if (MODE == MODE_FWD) {
        const int t0 = -open1(p, A, B + 1), tl = -ext1(p, A, B + 1);
        for (int c = 1 + lane; c <= N; c += WAVE) h0[c] = t0 + c * tl;
    } else {
        const int4 re = p.row[A + M];
        for (int c = 1 + lane; c <= N; c += WAVE) {
            const int J = B + N - c + 1;
            h0[c] = -(re.z & (re.x + p.col[J].x)) - (re.z & (p.pl2[B + N + 1] - p.pl2[J]));
        }
    }
    wave_fence();
}

// One node of the recursion (one call of progDiff) by one wave.  Every lane
// runs every statement here and writes the (identical) node fields: the node
// loop of cw_prf_mm must contain no lane-divergent branch, or the compiler's
// structurization can leave some lanes cycling through the loop header (whose
// shuffle then reads an inactive lane) while others wait at its exit.  Kept
// out of line for the same reason.
__device__ __attribute__((noinline)) void do_node(const Prob& p, int* nd, int is_root, int rmax, int* scr, int rowcap,
                        int* out_score, lint* stage)
{
    const int lane = lane_id();
    const int A = uni(nd[0]), B = uni(nd[1]), M = uni(nd[2]), N = uni(nd[3]);
    const int flags = uni(nd[4]);
    const int go1 = flags & 1, go2 = (flags >> 1) & 1;
    if (N <= 0 || M <= 1) {                                        // :540-600
        int midj = 0, midh;
        if (N <= 0) {
            midh = -gap_pen1(p, A, B, M);
        } else if (M <= 0) {
            midh = -gap_pen2(p, A, B, N);
        } else {
            const int midh0 = go1 == 0 ? -gap_pen1(p, A + 1, B + 1, N)
                                       : -gap_pen2(p, A + 1, B, 1) - gap_pen1(p, A + 1, B + 1, N);
            int bv = -BIG, bj = BIG;
            for (int j = 1 + lane; j <= N; j += WAVE) {
                const int v = -gap_pen1(p, A, B + 1, j - 1) + p.S[(int64_t)(B + j) * p.ldS + A + 1] -
                              gap_pen1(p, A + 1, B + j + 1, N - j);
                if (v > bv) { bv = v; bj = j; }
            }
            const int V = wave_max(bv);
            const int J = wave_min(bv == V ? bj : BIG);
            if (V > midh0) { midh = V; midj = J; }
            else midh = midh0;
        }
        nd[5] = 0;
        nd[6] = midj;
        nd[7] = -1;
        if (is_root) *out_score = midh;
        return;
    }
    const int midi = M / 2;
    const Bufs g{scr, scr + rowcap, scr + 2 * rowcap, scr + 3 * rowcap};
    int* fH = scr + 4 * rowcap;
    int* fE = scr + 5 * rowcap;
    int* rH = scr + 6 * rowcap;
    int* rK = scr + 7 * rowcap;

    PSweep F;
    F.p = p;
    F.nrows = midi; F.ncols = N;
    F.ib = A; F.id = 1; F.jb = B; F.jd = 1;
    F.lt0 = go1 ? -open2(p, A + 1, B) : 0;
    F.ltl = -ext2(p, A + 1, B);
    F.go2 = 0; F.jedge = 0;
    F.stage = stage; F.jrec0 = B;
    fill_top<MODE_FWD>(p, A, B, M, N, g.h0);
    psweep<MODE_FWD>(F, rmax, g, fH, fE);

    PSweep Rv;
    Rv.p = p;
    Rv.nrows = M - midi; Rv.ncols = N;
    Rv.ib = A + M + 1; Rv.id = -1; Rv.jb = B + N + 1; Rv.jd = -1;
    Rv.lt0 = 0; Rv.ltl = 0;
    Rv.go2 = go2; Rv.jedge = B + N;
    Rv.stage = stage; Rv.jrec0 = B;
    fill_top<MODE_REV>(p, A, B, M, N, g.h0);
    psweep<MODE_REV>(Rv, rmax, g, rH, rK);
    wave_fence();

    // midpoint (:756-776)
    const int HH0 = left_h<MODE_FWD>(F, midi);
    const int RRN = left_h<MODE_REV>(Rv, M - midi);
    int vmax = -BIG, jfirst = BIG, jc = -1, wmax = -BIG, jw = -1;
    for (int j = lane; j <= N; j += WAVE) {
        const int HHj = (j == 0) ? HH0 : fH[j];
        const int DDj = (j == 0) ? HH0 : fE[j];
        const int gS = open2(p, A + midi + 1, B + j);
        int RRj, SSj;
        if (j == N) { RRj = RRN; SSj = RRN; }
        else { RRj = rH[N - j]; SSj = rK[N - j] - gS; }
        const int v = HHj + RRj;
        const int w = DDj + SSj + gS;
        if (v > vmax) { vmax = v; jfirst = j; jc = -1; }
        if (v == vmax && HHj != DDj && RRj == SSj) jc = j;
        if (w >= wmax) { wmax = w; jw = j; }
    }
    const int Mv = wave_max(vmax);
    const int J0 = wave_min(vmax == Mv ? jfirst : BIG);
    const int JC = wave_max(vmax == Mv ? jc : -1);
    const int Wv = wave_max(wmax);
    const int JW = wave_max(wmax == Wv ? jw : -1);
    int type, midj, midh;
    if (Wv > Mv) { type = 2; midj = JW; midh = Wv; }
    else { type = 1; midj = J0 > JC ? J0 : JC; midh = Mv; }
    nd[5] = type;
    nd[6] = midj;
    if (is_root) *out_score = midh;
}

} // namespace cwp

using namespace cwp;

__device__ static Prob make_prob(const PrfArgs& a, const PrfProb& pp)
{
    Prob p;
    p.L1 = pp.L1;
    p.L2 = pp.L2;
    p.ldS = pp.ldS;
    p.row = (const int4*)(a.pool + pp.row);
    p.col = (const int4*)(a.pool + pp.col);
    p.pl1 = a.pool + pp.pl1;
    p.pl2 = a.pool + pp.pl2;
    p.S = a.pool + pp.S;
    return p;
}

// S[J*ldS + I] = prfScore(I, J) for 1 <= I <= L1, 1 <= J <= L2.  Block = one
// 64x64 tile: lane = I within the tile, wave = J phase; profile-2 rows of the
// tile in LDS (the 27 scored columns: codes 0..24, 30, 31).
extern "C" __global__ void __launch_bounds__(256) cw_prf_score(PrfArgs a)
{
    __shared__ int p2t[64][28];
    const int tile = blockIdx.x;
    int lo = 0, hi = a.nprob;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.tile_base[mid] <= tile) lo = mid;
        else hi = mid;
    }
    const PrfProb pp = a.probs[lo];
    const int t = tile - a.tile_base[lo];
    const int nti = (pp.L1 + 63) >> 6;
    const int ti = t % nti, tj = t / nti;
    const int* P1 = a.pool + pp.prf1;
    const int* P2 = a.pool + pp.prf2;
    for (int k = threadIdx.x; k < 64 * 27; k += 256) {
        const int jj = k / 27, q = k % 27;
        const int J = tj * 64 + 1 + jj;
        const int code = q < 25 ? q : (q == 25 ? 30 : 31);
        p2t[jj][q] = (J <= pp.L2) ? P2[(int64_t)J * CWP_PRF_W + code] : 0;
    }
    __syncthreads();
    const int I = ti * 64 + (threadIdx.x & 63) + 1;
    if (I > pp.L1) return;
    int r1[27];
#pragma unroll
    for (int q = 0; q < 27; ++q) {
        const int code = q < 25 ? q : (q == 25 ? 30 : 31);
        r1[q] = P1[(int64_t)I * CWP_PRF_W + code];
    }
    int* S = a.pool + pp.S;
    for (int jj = threadIdx.x >> 6; jj < 64; jj += 4) {
        const int J = tj * 64 + 1 + jj;
        if (J > pp.L2) break;
        int s = 0;
#pragma unroll
        for (int q = 0; q < 27; ++q) s += r1[q] * p2t[jj][q];
        S[(int64_t)J * pp.ldS + I] = s / 10;
    }
}

// Myers-Miller recursion of one profile pair per workgroup, level by level.
extern "C" __global__ void __launch_bounds__(CWP_THREADS) cw_prf_mm(PrfArgs a)
{
    __shared__ int s_lev_b, s_lev_e, s_ctr, s_err;
    __shared__ int s_stage[CWP_WAVES * CWP_STAGE_INTS];
    const int prob = blockIdx.x;
    const PrfProb pp = a.probs[prob];
    const Prob p = make_prob(a, pp);
    int* nodes = a.pool + pp.nodes;
    const int wave = threadIdx.x / WAVE;
    const int lane = lane_id();
    int* scr = a.pool + pp.scratch + (int64_t)wave * CWP_SCR_ROWS * pp.rowcap;
    const int cap = pp.node_cap;
    if (threadIdx.x == 0) {
        int* nd = nodes;
        nd[0] = 0; nd[1] = 0; nd[2] = pp.L1; nd[3] = pp.L2;
        nd[4] = pp.flags & 3; nd[5] = 0; nd[6] = 0; nd[7] = -1;
        s_lev_b = 0;
        s_lev_e = 1;
        s_err = 0;
    }
    __syncthreads();
    int levels = 0;
    for (;;) {
        const int lb = s_lev_b, le = s_lev_e;
        if (lb >= le) break;
        __syncthreads();
        if (threadIdx.x == 0) s_ctr = lb;
        __syncthreads();
        for (;;) {
            const int got = atomicAdd(&s_ctr, lane == 0 ? 1 : 0);
            const int idx = uni(__shfl(got, 0, WAVE));
            if (idx >= le) break;
            do_node(p, nodes + (int64_t)idx * CWP_NODE_INTS, idx == 0, a.rmax, scr, pp.rowcap,
                    a.out + 4 * prob, (lint*)(s_stage + wave * CWP_STAGE_INTS));
        }
        __syncthreads();
        if (wave == 0) {
            int next = le;
            for (int base = lb; base < le; base += WAVE) {
                const int idx = base + lane;
                int* nd = nodes + (int64_t)idx * CWP_NODE_INTS;
                const int type = idx < le ? nd[5] : 0;
                const uint64_t bal = __ballot(type != 0);
                const int before = __popcll(bal & ((1ull << lane) - 1ull));
                const int slot = next + 2 * before;
                if (type != 0) {
                    if (slot + 2 <= cap) {
                        const int A = nd[0], B = nd[1], M = nd[2], N = nd[3], fl = nd[4];
                        const int go1 = fl & 1, go2 = (fl >> 1) & 1, midj = nd[6];
                        const int midi = M / 2;
                        int* c0 = nodes + (int64_t)slot * CWP_NODE_INTS;
                        int* c1 = c0 + CWP_NODE_INTS;
                        if (type == 1) {
                            c0[0] = A; c0[1] = B; c0[2] = midi; c0[3] = midj; c0[4] = go1 | 2;
                            c1[0] = A + midi; c1[1] = B + midj; c1[2] = M - midi; c1[3] = N - midj;
                            c1[4] = 1 | (go2 << 1);
                        } else {
                            c0[0] = A; c0[1] = B; c0[2] = midi - 1; c0[3] = midj; c0[4] = go1;
                            c1[0] = A + midi + 1; c1[1] = B + midj; c1[2] = M - midi - 1; c1[3] = N - midj;
                            c1[4] = go2 << 1;
                        }
                        c0[5] = c1[5] = 0;
                        c0[6] = c1[6] = 0;
                        c0[7] = c1[7] = -1;
                        nd[7] = slot;
                    } else {
                        s_err = 1;
                    }
                }
                next += 2 * __popcll(bal);
            }
            s_lev_b = le;                           // same value from every lane
            s_lev_e = next < cap ? next : cap;
        }
        __syncthreads();
        ++levels;
    }
    if (threadIdx.x == 0) {
        a.out[4 * prob + 1] = s_lev_e;
        a.out[4 * prob + 2] = s_err;
        a.out[4 * prob + 3] = levels;
    }
}

// ---- host-side launchers (used by cwhip_prf.cpp) ---------------------------
hipError_t cwp_launch_score(int tiles, hipStream_t st, const PrfArgs& a)
{
    hipLaunchKernelGGL(cw_prf_score, dim3(tiles), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t cwp_launch_mm(int nprob, hipStream_t st, const PrfArgs& a)
{
    hipLaunchKernelGGL(cw_prf_mm, dim3(nprob), dim3(CWP_THREADS), 0, st, a);
    return hipGetLastError();
}
