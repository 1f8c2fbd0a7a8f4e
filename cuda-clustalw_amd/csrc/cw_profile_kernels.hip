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
//   cw_prf_wide   (batches of up to CWP_WIDE_MAXP pairs) the two half sweeps of
//                 each large root as pipelined stripes of one-wave workgroups
//                 on separate CUs.
//   cw_prf_level  (same batches) the half sweeps of the next levels' nodes, one
//                 wave per half over the chip.  cw_prf_mm then finishes one
//                 level per launch (leaves, local subtrees, midpoints,
//                 children) and, in its last launch, sweeps the remaining
//                 levels itself.  A lone wave on a CU sweeps about twice as many
//                 cells per cycle as one of the four that share a CU's texture
//                 path in cw_prf_mm.
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
// lane l receives src[l+1]; lane 63 receives 0 (one DPP move, no copy of old)
__device__ __forceinline__ int shl1z(int src) {
    return __builtin_amdgcn_mov_dpp(src, 0x130, 0xf, 0xf, true);
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
// a + b + c as one v_add3_u32 (the compiler's reassociation would otherwise
// split it into two adds)
__device__ __forceinline__ int add3(int a, int b, int c) {
    int r;
    asm("v_add3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
// -x, computed no earlier than `after` exists: keeps the scheduler from
// hoisting the negation of a prefetched value (and with it the wait for the
// load) above the step that uses it
__device__ __forceinline__ int neg_after(int x, int after) {
    int r;
    asm("v_sub_u32 %0, 0, %1" : "=v"(r) : "v"(x), "v"(after));
    return r;
}

// LDS address space: reads through these compile to ds_read_* rather than flat loads
#define CWP_LDS __attribute__((address_space(3)))
typedef CWP_LDS int lint;
__device__ __forceinline__ int lds_add(lint* q, int v)
{
    return __hip_atomic_fetch_add(q, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Global address space: these arrays are read with global (not flat) loads,
// which LDS waits (lgkmcnt) do not wait for.
#define CWP_GLB __attribute__((address_space(1)))
typedef CWP_GLB const int gcint;
typedef CWP_GLB const int4 gcint4;
typedef CWP_GLB int gint;

// An int4 read through a global-address-space pointer.
__device__ __forceinline__ int4 ld4(gcint4* q) { return make_int4(q->x, q->y, q->z, q->w); }

// One profile pair as the device sees it.
struct Prob {
    int L1, L2, ldS;
    gcint4* row;       // {g1, l1, m1, g1+l1}[I], I = 0..L1+1
    gcint4* col;       // {g2, l2, m2, g2+l2}[J], J = 0..L2+1
    gcint* pl1;        // pl1[x] = sum_{y<x} l1[y]
    gcint* pl2;
    gcint* S;          // S[J*ldS + I]
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
    lint* err;         // set to 1 on a violated precondition (reported, not trapped)
};

template <int MODE>
__device__ __forceinline__ int left_h(const PSweep& W, int i) {
    if (i <= 0) return 0;
    if (MODE == MODE_FWD) return W.lt0 + i * W.ltl;
    const int I = W.ib + W.id * i;
    const int4 ce = ld4(&W.p.col[W.jedge]);
    const int g = W.go2 ? -(ce.z & (W.p.row[I].x + ce.x)) : 0;
    return g - (ce.z & (W.p.pl1[W.ib] - W.p.pl1[I]));
}

template <bool V> struct Tag { static constexpr bool value = V; };
template <int V> struct Par { static constexpr int value = V; };

// f(Par<0>{}), f(Par<1>{}), ..., f(Par<N-1>{}) (compile-time unrolled)
template <int N, int I = 0>
struct Unroll {
    template <typename F>
    __device__ __forceinline__ static void run(F&& f)
    {
        if constexpr (I < N) {
            f(Par<I>{});
            Unroll<N, I + 1>::run(f);
        }
    }
};

// Two waves sharing one half sweep: stripe j reads the bottom row stripe j-1
// writes, 64 columns at a time.  `pub` counts the chunks a stripe has flushed
// to memory; the next stripe waits on it (`wait`) before it loads a chunk.  A
// wait that outlasts ~2^24 sleeps sets *err and goes on, so a mistake shows up
// as a reported error, not a hang.
struct StripeSync {
    const int* wait;   // generic pointers, tested against null.  Stripes of one
    int* pub;          // workgroup: the counters and the error flag are in LDS
    int* err;          // (offset 0 is a valid LDS address) and are accessed as
                       // such; stripes of different workgroups: global memory
    int limit;         // the latter's wait bound, in sleeps
};

// Stripes of different workgroups (cw_prf_wide) may run on different XCDs,
// whose L2 caches are not coherent with each other.  Everything they hand over
// (chunk rows and counters) is written and read with relaxed agent-scope
// atomics, which go past the non-coherent caches for that one access (global
// loads and stores with sc1).  The consumer takes a chunk with an agent-scope
// acquire fence after its counter load.  The producer's side is done by hand:
// its chunk stores are themselves such atomics, so it only has to wait for
// them to complete (vmcnt(0)) before it stores the counter.  An agent-scope
// release fence would add a write-back of the whole L2 (buffer_wbl2) per
// chunk for lines this kernel never leaves dirty: 2 ms more per cwhip_msa on
// 1000 x ~1002, measured.
__device__ __forceinline__ int ld_agent(const int* q)
{
    return __hip_atomic_load((CWP_GLB int*)q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(int* q, int v)
{
    __hip_atomic_store((CWP_GLB int*)q, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima): every store
// this wave issued has completed
__device__ __forceinline__ void wait_vm_stores()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// wait_chunks for a producer in another workgroup.  A producer that never
// publishes (it failed its own wait) ends the wait through the shared error
// flag, so one failure stops every stripe of the launch within one timeout.
__device__ __forceinline__ void wait_chunks_glb(const StripeSync& sy, int need)
{
    int spins = 0;
    while (uni(ld_agent(sy.wait)) < need) {
        if (uni(ld_agent(sy.err)) != 0) break;
        __builtin_amdgcn_s_sleep(2);
        if (++spins > sy.limit) {
            st_agent(sy.err, 1);
            break;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

__device__ __forceinline__ void wait_chunks(const StripeSync& sy, int need)
{
    int spins = 0;
    while (uni(*(const volatile lint*)sy.wait) < need) {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > (1 << 24)) {
            *(lint*)sy.err = 1;                          // same value from every lane
            break;
        }
    }
    wave_fence();
}

// One stripe: local rows row0+1 .. row0+P, R rows per lane (lanes >= k1 hold
// R-1 rows behind a pass-through row).  The top boundary row comes from inH
// (inE, or -inf on the first stripe); lane 63's rows leave through outH/outE.
// DIRECT: R == 1, P < 64, lane P-1 writes its row straight to out.
//
// Rows and columns of a half sweep run the same way: MODE_FWD scores row
// I = ib + i and column J = jb + c, MODE_REV I = ib - i and J = jb - c.  The
// row gap mask m1 is all ones on every row a half sweep reads (forward rows
// A+1 .. A+M/2, reverse records A+M/2 .. A+M-1: never record 0 or L1), so the
// recurrences leave it out (a stripe that finds otherwise sets *W.err); the
// column mask m2 can be 0 only at the node's last column c = ncols, so the
// blocks in which every lane stays left of that column leave it out too.
//
// At step t lane L works on column c = t - L.  Its column record and its R
// scores S(I_r, J) are loaded D steps ahead into buffer t % D, which step t-D
// refilled after its own use (D = 8 up to four rows per lane, else 4: a lone
// wave on a CU waits out most of a score load's latency otherwise).  MODE_REV
// also needs col[J-1], the record of column c+1: the one the lane loaded for
// step t+1.
// X: the stripe above and below run in other workgroups (chunk rows and
// counters through agent-scope accesses, counters in global memory).
template <int R, int MODE, bool DIRECT, bool X = false>
__device__ __forceinline__ void pstripe(const PSweep& W, int row0, int P, int first,
                                        const int* inH, const int* inE, int* outH, int* outE,
                                        StripeSync sy)
{
    constexpr int DIR = (MODE == MODE_FWD) ? 1 : -1;
    constexpr int D = R <= 4 ? 8 : 4;
    const int lane = lane_id();
    // wave-uniform values in scalar registers, so the control of the step
    // loops (chunk turns, flushes, bounds) runs on the scalar unit
    const int ncols = uni(W.ncols);
    const int nrows = uni(W.nrows);
    P = uni(P);
    first = uni(first);
    row0 = uni(row0);
    const int ib = uni(W.ib), jb = uni(W.jb);
    const int k1 = DIRECT ? P : P - WAVE * (R - 1);
    const bool dummy = (!DIRECT) && (R > 1) && (lane >= k1);
    const bool lane_on = DIRECT ? (lane < P) : true;
    const int d = dummy ? 1 : 0;
    const int firstrow = row0 + 1 + (lane < k1 ? lane * R : k1 * R + (lane - k1) * (R - 1));
    const Prob& p = W.p;
    const unsigned ldS = (unsigned)uni(p.ldS);

    // per row: -gap-open terms, LENCOL, the left column H and F
    int ngF[R], ngE[R], lE[R], hL[R], f[R];
    int mall = -1;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = firstrow + r - d;
        const int ii = i < 1 ? 1 : (i > nrows ? nrows : i);
        const int I = ib + DIR * ii;
        const int4 rf = ld4(&p.row[MODE == MODE_FWD ? I : I - 1]);
        const int4 re = ld4(&p.row[I]);
        ngF[r] = -rf.x;
        mall &= rf.z;
        ngE[r] = (MODE == MODE_FWD) ? -re.w : -re.x;
        lE[r] = re.y;
        hL[r] = left_h<MODE>(W, i < nrows ? i : nrows);   // idle lanes: any in-range row
        f[r] = NEG;
    }
    if (!__all(mall == -1)) *W.err = 1;                  // same value from every lane
    int dprev = left_h<MODE>(W, firstrow - 1 < nrows ? firstrow - 1 : nrows);
    // S row of the lane's first row (rows r follow at I0 + DIR*r; a multi-row
    // stripe's rows, its pass-through row included, are all in 0 .. nrows)
    const int i0 = DIRECT ? (firstrow > nrows ? nrows : firstrow) : firstrow - d;
    const unsigned sI0 = (unsigned)(ib + DIR * i0);

    // lane-0 input chunks: lane l of chunk ci holds column 64*ci + 1 + l
    auto chunk = [&](int ci, int& h, int& e) {
        if (!DIRECT && sy.wait) {
            if constexpr (X) wait_chunks_glb(sy, ci + 1);
            else wait_chunks(sy, ci + 1);
        }
        const int col = 64 * ci + 1 + lane;
        if (col <= ncols) {
            if (X && sy.wait) {                         // written by the workgroup above
                h = ld_agent(inH + col);
                e = ld_agent(inE + col);
            } else {
                h = ((gcint*)inH)[col];
                e = first ? NEG : ((gcint*)inE)[col];
            }
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

    // The column record and scores of step t go into rec/sv[t % D], read
    // through pointers cp/sp[t % D] to the record and to S(I0, J).  Columns
    // are clamped to 0 .. ncols+1, which keeps J and every S index inside the
    // pair's arrays.  Within a block whose lanes all stay on columns
    // 1 .. ncols-1 and that starts at t0 with t0 + 62 + D <= ncols, no load is
    // clamped (t+D - lane <= ncols+1), so a pointer moves on by D columns per
    // use.
    int4 rec[D];
    int sv[D][R];
    gcint4* cp[D];
    gcint* sp[D];
    auto load = [&](int t, auto par, auto from_scratch) {
        constexpr int Q = decltype(par)::value;
        if constexpr (decltype(from_scratch)::value) {
            const int c = min(max(t - lane, 0), ncols + 1);
            const int J = jb + DIR * c;
            cp[Q] = p.col + J;
            sp[Q] = p.S + (__umul24((unsigned)J, ldS) + sI0);
        } else {
            cp[Q] += D * DIR;
            sp[Q] += D * DIR * (int)ldS;
        }
        rec[Q] = ld4(cp[Q]);
#pragma unroll
        for (int r = 0; r < R; ++r) sv[Q][r] = sp[Q][DIR * r];
    };
    Unroll<D>::run([&](auto i) {
        constexpr int I = decltype(i)::value;
        load(I + 1, Par<(I + 1) % D>{}, Tag<true>{});
    });

    // par: t % D.  masked: some lane may be off its columns, or on the last
    // one (exec masking and the column mask).  flush: this step may end an
    // output chunk (every 64th output column, and the last one).
    auto step = [&](int t, auto par, auto masked, auto flush) {
        constexpr int Q = decltype(par)::value;
        constexpr bool MK = decltype(masked)::value;
        const int nH = shl1z(cinH);                     // chunk rotation for the next step (lane 63's
        const int nE = shl1z(cinE);                     // element is past the chunk's last use)
        const int inHv = shr1(cinH, outHv);             // lane 0: chunk element; others: lane above
        const int inEv = shr1(cinE, outEv);
        cinH = nH;
        cinE = nE;
        const int4 cF = rec[Q];
        const int4 cE = (MODE == MODE_FWD) ? rec[Q] : rec[(Q + 1) % D];
        const int* const sc = sv[Q];
        const int c = t - lane;
        bool act = true;
        if constexpr (MK) act = lane_on && c >= 1 && c <= ncols;
        if (act) {
            int hup = inHv, eup = inEv, dg = dprev;
            dprev = inHv;
            const int nFw = neg_after(cF.w, inHv), nFx = neg_after(cF.x, inHv);
            const int nEx = neg_after(cE.x, inHv);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                int h, e;
                if constexpr (MODE == MODE_FWD) {
                    const int fn = max(add3(hL[r], ngF[r], nFw), f[r] - cF.y);
                    const int en = MK ? max(hup - ((cE.x - ngE[r]) & cE.z), eup - (lE[r] & cE.z))
                                      : max(add3(hup, ngE[r], nEx), eup - lE[r]);
                    h = max3(dg + sc[r], fn, en);
                    f[r] = fn;
                    e = en;
                } else {
                    const int g = max(hL[r], f[r]) - cF.y;
                    const int k = MK ? max(hup, eup) - (lE[r] & cE.z) : max(hup, eup) - lE[r];
                    const int ke = MK ? k - ((cE.x - ngE[r]) & cE.z) : add3(k, ngE[r], nEx);
                    h = max3(dg + sc[r], add3(g, ngF[r], nFx), ke);
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
                ((gint*)outH)[c] = hup;
                ((gint*)outE)[c] = eup;
            }
        }
        load(t + D, par, Tag<MK>{});                   // buffer Q is free again
        if (!DIRECT) {
            coutH = shl1(outHv, coutH);
            coutE = shl1(outEv, coutE);
            const int oc = t - 63;
            if (decltype(flush)::value && oc >= 1 && (((oc & 63) == 0) || oc == ncols)) {
                const int col = oc - 63 + lane;
                const int cstart = ((oc - 1) & ~63) + 1;
                if (col >= cstart) {
                    if constexpr (X) {
                        st_agent(outH + col, coutH);
                        st_agent(outE + col, coutE);
                    } else {
                        ((gint*)outH)[col] = coutH;
                        ((gint*)outE)[col] = coutE;
                    }
                }
                if (sy.pub) {
                    if constexpr (X) {
                        wait_vm_stores();
                        st_agent(sy.pub, (oc + 63) >> 6);
                    } else {
                        wave_fence();
                        *(volatile lint*)sy.pub = (oc + 63) >> 6;   // same value from every lane
                    }
                }
            }
        }
    };

    // Steps run in blocks of 64: block b covers t = 64b+1 .. 64b+64 (t = 1
    // mod D at its start, so steps take buffers 1, 2, .., D-1, 0 in turn), the
    // lane-0 input chunk turns at its start.  A block in which every lane
    // stays on columns 1 .. ncols-1 and no load is clamped (t0 >= 64,
    // t0 + 62 + D <= ncols) runs without exec masking or column mask and
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
        if (!DIRECT && t0 >= 64 && t0 + 62 + D <= ncols) {
#pragma unroll 1
            for (int k = 0; k < 64 - D; k += D)
                Unroll<D>::run([&](auto i) {
                    constexpr int I = decltype(i)::value;
                    step(t0 + k + I, Par<(I + 1) % D>{}, Tag<false>{}, Tag<false>{});
                });
            Unroll<D>::run([&](auto i) {                 // the flush after step t0 + 62
                constexpr int I = decltype(i)::value;
                step(t0 + 64 - D + I, Par<(I + 1) % D>{}, Tag<false>{}, Tag<64 - D + I == 62>{});
            });
        } else {
            int t = t0;
#pragma unroll 1
            for (; t + D - 1 <= t1; t += D)
                Unroll<D>::run([&](auto i) {
                    constexpr int I = decltype(i)::value;
                    step(t + I, Par<(I + 1) % D>{}, Tag<true>{}, Tag<true>{});
                });
            Unroll<D - 1>::run([&](auto i) {
                constexpr int I = decltype(i)::value;
                if (t + I <= t1) step(t + I, Par<(I + 1) % D>{}, Tag<true>{}, Tag<true>{});
            });
        }
    }
}

// Every stripe runs out of line: its loads' waits are then placed from the
// stripe's own accesses alone (all global or LDS; a flat access pending in
// the caller would make every wait for a load a wait for all of them).
template <int R, int MODE, bool DIRECT = false, bool X = false>
__device__ __attribute__((noinline)) void pstripe_call(const PSweep W, int row0, int P, int first,
                                                       const int* inH, const int* inE, int* outH,
                                                       int* outE, StripeSync sy)
{
    pstripe<R, MODE, DIRECT, X>(W, row0, P, first, inH, inE, outH, outE, sy);
}

// A stripe of 64..512 rows whose neighbours run in other workgroups.
template <int MODE>
__device__ void run_pstripe_x(int R, const PSweep& W, int row0, int P, int first, const int* inH,
                              const int* inE, int* outH, int* outE, StripeSync sy)
{
    switch (R) {
    case 1: pstripe_call<1, MODE, false, true>(W, row0, P, first, inH, inE, outH, outE, sy); break;
    case 2: pstripe_call<2, MODE, false, true>(W, row0, P, first, inH, inE, outH, outE, sy); break;
    case 3: pstripe_call<3, MODE, false, true>(W, row0, P, first, inH, inE, outH, outE, sy); break;
    case 4: pstripe_call<4, MODE, false, true>(W, row0, P, first, inH, inE, outH, outE, sy); break;
    case 5: pstripe_call<5, MODE, false, true>(W, row0, P, first, inH, inE, outH, outE, sy); break;
    case 6: pstripe_call<6, MODE, false, true>(W, row0, P, first, inH, inE, outH, outE, sy); break;
    case 7: pstripe_call<7, MODE, false, true>(W, row0, P, first, inH, inE, outH, outE, sy); break;
    default: pstripe_call<8, MODE, false, true>(W, row0, P, first, inH, inE, outH, outE, sy); break;
    }
}

template <int MODE>
__device__ void run_pstripe(int R, bool direct, const PSweep& W, int row0, int P, int first,
                            const int* inH, const int* inE, int* outH, int* outE,
                            StripeSync sy = StripeSync{nullptr, nullptr, nullptr})
{
    if (direct) { pstripe_call<1, MODE, true>(W, row0, P, first, inH, inE, outH, outE, sy); return; }
    switch (R) {
    case 1: pstripe_call<1, MODE>(W, row0, P, first, inH, inE, outH, outE, sy); break;
    case 2: pstripe_call<2, MODE>(W, row0, P, first, inH, inE, outH, outE, sy); break;
    case 3: pstripe_call<3, MODE>(W, row0, P, first, inH, inE, outH, outE, sy); break;
    case 4: pstripe_call<4, MODE>(W, row0, P, first, inH, inE, outH, outE, sy); break;
    case 5: pstripe_call<5, MODE>(W, row0, P, first, inH, inE, outH, outE, sy); break;
    case 6: pstripe_call<6, MODE>(W, row0, P, first, inH, inE, outH, outE, sy); break;
    case 7: pstripe_call<7, MODE>(W, row0, P, first, inH, inE, outH, outE, sy); break;
    default: pstripe_call<8, MODE>(W, row0, P, first, inH, inE, outH, outE, sy); break;
    }
}

struct Bufs { int *h0, *e0, *h1, *e1; };

// Whole half sweep in balanced stripes; the top boundary row must already be
// in g.h0[1..ncols].  The final row goes to finH/finE[1..ncols].
template <int MODE>
__device__ void psweep(const PSweep& W, int rmax, const Bufs& g, int* finH, int* finE)
{
    const int nrows = W.nrows;
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

// Member m of an S-wave group runs stripes m, m+S, ... of a half sweep (at
// least S stripes, each of >= 64 rows).  Member regions (CWP_SCR_ROWS rows of
// rowcap each): rows 0/1 the top row (member 0, written by fill_top), rows
// 2/3 and 4/5 two output slots used in turn; stripe j of member j%S writes
// slot (j/S)&1 and stripe j+1 reads it.  A member starts stripe j+2S only
// after its stripe j+S, which runs behind stripe j+1, so a slot is never
// overwritten before it has been read.  prog[j] counts stripe j's flushed
// chunks.
constexpr int kMaxSplitStripes = 64;
constexpr int kMaxSplitNodes = CWP_WAVES / 4;    // nodes of a level whose halves can be split

template <int MODE>
__device__ void psweep_part(const PSweep& W, int rmax, int m, int S, int* const* regions, int rowcap,
                            int* finH, int* finE, int* prog, int* err)
{
    const int nrows = W.nrows;
    const int per = WAVE * rmax;
    int nst = uni((nrows + per - 1) / per);
    if (nst < S) nst = S;
    const int base = nrows / nst, rem = nrows % nst;
    int row0 = 0;
    for (int j = 0; j < nst; ++j) {
        const int P = uni(base + (j < rem ? 1 : 0));
        if (j % S == m) {
            const int R = uni((P + WAVE - 1) / WAVE);
            const int* inH;
            const int* inE;
            if (j == 0) {
                inH = regions[0];
                inE = regions[0] + rowcap;
            } else {
                const int slot = ((j - 1) / S) & 1;
                inH = regions[(j - 1) % S] + (2 + 2 * slot) * rowcap;
                inE = inH + rowcap;
            }
            int* oH;
            int* oE;
            if (j == nst - 1) {
                oH = finH;
                oE = finE;
            } else {
                oH = regions[m] + (2 + 2 * ((j / S) & 1)) * rowcap;
                oE = oH + rowcap;
            }
            const StripeSync sy{j > 0 ? prog + (j - 1) : nullptr, j < nst - 1 ? prog + j : nullptr, err};
            run_pstripe<MODE>(R, false, W, row0, P, j == 0, inH, inE, oH, oE, sy);
            wave_fence();
        }
        row0 += P;
    }
}

// Member m of the G one-wave workgroups of a half sweep in cw_prf_wide runs
// stripe m of G (64..512 rows each).  Member regions as in psweep_part, one
// output slot each (member r's region at reg0 + r * CWP_SCR_ROWS * rowcap);
// prog[j] counts stripe j's flushed chunks, in global memory.  Stripe m waits
// only on stripe m-1, i.e. on a workgroup of lower index.
template <int MODE>
__device__ void psweep_wide(const PSweep& W, int m, int G, int* reg0, int rowcap, int* finH, int* finE,
                            int* prog, int* err, int limit)
{
    const int nrows = W.nrows;
    const int base = nrows / G, rem = nrows % G;
    const int row0 = uni(m * base + (m < rem ? m : rem));
    const int P = uni(base + (m < rem ? 1 : 0));
    const int R = uni((P + WAVE - 1) / WAVE);
    const int64_t stride = (int64_t)CWP_SCR_ROWS * rowcap;
    const int* inH = m == 0 ? reg0 : reg0 + (m - 1) * stride + 2 * rowcap;
    const int* inE = inH + rowcap;
    int* oH = m == G - 1 ? finH : reg0 + m * stride + 2 * rowcap;
    int* oE = m == G - 1 ? finE : oH + rowcap;
    const StripeSync sy{m > 0 ? prog + (m - 1) : nullptr, m < G - 1 ? prog + m : nullptr, err, limit};
    run_pstripe_x<MODE>(R, W, row0, P, m == 0, inH, inE, oH, oE, sy);
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
        const int4 re = ld4(&p.row[A + M]);
        for (int c = 1 + lane; c <= N; c += WAVE) {
            const int J = B + N - c + 1;
            h0[c] = -(re.z & (re.x + p.col[J].x)) - (re.z & (p.pl2[B + N + 1] - p.pl2[J]));
        }
    }
    wave_fence();
}

// A node of the recursion (one call of progDiff) outside the local subtrees is
// done in two phases by whole waves: its two half sweeps as independent tasks
// (so two waves work on one node), then, after a workgroup barrier, its
// midpoint.  The half sweeps leave their final rows in the level buffers at
// offset B + k (k = the node's position in its level): the nodes of a level
// cover disjoint column ranges [B+1, B+N] in order, so these slices do not
// overlap.  Every lane runs every statement of these functions and writes the
// (identical) node fields: the task loops of cw_prf_mm must contain no
// lane-divergent branch, or the compiler's structurization can leave some lanes
// cycling through the loop header (whose shuffle then reads an inactive lane)
// while others wait at its exit.  Kept out of line for the same reason.

struct LevelBufs {
    int* fH;   // forward final row H, [B+k+j], j = 1..N
    int* fE;   // forward final row E
    int* rH;   // reverse final row H, [B+k+c], reversed column c = 1..N
    int* rK;   // reverse final row K
};

__device__ __forceinline__ LevelBufs level_bufs(int* lvl, int lcap, int B, int k)
{
    return LevelBufs{lvl + B + k, lvl + lcap + B + k, lvl + 2 * lcap + B + k, lvl + 3 * lcap + B + k};
}

// Leaves (:540-600): N <= 0, M <= 0, or M == 1.
__device__ __attribute__((noinline)) void do_leaf(const Prob& p, int* nd, int is_root, int* out_score)
{
    const int lane = lane_id();
    const int A = uni(nd[0]), B = uni(nd[1]), M = uni(nd[2]), N = uni(nd[3]);
    const int go1 = uni(nd[4]) & 1;
    int midj = 0, midh;
    if (N <= 0) {
        midh = is_root ? -gap_pen1(p, A, B, M) : 0;      // only the root's score is kept
    } else if (M <= 0) {
        midh = is_root ? -gap_pen2(p, A, B, N) : 0;
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
}

// The forward (half 0, rows 1..midi) or reverse (half 1, rows M..midi+1)
// half sweep of node (A, B, M, N) with gap-open flags `flags`.
__device__ __forceinline__ PSweep half_def(const Prob& p, int A, int B, int M, int N, int flags, int half,
                                          int* err)
{
    const int go1 = flags & 1, go2 = (flags >> 1) & 1;
    const int midi = M / 2;
    PSweep W;
    W.p = p;
    W.ncols = N;
    W.err = (lint*)err;
    if (half == 0) {
        W.nrows = midi;
        W.ib = A; W.id = 1; W.jb = B; W.jd = 1;
        W.lt0 = go1 ? -open2(p, A + 1, B) : 0;
        W.ltl = -ext2(p, A + 1, B);
        W.go2 = 0; W.jedge = 0;
    } else {
        W.nrows = M - midi;
        W.ib = A + M + 1; W.id = -1; W.jb = B + N + 1; W.jd = -1;
        W.lt0 = 0; W.ltl = 0;
        W.go2 = go2; W.jedge = B + N;
    }
    return W;
}

// Forward or reverse sweep of node nd (M >= 2, N >= 1) into the level buffers.
__device__ __attribute__((noinline)) void do_half(const Prob& p, const int* nd, int half, int rmax,
                                                  int* scr, int rowcap, LevelBufs lv, int m, int S,
                                                  int* const* regions, int* prog, int* err)
{
    const int A = uni(nd[0]), B = uni(nd[1]), M = uni(nd[2]), N = uni(nd[3]);
    const int flags = uni(nd[4]);
    const Bufs g{scr, scr + rowcap, scr + 2 * rowcap, scr + 3 * rowcap};
    if (half == 0) {
        const PSweep F = half_def(p, A, B, M, N, flags, 0, err);
        if (S > 1) {
            if (m == 0) fill_top<MODE_FWD>(p, A, B, M, N, regions[0]);
            psweep_part<MODE_FWD>(F, rmax, m, S, regions, rowcap, lv.fH, lv.fE, prog, err);
        } else {
            fill_top<MODE_FWD>(p, A, B, M, N, g.h0);
            psweep<MODE_FWD>(F, rmax, g, lv.fH, lv.fE);
        }
    } else {
        const PSweep Rv = half_def(p, A, B, M, N, flags, 1, err);
        if (S > 1) {
            if (m == 0) fill_top<MODE_REV>(p, A, B, M, N, regions[0]);
            psweep_part<MODE_REV>(Rv, rmax, m, S, regions, rowcap, lv.rH, lv.rK, prog, err);
        } else {
            fill_top<MODE_REV>(p, A, B, M, N, g.h0);
            psweep<MODE_REV>(Rv, rmax, g, lv.rH, lv.rK);
        }
    }
    wave_fence();
}

// The midpoint of node nd (:756-776) from its half sweeps' final rows.
__device__ __attribute__((noinline)) void do_mid(const Prob& p, int* nd, int is_root, LevelBufs lv,
                                                 int* out_score)
{
    const int lane = lane_id();
    const int A = uni(nd[0]), B = uni(nd[1]), M = uni(nd[2]), N = uni(nd[3]);
    const int flags = uni(nd[4]);
    const int go1 = flags & 1, go2 = (flags >> 1) & 1;
    const int midi = M / 2;
    // HH[0]: left column of forward row midi; RR[N]: left column of reverse row M-midi
    const int HH0 = (go1 ? -open2(p, A + 1, B) : 0) - midi * ext2(p, A + 1, B);
    const int Ie = A + midi + 1;
    const int4 ce = ld4(&p.col[B + N]);
    const int RRN = (go2 ? -(ce.z & (p.row[Ie].x + ce.x)) : 0) - (ce.z & (p.pl1[A + M + 1] - p.pl1[Ie]));
    const int* fH = lv.fH;
    const int* fE = lv.fE;
    const int* rH = lv.rH;
    const int* rK = lv.rK;
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

// The two children of a split node (:796-809): type 1 splits at (midi, midj),
// type 2 through the gap of rows midi..midi+1 (the del(2) between them is the
// host's).  Records: A B M N flags type midj child.
template <typename Ptr>
__device__ __forceinline__ void make_children(Ptr c0, Ptr c1, int A, int B, int M, int N, int fl,
                                              int type, int midj)
{
    const int go1 = fl & 1, go2 = (fl >> 1) & 1;
    const int midi = M / 2;
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
}

// ---- subtrees solved in LDS ---------------------------------------------------
// A node with at most 2*64 rows and 256 columns whose data fits the wave's
// CWP_LOCAL_INTS is solved with all its descendants by one wave, depth first,
// without returning to the level loop.  Its profile rows, column records,
// prefix sums and score block are copied to LDS once; every node below it
// reads only rows A0..A0+M0+1 and columns B0..B0+N0+1 (the leaves' N<=0 / M<=0
// scores, which read further, are needed for the root only).  Records of the
// subtree stay in LDS and go to the node table in one pass at the end; child
// slots come from the workgroup's slot counter.
//
// A local node runs both half sweeps at once with lanes owning columns: lanes
// 0..31 the forward half (rows 1..midi from the corner (A,B)), lanes 32..63
// the reverse half (rows 1..M-midi from (A+M+1,B+N+1)), C = 1, 2, 4 or 8
// columns per lane.  At step t lane L works on row t-L of its half; the H and
// F (forward) or G (reverse) of its last column move to lane L+1 by DPP.  The
// top rows and left columns are the reference's boundary values (:636-693),
// computed in registers.

constexpr int kLocalRec = 9;       // A B M N flags type midj child slot
constexpr int kLocalStack = 64;    // depth-first stack (record indices)
constexpr int kLocalMaxRows = 2 * WAVE;
constexpr int kLocalMaxCols = 32 * 8;
constexpr int kLocalFlag = 4;      // flags bit: subtree solved locally (level loop skips it)

// A profile pair's data for rows A0..A0+M0+1 and columns B0..B0+N0+1 in LDS;
// accessors take global indices.
struct LProb {
    int L1, L2, A0, B0, M0;
    lint* rw;          // row records, 4 ints each
    lint* cl;          // column records, 4 ints each
    lint* p1;          // pl1
    lint* p2;          // pl2
    lint* sc;          // scores, sc[(J-B0-1)*M0 + I-A0-1]
    __device__ int4 row(int I) const {
        const lint* r = rw + 4 * (I - A0);
        return make_int4(r[0], r[1], r[2], r[3]);
    }
    __device__ int4 col(int J) const {
        const lint* c = cl + 4 * (J - B0);
        return make_int4(c[0], c[1], c[2], c[3]);
    }
    __device__ int pl1(int x) const { return p1[x - A0]; }
    __device__ int pl2(int x) const { return p2[x - B0]; }
    __device__ int S(int I, int J) const { return sc[(J - B0 - 1) * M0 + (I - A0 - 1)]; }
};

// gapPenalty1/2 (:1166-1189, :1239-1262) on the LDS copy
__device__ __forceinline__ int lgap_pen1(const LProb& q, int i, int j, int k) {
    if (k <= 0) return 0;
    const int4 r = q.row(i);
    if (r.z == 0) return 0;
    const int e = j + k < q.L2 ? j + k : q.L2;
    return r.x + q.col(j).x + (j < q.L2 ? q.pl2(e) - q.pl2(j) : 0);
}
__device__ __forceinline__ int lgap_pen2(const LProb& q, int i, int j, int k) {
    if (k <= 0) return 0;
    const int4 c = q.col(j);
    if (c.z == 0) return 0;
    const int e = i + k < q.L1 ? i + k : q.L1;
    return q.row(i).x + c.x + (i < q.L1 ? q.pl1(e) - q.pl1(i) : 0);
}

__device__ __forceinline__ int local_ints(int M0, int N0)
{
    return 4 * (M0 + 2) + 4 * (N0 + 2) + (M0 + 2) + (N0 + 2) + M0 * N0 + 4 * (N0 + 2) +
           (kLocalRec + 1) * (2 * M0 + 4) + kLocalStack + 2;
}

__device__ __forceinline__ bool local_fits(int M, int N)
{
    return M >= 2 && M <= kLocalMaxRows && N >= 1 && N <= kLocalMaxCols &&
           local_ints(M, N) <= CWP_LOCAL_INTS;
}

// The final rows of both halves of node (A,B,M,N): fH/fE[j] = HH/DD of row
// midi (j = 1..N), rH/rK[c] = RR and K of reverse row M-midi at reversed
// column c (j = N+1-c).
template <int C>
__device__ void lhalves(const LProb& q, int A, int B, int M, int N, int go1, int go2, lint* fH,
                        lint* fE, lint* rH, lint* rK)
{
    const int lane = lane_id();
    const int grp = lane >> 5;                 // 0: forward half, 1: reverse half
    const int L = lane & 31;
    const int midi = M / 2;
    const int rows = grp ? (M - midi) : midi;
    const int nl = (N + C - 1) / C;            // lanes in use per half
    // forward top row H(0,j) = t0 + j*tl; left column H(i,0) = lt0 + i*ltl
    const int4 ra = q.row(A), rb1 = q.row(A + 1);
    const int4 cb = q.col(B), cb1 = q.col(B + 1);
    const int t0 = -(ra.z & (ra.x + cb1.x)), tl = -(ra.z & cb1.y);
    const int lt0 = go1 ? -(cb.z & (rb1.x + cb.x)) : 0, ltl = -(cb.z & rb1.y);
    // reverse top row: -(m1 & (g1 + g2[J])) - (m1 & (pl2[B+N+1] - pl2[J])) at row A+M;
    // left column: (go2 ? -open2(I, B+N) : 0) - (m2[B+N] & (pl1[A+M+1] - pl1[I]))
    const int4 rm = q.row(A + M);
    const int4 ce = q.col(B + N);
    const int p2e = q.pl2(B + N + 1), p1e = q.pl1(A + M + 1);
    int4 cF[C], cE[C];
    int hup[C], vst[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int jj = L * C + c + 1;
        const int jc = jj <= N ? jj : N;
        const int J = grp ? (B + N + 1 - jc) : (B + jc);
        cF[c] = q.col(J);
        cE[c] = grp ? q.col(J - 1) : cF[c];
        hup[c] = grp ? (-(rm.z & (rm.x + cF[c].x)) - (rm.z & (p2e - q.pl2(J)))) : (t0 + jj * tl);
        vst[c] = NEG;
    }
    int dleft;
    if (L == 0) {
        dleft = 0;
    } else {
        const int jl = L * C;                  // H(0, jl) of this lane's left neighbour column
        const int Jl = grp ? (B + N + 1 - (jl <= N ? jl : N)) : 0;
        dleft = grp ? (-(rm.z & (rm.x + q.col(Jl).x)) - (rm.z & (p2e - q.pl2(Jl)))) : (t0 + jl * tl);
    }
    // S offsets of the lane's columns: S(I, J_c) = sc[so[c] + I]
    int so[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int jj = L * C + c + 1;
        const int jc = jj <= N ? jj : N;
        const int J = grp ? (B + N + 1 - jc) : (B + jc);
        so[c] = (J - q.B0 - 1) * q.M0 - q.A0 - 1;
    }
    // row data of the lane's row (records, pl1, scores), one step ahead
    auto rowI = [&](int i) {
        const int ic = i < 1 ? 1 : (i > rows ? rows : i);
        return grp ? (A + M + 1 - ic) : (A + ic);
    };
    int4 nrf, nre;
    int npl;
    int ns[C];
    auto fetch = [&](int i) {
        const int I = rowI(i);
        nre = q.row(I);
        nrf = q.row(grp ? I - 1 : I);
        npl = q.pl1(I);
#pragma unroll
        for (int c = 0; c < C; ++c) ns[c] = q.sc[so[c] + I];
    };
    fetch(1 - L);
    // H of the left column at the lane-0 rows: forward lt0 + i*ltl, reverse
    // (go2 ? -open2(I, B+N) : 0) - (m2[B+N] & (pl1[A+M+1] - pl1[I])) with I the
    // lane's own row, whose record and pl1 the fetch brought; lprev: row i-1
    const int go2m = go2 ? -1 : 0;
    int lprev = 0;
    int hl_out = 0, fl_out = NEG;
    const int T = (midi > M - midi ? midi : M - midi) + nl - 1;
    for (int t = 1; t <= T; ++t) {
        int hl = shr1(0, hl_out);
        int fl = shr1(NEG, fl_out);
        const int i = t - L;
        const int4 rf = nrf, re = nre;
        const int plc = npl;
        int s[C];
#pragma unroll
        for (int c = 0; c < C; ++c) s[c] = ns[c];
        fetch(i + 1);
        if (L == 0) {
            const int lv = i <= 0 ? 0
                           : grp ? (-(go2m & ce.z & (re.x + ce.x)) - (ce.z & (p1e - plc)))
                                 : lt0 + i * ltl;
            hl = lv;
            fl = NEG;
            dleft = lprev;
            lprev = lv;
        }
        if (i >= 1 && i <= rows && L < nl) {
            int d = dleft, hc = hl, fc = fl;
            // one recurrence for both halves (the gap masks are 0 or all ones, so
            // (x + y) & m = (x & m) + (y & m) and max(a - o - e, b - e) = max(a - o, b) - e):
            //   forward: F = max(H - o1, F) - e1 enters H as F
            //   reverse: G = max(H, G) - e1     enters H as G - o1   (likewise E / K)
            // fm selects where the opening penalty goes
            const int fm = grp ? 0 : -1;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int o1 = (rf.x + cF[c].x) & rf.z, e1 = cF[c].y & rf.z;
                const int o2 = (re.x + cE[c].x) & cE[c].z, e2 = re.y & cE[c].z;
                const int st1 = max(hc - (o1 & fm), fc) - e1;
                const int st2 = max(hup[c] - (o2 & fm), vst[c]) - e2;
                const int h = max3(d + s[c], st1 - (o1 & ~fm), st2 - (o2 & ~fm));
                vst[c] = st2;
                fc = st1;
                d = hup[c];
                hup[c] = h;
                hc = h;
            }
            dleft = hl;
            hl_out = hc;
            fl_out = fc;
        }
    }
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int jj = L * C + c + 1;
        if (jj <= N) {
            if (grp == 0) { fH[jj] = hup[c]; fE[jj] = vst[c]; }
            else { rH[jj] = hup[c]; rK[jj] = vst[c]; }
        }
    }
}

// One node of a local subtree (progDiff on LDS data): leaves as in do_leaf,
// internal nodes by lhalves and the midpoint of :756-776.  Writes type and
// midj to the record; *midh_out gets the score (used for the root only).
__device__ __attribute__((noinline)) void local_node(const LProb& q, lint* rd, int is_root, lint* rows4,
                                                     int rstride, int* out_score)
{
    const int lane = lane_id();
    const int A = uni(rd[0]), B = uni(rd[1]), M = uni(rd[2]), N = uni(rd[3]);
    const int flags = uni(rd[4]);
    const int go1 = flags & 1, go2 = (flags >> 1) & 1;
    if (N <= 0 || M <= 1) {                                        // :540-600
        int midj = 0, midh = 0;
        if (N <= 0 || M <= 0) {
            // only the root's score is kept; a local root has M >= 2 and N >= 1
        } else {
            const int midh0 = go1 == 0 ? -lgap_pen1(q, A + 1, B + 1, N)
                                       : -lgap_pen2(q, A + 1, B, 1) - lgap_pen1(q, A + 1, B + 1, N);
            int bv = -BIG, bj = BIG;
            for (int j = 1 + lane; j <= N; j += WAVE) {
                const int v = -lgap_pen1(q, A, B + 1, j - 1) + q.S(A + 1, B + j) -
                              lgap_pen1(q, A + 1, B + j + 1, N - j);
                if (v > bv) { bv = v; bj = j; }
            }
            const int V = wave_max(bv);
            const int J = wave_min(bv == V ? bj : BIG);
            if (V > midh0) { midh = V; midj = J; }
            else midh = midh0;
        }
        rd[5] = 0;
        rd[6] = midj;
        rd[7] = -1;
        if (is_root) *out_score = midh;
        return;
    }
    const int midi = M / 2;
    lint* fH = rows4;
    lint* fE = rows4 + rstride;
    lint* rH = rows4 + 2 * rstride;
    lint* rK = rows4 + 3 * rstride;
    if (N <= 32) lhalves<1>(q, A, B, M, N, go1, go2, fH, fE, rH, rK);
    else if (N <= 64) lhalves<2>(q, A, B, M, N, go1, go2, fH, fE, rH, rK);
    else if (N <= 128) lhalves<4>(q, A, B, M, N, go1, go2, fH, fE, rH, rK);
    else lhalves<8>(q, A, B, M, N, go1, go2, fH, fE, rH, rK);
    wave_fence();
    // HH[0] = left column of forward row midi; RR[N] = left column of reverse row M-midi
    const int4 rb1 = q.row(A + 1);
    const int4 cb = q.col(B), ce = q.col(B + N);
    const int HH0 = (go1 ? -(cb.z & (rb1.x + cb.x)) : 0) + midi * (-(cb.z & rb1.y));
    const int Ie = A + midi + 1;                 // I of reverse row M-midi
    const int RRN = (go2 ? -(ce.z & (q.row(Ie).x + ce.x)) : 0) - (ce.z & (q.pl1(A + M + 1) - q.pl1(Ie)));
    const int4 rmid = q.row(A + midi + 1);
    int vmax = -BIG, jfirst = BIG, jc = -1, wmax = -BIG, jw = -1;
    for (int j = lane; j <= N; j += WAVE) {
        const int HHj = (j == 0) ? HH0 : fH[j];
        const int DDj = (j == 0) ? HH0 : fE[j];
        const int4 cj = q.col(B + j);
        const int gS = cj.z & (rmid.x + cj.x);
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
    rd[5] = type;
    rd[6] = midj;
    if (is_root) *out_score = midh;
}

// ---- tiny nodes, one lane each -----------------------------------------------
// Leaves and nodes of at most kTinyRows x kTinyCols are not worth a wave: the
// wave keeps them in an LDS queue and every lane runs one of them as scalar
// code, the columns unrolled in registers (column j of the forward half and of
// the reverse half at the same index), the midpoint as the reference's own
// sequential scan over j = 0..N.  A split node appends its two children to
// the queue.

constexpr int kTinyRows = 16;
constexpr int kTinyCols = 16;

__device__ __forceinline__ bool is_tiny(int M, int N)
{
    return M <= 1 || N <= 0 || (M <= kTinyRows && N <= kTinyCols);
}

// One node by one lane.  ctr[0]: records used, ctr[1]: queue tail.  A split
// node's children are records of the subtree (node-table slots come at the
// subtree's write-out).
__device__ __forceinline__ void tiny_node(const LProb& q, lint* recs, int ri, lint* lq, lint* ctr,
                                          int rec_cap, int* s_err, int* out_score)
{
    lint* rd = recs + kLocalRec * ri;
    const int A = rd[0], B = rd[1], M = rd[2], N = rd[3], fl = rd[4];
    const int go1 = fl & 1, go2 = (fl >> 1) & 1;
    const bool root = rd[8] == 0;
    if (N <= 0 || M <= 1) {                                        // :540-600
        int midj = 0, midh = 0;
        if (N > 0 && M > 0) {
            const int midh0 = go1 == 0 ? -lgap_pen1(q, A + 1, B + 1, N)
                                       : -lgap_pen2(q, A + 1, B, 1) - lgap_pen1(q, A + 1, B + 1, N);
            int bv = -BIG, bj = 0;
            for (int j = 1; j <= N; ++j) {
                const int v = -lgap_pen1(q, A, B + 1, j - 1) + q.S(A + 1, B + j) -
                              lgap_pen1(q, A + 1, B + j + 1, N - j);
                if (v > bv) { bv = v; bj = j; }
            }
            if (bv > midh0) { midh = bv; midj = bj; }
            else midh = midh0;
        }
        rd[5] = 0;
        rd[6] = midj;
        rd[7] = -1;
        if (root) *out_score = midh;
    } else {
        constexpr int NT = kTinyCols;
        const int midi = M / 2;
        int cx[NT + 1], cy[NT + 1], cz[NT + 1];          // g2, l2, m2 of column B+k
#pragma unroll
        for (int k = 0; k <= NT; ++k) {
            const int4 c = q.col(B + (k <= N ? k : N));
            cx[k] = c.x;
            cy[k] = c.y;
            cz[k] = c.z;
        }
        // forward half (:636-680): rows 1..midi from corner (A,B)
        const int4 ra = q.row(A), rb1 = q.row(A + 1);
        const int t0 = -(ra.z & (ra.x + cx[1])), tl = -(ra.z & cy[1]);
        const int lt0 = go1 ? -(cz[0] & (rb1.x + cx[0])) : 0, ltl = -(cz[0] & rb1.y);
        int H[NT + 1], E[NT + 1];
        H[0] = 0;
        E[0] = NEG;
#pragma unroll
        for (int k = 1; k <= NT; ++k) {
            H[k] = t0 + k * tl;
            E[k] = NEG;
        }
        for (int i = 1; i <= midi; ++i) {
            const int4 r = q.row(A + i);
            int d = H[0], hc = lt0 + i * ltl, f = NEG;
            H[0] = hc;
#pragma unroll
            for (int k = 1; k <= NT; ++k) {
                if (k <= N) {
                    const int sv = q.S(A + i, B + k);
                    const int fn = max(hc - ((r.x + cx[k] + cy[k]) & r.z), f - (cy[k] & r.z));
                    const int en = max(H[k] - ((r.w + cx[k]) & cz[k]), E[k] - (r.y & cz[k]));
                    const int h = max3(d + sv, fn, en);
                    d = H[k];
                    H[k] = h;
                    E[k] = en;
                    hc = h;
                    f = fn;
                }
            }
        }
        // reverse half (:684-747): rows M..midi+1 from corner (A+M+1,B+N+1);
        // RH/RK[k] hold column B+k
        const int4 rm = q.row(A + M);
        const int4 ce = q.col(B + N);
        const int p1e = q.pl1(A + M + 1), p2e = q.pl2(B + N + 1);
        auto rleft = [&](int r) -> int {
            if (r <= 0) return 0;
            const int I = A + M + 1 - r;
            const int g = go2 ? -(ce.z & (q.row(I).x + ce.x)) : 0;
            return g - (ce.z & (p1e - q.pl1(I)));
        };
        int RH[NT + 2], RK[NT + 2];
        RH[0] = RK[0] = RH[NT + 1] = RK[NT + 1] = 0;
#pragma unroll
        for (int k = 1; k <= NT; ++k) {
            const int kk = k <= N ? k : N;
            RH[k] = -(rm.z & (rm.x + cx[k])) - (rm.z & (p2e - q.pl2(B + kk)));
            RK[k] = NEG;
        }
        const int R = M - midi;
        for (int r = 1; r <= R; ++r) {
            const int I = A + M + 1 - r;
            const int4 rf = q.row(I - 1), re = q.row(I);
            int hc = rleft(r), g = NEG, d = rleft(r - 1);
#pragma unroll
            for (int k = NT; k >= 1; --k) {
                if (k <= N) {
                    const int sv = q.S(I, B + k);
                    const int gn = max(hc, g) - (cy[k] & rf.z);
                    const int kn = max(RH[k], RK[k]) - (re.y & cz[k - 1]);
                    const int h = max3(d + sv, gn - ((rf.x + cx[k]) & rf.z),
                                       kn - ((re.x + cx[k - 1]) & cz[k - 1]));
                    d = RH[k];
                    RH[k] = h;
                    RK[k] = kn;
                    hc = h;
                    g = gn;
                }
            }
        }
        // midpoint (:756-776) in the reference's scan order
        const int HH0 = H[0];
        const int RRN = rleft(R);
        const int gx = q.row(A + midi + 1).x;
        int vmax = -BIG, jfirst = 0, jc = -1, wmax = -BIG, jw = -1;
#pragma unroll
        for (int j = 0; j <= NT; ++j) {
            if (j <= N) {
                const int HHj = (j == 0) ? HH0 : H[j];
                const int DDj = (j == 0) ? HH0 : E[j];
                const int gS = cz[j] & (gx + cx[j]);
                int RRj, SSj;
                if (j == N) { RRj = RRN; SSj = RRN; }
                else { RRj = RH[j + 1]; SSj = RK[j + 1] - gS; }
                const int v = HHj + RRj;
                const int w = DDj + SSj + gS;
                if (v > vmax) { vmax = v; jfirst = j; jc = -1; }
                if (v == vmax && HHj != DDj && RRj == SSj) jc = j;
                if (w >= wmax) { wmax = w; jw = j; }
            }
        }
        int type, midj, midh;
        if (wmax > vmax) { type = 2; midj = jw; midh = wmax; }
        else { type = 1; midj = jfirst > jc ? jfirst : jc; midh = vmax; }
        rd[5] = type;
        rd[6] = midj;
        rd[7] = -1;
        if (root) *out_score = midh;
        const int nr = lds_add(&ctr[0], 2);
        const int qp = lds_add(&ctr[1], 2);
        if (nr + 2 > rec_cap || qp + 2 > rec_cap) {
            *s_err = 1;
        } else {
            lint* c0 = recs + kLocalRec * nr;
            lint* c1 = c0 + kLocalRec;
            make_children(c0, c1, A, B, M, N, fl, type, midj);
            c0[8] = -1;                              // not the problem's root
            c1[8] = -1;
            rd[7] = nr;                              // record index until the write-out
            lq[qp] = nr;
            lq[qp + 1] = nr + 1;
        }
    }
}

__device__ __attribute__((noinline)) void local_subtree(const Prob& p, int* nodes, int root, int* base,
                                                        int slot0, int* s_err, int* out_score)
{
    const int lane = lane_id();
    gcint* rnd = (gcint*)nodes + (int64_t)root * CWP_NODE_INTS;
    const int A0 = uni(rnd[0]), B0 = uni(rnd[1]), M0 = uni(rnd[2]), N0 = uni(rnd[3]);
    const int fl0 = uni(rnd[4]);
    lint* lb = (lint*)base;
    lint* lrow = lb;                                 // rows A0 .. A0+M0+1, 4 ints each
    lint* lcol = lrow + 4 * (M0 + 2);                // columns B0 .. B0+N0+1
    lint* lpl1 = lcol + 4 * (N0 + 2);
    lint* lpl2 = lpl1 + (M0 + 2);
    lint* lS = lpl2 + (N0 + 2);                      // lS[(J-B0-1)*M0 + I-A0-1]
    lint* lout = lS + M0 * N0;                       // final rows of a node: 4 x (N0 + 2)
    lint* recs = lb + (4 * (M0 + 2) + 4 * (N0 + 2) + (M0 + 2) + (N0 + 2) + M0 * N0 + 4 * (N0 + 2));
    const int rec_cap = 2 * M0 + 4;
    lint* stk = recs + kLocalRec * rec_cap;
    lint* lq = stk + kLocalStack;                    // tiny-node queue (record indices)
    lint* ctr = lq + rec_cap;                        // records used, queue tail

    // copy in, 16 loads in flight per lane before the LDS stores
    for (int x = lane; x < M0 + 2; x += WAVE) {
        const int4 r = ld4(&p.row[A0 + x]);
        lrow[4 * x] = r.x; lrow[4 * x + 1] = r.y; lrow[4 * x + 2] = r.z; lrow[4 * x + 3] = r.w;
        lpl1[x] = p.pl1[A0 + x];
    }
    for (int x = lane; x < N0 + 2; x += WAVE) {
        const int4 c = ld4(&p.col[B0 + x]);
        lcol[4 * x] = c.x; lcol[4 * x + 1] = c.y; lcol[4 * x + 2] = c.z; lcol[4 * x + 3] = c.w;
        lpl2[x] = p.pl2[B0 + x];
    }
    const int tot = M0 * N0;
    for (int k0 = 0; k0 < tot; k0 += 16 * WAVE) {
        int v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int k = k0 + u * WAVE + lane;
            const int c = k / M0, r = k - c * M0;
            v[u] = k < tot ? p.S[(int64_t)(B0 + 1 + c) * p.ldS + A0 + 1 + r] : 0;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int k = k0 + u * WAVE + lane;
            if (k < tot) lS[k] = v[u];
        }
    }
    LProb q;
    q.L1 = p.L1; q.L2 = p.L2; q.A0 = A0; q.B0 = B0; q.M0 = M0;
    q.rw = lrow; q.cl = lcol; q.p1 = lpl1; q.p2 = lpl2; q.sc = lS;
    lint* rows4 = lout;                              // 4 rows of stride N0 + 2

    recs[0] = A0; recs[1] = B0; recs[2] = M0; recs[3] = N0; recs[4] = fl0;
    recs[5] = 0; recs[6] = 0; recs[7] = -1; recs[8] = root;
    stk[0] = 0;
    wave_fence();
    int sp = 1, nrec = 1, nq = 0;
    while (sp > 0) {
        sp -= 1;
        const int ri = uni(stk[sp]);
        lint* rd = recs + kLocalRec * ri;
        if (is_tiny(uni(rd[2]), uni(rd[3]))) {
            lq[nq] = ri;                             // same value from every lane
            nq += 1;
            wave_fence();
            continue;
        }
        local_node(q, rd, uni(rd[8]) == 0, rows4, N0 + 2, out_score);
        wave_fence();
        const int type = uni(rd[5]);
        if (type != 0) {
            if (nrec + 2 > rec_cap || sp + 2 > kLocalStack) {
                *s_err = 1;                          // same value from every lane
                break;
            }
            lint* c0 = recs + kLocalRec * nrec;
            lint* c1 = c0 + kLocalRec;
            make_children(c0, c1, uni(rd[0]), uni(rd[1]), uni(rd[2]), uni(rd[3]), uni(rd[4]), type,
                          uni(rd[6]));
            c0[8] = -1;                              // not the problem's root
            c1[8] = -1;
            rd[7] = nrec;                            // record index until the write-out
            stk[sp] = nrec + 1;
            stk[sp + 1] = nrec;
            sp += 2;
            nrec += 2;
            wave_fence();
        }
    }
    // the tiny nodes, one per lane, until the queue is empty
    ctr[0] = nrec;
    ctr[1] = nq;
    wave_fence();
    int head = 0;
    for (;;) {
        const int tail = uni(ctr[1]);
        if (head >= tail) break;
        const int take = tail - head < WAVE ? tail - head : WAVE;
        if (lane < take)
            tiny_node(q, recs, lq[head + lane], lq, ctr, rec_cap, s_err, out_score);
        head += take;
        wave_fence();
    }
    nrec = uni(ctr[0]);
    // records out to the node table: the root keeps its slot (flagged as
    // done), records 1 .. nrec-1 take slots slot0 .. slot0+nrec-2 of the
    // 2 M0 - 2 that cw_prf_mm reserved, and child links become slots
    if (nrec - 1 > 2 * M0 - 2) {
        *s_err = 1;                                  // same value from every lane
        return;
    }
    for (int k = lane; k < nrec * CWP_NODE_INTS; k += WAVE) {
        const int r = k >> 3, f = k & 7;
        const lint* rc = recs + kLocalRec * r;
        int v = rc[f];
        if (f == 4 && r == 0) v |= kLocalFlag;
        if (f == 7 && v >= 0) v = slot0 + v - 1;
        ((gint*)nodes)[(int64_t)(r == 0 ? root : slot0 + r - 1) * CWP_NODE_INTS + f] = v;
    }
    wave_fence();
}

} // namespace cwp

using namespace cwp;

__device__ static Prob make_prob(const PrfArgs& a, const PrfProb& pp)
{
    Prob p;
    p.L1 = pp.L1;
    p.L2 = pp.L2;
    p.ldS = pp.ldS;
    p.row = (gcint4*)(a.pool + pp.row);
    p.col = (gcint4*)(a.pool + pp.col);
    p.pl1 = (gcint*)(a.pool + pp.pl1);
    p.pl2 = (gcint*)(a.pool + pp.pl2);
    p.S = (gcint*)(a.pool + pp.S);
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
// A subtree small enough for one wave's LDS is left to cw_prf_local: its root
// goes on the batch's list (lane 0 of the wave appends it).
// A subtree of a root with M >= 2 rows has at most 2M - 1 nodes (a split
// node's children have at most its rows between them; leaves have M <= 1), so
// its records other than the root fit 2M - 2 node-table slots, reserved here
// from the workgroup's count.  Reserving that much for every local subtree
// keeps the table at most 2 L1 - 1 nodes, as for the tree itself.
__device__ __attribute__((noinline)) static void defer_local(const PrfArgs& a, int prob, int idx, int M,
                                                             int* s_alloc, int cap, int* s_err)
{
    const int lane = lane_id();
    const int got = atomicAdd(s_alloc, lane == 0 ? 2 * M - 2 : 0);
    const int slot0 = uni(__shfl(got, 0, WAVE));
    if (slot0 + 2 * M - 2 > cap) {
        *s_err = 1;                                  // same value from every lane
        return;
    }
    if (lane == 0) {
        const int i = atomicAdd(&a.local_ctr[0], 1);
        a.local_list[3 * i] = prob;
        a.local_list[3 * i + 1] = idx;
        a.local_list[3 * i + 2] = slot0;
    }
}

extern "C" __global__ void __launch_bounds__(CWP_THREADS) cw_prf_mm(PrfArgs a)
{
    __shared__ int s_lev_b, s_lev_e, s_ctr, s_err, s_alloc;
    __shared__ int s_prog[kMaxSplitNodes * 2 * kMaxSplitStripes];
    const int prob = blockIdx.x;
    const PrfProb pp = a.probs[prob];
    const Prob p = make_prob(a, pp);
    int* nodes = a.pool + pp.nodes;
    const int wave = threadIdx.x / WAVE;
    const int lane = lane_id();
    int* scr = a.pool + pp.scratch + (int64_t)wave * CWP_SCR_ROWS * pp.rowcap;
    const int cap = pp.node_cap;
    int* state = a.state + 4 * prob;
    if (threadIdx.x == 0) {
        if (a.resume) {                          // the level a previous launch stopped at
            s_lev_b = state[0];
            s_lev_e = state[1];
            s_alloc = state[2];
            s_err = a.out[4 * prob + 2];
        } else {
            int* nd = nodes;
            nd[0] = 0; nd[1] = 0; nd[2] = pp.L1; nd[3] = pp.L2;
            nd[4] = pp.flags & 3; nd[5] = 0; nd[6] = 0; nd[7] = -1;
            s_lev_b = 0;
            s_lev_e = 1;
            s_alloc = 1;
            s_err = 0;
        }
    }
    __syncthreads();
    int* lvl = a.pool + pp.lvl;
    const int lcap = pp.L2 + 2 + cap;
    int levels = a.resume ? uni(state[3]) : 0;
    for (;;) {
        const int lb = s_lev_b, le = s_lev_e;
        if (lb >= le) break;
        // the half sweeps of this level are already in the level buffers:
        // the root's from cw_prf_wide, a later level's from cw_prf_level
        const bool swept = lb == 0 ? (pp.flags & CWP_ROOT_WIDE) != 0 : a.ext != 0;
        __syncthreads();
        for (int k = threadIdx.x; k < kMaxSplitNodes * 2 * kMaxSplitStripes; k += blockDim.x) s_prog[k] = 0;
        if (threadIdx.x == 0) s_ctr = 0;
        __syncthreads();
        // A level of at most kMaxSplitNodes large nodes gives each half to S
        // waves (pipelined stripes), S = CWP_WAVES / (2 x nodes) at most, so
        // every task has a wave of its own and all waves work.  Every stripe
        // keeps >= 64 rows (a stripe of one row per lane must fill the wave):
        // the smaller half has >= 64 S rows, and a stripe holds up to 128 or
        // more.
        int S = 1;
        if (!swept && le - lb <= kMaxSplitNodes) {
            int sk = CWP_WAVES / (2 * (le - lb));
            for (int k = lb; k < le; ++k) {
                const int* ndk = nodes + (int64_t)k * CWP_NODE_INTS;
                const int M0 = uni(ndk[2]), N0 = uni(ndk[3]);
                const int per = WAVE * a.rmax;
                const int half_rows = M0 - M0 / 2;
                int fit = (M0 / 2) / WAVE;
                if (local_fits(M0, N0) || N0 < 1 || a.rmax < 2 || (half_rows + per - 1) / per > kMaxSplitStripes)
                    fit = 1;
                sk = sk < fit ? sk : fit;
            }
            if (sk >= 2) S = sk;
        }
        // tasks 2S k .. 2S k + 2S - 1: the halves of node lb+k, S members each
        // (a local subtree or a leaf is one task)
        const int ntask = 2 * S * (le - lb);
        for (;;) {
            const int got = atomicAdd(&s_ctr, lane == 0 ? 1 : 0);
            const int t = uni(__shfl(got, 0, WAVE));
            if (t >= ntask) break;
            const int idx = lb + t / (2 * S), within = t % (2 * S);
            const int half = within / S, m = within % S;
            int* nd = nodes + (int64_t)idx * CWP_NODE_INTS;
            const int M = uni(nd[2]), N = uni(nd[3]);
            if (local_fits(M, N)) {
                if (within == 0) defer_local(a, prob, idx, M, &s_alloc, cap, &s_err);
            } else if (N <= 0 || M <= 1) {
                if (within == 0) do_leaf(p, nd, idx == 0, a.out + 4 * prob);
            } else if (!swept) {
                // a split half's members use the scratch of waves
                // (2 (idx - lb) + half) S .. + S - 1 (no wave sweeps alone then)
                const int grp = 2 * (idx - lb) + half;
                int* regions[CWP_WAVES / 2];
                for (int r = 0; r < CWP_WAVES / 2; ++r)
                    regions[r] = a.pool + pp.scratch +
                                 (int64_t)((grp * S + r) % CWP_WAVES) * CWP_SCR_ROWS * pp.rowcap;
                do_half(p, nd, half, a.rmax, scr, pp.rowcap, level_bufs(lvl, lcap, uni(nd[1]), idx - lb), m,
                        S, regions, s_prog + (S > 1 ? grp : 0) * kMaxSplitStripes, &s_err);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) s_ctr = lb;
        __syncthreads();
        for (;;) {
            const int got = atomicAdd(&s_ctr, lane == 0 ? 1 : 0);
            const int idx = uni(__shfl(got, 0, WAVE));
            if (idx >= le) break;
            int* nd = nodes + (int64_t)idx * CWP_NODE_INTS;
            const int M = uni(nd[2]), N = uni(nd[3]);
            if (!local_fits(M, N) && N >= 1 && M >= 2)
                do_mid(p, nd, idx == 0, level_bufs(lvl, lcap, uni(nd[1]), idx - lb), a.out + 4 * prob);
        }
        __syncthreads();
        if (wave == 0) {
            const int first = s_alloc;
            int next = first;
            for (int base = lb; base < le; base += WAVE) {
                const int idx = base + lane;
                int* nd = nodes + (int64_t)idx * CWP_NODE_INTS;
                const int type = (idx < le && !(nd[4] & kLocalFlag)) ? nd[5] : 0;
                const uint64_t bal = __ballot(type != 0);
                const int before = __popcll(bal & ((1ull << lane) - 1ull));
                const int slot = next + 2 * before;
                if (type != 0) {
                    if (slot + 2 <= cap) {
                        int* c0 = nodes + (int64_t)slot * CWP_NODE_INTS;
                        make_children(c0, c0 + CWP_NODE_INTS, nd[0], nd[1], nd[2], nd[3], nd[4], type,
                                      nd[6]);
                        nd[7] = slot;
                    } else {
                        s_err = 1;
                    }
                }
                next += 2 * __popcll(bal);
            }
            s_lev_b = first;                        // same value from every lane
            s_lev_e = next < cap ? next : cap;
            s_alloc = next < cap ? next : cap;
        }
        __syncthreads();
        ++levels;
        if (a.ext) break;                        // one level per launch: the next one's sweeps come first
    }
    if (threadIdx.x == 0) {
        state[0] = s_lev_b;
        state[1] = s_lev_e;
        state[2] = s_alloc;
        state[3] = levels;
        a.out[4 * prob + 1] = s_alloc;
        a.out[4 * prob + 2] = s_err;
        a.out[4 * prob + 3] = levels;
    }
}

// The half sweeps of one level of every pair of the batch, one wave per half
// (cw_prf_mm with PrfArgs::ext then finishes the level): workgroup b works on
// pair b / T, half (b % T) & 1 of node lev_b + (b % T) / 2 of the level the
// pair's state names, if that node is neither a leaf nor a local subtree.  T
// covers the level's nodes (at most 2^k at level k) and is at most the
// scratch sets a pair has.
extern "C" __global__ void __launch_bounds__(WAVE) cw_prf_level(PrfArgs a, int T)
{
    __shared__ int s_err;
    const int prob = blockIdx.x / T, t = blockIdx.x % T;
    const int* state = a.state + 4 * prob;
    const int lb = uni(state[0]), le = uni(state[1]);
    const int idx = lb + t / 2, half = t & 1;
    if (idx >= le) return;                       // the whole (one-wave) workgroup
    const PrfProb pp = a.probs[prob];
    int* nd = a.pool + pp.nodes + (int64_t)idx * CWP_NODE_INTS;
    const int M = uni(nd[2]), N = uni(nd[3]);
    if (local_fits(M, N) || N <= 0 || M <= 1) return;
    s_err = 0;                                   // every lane writes the same value
    __syncthreads();
    const Prob p = make_prob(a, pp);
    int* scr = a.pool + pp.scratch + (int64_t)t * CWP_SCR_ROWS * pp.rowcap;
    const LevelBufs lv = level_bufs(a.pool + pp.lvl, pp.L2 + 2 + pp.node_cap, uni(nd[1]), idx - lb);
    do_half(p, nd, half, a.rmax, scr, pp.rowcap, lv, 0, 1, nullptr, nullptr, &s_err);
    if (uni(s_err) != 0) a.out[4 * prob + 2] = 1;
}

// The half sweeps of the roots of large profile pairs (PrfProb::flags has
// CWP_ROOT_WIDE), before cw_prf_mm: one wave per workgroup, G workgroups per
// half, each running one stripe, so a pair's level 0 is spread over 2G CUs
// instead of the four waves of its cw_prf_mm workgroup.  Task b = {problem,
// half, member m, G}; the members of a half are consecutive workgroups.  The
// final rows go where cw_prf_mm's level 0 would leave them.
extern "C" __global__ void __launch_bounds__(WAVE) cw_prf_wide(PrfArgs a, WideArgs w)
{
    __shared__ int s_err;
    const int* tk = w.tasks + 4 * blockIdx.x;
    const int prob = uni(tk[0]), half = uni(tk[1]), m = uni(tk[2]), G = uni(tk[3]);
    s_err = 0;                                   // one wave: every lane writes the same value
    __syncthreads();
    const PrfProb pp = a.probs[prob];
    const Prob p = make_prob(a, pp);
    const int M = pp.L1, N = pp.L2;
    const int rowcap = pp.rowcap;
    int* reg0 = a.pool + pp.scratch + (int64_t)half * G * CWP_SCR_ROWS * rowcap;
    const LevelBufs lv = level_bufs(a.pool + pp.lvl, pp.L2 + 2 + pp.node_cap, 0, 0);
    int* prog = w.prog + (blockIdx.x - m);
    const PSweep W = half_def(p, 0, 0, M, N, pp.flags & 3, half, &s_err);
    if (half == 0) {
        if (m == 0) fill_top<MODE_FWD>(p, 0, 0, M, N, reg0);
        psweep_wide<MODE_FWD>(W, m, G, reg0, rowcap, lv.fH, lv.fE, prog, w.err, w.wait_limit);
    } else {
        if (m == 0) fill_top<MODE_REV>(p, 0, 0, M, N, reg0);
        psweep_wide<MODE_REV>(W, m, G, reg0, rowcap, lv.rH, lv.rK, prog, w.err, w.wait_limit);
    }
    wave_fence();
    if (uni(s_err) != 0) st_agent(w.err, 1);
}

// The subtrees cw_prf_mm listed, each solved by one wave in its LDS
// (local_subtree), taken from the list by the waves of all workgroups; node
// slots come from the problem's node count, which cw_prf_mm left in out[1].
extern "C" __global__ void __launch_bounds__(64 * CWP_LOCAL_WAVES) cw_prf_local(PrfArgs a)
{
    __shared__ __align__(16) int s_local[CWP_LOCAL_WAVES * CWP_LOCAL_INTS];
    const int wave = threadIdx.x / WAVE;
    const int lane = lane_id();
    const int total = uni(a.local_ctr[0]);
    for (;;) {
        const int got = atomicAdd(&a.local_ctr[1], lane == 0 ? 1 : 0);
        const int i = uni(__shfl(got, 0, WAVE));
        if (i >= total) break;
        const int prob = uni(a.local_list[3 * i]), idx = uni(a.local_list[3 * i + 1]);
        const int slot0 = uni(a.local_list[3 * i + 2]);
        const PrfProb pp = a.probs[prob];
        const Prob p = make_prob(a, pp);
        local_subtree(p, a.pool + pp.nodes, idx, s_local + wave * CWP_LOCAL_INTS, slot0,
                      a.out + 4 * prob + 2, a.out + 4 * prob);
    }
}

// ---- profile construction ---------------------------------------------------
// A code is a residue when it is 0..maxAA (24); gap codes are 30 and 31.
__device__ static inline bool pgap(int c) { return c > 24; }

__device__ static inline bool phyd(uint32_t mask, int c) { return !pgap(c) && ((mask >> c) & 1u); }

// (int) of a float as the reference's x86-64 build converts it (cvttss2si):
// NaN and values outside the int range give INT_MIN, where the GPU's
// conversion would clamp.  Reached when a group's weights sum to 0.
__device__ static inline int f2i_host(float v)
{
    return (v >= -2147483648.0f && v < 2147483648.0f) ? (int)v : (int)0x80000000u;
}

// The per-column sums over a group's rows that ProfileBase::calcGapCoeff
// gathers (gap counts between each row's first and last residue when end gaps
// are not counted, :129-160; hydrophilic stretches of more than 3 residues,
// calcHydrophilicPen :398-459; residue-specific penalties from the column's
// residue counts, calcResidueSpecificPen :360-391; the +-5 window identity of a
// two-row group, calcVPenalties :310-353), and the residue columns of the
// profile: ProfileWithSub::calcProfileWithSub (ProfileWithSub.cpp:42-147) for
// profile 1, ProfileStandard::calcStandardProfile (ProfileStandard.cpp:41-117)
// for profile 2.  A workgroup takes CWP_BUILD_COLS columns; each column's rows
// are split over CWP_BUILD_SLICES threads (one per wave) whose code counts meet
// in LDS.  All sums are integer, so the order over rows does not matter.
extern "C" __global__ void __launch_bounds__(CWP_BUILD_COLS * CWP_BUILD_SLICES) cw_prf_cols(BuildArgs a)
{
    constexpr int C = CWP_BUILD_COLS, NS = CWP_BUILD_SLICES;
    __shared__ int s_ws[NS * 32 * C];   // weighted code counts, [slice][code][column of the block]
    __shared__ int s_cn[NS * 32 * C];   // plain code counts
    __shared__ int s_gh[NS * 2 * C];    // gap counts and hydrophilic sums, [slice][2][column]
    int lo = 0, hi = a.nside - 1;       // the side of this block: the last one with col0 <= blockIdx
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (a.sides[mid].col0 <= (int)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const PrfSide sd = a.sides[lo];
    const int t = threadIdx.x % C, sl = threadIdx.x / C;
    const int pos = ((int)blockIdx.x - sd.col0) * C + t;
    const bool live = pos < sd.L;
    const int n = sd.n, Lp = sd.L + 2;
    int* ws = s_ws + sl * 32 * C;
    int* cn = s_cn + sl * 32 * C;
    for (int c = 0; c < 32; ++c) {
        ws[c * C + t] = 0;
        cn[c * C + t] = 0;
    }
    const uint8_t* base = a.rows + sd.rows;
    const int32_t* rinfo = a.pool + sd.rinfo;
    const bool hydp = (sd.flags & CWP_F_HYD) != 0;
    int gapc = 0, hyd = 0;
    const int k0 = (int)((int64_t)n * sl / NS), k1 = (int)((int64_t)n * (sl + 1) / NS);
    // A hydrophilic residue counts when it lies inside a stretch of more than
    // 3 hydrophilic residues (the reference looks up to 3 to the left, then
    // right for the rest).  A slice is one wave whose 64 lanes hold 64
    // consecutive columns of the row, so the stretch test reads the row's
    // hydrophilic flags from two ballots: the lanes' own columns and the 3
    // columns past each end of the block (lanes 0-2 and 61-63 load those).
    // Columns outside 0 .. L-1 count as not hydrophilic.
    const int col0 = pos - t;                // the block's first column
    const int ecol = t < 3 ? col0 - 3 + t : (t >= 61 ? col0 + 3 + t : -1);
    const bool eload = ecol >= 0 && ecol < sd.L;
    auto hyd_run = [&](const uint8_t* r, int c) -> bool {
        const bool h = phyd(a.hyd_mask, c);
        const uint64_t mid = __ballot(h);
        const uint64_t edge = __ballot(eload && phyd(a.hyd_mask, r[eload ? ecol : 0]));
        // bits 0-2: columns col0-3 .. col0-1; 3-66: the lanes; 67-69: col0+64 .. col0+66
        const uint64_t lo = (edge & 7ull) | (mid << 3);
        const uint64_t hi = (mid >> 61) | ((edge >> 61) << 3);
        // this lane's window: columns pos-3 .. pos+3 at bits 0 .. 6
        const uint64_t w = (t == 0 ? lo : ((lo >> t) | (hi << (64 - t)))) & 0x7Full;
        return h && ((w & 0x0F) == 0x0F || (w & 0x1E) == 0x1E || (w & 0x3C) == 0x3C ||
                     (w & 0x78) == 0x78);
    };
    auto one = [&](const uint8_t* r, int c, int wk, int f, int e) {
        if (c < 32) {
            ws[c * C + t] += wk;
            cn[c * C + t] += 1;
        }
        if (pgap(c) && pos >= f && pos < e) ++gapc;
        if (hydp && hyd_run(r, c)) hyd += 100;
    };
    if (live) {
        int k = k0;
        for (; k + 8 <= k1; k += 8) {        // eight rows' loads in flight
            int cv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) cv[u] = base[(int64_t)(k + u) * sd.stride + pos];
            const int* ri = rinfo + 3 * k;
#pragma unroll
            for (int u = 0; u < 8; ++u)
                one(base + (int64_t)(k + u) * sd.stride, cv[u], ri[3 * u], ri[3 * u + 1], ri[3 * u + 2]);
        }
        for (; k < k1; ++k) {
            const uint8_t* r = base + (int64_t)k * sd.stride;
            one(r, r[pos], rinfo[3 * k], rinfo[3 * k + 1], rinfo[3 * k + 2]);
        }
    }
    s_gh[(sl * 2) * C + t] = gapc;
    s_gh[(sl * 2 + 1) * C + t] = hyd;
    __syncthreads();
    // slice 0's counts become the column's: each slice adds up 8 of the 32 codes
    for (int c = sl * 8; c < sl * 8 + 8; ++c) {
        int w = 0, m = 0;
        for (int q = 0; q < NS; ++q) {
            w += s_ws[(q * 32 + c) * C + t];
            m += s_cn[(q * 32 + c) * C + t];
        }
        s_ws[c * C + t] = w;
        s_cn[c * C + t] = m;
    }
    gapc = hyd = 0;
    for (int q = 0; q < NS; ++q) {
        gapc += s_gh[(q * 2) * C + t];
        hyd += s_gh[(q * 2 + 1) * C + t];
    }
    __syncthreads();
    if (!live) return;                       // no barrier below
    int32_t* row = a.pool + sd.prf + (int64_t)(pos + 1) * CWP_PRF_W;
    if (sl == 0) {
        int32_t* tmp = a.pool + sd.tmp;
        int hw = 0, rw = 0, vw = 0;
        if (hydp) hw = (int)((float)hyd / (float)n);
        if (sd.flags & CWP_F_PREF) {
            int w = 0;
            for (int c = 0; c <= 24; ++c) w += s_cn[c * C + t] * a.pen[c];
            rw = w / n;
        }
        if (sd.flags & CWP_F_VAR) {          // two rows
            const uint8_t* ra = base;
            const uint8_t* rb = base + sd.stride;
            int tt = 0, w = 0;
            for (int j = pos - 5; j < pos + 5; ++j) {
                if (j < 0 || j >= sd.L) continue;
                const int x = ra[j], y = rb[j];
                if (pgap(x) || pgap(y)) continue;
                w += (x == y) ? 1 : 0;
                ++tt;
            }
            w += tt;
            w = tt > 0 ? (w * 100) / (2 * tt) : 100;
            vw = w < 50 ? 50 : w;
        }
        tmp[pos] = gapc;
        tmp[Lp + pos] = hw;
        tmp[2 * Lp + pos] = rw;
        tmp[3 * Lp + pos] = vw;
        for (int c = 25; c < 30; ++c) row[c] = 0;
        for (int c = 32; c < CWP_PRF_W; ++c) row[c] = 0;
    }
    // the 27 residue columns (codes 0..24, 30, 31), 7 per slice
    const int q0 = sl * 7, q1 = q0 + 7 < 27 ? q0 + 7 : 27;
    if (sd.flags & CWP_F_SUB) {
        const int32_t* mat = a.pool + sd.mat;
        if (gapc == n) {                     // an all-gap column
            for (int q = q0; q < q1; ++q) {
                const int res = q < 25 ? q : q + 5;
                row[res] = mat[res * 32 + 30];        // mat[res][GAP1]
            }
        } else {
            int wt[27];
            for (int u = 0; u < 27; ++u) wt[u] = s_ws[(u < 25 ? u : u + 5) * C + t];
            const float scale = (float)(n - gapc) / (float)n;
            for (int q = q0; q < q1; ++q) {
                const int res = q < 25 ? q : q + 5;
                int f = 0;
                for (int u = 0; u < 27; ++u) f += wt[u] * mat[(u < 25 ? u : u + 5) * 32 + res];
                row[res] = f2i_host(((float)f / (float)sd.sum2) * scale);
            }
        }
    } else {
        for (int q = q0; q < q1; ++q) {
            const int res = q < 25 ? q : q + 5;
            row[res] = sd.sum2 == 0 ? 0 : (int)(10 * (float)s_ws[res * C + t] / (float)sd.sum2);
        }
    }
}

// One workgroup per profile: the gap-distance windows of calcGapCoeff
// (:180-230), the per-column gap opening and extension of localPenalty
// (:468-512) and :232-297, the end rows (:283-297), then the DP's per-column
// records and the LENCOL prefix sums.
//
// The reference walks the gap counts left to right: a column with gaps keeps
// its count; a run of such columns marks the gdist-1 gap-free columns before it
// (-1 next to the run, -2, ...) and, up to the next run, the gdist gap-free
// columns after it (-1, -2, ..., -gdist); a later run's marks overwrite an
// earlier run's.  So a gap-free column p takes p - s for the last run start s
// in (p, p + gdist - 1], else -(p - e) - 1 for the end e of the run to its left
// when p - e < gdist, else 0.
extern "C" __global__ void __launch_bounds__(256) cw_prf_gapcols(BuildArgs a)
{
    __shared__ int s_scan[256];
    __shared__ int s_carry;
    const PrfSide sd = a.sides[blockIdx.x];
    const int L = sd.L, Lp = L + 2, n = sd.n, gd = sd.gdist, tid = threadIdx.x;
    const bool dna = (sd.flags & CWP_F_DNA) != 0, egp = (sd.flags & CWP_F_EGP) != 0;
    const int32_t* gaps = a.pool + sd.tmp;
    const int32_t* hydW = gaps + Lp;
    const int32_t* resW = gaps + 2 * Lp;
    const int32_t* vW = gaps + 3 * Lp;
    int32_t* prf = a.pool + sd.prf;
    for (int pos = tid; pos < L; pos += blockDim.x) {
        const int g0 = gaps[pos];
        int gp = g0;
        if (!dna && gd > 0 && g0 <= 0) {
            gp = 0;
            int sstar = -1;
            for (int s = pos + 1; s <= pos + gd - 1 && s < L; ++s)
                if (gaps[s] > 0 && gaps[s - 1] <= 0) sstar = s;
            if (sstar >= 0) {
                gp = pos - sstar;
            } else {
                for (int d = 0; d < gd && pos - d >= 1; ++d) {
                    if (gaps[pos - d] > 0) break;
                    if (gaps[pos - d - 1] > 0) {
                        gp = -d - 1;
                        break;
                    }
                }
            }
        }
        int g, l;
        if (gp <= 0) {
            if (!dna) {
                bool h = false;
                float gw = 1.0f;
                if (sd.flags & CWP_F_VAR) gw *= vW[pos] / 100.0;
                if ((sd.flags & CWP_F_HYD) && hydW[pos] > 0) {
                    gw *= 0.5;
                    h = true;
                }
                if ((sd.flags & CWP_F_PREF) && !h) gw *= (resW[pos] / 100.0);
                gw *= sd.gapcoef;
                g = (int)gw;
            } else {
                g = sd.gapcoef;
            }
            l = sd.lencoef;
            if (gp < 0) g = (int)((g * (2.0 + 2.0 * (gd + gp) / gd)));
        } else {
            const float scale = ((float)(n - g0) / (float)n) * 1.0f;
            g = (int)(scale * sd.gapcoef);
            l = (int)(0.5 * sd.lencoef);
        }
        if (g <= 0) g = 1;
        if (l <= 0) l = 1;
        if (!egp && pos + 1 == L) g = l = 0;
        prf[(int64_t)(pos + 1) * CWP_PRF_W + CWP_GAPCOL] = g;
        prf[(int64_t)(pos + 1) * CWP_PRF_W + CWP_LENCOL] = l;
    }
    if (tid < CWP_PRF_W) {
        int v0 = 0;
        if (egp && tid == CWP_GAPCOL) v0 = sd.gapcoef;
        if (egp && tid == CWP_LENCOL) v0 = sd.lencoef;
        prf[tid] = v0;                                   // row 0
        prf[(int64_t)(L + 1) * CWP_PRF_W + tid] = 0;     // row L + 1
    }
    if (tid == 0) s_carry = 0;
    __syncthreads();
    int32_t* rec = a.pool + sd.rec;
    int32_t* pl = a.pool + sd.pl;
    for (int b0 = 0; b0 < Lp; b0 += 256) {
        const int x = b0 + tid;
        int l = 0;
        if (x < Lp) {
            const int g = prf[(int64_t)x * CWP_PRF_W + CWP_GAPCOL];
            l = prf[(int64_t)x * CWP_PRF_W + CWP_LENCOL];
            const bool end = !egp && (x == 0 || x == L);
            rec[4 * x + 0] = g;
            rec[4 * x + 1] = l;
            rec[4 * x + 2] = end ? 0 : -1;
            rec[4 * x + 3] = g + l;
        }
        s_scan[tid] = l;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int v = tid >= off ? s_scan[tid - off] : 0;
            __syncthreads();
            s_scan[tid] += v;
            __syncthreads();
        }
        if (x < Lp) pl[x] = s_carry + s_scan[tid] - l;
        __syncthreads();
        if (tid == 255) s_carry += s_scan[255];
        __syncthreads();
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

hipError_t cwp_launch_wide(int ntask, hipStream_t st, const PrfArgs& a, const WideArgs& w)
{
    hipLaunchKernelGGL(cw_prf_wide, dim3(ntask), dim3(WAVE), 0, st, a, w);
    return hipGetLastError();
}

hipError_t cwp_launch_level(int nprob, int T, hipStream_t st, const PrfArgs& a)
{
    hipLaunchKernelGGL(cw_prf_level, dim3(nprob * T), dim3(WAVE), 0, st, a, T);
    return hipGetLastError();
}

hipError_t cwp_launch_local(int groups, hipStream_t st, const PrfArgs& a)
{
    hipLaunchKernelGGL(cw_prf_local, dim3(groups), dim3(64 * CWP_LOCAL_WAVES), 0, st, a);
    return hipGetLastError();
}

hipError_t cwp_launch_build(int ncolblk, int nside, hipStream_t st, const BuildArgs& a)
{
    if (ncolblk > 0)
        hipLaunchKernelGGL(cw_prf_cols, dim3(ncolblk), dim3(CWP_BUILD_COLS * CWP_BUILD_SLICES), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (nside > 0) hipLaunchKernelGGL(cw_prf_gapcols, dim3(nside), dim3(256), 0, st, a);
    return hipGetLastError();
}
