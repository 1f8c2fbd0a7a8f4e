// cw_nj_kernels.hip -- neighbour-joining on the device, over the DistMatrix of
// a pairwise run that never leaves device memory.
//
//   cw_nj_matrix  the (n+1)^2 double DistMatrix from the pairs' identity counts:
//                 FullPairwiseAlign's float distance (FullPairwiseAlign.cpp:
//                 172-181), stored by both SetAt calls (:184-185).
//   cw_nj         NJTree::generateTree (tree/NJTree.cpp:161-751) by one
//                 workgroup: per join, the live nodes' row and column sums, the
//                 pair with the smallest NJ total, the branch lengths, the
//                 merged node's distances and the treeDesc row of the join.
//
// The tree's bytes depend on every double rounding, so each value is computed
// by the reference's expression in the reference's order: the sums that run
// over the live nodes (sumd, the initial row and column sums, the merged node's
// sums) are sequential loops in list order on one lane; the per-node updates
// and the pair scan are independent per element and run in parallel, and the
// scan keeps the reference's winner (the smallest total, the first in its scan
// order: column jj ascending, then row ii ascending) by reducing on (total,
// scan position).  Built with FP contraction off, like the host code.
#include <hip/hip_runtime.h>
#include <climits>
#include <cmath>

#include "cw_nj_kernels.h"

namespace cwn {

constexpr int kThreads = 1024;
constexpr int kWaves = kThreads / 64;
constexpr int kPerThread = (CW_NJ_MAX_SEQS + kThreads - 1) / kThreads;   // list entries per thread
static_assert((CW_NJ_MAX_SEQS + 1) * (6 * sizeof(double) + sizeof(int)) + 1024 <= 160 * 1024,
              "per-node arrays exceed the LDS");

// FullPairwiseAlign.cpp:172-181 (a refused pair's distance is 1.0)
__device__ __forceinline__ float pair_distance(int count, int len1, int len2)
{
    if (count < 0) return 1.0f;
    float mm = (float)(100.0 * (float)count);
    if (len1 == 0 || len2 == 0) mm = 0;
    else mm /= (float)(len1 < len2 ? len1 : len2);
    return ((float)100.0 - mm) / (float)100.0;
}

extern "C" __global__ void __launch_bounds__(256) cw_nj_matrix(NjArgs a)
{
    const int ld = a.n + 1;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < a.npairs;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int si = a.pairs[2 * k], sj = a.pairs[2 * k + 1];
        if ((unsigned)si >= (unsigned)a.n || (unsigned)sj >= (unsigned)a.n) continue;
        const double d = (double)pair_distance(a.count[k], a.gapfree[si], a.gapfree[sj]);
        a.D[(int64_t)(si + 1) * ld + (sj + 1)] = d;
        a.D[(int64_t)(sj + 1) * ld + (si + 1)] = d;
    }
}

// scan position k of the pair (r, q), r < q: k = q (q - 1) / 2 + r
__device__ __forceinline__ void pair_of(long long k, int& q, int& r)
{
    int qq = (int)((1.0 + sqrt(1.0 + 8.0 * (double)k)) * 0.5);
    while ((long long)qq * (qq - 1) / 2 > k) --qq;
    while ((long long)(qq + 1) * qq / 2 <= k) ++qq;
    q = qq;
    r = (int)(k - (long long)qq * (qq - 1) / 2);
}

// s + x[a] + x[a + 1] + ... + x[b - 1], added one by one in that order (the
// reference's rounding); the loads run eight values ahead of the adds
__device__ double seq_sum(const double* x, int a, int b, double s)
{
    int q = a;
    if (q + 8 <= b) {
        double cur[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) cur[v] = x[q + v];
        for (; q + 16 <= b; q += 8) {
            double nxt[8];
#pragma unroll
            for (int v = 0; v < 8; ++v) nxt[v] = x[q + 8 + v];
#pragma unroll
            for (int v = 0; v < 8; ++v) s = s + cur[v];
#pragma unroll
            for (int v = 0; v < 8; ++v) cur[v] = nxt[v];
        }
#pragma unroll
        for (int v = 0; v < 8; ++v) s = s + cur[v];
        q += 8;
    }
    for (; q < b; ++q) s = s + x[q];
    return s;
}

// the largest i in 1 .. upto with desc(i, node) == 1 (0 if none), workgroup-wide
__device__ int last_row_with(const int32_t* desc, int ld, int node, int upto, int* s_best)
{
    if (threadIdx.x == 0) *s_best = 0;
    __syncthreads();
    int best = 0;
    for (int i = 1 + threadIdx.x; i <= upto; i += kThreads)
        if (desc[(int64_t)i * ld + node] == 1) best = i;
    if (best) atomicMax(s_best, best);
    __syncthreads();
    const int b = *s_best;
    __syncthreads();
    return b;
}

extern "C" __global__ void __launch_bounds__(kThreads) cw_nj(NjArgs a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = a.n, ld = n + 1;
    // by node
    double* sumCols = (double*)smem;
    double* sumRows = sumCols + (n + 1);
    double* av = sumRows + (n + 1);
    // by position in the live list (lv[q] is the node at position q)
    double* dq = av + (n + 1);                       // sumCols + sumRows of lv[q]
    double* cq = dq + (n + 1);                       // sumCols of lv[q]: the terms of sumd
    double* jq = cq + (n + 1);                       // the merged node's distance to lv[q]
    int* lv = (int*)(jq + (n + 1));                  // live nodes in list (ascending) order
    __shared__ double red_v[kWaves];
    __shared__ long long red_k[kWaves];
    __shared__ double s_sumd, s_part[3];
    __shared__ int s_mini, s_minj, s_qpos, s_rpos, s_type[2], s_best;
    double* D = a.D;
    int32_t* desc = a.desc;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    for (int i = tid; i <= n; i += kThreads) {
        sumCols[i] = sumRows[i] = av[i] = 0.0;
        if (i < n) lv[i] = i + 1;
    }
    for (int i = 1 + tid; i <= n; i += kThreads) D[(int64_t)i * ld + i] = 0.0;
    __syncthreads();
    // :244-274 (the matrix is symmetric, so D(p, j) is read as D(j, p))
    for (int j = 1 + tid; j <= n; j += kThreads) {
        const double* row = D + (int64_t)j * ld;
        double tmp = 0.0;
        for (int p = 1; p < j; ++p) tmp = tmp + row[p];
        sumRows[j] = tmp;
        tmp = 0.0;
        for (int p = j + 1; p <= n; ++p) tmp += row[p];
        sumCols[j] = tmp;
        cq[j - 1] = sumCols[j];
        dq[j - 1] = sumCols[j] + sumRows[j];
    }
    __syncthreads();

    double fnseqs = (double)n;
    int m = n;
    // sumd: the live nodes' column sums added in list order (:282-287); for
    // the later joins it is formed at the end of the join before
    if (tid == 0) s_sumd = seq_sum(cq, 0, m, 0.0);
    __syncthreads();
    for (int nc = 1; nc <= n - 3; ++nc) {
        const double twosumd = 2.0 * s_sumd;          // exact: a power-of-two scale
        const double fnseqs2 = fnseqs - 2.0;
        // The scan (:289-330): the total of pair (lv[r], lv[q]), r < q, at scan
        // position k = q (q - 1) / 2 + r.  The work is cut into runs of rows q
        // for one block of 64 positions r (lane l takes r = 64 rb + l): block rb
        // has rows 64 rb + 1 .. m - 1, and the blocks' rows, laid end to end,
        // are split evenly over the waves.  A lane meets its pairs of one run
        // in ascending k, so a strict "<" keeps the first of equal totals; runs
        // are merged on (total, k).  Row q's node and sum are the same for
        // every lane (LDS broadcasts), the lanes read consecutive entries of
        // row lv[q], and 16 rows' loads are issued before their totals are
        // compared, in q order.
        const double bound = 99999.0 * 2.0 * fnseqs2;
        double tm = bound;
        long long bk = LLONG_MAX;
        {
            const int nrb = (m - 1 + 63) / 64;           // blocks with a row: 64 rb + 1 < m
            const int total_rows = nrb * (m - 1) - 32 * nrb * (nrb - 1);
            const int s0 = (int)((long long)total_rows * wave / kWaves);
            const int s1 = (int)((long long)total_rows * (wave + 1) / kWaves);
            int rb = 0, base = 0;                        // rows of the blocks before rb
            while (rb < nrb && base + (m - 1 - 64 * rb) <= s0) {
                base += m - 1 - 64 * rb;
                ++rb;
            }
            int pos = s0;
            while (pos < s1) {
                const int len = m - 1 - 64 * rb;
                const int q0 = 64 * rb + 1 + (pos - base);
                const int q1 = 64 * rb + 1 + (s1 - base < len ? s1 - base : len);
                const int r = 64 * rb + lane;
                const int ii = r < m ? lv[r] : 0;        // column 0 is in bounds
                const double dr = r < m ? dq[r] : 0.0;
                const double* col = D + ii;
                double tu = bound;
                int bq = -1;
                // 16 rows: their totals, the smallest of them, and only when
                // that beats the lane's best, the first row that has it.  Rows
                // not above the lane's position (the first 64 rows of a block)
                // take no part: +inf.
                auto rows16 = [&](int q, bool diag) {
                    int jv[16];
                    double t[16];
#pragma unroll
                    for (int v = 0; v < 16; ++v) jv[v] = lv[q + v];
#pragma unroll
                    for (int v = 0; v < 16; ++v) t[v] = col[(int64_t)jv[v] * ld];
#pragma unroll
                    for (int v = 0; v < 16; ++v) {
                        t[v] = fnseqs2 * t[v] + twosumd - dr - dq[q + v];
                        if (diag && !(r < q + v)) t[v] = HUGE_VAL;
                    }
                    double mn = t[0];
#pragma unroll
                    for (int v = 1; v < 16; ++v) mn = fmin(mn, t[v]);
                    if (mn < tu) {
                        int f = 15;
#pragma unroll
                        for (int v = 14; v >= 0; --v)
                            if (t[v] == mn) f = v;
                        tu = mn;
                        bq = q + f;
                    }
                };
                int q = q0;
                const int qd = 64 * rb + 64;                 // rows from here on: every lane
                for (; q + 16 <= q1 && q < qd; q += 16) rows16(q, true);
                for (; q + 16 <= q1; q += 16) rows16(q, false);
                for (; q < q1; ++q) {
                    const double total = fnseqs2 * col[(int64_t)lv[q] * ld] + twosumd - dr - dq[q];
                    if (r < q && total < tu) {
                        tu = total;
                        bq = q;
                    }
                }
                if (bq >= 0) {
                    const long long k = (long long)bq * (bq - 1) / 2 + r;
                    if (tu < tm || (tu == tm && k < bk)) {
                        tm = tu;
                        bk = k;
                    }
                }
                pos = base + (s1 - base < len ? s1 - base : len);
                base += len;
                ++rb;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_xor(tm, o, 64);
            const long long ok = __shfl_xor(bk, o, 64);
            if (ov < tm || (ov == tm && ok < bk)) {
                tm = ov;
                bk = ok;
            }
        }
        if (lane == 0) {
            red_v[wave] = tm;
            red_k[wave] = bk;
        }
        __syncthreads();
        if (tid == 0) {
            double v = red_v[0];
            long long kk = red_k[0];
            for (int w = 1; w < kWaves; ++w)
                if (red_v[w] < v || (red_v[w] == v && red_k[w] < kk)) {
                    v = red_v[w];
                    kk = red_k[w];
                }
            int mini = 0, minj = 0, qpos = -1;
            if (kk == LLONG_MAX) {
                *a.status = 1;                       // no pair below the bound: left to the host
            } else {
                int q, r;
                pair_of(kk, q, r);
                mini = lv[r];
                minj = lv[q];
                qpos = q;
                double dio = sumCols[mini] + sumRows[mini];
                double djo = sumCols[minj] + sumRows[minj];
                double dmin = D[(int64_t)mini * ld + minj];
                dio = (dio - dmin) / fnseqs2;
                djo = (djo - dmin) / fnseqs2;
                double bi = (dmin + dio - djo) * 0.5;
                double bj = dmin - bi;
                bi = bi - av[mini];
                bj = bj - av[minj];
                s_type[0] = av[mini] > 0.0 ? 0 : 1;
                s_type[1] = av[minj] > 0.0 ? 0 : 1;
                if (fabs(bi) < 0.0001) bi = 0.0;
                if (fabs(bj) < 0.0001) bj = 0.0;
                a.left[nc] = bi;
                a.right[nc] = bj;
                if (dmin <= 0.0) dmin = 0.000001;
                av[mini] = dmin * 0.5;
            }
            s_mini = mini;
            s_minj = minj;
            s_qpos = qpos;
            s_rpos = qpos < 0 ? -1 : (int)(kk - (long long)qpos * (qpos - 1) / 2);
        }
        __syncthreads();
        const int mini = s_mini, minj = s_minj, qpos = s_qpos, qm = s_rpos;
        if (qpos < 0) return;                       // status set; every thread leaves here
        // treeDesc row nc: the joined nodes' leaves (:501-535)
        for (int i = 1 + tid; i <= n; i += kThreads) desc[(int64_t)nc * ld + i] = 0;
        __syncthreads();
        for (int h = 0; h < 2; ++h) {
            const int node = h ? minj : mini;
            if (s_type[h] == 0) {
                const int src = last_row_with(desc, ld, node, nc - 1, &s_best);
                if (src)
                    for (int j = 1 + tid; j <= n; j += kThreads)
                        if (desc[(int64_t)src * ld + j] == 1) desc[(int64_t)nc * ld + j] = 1;
            } else if (tid == 0) {
                desc[(int64_t)nc * ld + node] = 1;
            }
        }
        fnseqs = fnseqs - 1.0;
        // minj leaves the list: every entry after it moves down by one
        int mv[kPerThread];
#pragma unroll
        for (int u = 0; u < kPerThread; ++u) {
            const int q = tid + u * kThreads;
            mv[u] = (q >= qpos && q < m - 1) ? lv[q + 1] : 0;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kPerThread; ++u) {
            const int q = tid + u * kThreads;
            if (q >= qpos && q < m - 1) lv[q] = mv[u];
        }
        --m;
        __syncthreads();
        // :558-605, per live node (each touches only its own entries); the
        // merged node's own sums are redone below
        for (int q = tid; q < m; q += kThreads) {
            const int j = lv[q];
            double tdi = 0.0, tdj = 0.0;
            double sc = sumCols[j], sr = sumRows[j];
            if (j < minj) {
                tdj = D[(int64_t)j * ld + minj];
                sc -= tdj;
            } else if (j > minj) {
                tdj = D[(int64_t)minj * ld + j];
                sr -= tdj;
            }
            if (j < mini) {
                tdi = D[(int64_t)j * ld + mini];
                sc -= tdi;
            } else if (j > mini) {
                tdi = D[(int64_t)mini * ld + j];
                sr -= tdi;
            }
            const double jn = (tdj + tdi) * 0.5;
            jq[q] = jn;
            if (j < mini) {
                D[(int64_t)j * ld + mini] = jn;
                D[(int64_t)mini * ld + j] = jn;
                sc += jn;
            } else if (j > mini) {
                D[(int64_t)mini * ld + j] = jn;
                D[(int64_t)j * ld + mini] = jn;
                sr += jn;
            }
            sumCols[j] = sc;
            sumRows[j] = sr;
            cq[q] = sc;
            dq[q] = sc + sr;
        }
        __syncthreads();
        // The merged node's sums over the live nodes before and after it in
        // the list (mini keeps its position qm: minj was after it), and the
        // next join's sumd, each a sequential sum in list order; the three
        // independent chains run on three waves.  minj's row and column
        // (zeroed by the reference) are never read again.
        if (lane == 0 && wave < 3) {
            if (wave == 0) s_part[0] = seq_sum(jq, 0, qm, 0.0);          // sumRows[mini]
            else if (wave == 1) s_part[1] = seq_sum(jq, qm + 1, m, 0.0); // sumCols[mini]
            else s_part[2] = seq_sum(cq, 0, qm, 0.0);                    // sumd up to mini
        }
        __syncthreads();
        if (tid == 0) {
            sumRows[mini] = s_part[0];
            sumCols[mini] = s_part[1];
            sumCols[minj] = sumRows[minj] = 0.0;
            cq[qm] = sumCols[mini];
            dq[qm] = sumCols[mini] + sumRows[mini];
            s_sumd = seq_sum(cq, qm, m, s_part[2]);
        }
        __syncthreads();
    }
    // the last three nodes: a trichotomy (:615-741)
    __shared__ int s_l[4];
    if (tid == 0) {
        for (int i = 0; i < 3; ++i) s_l[i + 1] = lv[i];
        const int l1 = s_l[1], l2 = s_l[2], l3 = s_l[3];
        const double b1 = (D[(int64_t)l1 * ld + l2] + D[(int64_t)l1 * ld + l3] - D[(int64_t)l2 * ld + l3]) * 0.5;
        const double b2 = D[(int64_t)l1 * ld + l2] - b1;
        const double b3 = D[(int64_t)l1 * ld + l3] - b1;
        double br[4] = {0, b1 - av[l1], b2 - av[l2], b3 - av[l3]};
        for (int kk = 1; kk <= 3; ++kk)
            if (fabs(br[kk]) < 0.0001) br[kk] = 0.0;
        a.left[n - 2] = br[1];
        a.left[n - 1] = br[2];
        a.left[n] = br[3];
    }
    for (int i = 1 + tid; i <= n; i += kThreads) desc[(int64_t)(n - 2) * ld + i] = 0;
    __syncthreads();
    for (int i = 1; i <= 3; ++i) {
        const int node = s_l[i];
        if (av[node] > 0.0) {
            const int src = last_row_with(desc, ld, node, n - 3, &s_best);
            if (src)
                for (int j = 1 + tid; j <= n; j += kThreads)
                    if (desc[(int64_t)src * ld + j] == 1) desc[(int64_t)(n - 2) * ld + j] = i;
        } else if (tid == 0) {
            desc[(int64_t)(n - 2) * ld + node] = i;
        }
        __syncthreads();
    }
}

}  // namespace cwn

size_t cw_nj_lds_bytes(int n)
{
    return (size_t)(n + 1) * (6 * sizeof(double) + sizeof(int));
}

hipError_t cw_launch_nj(const NjArgs& a, hipStream_t st)
{
    const int ld = a.n + 1;
    hipError_t e = hipMemsetAsync(a.status, 0, sizeof(int32_t), st);
    if (e != hipSuccess) return e;
    if (a.pairs) {                                  // else the caller filled D
        e = hipMemsetAsync(a.D, 0, (size_t)ld * ld * sizeof(double), st);
        if (e != hipSuccess) return e;
        const int64_t nb = (a.npairs + 255) / 256;
        const int blocks = (int)(nb < 4096 ? nb : 4096);
        if (blocks > 0) hipLaunchKernelGGL(cwn::cw_nj_matrix, dim3(blocks), dim3(256), 0, st, a);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    const size_t lds = cw_nj_lds_bytes(a.n);
    e = hipFuncSetAttribute((const void*)cwn::cw_nj, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(cwn::cw_nj, dim3(1), dim3(cwn::kThreads), lds, st, a);
    return hipGetLastError();
}
