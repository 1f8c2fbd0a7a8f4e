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

#include "cw_nj_kernels.h"

namespace cwn {

constexpr int kThreads = 1024;
constexpr int kWaves = kThreads / 64;
constexpr int kPerThread = (CW_NJ_MAX_SEQS + kThreads - 1) / kThreads;   // list entries per thread

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
    double* sumCols = (double*)smem;
    double* sumRows = sumCols + (n + 1);
    double* av = sumRows + (n + 1);
    double* join = av + (n + 1);
    double* dq = join + (n + 1);
    int* lv = (int*)(dq + (n + 1));                  // live nodes in list (ascending) order
    __shared__ double red_v[kWaves];
    __shared__ long long red_k[kWaves];
    __shared__ double s_sumd;
    __shared__ int s_mini, s_minj, s_qpos, s_type[2], s_best;
    double* D = a.D;
    int32_t* desc = a.desc;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    for (int i = tid; i <= n; i += kThreads) {
        sumCols[i] = sumRows[i] = av[i] = join[i] = 0.0;
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
    }
    __syncthreads();

    double fnseqs = (double)n;
    int m = n;
    for (int nc = 1; nc <= n - 3; ++nc) {
        for (int q = tid; q < m; q += kThreads) {
            const int j = lv[q];
            dq[q] = sumCols[j] + sumRows[j];
        }
        if (tid == 0) {
            double s = 0.0;
            for (int q = 0; q < m; ++q) s += sumCols[lv[q]];
            s_sumd = s;
        }
        __syncthreads();
        const double sumd = s_sumd;
        const double fnseqs2 = fnseqs - 2.0;
        // the scan: total of pair (lv[r], lv[q]), r < q, at position k
        const long long npr = (long long)m * (m - 1) / 2;
        double tm = 99999.0 * 2.0 * fnseqs2;
        long long bk = LLONG_MAX;
        long long k = tid;
        if (k < npr) {
            int q, r;
            pair_of(k, q, r);
            for (; k < npr; k += kThreads) {
                const int jj = lv[q], ii = lv[r];
                const double total = fnseqs2 * D[(int64_t)jj * ld + ii] + 2.0 * sumd - dq[r] - dq[q];
                if (total < tm) {
                    tm = total;
                    bk = k;
                }
                r += kThreads;
                while (r >= q) {
                    r -= q;
                    ++q;
                }
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
        }
        __syncthreads();
        const int mini = s_mini, minj = s_minj, qpos = s_qpos;
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
        // :558-605, per live node (each touches only its own entries)
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
            join[j] = jn;
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
        }
        __syncthreads();
        if (tid == 0) {
            double da = 0.0;
            int q = 0;
            for (; q < m && lv[q] < mini; ++q) da = da + join[lv[q]];
            sumRows[mini] = da;
            if (q < m && lv[q] == mini) ++q;
            da = 0.0;
            for (; q < m; ++q) da = da + join[lv[q]];
            sumCols[mini] = da;
            sumCols[minj] = sumRows[minj] = 0.0;
        }
        __syncthreads();
        for (int j = 1 + tid; j <= n; j += kThreads) {
            D[(int64_t)minj * ld + j] = 0.0;
            D[(int64_t)j * ld + minj] = 0.0;
            join[j] = 0.0;
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
    return (size_t)(n + 1) * (5 * sizeof(double) + sizeof(int));
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
