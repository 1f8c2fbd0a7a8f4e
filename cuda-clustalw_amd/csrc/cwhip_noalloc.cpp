// cwhip_noalloc.cpp -- the batched pairwise stage for callers that cannot
// allocate device memory (cwhip_pairwise_noalloc in include/cwhip.h).
//
// The kernels are the batched path's (cw_launch_batched: queue reset, length
// binning, the persistent kernel); only their work space differs.  It is the
// file-scope device array `scratch` below.  Host code never passes the array's
// name to a kernel or a copy: on the host that name is a stand-in, not a device
// address.  The device address comes from hipGetSymbolAddress, once per device
// with that device current, and is kept per device.
//
// Layout of the array (int32 units):
//   [0]                    queue counters of the persistent kernel
//   [64, 64 + BINS_INTS)   bin counts and cursors, the long part's size and the total
//   [kOrderAt, + npairs)   the binned work order
//   [rows_at, ...)         per wave 8 DP boundary rows of rowcap ints
// The queue and the bins are cleared by the first kernel of every call; the
// order and the rows are written before they are read.
#include <hip/hip_runtime.h>

#include <mutex>

#include "cwhip_internal.h"

hipError_t cw_launch_batched(int blocks, int threads, size_t lds, hipStream_t st, const DevArgs& a,
                             const CwSplitShape* s);
hipError_t cw_occupancy(int threads, size_t lds, int* plain_blocks, int* pers_blocks);

namespace {

constexpr int64_t kScratchInts = CWHIP_NOALLOC_SCRATCH_BYTES / 4;
constexpr int64_t kBinsAt = 64;
constexpr int64_t kOrderAt = (kBinsAt + CW_BINS_INTS + 63) & ~int64_t(63);
constexpr int kWavesPerGroup = CW_MAX_THREADS / 64;
constexpr int kMaxDevices = 64;

static_assert(kOrderAt + CWHIP_NOALLOC_MAX_PAIRS < kScratchInts, "pair order must fit the array");

static __device__ __align__(256) int32_t scratch[kScratchInts];

struct DeviceScratch {
    bool ready = false;
    int32_t* ptr = nullptr;    // device address of `scratch` on this device
    int num_cu = 0;
};

std::mutex g_mu;
DeviceScratch g_dev[kMaxDevices];

// Once per device, with that device current: the array's device address and
// the CU count.  Nothing is allocated (the array is part of the code object).
int device_scratch(int dev, DeviceScratch* out)
{
    if (dev < 0 || dev >= kMaxDevices) return CWHIP_E_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceScratch& d = g_dev[dev];
    if (!d.ready) {
        int32_t* ptr = nullptr;
        if (hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return CWHIP_E_HIP;
        int ncu = 0;
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
            return CWHIP_E_HIP;
        d.ptr = ptr;
        d.num_cu = ncu;
        d.ready = true;
    }
    *out = d;
    return CWHIP_OK;
}

} // namespace

extern "C" int cwhip_pairwise_noalloc(const cwhip_pair_dev_params* p, const uint8_t* codes,
                                      const int64_t* offsets, const int32_t* lens, int nseq,
                                      const int32_t* pairs, int64_t npairs, int32_t* out_count,
                                      int32_t* out_status, void* stream)
{
    if (!p || !p->matrix || !p->gap_open_by_len || !out_status || nseq < 0 || npairs < 0 ||
        p->max_len < 0)
        return CWHIP_E_ARG;
    if (npairs > 0 && (!codes || !offsets || !lens || !pairs || !out_count)) return CWHIP_E_ARG;
    if (p->max_len > CWHIP_MAX_LEN) return CWHIP_E_TOOLONG;
    if (npairs > CWHIP_NOALLOC_MAX_PAIRS) return CWHIP_E_NOMEM;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return CWHIP_E_HIP;
    DeviceScratch ds;
    const int rc = device_scratch(dev, &ds);
    if (rc) return rc;

    DevArgs a{};
    a.codes = codes;
    a.offs = offsets;
    a.lens = lens;
    a.pairs = pairs;
    a.order = nullptr;
    a.pair_base = 0;
    a.pair_end = npairs;
    a.go_by_len = p->gap_open_by_len;
    a.matrix = p->matrix;
    a.ge = p->gap_extend;
    a.gap1 = p->gap_pos1;
    a.gap2 = p->gap_pos2;
    a.rmax = CW_RMAX;
    a.maxlen = p->max_len;
    a.nseq = nseq;
    a.lsplit = 0x7fffffff;                  // one launch over the whole queue
    a.qpart = 0;
    a.lcap = cw_lcap(p->max_len);
    a.wave_lds = cw_wave_lds_bytes(a.lcap);
    a.rowcap = ((p->max_len + 2) + 63) & ~63;
    a.scratch_stride = 8 * (int64_t)a.rowcap;
    a.out_count = out_count;
    a.out_info = nullptr;
    a.status = out_status;
    a.queue = (unsigned int*)ds.ptr;
    a.bins = ds.ptr + kBinsAt;
    a.order_out = ds.ptr + kOrderAt;
    for (int c = 0; c < CW_NCLASS; ++c) {
        a.rmax_cls[c] = CW_RMAX;
        a.colcap_cls[c] = CW_SMALL_COLS;
    }
    const int64_t rows_at = (kOrderAt + npairs + 63) & ~int64_t(63);
    a.scratch = ds.ptr + rows_at;

    // resident workgroups: as many as fit on the chip, never more waves than
    // pairs, and no more than the DP rows the rest of the array holds
    const int wpg = cw_fit_waves(kWavesPerGroup, a.wave_lds);
    if (wpg < 1) return CWHIP_E_TOOLONG;
    const int threads = 64 * wpg;
    const size_t lds = CW_TAB_BYTES + (size_t)wpg * a.wave_lds;
    int occ_plain = 0, occ_pers = 0;
    if (cw_occupancy(threads, lds, &occ_plain, &occ_pers) != hipSuccess) return CWHIP_E_HIP;
    if (occ_pers < 1) return CWHIP_E_TOOLONG;
    const int64_t fit = (kScratchInts - rows_at) / (a.scratch_stride * wpg);
    if (fit < 1) return CWHIP_E_NOMEM;
    int64_t blocks = (int64_t)ds.num_cu * occ_pers;
    blocks = std::min(blocks, (npairs + wpg - 1) / wpg);
    blocks = std::min(blocks, fit);
    if (blocks < 1) blocks = 1;   // npairs == 0: the reset still clears out_status
    return cw_launch_batched((int)blocks, threads, lds, (hipStream_t)stream, a, nullptr) == hipSuccess
               ? CWHIP_OK
               : CWHIP_E_HIP;
}
