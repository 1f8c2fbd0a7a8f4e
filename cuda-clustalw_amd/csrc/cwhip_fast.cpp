// cwhip_fast.cpp -- host side of the k-tuple distances (cwhip_fast_pairwise_range).
//
// FastPairwiseAlign::pairwiseAlign (pairwise/FastPairwiseAlign.cpp:20-145):
// the pairs of the range in the reference's order, one wave of cw_fast_wave
// per pair with its arrays in LDS (one lane of cw_fast_pairs per pair, arrays
// in HBM, where they do not fit), the distances back in that order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "cw_fast_kernels.h"
#include "cwhip_internal.h"

// cw_fast_wave with `per_wave` LDS bytes per pair: as many waves per CU as
// LDS holds (at most 16), grouped into workgroups of two to four waves so that
// the waves of a CU are spread over its four SIMDs (with one-wave workgroups,
// three waves per CU ran only 8% faster than two: they shared SIMDs).  The
// group size is the one that leaves the most waves resident (7 fit: two groups
// of 3, not one of 4); only when a single wave fits is it a group of one.
// Each wave takes pairs from the counter on its own.
static int waves_per_group(int per_cu)
{
    if (per_cu < 2) return 1;
    int best = 2;
    for (int w = 3; w <= 4; ++w)
        if (w * (per_cu / w) >= best * (per_cu / best)) best = w;
    return std::min(best, per_cu);
}

static hipError_t launch_waves(cwhip_ctx* ctx, FastArgs& a, size_t per_wave, int64_t np)
{
    const size_t lw = (per_wave + 15) / 16 * 16;
    const int per_cu = std::max<int>(1, std::min<int>(16, (int)((160 * 1024) / lw)));
    const int wpg = waves_per_group(per_cu);
    const int64_t groups = (int64_t)std::max(1, ctx->num_cu) * (per_cu / wpg);
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((np + wpg - 1) / wpg, groups));
    a.lds_wave = (int64_t)lw;
    return cwf_launch_wave((int)blocks, wpg, (size_t)wpg * lw, ctx->stream, a);
}

extern "C" {

int cwhip_fast_params_default(int dna, cwhip_fast_params* fp)
{
    if (!fp) return CWHIP_E_ARG;
    std::memset(fp, 0, sizeof *fp);
    fp->dna = dna ? 1 : 0;                     // UserParameters.cpp:86-94
    fp->ktup = dna ? 2 : 1;
    fp->window = dna ? 4 : 5;
    fp->wind_gap = dna ? 5 : 3;
    fp->signif = dna ? 4 : 5;
    fp->percent = 1;
    fp->max_aln = 0;
    return CWHIP_OK;
}

int cwhip_fast_pairwise_range(cwhip_ctx* ctx, const cwhip_fast_params* fp, int i_start, int i_end,
                              int j_start, int j_end, int32_t* out_pairs, double* out_dist,
                              int64_t* out_npairs)
{
    if (!ctx || !fp || !out_npairs) return fail(ctx, CWHIP_E_ARG, "bad k-tuple arguments");
    // the reference's fatal range checks (:27-37), and indices within the loaded sequences
    if (i_start < 0 || i_end < i_start || j_start < 0 || j_end < j_start || i_end > ctx->nseq ||
        j_end > ctx->nseq)
        return fail(ctx, CWHIP_E_ARG, "the range for pairwise alignment is incorrect");
    const int maxk = fp->dna ? 4 : 2;
    if (fp->ktup < 1 || fp->ktup > maxk || fp->window < 0 || fp->signif < 1)
        return fail(ctx, CWHIP_E_ARG, "bad k-tuple parameters");
    std::vector<int32_t> pairs;
    for (int i = i_start + 1; i <= i_end; ++i)                      // 1-based, :64-80
        for (int j = std::max(j_start + 2, i + 1); j <= j_end; ++j) {
            pairs.push_back(i - 1);
            pairs.push_back(j - 1);
        }
    const int64_t np = (int64_t)pairs.size() / 2;
    *out_npairs = np;
    if (out_pairs && np) std::memcpy(out_pairs, pairs.data(), pairs.size() * sizeof(int32_t));
    if (np == 0) return CWHIP_OK;
    if (!out_dist) return fail(ctx, CWHIP_E_ARG, "null distance buffer");
    HIPCHK(hipSetDevice(ctx->device));
    int lmax = 0;
    for (int s = 0; s < ctx->nseq; ++s) lmax = std::max(lmax, (int)ctx->h_lens[s]);
    const int max_aln = fp->max_aln > 0 ? fp->max_aln : 2 * lmax;  // Alignment::getMaxAlnLength
    int limit = 1;
    for (int q = 0; q < fp->ktup; ++q) limit *= fp->dna ? 4 : 25;
    FastArgs a;
    a.dna = fp->dna;
    a.ktup = fp->ktup;
    a.window = fp->window;
    a.wind_gap = fp->wind_gap;
    a.signif = fp->signif;
    a.percent = fp->percent ? 1 : 0;
    a.max_aln = max_aln;
    a.limit = limit;
    // per-lane rows; diagonals 1 .. l1 + l2 - 1, fragments 1 .. 2 max_aln - 1
    const int64_t dnum = std::max<int64_t>(2 * (int64_t)max_aln + 1, 2 * (int64_t)lmax + 2);
    int64_t rows = 0;
    a.o_zza = rows; rows += lmax + 2;
    a.o_zzb = rows; rows += lmax + 2;
    a.o_zzc = rows; rows += limit + 2;
    a.o_zzd = rows; rows += limit + 2;
    a.o_displ = rows; rows += dnum;
    a.o_slopes = rows; rows += dnum;
    a.o_diag = rows; rows += dnum;
    a.acc_len = 2 * (int64_t)max_aln + 2;
    a.o_acc = rows; rows += 4 * a.acc_len;
    // lanes: up to one per wave slot of the chip, within 8 GiB of scratch
    const int64_t cap_lanes = std::max<int64_t>(256, ((int64_t)8 << 30) / ((rows + 32) * 4) / 256 * 256);
    int64_t lanes = std::min<int64_t>((int64_t)std::max(1, ctx->num_cu) * 4 * 64, cap_lanes);
    lanes = std::min<int64_t>(lanes, (np + 255) / 256 * 256);
    a.lanes = lanes;
    a.lane_ints = (rows + 31) / 32 * 32;                            // 128-byte aligned lanes
    // one wave per pair with its arrays in LDS where they fit (16-bit positions
    // and fragment numbers), else one lane per pair with its arrays in HBM
    // (a fragment's score is at most ktup times its start in sequence 1 when the
    // window gap is not negative, which bounds the score buckets and the 16-bit
    // scores)
    const size_t wlds = cwf_wave_lds(lmax, limit, a.acc_len, fp->ktup);
    const bool lane_kernel = !(wlds <= 160 * 1024 && 2 * (int64_t)max_aln + 2 <= 32767 && lmax <= 16000 &&
                               fp->wind_gap >= 0);
    if (lane_kernel) HIPCHK(ctx->f_scratch.ensure((size_t)(a.lane_ints * lanes)));
    HIPCHK(ctx->f_pairs.ensure(pairs.size()));
    HIPCHK(ctx->f_dist.ensure((size_t)np));
    HIPCHK(ctx->f_ctr.ensure(2));
    HIPCHK(hipMemcpyAsync(ctx->f_pairs.p, pairs.data(), pairs.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                          ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->f_ctr.p, 0, 2 * sizeof(unsigned int), ctx->stream));
    a.codes = ctx->d_codes.p;
    a.offs = ctx->d_offs.p;
    a.lens = ctx->d_lens.p;
    a.pairs = ctx->f_pairs.p;
    a.npairs = np;
    a.out_dist = ctx->f_dist.p;
    a.scratch = lane_kernel ? ctx->f_scratch.p : nullptr;
    a.ovf_list = nullptr;
    a.ovf_n = nullptr;
    a.plist = nullptr;
    a.plist_n = nullptr;
    a.lds_wave = 0;
    a.next = ctx->f_ctr.p;
    a.wL = lmax;
    a.wD = 2 * lmax + 2;
    HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    if (!lane_kernel) {
        // Where the slots for every fragment a pair may make (2 maxAlnLength)
        // leave room for two waves per CU only, a first pass runs three with
        // at least two thirds of the slots; the few pairs that make more
        // fragments than that are listed and run again with all the slots.
        const int64_t full = a.acc_len;
        const size_t third = (size_t)(160 * 1024 / 3) / 512 * 512;   // LDS is granted in 512-byte blocks
        const size_t fixed = cwf_wave_lds(lmax, limit, 0, fp->ktup);
        int64_t cap1 = full;
        if (wlds > third && fixed < third) {
            const int64_t c = (int64_t)((third - fixed) / 6) & ~(int64_t)1;
            if (3 * c >= 2 * full) cap1 = c;
        }
        if (cap1 < full) {
            HIPCHK(ctx->f_scratch.ensure((size_t)np));
            a.acc_len = cap1;
            a.ovf_list = ctx->f_scratch.p;
            a.ovf_n = ctx->f_ctr.p + 1;
            HIPCHK(launch_waves(ctx, a, cwf_wave_lds(lmax, limit, cap1, fp->ktup), np));
            HIPCHK(hipMemsetAsync(ctx->f_ctr.p, 0, sizeof(unsigned int), ctx->stream));
            a.acc_len = full;
            a.ovf_list = nullptr;
            a.ovf_n = nullptr;
            a.plist = ctx->f_scratch.p;
            a.plist_n = ctx->f_ctr.p + 1;
        }
        HIPCHK(launch_waves(ctx, a, wlds, np));
    } else {
        HIPCHK(cwf_launch((int)(lanes / 256), ctx->stream, a));
    }
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->fast_ms = ms;
    HIPCHK(hipMemcpyAsync(out_dist, ctx->f_dist.p, np * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return CWHIP_OK;
}

float cwhip_fast_last_ms(const cwhip_ctx* ctx) { return ctx ? ctx->fast_ms : 0.f; }

}  // extern "C"
