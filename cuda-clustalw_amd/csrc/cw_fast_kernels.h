// Shared between the gfx950 k-tuple distance kernel and its host launcher.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// One launch of cw_fast_pairs: the pairs (0-based sequence indices, the
// reference's visiting order) and the parameters of FastPairwiseAlign.  Lane
// g's arrays are lane_ints int32 at scratch + g * lane_ints; o_* are their
// offsets there.
struct FastArgs {
    const uint8_t* codes;
    const int64_t* offs;
    const int32_t* lens;
    const int32_t* pairs;
    int64_t npairs;
    double* out_dist;
    int32_t* scratch;
    int64_t lanes, lane_ints;
    unsigned int* next;       // pair counter (0 at launch)
    int32_t dna, ktup, window, wind_gap, signif, percent, max_aln;
    int32_t limit;            // k-tuple codes: 25^ktup (protein) or 4^ktup (DNA)
    int64_t o_zza, o_zzb;     // per position: the previous position with its k-tuple (sequence 1, 2)
    int64_t o_zzc, o_zzd;     // per k-tuple code: its last position (sequence 1, 2)
    int64_t o_displ, o_slopes, o_diag;   // per diagonal
    int64_t o_acc, acc_len;   // fragments: four rows of acc_len
};

hipError_t cwf_launch(int blocks, hipStream_t st, const FastArgs& a);
