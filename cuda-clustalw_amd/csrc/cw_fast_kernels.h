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
    // cw_fast_wave in two passes: the first lists the pairs that run out of
    // fragment slots (ovf_list, count ovf_n; null: none can), the second takes
    // those (plist, count plist_n; null: work item k is pair k)
    int32_t* ovf_list;
    unsigned int* ovf_n;
    const int32_t* plist;
    const unsigned int* plist_n;
    int64_t lds_wave;         // cw_fast_wave: LDS bytes of each wave of a workgroup (16-byte multiple)
    int32_t dna, ktup, window, wind_gap, signif, percent, max_aln;
    int32_t limit;            // k-tuple codes: 25^ktup (protein) or 4^ktup (DNA)
    int64_t o_zza, o_zzb;     // per position: the previous position with its k-tuple (sequence 1, 2)
    int64_t o_zzc, o_zzd;     // per k-tuple code: its last position (sequence 1, 2)
    int64_t o_displ, o_slopes, o_diag;   // per diagonal
    int64_t o_acc, acc_len;   // fragments: four rows of acc_len
    int32_t wL, wD;           // cw_fast_wave: longest sequence, diagonal slots (2 wL + 2)
};

hipError_t cwf_launch(int blocks, hipStream_t st, const FastArgs& a);
hipError_t cwf_launch_wave(int blocks, int waves, size_t lds_bytes, hipStream_t st, const FastArgs& a);

// LDS bytes cw_fast_wave needs for sequences of up to L residues and acc_len
// (even) fragment slots: the fragments (6 bytes each), the diagonal counts and list (int16), the k-tuple
// codes of both sequences, then one region that holds first the position
// lists and the diagonal flags, later the fragment score buckets (scores
// 0 .. ktup L, 4 bytes each).
inline size_t cwf_wave_lds(int L, int limit, int64_t acc_len, int ktup)
{
    const int64_t D = 2 * (int64_t)L + 2;
    const int64_t lists = 4 * ((int64_t)L + 2) + 4 * ((int64_t)limit + 2) + D;
    const int64_t buckets = 4 * ((int64_t)ktup * L + 2);
    return (size_t)(6 * acc_len + 4 * D + 4 * ((int64_t)L + 2) + (lists > buckets ? lists : buckets) + 16);
}
