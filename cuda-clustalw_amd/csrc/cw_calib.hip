// Counter calibration for the int32 VALU roof: kernels whose wave-instruction
// count is known exactly, timed with HIP events. Each lane runs eight
// independent register chains of one kind of int32 VALU instruction (or of a
// mix: add / max alternating, or the forward DP loop's opcode mix), 32
// instructions per loop trip,
// so a launch of W waves and T trips issues W * T * 32 VALU instructions in
// its loop. Their SQ_INSTS_VALU readings against that count calibrate the
// counter; their instructions per cycle per SIMD (with the clock from
// SQ_BUSY_CYCLES) give the issue ceiling the DP kernels are measured against.
// Built as its own library (libcwcal.so); nothing in libcwhip.so calls it.
#include <hip/hip_runtime.h>

namespace {

// Instruction kinds: 0 v_add_u32, 1 v_max_i32, 2 v_add3_u32, 3 v_max3_i32,
// 4 v_mov_b32 with a DPP row shift, 5 v_subrev_u32, 6 v_cndmask_b32 (on VCC),
// 7 v_cmp_gt_i32 (to VCC), 8 v_mov_b32.
template <int KIND>
__device__ __forceinline__ void emit(int& x, int b)
{
    if constexpr (KIND == 0) asm volatile("v_add_u32 %0, %0, %1" : "+v"(x) : "v"(b));
    else if constexpr (KIND == 1) asm volatile("v_max_i32 %0, %0, %1" : "+v"(x) : "v"(b));
    else if constexpr (KIND == 2) asm volatile("v_add3_u32 %0, %0, %1, %1" : "+v"(x) : "v"(b));
    else if constexpr (KIND == 3) asm volatile("v_max3_i32 %0, %0, %1, %1" : "+v"(x) : "v"(b));
    else if constexpr (KIND == 4) asm volatile("v_mov_b32_dpp %0, %0 row_shr:1 bound_ctrl:0" : "+v"(x));
    else if constexpr (KIND == 5) asm volatile("v_subrev_u32 %0, %1, %0" : "+v"(x) : "v"(b));
    else if constexpr (KIND == 6) asm volatile("v_cndmask_b32 %0, %0, %1, vcc" : "+v"(x) : "v"(b) : "vcc");
    else if constexpr (KIND == 7) asm volatile("v_cmp_gt_i32 vcc, %0, %1" : : "v"(x), "v"(b) : "vcc");
    else asm volatile("v_mov_b32 %0, %1" : "=v"(x) : "v"(b));
}

// The forward DP loop's opcode mix, 32 to a loop trip (max 8, subrev 7,
// add 5, cndmask 4, max3 2, cmp 2, DPP mov 2, mov 2).
constexpr int kDpMix[32] = {1, 5, 0, 1, 6, 5, 1, 7, 0, 1, 3, 5, 4, 1, 6, 0,
                            5, 1, 8, 6, 3, 5, 0, 1, 4, 5, 6, 1, 0, 7, 8, 5};

// OP 0 alternates add and max, OPs 1..6 are kinds 0..5 alone (see emit),
// OP 7 is the DP mix. (Kinds 6..8 alone are not offered: the compiler pads
// back-to-back inline-asm VCC and move instructions with s_nop, which would
// time the padding rather than the instruction.) K is the instruction's index in the loop trip; it
// works on chain K % 8, so an instruction and the next one on its chain have
// seven others between them.
template <int OP, int K>
__device__ __forceinline__ void op1(int (&a)[8], int b)
{
    if constexpr (OP == 0) emit<((K >> 3) & 1) ? 1 : 0>(a[K & 7], b);
    else if constexpr (OP <= 6) emit<OP - 1>(a[K & 7], b);
    else emit<kDpMix[K]>(a[K & 7], b);
    if constexpr (K < 31) op1<OP, K + 1>(a, b);
}

template <int OP>
__global__ void __launch_bounds__(256) cw_calib_valu(int* out, int trips, int b)
{
    int a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = threadIdx.x + k;
    for (int i = 0; i < trips; ++i) op1<OP, 0>(a, b);
    out[blockIdx.x * blockDim.x + threadIdx.x] = a[0] ^ a[1] ^ a[2] ^ a[3] ^ a[4] ^ a[5] ^ a[6] ^ a[7];
}

typedef void (*calib_kernel)(int*, int, int);
constexpr int kOps = 8;
const calib_kernel kKernels[kOps] = {cw_calib_valu<0>, cw_calib_valu<1>, cw_calib_valu<2>, cw_calib_valu<3>,
                                     cw_calib_valu<4>, cw_calib_valu<5>, cw_calib_valu<6>, cw_calib_valu<7>};

}  // namespace

extern "C" {

// One timed launch of calibration kernel `op` (0..7, see op1) on the current
// device with waves_per_cu waves on every CU (a multiple of 4 up to 32:
// 256-lane workgroups). Returns 0 and fills *ms (event time of the launch),
// *loop_instrs (wave instructions issued by the loop) and *lane_ops
// (loop_instrs * 64); a negative value on a bad argument or a HIP error.
int cwcal_valu(int op, int trips, int waves_per_cu, double* ms, double* loop_instrs, double* lane_ops)
{
    if (op < 0 || op >= kOps || trips <= 0 || waves_per_cu <= 0 || waves_per_cu > 32 || (waves_per_cu & 3) ||
        !ms || !loop_instrs || !lane_ops)
        return -1;
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return -2;
    const int blocks = prop.multiProcessorCount * waves_per_cu / 4;
    int* out = nullptr;
    if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(int)) != hipSuccess) return -3;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    float t = 0.f;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = -4;
    } else {
        hipLaunchKernelGGL(kKernels[op], dim3(blocks), dim3(256), 0, 0, out, 16, 1);   // warm-up
        if (hipGetLastError() != hipSuccess || hipEventRecord(e0, 0) != hipSuccess) rc = -5;
        hipLaunchKernelGGL(kKernels[op], dim3(blocks), dim3(256), 0, 0, out, trips, 1);
        if (hipGetLastError() != hipSuccess || hipEventRecord(e1, 0) != hipSuccess) rc = -5;
        if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess)) rc = -5;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (rc) return rc;
    *ms = t;
    *loop_instrs = (double)blocks * 4 * trips * 32.0;
    *lane_ops = *loop_instrs * 64.0;
    return 0;
}

}  // extern "C"
