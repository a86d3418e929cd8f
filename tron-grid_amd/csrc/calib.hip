// calib.hip — device calibration for the benchmark's roofline (not part of the SVGF ABI).
//
// Measures sustained wave64 VALU issue rates of the GPU: every thread of 8 workgroups per CU
// (8 waves per SIMD) runs 16 independent register chains, so issue, not latency, bounds the
// kernel. The chains hold v_fma_f32 only, or one transcendental kind (v_exp_f32, v_log_f32,
// v_rcp_f32, v_sqrt_f32), or v_exp_f32 and v_fma_f32 mixed 1:1 and 1:7 (the a-trous taps issue
// two transcendentals among 16 instructions). Rates are wave64 instructions per second over the
// whole device, best of 5 timed launches (hipEvents). Each thread stores one float, in bounds
// of the buffer the host allocates for exactly the launched threads.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int CHAINS = 16;
constexpr int BLOCK = 256;
constexpr int OPS = 7;

template <int OP>
__global__ __launch_bounds__(BLOCK) void k_valu(float* out, float a, float b, int iters)
{
    float x[CHAINS];
#pragma unroll
    for (int i = 0; i < CHAINS; ++i) x[i] = a + 0.001f * (float)(threadIdx.x + i);
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < CHAINS; ++i) {
            if constexpr (OP == 0) x[i] = fmaf(x[i], a, b);
            else if constexpr (OP == 1) x[i] = __builtin_amdgcn_exp2f(x[i]);
            else if constexpr (OP == 2) x[i] = __builtin_amdgcn_logf(x[i]);
            else if constexpr (OP == 3) x[i] = __builtin_amdgcn_rcpf(x[i]);
            else if constexpr (OP == 4) x[i] = __builtin_amdgcn_sqrtf(x[i]);
            else if constexpr (OP == 5) x[i] = (i & 1) ? __builtin_amdgcn_exp2f(x[i]) : fmaf(x[i], a, b);
            else x[i] = (i % 8 == 0) ? __builtin_amdgcn_exp2f(x[i]) : fmaf(x[i], a, b);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < CHAINS; ++i) s += x[i];
    out[blockIdx.x * BLOCK + threadIdx.x] = s;
}

template <int OP>
void launch(float* out, int blocks, int iters)
{
    hipLaunchKernelGGL(k_valu<OP>, dim3(blocks), dim3(BLOCK), 0, 0, out, 0.999f, 0.0005f, iters);
}

void launch_op(int op, float* out, int blocks, int iters)
{
    switch (op) {
    case 0: launch<0>(out, blocks, iters); break;
    case 1: launch<1>(out, blocks, iters); break;
    case 2: launch<2>(out, blocks, iters); break;
    case 3: launch<3>(out, blocks, iters); break;
    case 4: launch<4>(out, blocks, iters); break;
    case 5: launch<5>(out, blocks, iters); break;
    default: launch<6>(out, blocks, iters); break;
    }
}

}  // namespace

extern "C" __attribute__((visibility("default")))
// Sustained wave64 instructions per second of chains of kind `op` on device `device`: 0 FMA,
// 1 exp2, 2 log2, 3 reciprocal, 4 square root, 5 exp2 and FMA alternating, 6 one exp2 in eight
// instructions. Returns 0 on success, else the HIP error (hipErrorInvalidValue for a bad op).
int svgf_calib_valu_op(int device, int op, double* wave_instr_per_s)
{
    if (!wave_instr_per_s || op < 0 || op >= OPS) return (int)hipErrorInvalidValue;
    hipError_t e = hipSetDevice(device);
    int cus = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) return (int)e;
    const int blocks = cus * 8, iters = 4096;
    float* out = nullptr;
    if ((e = hipMalloc(&out, (size_t)blocks * BLOCK * sizeof(float))) != hipSuccess) return (int)e;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float best = 1e30f;
    if ((e = hipEventCreate(&e0)) == hipSuccess && (e = hipEventCreate(&e1)) == hipSuccess) {
        launch_op(op, out, blocks, iters);  // warm-up
        e = hipDeviceSynchronize();
        for (int r = 0; r < 5 && e == hipSuccess; ++r) {
            if ((e = hipEventRecord(e0, 0)) != hipSuccess) break;
            launch_op(op, out, blocks, iters);
            if ((e = hipEventRecord(e1, 0)) != hipSuccess) break;
            if ((e = hipEventSynchronize(e1)) != hipSuccess) break;
            float ms = 0.0f;
            if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) break;
            best = ms < best ? ms : best;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    if (e != hipSuccess) return (int)e;
    const double instr = (double)blocks * (BLOCK / 64) * iters * CHAINS;  // wave64 instructions
    *wave_instr_per_s = instr / (best * 1e-3);
    return 0;
}
