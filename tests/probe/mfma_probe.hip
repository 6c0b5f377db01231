// Lane-layout probe for the MFMA builtins the GETT kernels use (test only).
// One wave.  For every lane p of the first operand:
//   run p:      a = (lane == p), b = lane + 1   -> row of A-lane p, and which
//                                                 B-lane feeds each column
//   run 64+p:   a = 1, b = (lane == p)          -> column of B-lane p
// Every run writes the 64 lanes x NR accumulator registers to
// out[(run * 64 + lane) * NR + r] (as double).  A probe with NR registers and
// 128 runs writes 128 * 64 * NR doubles; the host buffer is that size.
#include <hip/hip_runtime.h>

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__global__ void probe_f64_16x16x4(double* out)
{
    const int lane = threadIdx.x;
    for (int run = 0; run < 128; ++run) {
        const int p = run & 63;
        const double a = run < 64 ? (lane == p ? 1.0 : 0.0) : 1.0;
        const double b = run < 64 ? (double)(lane + 1) : (lane == p ? 1.0 : 0.0);
        f64x4 acc = {0, 0, 0, 0};
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        for (int r = 0; r < 4; ++r) out[(run * 64 + lane) * 4 + r] = acc[r];
    }
}

__global__ void probe_f32_32x32x2(double* out)
{
    const int lane = threadIdx.x;
    for (int run = 0; run < 128; ++run) {
        const int p = run & 63;
        const float a = run < 64 ? (lane == p ? 1.0f : 0.0f) : 1.0f;
        const float b = run < 64 ? (float)(lane + 1) : (lane == p ? 1.0f : 0.0f);
        f32x16 acc = {};
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
        for (int r = 0; r < 16; ++r) out[(run * 64 + lane) * 16 + r] = acc[r];
    }
}

// bf16 32x32x16: 8 values per lane; item e of lane p is probed as "lane"
// 8 * p + e (512 input slots, of which runs cover item 0 of each lane and
// every item of lane 0 and 32 -- enough to fix the k order).
__global__ void probe_bf16_32x32x16(double* out)
{
    const int lane = threadIdx.x;
    for (int run = 0; run < 128; ++run) {
        const int p = run & 63;
        bf16x8 a, b;
        for (int e = 0; e < 8; ++e) {
            const int slot_a = run < 64 ? (lane == p && e == 0) : 1;
            a[e] = (__bf16)(float)slot_a;
            // b: distinct small integers per (lane, item) for the first 64
            // runs: value = 1 + lane + 64 * e would not be exact in bf16, so
            // use item 0 only: 1 + lane (exact up to 256)
            const float bv = run < 64 ? (e == 0 ? (float)(lane + 1) : 0.0f) : ((lane == p && e == 0) ? 1.0f : 0.0f);
            b[e] = (__bf16)bv;
        }
        f32x16 acc = {};
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        for (int r = 0; r < 16; ++r) out[(run * 64 + lane) * 16 + r] = acc[r];
    }
}

extern "C" int tk_probe(int which, double* host_out)
{
    const int nr = which == 0 ? 4 : 16;
    const size_t n = (size_t)128 * 64 * nr;
    double* d = nullptr;
    if (hipMalloc(&d, n * sizeof(double)) != hipSuccess) return 1;
    if (which == 0) hipLaunchKernelGGL(probe_f64_16x16x4, dim3(1), dim3(64), 0, 0, d);
    else if (which == 1) hipLaunchKernelGGL(probe_f32_32x32x2, dim3(1), dim3(64), 0, 0, d);
    else hipLaunchKernelGGL(probe_bf16_32x32x16, dim3(1), dim3(64), 0, 0, d);
    hipError_t e = hipMemcpy(host_out, d, n * sizeof(double), hipMemcpyDeviceToHost);
    hipFree(d);
    return e == hipSuccess ? 0 : 2;
}
