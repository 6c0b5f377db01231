// Element-wise comparison of two results of one product, for the
// self-checking executor (kind 12).
//
// The rule is the reference's compare_tensors_{s,d,c,z}
// (reference_implementation/tblis_bindings/tblis_bind.cpp:577-640): two
// values a, b differ when |(a - b) / max(a, b)| > 0.00005, with the signed
// max and the division in the storage precision (a NaN quotient, e.g. 0 / 0,
// does not count); complex values compare each component that way, and
// binary64 complex values also need an absolute difference above 1e-12 in
// one component.  binary16 / bfloat16 values are compared as binary32.
#include "tk_device.hpp"

namespace tk {

template <typename R>
__device__ __forceinline__ bool rel_over(R a, R b)
{
    const R q = (a - b) / (a > b ? a : b);
    return (double)(q < (R)0 ? -q : q) > 0.00005;
}

__device__ __forceinline__ bool differs(int type, const void* x, const void* y, int64_t i)
{
    switch (type) {
    case S_F32: return rel_over(((const float*)x)[i], ((const float*)y)[i]);
    case S_F64: return rel_over(((const double*)x)[i], ((const double*)y)[i]);
    case S_C32: {
        const float* a = (const float*)x + 2 * i;
        const float* b = (const float*)y + 2 * i;
        return rel_over(a[0], b[0]) || rel_over(a[1], b[1]);
    }
    case S_C64: {
        const double* a = (const double*)x + 2 * i;
        const double* b = (const double*)y + 2 * i;
        const double dr = a[0] - b[0], di = a[1] - b[1];
        return (rel_over(a[0], b[0]) || rel_over(a[1], b[1])) &&
               ((dr < 0 ? -dr : dr) > 1e-12 || (di < 0 ? -di : di) > 1e-12);
    }
    case S_F16: return rel_over((float)((const _Float16*)x)[i], (float)((const _Float16*)y)[i]);
    case S_BF16:
        return rel_over(bf16_bits_to_float(((const uint16_t*)x)[i]), bf16_bits_to_float(((const uint16_t*)y)[i]));
    default: return true;
    }
}

__global__ __launch_bounds__(256) void k_compare(tk_compare c)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < c.n; i += (int64_t)gridDim.x * 256) {
        if (differs(c.type, c.x, c.y, i)) {
            const unsigned long long slot = atomicAdd(c.count, 1ull);
            if (slot < TK_COMPARE_KEEP) c.first[slot] = i;
        }
    }
}

}  // namespace tk

extern "C" int tk_launch_compare(const tk_compare* c, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipMemsetAsync(c->count, 0, sizeof(unsigned long long), st);
    if (e != hipSuccess) return (int)e;
    if (c->n <= 0) return 0;
    const int64_t blocks = (c->n + 255) / 256;
    hipLaunchKernelGGL(tk::k_compare, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0, st, *c);
    return (int)hipGetLastError();
}
