// Bandwidth-bound pre- and post-passes for gfx950.
//
// k_prep: unary pre-reduction and storage -> compute conversion.
//   dst[i0 + n0 * (i1 + n1 * i2)] = sum_u op(src[t0[i0] + t1[i1] + t2[i2] + tu[u]])
//   The sum over u runs in ascending u, i.e. in the order of the reference's
//   unary odometer (product.c:370-391), so the reduced value is the same
//   number the reference's sum_A / sum_B holds when the compute type is the
//   reference's sum type; a 16-bit sum (round != 0) is rounded at every
//   addition as the reference's _Float16 sum is.  One lane per output; consecutive lanes take
//   consecutive i0, which is the unit-stride direction the host picks.
// k_prep2: A's and B's pre-passes in one launch.
// k_scatter: compute-type dense buffer -> D of another storage type.
#include "tk_device.hpp"

namespace tk {

template <typename R, bool CPLX>
__device__ __forceinline__ void prep_body(const tk_prep& p, int64_t block, int64_t nblocks)
{
    const int64_t total = p.n0 * p.n1 * p.n2;
    const bool wide_src = p.src_type == S_F64 || p.src_type == S_C64;
    for (int64_t o = block * 256 + threadIdx.x; o < total; o += nblocks * 256) {
        const int64_t i0 = o % p.n0, t = o / p.n0, i1 = t % p.n1, i2 = t / p.n1;
        const int64_t base = p.t0[i0] + p.t1[i1] + p.t2[i2];
        R sr = 0, si = 0;
        // loads in batches of UB (all in flight together), added in u order
        constexpr int UB = 8;
        for (int64_t u0 = 0; u0 < p.nu; u0 += UB) {
            double re[UB], im[UB];
#pragma unroll
            for (int j = 0; j < UB; ++j) {
                re[j] = 0; im[j] = 0;
                if (u0 + j < p.nu) load_storage(p.src, p.src_type, base + p.tu[u0 + j], re[j], im[j]);
            }
#pragma unroll
            for (int j = 0; j < UB; ++j) {
                if (u0 + j >= p.nu) break;
                if (p.round) {
                    // a 16-bit sum, as the reference's _Float16 / __bf16 sum_A:
                    // sum = round16(sum + x), the addition in binary32
                    // (binary64 for a binary64 source) and x unrounded
                    double tr = (double)sr + re[j], ti = (double)si + (p.conj ? -im[j] : im[j]);
                    if (!wide_src) { tr = (double)(float)tr; ti = (double)(float)ti; }
                    sr = (R)round16_d(tr, p.round);
                    if (CPLX) si = (R)round16_d(ti, p.round);
                    continue;
                }
                R vr = (R)re[j], vi = (R)im[j];
                if (p.conj) vi = -vi;
                sr += vr;
                if (CPLX) si += vi;
            }
        }
        Elem<R, CPLX>::st(p.dst, o, Val<R, CPLX>{sr, si});
    }
}

template <typename R, bool CPLX>
__global__ __launch_bounds__(256) void k_prep(tk_prep p)
{
    prep_body<R, CPLX>(p, blockIdx.x, gridDim.x);
}

// Two independent pre-passes (A's and B's) in one launch: workgroups
// [0, ga) take the first, the rest the second.
template <typename R, bool CPLX>
__global__ __launch_bounds__(256) void k_prep2(tk_prep a, tk_prep b, int ga)
{
    if ((int)blockIdx.x < ga) prep_body<R, CPLX>(a, blockIdx.x, ga);
    else prep_body<R, CPLX>(b, blockIdx.x - ga, gridDim.x - ga);
}

template <typename R, bool CPLX>
__global__ __launch_bounds__(256) void k_scatter(tk_scatter s)
{
    const int64_t total = s.M * s.N * s.L;
    for (int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x; o < total; o += (int64_t)gridDim.x * 256) {
        const int64_t m = o % s.M, t = o / s.M, n = t % s.N, l = t / s.N;
        Val<R, CPLX> v = Elem<R, CPLX>::ld(s.src, o);
        store_storage(s.dst, s.dst_type, s.lD[l] + s.mD[m] + s.nD[n], (double)v.re, CPLX ? (double)v.im : 0.0);
    }
}

static unsigned grid_for(int64_t total)
{
    int64_t b = (total + 255) / 256;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (unsigned)b;
}

}  // namespace tk

using namespace tk;

extern "C" int tk_launch_prep(const tk_prep* pp, void* stream)
{
    const tk_prep& p = *pp;
    hipStream_t st = (hipStream_t)stream;
    const int64_t total = p.n0 * p.n1 * p.n2;
    if (total <= 0) return 0;
    const dim3 grid(grid_for(total)), block(256);
    switch (p.ctype) {
    case TK_F32: hipLaunchKernelGGL((k_prep<float, false>), grid, block, 0, st, p); break;
    case TK_F64: hipLaunchKernelGGL((k_prep<double, false>), grid, block, 0, st, p); break;
    case TK_C32: hipLaunchKernelGGL((k_prep<float, true>), grid, block, 0, st, p); break;
    case TK_C64: hipLaunchKernelGGL((k_prep<double, true>), grid, block, 0, st, p); break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

extern "C" int tk_launch_prep2(const tk_prep* pa, const tk_prep* pb, void* stream)
{
    const tk_prep &a = *pa, &b = *pb;
    hipStream_t st = (hipStream_t)stream;
    const int64_t ta = a.n0 * a.n1 * a.n2, tb = b.n0 * b.n1 * b.n2;
    if (ta <= 0) return tk_launch_prep(pb, stream);
    if (tb <= 0 || a.ctype != b.ctype) {
        int rc = tk_launch_prep(pa, stream);
        return rc ? rc : tk_launch_prep(pb, stream);
    }
    const unsigned ga = grid_for(ta), gb = grid_for(tb);
    const dim3 grid(ga + gb), block(256);
    switch (a.ctype) {
    case TK_F32: hipLaunchKernelGGL((k_prep2<float, false>), grid, block, 0, st, a, b, (int)ga); break;
    case TK_F64: hipLaunchKernelGGL((k_prep2<double, false>), grid, block, 0, st, a, b, (int)ga); break;
    case TK_C32: hipLaunchKernelGGL((k_prep2<float, true>), grid, block, 0, st, a, b, (int)ga); break;
    case TK_C64: hipLaunchKernelGGL((k_prep2<double, true>), grid, block, 0, st, a, b, (int)ga); break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

extern "C" int tk_launch_scatter(const tk_scatter* sp, void* stream)
{
    const tk_scatter& s = *sp;
    hipStream_t st = (hipStream_t)stream;
    const int64_t total = s.M * s.N * s.L;
    if (total <= 0) return 0;
    const dim3 grid(grid_for(total)), block(256);
    switch (s.ctype) {
    case TK_F32: hipLaunchKernelGGL((k_scatter<float, false>), grid, block, 0, st, s); break;
    case TK_F64: hipLaunchKernelGGL((k_scatter<double, false>), grid, block, 0, st, s); break;
    case TK_C32: hipLaunchKernelGGL((k_scatter<float, true>), grid, block, 0, st, s); break;
    case TK_C64: hipLaunchKernelGGL((k_scatter<double, true>), grid, block, 0, st, s); break;
    default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}
