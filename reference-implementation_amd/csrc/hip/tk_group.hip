// Persistent launch over many uneven FP32 products (gfx950).
//
// One grid of `nblocks` workgroups; each takes the next tile index from an
// atomic counter, finds its product by binary search over the inclusive tile
// prefix, runs the tile and goes back for another until the counter passes
// the total.  Every workgroup reaches that exit.  Products are sorted by cost
// on the host, so long tiles start first.
//
//  thin tile  (S <= 16): 256 columns g, all s.  The S x 64 chunk of X and the
//             64 x 256 chunk of Y go through LDS (Y loaded along whichever of
//             k or g has unit stride, so global reads are coalesced); each lane
//             then owns one column g and accumulates S outputs with FMA.  This
//             is the HBM-streaming shape of Tucker mode products.
//  mfma tile  64 x 64 x 16 per step, 4 waves each one 32 x 32 block of
//             v_mfma_f32_32x32x2f32; operands through LDS with table loads.
#include "tk_device.hpp"

namespace tk {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int GT = 256;                   // thin tile columns
constexpr int KC = 64;                    // thin K chunk
constexpr int YP = GT + 1;                // LDS pitch of the Y chunk
constexpr int SMAX = 16;
constexpr int MT = 64, MK = 16, MP = MT + 4;

__device__ __forceinline__ float rnd_in(float x, int mode)
{
    if (mode == TK_MODE_BF16) return round_bf16(x);
    if (mode == TK_MODE_F16) return round_f16(x);
    return x;
}

__device__ __forceinline__ void group_epilogue(const tk_prob& p, int64_t l, int64_t s, int64_t g, float acc)
{
    float out = p.alpha * acc;
    if (!p.beta_zero) {
        float c = p.C[p.lC[l] + p.sC[s] + p.gC[g]];
        if (p.round_C) c = rnd_in(c, p.round);
        out += p.beta * c;
    }
    p.D[p.lD[l] + p.sD[s] + p.gD[g]] = out;
}

template <int SB>
__device__ __forceinline__ void thin_tile(const tk_prob& p, int64_t l, int64_t g0, float* lds)
{
    float* Yl = lds;               // [KC][YP]
    float* Xl = lds + KC * YP;     // [SMAX][KC]
    const int tid = threadIdx.x;
    float acc[SB];
#pragma unroll
    for (int s = 0; s < SB; ++s) acc[s] = 0.0f;
    const int64_t lx = p.lX[l], ly = p.lY[l];
    for (int64_t k0 = 0; k0 < p.K; k0 += KC) {
        const int kc = (int)(p.K - k0 < KC ? p.K - k0 : KC);
        __syncthreads();
        for (int e = tid; e < SB * KC; e += 256) {
            const int s = e / KC, k = e % KC;
            float v = 0.0f;
            if (s < p.S && k < kc) v = rnd_in(p.X[lx + p.sX[s] + p.kX[k0 + k]], p.round);
            Xl[s * KC + k] = v;
        }
        if (p.y_kfast) {
            for (int e = tid; e < KC * GT; e += 256) {
                const int k = e % KC, gg = e / KC;
                float v = 0.0f;
                if (k < kc && g0 + gg < p.G) v = rnd_in(p.Y[ly + p.kY[k0 + k] + p.gY[g0 + gg]], p.round);
                Yl[k * YP + gg] = v;
            }
        } else {
            for (int e = tid; e < KC * GT; e += 256) {
                const int gg = e % GT, k = e / GT;
                float v = 0.0f;
                if (k < kc && g0 + gg < p.G) v = rnd_in(p.Y[ly + p.kY[k0 + k] + p.gY[g0 + gg]], p.round);
                Yl[k * YP + gg] = v;
            }
        }
        __syncthreads();
        for (int k = 0; k < kc; ++k) {
            const float y = Yl[k * YP + tid];
#pragma unroll
            for (int s = 0; s < SB; ++s) acc[s] += Xl[s * KC + k] * y;
        }
    }
    const int64_t g = g0 + tid;
    if (g < p.G) {
#pragma unroll
        for (int s = 0; s < SB; ++s)
            if (s < p.S) group_epilogue(p, l, s, g, acc[s]);
    }
}

__device__ __forceinline__ void mfma_tile(const tk_prob& p, int64_t l, int64_t m0, int64_t n0, float* lds)
{
    float* As = lds;               // [MK][MP]
    float* Bs = lds + MK * MP;     // [MK][MP]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1, il = lane & 31, kl = lane >> 5;
    const int64_t lx = p.lX[l], ly = p.lY[l];
    f32x16 acc = {};
    for (int64_t k0 = 0; k0 < p.K; k0 += MK) {
        __syncthreads();
#pragma unroll
        for (int t = 0; t < (MT * MK) / 256; ++t) {
            const int e = tid + 256 * t;
            const int m = e % MT, k = e / MT;
            float v = 0.0f;
            if (m0 + m < p.S && k0 + k < p.K) v = rnd_in(p.X[lx + p.sX[m0 + m] + p.kX[k0 + k]], p.round);
            As[k * MP + m] = v;
        }
#pragma unroll
        for (int t = 0; t < (MT * MK) / 256; ++t) {
            const int e = tid + 256 * t;
            int n, k;
            if (p.y_kfast) { k = e % MK; n = e / MK; }
            else { n = e % MT; k = e / MT; }
            float v = 0.0f;
            if (n0 + n < p.G && k0 + k < p.K) v = rnd_in(p.Y[ly + p.gY[n0 + n] + p.kY[k0 + k]], p.round);
            Bs[k * MP + n] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < MK; kk += 2) {
            const float a = As[(kk + kl) * MP + wm * 32 + il];
            const float b = Bs[(kk + kl) * MP + wn * 32 + il];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(b, a, acc, 0, 0, 0);
        }
    }
    const int64_t m = m0 + wm * 32 + il;
    if (m < p.S) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t n = n0 + wn * 32 + 8 * (r >> 2) + 4 * kl + (r & 3);
            if (n < p.G) group_epilogue(p, l, m, n, acc[r]);
        }
    }
}

__global__ __launch_bounds__(256) void k_group(const tk_prob* __restrict__ probs, const int64_t* __restrict__ tile_end,
                                               int nprobs, int64_t total, unsigned long long* counter)
{
    __shared__ __attribute__((aligned(16))) float lds[KC * YP + SMAX * KC];
    __shared__ int64_t next;
    for (;;) {
        if (threadIdx.x == 0) next = (int64_t)atomicAdd(counter, 1ull);
        __syncthreads();
        const int64_t t = next;
        __syncthreads();
        if (t >= total) break;
        int lo = 0, hi = nprobs - 1;                 // first problem with tile_end > t
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (tile_end[mid] > t) hi = mid; else lo = mid + 1;
        }
        const tk_prob& p = probs[lo];
        const int64_t local = t - (lo ? tile_end[lo - 1] : 0);
        const int64_t per_l = p.tiles_s * p.tiles_g;
        const int64_t l = local / per_l, rem = local % per_l;
        const int64_t ts = rem / p.tiles_g, tg = rem % p.tiles_g;
        if (p.cls == TK_GROUP_THIN) {
            if (p.S <= 4) thin_tile<4>(p, l, tg * GT, lds);
            else if (p.S <= 8) thin_tile<8>(p, l, tg * GT, lds);
            else thin_tile<16>(p, l, tg * GT, lds);
        } else {
            mfma_tile(p, l, ts * MT, tg * MT, lds);
        }
    }
}

__global__ void k_zero_counter(unsigned long long* counter) { *counter = 0ull; }

}  // namespace tk

using namespace tk;

extern "C" int tk_launch_group(const tk_prob* probs, const int64_t* tile_end, int nprobs, int64_t total,
                               unsigned long long* counter, int nblocks, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (nprobs <= 0 || total <= 0) return 0;
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, st, counter);
    hipLaunchKernelGGL(k_group, dim3((unsigned)nblocks), dim3(256), 0, st, probs, tile_end, nprobs, total, counter);
    return (int)hipGetLastError();
}
