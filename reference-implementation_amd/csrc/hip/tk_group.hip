// Persistent launch over many uneven FP32 products (gfx950).
//
// One grid of `nblocks` workgroups; each takes the next tile index from an
// atomic counter, finds its product through the tile -> product map (or a
// binary search over the inclusive tile prefix when no map is given), runs
// the tile and goes back for another until its queue is drained, then
// drains the other queues.  Every workgroup reaches that exit.  Products are
// sorted by cost on the host, so long tiles start first; product groups keep
// one queue per XCD (a product's tiles on one XCD's L2).
//
//  thin tile  (S <= 16): 256 columns g, all s.  The S x 16 chunk of X and the
//             16 x 256 chunk of Y go through LDS (Y loaded along whichever of
//             k or g has unit stride, so global reads are coalesced); each lane
//             then owns one column g and accumulates S outputs with FMA.  This
//             is the HBM-streaming shape of Tucker mode products.
//  mfma tile  64 x 64 x 32 per step, 4 waves each one 32 x 32 block of
//             v_mfma_f32_32x32x2f32; operands through LDS; row offsets in
//             registers and K offsets staged in LDS per 256-wide chunk, so a
//             step's loads are one memory round trip, issued a step ahead.
#include "tk_device.hpp"

namespace tk {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int GT = 256;                   // thin tile columns
constexpr int KC = 16;                    // thin K chunk (small LDS: six workgroups per CU)
constexpr int YP = GT + 1;                // LDS pitch of the Y chunk
constexpr int SMAX = 16;
constexpr int MT = 64, MK = 32, MP = MT + 4;
constexpr int NBC = MT / 16;              // B columns per lane when B is read along K
constexpr int KO = 256;                   // K offsets staged in LDS per chunk (mfma tile)
// the mfma tile's A and B steps share the thin tile's LDS array
static_assert(2 * MK * MP <= KC * YP + SMAX * KC, "mfma tile staging must fit the shared LDS array");

// Pointers read from a descriptor in memory are generic to the compiler, and
// generic (FLAT) loads wait conservatively; the operands and tables all live
// in global memory, so they are accessed through address-space-1 pointers.
#define TK_GLOBAL __attribute__((address_space(1)))
typedef const float TK_GLOBAL* gf;
typedef float TK_GLOBAL* gfw;
typedef const int64_t TK_GLOBAL* gi;

struct GView {
    int64_t S, G, K;
    gf X, Y, C;
    gfw D;
    gi sX, kX, lX, gY, kY, lY, sD, gD, lD, sC, gC, lC;
    float alpha, beta;
    int beta_zero, round, round_C, y_kfast;
    __device__ __forceinline__ explicit GView(const tk_prob& p)
        : S(p.S), G(p.G), K(p.K), X((gf)p.X), Y((gf)p.Y), C((gf)p.C), D((gfw)p.D), sX((gi)p.sX), kX((gi)p.kX),
          lX((gi)p.lX), gY((gi)p.gY), kY((gi)p.kY), lY((gi)p.lY), sD((gi)p.sD), gD((gi)p.gD), lD((gi)p.lD),
          sC((gi)p.sC), gC((gi)p.gC), lC((gi)p.lC), alpha(p.alpha), beta(p.beta), beta_zero(p.beta_zero),
          round(p.round), round_C(p.round_C), y_kfast(p.y_kfast)
    {
    }
};

__device__ __forceinline__ float rnd_in(float x, int mode)
{
    if (mode == TK_MODE_BF16) return round_bf16(x);
    if (mode == TK_MODE_F16) return round_f16(x);
    return x;
}

// C value of output (l, s, g) as the epilogue scales it.  Tiles read all the
// C values of a lane before any of its D stores: C and D may alias, so a
// load issued after a store has to wait for it.
__device__ __forceinline__ float group_c(const GView& p, int64_t l, int64_t s, int64_t g)
{
    if (p.beta_zero) return 0.0f;
    float c = p.C[p.lC[l] + p.sC[s] + p.gC[g]];
    return p.round_C ? rnd_in(c, p.round) : c;
}

__device__ __forceinline__ void group_store(const GView& p, int64_t l, int64_t s, int64_t g, float acc, float c)
{
    float out = p.alpha * acc;
    if (!p.beta_zero) out += p.beta * c;
    p.D[p.lD[l] + p.sD[s] + p.gD[g]] = out;
}

template <int SB>
__device__ __forceinline__ void thin_tile(const GView& p, int64_t l, int64_t g0, float* lds, int64_t* goff)
{
    float* Yl = lds;               // [KC][YP]
    float* Xl = lds + KC * YP;     // [SMAX][KC]
    const int tid = threadIdx.x;
    float acc[SB];
#pragma unroll
    for (int s = 0; s < SB; ++s) acc[s] = 0.0f;
    const int64_t lx = p.lX[l], ly = p.lY[l];
    const int gn = (int)(p.G - g0 < GT ? p.G - g0 : GT);
    // column offsets of this tile, once (the K loop reuses them)
    __syncthreads();
    goff[tid] = tid < gn ? p.gY[g0 + tid] : 0;
    // dense Y (K consecutive, columns K apart): the tile is one contiguous block
    const bool dense = p.y_kfast == 2;
    for (int64_t k0 = 0; k0 < p.K; k0 += KC) {
        const int kc = (int)(p.K - k0 < KC ? p.K - k0 : KC);
        __syncthreads();
        for (int e = tid; e < SB * KC; e += 256) {
            const int s = e / KC, k = e % KC;
            float v = 0.0f;
            if (s < p.S && k < kc) v = rnd_in(p.X[lx + p.sX[s] + p.kX[k0 + k]], p.round);
            Xl[s * KC + k] = v;
        }
        if (dense) {
            // element (k, gg) at ly + goff[0] + gg * K + k0 + k
            gf base = p.Y + ly + goff[0] + k0;
#pragma unroll 8
            for (int e = tid; e < KC * GT; e += 256) {
                const int k = e % KC, gg = e / KC;
                float v = 0.0f;
                if (k < kc && gg < gn) v = rnd_in(base[(int64_t)gg * p.K + k], p.round);
                Yl[k * YP + gg] = v;
            }
        } else if (p.y_kfast) {
            const int k = tid % KC;
            const int64_t ko = k < kc ? p.kY[k0 + k] : 0;
#pragma unroll 8
            for (int gg = tid / KC; gg < GT; gg += 256 / KC) {
                float v = 0.0f;
                if (k < kc && gg < gn) v = rnd_in(p.Y[ly + ko + goff[gg]], p.round);
                Yl[k * YP + gg] = v;
            }
        } else {
            const int gg = tid;
#pragma unroll 8
            for (int k = 0; k < KC; ++k) {
                float v = 0.0f;
                if (k < kc && gg < gn) v = rnd_in(p.Y[ly + p.kY[k0 + k] + goff[gg]], p.round);
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
        float cv[SB];
#pragma unroll
        for (int s = 0; s < SB; ++s) cv[s] = s < p.S ? group_c(p, l, s, g) : 0.0f;
#pragma unroll
        for (int s = 0; s < SB; ++s)
            if (s < p.S) group_store(p, l, s, g, acc[s], cv[s]);
    }
}

__device__ __forceinline__ void mfma_tile(const GView& p, int64_t l, int64_t m0, int64_t n0, float* lds,
                                          int64_t* kox, int64_t* koy)
{
    float* As = lds;               // [MK][MP]
    float* Bs = lds + MK * MP;     // [MK][MP]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1, il = lane & 31, kl = lane >> 5;
    const int64_t lx = p.lX[l], ly = p.lY[l];
    // fixed rows of this lane: A element (tid % 64, tid / 64 + 4 t).  B
    // element (n, k) with the unit-stride index across lanes: along n, lane
    // column tid % 64 and k = tid / 64 + 4 t; along k, 16 lanes take 16
    // consecutive k of a column, so a lane touches NBC columns
    // (n = tid / 16 + 16 c) at k = tid % 16 + 16 j.
    const int am = tid % MT;
    const bool arow_ok = m0 + am < p.S;
    const int64_t arow = arow_ok ? lx + p.sX[m0 + am] : 0;
    const bool bk = p.y_kfast != 0;
    int64_t bcol[NBC];
    bool bok[NBC];                      // offsets may be negative: validity kept apart
#pragma unroll
    for (int c = 0; c < NBC; ++c) {
        const int n = bk ? tid / 16 + 16 * c : tid % MT;
        bok[c] = n0 + n < p.G;
        bcol[c] = bok[c] ? ly + p.gY[n0 + n] : 0;
    }
    // B element t of this lane: column slot, n and k
    auto bslot = [&](int t) { return bk ? t % NBC : 0; };
    auto bn_of = [&](int t) { return bk ? tid / 16 + 16 * (t % NBC) : tid % MT; };
    auto bk_of = [&](int t) { return bk ? tid % 16 + 16 * (t / NBC) : tid / MT + (256 / MT) * t; };
    float ra[(MT * MK) / 256], rb[(MT * MK) / 256];
    // loads of one K step: offsets of K come from LDS (staged per KO-chunk)
    auto fetch = [&](int64_t k0) {
        const int64_t kb = k0 - (k0 / KO) * KO;
#pragma unroll
        for (int t = 0; t < (MT * MK) / 256; ++t) {
            const int k = tid / MT + (256 / MT) * t;
            ra[t] = (arow_ok && k0 + k < p.K) ? rnd_in(p.X[arow + kox[kb + k]], p.round) : 0.0f;
        }
#pragma unroll
        for (int t = 0; t < (MT * MK) / 256; ++t) {
            const int k = bk_of(t), c = bslot(t);
            rb[t] = (bok[c] && k0 + k < p.K) ? rnd_in(p.Y[bcol[c] + koy[kb + k]], p.round) : 0.0f;
        }
    };
    auto stage = [&](int64_t k0) {      // K offsets for k0 .. k0 + KO
        __syncthreads();
        for (int k = tid; k < KO; k += 256) {
            const bool in = k0 + k < p.K;
            kox[k] = in ? p.kX[k0 + k] : 0;
            koy[k] = in ? p.kY[k0 + k] : 0;
        }
        __syncthreads();
    };
    f32x16 acc = {};
    if (p.K > 0) {
        stage(0);
        fetch(0);
    }
    for (int64_t k0 = 0; k0 < p.K; k0 += MK) {
        __syncthreads();
#pragma unroll
        for (int t = 0; t < (MT * MK) / 256; ++t) As[(tid / MT + (256 / MT) * t) * MP + am] = ra[t];
#pragma unroll
        for (int t = 0; t < (MT * MK) / 256; ++t) Bs[bk_of(t) * MP + bn_of(t)] = rb[t];
        __syncthreads();
        const int64_t kn = k0 + MK;
        if (kn < p.K) {
            if (kn % KO == 0) {
                // the offsets of the next chunk are needed before its first loads
                stage(kn);
            }
            fetch(kn);                  // next step's loads in flight during the MFMAs
        }
#pragma unroll
        for (int kk = 0; kk < MK; kk += 2) {
            const float a = As[(kk + kl) * MP + wm * 32 + il];
            const float b = Bs[(kk + kl) * MP + wn * 32 + il];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(b, a, acc, 0, 0, 0);
        }
    }
    const int64_t m = m0 + wm * 32 + il;
    if (m < p.S) {
        float cv[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t n = n0 + wn * 32 + 8 * (r >> 2) + 4 * kl + (r & 3);
            cv[r] = n < p.G ? group_c(p, l, m, n) : 0.0f;
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t n = n0 + wn * 32 + 8 * (r >> 2) + 4 * kl + (r & 3);
            if (n < p.G) group_store(p, l, m, n, acc[r], cv[r]);
        }
    }
}

// Tiles come from `nq` queues: queue q holds tiles [qbeg[q], qbeg[q + 1])
// (one queue over [0, total) when qbeg is null) and has its own counter,
// counter[32 q].  A workgroup starts with queue blockIdx.x % nq -- with one
// queue per XCD that is its own XCD's, since workgroups are dispatched to the
// XCDs round-robin, so a product's tiles share that XCD's L2 -- and then
// takes what is left in the others, in order, so every workgroup ends after
// every queue is empty.
__global__ __launch_bounds__(256, TK_GROUP_WG_PER_CU) void k_group(const tk_prob* __restrict__ probs, const int64_t* __restrict__ tile_end,
                                               const int* __restrict__ tile_prob, int nprobs, int64_t total,
                                               const int64_t* __restrict__ qbeg, int nq,
                                               unsigned long long* counter)
{
    __shared__ __attribute__((aligned(16))) float lds[KC * YP + SMAX * KC];
    __shared__ int64_t goff[GT];
    __shared__ int64_t kox[KO], koy[KO];
    __shared__ int64_t next;
    const int home = (int)(blockIdx.x % (unsigned)nq);
    for (int qi = 0; qi < nq; ++qi) {
        const int q = (home + qi) % nq;
        const int64_t qb = qbeg ? qbeg[q] : 0, qe = qbeg ? qbeg[q + 1] : total;
        for (;;) {
            if (threadIdx.x == 0) next = qb + (int64_t)atomicAdd(counter + 32 * q, 1ull);
            __syncthreads();
            const int64_t t = next;
            __syncthreads();
            if (t >= qe) break;
            int lo;
            if (tile_prob) {
                lo = tile_prob[t];                       // direct map, one load
            } else {
                int a = 0, b = nprobs - 1;               // first problem with tile_end > t
                while (a < b) {
                    const int mid = (a + b) >> 1;
                    if (tile_end[mid] > t) b = mid; else a = mid + 1;
                }
                lo = a;
            }
            const tk_prob& pd = probs[lo];
            const GView p(pd);
            const int64_t local = t - (lo ? tile_end[lo - 1] : 0);
            const int64_t tiles_s = pd.tiles_s, tiles_g = pd.tiles_g;
            const int cls = pd.cls;
            const int64_t per_l = tiles_s * tiles_g;
            const int64_t l = local / per_l, rem = local % per_l;
            const int64_t ts = rem / tiles_g, tg = rem % tiles_g;
            if (cls == TK_GROUP_THIN) {
                if (p.S <= 4) thin_tile<4>(p, l, tg * GT, lds, goff);
                else if (p.S <= 8) thin_tile<8>(p, l, tg * GT, lds, goff);
                else thin_tile<16>(p, l, tg * GT, lds, goff);
            } else {
                mfma_tile(p, l, ts * MT, tg * MT, lds, kox, koy);
            }
        }
    }
}

// Fill kernel (graph-capturable, unlike a memset node): counters 0 .. nq-1.
__global__ void k_zero_counters(unsigned long long* counter, int nq)
{
    if ((int)threadIdx.x < nq) counter[32 * threadIdx.x] = 0ull;
}

}  // namespace tk

using namespace tk;

extern "C" int tk_launch_group_q(const tk_prob* probs, const int64_t* tile_end, const int* tile_prob, int nprobs,
                                 int64_t total, const int64_t* qbeg, int nq, unsigned long long* counter, int nblocks,
                                 void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (nprobs <= 0 || total <= 0) return 0;
    if (nq < 1 || nq > 64 || (nq > 1 && !qbeg)) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, counter, nq);
    hipLaunchKernelGGL(k_group, dim3((unsigned)nblocks), dim3(256), 0, st, probs, tile_end, tile_prob, nprobs, total,
                       nq > 1 ? qbeg : (const int64_t*)nullptr, nq, counter);
    return (int)hipGetLastError();
}

extern "C" int tk_launch_group(const tk_prob* probs, const int64_t* tile_end, const int* tile_prob, int nprobs,
                               int64_t total, unsigned long long* counter, int nblocks, void* stream)
{
    return tk_launch_group_q(probs, tile_end, tile_prob, nprobs, total, nullptr, 1, counter, nblocks, stream);
}
