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
//  mfma tile  128 x 128 outputs, four independent 64 x 64 wave blocks over
//             the whole K, 2 x 2 MFMA blocks each (FP32, or BF16 / F16
//             inputs for those precisions) fed straight from global memory
//             (K offsets staged in LDS per 256-deep chunk).
#include "tk_device.hpp"

#include <type_traits>

namespace tk {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int GT = 256;                   // thin tile columns
constexpr int KC = 16;                    // thin K chunk (its X chunk is one element per lane: SMAX KC <= 256)
constexpr int YP = GT + 1;                // LDS pitch of the Y chunk
constexpr int SMAX = 16;
constexpr int MT = TK_GROUP_MT;           // mfma tile edge
constexpr int KO = 256;                   // K offsets staged in LDS per chunk (mfma tile)
static_assert(MT == 128, "the mfma tile is four 64 x 64 wave blocks");
static_assert(SMAX * KC <= 256, "a thin tile loads its X chunk one element per lane");

// Pointers read from a descriptor in memory are generic to the compiler, and
// generic (FLAT) loads wait conservatively; the operands and tables all live
// in global memory, so they are accessed through address-space-1 pointers.
#define TK_GLOBAL __attribute__((address_space(1)))
#define TK_CONST __attribute__((address_space(4)))
typedef const float TK_GLOBAL* gf;
typedef float TK_GLOBAL* gfw;
typedef const int64_t TK_GLOBAL* gi;

// A product's descriptor, read field by field where a field is used (scalar
// loads through the constant address space) rather than copied whole into
// registers for the tile: the persistent loop has too few scalar registers
// to hold every table pointer at once.
struct GView {
    const tk_prob TK_CONST* d;
    __device__ __forceinline__ explicit GView(const tk_prob TK_CONST* p) : d(p) {}
    __device__ __forceinline__ int64_t S() const { return d->S; }
    __device__ __forceinline__ int64_t G() const { return d->G; }
    __device__ __forceinline__ int64_t K() const { return d->K; }
    __device__ __forceinline__ int64_t L() const { return d->L; }
    __device__ __forceinline__ gf X() const { return (gf)d->X; }
    __device__ __forceinline__ gf Y() const { return (gf)d->Y; }
    __device__ __forceinline__ gf C() const { return (gf)d->C; }
    __device__ __forceinline__ gfw D() const { return (gfw)d->D; }
    __device__ __forceinline__ gi sX() const { return (gi)d->sX; }
    __device__ __forceinline__ gi kX() const { return (gi)d->kX; }
    __device__ __forceinline__ gi lX() const { return (gi)d->lX; }
    __device__ __forceinline__ gi gY() const { return (gi)d->gY; }
    __device__ __forceinline__ gi kY() const { return (gi)d->kY; }
    __device__ __forceinline__ gi lY() const { return (gi)d->lY; }
    __device__ __forceinline__ gi sD() const { return (gi)d->sD; }
    __device__ __forceinline__ gi gD() const { return (gi)d->gD; }
    __device__ __forceinline__ gi lD() const { return (gi)d->lD; }
    __device__ __forceinline__ gi sC() const { return (gi)d->sC; }
    __device__ __forceinline__ gi gC() const { return (gi)d->gC; }
    __device__ __forceinline__ gi lC() const { return (gi)d->lC; }
    __device__ __forceinline__ float alpha() const { return d->alpha; }
    __device__ __forceinline__ float beta() const { return d->beta; }
    __device__ __forceinline__ int beta_zero() const { return d->beta_zero; }
    __device__ __forceinline__ int round() const { return d->round; }
    __device__ __forceinline__ int round_C() const { return d->round_C; }
    __device__ __forceinline__ int y_kfast() const { return d->y_kfast; }
};

__device__ __forceinline__ float rnd_in(float x, int mode)
{
    if (mode == TK_MODE_BF16) return round_bf16(x);
    if (mode == TK_MODE_F16) return round_f16(x);
    return x;
}

// C value at element offset `off` as the epilogue scales it.  Tiles read all
// the C values of a lane before any of its D stores: C and D may alias, so a
// load issued after a store has to wait for it.
__device__ __forceinline__ float group_c(const GView& p, int64_t off)
{
    if (p.beta_zero()) return 0.0f;
    const float c = p.C()[off];
    return p.round_C() ? rnd_in(c, p.round()) : c;
}

__device__ __forceinline__ void group_store(const GView& p, int64_t off, float acc, float c)
{
    float out = p.alpha() * acc;
    if (!p.beta_zero()) out += p.beta() * c;
    p.D()[off] = out;
}

template <int RM>
__device__ __forceinline__ float rnd_m(float x)
{
    if (RM == TK_MODE_BF16) return round_bf16(x);
    if (RM == TK_MODE_F16) return round_f16(x);
    return x;
}

// Thin tile.  X's chunk (this lane: s = tid / KC, k = tid % KC) and Y's
// chunk go through LDS.  Dense Y (K consecutive, columns K apart: the tile is
// one contiguous block) is prefetched: a chunk's loads, unconditional at
// valid addresses and masked when stored, are issued into registers right
// after the previous chunk reached LDS, so they are in flight during its
// arithmetic.  Other Y layouts load each chunk when it is needed.  The
// offsets of C and D are read with the column offsets at the start (the
// lane's column in registers, the rows in LDS `soff`), so the epilogue waits
// on no table; so are X's first KO K offsets (into `kxs`), so a chunk's X
// load depends on no table load of its own.
template <int SB, int RM>
__device__ __forceinline__ void thin_tile(const GView& p, int64_t l, int64_t g0, float* lds, int64_t* goff,
                                          int64_t* soff, int64_t* kxs)
{
    constexpr int NY = KC * GT / 256;           // Y elements per lane and chunk
    float* Yl = lds;               // [KC][YP]
    float* Xl = lds + KC * YP;     // [KC][SMAX]: a k's rows contiguous, read four at a time
    const int tid = threadIdx.x;
    float acc[SB];
#pragma unroll
    for (int s = 0; s < SB; ++s) acc[s] = 0.0f;
    const int64_t lx = p.lX()[l], ly = p.lY()[l];
    const int gn = (int)(p.G() - g0 < GT ? p.G() - g0 : GT);
    const int gl = tid < gn ? tid : 0;
    const int nks = (int)(p.K() < KO ? p.K() : KO);
    const int xs = tid / KC, xk = tid % KC;
    const bool x_row = tid < SB * KC && xs < p.S();
    const int si = tid < SB && tid < p.S() ? tid : 0;
    // every table read of the tile, issued together at valid indices (C's
    // tables exist whether or not beta is 0) before any of them is waited
    // for: the column offsets of Y, D and C, the rows of D and C, X's K
    // offsets and this lane's X row
    const int64_t gy = p.gY()[g0 + gl], gd = p.gD()[g0 + gl], gc = p.gC()[g0 + gl];
    const int64_t ldl = p.lD()[l], lcl = p.lC()[l];
    const int64_t sd = p.sD()[si], sc = p.sC()[si];
    const int64_t kx0 = p.kX()[tid < nks ? tid : 0];
    const int64_t sx = p.sX()[x_row ? xs : 0];
    __syncthreads();
    goff[tid] = gy;
    if (tid < SB) {
        soff[tid] = sd;
        soff[SMAX + tid] = p.beta_zero() ? 0 : sc;
    }
    if (tid < nks) kxs[tid] = kx0;
    for (int k = tid + 256; k < nks; k += 256) kxs[k] = p.kX()[k];
    const int64_t dcol = ldl + gd;
    const int64_t ccol = p.beta_zero() ? 0 : lcl + gc;
    const int64_t xrow = lx + sx;
    __syncthreads();
    // X's K offset of this lane's chunk element (staged for k < KO)
    auto kx_of = [&](int64_t k) -> int64_t { return k < KO ? kxs[k] : p.kX()[k]; };
    const bool dense = p.y_kfast() == 2;
    const int64_t g0off = goff[0];
    float yv[NY], xv = 0.0f;
    // dense chunk loads: element (k, gg) at ly + g0off + gg K + k0 + k, one
    // base per chunk and 32-bit offsets (the host takes this path only when
    // K < 2^21)
    auto fetch_dense = [&](int64_t k0) {
        const int kc = (int)(p.K() - k0 < KC ? p.K() - k0 : KC);
        xv = p.X()[xrow + kx_of(k0 + (xk < kc ? xk : 0))];
        gf yb = p.Y() + ly + g0off + k0;
        const uint32_t kk = (uint32_t)(xk < kc ? xk : 0);
#pragma unroll
        for (int i = 0; i < NY; ++i) {
            const int gg = tid / KC + (256 / KC) * i;
            yv[i] = yb[(uint32_t)(gg < gn ? gg : 0) * (uint32_t)p.K() + kk];
        }
    };
    if (dense && p.K() > 0) fetch_dense(0);
    for (int64_t k0 = 0; k0 < p.K(); k0 += KC) {
        const int kc = (int)(p.K() - k0 < KC ? p.K() - k0 : KC);
        __syncthreads();                          // the previous chunk's LDS reads are done
        if (!dense) xv = p.X()[xrow + kx_of(k0 + (xk < kc ? xk : 0))];
        if (tid < SB * KC) Xl[xk * SMAX + xs] = (x_row && xk < kc) ? rnd_m<RM>(xv) : 0.0f;
        if (dense) {
#pragma unroll
            for (int i = 0; i < NY; ++i) {
                const int gg = tid / KC + (256 / KC) * i;
                Yl[xk * YP + gg] = (xk < kc && gg < gn) ? rnd_m<RM>(yv[i]) : 0.0f;
            }
        } else if (p.y_kfast()) {
            const int k = tid % KC;
            const int64_t ko = p.kY()[k0 + (k < kc ? k : 0)];
#pragma unroll 8
            for (int gg = tid / KC; gg < GT; gg += 256 / KC) {
                float v = 0.0f;
                if (k < kc && gg < gn) v = rnd_m<RM>(p.Y()[ly + ko + goff[gg]]);
                Yl[k * YP + gg] = v;
            }
        } else {
            const int gg = tid;
#pragma unroll 8
            for (int k = 0; k < KC; ++k) {
                float v = 0.0f;
                if (k < kc && gg < gn) v = rnd_m<RM>(p.Y()[ly + p.kY()[k0 + k] + goff[gg]]);
                Yl[k * YP + gg] = v;
            }
        }
        __syncthreads();
        // the registers are free again: the next dense chunk's loads go out
        // now and are in flight during this chunk's arithmetic
        if (dense && k0 + KC < p.K()) fetch_dense(k0 + KC);
        for (int k = 0; k < kc; ++k) {
            const float y = Yl[k * YP + tid];
#pragma unroll
            for (int s4 = 0; s4 < SB / 4; ++s4) {
                const float4 x4 = reinterpret_cast<const float4*>(Xl + k * SMAX)[s4];
                acc[4 * s4] += x4.x * y;
                acc[4 * s4 + 1] += x4.y * y;
                acc[4 * s4 + 2] += x4.z * y;
                acc[4 * s4 + 3] += x4.w * y;
            }
        }
    }
    if (tid < gn) {
        float cv[SB];
#pragma unroll
        for (int s = 0; s < SB; ++s) cv[s] = s < p.S() ? group_c(p, ccol + soff[SMAX + s]) : 0.0f;
#pragma unroll
        for (int s = 0; s < SB; ++s)
            if (s < p.S()) group_store(p, dcol + soff[s], acc[s], cv[s]);
    }
}

// MFMA tile: 128 x 128 outputs, four independent 64 x 64 wave blocks (wave
// w: rows 64 (w & 1), columns 64 (w >> 1)), each over the whole K with no
// exchange between the waves.  A wave feeds its 2 x 2 MFMA blocks with
// operands loaded from global memory straight into the MFMA lane layout --
// two rows of A and two columns of B per lane and K step, so one load per
// MFMA -- a batch of loads in flight at a time, unconditional at valid
// addresses and masked afterwards.  The K offsets, shared by the four waves,
// are staged in LDS per 256-deep chunk.  Gathered loads are the bound of
// these tiles (the texture addresser takes one 64-lane address set per load
// instruction), so the tile is as large as the accumulators allow.
//   FP32 (RM = 0): v_mfma_f32_32x32x2f32, lane l supplies row l % 32 at
//     k = 2 j + l / 32;
//   BF16 / F16 inputs (RM = TK_MODE_BF16 / TK_MODE_F16, FP32 accumulate):
//     v_mfma_f32_32x32x16_{bf16,f16}, lane l supplies row l % 32 at the 8
//     consecutive k = 16 j + 8 (l / 32) + 0..7, each rounded once (RNE).
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

template <int RM>
__device__ __forceinline__ uint32_t pack16(float lo, float hi)
{
    if (RM == TK_MODE_BF16) {
        const uint32_t a = __float_as_uint(round_bf16(lo)) >> 16, b = __float_as_uint(round_bf16(hi)) & 0xffff0000u;
        return a | b;
    }
    const _Float16 x = (_Float16)lo, y = (_Float16)hi;
    return (uint32_t)__builtin_bit_cast(uint16_t, x) | ((uint32_t)__builtin_bit_cast(uint16_t, y) << 16);
}

// NA (NB): the wave's 32-row (column) blocks that hold an element of the
// product (1 or 2): a block wholly past the edge is neither loaded nor
// multiplied.
template <int RM, int NA, int NB>
__device__ __forceinline__ void mfma_chunk(gf Xb, gf Yb, const int32_t (&arow)[2], const int32_t (&bcol)[2],
                                           const bool (&a_ok)[2], const bool (&b_ok)[2], const int32_t* kox,
                                           const int32_t* koy, int kc, int kl, f32x16 (&acc)[2][2])
{
    if (RM == 0) {
        // two register sets: the loads of the next JB steps are in flight
        // while the MFMAs of the current ones run
        constexpr int JB = 4;
        float a0[JB][2], b0[JB][2], a1[JB][2], b1[JB][2];
        auto issue = [&](int j0, float (&a)[JB][2], float (&b)[JB][2]) {
#pragma unroll
            for (int j = 0; j < JB; ++j) {
                const int k = j0 + 2 * j + kl;
                const int32_t ka = kox[k < kc ? k : 0], kb = koy[k < kc ? k : 0];
#pragma unroll
                for (int i = 0; i < NA; ++i) a[j][i] = Xb[(uint32_t)(arow[i] + ka)];
#pragma unroll
                for (int i = 0; i < NB; ++i) b[j][i] = Yb[(uint32_t)(bcol[i] + kb)];
            }
        };
        auto compute = [&](int j0, const float (&a)[JB][2], const float (&b)[JB][2]) {
#pragma unroll
            for (int j = 0; j < JB; ++j) {
                const bool in = j0 + 2 * j + kl < kc;
#pragma unroll
                for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                    for (int ia = 0; ia < NA; ++ia)
                        acc[ib][ia] = __builtin_amdgcn_mfma_f32_32x32x2f32((in && b_ok[ib]) ? b[j][ib] : 0.0f,
                                                                           (in && a_ok[ia]) ? a[j][ia] : 0.0f,
                                                                           acc[ib][ia], 0, 0, 0);
            }
        };
        issue(0, a0, b0);
        for (int j0 = 0; j0 < kc; j0 += 4 * JB) {
            const int j1 = j0 + 2 * JB, j2 = j1 + 2 * JB;
            if (j1 < kc) issue(j1, a1, b1);
            compute(j0, a0, b0);
            if (j1 >= kc) break;
            if (j2 < kc) issue(j2, a0, b0);
            compute(j1, a1, b1);
        }
    } else {
        // 16-deep steps: A's 16 loads packed, then B's (a second register
        // set for the next step's loads would spill beside 64 accumulators)
        for (int j0 = 0; j0 < kc; j0 += 16) {
            uint4 ua[2], ub[2];
            auto load_pack = [&](auto NI_, gf base, const int32_t (&off)[2], const bool (&ok)[2], const int32_t* ko,
                                 uint4 (&out)[2]) {
                constexpr int NI = decltype(NI_)::value;
                float v[2][8];
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const int k = j0 + 8 * kl + t;
                    const int32_t kk = ko[k < kc ? k : 0];
#pragma unroll
                    for (int i = 0; i < NI; ++i) v[i][t] = base[(uint32_t)(off[i] + kk)];
                }
#pragma unroll
                for (int i = 0; i < NI; ++i) {
                    uint32_t pk[4];
#pragma unroll
                    for (int t = 0; t < 8; t += 2) {
                        const int k = j0 + 8 * kl + t;
                        pk[t / 2] = pack16<RM>((ok[i] && k < kc) ? v[i][t] : 0.0f, (ok[i] && k + 1 < kc) ? v[i][t + 1] : 0.0f);
                    }
                    out[i] = make_uint4(pk[0], pk[1], pk[2], pk[3]);
                }
            };
            load_pack(std::integral_constant<int, NA>{}, Xb, arow, a_ok, kox, ua);
            load_pack(std::integral_constant<int, NB>{}, Yb, bcol, b_ok, koy, ub);
#pragma unroll
            for (int ib = 0; ib < NB; ++ib)
#pragma unroll
                for (int ia = 0; ia < NA; ++ia) {
                    if (RM == TK_MODE_BF16)
                        acc[ib][ia] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            __builtin_bit_cast(bf16x8, ub[ib]), __builtin_bit_cast(bf16x8, ua[ia]), acc[ib][ia], 0, 0, 0);
                    else
                        acc[ib][ia] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                            __builtin_bit_cast(f16x8, ub[ib]), __builtin_bit_cast(f16x8, ua[ia]), acc[ib][ia], 0, 0, 0);
                }
        }
    }
}

template <int RM>
__device__ __forceinline__ void mfma_tile(const GView& p, int64_t l, int64_t m0, int64_t n0, int64_t* kox,
                                          int64_t* koy, int64_t* eoff)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int il = lane & 31, kl = lane >> 5;
    const int64_t lx = p.lX()[l], ly = p.lY()[l];
    const int64_t wm0 = m0 + 64 * (wave & 1), wn0 = n0 + 64 * (wave >> 1);
    // a wave whose block lies past the edge only helps stage the K offsets;
    // a2 (b2): the wave's second 32-row (column) block holds an element
    const bool active = wm0 < p.S() && wn0 < p.G();
    const bool a2 = wm0 + 32 < p.S(), b2 = wn0 + 32 < p.G();
    // operands addressed from their lowest readable element with 32-bit
    // offsets (the host checked every element lies within 2^29 of it)
    const int64_t xlo = p.d->xlo, ylo = p.d->ylo;
    gf Xb = p.X() + xlo, Yb = p.Y() + ylo;
    int32_t arow[2], bcol[2];
    bool a_ok[2], b_ok[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int64_t m = wm0 + 32 * i + il, n = wn0 + 32 * i + il;
        a_ok[i] = m < p.S();
        b_ok[i] = n < p.G();
        arow[i] = (int32_t)(lx + p.sX()[a_ok[i] ? m : 0] - xlo);
        bcol[i] = (int32_t)(ly + p.gY()[b_ok[i] ? n : 0] - ylo);
    }
    // the tile's row and column offsets of C and D, read now into LDS (the
    // epilogue then waits on no table): eoff = C rows | D rows | C columns |
    // D columns.  Every lane loads (lanes MT.. repeat the first MT, at valid
    // indices) and the first MT store, so no load sits in a branch.
    {
        const int et = tid & (MT - 1);
        const int64_t m = m0 + et < p.S() ? m0 + et : 0, n = n0 + et < p.G() ? n0 + et : 0;
        const int64_t esc = p.sC()[m], esd = p.sD()[m], egc = p.gC()[n], egd = p.gD()[n];
        if (tid < MT) {
            eoff[tid] = p.beta_zero() ? 0 : esc;
            eoff[MT + tid] = esd;
            eoff[2 * MT + tid] = p.beta_zero() ? 0 : egc;
            eoff[3 * MT + tid] = egd;
        }
    }
    f32x16 acc[2][2];
#pragma unroll
    for (int ib = 0; ib < 2; ++ib)
#pragma unroll
        for (int ia = 0; ia < 2; ++ia) acc[ib][ia] = f32x16{};
    int32_t* kx = reinterpret_cast<int32_t*>(kox);
    int32_t* ky = reinterpret_cast<int32_t*>(koy);
    for (int64_t k0 = 0; k0 < p.K(); k0 += KO) {
        const int kc = (int)(p.K() - k0 < KO ? p.K() - k0 : KO);
        __syncthreads();                         // the previous chunk's offsets are no longer read
        for (int k = tid; k < kc; k += 256) {
            kx[k] = (int32_t)p.kX()[k0 + k];
            ky[k] = (int32_t)p.kY()[k0 + k];
        }
        __syncthreads();
        if (!active) continue;
        if (a2 && b2) mfma_chunk<RM, 2, 2>(Xb, Yb, arow, bcol, a_ok, b_ok, kx, ky, kc, kl, acc);
        else if (a2) mfma_chunk<RM, 2, 1>(Xb, Yb, arow, bcol, a_ok, b_ok, kx, ky, kc, kl, acc);
        else if (b2) mfma_chunk<RM, 1, 2>(Xb, Yb, arow, bcol, a_ok, b_ok, kx, ky, kc, kl, acc);
        else mfma_chunk<RM, 1, 1>(Xb, Yb, arow, bcol, a_ok, b_ok, kx, ky, kc, kl, acc);
    }
    __syncthreads();                             // eoff is visible (also when K is 0)
    if (!active) return;
    const int64_t lc = p.beta_zero() ? 0 : p.lC()[l], ld = p.lD()[l];
    // acc[ib][ia][r] of lane holds out[wn0 + 32 ib + 8 (r / 4) + 4 kl + r % 4][wm0 + 32 ia + il];
    // per 32 x 32 block, the lane's C values are read before its D stores
#pragma unroll
    for (int ia = 0; ia < 2; ++ia) {
        if (!a_ok[ia]) continue;
        const int mt = (int)(wm0 - m0) + 32 * ia + il;
        const int64_t rc = lc + eoff[mt], rd = ld + eoff[MT + mt];
#pragma unroll
        for (int ib = 0; ib < 2; ++ib) {
            float cv[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int nt = (int)(wn0 - n0) + 32 * ib + 8 * (r >> 2) + 4 * kl + (r & 3);
                float c = 0.0f;
                if (!p.beta_zero() && n0 + nt < p.G()) {
                    c = p.C()[rc + eoff[2 * MT + nt]];
                    if (p.round_C()) c = rnd_in(c, p.round());
                }
                cv[r] = c;
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int nt = (int)(wn0 - n0) + 32 * ib + 8 * (r >> 2) + 4 * kl + (r & 3);
                if (n0 + nt < p.G()) {
                    float out = p.alpha() * acc[ib][ia][r];
                    if (!p.beta_zero()) out += p.beta() * cv[r];
                    p.D()[rd + eoff[3 * MT + nt]] = out;
                }
            }
        }
    }
}

// Thin tile of a product whose tables are all index * stride, over a dense
// Y (element (k, g) at g K + k): thin_tile's data flow with every offset
// computed from the descriptor's strides, so the tile's first loads are its
// data.  Y's chunk loads are issued one chunk ahead, as in thin_tile.
template <int SB, int RM>
__device__ __forceinline__ void thin_affine(const GView& p, int64_t l, int64_t g0, float* lds)
{
    constexpr int NY = KC * GT / 256;           // Y elements per lane and chunk
    float* Yl = lds;               // [KC][YP]
    float* Xl = lds + KC * YP;     // [KC][SMAX]: a k's rows contiguous, read four at a time
    const int tid = threadIdx.x;
    const tk_prob TK_CONST* d = p.d;
    const int64_t K = p.K(), S = p.S();
    float acc[SB];
#pragma unroll
    for (int s = 0; s < SB; ++s) acc[s] = 0.0f;
    const int gn = (int)(p.G() - g0 < GT ? p.G() - g0 : GT);
    const int xs = tid / KC, xk = tid % KC;
    const bool x_row = tid < SB * KC && xs < S;
    const int64_t skX = d->skX;
    gf xb = p.X() + l * d->slX + (int64_t)(x_row ? xs : 0) * d->ssX;
    // 32-bit offsets within the tile's block: the host takes the dense path
    // only when K < 2^21
    gf yb0 = p.Y() + l * d->slY + g0 * K;
    float yv[NY], xv = 0.0f;
    auto fetch = [&](int64_t k0) {
        const int kc = (int)(K - k0 < KC ? K - k0 : KC);
        const uint32_t kk = (uint32_t)(xk < kc ? xk : 0);
        xv = xb[(k0 + kk) * skX];
        gf yb = yb0 + k0;
#pragma unroll
        for (int i = 0; i < NY; ++i) {
            const int gg = tid / KC + (256 / KC) * i;
            yv[i] = yb[(uint32_t)(gg < gn ? gg : 0) * (uint32_t)K + kk];
        }
    };
    if (K > 0) fetch(0);
    for (int64_t k0 = 0; k0 < K; k0 += KC) {
        const int kc = (int)(K - k0 < KC ? K - k0 : KC);
        __syncthreads();                          // the previous chunk's LDS reads are done
        if (tid < SB * KC) Xl[xk * SMAX + xs] = (x_row && xk < kc) ? rnd_m<RM>(xv) : 0.0f;
#pragma unroll
        for (int i = 0; i < NY; ++i) {
            const int gg = tid / KC + (256 / KC) * i;
            Yl[xk * YP + gg] = (xk < kc && gg < gn) ? rnd_m<RM>(yv[i]) : 0.0f;
        }
        __syncthreads();
        if (k0 + KC < K) fetch(k0 + KC);
        for (int k = 0; k < kc; ++k) {
            const float y = Yl[k * YP + tid];
#pragma unroll
            for (int s4 = 0; s4 < SB / 4; ++s4) {
                const float4 x4 = reinterpret_cast<const float4*>(Xl + k * SMAX)[s4];
                acc[4 * s4] += x4.x * y;
                acc[4 * s4 + 1] += x4.y * y;
                acc[4 * s4 + 2] += x4.z * y;
                acc[4 * s4 + 3] += x4.w * y;
            }
        }
    }
    if (tid < gn) {
        const int64_t g = g0 + tid;
        const int64_t dcol = l * d->slD + g * d->sgD;
        const int64_t ccol = p.beta_zero() ? 0 : l * d->slC + g * d->sgC;
        const int64_t ssD = d->ssD, ssC = d->ssC;
        float cv[SB];
#pragma unroll
        for (int s = 0; s < SB; ++s) cv[s] = s < S ? group_c(p, ccol + s * ssC) : 0.0f;
#pragma unroll
        for (int s = 0; s < SB; ++s)
            if (s < S) group_store(p, dcol + s * ssD, acc[s], cv[s]);
    }
}

template <int RM>
__device__ __forceinline__ void thin_dispatch(const GView& p, int64_t l, int64_t g0, float* lds, int64_t* goff,
                                              int64_t* soff, int64_t* kxs)
{
    if (p.d->affine) {
        if (p.S() <= 4) thin_affine<4, RM>(p, l, g0, lds);
        else if (p.S() <= 8) thin_affine<8, RM>(p, l, g0, lds);
        else thin_affine<16, RM>(p, l, g0, lds);
        return;
    }
    if (p.S() <= 4) thin_tile<4, RM>(p, l, g0, lds, goff, soff, kxs);
    else if (p.S() <= 8) thin_tile<8, RM>(p, l, g0, lds, goff, soff, kxs);
    else thin_tile<16, RM>(p, l, g0, lds, goff, soff, kxs);
}

__device__ __forceinline__ int64_t uniform64(int64_t v)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// A tile of the launch: its product and its (l, ts, tg) position.
struct TileRef {
    int lo, l, ts, tg;
};

// The decoded tile record (one 16-byte load), or without records the first
// product whose inclusive tile end is past t (binary search) and the
// position from its tile counts.
__device__ __forceinline__ TileRef tile_ref(const tk_prob* __restrict__ probs, const int64_t* __restrict__ tile_end,
                                            const tk_tile* __restrict__ tiles, int nprobs, int64_t t)
{
    TileRef r;
    if (tiles) {
        typedef const int32_t TK_CONST* cint;
        const cint v = (cint)(tiles + t);
        r.lo = __builtin_amdgcn_readfirstlane(v[0]);
        r.l = __builtin_amdgcn_readfirstlane(v[1]);
        r.ts = __builtin_amdgcn_readfirstlane(v[2]);
        r.tg = __builtin_amdgcn_readfirstlane(v[3]);
        return r;
    }
    int a = 0, b = nprobs - 1;
    while (a < b) {
        const int mid = (a + b) >> 1;
        if (tile_end[mid] > t) b = mid; else a = mid + 1;
    }
    r.lo = __builtin_amdgcn_readfirstlane(a);
    const tk_prob TK_CONST* pd = (const tk_prob TK_CONST*)probs + r.lo;
    const int64_t local = t - (r.lo ? tile_end[r.lo - 1] : 0);
    const int64_t per_l = pd->tiles_s * pd->tiles_g;
    const int64_t rem = local % per_l;
    r.l = (int)(local / per_l);
    r.ts = (int)(rem / pd->tiles_g);
    r.tg = (int)(rem % pd->tiles_g);
    return r;
}

// Tiles come from `nq` queues: queue q holds tiles [qbeg[q], qbeg[q + 1])
// (one queue over [0, total) when qbeg is null) and has its own counter,
// counter[32 q].  A workgroup starts with queue blockIdx.x % nq -- with one
// queue per XCD that is its own XCD's, since workgroups are dispatched to the
// XCDs round-robin, so a product's tiles share that XCD's L2 -- and then
// takes what is left in the others, in order, so every workgroup ends after
// every queue is empty.  A workgroup takes the index of its next tile from
// the counter when it starts a tile.  The thin-tile kernel, given decoded
// tile records, also holds the tile after the one it runs (its first atomic
// takes two) and loads that tile's record while the current tile runs.
// MFMA tiles are long enough to take their record at the start, and a
// second held tile would lengthen the tail of their launch.  An index at or past the
// queue's end is dropped; every index below it is run by the workgroup that
// took it.
// KIND (TK_GROUP_KIND_*) restricts the tiles a kernel can run, so the thin-
// only kernel carries neither the MFMA tile's registers nor its LDS.
template <int KIND>
__global__ __launch_bounds__(256, KIND == TK_GROUP_KIND_THIN ? TK_GROUP_WG_PER_CU_THIN : TK_GROUP_WG_PER_CU)
void k_group(const tk_prob* __restrict__ probs, const int64_t* __restrict__ tile_end,
             const tk_tile* __restrict__ tiles, int nprobs, int64_t total, const int64_t* __restrict__ qbeg, int nq,
             unsigned long long* counter)
{
    __shared__ __attribute__((aligned(16))) float lds[KC * YP + SMAX * KC];
    __shared__ int64_t goff[GT];
    __shared__ int64_t soff[2 * SMAX];
    __shared__ int64_t kox[KO], koy[KO];
    __shared__ int64_t eoff[4 * MT];
    __shared__ int64_t next;
    const int home = (int)(blockIdx.x % (unsigned)nq);
    const bool prefetch = KIND == TK_GROUP_KIND_THIN && tiles;
    for (int qi = 0; qi < nq; ++qi) {
        const int q = (home + qi) % nq;
        const int64_t qb = qbeg ? qbeg[q] : 0, qe = qbeg ? qbeg[q + 1] : total;
        if (threadIdx.x == 0) next = qb + (int64_t)atomicAdd(counter + 32 * q, prefetch ? 2ull : 1ull);
        __syncthreads();
        // uniform over the workgroup: said so, so the descriptor below is
        // read with scalar loads into scalar registers
        int64_t t = uniform64(next), t1 = t + 1;
        __syncthreads();
        TileRef cur{};
        if (t < qe) cur = tile_ref(probs, tile_end, tiles, nprobs, t);
        while (t < qe) {
            // the next tile's record, in flight during this tile (records
            // only: the binary search would wait on its loads here)
            TileRef nxt{};
            if (prefetch && t1 < qe) nxt = tile_ref(probs, tile_end, tiles, nprobs, t1);
            // the index of a later tile (the next one, or with prefetching
            // the one after it) is taken now and kept in thread 0's
            // registers until this tile is done
            unsigned long long ahead = 0;
            if (threadIdx.x == 0) ahead = atomicAdd(counter + 32 * q, 1ull);
            // the descriptor through the constant address space: its scalar
            // loads are invariant, so they are re-issued where a field is
            // used instead of holding every field in registers for the tile
            const tk_prob TK_CONST* pd = (const tk_prob TK_CONST*)probs + cur.lo;
            const GView p(pd);
            const int cls = pd->cls;
            const int64_t l = cur.l, ts = cur.ts, tg = cur.tg;   // tile records hold them as int32
            if (KIND == TK_GROUP_KIND_THIN || (KIND == TK_GROUP_KIND_ALL && cls == TK_GROUP_THIN)) {
                if (p.round() == TK_MODE_BF16) thin_dispatch<TK_MODE_BF16>(p, l, tg * GT, lds, goff, soff, kox);
                else if (p.round() == TK_MODE_F16) thin_dispatch<TK_MODE_F16>(p, l, tg * GT, lds, goff, soff, kox);
                else thin_dispatch<0>(p, l, tg * GT, lds, goff, soff, kox);
            } else {
                if (p.round() == TK_MODE_BF16) mfma_tile<TK_MODE_BF16>(p, l, ts * MT, tg * MT, kox, koy, eoff);
                else if (p.round() == TK_MODE_F16) mfma_tile<TK_MODE_F16>(p, l, ts * MT, tg * MT, kox, koy, eoff);
                else mfma_tile<0>(p, l, ts * MT, tg * MT, kox, koy, eoff);
            }
            if (threadIdx.x == 0) next = qb + (int64_t)ahead;
            __syncthreads();
            if (prefetch) {
                t = t1;
                t1 = uniform64(next);
            } else {
                t = uniform64(next);
            }
            __syncthreads();
            if (t < qe) cur = prefetch ? nxt : tile_ref(probs, tile_end, tiles, nprobs, t);
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

extern "C" int tk_launch_group_kind(int kind, const tk_prob* probs, const int64_t* tile_end, const tk_tile* tiles,
                                    int nprobs, int64_t total, const int64_t* qbeg, int nq, unsigned long long* counter,
                                    int nblocks, void* stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (nprobs <= 0 || total <= 0) return 0;
    if (nq < 1 || nq > 64 || (nq > 1 && !qbeg)) return (int)hipErrorInvalidValue;
    const int64_t* q = nq > 1 ? qbeg : (const int64_t*)nullptr;
    hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, counter, nq);
    switch (kind) {
    case TK_GROUP_KIND_THIN:
        hipLaunchKernelGGL(k_group<TK_GROUP_KIND_THIN>, dim3((unsigned)nblocks), dim3(256), 0, st, probs, tile_end,
                           tiles, nprobs, total, q, nq, counter);
        break;
    case TK_GROUP_KIND_MFMA:
        hipLaunchKernelGGL(k_group<TK_GROUP_KIND_MFMA>, dim3((unsigned)nblocks), dim3(256), 0, st, probs, tile_end,
                           tiles, nprobs, total, q, nq, counter);
        break;
    default:
        hipLaunchKernelGGL(k_group<TK_GROUP_KIND_ALL>, dim3((unsigned)nblocks), dim3(256), 0, st, probs, tile_end,
                           tiles, nprobs, total, q, nq, counter);
        break;
    }
    return (int)hipGetLastError();
}

extern "C" int tk_launch_group_q(const tk_prob* probs, const int64_t* tile_end, const tk_tile* tiles, int nprobs,
                                 int64_t total, const int64_t* qbeg, int nq, unsigned long long* counter, int nblocks,
                                 void* stream)
{
    return tk_launch_group_kind(TK_GROUP_KIND_ALL, probs, tile_end, tiles, nprobs, total, qbeg, nq, counter,
                                nblocks, stream);
}

extern "C" int tk_launch_group(const tk_prob* probs, const int64_t* tile_end, const tk_tile* tiles, int nprobs,
                               int64_t total, unsigned long long* counter, int nblocks, void* stream)
{
    return tk_launch_group_q(probs, tile_end, tiles, nprobs, total, nullptr, 1, counter, nblocks, stream);
}
