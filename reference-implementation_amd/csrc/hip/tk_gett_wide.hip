// k_gett_wide: split-BF16 x3 GETT (FP32 data, x = hi + lo in bfloat16,
// hi*hi + hi*lo + lo*hi on v_mfma_f32_32x32x16_bf16) with 256 x 128
// workgroup tiles for gfx950.
//
// The split form reads each MFMA operand twice as often as a plain 16-bit
// GETT, so with the 64 x 64 wave tiles of k_gett_lp it is bound by LDS.
// Here every wave owns a 128 x 64 block of D (4 x 2 blocks of 32 x 32): per
// MFMA it reads 25 % fewer LDS bytes, and the workgroup reads A and B from
// L2 half as often per flop.  The price is one workgroup per CU (two
// 16-bit planes of a 256 x 32 and a 128 x 32 tile, double-buffered: 96 KB
// of LDS), so the conversion of the next tile and the global loads are
// interleaved with the MFMAs inside each wave (see k_gett_wide).
//
// LDS holds each tile as 8-k groups: 16 bytes (8 consecutive k of one row)
// per slot, slot (row, kg) at kg * ROWS + row with the row's low four bits
// XOR-ed by (row / 16) % 2 and 2 ((kg + row / 32) % 2).  An MFMA operand is
// one 16-byte read; the 16 lanes reading 16 consecutive rows hit 16
// distinct slots.  Within a wave's reads the XOR depends on the lane and on
// whether the 32-row block is odd, so a wave reads each operand through two
// base registers and immediate offsets.  The load mappings below put 16
// distinct slots under every 16 lanes' stores (32 lanes' for 8-byte stores).
//
// The products of each step run in the same order as k_gett_lp's, so both
// kernels give the same bits.
#include "tk_device.hpp"

#include <utility>

namespace tk {

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));

constexpr int WM = 256, WK = 32;

__device__ __forceinline__ f32x16 wmfma(u32x4 b, u32x4 a, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, b), __builtin_bit_cast(bf16x8, a), c,
                                                   0, 0, 0);
}

// 16-byte LDS slot of (row, 8-k group kg) in a tile of ROWS rows
template <int ROWS>
__device__ __forceinline__ int wslot(int row, int kg)
{
    return kg * ROWS + (row & ~15) + ((row & 15) ^ ((row >> 4) & 1) ^ (((kg ^ (row >> 5)) & 1) << 1));
}

// Two values (consecutive k of one row) -> hi = bf16(x) and lo = bf16(x - hi)
// (x - hi is exact), one packed word each; the hi values come back exactly
// from the packed bits.
__device__ __forceinline__ void wpack2(float x0, float x1, uint32_t& hi, uint32_t& lo)
{
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const uint32_t h = __builtin_bit_cast(uint32_t, bf16x2{(__bf16)x0, (__bf16)x1});
    hi = h;
    const float b0 = __uint_as_float(h << 16), b1 = __uint_as_float(h & 0xffff0000u);
    lo = __builtin_bit_cast(uint32_t, bf16x2{(__bf16)(x0 - b0), (__bf16)(x1 - b1)});
}

// One ROWS x 32 operand tile, read by 256 lanes: the rows' offsets (fixed
// for the tile), the K-table entries of the next tile to load, and STAGES
// register stages of raw FP32 data.  Lane t of the workgroup:
//   CONTIG_ROW, 256 rows: rows 4 i .. 4 i + 3 with i = t % 64 at the 8 k of
//     group kg = t / 64, one float4 (4 rows) per k; the wave's K-table
//     reads are uniform (scalar loads).
//   CONTIG_ROW, 128 rows: rows 4 i .. + 3 with i = t % 16 + 16 ((t / 32) % 2)
//     at the 4 k of quad q = (t / 16) % 2 + 2 (t / 64), one float4 per k,
//     8-byte stores.
//   CONTIG_K: rows t % 16 + 16 (t / 64) + 64 s at the 8 k of group
//     kg = (t / 16) % 4, two float4 along k per row.
template <int LMODE, int ROWS, int STAGES>
struct WOp {
    static constexpr int NV = ROWS / 32;                                  // float4 per lane and stage
    static constexpr int NR = LMODE == TK_CONTIG_ROW ? 1 : ROWS / 64;     // row offsets held
    static constexpr int NK = LMODE == TK_CONTIG_ROW ? NV : 2;            // K-table entries per tile
    float4 v[STAGES][NV];
    int64_t roff[NR];
    int64_t kv[NK];

    // CONTIG_ROW: first of the lane's four rows; its k group (256 rows) or
    // k quad (128 rows)
    __device__ __forceinline__ static int ri()
    {
        const int t = threadIdx.x;
        return 4 * (ROWS == 256 ? (t & 63) : (t & 15) + 16 * ((t >> 5) & 1));
    }
    __device__ __forceinline__ static int rq()
    {
        const int t = threadIdx.x;
        if (ROWS == 256) return __builtin_amdgcn_readfirstlane(t >> 6);
        return ((t >> 4) & 1) + 2 * (t >> 6);
    }
    // CONTIG_K: row of row offset s; the lane's k group
    __device__ __forceinline__ static int rk(int s) { return (threadIdx.x & 15) + 16 * (threadIdx.x >> 6) + 64 * s; }
    __device__ __forceinline__ static int kgk() { return (threadIdx.x >> 4) & 3; }
    // tile-relative k of K-table entry e
    __device__ __forceinline__ static int kidx(int e)
    {
        if (LMODE == TK_CONTIG_K) return 8 * kgk() + 4 * e;
        return (ROWS == 256 ? 8 : 4) * rq() + e;
    }

    __device__ __forceinline__ void rows(const int64_t* rt, int64_t r0, int64_t Rn)
    {
        if (LMODE == TK_CONTIG_ROW) {
            const int64_t r = r0 + ri();
            roff[0] = r < Rn ? rt[r] : 0;
        } else {
#pragma unroll
            for (int s = 0; s < NR; ++s) {
                const int64_t r = r0 + rk(s);
                roff[s] = r < Rn ? rt[r] : 0;
            }
        }
    }

    template <bool CHECK>
    __device__ __forceinline__ void ktab(const int64_t* kt, int64_t k0, int64_t Kn)
    {
        // 256-row CONTIG_ROW: the index is uniform over the wave, read
        // through the constant address space into scalar registers
        typedef const __attribute__((address_space(4))) int64_t* ctab;
#pragma unroll
        for (int e = 0; e < NK; ++e) {
            const int64_t k = k0 + kidx(e);
            if (LMODE == TK_CONTIG_ROW && ROWS == 256) kv[e] = (!CHECK || k < Kn) ? ((ctab)kt)[k] : 0;
            else kv[e] = (!CHECK || k < Kn) ? kt[k] : 0;
        }
    }

    // raw load e (< NV) of the tile at k0 (whose K-table entries are in
    // kv) into stage s.  CHECK: rows past Rn read row offset 0 and k past
    // Kn read K-table entry 0 (ktab), so every 16-byte load reads an aligned
    // group whose first element is an element of the operand; the values
    // outside the tile's valid rows and k are then replaced by zero.
    template <bool CHECK>
    __device__ __forceinline__ void load(int s, int e, const float* base, int64_t r0, int64_t k0, int64_t Rn,
                                         int64_t Kn)
    {
        if (LMODE == TK_CONTIG_ROW) {
            const int64_t r = r0 + ri();
            float4 q = *reinterpret_cast<const float4*>(base + roff[0] + kv[e]);
            if (CHECK) {
                const bool kin = k0 + kidx(e) < Kn;
                q.x = kin && r < Rn ? q.x : 0.0f;
                q.y = kin && r + 1 < Rn ? q.y : 0.0f;
                q.z = kin && r + 2 < Rn ? q.z : 0.0f;
                q.w = kin && r + 3 < Rn ? q.w : 0.0f;
            }
            v[s][e] = q;
        } else {
            const int si = e >> 1, h = e & 1;
            const bool rin = r0 + rk(si) < Rn;
            float4 q = *reinterpret_cast<const float4*>(base + roff[si] + kv[h]);
            if (CHECK) {
                const int64_t k = k0 + kidx(h);
                q.x = rin && k < Kn ? q.x : 0.0f;
                q.y = rin && k + 1 < Kn ? q.y : 0.0f;
                q.z = rin && k + 2 < Kn ? q.z : 0.0f;
                q.w = rin && k + 3 < Kn ? q.w : 0.0f;
            }
            v[s][e] = q;
        }
    }
    template <bool CHECK>
    __device__ __forceinline__ void load_all(int s, const float* base, int64_t r0, int64_t k0, int64_t Rn, int64_t Kn)
    {
#pragma unroll
        for (int e = 0; e < NV; ++e) load<CHECK>(s, e, base, r0, k0, Rn, Kn);
    }

    // Conversion units of a stage: 8 values of one row (one 16-byte slot),
    // or for the 128-row CONTIG_ROW tile 4 values (half a slot); a unit is
    // converted pair by pair (NP pairs) into ph / pl and then stored.
    static constexpr int NU = LMODE == TK_CONTIG_ROW ? 4 : NR;
    static constexpr int NP = LMODE == TK_CONTIG_ROW && ROWS == 128 ? 2 : 4;
    uint32_t ph[4], pl[4];

    // LDS slot of unit u within a plane (8-byte half index for the 128-row
    // CONTIG_ROW tile: slot * 2 + half)
    __device__ __forceinline__ static int unit_slot(int u)
    {
        if (LMODE == TK_CONTIG_ROW && ROWS == 256) return wslot<ROWS>(ri() + u, rq());
        if (LMODE == TK_CONTIG_ROW) return 2 * wslot<ROWS>(ri() + u, rq() >> 1) + (rq() & 1);
        return wslot<ROWS>(rk(u), kgk());
    }

    // pair p of unit u of stage s (values of consecutive k 2p, 2p + 1)
    __device__ __forceinline__ void pair(int s, int u, int p)
    {
        float x0, x1;
        if (LMODE == TK_CONTIG_ROW) {
            x0 = v[s][2 * p][u];
            x1 = v[s][2 * p + 1][u];
        } else {
            x0 = v[s][2 * u + (p >> 1)][2 * (p & 1)];
            x1 = v[s][2 * u + (p >> 1)][2 * (p & 1) + 1];
        }
        wpack2(x0, x1, ph[p], pl[p]);
    }

    // the converted unit into plane 0 (hi) and plane 1 (lo) at `slot`;
    // `stride` is the distance between the planes in 16-byte units
    __device__ __forceinline__ void put(u32x4* plane, int stride, int slot) const
    {
        if (NP == 2) {
            reinterpret_cast<u32x2*>(plane)[slot] = u32x2{ph[0], ph[1]};
            reinterpret_cast<u32x2*>(plane + stride)[slot] = u32x2{pl[0], pl[1]};
        } else {
            plane[slot] = u32x4{ph[0], ph[1], ph[2], ph[3]};
            plane[stride + slot] = u32x4{pl[0], pl[1], pl[2], pl[3]};
        }
    }
};

// Slot of the MFMA operand of 32-row block `blk` (block rows blk * 32 ..
// + 31 of the tile) at k step kk for this lane, as base[blk % 2] + the
// block's and k step's immediate: lane (il, kl) reads row blk * 32 + il
// (within its wave's rows `w0`) at k group kk / 8 + kl.
template <int ROWS>
__device__ __forceinline__ int read_base(int w0, int parity, int il, int kl)
{
    return kl * ROWS + w0 + (il & 16) + ((il & 15) ^ ((il >> 4) & 1) ^ (((kl ^ parity) & 1) << 1));
}

template <int N> struct IC {
    static constexpr int value = N;
};

// f(IC<0>{}), f(IC<1>{}), ..., in order: a slot index usable as a constant
template <class F, int... I>
__device__ __forceinline__ void for_each_slot(F&& f, std::integer_sequence<int, I...>)
{
    (f(IC<I>{}), ...);
}

// D tile of 256 x WN = 2 x 2 waves of 128 x WN / 2 (4 x FN MFMA blocks of
// 32 x 32, FN = WN / 64; launched with WN = 128: at WN = 256 the 256
// accumulators per lane fill the accumulation registers and spill).  Tile kt's K step runs the MFMAs of LDS
// buffer kt % 2 while tile kt + 1's raw data (one register stage per
// operand, loaded during the previous step) is converted into LDS buffer
// (kt + 1) % 2 and tile kt + 2 is loaded into the freed registers.  The step
// is 6 NB MFMA slots (three products of the split pieces for each of two
// 16-k steps, NB = 4 FN blocks each) fenced by sched_barrier: a wave cannot
// issue while its next MFMA waits for the matrix core, so each slot pairs
// one MFMA with a piece of work sized to run beside it: B's conversion a
// pair of values per slot, then B's loads one per slot beside A's
// conversion, the second k step's operand reads once the first's lo pieces
// are used up, A's loads after its conversion, then the next K-table
// entries.  A load has about two thirds of a step to arrive.  Interior tiles
// run steps in pairs without branches; edge tiles and the last steps run
// with bounds checks.
template <int AL, int BL, int WN>
__global__ __launch_bounds__(256, 1) void k_gett_wide(tk_gett g)
{
    constexpr int P = 2;                               // 16-bit planes per operand
    constexpr int FM = 4, FN = WN / 64, S = 32;
    constexpr int AU = 4 * WM, BU = 4 * WN;            // 16-byte slots of an A / B plane
    // [buffer][A plane 0, A plane 1, B plane 0, B plane 1]
    __shared__ u32x4 lds[2][P * (AU + BU)];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int il = lane & 31, kl = lane >> 5;

    const int64_t ntm = (g.M + WM - 1) / WM, ntn = (g.N + WN - 1) / WN;
    const int64_t tile = xcd_tile(blockIdx.x, ntm * ntn);
    const int64_t m0 = (tile / ntn) * WM, n0 = (tile % ntn) * WN;
    const bool interior = m0 + WM <= g.M && n0 + WN <= g.N;
    const int64_t nk = (g.K + WK - 1) / WK, nk_full = g.K / WK;

    using TA = WOp<AL, WM, 1>;
    using TB = WOp<BL, WN, 1>;
    constexpr int NUA = TA::NU, NUB = TB::NU;         // conversion units per tile
    constexpr int PA = NUA * TA::NP, PB = NUB * TB::NP;   // conversion pairs per tile
    // the slot plan: B pairs [0, PB), B loads [PB, PB + NV_B), A pairs
    // [PB, PB + PA), A loads [LA, LA + NV_A), the K-table after them, the
    // second k step's operand reads two per slot from RD (the first k step's
    // lo pieces are used up by then)
    constexpr int NB = FN * FM, NS = 6 * NB, LA = PB + PA, RD = 2 * NB, NRD = 2 * (FN + FM);
    static_assert(LA + TA::NV < NS && PB + TB::NV <= LA + TA::NV && RD + NRD / 2 <= NS, "slot plan");
    TA ta;
    TB tb;
    ta.rows(g.mA, m0, g.M);
    tb.rows(g.nB, n0, g.N);
    int sa[NUA], sb[NUB];
#pragma unroll
    for (int u = 0; u < NUA; ++u) sa[u] = TA::unit_slot(u);
#pragma unroll
    for (int u = 0; u < NUB; ++u) sb[u] = TB::unit_slot(u);
    const int ra[2] = {read_base<WM>(wm * FM * S, 0, il, kl), read_base<WM>(wm * FM * S, 1, il, kl)};
    const int rb[2] = {read_base<WN>(wn * FN * S, 0, il, kl), read_base<WN>(wn * FN * S, 1, il, kl)};

    for (int64_t l = blockIdx.y; l < g.L; l += gridDim.y) {
        const float* Ab = (const float*)g.A + g.lA[l];
        const float* Bb = (const float*)g.B + g.lB[l];
        f32x16 acc[FN][FM];
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
#pragma unroll
            for (int fm = 0; fm < FM; ++fm) acc[fn][fm] = f32x16{};

        auto ktab = [&](auto CHK, int64_t t) __attribute__((always_inline)) {
            ta.template ktab<decltype(CHK)::value>(g.kA, t * WK, g.K);
            tb.template ktab<decltype(CHK)::value>(g.kB, t * WK, g.K);
        };
        // pair i of A's (B's) conversion into LDS buffer buf; the unit is
        // stored after its last pair
        auto a_pair = [&](int buf, int i) __attribute__((always_inline)) {
            const int u = i / TA::NP, p = i % TA::NP;
            ta.pair(0, u, p);
            if (p == TA::NP - 1) ta.put(&lds[buf][0], AU, sa[u]);
        };
        auto b_pair = [&](int buf, int i) __attribute__((always_inline)) {
            const int u = i / TB::NP, p = i % TB::NP;
            tb.pair(0, u, p);
            if (p == TB::NP - 1) tb.put(&lds[buf][P * AU], BU, sb[u]);
        };
        // operands of k step kk (0 or 16) of LDS buffer `cur`: lanes 0-31
        // give k kk .. kk + 7, lanes 32-63 k kk + 8 .. kk + 15
        auto operand_b = [&](int cur, int kk, int q, int fn) __attribute__((always_inline)) {
            return lds[cur][P * AU + q * BU + rb[fn & 1] + (kk >> 3) * WN + fn * S];
        };
        auto operand_a = [&](int cur, int kk, int q, int fm) __attribute__((always_inline)) {
            return lds[cur][q * AU + ra[fm & 1] + (kk >> 3) * WM + fm * S];
        };

        // one K step; PAR = kt % 2; CHECK: bounds-checked loads and guarded
        // conversion
        auto step = [&](auto PAR_, auto CHK, int64_t kt) __attribute__((always_inline)) {
            constexpr int PAR = decltype(PAR_)::value;
            constexpr bool CHECK = decltype(CHK)::value;
            // [k step][plane][block]; the split pieces' products, smallest
            // first: hi(A) lo(B), lo(A) hi(B), hi hi
            u32x4 a[2][P][FM], b[2][P][FN];
            constexpr int PBQ[3] = {1, 0, 0};
            constexpr int PAQ[3] = {0, 1, 0};
            const bool next = !CHECK || kt + 1 < nk, next2 = !CHECK || kt + 2 < nk;
#pragma unroll
            for (int q = 0; q < P; ++q) {
#pragma unroll
                for (int fn = 0; fn < FN; ++fn) b[0][q][fn] = operand_b(PAR, 0, q, fn);
#pragma unroll
                for (int fm = 0; fm < FM; ++fm) a[0][q][fm] = operand_a(PAR, 0, q, fm);
            }
            // read r of the second k step, in the order its products use
            // them: B lo, A hi, B hi, A lo
            auto read2 = [&](auto R) __attribute__((always_inline)) {
                constexpr int r = decltype(R)::value;
                if constexpr (r < FN) b[1][1][r] = operand_b(PAR, 16, 1, r);
                else if constexpr (r < FN + FM) a[1][0][r - FN] = operand_a(PAR, 16, 0, r - FN);
                else if constexpr (r < 2 * FN + FM) b[1][0][r - FN - FM] = operand_b(PAR, 16, 0, r - FN - FM);
                else a[1][1][r - 2 * FN - FM] = operand_a(PAR, 16, 1, r - 2 * FN - FM);
            };
            __builtin_amdgcn_sched_barrier(0);
            for_each_slot([&](auto SLOT) __attribute__((always_inline)) {
                constexpr int slot = decltype(SLOT)::value;
                constexpr int h = slot / (3 * NB), pr = (slot % (3 * NB)) / NB;
                constexpr int fn = (slot % NB) / FM, fm = slot % FM;
                acc[fn][fm] = wmfma(b[h][PBQ[pr]][fn], a[h][PAQ[pr]][fm], acc[fn][fm]);
                if constexpr (slot < PB) {
                    if (next) b_pair(PAR ^ 1, slot);
                } else if constexpr (slot < PB + PA) {
                    if (next) a_pair(PAR ^ 1, slot - PB);
                }
                if constexpr (slot >= PB && slot < PB + TB::NV)
                    if (next2) tb.template load<CHECK>(0, slot - PB, Bb, n0, (kt + 2) * WK, g.N, g.K);
                if constexpr (slot >= RD && slot < RD + NRD / 2) {
                    read2(IC<2 * (slot - RD)>{});
                    read2(IC<2 * (slot - RD) + 1>{});
                }
                if constexpr (slot >= LA && slot < LA + TA::NV)
                    if (next2) ta.template load<CHECK>(0, slot - LA, Ab, m0, (kt + 2) * WK, g.M, g.K);
                if constexpr (slot == LA + TA::NV)
                    if (!CHECK || kt + 3 < nk) ktab(CHK, kt + 3);
                __builtin_amdgcn_sched_barrier(0);
            }, std::make_integer_sequence<int, NS>{});
            __syncthreads();
        };

        // prologue: tile 0 converted into LDS buffer 0, tile 1 in the
        // register stages, tile 2's K-table entries
        int64_t kt = 0;
        if (nk > 0) {
            ktab(IC<1>{}, 0);
            ta.template load_all<true>(0, Ab, m0, 0, g.M, g.K);
            tb.template load_all<true>(0, Bb, n0, 0, g.N, g.K);
#pragma unroll
            for (int i = 0; i < PA; ++i) a_pair(0, i);
#pragma unroll
            for (int i = 0; i < PB; ++i) b_pair(0, i);
            if (nk > 1) {
                ktab(IC<1>{}, 1);
                ta.template load_all<true>(0, Ab, m0, WK, g.M, g.K);
                tb.template load_all<true>(0, Bb, n0, WK, g.N, g.K);
            }
            if (nk > 2) ktab(IC<1>{}, 2);
        }
        __syncthreads();
        if (interior) {
            for (; kt + 4 < nk_full; kt += 2) {
                step(IC<0>{}, IC<0>{}, kt);
                step(IC<1>{}, IC<0>{}, kt + 1);
            }
        }
        for (; kt < nk; ++kt) {
            if (kt & 1) step(IC<1>{}, IC<1>{}, kt);
            else step(IC<0>{}, IC<1>{}, kt);
        }

        // epilogue: lane (il, kl) of wave (wm, wn) holds rows m = ... + il and
        // columns n = ... + 8 (r / 4) + 4 kl + r % 4 of each 32 x 32 block
        const float alpha = (float)g.alpha[0];
        const int64_t dl = g.lD[l], cl = g.beta_zero ? 0 : g.lC[l];
#pragma unroll
        for (int fm = 0; fm < FM; ++fm) {
            const int64_t m = m0 + wm * FM * S + fm * S + il;
            if (m >= g.M) continue;
            const int64_t dm = dl + g.mD[m];
            if (g.beta_zero) {
#pragma unroll
                for (int fn = 0; fn < FN; ++fn)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        const int64_t n = n0 + wn * FN * S + fn * S + 8 * (r >> 2) + 4 * kl + (r & 3);
                        if (n < g.N) ((float*)g.D)[dm + g.nD[n]] = alpha * acc[fn][fm][r];
                    }
                continue;
            }
            // all C reads of this row block before its D writes
            const int64_t cm = cl + g.mC[m];
#pragma unroll
            for (int fn = 0; fn < FN; ++fn) {
                float cv[16];
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int64_t n = n0 + wn * FN * S + fn * S + 8 * (r >> 2) + 4 * kl + (r & 3);
                    cv[r] = n < g.N ? epilogue_c<float, false>(g, cm + g.nC[n]).re : 0.0f;
                }
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int64_t n = n0 + wn * FN * S + fn * S + 8 * (r >> 2) + 4 * kl + (r & 3);
                    if (n < g.N)
                        epilogue_d<float, false>(g, dm + g.nD[n], acc[fn][fm][r], 0.0f,
                                                 Val<float, false>{cv[r], 0.0f});
                }
            }
        }
    }
}

__host__ __device__ inline int64_t wide_tiles(const tk_gett& g, int wn)
{
    return ((g.M + WM - 1) / WM) * ((g.N + wn - 1) / wn) * g.L;
}

template <int AL, int BL>
hipError_t launch(const tk_gett& g, hipStream_t st)
{
    const dim3 grid((unsigned)(wide_tiles(g, 128) / g.L), (unsigned)(g.L < 65535 ? g.L : 65535)), block(256);
    hipLaunchKernelGGL((k_gett_wide<AL, BL, 128>), grid, block, 0, st, g);
    return hipGetLastError();
}

}  // namespace

}  // namespace tk

using namespace tk;

// 256 x 128 tiles for split-BF16 when both operands load along a
// contiguous direction, K is not split and the tiles cover every CU.
extern "C" int tk_gett_wide_fits(const tk_gett* g)
{
    if (g->ctype != TK_F32 || g->mode != TK_MODE_SPLIT_BF16 || !g->a_vec || !g->b_vec || g->ksplit > 1) return 0;
    if (g->a_load == TK_GATHER || g->b_load == TK_GATHER) return 0;
    return wide_tiles(*g, 128) >= (int64_t)(g->cus > 0 ? g->cus : 256);
}

extern "C" int tk_launch_gett_wide(const tk_gett* gp, void* stream)
{
    const tk_gett& g = *gp;
    hipStream_t st = (hipStream_t)stream;
    if (g.a_load == TK_CONTIG_ROW && g.b_load == TK_CONTIG_ROW) return (int)launch<TK_CONTIG_ROW, TK_CONTIG_ROW>(g, st);
    if (g.a_load == TK_CONTIG_ROW && g.b_load == TK_CONTIG_K) return (int)launch<TK_CONTIG_ROW, TK_CONTIG_K>(g, st);
    if (g.a_load == TK_CONTIG_K && g.b_load == TK_CONTIG_ROW) return (int)launch<TK_CONTIG_K, TK_CONTIG_ROW>(g, st);
    if (g.a_load == TK_CONTIG_K && g.b_load == TK_CONTIG_K) return (int)launch<TK_CONTIG_K, TK_CONTIG_K>(g, st);
    return (int)hipErrorInvalidValue;
}
