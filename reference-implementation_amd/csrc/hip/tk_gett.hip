// Strided GETT kernels for gfx950 (MI355X, CDNA4, wave64).
//
// D[l, m, n] = op_D(alpha * sum_k A'[l, m, k] B'[l, k, n] + beta * op_C(C[l, m, n]))
// with every operand addressed through per-group int64 offset tables (tk.h).
//
//  k_simple   one lane per output element, FMA over K: thin and
//             Hadamard-only products where an MFMA tile would be mostly empty.
//  k_splitk   one workgroup per output element, K split over 256 lanes and
//             reduced through LDS: very few outputs with a long K.
//  k_gett_f32 FP32, v_mfma_f32_32x32x2f32 on 128 x 128 x 32 tiles with the
//             load modes as template parameters.
//  k_gett_lp  FP32 data with 16-bit MFMA inputs (BF16, F16, split-BF16):
//             converted once on the way into 16-bit K-major LDS tiles,
//             v_mfma_f32_32x32x16_{bf16,f16}.
//             Both run split-K (partial sums + k_splitk_reduce) when their
//             128 x 128 tiles alone cannot fill the GPU.
//  k_mfma     LDS-tiled MFMA GETT, 256 lanes = 2x2 waves, double-buffered
//             LDS, register prefetch of the next K tile, XCD-aware tile order.
//             FP64: v_mfma_f64_16x16x4f64; complex: four real MFMAs per step
//             on split re/im LDS planes.  Large or 4x smaller tiles by how
//             many workgroups the problem gives.
//             The MFMA's first operand is the B tile and its second the A
//             tile, so output lanes run along M, the unit-stride direction of
//             D when the host has oriented the problem that way.
#include "tk_device.hpp"

namespace tk {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// scalar multiply-accumulate with the input modes
// ---------------------------------------------------------------------------
template <typename R, bool CPLX>
__device__ __forceinline__ void mac(int mode, Val<R, CPLX> a, Val<R, CPLX> b, R& sr, R& si)
{
    if (mode == TK_MODE_SPLIT_BF16) {
        // x = hi + lo with hi = bf16(x), lo = bf16(x - hi); three products
        auto split = [](R x, R& hi, R& lo) { hi = (R)round_bf16((float)x); lo = (R)round_bf16((float)(x - hi)); };
        R ah, al, bh, bl;
        split(a.re, ah, al); split(b.re, bh, bl);
        R rr = ah * bh + (ah * bl + al * bh);
        if (!CPLX) { sr += rr; return; }
        R aih, ail, bih, bil;
        split(a.im, aih, ail); split(b.im, bih, bil);
        R ii = aih * bih + (aih * bil + ail * bih);
        R ri = ah * bih + (ah * bil + al * bih);
        R ir = aih * bh + (aih * bl + ail * bh);
        sr += rr - ii; si += ri + ir;
        return;
    }
    if (mode == TK_MODE_SPLIT6_BF16) {
        // x = h + m + l, each a bfloat16 of the remainder; six products
        auto split = [](R x, R& h, R& m, R& l) {
            h = (R)round_bf16((float)x);
            m = (R)round_bf16((float)(x - h));
            l = (R)round_bf16((float)(x - h - m));
        };
        auto prod = [](R ah, R am, R al, R bh, R bm, R bl) {
            return ah * bh + ((ah * bm + am * bh) + ((ah * bl + al * bh) + am * bm));
        };
        R ah, am, al, bh, bm, bl;
        split(a.re, ah, am, al); split(b.re, bh, bm, bl);
        const R rr = prod(ah, am, al, bh, bm, bl);
        if (!CPLX) { sr += rr; return; }
        R aih, aim, ail, bih, bim, bil;
        split(a.im, aih, aim, ail); split(b.im, bih, bim, bil);
        const R ii = prod(aih, aim, ail, bih, bim, bil);
        const R ri = prod(ah, am, al, bih, bim, bil);
        const R ir = prod(aih, aim, ail, bh, bm, bl);
        sr += rr - ii; si += ri + ir;
        return;
    }
    if (mode == TK_MODE_BF16 || mode == TK_MODE_F16) {
        a.re = round_mode(a.re, mode); b.re = round_mode(b.re, mode);
        if (CPLX) { a.im = round_mode(a.im, mode); b.im = round_mode(b.im, mode); }
    }
    if (CPLX) {
        sr += a.re * b.re - a.im * b.im;
        si += a.re * b.im + a.im * b.re;
    } else {
        sr += a.re * b.re;
    }
}

// ---------------------------------------------------------------------------
// k_simple
// ---------------------------------------------------------------------------
template <typename R, bool CPLX>
__global__ __launch_bounds__(256) void k_simple(tk_gett g)
{
    const int64_t total = g.M * g.N * g.L;
    for (int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x; o < total; o += (int64_t)gridDim.x * 256) {
        const int64_t m = o % g.M, t = o / g.M, n = t % g.N, l = t / g.N;
        const int64_t ao = g.lA[l] + g.mA[m], bo = g.lB[l] + g.nB[n];
        R sr = 0, si = 0;
        for (int64_t k = 0; k < g.K; ++k) {
            Val<R, CPLX> a = Elem<R, CPLX>::ld(g.A, ao + g.kA[k]);
            Val<R, CPLX> b = Elem<R, CPLX>::ld(g.B, bo + g.kB[k]);
            if (CPLX && g.conjA) a.im = -a.im;
            if (CPLX && g.conjB) b.im = -b.im;
            mac<R, CPLX>(g.mode, a, b, sr, si);
        }
        epilogue<R, CPLX>(g, l, m, n, sr, si);
    }
}

// ---------------------------------------------------------------------------
// k_splitk: grid.x = number of outputs
// ---------------------------------------------------------------------------
template <typename R, bool CPLX>
__global__ __launch_bounds__(256) void k_splitk(tk_gett g)
{
    __shared__ R red[2][256];
    const int64_t o = blockIdx.x;
    const int64_t m = o % g.M, t = o / g.M, n = t % g.N, l = t / g.N;
    const int64_t ao = g.lA[l] + g.mA[m], bo = g.lB[l] + g.nB[n];
    R sr = 0, si = 0;
    for (int64_t k = threadIdx.x; k < g.K; k += 256) {
        Val<R, CPLX> a = Elem<R, CPLX>::ld(g.A, ao + g.kA[k]);
        Val<R, CPLX> b = Elem<R, CPLX>::ld(g.B, bo + g.kB[k]);
        if (CPLX && g.conjA) a.im = -a.im;
        if (CPLX && g.conjB) b.im = -b.im;
        mac<R, CPLX>(g.mode, a, b, sr, si);
    }
    red[0][threadIdx.x] = sr;
    red[1][threadIdx.x] = si;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) epilogue<R, CPLX>(g, l, m, n, red[0][0], red[1][0]);
}

// ---------------------------------------------------------------------------
// MFMA primitives
// ---------------------------------------------------------------------------
template <typename R, int S> struct Mfma;
template <> struct Mfma<float, 32> {
    typedef f32x16 acc_t;
    static constexpr int NR = 16;
    static __device__ __forceinline__ acc_t run(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
    // row of accumulator register r in lane `lane` (32x32 output block)
    static __device__ __forceinline__ int row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }
};
template <> struct Mfma<float, 16> {
    typedef f32x4 acc_t;
    static constexpr int NR = 4;
    static __device__ __forceinline__ acc_t run(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int r, int lane) { return 4 * (lane >> 4) + r; }
};
template <> struct Mfma<double, 16> {
    typedef f64x4 acc_t;
    static constexpr int NR = 4;
    static __device__ __forceinline__ acc_t run(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // FP64 accumulators are spread with stride 4 over the rows: register r
    // of lane l holds row l / 16 + 4 r (the FP32 16x16x4 form keeps 4
    // consecutive rows per lane instead)
    static __device__ __forceinline__ int row(int r, int lane) { return (lane >> 4) + 4 * r; }
};

// ---------------------------------------------------------------------------
// tile loader: BR rows (M for A, N for B) x BK contracted, into registers
// ---------------------------------------------------------------------------
template <typename R, bool CPLX, int BR, int BK>
struct TileLoader {
    static constexpr int EW = CPLX ? 2 : 1;                 // R words per element
    static constexpr int VE = 16 / (EW * (int)sizeof(R));   // elements per 16-byte vector
    static constexpr int NS = (BR * BK) / (256 * VE);        // vector slots per lane
    static_assert(NS * 256 * VE == BR * BK, "tile must split evenly over 256 lanes");

    R v[NS][VE * EW];

    // mode: TK_GATHER / TK_CONTIG_ROW / TK_CONTIG_K; vec: 16-byte loads allowed
    __device__ __forceinline__ void load(const R* base, const int64_t* rt, const int64_t* kt, int64_t r0,
                                         int64_t k0, int64_t Rn, int64_t Kn, int mode, int vec)
    {
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            const int s = threadIdx.x + 256 * i;
            if (mode == TK_CONTIG_K) {
                const int kk = (s % (BK / VE)) * VE, rr = s / (BK / VE);
                const int64_t r = r0 + rr, k = k0 + kk;
                if (r < Rn && k + VE <= Kn && vec) {
                    const R* p = base + EW * (rt[r] + kt[k]);
                    load16(p, v[i]);
                } else {
#pragma unroll
                    for (int j = 0; j < VE; ++j) {
                        const bool in = r < Rn && k + j < Kn;
                        const int64_t off = in ? EW * (rt[r] + kt[k + j]) : 0;
#pragma unroll
                        for (int w = 0; w < EW; ++w) v[i][j * EW + w] = in ? base[off + w] : (R)0;
                    }
                }
            } else {
                const int rr = (s % (BR / VE)) * VE, kk = s / (BR / VE);
                const int64_t r = r0 + rr, k = k0 + kk;
                if (mode == TK_CONTIG_ROW && vec && k < Kn && r + VE <= Rn) {
                    const R* p = base + EW * (rt[r] + kt[k]);
                    load16(p, v[i]);
                } else {
#pragma unroll
                    for (int j = 0; j < VE; ++j) {
                        const bool in = r + j < Rn && k < Kn;
                        const int64_t off = in ? EW * (rt[r + j] + kt[k]) : 0;
#pragma unroll
                        for (int w = 0; w < EW; ++w) v[i][j * EW + w] = in ? base[off + w] : (R)0;
                    }
                }
            }
        }
    }

    static __device__ __forceinline__ void load16(const R* p, R* out)
    {
        const float4 q = *reinterpret_cast<const float4*>(p);
        R tmp[16 / sizeof(R)];
        __builtin_memcpy(tmp, &q, 16);
#pragma unroll
        for (int w = 0; w < (int)(16 / sizeof(R)); ++w) out[w] = tmp[w];
    }

    // LDS image: planes [EW][BK][BR + PAD], row index fastest
    template <int LD>
    __device__ __forceinline__ void store(R* lds, int mode, int conj)
    {
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            const int s = threadIdx.x + 256 * i;
            int rr, kk;
            if (mode == TK_CONTIG_K) { kk = (s % (BK / VE)) * VE; rr = s / (BK / VE); }
            else { rr = (s % (BR / VE)) * VE; kk = s / (BR / VE); }
#pragma unroll
            for (int j = 0; j < VE; ++j) {
                const int r = mode == TK_CONTIG_K ? rr : rr + j;
                const int k = mode == TK_CONTIG_K ? kk + j : kk;
                lds[k * LD + r] = v[i][j * EW];
                if (CPLX) {
                    const R im = v[i][j * EW + 1];
                    lds[BK * LD + k * LD + r] = conj ? -im : im;
                }
            }
        }
    }
};

// ---------------------------------------------------------------------------
// k_mfma
// ---------------------------------------------------------------------------
template <typename R, bool CPLX, int S, int FM, int FN, int BK>
__global__ __launch_bounds__(256) void k_mfma(tk_gett g)
{
    typedef Mfma<R, S> MF;
    typedef typename MF::acc_t acc_t;
    constexpr int BM = 2 * FM * S, BN = 2 * FN * S;
    constexpr int P = CPLX ? 2 : 1;
    constexpr int PAD = 4;
    constexpr int LDA = BM + PAD, LDB = BN + PAD;
    constexpr int KSTEP = 64 / S;
    __shared__ R As[2][P * BK * LDA];
    __shared__ R Bs[2][P * BK * LDB];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int il = lane % S, kl = lane / S;

    const int64_t ntm = (g.M + BM - 1) / BM, ntn = (g.N + BN - 1) / BN;
    const int64_t tile = xcd_tile(blockIdx.x, ntm * ntn);
    const int64_t m0 = (tile / ntn) * BM, n0 = (tile % ntn) * BN;

    for (int64_t l = blockIdx.y; l < g.L; l += gridDim.y) {
        const R* Ab = (const R*)g.A + (CPLX ? 2 : 1) * g.lA[l];
        const R* Bb = (const R*)g.B + (CPLX ? 2 : 1) * g.lB[l];

        acc_t acc[FN][FM], acci[FN][FM];
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
#pragma unroll
            for (int fm = 0; fm < FM; ++fm) {
                acc[fn][fm] = acc_t{};
                acci[fn][fm] = acc_t{};
            }

        TileLoader<R, CPLX, BM, BK> la;
        TileLoader<R, CPLX, BN, BK> lb;
        const int64_t nk = (g.K + BK - 1) / BK;
        if (nk > 0) {
            la.load(Ab, g.mA, g.kA, m0, 0, g.M, g.K, g.a_load, g.a_vec);
            lb.load(Bb, g.nB, g.kB, n0, 0, g.N, g.K, g.b_load, g.b_vec);
            la.template store<LDA>(As[0], g.a_load, g.conjA);
            lb.template store<LDB>(Bs[0], g.b_load, g.conjB);
        }
        __syncthreads();
        for (int64_t kt = 0; kt < nk; ++kt) {
            const int cur = kt & 1;
            if (kt + 1 < nk) {
                la.load(Ab, g.mA, g.kA, m0, (kt + 1) * BK, g.M, g.K, g.a_load, g.a_vec);
                lb.load(Bb, g.nB, g.kB, n0, (kt + 1) * BK, g.N, g.K, g.b_load, g.b_vec);
            }
            const R* as = As[cur];
            const R* bs = Bs[cur];
#pragma unroll
            for (int kk = 0; kk < BK; kk += KSTEP) {
                R a[FM], b[FN], ai[FM], bi[FN];
#pragma unroll
                for (int fm = 0; fm < FM; ++fm) {
                    a[fm] = as[(kk + kl) * LDA + wm * FM * S + fm * S + il];
                    if (CPLX) ai[fm] = as[BK * LDA + (kk + kl) * LDA + wm * FM * S + fm * S + il];
                }
#pragma unroll
                for (int fn = 0; fn < FN; ++fn) {
                    b[fn] = bs[(kk + kl) * LDB + wn * FN * S + fn * S + il];
                    if (CPLX) bi[fn] = bs[BK * LDB + (kk + kl) * LDB + wn * FN * S + fn * S + il];
                }
#pragma unroll
                for (int fn = 0; fn < FN; ++fn)
#pragma unroll
                    for (int fm = 0; fm < FM; ++fm) {
                        acc[fn][fm] = MF::run(b[fn], a[fm], acc[fn][fm]);
                        if (CPLX) {
                            acc[fn][fm] = MF::run(-bi[fn], ai[fm], acc[fn][fm]);
                            acci[fn][fm] = MF::run(bi[fn], a[fm], acci[fn][fm]);
                            acci[fn][fm] = MF::run(b[fn], ai[fm], acci[fn][fm]);
                        }
                    }
            }
            if (kt + 1 < nk) {
                la.template store<LDA>(As[cur ^ 1], g.a_load, g.conjA);
                lb.template store<LDB>(Bs[cur ^ 1], g.b_load, g.conjB);
            }
            __syncthreads();
        }

        // table offsets hoisted: per batch index, per row m (fixed per lane
        // and fm), per column n (fixed per fn and register)
        const int64_t lc = g.beta_zero ? 0 : g.lC[l], ld = g.lD[l];
        int64_t nc[FN][MF::NR], nd[FN][MF::NR];
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
#pragma unroll
            for (int r = 0; r < MF::NR; ++r) {
                const int64_t n = n0 + wn * FN * S + fn * S + MF::row(r, lane);
                const bool in = n < g.N;
                nd[fn][r] = in ? g.nD[n] : 0;
                nc[fn][r] = (in && !g.beta_zero) ? g.nC[n] : 0;
            }
#pragma unroll
        for (int fm = 0; fm < FM; ++fm) {
            const int64_t m = m0 + wm * FM * S + fm * S + il;
            if (m >= g.M) continue;
            const int64_t mc = lc + (g.beta_zero ? 0 : g.mC[m]), md = ld + g.mD[m];
            Val<R, CPLX> cv[FN][MF::NR];
#pragma unroll
            for (int fn = 0; fn < FN; ++fn)
#pragma unroll
                for (int r = 0; r < MF::NR; ++r) {
                    const int64_t n = n0 + wn * FN * S + fn * S + MF::row(r, lane);
                    cv[fn][r] = n < g.N ? epilogue_c<R, CPLX>(g, mc + nc[fn][r]) : Val<R, CPLX>{(R)0, (R)0};
                }
#pragma unroll
            for (int fn = 0; fn < FN; ++fn)
#pragma unroll
                for (int r = 0; r < MF::NR; ++r) {
                    const int64_t n = n0 + wn * FN * S + fn * S + MF::row(r, lane);
                    if (n < g.N)
                        epilogue_d<R, CPLX>(g, md + nd[fn][r], acc[fn][fm][r], CPLX ? acci[fn][fm][r] : (R)0,
                                            cv[fn][r]);
                }
        }
    }
}

// ---------------------------------------------------------------------------
// 16-bit MFMA (gfx950 v_mfma_f32_32x32x16_{bf16,f16}): lane l holds row
// l % 32 and k = 8 * (l / 32) .. +7 of each operand; FP32 accumulate.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

template <int MODE>
__device__ __forceinline__ f32x16 mfma16(uint4 b, uint4 a, f32x16 c)
{
    if (MODE == TK_MODE_F16 || MODE == TK_MODE_SPLIT_F16)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, b), __builtin_bit_cast(f16x8, a), c,
                                                      0, 0, 0);
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, b), __builtin_bit_cast(bf16x8, a), c,
                                                   0, 0, 0);
}

// ---------------------------------------------------------------------------
// k_gett_f32: FP32 GETT specialised on how A and B tiles are read.
//   AL, BL: TK_CONTIG_ROW (16-byte loads along M / N), TK_CONTIG_K (16-byte
//   loads along K) or TK_GATHER (per-element table lookups).
// 128 x 128 x 32 tiles, 2 x 2 waves of 64 x 64 (2 x 2 v_mfma_f32_32x32x2f32
// blocks each), LDS double-buffered (2 x 33 KB), one barrier per K tile, the
// next tile's global loads in flight during the current tile's MFMAs.  Row
// offsets of the rows a lane loads are read from the tables once, before the
// K loop; interior tiles load without bounds checks.
// ---------------------------------------------------------------------------
template <int MODE, int BR, int BK>
struct F32Tile {
    static constexpr int NS = (BR * BK) / (256 * 4);   // 16-byte slots per lane
    float4 v[NS];
    int64_t roff[MODE == TK_CONTIG_ROW ? 1 : (MODE == TK_CONTIG_K ? NS : 4)];

    // rows this lane loads (relative to the tile) -- fixed for the whole K loop
    __device__ __forceinline__ static int row_of(int i, int j)
    {
        const int s = threadIdx.x + 256 * i;
        if (MODE == TK_CONTIG_K) return s / (BK / 4);
        return (s % (BR / 4)) * 4 + j;
    }
    __device__ __forceinline__ static int k_of(int i, int j)
    {
        const int s = threadIdx.x + 256 * i;
        if (MODE == TK_CONTIG_K) return (s % (BK / 4)) * 4 + j;
        return s / (BR / 4);
    }

    __device__ __forceinline__ void rows(const int64_t* rt, int64_t r0, int64_t Rn)
    {
        if (MODE == TK_CONTIG_ROW) {
            const int64_t r = r0 + row_of(0, 0);
            roff[0] = r < Rn ? rt[r] : 0;
        } else if (MODE == TK_CONTIG_K) {
#pragma unroll
            for (int i = 0; i < NS; ++i) {
                const int64_t r = r0 + row_of(i, 0);
                roff[i] = r < Rn ? rt[r] : 0;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int64_t r = r0 + row_of(0, j);
                roff[j] = r < Rn ? rt[r] : 0;
            }
        }
    }

    template <bool CHECK>
    __device__ __forceinline__ void load(const float* base, const int64_t* kt, int64_t r0, int64_t k0, int64_t Rn,
                                         int64_t Kn)
    {
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            if (MODE == TK_CONTIG_ROW) {
                const int64_t r = r0 + row_of(i, 0), k = k0 + k_of(i, 0);
                if (!CHECK || (k < Kn && r + 4 <= Rn)) {
                    v[i] = *reinterpret_cast<const float4*>(base + roff[0] + kt[k]);
                } else {
                    float t[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) t[j] = (k < Kn && r + j < Rn) ? base[roff[0] + kt[k] + j] : 0.0f;
                    v[i] = make_float4(t[0], t[1], t[2], t[3]);
                }
            } else if (MODE == TK_CONTIG_K) {
                const int64_t r = r0 + row_of(i, 0), k = k0 + k_of(i, 0);
                if (!CHECK || (r < Rn && k + 4 <= Kn)) {
                    v[i] = *reinterpret_cast<const float4*>(base + roff[i] + kt[k]);
                } else {
                    float t[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) t[j] = (r < Rn && k + j < Kn) ? base[roff[i] + kt[k + j]] : 0.0f;
                    v[i] = make_float4(t[0], t[1], t[2], t[3]);
                }
            } else {
                const int64_t k = k0 + k_of(i, 0);
                float t[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int64_t r = r0 + row_of(i, j);
                    t[j] = (!CHECK || (k < Kn && r < Rn)) ? base[roff[j] + kt[k]] : 0.0f;
                }
                v[i] = make_float4(t[0], t[1], t[2], t[3]);
            }
        }
    }

    template <int LD>
    __device__ __forceinline__ void store(float* lds) const
    {
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            if (MODE == TK_CONTIG_K) {
                const int r = row_of(i, 0), k = k_of(i, 0);
                lds[(k + 0) * LD + r] = v[i].x;
                lds[(k + 1) * LD + r] = v[i].y;
                lds[(k + 2) * LD + r] = v[i].z;
                lds[(k + 3) * LD + r] = v[i].w;
            } else {
                *reinterpret_cast<float4*>(&lds[k_of(i, 0) * LD + row_of(i, 0)]) = v[i];
            }
        }
    }
};

// Epilogue of a 128 x 128 FP32-accumulator tile (2 x 2 waves of 2 x 2
// v_mfma 32x32 blocks): SPLIT stores the raw partial sums for
// k_splitk_reduce, otherwise D = alpha * acc + beta * C (op_C / rounding per
// the plan) through the offset tables.
template <bool SPLIT>
__device__ __forceinline__ void tile_epilogue(const tk_gett& g, const f32x16 (&acc)[2][2], int64_t l, int64_t m0,
                                              int64_t n0, int64_t tile, int64_t ntiles, int wave, int lane)
{
    constexpr int S = 32, FM = 2, FN = 2;
    const int wm = wave & 1, wn = wave >> 1;
    const int il = lane & 31, kl = lane >> 5;
    if (SPLIT) {
        // raw partial sums in accumulator order: per (split, l, tile,
        // wave) 16 float4 slots x 64 lanes, one 1 KB store per slot
        float4* part = reinterpret_cast<float4*>(g.partial) +
                       ((((int64_t)blockIdx.z * g.L + l) * ntiles + tile) * 4 + wave) * (16 * 64) + lane;
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
#pragma unroll
            for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    part[((fn * FM + fm) * 4 + q) * 64] =
                        make_float4(acc[fn][fm][4 * q], acc[fn][fm][4 * q + 1], acc[fn][fm][4 * q + 2],
                                    acc[fn][fm][4 * q + 3]);
        return;
    }
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
                    epilogue_d<float, false>(g, dm + g.nD[n], acc[fn][fm][r], 0.0f, Val<float, false>{cv[r], 0.0f});
            }
        }
    }
}

template <int AL, int BL, int MODE, bool SPLIT>
__global__ __launch_bounds__(256, 2) void k_gett_f32(tk_gett g)
{
    static_assert(MODE == TK_MODE_PLAIN, "16-bit input modes run k_gett_lp");
    constexpr int S = 32, FM = 2, FN = 2, BM = 128, BN = 128, BK = 32;
    constexpr int LDA = BM + 4, LDB = BN + 4;
    __shared__ __attribute__((aligned(16))) float As[2][BK * LDA];
    __shared__ __attribute__((aligned(16))) float Bs[2][BK * LDB];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int il = lane & 31, kl = lane >> 5;

    const int64_t ntm = (g.M + BM - 1) / BM, ntn = (g.N + BN - 1) / BN;
    const int64_t tile = xcd_tile(blockIdx.x, ntm * ntn);
    const int64_t m0 = (tile / ntn) * BM, n0 = (tile % ntn) * BN;
    const bool interior = m0 + BM <= g.M && n0 + BN <= g.N;
    // SPLIT: this workgroup covers K in [kb, ke) and writes raw partial sums
    const int64_t kb = SPLIT ? (int64_t)blockIdx.z * g.kchunk : 0;
    const int64_t ke = SPLIT ? (kb + g.kchunk < g.K ? kb + g.kchunk : g.K) : g.K;
    const int64_t klen = ke > kb ? ke - kb : 0;
    const int64_t nk = (klen + BK - 1) / BK, nk_full = klen / BK;

    F32Tile<AL, BM, BK> la;
    F32Tile<BL, BN, BK> lb;
    la.rows(g.mA, m0, g.M);
    lb.rows(g.nB, n0, g.N);

    for (int64_t l = blockIdx.y; l < g.L; l += gridDim.y) {
        const float* Ab = (const float*)g.A + g.lA[l];
        const float* Bb = (const float*)g.B + g.lB[l];
        f32x16 acc[FN][FM];
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
#pragma unroll
            for (int fm = 0; fm < FM; ++fm) acc[fn][fm] = f32x16{};

        auto fetch = [&](int64_t kt) {
            if (interior && kt < nk_full) {
                la.template load<false>(Ab, g.kA, m0, kb + kt * BK, g.M, ke);
                lb.template load<false>(Bb, g.kB, n0, kb + kt * BK, g.N, ke);
            } else {
                la.template load<true>(Ab, g.kA, m0, kb + kt * BK, g.M, ke);
                lb.template load<true>(Bb, g.kB, n0, kb + kt * BK, g.N, ke);
            }
        };
        if (nk > 0) {
            fetch(0);
            la.template store<LDA>(As[0]);
            lb.template store<LDB>(Bs[0]);
        }
        __syncthreads();
        for (int64_t kt = 0; kt < nk; ++kt) {
            const int cur = kt & 1;
            if (kt + 1 < nk) fetch(kt + 1);
            const float* as = As[cur];
            const float* bs = Bs[cur];
            auto mfmas = [&](int k_begin, int k_end) {
#pragma unroll
                for (int kk = k_begin; kk < k_end; kk += 2) {
                    float a[FM], b[FN];
#pragma unroll
                    for (int fm = 0; fm < FM; ++fm) a[fm] = as[(kk + kl) * LDA + wm * FM * S + fm * S + il];
#pragma unroll
                    for (int fn = 0; fn < FN; ++fn) b[fn] = bs[(kk + kl) * LDB + wn * FN * S + fn * S + il];
#pragma unroll
                    for (int fn = 0; fn < FN; ++fn)
#pragma unroll
                        for (int fm = 0; fm < FM; ++fm)
                            acc[fn][fm] = __builtin_amdgcn_mfma_f32_32x32x2f32(b[fn], a[fm], acc[fn][fm], 0, 0, 0);
                }
            };
            // the other LDS buffer has been free since the last barrier: the
            // next tile goes there half-way through this tile's MFMAs, so
            // the stores overlap the MFMA pipe instead of following it.  The
            // MFMA stretches run at raised wave priority, so the SIMD's other
            // wave does its loads and stores in their shadow.
            __builtin_amdgcn_s_setprio(1);
            mfmas(0, BK / 2);
            __builtin_amdgcn_s_setprio(0);
            if (kt + 1 < nk) {
                la.template store<LDA>(As[cur ^ 1]);
                lb.template store<LDB>(Bs[cur ^ 1]);
            }
            __builtin_amdgcn_s_setprio(1);
            mfmas(BK / 2, BK);
            __builtin_amdgcn_s_setprio(0);
            __syncthreads();
        }

        tile_epilogue<SPLIT>(g, acc, l, m0, n0, tile, ntm * ntn, wave, lane);
    }
}

// ---------------------------------------------------------------------------
// k_gett_lp: FP32 data computed with 16-bit MFMA inputs (BF16, F16, split-BF16)
//
// Each element is converted once, on its way into LDS, and LDS holds 16-bit
// K-major tiles: row r of a 128 x 32 tile is 32 consecutive k at pitch LDK,
// so an MFMA operand (8 consecutive k of one row) is one 16-byte read.  The
// 4-k groups of a row are XOR-swizzled by (row / 16) % 4 (times 2, keeping
// the two groups an operand reads adjacent): the 8-byte writes of 16 lanes
// that start 16 rows apart then fall on distinct banks, and the 16-byte
// reads of 8 consecutive rows are conflict-free with or without it.
// Split-BF16 keeps a hi and a lo plane (x = hi + lo, hi = bf16(x),
// lo = bf16(x - hi)) and runs hi*hi + hi*lo + lo*hi.
// Each lane loads one 4-row x 4-k FP32 sub-block per operand per K tile:
// float4 along rows (CONTIG_ROW), along k (CONTIG_K) or 16 table gathers.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lp_swz(int row, int q) { return q ^ (2 * ((row >> 4) & 3)); }

template <int MODE>
__device__ __forceinline__ void lp_pack(const float (&x)[4], uint2& hi, uint2& lo, uint2& lo2)
{
    if (MODE == TK_MODE_F16 || MODE == TK_MODE_SPLIT_F16) {
        typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
        f16x4 h, l;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            h[e] = (_Float16)x[e];                               // round to nearest even
            if (MODE == TK_MODE_SPLIT_F16) l[e] = (_Float16)(x[e] - (float)h[e]);   // exact remainder
        }
        hi = __builtin_bit_cast(uint2, h);
        if (MODE == TK_MODE_SPLIT_F16) lo = __builtin_bit_cast(uint2, l);
    } else {
        typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
        bf16x4 h, l, l2;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            h[e] = (__bf16)x[e];                                 // v_cvt_pk_bf16_f32, round to nearest even
            if (MODE == TK_MODE_SPLIT_BF16 || MODE == TK_MODE_SPLIT6_BF16) {
                const float r = x[e] - (float)h[e];              // exact
                l[e] = (__bf16)r;
                if (MODE == TK_MODE_SPLIT6_BF16) l2[e] = (__bf16)(r - (float)l[e]);
            }
        }
        hi = __builtin_bit_cast(uint2, h);
        if (MODE == TK_MODE_SPLIT_BF16 || MODE == TK_MODE_SPLIT6_BF16) lo = __builtin_bit_cast(uint2, l);
        if (MODE == TK_MODE_SPLIT6_BF16) lo2 = __builtin_bit_cast(uint2, l2);
    }
}

// BK-deep tile of 128 rows; LDS pitch BK + 8 (16-byte aligned rows whose
// 16-byte reads by 8 consecutive rows hit distinct banks).
template <int LMODE, int BK>
struct LpTile {
    static constexpr int BR = 128, LDK = BK + 8;
    static constexpr int NB = BR * BK / (256 * 16);  // 4 x 4 sub-blocks per lane
    float v[NB][4][4];                               // [sub-block][row j][k e]
    int64_t roff[LMODE == TK_CONTIG_ROW ? 1 : 4];    // the lane's 4 rows (shared by its sub-blocks)

    // sub-block b of this lane: rows 4 * gi() .. +3, k 4 * gk(b) .. +3.
    // Lanes adjacent in memory are adjacent in the lane index: along rows
    // (CONTIG_ROW, GATHER) 32 lanes cover the 128 rows, along k (CONTIG_K)
    // 8 lanes cover 32 consecutive k.
    __device__ __forceinline__ static int gi() { return LMODE == TK_CONTIG_K ? threadIdx.x / 8 : threadIdx.x % 32; }
    __device__ __forceinline__ static int gk(int b)
    {
        return LMODE == TK_CONTIG_K ? threadIdx.x % 8 + 8 * b : threadIdx.x / 32 + 8 * b;
    }

    __device__ __forceinline__ void rows(const int64_t* rt, int64_t r0, int64_t Rn)
    {
        const int64_t r = r0 + 4 * gi();
        if (LMODE == TK_CONTIG_ROW) {
            roff[0] = r < Rn ? rt[r] : 0;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) roff[j] = r + j < Rn ? rt[r + j] : 0;
        }
    }

    template <bool CHECK>
    __device__ __forceinline__ void load(const float* base, const int64_t* kt, int64_t r0, int64_t k0, int64_t Rn,
                                         int64_t Kn)
    {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const int64_t r = r0 + 4 * gi(), k = k0 + 4 * gk(b);
            const int64_t* ro = roff;
            if (LMODE == TK_CONTIG_ROW) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int64_t kk = k + e;
                    if (!CHECK || (kk < Kn && r + 4 <= Rn)) {
                        const float4 q = *reinterpret_cast<const float4*>(base + ro[0] + kt[kk]);
                        v[b][0][e] = q.x; v[b][1][e] = q.y; v[b][2][e] = q.z; v[b][3][e] = q.w;
                    } else {
#pragma unroll
                        for (int j = 0; j < 4; ++j)
                            v[b][j][e] = (kk < Kn && r + j < Rn) ? base[ro[0] + kt[kk] + j] : 0.0f;
                    }
                }
            } else if (LMODE == TK_CONTIG_K) {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (!CHECK || (r + j < Rn && k + 4 <= Kn)) {
                        const float4 q = *reinterpret_cast<const float4*>(base + ro[j] + kt[k]);
                        v[b][j][0] = q.x; v[b][j][1] = q.y; v[b][j][2] = q.z; v[b][j][3] = q.w;
                    } else {
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            v[b][j][e] = (r + j < Rn && k + e < Kn) ? base[ro[j] + kt[k + e]] : 0.0f;
                    }
                }
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        v[b][j][e] = (!CHECK || (r + j < Rn && k + e < Kn)) ? base[ro[j] + kt[k + e]] : 0.0f;
            }
        }
    }

    // converted rows into the K-major 16-bit planes (hi; lo for split; the
    // third piece for the three-piece split).  Split-F16 first raises amax to
    // the bits of |x| (integer max of the sign-cleared bits: every NaN and
    // inf lands above every finite value) and multiplies by the scale.
    template <int MODE>
    __device__ __forceinline__ void store(uint16_t* hi, uint16_t* lo, uint16_t* lo2, float scale,
                                          uint32_t& amax) const
    {
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = 4 * gi() + j;
                const int off = row * LDK + 4 * lp_swz(row, gk(b));
                uint2 h, l, l2;
                if (MODE == TK_MODE_SPLIT_F16) {
                    float x[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const uint32_t u = __float_as_uint(v[b][j][e]) & 0x7fffffffu;
                        amax = u > amax ? u : amax;
                        x[e] = v[b][j][e] * scale;
                    }
                    lp_pack<MODE>(x, h, l, l2);
                } else {
                    lp_pack<MODE>(v[b][j], h, l, l2);
                }
                *reinterpret_cast<uint2*>(hi + off) = h;
                if (MODE == TK_MODE_SPLIT_BF16 || MODE == TK_MODE_SPLIT6_BF16 || MODE == TK_MODE_SPLIT_F16)
                    *reinterpret_cast<uint2*>(lo + off) = l;
                if (MODE == TK_MODE_SPLIT6_BF16) *reinterpret_cast<uint2*>(lo2 + off) = l2;
            }
    }
};

// MFMA operand: 8 consecutive k (kk .. kk + 7, kk a multiple of 8) of one row
template <int LDK>
__device__ __forceinline__ uint4 lp_operand(const uint16_t* plane, int row, int kk)
{
    return *reinterpret_cast<const uint4*>(plane + row * LDK + 4 * lp_swz(row, kk >> 2));
}

// One wave per SIMD for the three-piece split, and for split-F16 when
// neither operand loads rows contiguously (four row offsets per operand):
// at two waves those would spill.
template <int AL, int BL, int MODE>
constexpr int lp_waves()
{
    return MODE == TK_MODE_SPLIT6_BF16 ||
                   (MODE == TK_MODE_SPLIT_F16 && AL != TK_CONTIG_ROW && BL != TK_CONTIG_ROW)
               ? 1
               : 2;
}

template <int AL, int BL, int MODE, bool SPLIT>
__global__ __launch_bounds__(256, (lp_waves<AL, BL, MODE>())) void k_gett_lp(tk_gett g)
{
    // K tile: 64 deep (twice the bytes in flight per step) where two
    // workgroups per CU still fit in LDS; the split forms' planes need 32
    constexpr bool S3 = MODE == TK_MODE_SPLIT_BF16 || MODE == TK_MODE_SPLIT_F16, S6 = MODE == TK_MODE_SPLIT6_BF16;
    // The accurate forms carry more bits per product than an FP32 sum over
    // long K keeps, so their MFMA accumulators hold FOLD K tiles (128 k)
    // at a time and are then added into a total: split-F16 plainly (two
    // sum levels of about sqrt(K) terms each; the extra accumulator set
    // still fits two waves per SIMD), the three-piece split-BF16 with
    // compensation (Kahan).
    constexpr bool RANGE = MODE == TK_MODE_SPLIT_F16, TWO_LEVEL = S6 || RANGE;
    constexpr int FOLD = 4;
    uint32_t amax_a = 0, amax_b = 0;
    constexpr int S = 32, FM = 2, FN = 2, BM = 128, BN = 128, BK = (S3 || S6) ? 32 : 64;
    constexpr int LDK = BK + 8;
    constexpr int P = S6 ? 3 : S3 ? 2 : 1;                         // planes per operand
    constexpr int PLANE = BM * LDK;                                 // 16-bit elements per plane
    __shared__ __attribute__((aligned(16))) uint16_t lds[2][2 * P][PLANE];   // [buffer][A planes, B planes]

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int il = lane & 31, kl = lane >> 5;

    const int64_t ntm = (g.M + BM - 1) / BM, ntn = (g.N + BN - 1) / BN;
    const int64_t tile = xcd_tile(blockIdx.x, ntm * ntn);
    const int64_t m0 = (tile / ntn) * BM, n0 = (tile % ntn) * BN;
    const bool interior = m0 + BM <= g.M && n0 + BN <= g.N;
    // SPLIT: this workgroup covers K in [kb, ke) and writes raw partial sums
    const int64_t kb = SPLIT ? (int64_t)blockIdx.z * g.kchunk : 0;
    const int64_t ke = SPLIT ? (kb + g.kchunk < g.K ? kb + g.kchunk : g.K) : g.K;
    const int64_t klen = ke > kb ? ke - kb : 0;
    const int64_t nk = (klen + BK - 1) / BK, nk_full = klen / BK;

    LpTile<AL, BK> la0;
    LpTile<BL, BK> lb0;
    la0.rows(g.mA, m0, g.M);
    lb0.rows(g.nB, n0, g.N);

    for (int64_t l = blockIdx.y; l < g.L; l += gridDim.y) {
        const float* Ab = (const float*)g.A + g.lA[l];
        const float* Bb = (const float*)g.B + g.lB[l];
        f32x16 acc[FN][FM], sum[FN][TWO_LEVEL ? FM : 1], cmp[FN][S6 ? FM : 1];
#pragma unroll
        for (int fn = 0; fn < FN; ++fn)
#pragma unroll
            for (int fm = 0; fm < FM; ++fm) {
                acc[fn][fm] = f32x16{};
                if (TWO_LEVEL) sum[fn][fm] = f32x16{};
                if (S6) cmp[fn][fm] = f32x16{};
            }

        auto fetch = [&](LpTile<AL, BK>& ta, LpTile<BL, BK>& tb, int64_t kt) {
            if (interior && kt < nk_full) {
                ta.template load<false>(Ab, g.kA, m0, kb + kt * BK, g.M, ke);
                tb.template load<false>(Bb, g.kB, n0, kb + kt * BK, g.N, ke);
            } else {
                ta.template load<true>(Ab, g.kA, m0, kb + kt * BK, g.M, ke);
                tb.template load<true>(Bb, g.kB, n0, kb + kt * BK, g.N, ke);
            }
        };
        auto stash = [&](const LpTile<AL, BK>& ta, const LpTile<BL, BK>& tb, int buf) {
            ta.template store<MODE>(lds[buf][0], lds[buf][P > 1 ? 1 : 0], lds[buf][P > 2 ? 2 : 0], g.scale[0],
                                    amax_a);
            tb.template store<MODE>(lds[buf][P], lds[buf][P > 1 ? P + 1 : P], lds[buf][P > 2 ? P + 2 : P],
                                    g.scale[1], amax_b);
        };
        auto compute = [&](int cur) {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 16) {
                uint4 a[P][FM], b[P][FN];
#pragma unroll
                for (int q = 0; q < P; ++q) {
#pragma unroll
                    for (int fm = 0; fm < FM; ++fm)
                        a[q][fm] = lp_operand<LDK>(lds[cur][q], wm * FM * S + fm * S + il, kk + 8 * kl);
#pragma unroll
                    for (int fn = 0; fn < FN; ++fn)
                        b[q][fn] = lp_operand<LDK>(lds[cur][P + q], wn * FN * S + fn * S + il, kk + 8 * kl);
                }
                // products of the split pieces, smallest terms first; each
                // product over all FN x FM blocks before the next, so that
                // back-to-back MFMAs write different accumulators
                constexpr int NPROD = S6 ? 6 : S3 ? 3 : 1;
                constexpr int PB[6] = {1, 2, 0, 1, 0, 0};   // B piece (h = 0, m = 1, l = 2)
                constexpr int PA[6] = {1, 0, 2, 0, 1, 0};   // A piece: m*m, h*l, l*h, h*m, m*h, h*h
#pragma unroll
                for (int pr = 6 - NPROD; pr < 6; ++pr)
#pragma unroll
                    for (int fn = 0; fn < FN; ++fn)
#pragma unroll
                        for (int fm = 0; fm < FM; ++fm)
                            acc[fn][fm] = mfma16<MODE>(b[PB[pr]][fn], a[PA[pr]][fm], acc[fn][fm]);
            }
        };
        // the FP32 sum of the last FOLD K tiles into the total
        auto fold = [&]() {
#pragma unroll
            for (int fn = 0; fn < FN; ++fn)
#pragma unroll
                for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        if (S6) {
                            const float y = acc[fn][fm][r] - cmp[fn][fm][r];
                            const float t = sum[fn][fm][r] + y;
                            cmp[fn][fm][r] = (t - sum[fn][fm][r]) - y;
                            sum[fn][fm][r] = t;
                        } else {
                            sum[fn][fm][r] += acc[fn][fm][r];
                        }
                        acc[fn][fm][r] = 0.0f;
                    }
        };

        if (nk > 0) {
            fetch(la0, lb0, 0);
            stash(la0, lb0, 0);
        }
        __syncthreads();
        for (int64_t kt = 0; kt < nk; ++kt) {
            const int cur = kt & 1;
            if (kt + 1 < nk) fetch(la0, lb0, kt + 1);   // next tile's loads in flight during the MFMAs
            __builtin_amdgcn_s_setprio(1);               // MFMAs ahead of the other wave's conversions
            compute(cur);
            __builtin_amdgcn_s_setprio(0);
            if (kt + 1 < nk) stash(la0, lb0, cur ^ 1);
            if (TWO_LEVEL && ((kt + 1) % FOLD == 0 || kt + 1 == nk)) fold();
            __syncthreads();
        }
        if (TWO_LEVEL) {
#pragma unroll
            for (int fn = 0; fn < FN; ++fn)
#pragma unroll
                for (int fm = 0; fm < FM; ++fm)
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        acc[fn][fm][r] = S6 ? sum[fn][fm][r] - cmp[fn][fm][r] : sum[fn][fm][r];
        }
        tile_epilogue<SPLIT>(g, acc, l, m0, n0, tile, ntm * ntn, wave, lane);
    }
    if (RANGE && g.amax) {
        // the operands' max |x| seen by this wave, one atomic per wave
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t oa = __shfl_xor(amax_a, off, 64), ob = __shfl_xor(amax_b, off, 64);
            amax_a = oa > amax_a ? oa : amax_a;
            amax_b = ob > amax_b ? ob : amax_b;
        }
        if (lane == 0) {
            atomicMax(g.amax, amax_a);
            atomicMax(g.amax + 1, amax_b);
        }
    }
}

// ---------------------------------------------------------------------------
// k_small_f32: FP32 products with few outputs and a moderate K (the 256 x
// 256 outputs of a 16^6 four-index product at K = 256 give 128 x 128 tiles
// only 4 workgroups).
// One workgroup per 16 x 16 output tile, so the grid has M N L / 256
// workgroups; its 4 waves each take a quarter of K with
// v_mfma_f32_16x16x4f32 (lane l supplies row l % 16 at k = l / 16 of a
// 4-deep step).  Operands go from global memory straight into that lane
// layout: a chunk's K-table reads, then its data reads, are all in flight
// together.  The four partial tiles are summed through LDS in wave order
// (deterministic) and each lane writes one output.  One launch, no partial
// buffer and no second kernel.
// ---------------------------------------------------------------------------
template <int CH>
__global__ __launch_bounds__(256) void k_small_f32(tk_gett g)
{
    __shared__ float part[4][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t ntm = (g.M + 15) / 16, ntn = (g.N + 15) / 16;
    const int64_t tile = xcd_tile(blockIdx.x, ntm * ntn);
    const int64_t m0 = (tile / ntn) * 16, n0 = (tile % ntn) * 16;
    const int64_t m = m0 + (lane & 15), n = n0 + (lane & 15);
    const int kq = lane >> 4;
    const bool m_in = m < g.M, n_in = n < g.N;
    const int64_t ma = m_in ? g.mA[m] : 0, nb = n_in ? g.nB[n] : 0;
    // this wave's quarter of K, in 4-deep steps
    const int64_t steps = (g.K + 3) / 4, sq = (steps + 3) / 4;
    const int64_t s_begin = wave * sq, s_end = s_begin + sq < steps ? s_begin + sq : steps;

    // the output this thread writes after the reduction
    const int t = threadIdx.x;
    const int64_t mo = m0 + (t & 15), no = n0 + (t >> 4);
    const bool o_in = mo < g.M && no < g.N;
    const int64_t mi = o_in ? mo : 0, ni = o_in ? no : 0;   // valid table indices for every thread

    for (int64_t l = blockIdx.y; l < g.L; l += gridDim.y) {
        const float* Ab = (const float*)g.A + g.lA[l] + ma;
        const float* Bb = (const float*)g.B + g.lB[l] + nb;
        // the output's C value and D offset, read now: their table and C
        // loads are in flight with the K loop's, so the epilogue waits on none
        const int64_t coff = g.beta_zero ? 0 : g.lC[l] + g.mC[mi] + g.nC[ni];
        const int64_t doff = g.lD[l] + g.mD[mi] + g.nD[ni];
        const Val<float, false> cv = epilogue_c<float, false>(g, coff);
        f32x4 acc = f32x4{};
        for (int64_t s0 = s_begin; s0 < s_end; s0 += CH) {
            int64_t ka[CH], kb[CH];
            bool in[CH];
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                const int64_t k = 4 * (s0 + j) + kq;
                in[j] = s0 + j < s_end && k < g.K;
                const int64_t kk = in[j] ? k : 0;        // table entry 0 is valid whenever K > 0
                ka[j] = g.kA[kk];
                kb[j] = g.kB[kk];
            }
            // unconditional loads at valid addresses (offset 0 when out of
            // range), masked afterwards: no branch separates them
            float a[CH], b[CH];
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                a[j] = Ab[ka[j]];
                b[j] = Bb[kb[j]];
            }
#pragma unroll
            for (int j = 0; j < CH; ++j)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(in[j] && n_in ? b[j] : 0.0f, in[j] && m_in ? a[j] : 0.0f,
                                                            acc, 0, 0, 0);
        }
        // lane holds out[n0 + 4 (lane / 16) + r][m0 + lane % 16]
        __syncthreads();   // the previous batch index's reads of part are done
#pragma unroll
        for (int r = 0; r < 4; ++r) part[wave][(4 * kq + r) * 16 + (lane & 15)] = acc[r];
        __syncthreads();
        const float sum = ((part[0][t] + part[1][t]) + part[2][t]) + part[3][t];
        if (o_in) epilogue_d<float, false>(g, doff, sum, 0.0f, cv);
    }
}

// Sums the split-K partials of k_gett_f32<..., true> in split order
// (deterministic) and applies the epilogue.  Position p of a split's partial
// block is (l, tile, wave, slot, lane, element) with element fastest, the
// order the GETT kernel stored its accumulators in.
__global__ __launch_bounds__(256) void k_splitk_reduce(tk_gett g)
{
    const int64_t ntm = (g.M + 127) / 128, ntn = (g.N + 127) / 128;
    const int64_t per_split = g.L * ntm * ntn * 16384;
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < per_split; p += (int64_t)gridDim.x * 256) {
        const int e = (int)(p & 3), lane = (int)((p >> 2) & 63), slot = (int)((p >> 8) & 15), wave = (int)((p >> 12) & 3);
        const int64_t tl = p >> 14, tile = tl % (ntm * ntn), l = tl / (ntm * ntn);
        const int fn = slot >> 3, fm = (slot >> 2) & 1, r = 4 * (slot & 3) + e;
        const int64_t m = (tile / ntn) * 128 + (wave & 1) * 64 + fm * 32 + (lane & 31);
        const int64_t n = (tile % ntn) * 128 + (wave >> 1) * 64 + fn * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (m >= g.M || n >= g.N) continue;
        float s = 0.0f;
        for (int z = 0; z < g.ksplit; ++z) s += g.partial[z * per_split + p];
        epilogue<float, false>(g, l, m, n, s, 0.0f);
    }
}

// Name of the kernel template the calling thread launched last (read back by
// the host for TAPP_AMD_ATTR_LAST_KERNEL_NAME).
static thread_local char last_name[96];

static const char* load_name(int m)
{
    return m == TK_CONTIG_ROW ? "CONTIG_ROW" : m == TK_CONTIG_K ? "CONTIG_K" : "GATHER";
}

static const char* mode_name(int m)
{
    switch (m) {
    case TK_MODE_PLAIN: return "PLAIN";
    case TK_MODE_BF16: return "BF16";
    case TK_MODE_F16: return "F16";
    case TK_MODE_SPLIT_BF16: return "SPLIT_BF16";
    case TK_MODE_SPLIT_F16: return "SPLIT_F16";
    default: return "SPLIT6_BF16";
    }
}

template <int AL, int BL>
static hipError_t launch_gett_f32_modes(const tk_gett& g, hipStream_t st)
{
    const int64_t tiles = ((g.M + 127) / 128) * ((g.N + 127) / 128);
    const int64_t ly = g.L < 65535 ? g.L : 65535;
    const dim3 grid((unsigned)tiles, (unsigned)ly, (unsigned)(g.ksplit > 1 ? g.ksplit : 1)), block(256);
    const bool split = g.ksplit > 1;
#define TK_LAUNCH(KERNEL, MODE)                                                       \
    if (split) hipLaunchKernelGGL((KERNEL<AL, BL, MODE, true>), grid, block, 0, st, g); \
    else hipLaunchKernelGGL((KERNEL<AL, BL, MODE, false>), grid, block, 0, st, g)
    switch (g.mode) {
    case TK_MODE_PLAIN: TK_LAUNCH(k_gett_f32, TK_MODE_PLAIN); break;
    case TK_MODE_BF16: TK_LAUNCH(k_gett_lp, TK_MODE_BF16); break;
    case TK_MODE_F16: TK_LAUNCH(k_gett_lp, TK_MODE_F16); break;
    case TK_MODE_SPLIT_BF16: TK_LAUNCH(k_gett_lp, TK_MODE_SPLIT_BF16); break;
    case TK_MODE_SPLIT_F16: TK_LAUNCH(k_gett_lp, TK_MODE_SPLIT_F16); break;
    default: TK_LAUNCH(k_gett_lp, TK_MODE_SPLIT6_BF16); break;
    }
#undef TK_LAUNCH
    snprintf(last_name, sizeof last_name, "%s<%s, %s, %s%s>", g.mode == TK_MODE_PLAIN ? "k_gett_f32" : "k_gett_lp",
             load_name(AL), load_name(BL), mode_name(g.mode), split ? ", split-K" : "");
    if (g.ksplit > 1) {
        const int64_t total = tiles * g.L * 16384;
        int64_t blocks = (total + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL(k_splitk_reduce, dim3((unsigned)blocks), dim3(256), 0, st, g);
    }
    return hipGetLastError();
}

static hipError_t launch_gett_f32(const tk_gett& g, hipStream_t st)
{
    const int a = g.a_vec ? g.a_load : TK_GATHER;
    const int b = g.b_vec ? g.b_load : TK_GATHER;
#define TK_F32_CASE(X, Y) \
    if (a == X && b == Y) return launch_gett_f32_modes<X, Y>(g, st);
    TK_F32_CASE(TK_CONTIG_ROW, TK_CONTIG_ROW)
    TK_F32_CASE(TK_CONTIG_ROW, TK_CONTIG_K)
    TK_F32_CASE(TK_CONTIG_ROW, TK_GATHER)
    TK_F32_CASE(TK_CONTIG_K, TK_CONTIG_ROW)
    TK_F32_CASE(TK_CONTIG_K, TK_CONTIG_K)
    TK_F32_CASE(TK_CONTIG_K, TK_GATHER)
    TK_F32_CASE(TK_GATHER, TK_CONTIG_ROW)
    TK_F32_CASE(TK_GATHER, TK_CONTIG_K)
    TK_F32_CASE(TK_GATHER, TK_GATHER)
#undef TK_F32_CASE
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
template <typename R, bool CPLX>
static hipError_t launch_simple(const tk_gett& g, hipStream_t st)
{
    const int64_t total = g.M * g.N * g.L;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    snprintf(last_name, sizeof last_name, "k_simple<%s, %s>", sizeof(R) == 8 ? "double" : "float",
             CPLX ? "complex" : "real");
    hipLaunchKernelGGL((k_simple<R, CPLX>), dim3((unsigned)blocks), dim3(256), 0, st, g);
    return hipGetLastError();
}

template <typename R, bool CPLX>
static hipError_t launch_splitk(const tk_gett& g, hipStream_t st)
{
    const int64_t total = g.M * g.N * g.L;
    snprintf(last_name, sizeof last_name, "k_splitk<%s, %s>", sizeof(R) == 8 ? "double" : "float",
             CPLX ? "complex" : "real");
    hipLaunchKernelGGL((k_splitk<R, CPLX>), dim3((unsigned)total), dim3(256), 0, st, g);
    return hipGetLastError();
}

template <typename R, bool CPLX, int S, int FM, int FN, int BK>
static int64_t mfma_workgroups(const tk_gett& g)
{
    constexpr int BM = 2 * FM * S, BN = 2 * FN * S;
    return ((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN) * g.L;
}

template <typename R, bool CPLX, int S, int FM, int FN, int BK>
static hipError_t launch_mfma(const tk_gett& g, hipStream_t st)
{
    constexpr int BM = 2 * FM * S, BN = 2 * FN * S;
    const int64_t tiles = ((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN);
    const int64_t ly = g.L < 65535 ? g.L : 65535;
    snprintf(last_name, sizeof last_name, "k_mfma<%s, %s, %d, %d, %d, %d>", sizeof(R) == 8 ? "double" : "float",
             CPLX ? "complex" : "real", S, FM, FN, BK);
    hipLaunchKernelGGL((k_mfma<R, CPLX, S, FM, FN, BK>), dim3((unsigned)tiles, (unsigned)ly), dim3(256), 0, st, g);
    return hipGetLastError();
}

}  // namespace tk

using namespace tk;

extern "C" int tk_launch_gett(const tk_gett* gp, void* stream)
{
    const tk_gett& g = *gp;
    hipStream_t st = (hipStream_t)stream;
    if (g.M <= 0 || g.N <= 0 || g.L <= 0) return 0;
    const int kern = g.kernel;
    switch (g.ctype) {
    case TK_F32:
        if (kern == TK_KERNEL_SIMPLE) return launch_simple<float, false>(g, st);
        if (kern == TK_KERNEL_SPLITK) return launch_splitk<float, false>(g, st);
        if (kern == TK_KERNEL_SMALL && g.mode == TK_MODE_PLAIN) {
            const int64_t tiles = ((g.M + 15) / 16) * ((g.N + 15) / 16);
            const int64_t ly = g.L < 65535 ? g.L : 65535;
            snprintf(last_name, sizeof last_name, "k_small_f32<16>");
            hipLaunchKernelGGL((k_small_f32<16>), dim3((unsigned)tiles, (unsigned)ly), dim3(256), 0, st, g);
            return (int)hipGetLastError();
        }
        if (g.tile_variant != 2 && tk_gett_wide_fits(&g)) {
            snprintf(last_name, sizeof last_name, "k_gett_wide<%s, %s, 128>", load_name(g.a_load),
                     load_name(g.b_load));
            return tk_launch_gett_wide(&g, st);
        }
        return launch_gett_f32(g, st);
    // F64 and complex: large tiles while they give every CU 4 workgroups,
    // else 4x smaller ones (more workgroups and waves in flight for short-K
    // problems).  Tall tiles -- the small tile's columns, twice its rows --
    // run when asked for (tile variant 3, an autotuner candidate): for a
    // complex-double product with M = 40 and K = 24, one 64-row tile per
    // column block instead of a 32-row and an 8-row one measured 64.6 us
    // against 59.9 us, the halved workgroup count costing more than the
    // shorter chains save.
    case TK_F64: {
        if (kern == TK_KERNEL_SIMPLE) return launch_simple<double, false>(g, st);
        if (kern == TK_KERNEL_SPLITK) return launch_splitk<double, false>(g, st);
        const bool large = g.tile_variant == 1 ||
                           (g.tile_variant == 0 && mfma_workgroups<double, false, 16, 4, 4, 8>(g) >= 4 * (int64_t)g.cus);
        if (large) return launch_mfma<double, false, 16, 4, 4, 8>(g, st);
        if (g.tile_variant == 3) return launch_mfma<double, false, 16, 4, 2, 8>(g, st);
        return launch_mfma<double, false, 16, 2, 2, 8>(g, st);
    }
    case TK_C32: {
        if (kern == TK_KERNEL_SIMPLE || g.mode != TK_MODE_PLAIN) return launch_simple<float, true>(g, st);
        if (kern == TK_KERNEL_SPLITK) return launch_splitk<float, true>(g, st);
        const bool large = g.tile_variant == 1 ||
                           (g.tile_variant == 0 && mfma_workgroups<float, true, 32, 2, 2, 8>(g) >= 4 * (int64_t)g.cus);
        if (large) return launch_mfma<float, true, 32, 2, 2, 8>(g, st);
        if (g.tile_variant == 3) return launch_mfma<float, true, 32, 2, 1, 8>(g, st);
        return launch_mfma<float, true, 32, 1, 1, 8>(g, st);
    }
    case TK_C64: {
        if (kern == TK_KERNEL_SIMPLE) return launch_simple<double, true>(g, st);
        if (kern == TK_KERNEL_SPLITK) return launch_splitk<double, true>(g, st);
        const bool large = g.tile_variant == 1 ||
                           (g.tile_variant == 0 && mfma_workgroups<double, true, 16, 2, 2, 8>(g) >= 4 * (int64_t)g.cus);
        if (large) return launch_mfma<double, true, 16, 2, 2, 8>(g, st);
        if (g.tile_variant == 3) return launch_mfma<double, true, 16, 2, 1, 8>(g, st);
        return launch_mfma<double, true, 16, 1, 1, 8>(g, st);
    }
    default:
        return (int)hipErrorInvalidValue;
    }
}

extern "C" int tk_kernel_count(void) { return 16; }

extern "C" const char* tk_last_gett_name(void) { return last_name; }
