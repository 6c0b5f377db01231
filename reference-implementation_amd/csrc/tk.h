/*
 * Internal C ABI between the C host layer (csrc/host) and the gfx950 HIP
 * kernels (csrc/hip).  Plain C types only: device pointers, sizes, int64
 * offset tables.  Not part of the public API.
 *
 * Index model (shared by every kernel).  After the plan stage a product is
 *     D[l, m, n] = op_D( alpha * sum_k A'[l, m, k] * B'[l, k, n] + beta * op_C(C[l, m, n]) )
 * where each group (L = Hadamard/batch labels, M = labels of A and D only,
 * N = labels of B and D only, K = contracted labels) is linearised and every
 * operand reaches element (l, m, n, k) through one int64 offset table per
 * group: address = base + lX[l] + mX[m] + nX[n] + kX[k] (in elements).
 * Tables carry signed strides, diagonals (summed strides) and zero strides
 * unchanged.
 */
#ifndef TK_H_
#define TK_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* compute types */
enum { TK_F32 = 0, TK_F64 = 1, TK_C32 = 2, TK_C64 = 3 };

/* multiply modes for single-precision compute */
enum {
    TK_MODE_PLAIN = 0,     /* FP32 or FP64 arithmetic as stored */
    TK_MODE_F16 = 1,       /* operands rounded to binary16, FP16 MFMA, FP32 accumulate */
    TK_MODE_BF16 = 2,      /* operands rounded to bfloat16, BF16 MFMA, FP32 accumulate */
    TK_MODE_SPLIT_BF16 = 3, /* x = hi + lo in bfloat16, hi*hi + hi*lo + lo*hi, FP32 accumulate */
    TK_MODE_SPLIT6_BF16 = 4, /* x = h + m + l in bfloat16, six products (hh, hm, mh, hl, lh, mm),
                               FP32 sums of 128 k added with compensation (Kahan) */
    TK_MODE_SPLIT_F16 = 5   /* x * s = hi + lo in binary16 (s a power of two per operand),
                               hi*hi + hi*lo + lo*hi, FP32 sums of 128 k added into an FP32
                               total; the kernel reports max |x| of each operand so the host
                               can check the range */
};

/* how a kernel may read an operand tile */
enum {
    TK_GATHER = 0,      /* per-element table lookups */
    TK_CONTIG_ROW = 1,  /* unit step along the M (A, C, D) or N (B) group inside each tile */
    TK_CONTIG_K = 2     /* unit step along K inside each tile */
};

typedef struct tk_gett {
    int ctype;               /* TK_F32 .. TK_C64 */
    int mode;                /* TK_MODE_* (single precision only) */
    int64_t M, N, K, L;
    const void* A;           /* device addresses of element (0,...,0) */
    const void* B;
    const void* C;           /* NULL when beta == 0: C is never read */
    void* D;
    const int64_t *mA, *mC, *mD;   /* [M] */
    const int64_t *nB, *nC, *nD;   /* [N] */
    const int64_t *kA, *kB;        /* [K] */
    const int64_t *lA, *lB, *lC, *lD;  /* [L] */
    int conjA, conjB, conjC, conjD;
    int conjD_via_f32;       /* conj of the result goes through complex float */
    int round_C;             /* C rounded like the operands (low-precision modes) */
    double alpha[2], beta[2];/* (re, im), already in the compute precision */
    int beta_zero;
    int a_load, b_load;      /* TK_GATHER / TK_CONTIG_ROW / TK_CONTIG_K */
    int a_vec, b_vec;        /* vector width (elements) usable on the contiguous dim */
    int d_contig;            /* D (and C if c_contig) unit step along M inside each tile */
    int c_contig;
    int kernel;              /* chosen kernel variant, TK_KERNEL_* */
    int cus;                 /* compute units of the device (tile-size choice) */
    int tile_variant;        /* MFMA tiles: 0 by shape, 1 large, 2 small (F64/C32/C64 tile size;
                                FP32 split forms: 256 x 128 unless 2) */
    int ksplit;              /* > 1: K split over gridDim.z; partial sums go to `partial` */
    int64_t kchunk;          /* K length per split (multiple of the kernel's K tile) */
    float* partial;          /* [ksplit][L][N][M] FP32 partial sums (real compute only) */
    float scale[2];          /* TK_MODE_SPLIT_F16: A and B are multiplied by these before the split */
    uint32_t* amax;          /* TK_MODE_SPLIT_F16: [2] raised (atomic max) to the bits of max |A|,
                                max |B| (unscaled; NaN and inf compare above every finite value) */
} tk_gett;

enum {
    TK_KERNEL_AUTO = 0,
    TK_KERNEL_SIMPLE = 1,    /* one lane per output element, FMA over K */
    TK_KERNEL_MFMA = 2,      /* LDS-tiled MFMA GETT */
    TK_KERNEL_SPLITK = 3,    /* one workgroup per output, K split across lanes */
    TK_KERNEL_SMALL = 5      /* FP32: 16 x 16 MFMA tiles, K split over the workgroup's waves */
};

/* Pre-pass: reduce unary modes, apply conj and input rounding, convert the
 * storage type to the compute type and write a dense buffer
 *     dst[i0 + n0 * (i1 + n1 * i2)] = sum_u op(src[t0[i0] + t1[i1] + t2[i2] + tu[u]])
 * summing u in ascending order (the reference's odometer order). */
typedef struct tk_prep {
    int src_type;            /* TAPP_datatype of the source */
    int ctype;               /* compute type of dst */
    int conj;
    int round;               /* 0, TK_MODE_F16, TK_MODE_BF16: rounding of each loaded value */
    const void* src;
    void* dst;
    int64_t n0, n1, n2, nu;
    const int64_t *t0, *t1, *t2, *tu;
} tk_prep;

/* Post-pass for mixed storage: D[l, m, n] (type_D, D's tables) =
 * convert(src dense compute-type buffer, index m + M * (n + N * l)). */
typedef struct tk_scatter {
    int ctype;
    int dst_type;
    const void* src;
    void* dst;
    int64_t M, N, L;
    const int64_t *mD, *nD, *lD;
} tk_scatter;

/* One product of a persistent group launch (FP32 compute, no unary labels,
 * device operands).  Classes:
 *   TK_GROUP_THIN: D[l, s, g] = sum_k X[l, s, k] Y[l, k, g] with S <= 16; a
 *     workgroup takes 256 columns g and every s, streaming Y through LDS.
 *   TK_GROUP_MFMA: D[l, m, n] over TK_GROUP_MT x TK_GROUP_MT MFMA tiles (X = rows m,
 *     Y = cols n).
 * tiles = L * tiles_s * tiles_g; tile_end is the inclusive prefix over the
 * group's problem order. */
enum { TK_GROUP_THIN = 0, TK_GROUP_MFMA = 1 };

/* Which tiles a persistent launch runs: both classes (the serving loop's
 * batches), only thin tiles or only MFMA tiles (product groups launch the
 * two classes as two kernels, each at its own occupancy). */
enum { TK_GROUP_KIND_ALL = 0, TK_GROUP_KIND_THIN = 1, TK_GROUP_KIND_MFMA = 2 };
/* Persistent group workgroups per CU: the kernel's occupancy bound (launch
 * bounds) and the grid the host launches.  A kernel with MFMA tiles runs
 * three 256-lane workgroups (168 VGPRs each): the MFMA tile's four 64 x 64
 * wave blocks hold 64 accumulators per lane beside two batches of loads in
 * flight, which four per CU (128 VGPRs) would spill.  The thin-only kernel
 * fits 128 VGPRs and runs four, with more of the streamed Y in flight. */
#define TK_GROUP_WG_PER_CU 3
#define TK_GROUP_WG_PER_CU_THIN 4
/* Edge of a TK_GROUP_MFMA output tile. */
#define TK_GROUP_MT 128

typedef struct tk_prob {
    int64_t S, G, K, L;
    int64_t tiles_s, tiles_g;
    const float* X;
    const float* Y;
    const float* C;
    float* D;
    const int64_t *sX, *kX, *lX;
    const int64_t *gY, *kY, *lY;
    const int64_t *sD, *gD, *lD;
    const int64_t *sC, *gC, *lC;
    float alpha, beta;
    int beta_zero;
    int cls;
    int y_kfast;        /* 0: load Y along g, 1: along K (unit K step), 2: dense K x G block */
    int round;          /* 0 or TK_MODE_BF16 / TK_MODE_F16: inputs rounded (FP32 accumulate) */
    int round_C;
    int x_is_b;         /* X is B (host bookkeeping; not read by the kernel) */
    /* MFMA tiles: X + xlo (Y + ylo) is the lowest element the tile can read,
     * and every element lies within 2^29 of it, so offsets are 32-bit */
    int64_t xlo, ylo;
    /* Thin tiles over a dense Y (y_kfast 2) whose other tables are all
     * index * stride: affine = 1 and the strides below, so a tile reads no
     * table.  X (s, k, l) at s ssX + k skX + l slX, Y's l at l slY, D
     * (s, g, l) at s ssD + g sgD + l slD, C likewise. */
    int affine;
    int64_t ssX, skX, slX, slY, ssD, sgD, slD, ssC, sgC, slC;
} tk_prob;

/* A tile of a persistent launch, decoded on the host: its product (index
 * into the launch's descriptors), Hadamard index and row / column tile. */
typedef struct tk_tile {
    int32_t prob, l, ts, tg;
} tk_tile;

/* All launchers are asynchronous on `stream` (a hipStream_t) and return 0
 * or a hipError_t value. */
int tk_launch_gett(const tk_gett* g, void* stream);
/* Split-BF16 with 256 x 128 tiles (both operands loaded along a contiguous
 * direction, no split-K, at least one tile per CU): whether g qualifies,
 * and its launch.  tk_launch_gett takes it by itself unless
 * g->tile_variant is 2. */
int tk_gett_wide_fits(const tk_gett* g);
int tk_launch_gett_wide(const tk_gett* g, void* stream);
int tk_launch_prep(const tk_prep* p, void* stream);
/* A's and B's pre-passes (same compute type) as one launch */
int tk_launch_prep2(const tk_prep* a, const tk_prep* b, void* stream);
int tk_launch_scatter(const tk_scatter* s, void* stream);
/* Zeroes *counter, then runs the persistent group kernel: `nblocks`
 * workgroups take tiles from the atomic counter until `total` is reached.
 * tiles (optional, tk_tile[total]) decodes every tile; without it a
 * workgroup finds a tile's product by a binary search over tile_end. */
int tk_launch_group(const tk_prob* probs, const int64_t* tile_end, const tk_tile* tiles, int nprobs,
                    int64_t total, unsigned long long* counter, int nblocks, void* stream);
/* Same with `nq` tile queues: queue q holds tiles [qbeg[q], qbeg[q + 1])
 * (qbeg on the device, nq + 1 entries) and counts with counter[32 q]; a
 * workgroup drains queue blockIdx % nq first, then the others. */
int tk_launch_group_q(const tk_prob* probs, const int64_t* tile_end, const tk_tile* tiles, int nprobs,
                      int64_t total, const int64_t* qbeg, int nq, unsigned long long* counter, int nblocks,
                      void* stream);
/* Same, running only the tiles of `kind` (TK_GROUP_KIND_*): every product
 * given must be of that class. */
int tk_launch_group_kind(int kind, const tk_prob* probs, const int64_t* tile_end, const tk_tile* tiles, int nprobs,
                         int64_t total, const int64_t* qbeg, int nq, unsigned long long* counter, int nblocks,
                         void* stream);

/* Element-wise comparison of two results of one product (self-checking
 * executor): x[i] and y[i] (storage type `type`, TAPP_datatype) for i < n.
 * *count receives the number of differing i and first[] up to
 * TK_COMPARE_KEEP of them, in no particular order. */
#define TK_COMPARE_KEEP 16
typedef struct tk_compare {
    int type;
    const void* x;
    const void* y;
    int64_t n;
    unsigned long long* count;
    int64_t* first;          /* [TK_COMPARE_KEEP] */
} tk_compare;

int tk_launch_compare(const tk_compare* c, void* stream);

/* Device scratch of the allocation-free entry point: a file-scope device
 * array of TK_SCRATCH_BYTES per GPU (tk_scratch.hip).  *ptr receives its
 * device address on `device` (looked up once per device, then cached). */
#define TK_SCRATCH_BYTES ((size_t)64 << 20)
int tk_scratch_get(int device, void** ptr, size_t* bytes);

/* Number of kernel variants compiled in, for diagnostics. */
int tk_kernel_count(void);
/* Name of the kernel template the calling thread last launched through
 * tk_launch_gett. */
const char* tk_last_gett_name(void);

#ifdef __cplusplus
}
#endif

#endif /* TK_H_ */
