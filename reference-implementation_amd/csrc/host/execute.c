/*
 * Device orchestration of one product (the part of product.c:263-422 that
 * runs the loop nest, re-designed for the GPU).
 *
 * Offset tables: each group (L, M, N, K, UA, UB) is linearised with its
 * first mode fastest and every operand gets one int64 table per group
 * holding sum_i coord_i * stride_i.  Tables depend only on the layouts, so
 * they are built once per (plan, layout) and kept on the device in the
 * plan's cache; a call with unchanged layouts uploads nothing.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

#include <stdio.h>

/* Last device error of this thread, for TAPP_AMD_last_error_message. */
static _Thread_local char last_error[256];

static int device_error(const char* what, int err, int line)
{
    snprintf(last_error, sizeof last_error, "%s failed at execute.c:%d: %s (%d)", what, line,
             hipGetErrorString((hipError_t)err), err);
    if (getenv("TAPP_AMD_VERBOSE")) fprintf(stderr, "tapp-amd: %s\n", last_error);
    return TAPP_AMD_ERR_DEVICE;
}

#define HIPCHK(call)                                                            \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) return device_error(#call, (int)e_, __LINE__);    \
    } while (0)
#define LAUNCHCHK(call)                                                         \
    do {                                                                        \
        int e_ = (call);                                                        \
        if (e_ != 0) return device_error(#call, e_, __LINE__);                  \
    } while (0)

#define LAUNCHCHK_UNLOCK(call)                                                  \
    do {                                                                        \
        int e_ = (call);                                                        \
        if (e_ != 0) rc = device_error(#call, e_, __LINE__);                    \
    } while (0)

size_t TAPP_AMD_last_error_message(char* message, size_t maxlen)
{
    size_t len = strlen(last_error);
    if (maxlen == 0) return len;
    size_t n = maxlen - 1 < len ? maxlen - 1 : len;
    memcpy(message, last_error, n);
    message[n] = '\0';
    return n;
}

#define ALIGN256(x) (((x) + 255u) & ~(size_t)255u)

enum {
    T_mA, T_mC, T_mD, T_nB, T_nC, T_nD, T_kA, T_kB, T_lA, T_lB, T_lC, T_lD, T_uA, T_uB,
    /* dense layouts of the pre-pass buffers */
    T_dmA, T_dkA, T_dlA, T_dnB, T_dkB, T_dlB, T_dmD, T_dnD, T_dlD, T_zero,
    T_COUNT
};

struct tapp_plan_cache {
    uint64_t key;
    int ordA[3], ordB[3];    /* pre-pass order of (rows, K, L), fastest first */
    size_t off[T_COUNT];     /* element offset of each table in the arena */
    size_t total;
    int64_t* host;
    int64_t* dev;
    int device;
};

void tk_plan_cache_free(struct tapp_plan_cache* c)
{
    if (!c) return;
    if (c->dev) hipFree(c->dev);
    free(c->host);
    free(c);
}

const void* tk_plan_tables_id(const struct tapp_plan* plan, uint64_t* key)
{
    const struct tapp_plan_cache* c = plan->cache;
    *key = c ? c->key : 0;
    return c ? (const void*)c->dev : NULL;
}

static uint64_t mix(uint64_t h, uint64_t v)
{
    h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
    return h * 0xff51afd7ed558ccdull;
}

static uint64_t group_key(uint64_t h, const tk_group* g)
{
    h = mix(h, (uint64_t)g->n);
    for (int i = 0; i < g->n; i++) {
        h = mix(h, (uint64_t)g->ext[i]);
        for (int x = 0; x < 4; x++) h = mix(h, (uint64_t)g->s[x][i]);
    }
    return h;
}

static uint64_t problem_key(const tk_problem* P)
{
    uint64_t h = 0x1234567ull;
    h = group_key(h, &P->L); h = group_key(h, &P->M); h = group_key(h, &P->N);
    h = group_key(h, &P->K); h = group_key(h, &P->UA); h = group_key(h, &P->UB);
    h = mix(h, (uint64_t)P->nK);
    return h;
}

/* tab[i] = sum_j coord_j(i) * stride_j over the group's modes, first fastest */
static void fill_table(int64_t* tab, const tk_group* g, int x, int64_t n)
{
    if (n <= 0) return;
    tab[0] = 0;
    int64_t block = 1;
    for (int i = 0; i < g->n; i++) {
        for (int64_t c = 1; c < g->ext[i]; c++)
            for (int64_t r = 0; r < block; r++) tab[c * block + r] = tab[r] + c * g->s[x][i];
        block *= g->ext[i];
    }
}

static void fill_linear(int64_t* tab, int64_t n, int64_t step)
{
    for (int64_t i = 0; i < n; i++) tab[i] = i * step;
}

/* smallest |stride| of operand x over a group's modes (INT64_MAX if none) */
static int64_t min_stride(const tk_group* g, int x)
{
    int64_t m = INT64_MAX;
    for (int i = 0; i < g->n; i++) {
        int64_t v = llabs(g->s[x][i]);
        if (v < m) m = v;
    }
    return m;
}

/* Order in which the pre-pass walks (rows, K, L) of operand x: the group
 * holding the operand's smallest stride first, so consecutive lanes read
 * neighbouring elements; the dense buffer is written in the same order.
 * dense_step receives each group's step in that buffer. */
static void prep_order(const tk_group* rows, const tk_group* K, const tk_group* L, int x, const int64_t size[3],
                       int ord[3], int64_t dense_step[3])
{
    int64_t key[3] = {min_stride(rows, x), min_stride(K, x), min_stride(L, x)};
    ord[0] = 0; ord[1] = 1; ord[2] = 2;
    for (int i = 1; i < 3; i++)
        for (int j = i; j > 0 && key[ord[j]] < key[ord[j - 1]]; j--) {
            int t = ord[j]; ord[j] = ord[j - 1]; ord[j - 1] = t;
        }
    int64_t step = 1;
    for (int i = 0; i < 3; i++) { dense_step[ord[i]] = step; step *= size[ord[i]]; }
}

/* Element offset of each table in a plan's arena (off may be NULL); returns
 * the arena's length in elements. */
static size_t table_layout(const tk_problem* P, size_t* off)
{
    const int64_t M = P->nM, N = P->nN, K = P->nK, L = P->nL;
    const int64_t len[T_COUNT] = {
        [T_mA] = M, [T_mC] = M, [T_mD] = M, [T_nB] = N, [T_nC] = N, [T_nD] = N,
        [T_kA] = K, [T_kB] = K, [T_lA] = L, [T_lB] = L, [T_lC] = L, [T_lD] = L,
        [T_uA] = P->nUA, [T_uB] = P->nUB,
        [T_dmA] = M, [T_dkA] = K, [T_dlA] = L, [T_dnB] = N, [T_dkB] = K, [T_dlB] = L,
        [T_dmD] = M, [T_dnD] = N, [T_dlD] = L, [T_zero] = 1};
    size_t total = 0;
    for (int t = 0; t < T_COUNT; t++) {
        if (off) off[t] = total;
        total += (size_t)(len[t] > 0 ? len[t] : 1);
        total = (total + 31) & ~(size_t)31;   /* 256-byte aligned sub-tables */
    }
    return total;
}

/* The host side of a plan's tables: offsets, sizes and contents. */
static struct tapp_plan_cache* host_tables(const tk_problem* P, uint64_t key)
{
    struct tapp_plan_cache* c = calloc(1, sizeof *c);
    if (!c) return NULL;
    const int64_t M = P->nM, N = P->nN, K = P->nK, L = P->nL;
    const size_t total = table_layout(P, c->off);
    c->total = total;
    c->key = key;
    c->host = malloc(total * sizeof(int64_t));
    if (!c->host) { free(c); return NULL; }
    int64_t* h = c->host;
    fill_table(h + c->off[T_mA], &P->M, 0, M);
    fill_table(h + c->off[T_mC], &P->M, 2, M);
    fill_table(h + c->off[T_mD], &P->M, 3, M);
    fill_table(h + c->off[T_nB], &P->N, 1, N);
    fill_table(h + c->off[T_nC], &P->N, 2, N);
    fill_table(h + c->off[T_nD], &P->N, 3, N);
    fill_table(h + c->off[T_kA], &P->K, 0, K);
    fill_table(h + c->off[T_kB], &P->K, 1, K);
    fill_table(h + c->off[T_lA], &P->L, 0, L);
    fill_table(h + c->off[T_lB], &P->L, 1, L);
    fill_table(h + c->off[T_lC], &P->L, 2, L);
    fill_table(h + c->off[T_lD], &P->L, 3, L);
    fill_table(h + c->off[T_uA], &P->UA, 0, P->nUA);
    fill_table(h + c->off[T_uB], &P->UB, 1, P->nUB);
    {
        const int64_t sa[3] = {M, K, L}, sb[3] = {N, K, L};
        int64_t da[3], db[3];
        prep_order(&P->M, &P->K, &P->L, 0, sa, c->ordA, da);
        prep_order(&P->N, &P->K, &P->L, 1, sb, c->ordB, db);
        fill_linear(h + c->off[T_dmA], M, da[0]);
        fill_linear(h + c->off[T_dkA], K, da[1]);
        fill_linear(h + c->off[T_dlA], L, da[2]);
        fill_linear(h + c->off[T_dnB], N, db[0]);
        fill_linear(h + c->off[T_dkB], K, db[1]);
        fill_linear(h + c->off[T_dlB], L, db[2]);
    }
    fill_linear(h + c->off[T_dmD], M, 1);
    fill_linear(h + c->off[T_dnD], N, M);
    fill_linear(h + c->off[T_dlD], L, M * N);
    h[c->off[T_zero]] = 0;
    return c;
}

static int build_tables(struct tapp_plan* plan, const tk_problem* P, int device, struct tapp_plan_cache** out)
{
    uint64_t key = problem_key(P);
    struct tapp_plan_cache* c = plan->cache;
    if (c && c->key == key && c->device == device && c->dev) { *out = c; return 0; }
    if (c) { tk_plan_cache_free(c); plan->cache = NULL; }
    c = host_tables(P, key);
    if (!c) return TAPP_AMD_ERR_NOMEM;
    const size_t total = c->total;
    hipError_t e = hipMalloc((void**)&c->dev, total * sizeof(int64_t));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        free(c->host); free(c);
        device_error("hipMalloc(tables)", (int)e, __LINE__);
        return TAPP_AMD_ERR_NOMEM;
    }
    e = hipMemcpy(c->dev, c->host, total * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        tk_plan_cache_free(c);
        return device_error("hipMemcpy(tables)", (int)e, __LINE__);
    }
    c->device = device;
    plan->cache = c;
    *out = c;
    return 0;
}

/* Host tables only (no device copy), kept in the plan's second cache slot:
 * the allocation-free path uploads them into its scratch each call, and
 * never frees device memory when a layout changes. */
static int build_host_tables(struct tapp_plan* plan, const tk_problem* P, struct tapp_plan_cache** out)
{
    uint64_t key = problem_key(P);
    struct tapp_plan_cache* c = plan->hcache;
    if (c && c->key == key) { *out = c; return 0; }
    if (c) { tk_plan_cache_free(c); plan->hcache = NULL; }
    if (!(c = host_tables(P, key))) return TAPP_AMD_ERR_NOMEM;
    plan->hcache = c;
    *out = c;
    return 0;
}

/* Device tables of a planned problem (built or taken from the plan cache). */
int tk_problem_tables(struct tapp_plan* plan, const tk_problem* P, int device, const int64_t** mA,
                      const int64_t** kA, const int64_t** lA, const int64_t** nB, const int64_t** kB,
                      const int64_t** lB, const int64_t** mD, const int64_t** nD, const int64_t** lD,
                      const int64_t** mC, const int64_t** nC, const int64_t** lC)
{
    struct tapp_plan_cache* c;
    pthread_mutex_lock(&plan->mu);
    int rc = build_tables(plan, P, device, &c);
    pthread_mutex_unlock(&plan->mu);
    if (rc) return rc;
    const int64_t* T = c->dev;
    *mA = T + c->off[T_mA]; *kA = T + c->off[T_kA]; *lA = T + c->off[T_lA];
    *nB = T + c->off[T_nB]; *kB = T + c->off[T_kB]; *lB = T + c->off[T_lB];
    *mD = T + c->off[T_mD]; *nD = T + c->off[T_nD]; *lD = T + c->off[T_lD];
    *mC = T + c->off[T_mC]; *nC = T + c->off[T_nC]; *lC = T + c->off[T_lC];
    return 0;
}

/* Host copies of the K tables of A and B (valid after tk_problem_tables). */
void tk_problem_host_k(struct tapp_plan* plan, const int64_t** kA, const int64_t** kB)
{
    struct tapp_plan_cache* c = plan->cache;
    *kA = c->host + c->off[T_kA];
    *kB = c->host + c->off[T_kB];
}

/* Host copies of the M table of A and the N table of B. */
void tk_problem_host_mn(struct tapp_plan* plan, const int64_t** mA, const int64_t** nB)
{
    struct tapp_plan_cache* c = plan->cache;
    *mA = c->host + c->off[T_mA];
    *nB = c->host + c->off[T_nB];
}

/* --- residency ------------------------------------------------------------ */


static int ensure_ws(struct tapp_exec* ex, size_t bytes)
{
    if (bytes <= ex->ws_size) return 0;
    if (ex->ws) { hipStreamSynchronize(ex->stream); hipFree(ex->ws); ex->ws = NULL; ex->ws_size = 0; }
    size_t want = bytes + bytes / 4;
    hipError_t e = hipMalloc(&ex->ws, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        device_error("hipMalloc(workspace)", (int)e, __LINE__);
        return TAPP_AMD_ERR_NOMEM;
    }
    ex->ws_size = want;
    return 0;
}

/* --- autotuning of single products (tune.c holds the cache) -------------
 * A configuration packs the kernel family (bits 0-3), the F64/C32/C64 MFMA
 * tile size (bits 4-7: 0 by workgroup count, 1 large, 2 small) and whether a
 * split-K launch is turned off (bit 8). */
static void apply_cfg(tk_gett* g, int cfg, int64_t ks, int64_t kchunk, float* partial)
{
    g->kernel = cfg & 15;
    g->tile_variant = (cfg >> 4) & 15;
    const bool split = g->kernel == TK_KERNEL_MFMA && g->ctype == TK_F32 && ks > 1 && !((cfg >> 8) & 1);
    g->ksplit = split ? (int)ks : 0;
    g->kchunk = split ? kchunk : 0;
    g->partial = split ? partial : NULL;
}

typedef struct {
    tk_gett* g;
    hipStream_t st;
} gett_ctx;

static int launch_gett_ctx(void* c)
{
    gett_ctx* x = (gett_ctx*)c;
    return tk_launch_gett(x->g, x->st);
}

/* Candidates in a fixed order, each timed over 3 launches after a warm-up;
 * the fastest wins (ties keep the earlier one). */
static int tune_gett(struct tapp_exec* ex, tk_gett* g, int64_t ks, int64_t kchunk, float* partial, int64_t outs,
                     int* best)
{
    int cand[8], nc = 0;
    const double macs = (double)outs * (double)g->K;
    if (macs <= (double)(1 << 30)) cand[nc++] = TK_KERNEL_SIMPLE;
    if (g->ctype == TK_F32) {
        cand[nc++] = TK_KERNEL_MFMA;
        if (ks > 1) cand[nc++] = TK_KERNEL_MFMA | (1 << 8);
        if (tk_gett_wide_fits(g)) cand[nc++] = TK_KERNEL_MFMA | (2 << 4);
        const int64_t t16 = ((g->M + 15) / 16) * ((g->N + 15) / 16) * g->L;
        if (g->mode == TK_MODE_PLAIN && g->K <= 8192 && t16 <= 16 * (int64_t)ex->cus) cand[nc++] = TK_KERNEL_SMALL;
    } else {
        cand[nc++] = TK_KERNEL_MFMA | (1 << 4);
        cand[nc++] = TK_KERNEL_MFMA | (2 << 4);
        cand[nc++] = TK_KERNEL_MFMA | (3 << 4);
    }
    if (outs <= 65536 && g->K >= 64) cand[nc++] = TK_KERNEL_SPLITK;
    tk_gett t = *g;
    gett_ctx ctx = {&t, ex->stream};
    double best_us = -1.0;
    *best = g->kernel;
    for (int i = 0; i < nc; i++) {
        apply_cfg(&t, cand[i], ks, kchunk, partial);
        const double us = tk_tune_time(ex, launch_gett_ctx, &ctx, 3);
        if (us < 0) return device_error("autotune launch", (int)hipGetLastError(), __LINE__);
        if (best_us < 0 || us < best_us) { best_us = us; *best = cand[i]; }
    }
    return 0;
}

/* Measuring reruns the product: only when rerunning gives the same D, i.e.
 * C is not read or is not the memory D is written to. */
static bool tune_safe(const tk_problem* P, const void* const* src, bool needC)
{
    if (P->beta_zero || needC || !src[2]) return true;
    const int64_t ec = (int64_t)tk_type_size(P->type[2]), ed = (int64_t)tk_type_size(P->type[3]);
    const char* c0 = (const char*)src[2] + P->lo[2] * ec;
    const char* c1 = (const char*)src[2] + (P->hi[2] + 1) * ec;
    const char* d0 = (const char*)src[3] + P->lo[3] * ed;
    const char* d1 = (const char*)src[3] + (P->hi[3] + 1) * ed;
    return c1 <= d0 || d1 <= c0;
}

static size_t ctype_size(int ct)
{
    switch (ct) {
    case TK_F32: return 4;
    case TK_F64: return 8;
    case TK_C32: return 8;
    case TK_C64: return 16;
    default: return 0;
    }
}

static bool compatible(int storage, int ct)
{
    return (storage == 0 && ct == TK_F32) || (storage == 1 && ct == TK_F64) || (storage == 2 && ct == TK_C32) ||
           (storage == 3 && ct == TK_C64);
}

/* Unit step over groups of `ve` consecutive rows and every vector start a
 * multiple of `ve`, with the other tables multiples of `ve` and the base
 * 16-byte aligned: 16-byte loads along the rows are then exact. */
static bool vec_rows_ok(const int64_t* rt, int64_t n, const int64_t* kt, int64_t nk, const int64_t* lt,
                        int64_t nl, int ve, const void* base)
{
    if (ve <= 1) return false;
    if (((uintptr_t)base) % 16) return false;
    for (int64_t r = 0; r < n; r += ve) {
        if (rt[r] % ve) return false;
        for (int j = 1; j < ve && r + j < n; j++)
            if (rt[r + j] != rt[r] + j) return false;
    }
    for (int64_t k = 0; k < nk; k++) if (kt[k] % ve) return false;
    for (int64_t l = 0; l < nl; l++) if (lt[l] % ve) return false;
    return true;
}

typedef struct {
    const void* base;   /* device address of element (0, ..., 0) */
    bool staged;
} dptr;

/* Kernel family for a planned product: `force` unless it is
 * TK_KERNEL_AUTO, else by shape.  Symmetric in M and N, so it holds after
 * the orientation swap. */
int tk_kernel_for(const tk_problem* P, int force)
{
    const int64_t M = P->nM, N = P->nN, K = P->nK, outs = P->nM * P->nN * P->nL;
    int kern = force;
    if (kern == TK_KERNEL_AUTO) {
        if (K == 0) kern = TK_KERNEL_SIMPLE;
        else if (outs <= 4096 && K >= 1024) kern = TK_KERNEL_SPLITK;
        else if (M >= 16 && N >= 16 && K >= 4 && outs * K >= (1 << 16)) kern = TK_KERNEL_MFMA;
        else kern = TK_KERNEL_SIMPLE;
    }
    if (kern == TK_KERNEL_SPLITK && outs > (int64_t)1 << 31) kern = TK_KERNEL_SIMPLE;
    return kern;
}

/* Power-of-two scale exponent for split-F16 from an operand's max |x|
 * (bits as the kernel reports them): 0 when binary16 already holds the
 * operand at full precision (max in [2^-3, 2^15): hi and, for the largest
 * elements, lo are normal numbers, and nothing overflows), else e with
 * max * 2^e in [2^14, 2^15).  Returns false for NaN / inf and for maxima
 * below 2^-100, which take the FP32 kernel. */
static bool f16_scale_exp(uint32_t bits, int* e)
{
    *e = 0;
    if (bits >= 0x7f800000u) return false;
    float m;
    memcpy(&m, &bits, sizeof m);
    if (m == 0.0f || (m >= 0x1p-3f && m < 0x1p15f)) return true;
    if (m < 0x1p-100f) return false;
    int ex2;
    frexpf(m, &ex2);            /* m = f 2^ex2, f in [0.5, 1) */
    *e = 15 - ex2;
    return true;
}

/* Split-F16 GETT: one launch at scale 1 that also reports max |A| and
 * max |B|; when an operand is outside binary16's comfortable range a second
 * launch either scales the operands by powers of two (alpha divided by the
 * same, exactly) or, for NaN / inf / extreme magnitudes, runs the FP32
 * kernel.  Synchronous; D is rewritten completely by a second launch and C
 * is intact (the caller checked). */
static int split_f16_gett(tk_gett* g, uint32_t* amax, hipStream_t st)
{
    uint32_t seen[2];
    g->scale[0] = g->scale[1] = 1.0f;
    g->amax = amax;
    HIPCHK(hipMemsetAsync(amax, 0, 2 * sizeof(uint32_t), st));
    LAUNCHCHK(tk_launch_gett(g, st));
    HIPCHK(hipMemcpyAsync(seen, amax, sizeof seen, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    int ea, eb;
    bool ok = f16_scale_exp(seen[0], &ea) && f16_scale_exp(seen[1], &eb);
    if (ok && ea == 0 && eb == 0) return 0;
    const double alpha = ok ? ldexp(g->alpha[0], -(ea + eb)) : g->alpha[0];
    if (ok && alpha != 0.0 && (fabs(alpha) > FLT_MAX || fabs(alpha) < FLT_MIN)) ok = false;
    if (ok) {
        g->scale[0] = ldexpf(1.0f, ea);
        g->scale[1] = ldexpf(1.0f, eb);
        g->alpha[0] = alpha;
    } else {
        g->mode = TK_MODE_PLAIN;
        g->amax = NULL;
    }
    LAUNCHCHK(tk_launch_gett(g, st));
    return 0;
}

int tk_execute(struct tapp_plan* p, struct tapp_exec* ex, const void* alpha, const void* A, const void* B,
               const void* beta, const void* C, void* D)
{
    return tk_execute_on(p, ex, alpha, A, B, beta, C, D, true);
}

/* Device memory of one call: NULL = the plan's cached tables and the
 * executor's workspace; otherwise one caller-provided region that holds the
 * tables and every buffer. */
typedef struct {
    char* base;
    size_t bytes;
} region;

static int execute_on(struct tapp_plan* p, struct tapp_exec* ex, const void* alpha, const void* A, const void* B,
                      const void* beta, const void* C, void* D, bool sync, const region* scr);

/* With sync == false every step stays ordered on ex->stream and the call
 * returns after enqueueing: host operands must stay valid, and D is written,
 * once the stream has reached this point.  The plan is locked while its
 * offset tables are built or read and the work is enqueued. */
int tk_execute_on(struct tapp_plan* p, struct tapp_exec* ex, const void* alpha, const void* A, const void* B,
                  const void* beta, const void* C, void* D, bool sync)
{
    pthread_mutex_lock(&p->mu);
    int rc = execute_on(p, ex, alpha, A, B, beta, C, D, sync, NULL);
    pthread_mutex_unlock(&p->mu);
    return rc;
}

/* One lock per device: calls of the allocation-free path share the device's
 * scratch array, so each holds it until its kernels have finished. */
#define NOALLOC_DEVICES 64
static pthread_mutex_t noalloc_mu[NOALLOC_DEVICES] = {[0 ... NOALLOC_DEVICES - 1] = PTHREAD_MUTEX_INITIALIZER};

int tk_execute_noalloc(struct tapp_plan* p, struct tapp_exec* ex, const void* alpha, const void* A, const void* B,
                       const void* beta, const void* C, void* D)
{
    if (!ex || ex->device < 0 || ex->device >= NOALLOC_DEVICES) return TAPP_AMD_ERR_INVALID;
    pthread_mutex_lock(&noalloc_mu[ex->device]);
    region scr = {NULL, 0};
    int rc = 0;
    if (ex->have_stream) {
        void* base = NULL;
        LAUNCHCHK_UNLOCK(tk_scratch_get(ex->device, &base, &scr.bytes));
        scr.base = (char*)base;
    }
    if (!rc) {
        pthread_mutex_lock(&p->mu);
        rc = execute_on(p, ex, alpha, A, B, beta, C, D, true, &scr);
        pthread_mutex_unlock(&p->mu);
    }
    pthread_mutex_unlock(&noalloc_mu[ex->device]);
    return rc;
}

/* Where a call's operands live and what device memory it needs beyond
 * them: host operands staged over their spans, pre-pass / conversion
 * buffers, split-K partial sums and the split-F16 range word, laid out in
 * one workspace.  Shared by execute_on and the serving loop's admission. */
typedef struct {
    const void* src[4];          /* device address of each operand (resident copies resolved) */
    bool on_dev[4];
    size_t span[4];              /* bytes staged from host memory */
    bool need[4];                /* pre-pass (A, B, C) or conversion (D) buffer */
    size_t tmp[4];
    int kern, mode;
    bool may_tune;
    int64_t ks, kchunk;
    size_t part;                 /* split-K partials */
    size_t off_stage[4], off_tmp[4], off_part, off_amax, total;
} work_layout;

static void work_offsets(work_layout* w)
{
    size_t total = 0;
    for (int x = 0; x < 4; x++) { w->off_stage[x] = total; total = ALIGN256(total + w->span[x]); }
    for (int x = 0; x < 4; x++) { w->off_tmp[x] = total; total = ALIGN256(total + w->tmp[x]); }
    w->off_part = total;
    total = ALIGN256(total + w->part);
    w->off_amax = total;
    if (w->mode == TK_MODE_SPLIT_F16) total = ALIGN256(total + 2 * sizeof(uint32_t));
    w->total = total;
}

static int plan_work(const struct tapp_plan* p, const struct tapp_exec* ex, const tk_problem* Pp, const void* A,
                     const void* B, const void* C, void* D, bool sync, bool scratch, work_layout* w)
{
    const tk_problem P = *Pp;
    const int64_t M = P.nM, N = P.nN, K = P.nK, L = P.nL;
    const int ct = P.ctype;
    const size_t cs = ctype_size(ct);
    memset(w, 0, sizeof *w);
    /* device operands are used in place, host ranges resident on the plan's
     * handle are read from their copies (not D: it is written through the
     * caller's pointer), other host operands are staged over their span */
    const void* ops[4] = {A, B, P.beta_zero ? NULL : C, D};
    for (int x = 0; x < 4; x++) {
        w->src[x] = ops[x];
        w->on_dev[x] = true;
        if (!ops[x]) continue;
        const int64_t es = (int64_t)tk_type_size(P.type[x]);
        const void* r = x < 3 ? tk_resolve_operand(p->handle, ops[x], P.lo[x] * es, (P.hi[x] + 1) * es)
                              : (tk_is_device_ptr(ops[x]) ? ops[x] : NULL);
        if (r) { w->src[x] = r; continue; }
        w->on_dev[x] = false;
        if (P.hi[x] >= P.lo[x]) w->span[x] = (size_t)(P.hi[x] - P.lo[x] + 1) * (size_t)es;
        if (scratch) {
            snprintf(last_error, sizeof last_error,
                     "allocation-free execute: operand %c is host memory (device or resident operands only)", "ABCD"[x]);
            return TAPP_AMD_ERR_INVALID;
        }
    }
    w->need[0] = P.nUA > 1 || !compatible(P.type[0], ct) || (P.op_A && P.nUA > 1);
    w->need[1] = P.nUB > 1 || !compatible(P.type[1], ct);
    w->need[2] = !P.beta_zero && !compatible(P.type[2], ct);
    w->need[3] = !compatible(P.type[3], ct);
    w->tmp[0] = w->need[0] ? (size_t)(M * K * L) * cs : 0;
    w->tmp[1] = w->need[1] ? (size_t)(N * K * L) * cs : 0;
    w->tmp[2] = w->need[2] ? (size_t)(M * N * L) * cs : 0;
    w->tmp[3] = w->need[3] ? (size_t)(M * N * L) * cs : 0;

    /* kernel choice (symmetric in M and N, so it holds after the orientation
     * swap in execute_on) */
    int kern = tk_kernel_for(&P, ex->force_kernel);
    w->may_tune = ex->autotune && ex->force_kernel == TK_KERNEL_AUTO && K > 0 && P.mode != TK_MODE_SPLIT_F16;

    /* split-F16 runs on the real FP32 MFMA GETT only: the other kernels,
     * and complex products, compute in FP32, which is as accurate.  Its
     * range check may ask for a second launch, which needs a synchronous
     * call whose C is still intact after the first (tune_safe); otherwise
     * the three-piece split-BF16 gives the same accuracy in one launch. */
    int mode = P.mode;
    if (mode == TK_MODE_SPLIT_F16) {
        if (kern != TK_KERNEL_MFMA || ct != TK_F32) mode = TK_MODE_PLAIN;
        else if (!sync || !tune_safe(&P, w->src, w->need[2])) mode = TK_MODE_SPLIT6_BF16;
    }
    /* few outputs and a moderate K (a 16^6 four-index product: 4 tiles of 128 x 128): 16 x
     * 16 tiles with K split over each workgroup's waves, in one launch,
     * instead of split-K's partial sums and second kernel */
    if (kern == TK_KERNEL_MFMA && ct == TK_F32 && mode == TK_MODE_PLAIN && ex->force_kernel == TK_KERNEL_AUTO) {
        const int64_t t128 = ((M + 127) / 128) * ((N + 127) / 128) * L;
        const int64_t t16 = ((M + 15) / 16) * ((N + 15) / 16) * L;
        if (t128 < 2 * (int64_t)ex->cus && K <= 1024 && t16 <= 4 * (int64_t)ex->cus) kern = TK_KERNEL_SMALL;
    }
    if (kern == TK_KERNEL_SMALL && (ct != TK_F32 || mode != TK_MODE_PLAIN)) kern = TK_KERNEL_MFMA;
    /* split-K for the FP32-data MFMA kernel when its 128 x 128 tiles cannot
     * fill the GPU: workgroups take K slices (multiples of its 32-deep K
     * tile) and store FP32 partial sums of whole tiles, which a second
     * kernel adds in a fixed order */
    int64_t ks = 1, kchunk = 0;
    size_t part = 0;
    if ((kern == TK_KERNEL_MFMA || w->may_tune) && ct == TK_F32 && K >= 64) {
        const int64_t tiles = ((M + 127) / 128) * ((N + 127) / 128) * L;
        const int64_t target = 2 * (int64_t)ex->cus;
        if (tiles < target) {
            ks = (target + tiles - 1) / tiles;
            if (ks > (K + 31) / 32) ks = (K + 31) / 32;
            if (ks > 64) ks = 64;
            kchunk = ((K + ks - 1) / ks + 31) / 32 * 32;
            ks = (K + kchunk - 1) / kchunk;
            part = sizeof(float) * (size_t)(ks * tiles * 128 * 128);   /* whole tiles */
            if (ks < 2 || part > ((size_t)64 << 20)) { ks = 1; kchunk = 0; part = 0; }
        }
    }
    w->kern = kern;
    w->mode = mode;
    w->ks = ks;
    w->kchunk = kchunk;
    w->part = part;
    work_offsets(w);
    return 0;
}

int tk_call_bytes(struct tapp_plan* p, const struct tapp_exec* ex, const tk_problem* P, const void* A,
                  const void* B, const void* C, void* D, size_t* work, size_t* tables, bool* devops)
{
    work_layout w;
    int rc = plan_work(p, ex, P, A, B, C, D, false, false, &w);
    if (rc) return rc;
    *work = w.total;
    *devops = w.on_dev[0] && w.on_dev[1] && w.on_dev[2] && w.on_dev[3];
    *tables = tk_plan_needs_tables(p, ex) ? table_layout(P, NULL) * sizeof(int64_t) : 0;
    return 0;
}

bool tk_plan_needs_tables(struct tapp_plan* p, const struct tapp_exec* ex)
{
    tk_problem P;
    if (tk_plan_problem(p, ex, NULL, NULL, NULL, &P)) return false;
    pthread_mutex_lock(&p->mu);
    const struct tapp_plan_cache* c = p->cache;
    const bool have = c && c->key == problem_key(&P) && c->device == ex->device && c->dev;
    pthread_mutex_unlock(&p->mu);
    return !have;
}

int tk_exec_reserve_ws(struct tapp_exec* ex, size_t bytes)
{
    if (bytes <= ex->ws_size) return 0;
    if (ex->ws) { hipStreamSynchronize(ex->stream); hipFree(ex->ws); ex->ws = NULL; ex->ws_size = 0; }
    hipError_t e = hipMalloc(&ex->ws, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ex->ws = NULL;
        device_error("hipMalloc(workspace)", (int)e, __LINE__);
        return TAPP_AMD_ERR_NOMEM;
    }
    ex->ws_size = bytes;
    return 0;
}

void tk_exec_release_ws(struct tapp_exec* ex)
{
    if (!ex->ws) return;
    hipStreamSynchronize(ex->stream);
    hipFree(ex->ws);
    (void)hipGetLastError();
    ex->ws = NULL;
    ex->ws_size = 0;
}

static int execute_on(struct tapp_plan* p, struct tapp_exec* ex, const void* alpha, const void* A, const void* B,
                      const void* beta, const void* C, void* D, bool sync, const region* scr)
{
    tk_problem P;
    int rc = tk_plan_problem(p, ex, alpha, beta, C, &P);
    if (rc) return rc;
    if (P.empty) return 0;
    if (!ex->have_stream) return device_error("executor without a device stream", 0, __LINE__);
    if (!A || !B || !D || (!P.beta_zero && !C)) return TAPP_AMD_ERR_INVALID;
    HIPCHK(hipSetDevice(ex->device));

    struct tapp_plan_cache* tc;
    const int64_t* T;
    size_t tab_bytes = 0;
    if (scr) {
        if ((rc = build_host_tables(p, &P, &tc))) return rc;
        tab_bytes = ALIGN256(tc->total * sizeof(int64_t));
        T = (const int64_t*)scr->base;
    } else {
        if ((rc = build_tables(p, &P, ex->device, &tc))) return rc;
        T = tc->dev;
    }
    const int64_t* H = tc->host;
#define DT(t) (T + tc->off[t])
#define HT(t) (H + tc->off[t])

    const int64_t M = P.nM, N = P.nN, K = P.nK, L = P.nL;
    const int ct = P.ctype;
    const size_t cs = ctype_size(ct);

    work_layout w;
    if ((rc = plan_work(p, ex, &P, A, B, C, D, sync, scr != NULL, &w))) return rc;
    const void* const* src = w.src;
    const bool* on_dev = w.on_dev;
    const size_t* span_bytes = w.span;
    const bool needA = w.need[0], needB = w.need[1], needC = w.need[2], needD = w.need[3];
    const int64_t outs = M * N * L;
    int kern = w.kern;
    const int mode = w.mode;
    const bool may_tune = w.may_tune;
    int64_t ks = w.ks, kchunk = w.kchunk;
    /* scratch: split-K gives way first (its partials are the only optional
     * buffer), then the call is refused */
    if (scr && tab_bytes + w.total > scr->bytes && w.part) {
        w.part = 0;
        ks = 1;
        kchunk = 0;
        work_offsets(&w);
    }
    if (scr && tab_bytes + w.total > scr->bytes) {
        snprintf(last_error, sizeof last_error,
                 "allocation-free execute: needs %zu bytes of scratch, the device scratch holds %zu",
                 tab_bytes + w.total, scr->bytes);
        return TAPP_AMD_ERR_SCRATCH;
    }
    const size_t* off_stage = w.off_stage;
    const size_t* off_tmp = w.off_tmp;
    const size_t off_part = w.off_part, off_amax = w.off_amax, total = w.total;
    hipStream_t st = ex->stream;
    char* ws;
    if (scr) {
        ws = scr->base + tab_bytes;
        HIPCHK(hipMemcpyAsync(scr->base, tc->host, tc->total * sizeof(int64_t), hipMemcpyHostToDevice, st));
    } else {
        if (total && (rc = ensure_ws(ex, total))) return rc;
        ws = (char*)ex->ws;
    }

    /* operands on the device: element (0, ..., 0) addresses */
    dptr d[4];
    for (int x = 0; x < 4; x++) {
        d[x].staged = false;
        d[x].base = src[x];
        if (!src[x]) continue;
        const size_t es = tk_type_size(P.type[x]);
        if (!on_dev[x]) {
            if (span_bytes[x] == 0) { d[x].base = ws + off_stage[x]; continue; }
            char* buf = ws + off_stage[x];
            const char* hsrc = (const char*)src[x] + P.lo[x] * (int64_t)es;
            /* D is copied in too so elements the product does not write keep their values */
            HIPCHK(hipMemcpyAsync(buf, hsrc, span_bytes[x], hipMemcpyHostToDevice, st));
            d[x].base = buf - P.lo[x] * (int64_t)es;
            d[x].staged = true;
        }
        d[x].base = (const char*)d[x].base + P.pin[x] * (int64_t)es;
    }

    if (ex->timing) hipEventRecord(ex->ev0, st);

    /* pre-passes */
    tk_gett g;
    memset(&g, 0, sizeof g);
    g.ctype = ct;
    g.mode = mode;
    g.M = M; g.N = N; g.K = K; g.L = L;
    g.conjA = P.op_A; g.conjB = P.op_B; g.conjC = P.op_C; g.conjD = P.op_D;
    g.conjD_via_f32 = P.conjD_via_f32;
    g.round_C = P.round_C;
    g.alpha[0] = P.alpha[0]; g.alpha[1] = P.alpha[1];
    g.beta[0] = P.beta[0]; g.beta[1] = P.beta[1];
    g.beta_zero = P.beta_zero;

    g.A = d[0].base; g.mA = DT(T_mA); g.kA = DT(T_kA); g.lA = DT(T_lA);
    g.B = d[1].base; g.nB = DT(T_nB); g.kB = DT(T_kB); g.lB = DT(T_lB);
    g.C = d[2].base; g.mC = DT(T_mC); g.nC = DT(T_nC); g.lC = DT(T_lC);
    g.D = (void*)d[3].base; g.mD = DT(T_mD); g.nD = DT(T_nD); g.lD = DT(T_lD);
    const int64_t *hmA = HT(T_mA), *hkA = HT(T_kA), *hlA = HT(T_lA);
    const int64_t *hnB = HT(T_nB), *hkB = HT(T_kB), *hlB = HT(T_lB);

    /* A's and B's pre-passes (one launch when both are needed) */
    tk_prep prA, prB;
    if (needA && K > 0) {
        const int64_t* tabs[3] = {DT(T_mA), DT(T_kA), DT(T_lA)};
        const int64_t sizes[3] = {M, K, L};
        const int* o = tc->ordA;
        prA = (tk_prep){P.type[0], ct, P.op_A, P.sum_round[0],
                        d[0].base, ws + off_tmp[0], sizes[o[0]], sizes[o[1]], sizes[o[2]], P.nUA,
                        tabs[o[0]], tabs[o[1]], tabs[o[2]], DT(T_uA)};
        g.A = ws + off_tmp[0]; g.mA = DT(T_dmA); g.kA = DT(T_dkA); g.lA = DT(T_dlA); g.conjA = 0;
        hmA = HT(T_dmA); hkA = HT(T_dkA); hlA = HT(T_dlA);
    }
    if (needB && K > 0) {
        const int64_t* tabs[3] = {DT(T_nB), DT(T_kB), DT(T_lB)};
        const int64_t sizes[3] = {N, K, L};
        const int* o = tc->ordB;
        prB = (tk_prep){P.type[1], ct, P.op_B, P.sum_round[1],
                        d[1].base, ws + off_tmp[1], sizes[o[0]], sizes[o[1]], sizes[o[2]], P.nUB,
                        tabs[o[0]], tabs[o[1]], tabs[o[2]], DT(T_uB)};
        g.B = ws + off_tmp[1]; g.nB = DT(T_dnB); g.kB = DT(T_dkB); g.lB = DT(T_dlB); g.conjB = 0;
        hnB = HT(T_dnB); hkB = HT(T_dkB); hlB = HT(T_dlB);
    }
    if (needA && needB && K > 0) LAUNCHCHK(tk_launch_prep2(&prA, &prB, st));
    else if (needA && K > 0) LAUNCHCHK(tk_launch_prep(&prA, st));
    else if (needB && K > 0) LAUNCHCHK(tk_launch_prep(&prB, st));
    if (needC) {
        /* a C rounded to 16 bits is rounded once from its storage value */
        tk_prep pr = {P.type[2], ct, 0, P.round_C ? P.mode : 0, d[2].base, ws + off_tmp[2], M, N, L, 1,
                      DT(T_mC), DT(T_nC), DT(T_lC), DT(T_zero)};
        LAUNCHCHK(tk_launch_prep(&pr, st));
        g.C = ws + off_tmp[2]; g.mC = DT(T_dmD); g.nC = DT(T_dnD); g.lC = DT(T_dlD);
    }
    if (needD) {
        g.D = ws + off_tmp[3]; g.mD = DT(T_dmD); g.nD = DT(T_dnD); g.lD = DT(T_dlD);
    }

    /* orientation: output lanes run along M; swap the roles of A and B when
     * D's unit-stride mode belongs to N */
    {
        int64_t dm = INT64_MAX, dn = INT64_MAX;
        for (int i = 0; i < P.M.n; i++) if (llabs(P.M.s[3][i]) < dm) dm = llabs(P.M.s[3][i]);
        for (int i = 0; i < P.N.n; i++) if (llabs(P.N.s[3][i]) < dn) dn = llabs(P.N.s[3][i]);
        if (needD) { dm = 1; dn = M; }
        if (dn < dm) {
            tk_gett s = g;
            s.M = N; s.N = M;
            s.A = g.B; s.mA = g.nB; s.kA = g.kB; s.lA = g.lB; s.conjA = g.conjB;
            s.B = g.A; s.nB = g.mA; s.kB = g.kA; s.lB = g.lA; s.conjB = g.conjA;
            s.mC = g.nC; s.nC = g.mC; s.mD = g.nD; s.nD = g.mD;
            g = s;
            const int64_t *t1 = hmA, *t2 = hkA, *t3 = hlA;
            hmA = hnB; hkA = hkB; hlA = hlB;
            hnB = t1; hkB = t2; hlB = t3;
        }
    }

    g.kernel = kern;
    g.cus = ex->cus;
    g.tile_variant = ex->tiles;

    /* tile access patterns for the MFMA kernel */
    if (kern == TK_KERNEL_MFMA || may_tune) {
        const int ve = (int)(16 / cs) > 0 ? (int)(16 / cs) : 1;
        if (vec_rows_ok(hmA, g.M, hkA, g.K, hlA, g.L, ve, g.A)) { g.a_load = TK_CONTIG_ROW; g.a_vec = 1; }
        else if (vec_rows_ok(hkA, g.K, hmA, g.M, hlA, g.L, ve, g.A)) { g.a_load = TK_CONTIG_K; g.a_vec = 1; }
        else { g.a_load = TK_GATHER; g.a_vec = 0; }
        if (vec_rows_ok(hnB, g.N, hkB, g.K, hlB, g.L, ve, g.B)) { g.b_load = TK_CONTIG_ROW; g.b_vec = 1; }
        else if (vec_rows_ok(hkB, g.K, hnB, g.N, hlB, g.L, ve, g.B)) { g.b_load = TK_CONTIG_K; g.b_vec = 1; }
        else { g.b_load = TK_GATHER; g.b_vec = 0; }
    }

    if (kern == TK_KERNEL_MFMA && ks > 1) {
        g.ksplit = (int)ks;
        g.kchunk = kchunk;
        g.partial = (float*)(ws + off_part);
    }

    /* autotuned shape class (opt-in): a cached choice replaces the rules; an
     * unknown class is measured here when that is safe (synchronous call; C
     * not overlapping D unless beta is zero or C was converted) */
    if (may_tune) {
        const int64_t f[] = {1, ct, P.mode, tk_tune_bucket(g.M), tk_tune_bucket(g.N), tk_tune_bucket(K),
                             tk_tune_bucket(L), g.a_load, g.b_load, P.beta_zero, needD};
        const uint64_t key = tk_tune_key(f, (int)(sizeof f / sizeof f[0]));
        int cfg = 0;
        bool have = tk_tune_get(ex->device, key, &cfg);
        if (!have && ex->autotune == 1 && sync && !scr && tune_safe(&P, src, needC)) {
            if ((rc = tune_gett(ex, &g, ks, kchunk, (float*)(ws + off_part), outs, &cfg))) return rc;
            char desc[160];
            snprintf(desc, sizeof desc, "gett ct=%d mode=%d M=%lld N=%lld K=%lld L=%lld load=%d,%d beta0=%d",
                     ct, P.mode, (long long)g.M, (long long)g.N, (long long)K, (long long)L, g.a_load, g.b_load,
                     (int)P.beta_zero);
            tk_tune_put(ex->device, key, cfg, desc);
            have = true;
        }
        if (have) {
            apply_cfg(&g, cfg, ks, kchunk, (float*)(ws + off_part));
            kern = g.kernel;
        }
    }
    if (mode == TK_MODE_SPLIT_F16) {
        if ((rc = split_f16_gett(&g, (uint32_t*)(ws + off_amax), st))) return rc;
    } else {
        LAUNCHCHK(tk_launch_gett(&g, st));
    }
    ex->last_kernel = kern;
    snprintf(ex->last_kernel_name, sizeof ex->last_kernel_name, "%s", tk_last_gett_name());

    if (needD) {
        tk_scatter sc = {ct, P.type[3], ws + off_tmp[3], (void*)d[3].base, M, N, L,
                         DT(T_mD), DT(T_nD), DT(T_lD)};
        LAUNCHCHK(tk_launch_scatter(&sc, st));
    }
    if (ex->timing) hipEventRecord(ex->ev1, st);

    if (d[3].staged) {
        const size_t es = tk_type_size(P.type[3]);
        char* hdst = (char*)D + P.lo[3] * (int64_t)es;
        HIPCHK(hipMemcpyAsync(hdst, ws + off_stage[3], span_bytes[3], hipMemcpyDeviceToHost, st));
    }
    if (!sync) return 0;
    HIPCHK(hipStreamSynchronize(st));
    if (ex->timing) {
        float ms = 0;
        hipEventElapsedTime(&ms, ex->ev0, ex->ev1);
        ex->last_device_us = ms * 1000.0;
    }
    return 0;
#undef DT
#undef HT
}
