/*
 * Product groups: many independent products, each with its own plan and
 * operands, executed as one persistent launch (csrc/hip/tk_group.hip).
 *
 * The reference declares TAPP_execute_batched_product (product.h:66-75: one
 * plan, num_batches operand sets) but defines nothing; uneven batches have
 * no reference API at all.  A group is prepared once -- every product
 * validated with the codes of TAPP_execute_product, classified, its offset
 * tables cached, the descriptors uploaded and the zero-counter + persistent
 * kernel pair captured in a hipGraph -- and then replayed.  Products the
 * persistent kernel does not cover (non-FP32 compute, unary labels,
 * conversion pre-passes, host operands) run one by one after it, through the
 * same path as TAPP_execute_product.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

/* tile queues of a product group: one per XCD of an MI355X */
#define GROUP_QUEUES 8

/* One persistent launch of a group: the products of one tile class (their
 * descriptors a contiguous run of the group's), with its own tile queues. */
typedef struct {
    int kind;                  /* TK_GROUP_KIND_THIN or TK_GROUP_KIND_MFMA */
    int n;                     /* products */
    int64_t total;             /* tiles */
    int nblocks;
    tk_prob* dprobs;
    int64_t* dtile_end;        /* inclusive tile prefix over the launch's products */
    tk_tile* dtiles;           /* every tile decoded: product of the launch, l, ts, tg (or NULL) */
    int64_t* dqbeg;            /* first tile of each queue, GROUP_QUEUES + 1 entries */
    unsigned long long* dcounter;   /* one counter per queue, 256 bytes apart */
} group_launch;

struct tapp_group {
    uint32_t magic;
    struct tapp_exec* ex;
    int n;                     /* products in the persistent launches */
    int64_t total;             /* their tiles */
    void* dmem;                /* descriptors | per launch: tile ends, tile map, queue starts, counters */
    tk_prob* dprobs;           /* all n descriptors: the thin launch's, then the MFMA launch's */
    group_launch launch[2];
    int nlaunch;
    hipGraph_t graph;
    hipGraphExec_t gexec;
    bool have_graph;
    /* kept for a batched call's reuse: the descriptors in launch order
     * (pinned host memory) and the operand set each came from */
    tk_prob* hprobs;
    int* hsrc;
    /* products executed one by one */
    int nrest;
    struct tapp_plan** rest_plan;
    const void **rest_A, **rest_B, **rest_C;
    void** rest_D;
    /* alpha and beta of each, copied at creation (read as type_D: at most
     * 16 bytes, a binary64 complex) */
    unsigned char (*rest_scalar)[2][16];
};


/* How the persistent kernel reads Y: 0 = along columns (unit step across
 * g), 1 = along K (unit step in k), 2 = dense (k[i] = i and g[j] = j * K:
 * each tile of columns is one contiguous block). */
static int y_access(const int64_t* kt, int64_t K, const int64_t* gt, int64_t G)
{
    if (K <= 1) return 0;
    if (K >= ((int64_t)1 << 21)) return kt[1] - kt[0] == 1 ? 1 : 0;   /* the dense path's 32-bit tile offsets */
    for (int64_t k = 0; k < K; k++)
        if (kt[k] != kt[0] + k) return kt[1] - kt[0] == 1 ? 1 : 0;
    for (int64_t g = 0; g < G; g++)
        if (gt[g] != g * K) return 1;
    return kt[0] == 0 ? 2 : 1;
}

/* Whether operand x's offset table over group g (first mode fastest) is
 * index * stride: each mode's stride is the previous one times its extent.
 * The stride goes to *st (0 for an empty group). */
static bool affine_group(const tk_group* g, int x, int64_t* st)
{
    *st = g->n > 0 ? g->s[x][0] : 0;
    for (int i = 1; i < g->n; i++)
        if (g->s[x][i] != g->s[x][i - 1] * g->ext[i - 1]) return false;
    return true;
}

/* Thin tiles over a dense Y: the strides that replace X's, D's and C's
 * tables and the Hadamard tables when all of them are affine. */
static void describe_affine(const tk_problem* P, bool x_is_b, tk_prob* out)
{
    out->affine = 0;
    if (out->cls != TK_GROUP_THIN || out->y_kfast != 2) return;
    const int xi = x_is_b ? 1 : 0, yi = 1 - xi;
    const tk_group* gs = x_is_b ? &P->N : &P->M;     /* X's rows: D's s */
    const tk_group* gg = x_is_b ? &P->M : &P->N;     /* Y's columns: D's g */
    int64_t v[10] = {0};
    bool ok = affine_group(gs, xi, &v[0]) && affine_group(&P->K, xi, &v[1]) && affine_group(&P->L, xi, &v[2]) &&
              affine_group(&P->L, yi, &v[3]) && affine_group(gs, 3, &v[4]) && affine_group(gg, 3, &v[5]) &&
              affine_group(&P->L, 3, &v[6]);
    if (!P->beta_zero)
        ok = ok && affine_group(gs, 2, &v[7]) && affine_group(gg, 2, &v[8]) && affine_group(&P->L, 2, &v[9]);
    if (!ok) return;
    out->ssX = v[0]; out->skX = v[1]; out->slX = v[2]; out->slY = v[3];
    out->ssD = v[4]; out->sgD = v[5]; out->slD = v[6];
    out->ssC = v[7]; out->sgC = v[8]; out->slC = v[9];
    out->affine = 1;
}

typedef struct { tk_prob p; double cost; int idx; } ranked;

static int by_cost(const void* x, const void* y)
{
    const ranked* a = x; const ranked* b = y;
    return a->cost > b->cost ? -1 : a->cost < b->cost;
}

bool tk_group_may_run(const tk_problem* P)
{
    if (P->ctype != TK_F32 || P->mode == TK_MODE_SPLIT_BF16 || P->mode == TK_MODE_SPLIT6_BF16 ||
        P->mode == TK_MODE_SPLIT_F16)
        return false;
    if (P->nUA != 1 || P->nUB != 1) return false;
    for (int x = 0; x < 4; x++)
        if (P->type[x] != 0 && !(x == 2 && P->beta_zero)) return false;
    return true;
}

/* Descriptor of one eligible product; returns false when the product needs
 * the general path. */
bool tk_group_describe(struct tapp_plan* plan, struct tapp_exec* ex, const tk_problem* P, const void* A,
                       const void* B, const void* C, void* D, int force_cls, tk_prob* out, int* rc)
{
    *rc = 0;
    if (!tk_group_may_run(P)) return false;
    /* device operands, or host ranges resident on the plan's handle (A, B, C) */
    A = tk_resolve_operand(plan->handle, A, P->lo[0] * 4, (P->hi[0] + 1) * 4);
    B = tk_resolve_operand(plan->handle, B, P->lo[1] * 4, (P->hi[1] + 1) * 4);
    if (!P->beta_zero) C = tk_resolve_operand(plan->handle, C, P->lo[2] * 4, (P->hi[2] + 1) * 4);
    if (!A || !B || !tk_is_device_ptr(D) || (!P->beta_zero && !C)) return false;
    const int64_t *mA, *kA, *lA, *nB, *kB, *lB, *mD, *nD, *lD, *mC, *nC, *lC;
    if ((*rc = tk_problem_tables(plan, P, ex->device, &mA, &kA, &lA, &nB, &kB, &lB, &mD, &nD, &lD, &mC, &nC, &lC)))
        return false;
    const int64_t *hkA, *hkB, *hmA, *hnB;
    tk_problem_host_k(plan, &hkA, &hkB);
    tk_problem_host_mn(plan, &hmA, &hnB);
    memset(out, 0, sizeof *out);
    const float* a = (const float*)A + P->pin[0];
    const float* b = (const float*)B + P->pin[1];
    const float* c = C ? (const float*)C + P->pin[2] : NULL;
    float* d = (float*)D + P->pin[3];
    const int64_t M = P->nM, N = P->nN;
    int64_t dm = INT64_MAX, dn = INT64_MAX;
    for (int i = 0; i < P->M.n; i++) if (llabs(P->M.s[3][i]) < dm) dm = llabs(P->M.s[3][i]);
    for (int i = 0; i < P->N.n; i++) if (llabs(P->N.s[3][i]) < dn) dn = llabs(P->N.s[3][i]);
    bool x_is_b;
    const bool thin_ok = N <= 16 || M <= 16;
    if (thin_ok && force_cls != TK_GROUP_MFMA) {
        out->cls = TK_GROUP_THIN;
        x_is_b = N <= M;            /* the small side is X */
    } else {
        out->cls = TK_GROUP_MFMA;
        x_is_b = dn < dm;           /* MFMA lanes run along X's rows: put D's unit stride there */
    }
    if (x_is_b) {
        out->S = N; out->G = M;
        out->X = b; out->sX = nB; out->kX = kB; out->lX = lB;
        out->Y = a; out->gY = mA; out->kY = kA; out->lY = lA;
        out->sD = nD; out->gD = mD; out->sC = nC; out->gC = mC;
        out->y_kfast = y_access(hkA, P->nK, hmA, M);
    } else {
        out->S = M; out->G = N;
        out->X = a; out->sX = mA; out->kX = kA; out->lX = lA;
        out->Y = b; out->gY = nB; out->kY = kB; out->lY = lB;
        out->sD = mD; out->gD = nD; out->sC = mC; out->gC = nC;
        out->y_kfast = y_access(hkB, P->nK, hnB, N);
    }
    out->lD = lD; out->lC = lC;
    out->x_is_b = x_is_b;
    out->K = P->nK; out->L = P->nL;
    if (out->cls == TK_GROUP_MFMA) {
        /* 32-bit offsets from the lowest element of each operand */
        const int xi = x_is_b ? 1 : 0, yi = 1 - xi;
        const int64_t lim = (int64_t)1 << 29;
        if (P->hi[xi] - P->lo[xi] >= lim || P->hi[yi] - P->lo[yi] >= lim) return false;
        out->xlo = P->lo[xi] - P->pin[xi];
        out->ylo = P->lo[yi] - P->pin[yi];
    }
    out->C = c; out->D = d;
    out->alpha = (float)P->alpha[0];
    out->beta = (float)P->beta[0];
    out->beta_zero = P->beta_zero;
    out->round = P->mode;
    out->round_C = P->round_C;
    if (out->cls == TK_GROUP_THIN) { out->tiles_s = 1; out->tiles_g = (out->G + 255) / 256; }
    else { out->tiles_s = (out->S + TK_GROUP_MT - 1) / TK_GROUP_MT; out->tiles_g = (out->G + TK_GROUP_MT - 1) / TK_GROUP_MT; }
    describe_affine(P, x_is_b, out);
    return true;
}

static double tile_cost(const tk_prob* d)
{
    const double per_tile = d->cls == TK_GROUP_THIN ? 256.0 * (double)d->K * (double)(d->S + 1)
                                                    : (double)TK_GROUP_MT * TK_GROUP_MT * (double)d->K;
    return per_tile + 1.0;
}

/* --- autotuning: thin FMA tiles or MFMA tiles per shape class ----------- */

typedef struct {
    tk_prob* probs;
    int64_t* tile_end;
    unsigned long long* counter;
    int n;
    int64_t total;
    int nblocks;
    hipStream_t st;
} group_ctx;

static int launch_group_ctx(void* c)
{
    group_ctx* x = (group_ctx*)c;
    return tk_launch_group(x->probs, x->tile_end, NULL, x->n, x->total, x->counter, x->nblocks, x->st);
}

/* Device time of one persistent launch over `n` descriptors (in the order
 * given; the caller sorts them as the group will run them). */
static double time_descriptors(struct tapp_exec* ex, const tk_prob* h, int n)
{
    const int64_t nb = TK_GROUP_WG_PER_CU * (int64_t)ex->cus;
    int64_t* he = malloc(sizeof(int64_t) * (size_t)n);
    if (!he) return -1.0;
    int64_t total = 0;
    for (int i = 0; i < n; i++) { total += h[i].L * h[i].tiles_s * h[i].tiles_g; he[i] = total; }
    const size_t pb = sizeof(tk_prob) * (size_t)n, o_end = (pb + 255) & ~(size_t)255;
    const size_t o_cnt = (o_end + sizeof(int64_t) * (size_t)n + 255) & ~(size_t)255;
    void* dm = NULL;
    double us = -1.0;
    if (hipMalloc(&dm, o_cnt + 256) == hipSuccess &&
        hipMemcpy(dm, h, pb, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy((char*)dm + o_end, he, sizeof(int64_t) * (size_t)n, hipMemcpyHostToDevice) == hipSuccess) {
        group_ctx ctx = {(tk_prob*)dm, (int64_t*)((char*)dm + o_end), (unsigned long long*)((char*)dm + o_cnt), n,
                         total, (int)(total < nb ? total : nb), ex->stream};
        us = tk_tune_time(ex, launch_group_ctx, &ctx, 3);
    }
    (void)hipGetLastError();
    if (dm) { (void)hipStreamSynchronize(ex->stream); (void)hipFree(dm); }
    free(he);
    return us;
}

/* Measuring runs every product again: only without C that is D. */
static bool group_tune_safe(const tk_prob* d)
{
    if (d->beta_zero || !d->C) return true;
    return d->C != (const float*)d->D;
}

/* The group's launch with the products whose small-side bucket is in
 * `mask` on MFMA tiles (their MFMA descriptors in alt[]), sorted by cost as
 * create_group will order them. */
static double time_mask(struct tapp_exec* ex, const ranked* list, const tk_prob* alt, const int* bucket, int n,
                        unsigned mask, ranked* work, tk_prob* flat)
{
    for (int i = 0; i < n; i++) {
        work[i] = list[i];
        if (bucket[i] >= 0 && (mask >> bucket[i]) & 1u) {
            work[i].p = alt[i];
            work[i].cost = tile_cost(&alt[i]);
        }
    }
    qsort(work, (size_t)n, sizeof(ranked), by_cost);
    for (int i = 0; i < n; i++) flat[i] = work[i].p;
    return time_descriptors(ex, flat, n);
}

/* Thin FMA tiles or MFMA tiles for products whose smaller side allows both,
 * chosen per shape class -- the power-of-two bucket of the smaller side --
 * in the context of the whole group: starting from the rule (thin tiles),
 * each bucket in turn is moved to MFMA tiles and kept there if the whole
 * launch gets at least 3 % faster.  The result is cached under the group's
 * mix of shape classes (each product's small side, large side, K and L
 * buckets), so a group with the same mix is not measured again. */
static int tune_group(struct tapp_exec* ex, ranked* list, int n, const TAPP_tensor_product* plans,
                      const void* const* alpha, const void* const* A, const void* const* B,
                      const void* const* beta, const void* const* C, void* const* D)
{
    int* bucket = malloc(sizeof(int) * (size_t)n);
    uint64_t* cls = malloc(sizeof(uint64_t) * (size_t)n);
    if (!bucket || !cls) { free(bucket); free(cls); return TAPP_AMD_ERR_NOMEM; }
    int nthin = 0;
    bool safe = true;
    unsigned present = 0;
    for (int i = 0; i < n; i++) {
        const tk_prob* d = &list[i].p;
        bucket[i] = -1;
        if (!group_tune_safe(d)) safe = false;
        if (d->cls != TK_GROUP_THIN) continue;
        bucket[i] = (int)tk_tune_bucket(d->S);
        present |= 1u << bucket[i];
        const int64_t f[] = {tk_tune_bucket(d->S), tk_tune_bucket(d->G), tk_tune_bucket(d->K), tk_tune_bucket(d->L),
                             d->round};
        cls[nthin++] = tk_tune_key(f, (int)(sizeof f / sizeof f[0]));
    }
    int rc = 0;
    if (nthin == 0) goto out;
    /* key: the sorted multiset of thin-capable shape classes */
    for (int i = 1; i < nthin; i++)
        for (int j = i; j > 0 && cls[j] < cls[j - 1]; j--) { uint64_t t = cls[j]; cls[j] = cls[j - 1]; cls[j - 1] = t; }
    uint64_t mix = 1469598103934665603ull;
    for (int i = 0; i < nthin; i++) { mix ^= cls[i]; mix *= 1099511628211ull; }
    const int64_t kf[] = {3, (int64_t)mix, nthin};
    const uint64_t key = tk_tune_key(kf, 3);
    int mask = 0;
    if (!tk_tune_get(ex->device, key, &mask)) {
        if (ex->autotune != 1 || !safe) goto out;
        tk_prob* alt = malloc(sizeof(tk_prob) * (size_t)n);
        ranked* work = malloc(sizeof(ranked) * (size_t)n);
        tk_prob* flat = malloc(sizeof(tk_prob) * (size_t)n);
        if (!alt || !work || !flat) { free(alt); free(work); free(flat); rc = TAPP_AMD_ERR_NOMEM; goto out; }
        for (int i = 0; i < n && !rc; i++) {
            if (bucket[i] < 0) continue;
            const int o = list[i].idx;
            struct tapp_plan* p = (struct tapp_plan*)plans[o];
            const void* c = C ? C[o] : NULL;
            tk_problem P;
            int prc = tk_plan_problem(p, ex, alpha[o], beta[o], c, &P);
            if (!prc && !tk_group_describe(p, ex, &P, A[o], B[o], c, D[o], TK_GROUP_MFMA, &alt[i], &prc) && !prc)
                bucket[i] = -1;
            rc = prc;
        }
        double best = rc ? -1.0 : time_mask(ex, list, alt, bucket, n, 0u, work, flat);
        if (!rc && best < 0) rc = TAPP_AMD_ERR_DEVICE;
        double t0 = best;
        for (int b = 31; b >= 0 && !rc; b--) {
            if (!((present >> b) & 1u)) continue;
            const double t = time_mask(ex, list, alt, bucket, n, (unsigned)mask | (1u << b), work, flat);
            if (t < 0) { rc = TAPP_AMD_ERR_DEVICE; break; }
            if (t < 0.97 * best) { best = t; mask |= 1 << b; }
        }
        free(alt); free(work); free(flat);
        if (rc) goto out;
        char desc[160];
        snprintf(desc, sizeof desc, "group mix=%016llx thin-capable=%d rule=%.1fus tuned=%.1fus mfma-buckets=0x%x",
                 (unsigned long long)mix, nthin, t0, best, (unsigned)mask);
        tk_tune_put(ex->device, key, mask, desc);
    }
    /* products of the chosen buckets on MFMA tiles */
    for (int i = 0; i < n && !rc && mask; i++) {
        if (bucket[i] < 0 || !(((unsigned)mask >> bucket[i]) & 1u)) continue;
        const int o = list[i].idx;
        struct tapp_plan* p = (struct tapp_plan*)plans[o];
        const void* c = C ? C[o] : NULL;
        tk_problem P;
        tk_prob d;
        int prc = tk_plan_problem(p, ex, alpha[o], beta[o], c, &P);
        if (!prc && tk_group_describe(p, ex, &P, A[o], B[o], c, D[o], TK_GROUP_MFMA, &d, &prc)) {
            list[i].p = d;
            list[i].cost = tile_cost(&d);
        }
        rc = prc;
    }
out:
    free(bucket);
    free(cls);
    return rc;
}

static void group_free(struct tapp_group* g)
{
    if (!g) return;
    if (g->have_graph) { hipGraphExecDestroy(g->gexec); hipGraphDestroy(g->graph); }
    if (g->dmem) hipFree(g->dmem);
    if (g->hprobs) hipHostFree(g->hprobs);
    free(g->hsrc);
    free(g->rest_plan); free(g->rest_A); free(g->rest_B);
    free(g->rest_C); free(g->rest_D); free(g->rest_scalar);
    free(g);
}

static int create_group_on(TAPP_AMD_product_group* group, struct tapp_exec* ex, int count,
                           const TAPP_tensor_product* plans, const void* const* alpha, const void* const* A,
                           const void* const* B, const void* const* beta, const void* const* C, void* const* D,
                           bool capture, bool keep_host);

static int create_group(TAPP_AMD_product_group* group, TAPP_executor exec, int count,
                        const TAPP_tensor_product* plans, const void* const* alpha, const void* const* A,
                        const void* const* B, const void* const* beta, const void* const* C, void* const* D,
                        bool capture)
{
    if (!group || count < 0 || (count > 0 && (!plans || !alpha || !A || !B || !beta || !D)))
        return TAPP_AMD_ERR_INVALID;
    struct tapp_exec* ex = tk_exec_lookup(exec);
    if (!ex) ex = (exec && (*(const int*)exec == 1 || *(const int*)exec == 2 || *(const int*)exec == 12))
                      ? tk_default_exec() : NULL;
    if (!ex) return 33;
    tk_exec_lock(ex);
    int rc = create_group_on(group, ex, count, plans, alpha, A, B, beta, C, D, capture, false);
    tk_exec_unlock(ex);
    return rc;
}

static int launch_all(struct tapp_group* g, hipStream_t st)
{
    for (int i = 0; i < g->nlaunch; i++) {
        const group_launch* L = &g->launch[i];
        int rc = tk_launch_group_kind(L->kind, L->dprobs, L->dtile_end, L->dtiles, L->n, L->total, L->dqbeg,
                                      GROUP_QUEUES, L->dcounter, L->nblocks, st);
        if (rc) return rc;
    }
    return 0;
}

/* Lays the ranked products (cost order) out on the device as up to two
 * persistent launches -- thin tiles, then MFMA tiles, each kernel at its
 * own occupancy -- and captures them in a graph.  Within a launch every
 * product goes whole to the XCD queue with the least work so far, so its
 * tiles share one L2; the queues are concatenated, each in cost order. */
static int build_launches(struct tapp_group* g, struct tapp_exec* ex, const ranked* list, int n, bool capture,
                          bool keep_host)
{
    const size_t pb = sizeof(tk_prob) * (size_t)n, tb = sizeof(int64_t) * (size_t)n;
    tk_prob* hp = malloc(pb);
    int64_t* he = malloc(tb);
    int* bin = malloc(sizeof(int) * (size_t)n);
    int* src = malloc(sizeof(int) * (size_t)n);
    int* ord = malloc(sizeof(int) * (size_t)n);
    tk_tile* hm = NULL;
    int rc = 0;
    if (!hp || !he || !bin || !src || !ord) { rc = TAPP_AMD_ERR_NOMEM; goto out; }
    /* thin products first, then MFMA products, each run in cost order */
    int nthin = 0;
    for (int i = 0; i < n; i++) if (list[i].p.cls == TK_GROUP_THIN) ord[nthin++] = i;
    for (int i = 0, k = nthin; i < n; i++) if (list[i].p.cls != TK_GROUP_THIN) ord[k++] = i;
    const int first[3] = {0, nthin, n};
    int64_t qbeg[2][GROUP_QUEUES + 1];
    int64_t tot[2] = {0, 0};
    for (int c = 0; c < 2; c++) {
        double work[GROUP_QUEUES] = {0};
        for (int i = first[c]; i < first[c + 1]; i++) {
            const tk_prob* d = &list[ord[i]].p;
            int q = 0;
            for (int j = 1; j < GROUP_QUEUES; j++) if (work[j] < work[q]) q = j;
            bin[i] = q;
            work[q] += list[ord[i]].cost * (double)(d->L * d->tiles_s * d->tiles_g);
        }
        int64_t total = 0;
        for (int q = 0, k = first[c]; q < GROUP_QUEUES; q++) {
            qbeg[c][q] = total;
            for (int i = first[c]; i < first[c + 1]; i++) {
                if (bin[i] != q) continue;
                hp[k] = list[ord[i]].p;
                src[k] = list[ord[i]].idx;
                total += hp[k].L * hp[k].tiles_s * hp[k].tiles_g;
                he[k++] = total;
            }
        }
        qbeg[c][GROUP_QUEUES] = total;
        tot[c] = total;
    }
    g->total = tot[0] + tot[1];
    /* every tile decoded (product, l, ts, tg), so a workgroup knows its next
     * tile from one 16-byte load, issued while it runs the current one */
    const bool map = g->total <= INT32_MAX;
    const size_t mb = map ? sizeof(tk_tile) * (size_t)g->total : 0;
    if (map) {
        if (!(hm = malloc(mb > 0 ? mb : 1))) { rc = TAPP_AMD_ERR_NOMEM; goto out; }
        int64_t t = 0;
        for (int c = 0; c < 2; c++)
            for (int i = first[c]; i < first[c + 1]; i++)
                for (int64_t l = 0; l < hp[i].L; l++)
                    for (int64_t ts = 0; ts < hp[i].tiles_s; ts++)
                        for (int64_t tg = 0; tg < hp[i].tiles_g; tg++)
                            hm[t++] = (tk_tile){i - first[c], (int32_t)l, (int32_t)ts, (int32_t)tg};
    }
    const size_t o_end = (pb + 255) & ~(size_t)255, o_map = (o_end + tb + 255) & ~(size_t)255;
    const size_t o_q = (o_map + mb + 255) & ~(size_t)255;
    const size_t o_cnt = (o_q + sizeof qbeg + 255) & ~(size_t)255;
    hipError_t e = hipMalloc(&g->dmem, o_cnt + 2 * 256 * GROUP_QUEUES);
    if (e != hipSuccess) { (void)hipGetLastError(); g->dmem = NULL; rc = TAPP_AMD_ERR_NOMEM; goto out; }
    char* base = (char*)g->dmem;
    g->dprobs = (tk_prob*)base;
    e = hipMemcpy(base, hp, pb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(base + o_end, he, tb, hipMemcpyHostToDevice);
    if (e == hipSuccess && map) e = hipMemcpy(base + o_map, hm, mb, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(base + o_q, qbeg, sizeof qbeg, hipMemcpyHostToDevice);
    if (e != hipSuccess) { (void)hipGetLastError(); rc = TAPP_AMD_ERR_DEVICE; goto out; }
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ex->device) != hipSuccess || cus <= 0)
        cus = 256;
    g->nlaunch = 0;
    for (int c = 0; c < 2; c++) {
        if (first[c + 1] == first[c]) continue;
        group_launch* L = &g->launch[g->nlaunch++];
        L->kind = c == 0 ? TK_GROUP_KIND_THIN : TK_GROUP_KIND_MFMA;
        L->n = first[c + 1] - first[c];
        L->total = tot[c];
        const int64_t nb = (int64_t)cus * (c == 0 ? TK_GROUP_WG_PER_CU_THIN : TK_GROUP_WG_PER_CU);
        L->nblocks = (int)(tot[c] < nb ? tot[c] : nb);
        L->dprobs = g->dprobs + first[c];
        L->dtile_end = (int64_t*)(base + o_end) + first[c];
        L->dtiles = map ? (tk_tile*)(base + o_map) + (c == 0 ? 0 : tot[0]) : NULL;
        L->dqbeg = (int64_t*)(base + o_q) + c * (GROUP_QUEUES + 1);
        L->dcounter = (unsigned long long*)(base + o_cnt + (size_t)c * 256 * GROUP_QUEUES);
    }
    if (keep_host) {
        if (hipHostMalloc((void**)&g->hprobs, pb, hipHostMallocDefault) == hipSuccess) {
            memcpy(g->hprobs, hp, pb);
            g->hsrc = src;
            src = NULL;
        } else {
            (void)hipGetLastError();
            g->hprobs = NULL;
        }
    }
    /* capture zero-counters + persistent kernels once; replays reuse them */
    if (capture && hipStreamBeginCapture(ex->stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        const int lrc = launch_all(g, ex->stream);
        hipError_t ce = hipStreamEndCapture(ex->stream, &g->graph);
        if (lrc == 0 && ce == hipSuccess && hipGraphInstantiate(&g->gexec, g->graph, NULL, NULL, 0) == hipSuccess)
            g->have_graph = true;
        else if (ce == hipSuccess && g->graph) hipGraphDestroy(g->graph);
        (void)hipGetLastError();
    }
out:
    free(hp); free(he); free(bin); free(src); free(ord); free(hm);
    return rc;
}

static int create_group_on(TAPP_AMD_product_group* group, struct tapp_exec* ex, int count,
                           const TAPP_tensor_product* plans, const void* const* alpha, const void* const* A,
                           const void* const* B, const void* const* beta, const void* const* C, void* const* D,
                           bool capture, bool keep_host)
{
    if (!ex->have_stream) return TAPP_AMD_ERR_DEVICE;
    struct tapp_group* g = calloc(1, sizeof *g);
    if (!g) return TAPP_AMD_ERR_NOMEM;
    g->magic = TAPP_AMD_EXEC_MAGIC;
    g->ex = ex;
    ranked* list = malloc(sizeof(ranked) * (size_t)(count > 0 ? count : 1));
    g->rest_plan = malloc(sizeof(void*) * (size_t)(count > 0 ? count : 1));
    g->rest_A = malloc(sizeof(void*) * (size_t)(count > 0 ? count : 1));
    g->rest_B = malloc(sizeof(void*) * (size_t)(count > 0 ? count : 1));
    g->rest_C = malloc(sizeof(void*) * (size_t)(count > 0 ? count : 1));
    g->rest_D = malloc(sizeof(void*) * (size_t)(count > 0 ? count : 1));
    g->rest_scalar = calloc((size_t)(count > 0 ? count : 1), sizeof *g->rest_scalar);
    if (!list || !g->rest_plan || !g->rest_A || !g->rest_B || !g->rest_C || !g->rest_D || !g->rest_scalar) {
        free(list); group_free(g);
        return TAPP_AMD_ERR_NOMEM;
    }
    if (hipSetDevice(ex->device) != hipSuccess) { free(list); group_free(g); return TAPP_AMD_ERR_DEVICE; }
    int n = 0;
    for (int i = 0; i < count; i++) {
        struct tapp_plan* p = (struct tapp_plan*)plans[i];
        if (!p) { free(list); group_free(g); return TAPP_AMD_ERR_INVALID; }
        const void* c = C ? C[i] : NULL;
        tk_problem P;
        int rc = tk_plan_problem(p, ex, alpha[i], beta[i], c, &P);
        if (rc) { free(list); group_free(g); return rc; }
        if (P.empty) continue;
        tk_prob d;
        if (tk_group_describe(p, ex, &P, A[i], B[i], c, D[i], -1, &d, &rc)) {
            list[n].p = d;
            list[n].cost = tile_cost(&d);
            list[n].idx = i;
            n++;
        } else {
            if (rc) { free(list); group_free(g); return rc; }
            int r = g->nrest++;
            g->rest_plan[r] = p; g->rest_A[r] = A[i]; g->rest_B[r] = B[i];
            g->rest_C[r] = c; g->rest_D[r] = D[i];
            const size_t ss = tk_type_size(P.type[3]);
            if (alpha[i] && ss) memcpy(g->rest_scalar[r][0], alpha[i], ss);
            if (beta[i] && ss) memcpy(g->rest_scalar[r][1], beta[i], ss);
        }
    }
    if (ex->autotune && n > 0) {
        int trc = tune_group(ex, list, n, plans, alpha, A, B, beta, C, D);
        if (trc) { free(list); group_free(g); return trc; }
    }
    qsort(list, (size_t)n, sizeof(ranked), by_cost);
    g->n = n;
    if (n > 0) {
        int rc = build_launches(g, ex, list, n, capture, keep_host);
        if (rc) { free(list); group_free(g); return rc; }
    }
    free(list);
    *group = (TAPP_AMD_product_group)g;
    return 0;
}

TAPP_error TAPP_AMD_create_product_group(TAPP_AMD_product_group* group, TAPP_executor exec, int count,
                                         const TAPP_tensor_product* plans, const void* const* alpha,
                                         const void* const* A, const void* const* B, const void* const* beta,
                                         const void* const* C, void* const* D)
{
    return create_group(group, exec, count, plans, alpha, A, B, beta, C, D, true);
}

static int execute_group(struct tapp_group* g);

/* --- TAPP_execute_batched_product ------------------------------------------
 * One plan, num_batches operand sets, one alpha and beta (product.h:66-75),
 * run as one persistent launch.  The group built for a call is kept on the
 * plan when every operand set ran in the persistent launch from its own
 * device pointers; the next call with the same batch count on the same
 * device reuses it: the plan is validated once, only operand pointers that
 * changed are classified, their descriptors patched and uploaded (one copy
 * from pinned memory), and the captured graph replayed.  Anything else --
 * another batch count or device, a changed layout, precision or beta == 0,
 * a host or resident operand, a NULL pointer -- builds the group again,
 * with the codes and behaviour of the first call. */

struct tapp_batch_cache {
    int device, num;
    int mode, beta_zero, round_C;
    bool c_null;                 /* the call passed no C array */
    uint64_t key;                /* layout key and device base of the offset tables */
    const void* tables;          /* the descriptors point into */
    struct tapp_group* g;        /* n == num products, nothing one by one */
    const void **A, **B, **C;    /* operand pointers of the last call, per set */
    void** D;
};

void tk_batch_cache_free(struct tapp_batch_cache* b)
{
    if (!b) return;
    group_free(b->g);
    free(b->A); free(b->B); free(b->C); free(b->D);
    free(b);
}

/* Device allocations met during one call, so that operand pointers inside
 * one already classified are not looked up again. */
typedef struct {
    const char* lo[8];
    const char* hi[8];
    int n, next;
} range_memo;

static bool memo_is_device(range_memo* m, const void* p)
{
    const char* c = (const char*)p;
    for (int i = 0; i < m->n; i++)
        if (c >= m->lo[i] && c < m->hi[i]) return true;
    if (!tk_is_device_ptr(p)) return false;
    void* base = NULL;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (void*)p) == hipSuccess && size > 0) {
        const int s = m->n < 8 ? m->n++ : (m->next++ & 7);
        m->lo[s] = (const char*)base;
        m->hi[s] = (const char*)base + size;
    } else {
        (void)hipGetLastError();
    }
    return true;
}

static const char* shift(const void* now, const void* was, const void* field)
{
    return (const char*)field + ((const char*)now - (const char*)was);
}

/* Reuse of the plan's prepared batch; returns -1 when it does not apply. */
static int batched_reuse(struct tapp_plan* p, struct tapp_exec* ex, int num, const void* alpha,
                         const void** A, const void** B, const void* beta, const void** C, void** D)
{
    struct tapp_batch_cache* b = p->batch;
    if (!b || b->device != ex->device || b->num != num || b->c_null != (C == NULL)) return -1;
    tk_problem P;
    if (tk_plan_problem(p, ex, alpha, beta, C ? C[0] : NULL, &P) || P.empty) return -1;
    if (P.mode != b->mode || P.beta_zero != b->beta_zero || P.round_C != b->round_C) return -1;
    const int64_t *t[12];
    if (tk_problem_tables(p, &P, ex->device, &t[0], &t[1], &t[2], &t[3], &t[4], &t[5], &t[6], &t[7], &t[8], &t[9],
                          &t[10], &t[11]))
        return -1;
    uint64_t key;
    if (tk_plan_tables_id(p, &key) != b->tables || key != b->key) return -1;
    /* every changed pointer must be device memory (checked before anything
     * is patched, so a refusal leaves the cache as it was) */
    range_memo memo = {{0}, {0}, 0, 0};
    bool changed = false;
    for (int i = 0; i < num; i++) {
        if (!A[i] || !B[i] || !D[i] || (!P.beta_zero && !C[i])) return -1;
        if (A[i] != b->A[i]) { if (!memo_is_device(&memo, A[i])) return -1; changed = true; }
        if (B[i] != b->B[i]) { if (!memo_is_device(&memo, B[i])) return -1; changed = true; }
        if (D[i] != b->D[i]) { if (!memo_is_device(&memo, D[i])) return -1; changed = true; }
        if (!P.beta_zero && C[i] != b->C[i]) { if (!memo_is_device(&memo, C[i])) return -1; changed = true; }
    }
    struct tapp_group* g = b->g;
    const float al = (float)P.alpha[0], be = (float)P.beta[0];
    if (g->n > 0 && (al != g->hprobs[0].alpha || be != g->hprobs[0].beta)) changed = true;
    if (changed) {
        for (int k = 0; k < g->n; k++) {
            tk_prob* d = &g->hprobs[k];
            const int i = g->hsrc[k];
            const void *xa = d->x_is_b ? B[i] : A[i], *xw = d->x_is_b ? b->B[i] : b->A[i];
            const void *ya = d->x_is_b ? A[i] : B[i], *yw = d->x_is_b ? b->A[i] : b->B[i];
            d->X = (const float*)shift(xa, xw, d->X);
            d->Y = (const float*)shift(ya, yw, d->Y);
            d->D = (float*)shift(D[i], b->D[i], d->D);
            if (!P.beta_zero) d->C = (const float*)shift(C[i], b->C[i], d->C);
            d->alpha = al;
            d->beta = be;
        }
        for (int i = 0; i < num; i++) {
            b->A[i] = A[i]; b->B[i] = B[i]; b->D[i] = D[i];
            b->C[i] = C ? C[i] : NULL;
        }
        if (hipSetDevice(ex->device) != hipSuccess ||
            hipMemcpyAsync(g->dprobs, g->hprobs, sizeof(tk_prob) * (size_t)g->n, hipMemcpyHostToDevice,
                           ex->stream) != hipSuccess) {
            (void)hipGetLastError();
            p->batch = NULL;          /* device copy unknown: build again next time */
            tk_batch_cache_free(b);
            return TAPP_AMD_ERR_DEVICE;
        }
    }
    g->ex = ex;
    return execute_group(g);
}

/* Keeps `g` on the plan for the next call when every operand set runs in
 * the persistent launch from its caller's pointers; otherwise frees it. */
static void batched_keep(struct tapp_plan* p, struct tapp_exec* ex, struct tapp_group* g, int num,
                         const void* alpha, const void** A, const void** B, const void* beta, const void** C,
                         void** D)
{
    tk_problem P;
    uint64_t key = 0;
    const void* tables = tk_plan_tables_id(p, &key);
    bool ok = g->n == num && g->nrest == 0 && g->hprobs && g->hsrc && tables &&
              !tk_plan_problem(p, ex, alpha, beta, C ? C[0] : NULL, &P);
    for (int k = 0; ok && k < g->n; k++) {
        /* not a resident copy: the descriptor's pointers are the caller's */
        const tk_prob* d = &g->hprobs[k];
        const int i = g->hsrc[k], xi = d->x_is_b ? 1 : 0;
        const void* ops[2] = {A[i], B[i]};
        ok = d->X == (const float*)ops[xi] + P.pin[xi] && d->Y == (const float*)ops[1 - xi] + P.pin[1 - xi] &&
             d->D == (float*)D[i] + P.pin[3] && (P.beta_zero || d->C == (const float*)C[i] + P.pin[2]);
    }
    struct tapp_batch_cache* b = ok ? calloc(1, sizeof *b) : NULL;
    if (b) {
        b->A = malloc(sizeof(void*) * (size_t)num);
        b->B = malloc(sizeof(void*) * (size_t)num);
        b->C = malloc(sizeof(void*) * (size_t)num);
        b->D = malloc(sizeof(void*) * (size_t)num);
    }
    if (!b || !b->A || !b->B || !b->C || !b->D) {
        if (b) { free(b->A); free(b->B); free(b->C); free(b->D); free(b); }
        group_free(g);
        return;
    }
    b->device = ex->device;
    b->num = num;
    b->mode = P.mode;
    b->beta_zero = P.beta_zero;
    b->round_C = P.round_C;
    b->c_null = C == NULL;
    b->key = key;
    b->tables = tables;
    b->g = g;
    for (int i = 0; i < num; i++) {
        b->A[i] = A[i]; b->B[i] = B[i]; b->D[i] = D[i];
        b->C[i] = C ? C[i] : NULL;
    }
    tk_batch_cache_free(p->batch);
    p->batch = b;
}

TAPP_error TAPP_execute_batched_product(TAPP_tensor_product plan, TAPP_executor exec, TAPP_status* status,
                                        int num_batches, const void* alpha, const void** A, const void** B,
                                        const void* beta, const void** C, void** D)
{
    (void)status;
    if (!plan || num_batches < 0 || (num_batches > 0 && (!A || !B || !D))) return TAPP_AMD_ERR_INVALID;
    if (num_batches == 0) return 0;
    struct tapp_plan* p = (struct tapp_plan*)plan;
    struct tapp_exec* ex = tk_exec_lookup(exec);
    if (!ex) ex = (exec && (*(const int*)exec == 1 || *(const int*)exec == 2 || *(const int*)exec == 12))
                      ? tk_default_exec() : NULL;
    if (!ex) return 33;
    tk_exec_lock(ex);
    pthread_mutex_lock(&p->mu);
    int rc = batched_reuse(p, ex, num_batches, alpha, A, B, beta, C, D);
    ex->last_reused = rc >= 0;
    if (rc < 0) {
        size_t n = (size_t)num_batches;
        TAPP_tensor_product* plans = malloc(sizeof(TAPP_tensor_product) * n);
        const void** al = malloc(sizeof(void*) * n);
        const void** be = malloc(sizeof(void*) * n);
        if (!plans || !al || !be) {
            rc = TAPP_AMD_ERR_NOMEM;
        } else {
            for (size_t i = 0; i < n; i++) { plans[i] = plan; al[i] = alpha; be[i] = beta; }
            TAPP_AMD_product_group g = 0;
            rc = create_group_on(&g, ex, num_batches, plans, al, A, B, be, C, D, true, true);
            if (!rc) rc = execute_group((struct tapp_group*)g);
            if (g && !rc) {
                batched_keep(p, ex, (struct tapp_group*)g, num_batches, alpha, A, B, beta, C, D);
            } else if (g) {
                group_free((struct tapp_group*)g);
            }
        }
        free(plans); free(al); free(be);
    }
    pthread_mutex_unlock(&p->mu);
    tk_exec_unlock(ex);
    return rc;
}

TAPP_error TAPP_AMD_execute_product_group(TAPP_AMD_product_group group)
{
    struct tapp_group* g = (struct tapp_group*)group;
    if (!g || g->magic != TAPP_AMD_EXEC_MAGIC) return TAPP_AMD_ERR_INVALID;
    tk_exec_lock(g->ex);
    int rc = execute_group(g);
    tk_exec_unlock(g->ex);
    return rc;
}

static int execute_group(struct tapp_group* g)
{
    struct tapp_exec* ex = g->ex;
    if (hipSetDevice(ex->device) != hipSuccess) return TAPP_AMD_ERR_DEVICE;
    if (ex->timing) hipEventRecord(ex->ev0, ex->stream);
    if (g->n > 0) {
        if (g->have_graph) {
            if (hipGraphLaunch(g->gexec, ex->stream) != hipSuccess) return TAPP_AMD_ERR_DEVICE;
        } else if (launch_all(g, ex->stream)) {
            return TAPP_AMD_ERR_DEVICE;
        }
    }
    if (ex->timing) hipEventRecord(ex->ev1, ex->stream);
    for (int r = 0; r < g->nrest; r++) {
        int rc = tk_execute(g->rest_plan[r], ex, g->rest_scalar[r][0], g->rest_A[r], g->rest_B[r],
                            g->rest_scalar[r][1], g->rest_C[r], g->rest_D[r]);
        if (rc) return rc;
    }
    if (hipStreamSynchronize(ex->stream) != hipSuccess) return TAPP_AMD_ERR_DEVICE;
    if (ex->timing && g->n > 0) {
        float ms = 0;
        hipEventElapsedTime(&ms, ex->ev0, ex->ev1);
        ex->last_device_us = ms * 1000.0;
    }
    ex->last_kernel = TAPP_AMD_KERNEL_GROUP;
    snprintf(ex->last_kernel_name, sizeof ex->last_kernel_name, "k_group (%d persistent, %d one by one)", g->n,
             g->nrest);
    return 0;
}

TAPP_error TAPP_AMD_destroy_product_group(TAPP_AMD_product_group group)
{
    struct tapp_group* g = (struct tapp_group*)group;
    if (!g) return 0;
    if (g->ex && g->ex->have_stream) hipStreamSynchronize(g->ex->stream);
    g->magic = 0;
    group_free(g);
    return 0;
}

TAPP_error TAPP_AMD_group_info(TAPP_AMD_product_group group, int64_t info[3])
{
    struct tapp_group* g = (struct tapp_group*)group;
    if (!g || g->magic != TAPP_AMD_EXEC_MAGIC) return TAPP_AMD_ERR_INVALID;
    info[0] = g->n;
    info[1] = g->nrest;
    info[2] = g->total;
    return 0;
}
