/*
 * Plan stage of TAPP_execute_product: integer-only, runs on the host per
 * call (microseconds), result cached per plan by the executor.
 *
 * 1. Validation, same checks, order and codes as product.c:197-206
 *    (helpers product.c:757-885).
 * 2. Repeated labels fold into one mode whose stride is the sum of the
 *    occurrences' strides, first occurrence keeps its place
 *    (compress_repeated_indices, product.c:969-1010), on A, B, C and D.
 * 3. Classification (product.c:271-310): a label of D in A and B is a
 *    Hadamard (batch, group L) label; in A only, group M; in B only, group N.
 *    A label of A and B absent from D is contracted (group K).  A label in
 *    one operand only is unary (groups UA, UB, kept in the operand's order,
 *    which is the reference's summation order).  C is addressed through D's
 *    positions (product.c:348) because check 5 forced idx_C == idx_D.
 * 4. Compute type and element-op semantics per precision branch of the
 *    reference (alloc_accum :1012, create_prec_scalar :1232,
 *    sum_unary_contractions :1863, calculate_beta_C_prec :3156,
 *    calculate_op_D :10737).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

enum { T_F32 = 0, T_F64 = 1, T_C32 = 2, T_C64 = 3, T_F16 = 4, T_BF16 = 5 };

size_t tk_type_size(int t)
{
    switch (t) {
    case T_F32: return 4;
    case T_F64: return 8;
    case T_C32: return 8;
    case T_C64: return 16;
    case T_F16: return 2;
    case T_BF16: return 2;
    default: return 0;
    }
}

bool tk_type_complex(int t) { return t == T_C32 || t == T_C64; }

typedef struct {
    int type, nmode;
    const int64_t* ext;
    const int64_t* str;
    const int64_t* idx;
} view;

/* --- validation ---------------------------------------------------------- */

/* Every label of o appears in a or b (product.c:757-784). */
static int check_occurrence(const view* o, const view* a, const view* b, int code)
{
    for (int i = 0; i < o->nmode; i++) {
        bool found = false;
        for (int j = 0; j < a->nmode && !found; j++) found = o->idx[i] == a->idx[j];
        for (int j = 0; j < b->nmode && !found; j++) found = o->idx[i] == b->idx[j];
        if (!found) return code;
    }
    return 0;
}

/* Per mode of x: same label in x, then y, then z must have the same extent
 * (product.c:786-813). */
static int check_extents(const view* x, const view* y, const view* z, int cxx, int cxy, int cxz)
{
    for (int i = 0; i < x->nmode; i++) {
        for (int j = 0; j < x->nmode; j++)
            if (x->idx[i] == x->idx[j] && x->ext[i] != x->ext[j]) return cxx;
        for (int j = 0; j < y->nmode; j++)
            if (x->idx[i] == y->idx[j] && x->ext[i] != y->ext[j]) return cxy;
        for (int j = 0; j < z->nmode; j++)
            if (x->idx[i] == z->idx[j] && x->ext[i] != z->ext[j]) return cxz;
    }
    return 0;
}

/* C and D: same mode count, labels and extents position by position
 * (product.c:815-834). */
static int check_structure(const view* c, const view* d)
{
    if (c->nmode != d->nmode) return 5;
    for (int i = 0; i < d->nmode; i++) {
        if (d->idx[i] != c->idx[i]) return 6;
        if (d->ext[i] != c->ext[i]) return 7;
    }
    return 0;
}

/* D must not overlap itself (product.c:836-870).  With two or more modes a
 * zero stride fails; otherwise (|stride|, extent) pairs sorted by stride,
 * ties by extent, must satisfy s[i+1] >= s[i] * e[i].  The reference sorts
 * through 32-bit int temporaries (product.c:905-918), so the compared
 * values are the low 32 bits of |stride| and extent; kept here so that the
 * accepted set of layouts is the reference's. */
typedef struct { int32_t s, e; } se;
static int cmp_se(const void* x, const void* y)
{
    const se* a = x; const se* b = y;
    if (a->s != b->s) return a->s < b->s ? -1 : 1;
    return a->e < b->e ? -1 : a->e > b->e;
}
static int check_aliasing(const view* d)
{
    int n = d->nmode;
    if (n <= 1) return 0;
    for (int i = 0; i < n; i++) if (d->str[i] == 0) return 8;
    se small[16];
    se* v = n <= 16 ? small : malloc(sizeof(se) * (size_t)n);
    if (!v) return TAPP_AMD_ERR_NOMEM;
    for (int i = 0; i < n; i++) {
        uint64_t s = (uint64_t)(d->str[i] < 0 ? -d->str[i] : d->str[i]);
        v[i].s = (int32_t)(uint32_t)s;
        v[i].e = (int32_t)(uint32_t)(uint64_t)d->ext[i];
    }
    qsort(v, (size_t)n, sizeof(se), cmp_se);
    int rc = 0;
    for (int i = 0; i + 1 < n; i++)
        if ((int64_t)v[i + 1].s < (int64_t)v[i].s * (int64_t)v[i].e) { rc = 8; break; }
    if (v != small) free(v);
    return rc;
}

/* Scalar of storage type t as (re, im) in double. */
static void scalar_value(const void* s, int t, double* re, double* im)
{
    *re = 0; *im = 0;
    if (!s) return;
    switch (t) {
    case T_F32: *re = *(const float*)s; break;
    case T_F64: *re = *(const double*)s; break;
    case T_C32: *re = ((const float*)s)[0]; *im = ((const float*)s)[1]; break;
    case T_C64: *re = ((const double*)s)[0]; *im = ((const double*)s)[1]; break;
    case T_F16: {
        uint16_t h = *(const uint16_t*)s;
        int e = (h >> 10) & 0x1f, m = h & 0x3ff;
        double v = e == 0 ? ldexp(m, -24) : e == 31 ? (m ? NAN : INFINITY) : ldexp(m + 1024, e - 25);
        *re = (h & 0x8000) ? -v : v;
    } break;
    case T_BF16: {
        uint32_t u = (uint32_t)(*(const uint16_t*)s) << 16;
        float f;
        memcpy(&f, &u, 4);
        *re = f;
    } break;
    default: break;
    }
}

/* Round-to-nearest-even of a binary64 value to binary16 (mode TK_MODE_F16) or
 * bfloat16 (TK_MODE_BF16) in one step, as the reference's conversion of a
 * float or double scalar to _Float16 does (create_prec_scalar, :1232-1490). */
static double rnd16_d(double x, int mode)
{
    if (!isfinite(x) || x == 0.0) return x;
    const int p = mode == TK_MODE_F16 ? 11 : 8, emin = mode == TK_MODE_F16 ? -14 : -126;
    const double ovf = mode == TK_MODE_F16 ? 65520.0 : ldexp(2.0 - ldexp(1.0, -8), 127);
    int e;
    frexp(fabs(x), &e);
    int q = e - p;
    if (q < emin - p + 1) q = emin - p + 1;
    double r = ldexp(rint(ldexp(fabs(x), -q)), q);
    if (r >= ovf) r = INFINITY;
    return copysign(r, x);
}

/* --- groups --------------------------------------------------------------- */

typedef struct {
    int n;
    int64_t idx[TK_MAXG], ext[TK_MAXG], str[TK_MAXG];
} cmodes;

static int compress(const view* v, cmodes* c)
{
    c->n = 0;
    for (int i = 0; i < v->nmode; i++) {
        bool seen = false;
        for (int j = 0; j < c->n; j++) seen |= c->idx[j] == v->idx[i];
        if (seen) continue;
        if (c->n >= TK_MAXG) return TAPP_AMD_ERR_UNSUPPORTED;
        int64_t s = 0;
        for (int j = 0; j < v->nmode; j++) if (v->idx[j] == v->idx[i]) s += v->str[j];
        c->idx[c->n] = v->idx[i];
        c->ext[c->n] = v->ext[i];
        c->str[c->n] = s;
        c->n++;
    }
    return 0;
}

static int findc(const cmodes* c, int64_t label)
{
    for (int j = 0; j < c->n; j++) if (c->idx[j] == label) return j;
    return -1;
}

static void add_mode(tk_group* g, int64_t ext, int64_t sa, int64_t sb, int64_t sc, int64_t sd)
{
    g->ext[g->n] = ext;
    g->s[0][g->n] = sa; g->s[1][g->n] = sb; g->s[2][g->n] = sc; g->s[3][g->n] = sd;
    g->n++;
}

static int64_t group_size(const tk_group* g)
{
    int64_t s = 1;
    for (int i = 0; i < g->n; i++) s *= g->ext[i];
    return s;
}

/* stable insertion sort of a group's modes by |stride of operand key|,
 * ties by |stride of operand tie| */
static void sort_group(tk_group* g, int key, int tie)
{
    for (int i = 1; i < g->n; i++) {
        for (int j = i; j > 0; j--) {
            int64_t a = llabs(g->s[key][j - 1]), b = llabs(g->s[key][j]);
            int64_t ta = llabs(g->s[tie][j - 1]), tb = llabs(g->s[tie][j]);
            if (a < b || (a == b && ta <= tb)) break;
            int64_t t = g->ext[j]; g->ext[j] = g->ext[j - 1]; g->ext[j - 1] = t;
            for (int x = 0; x < 4; x++) { t = g->s[x][j]; g->s[x][j] = g->s[x][j - 1]; g->s[x][j - 1] = t; }
        }
    }
}

/* Drops extent-1 modes, and pins zero-stride modes of D with extent > 1 at
 * their last coordinate: the reference writes that D element once per
 * coordinate in loop order (first coordinate fastest), so the value that
 * remains is the one computed at coordinate extent-1. */
static void simplify_group(tk_group* g, bool is_output, int64_t pin[4])
{
    int w = 0;
    for (int i = 0; i < g->n; i++) {
        if (g->ext[i] == 1) continue;
        if (is_output && g->s[3][i] == 0 && g->ext[i] > 1) {
            for (int x = 0; x < 4; x++) pin[x] += (g->ext[i] - 1) * g->s[x][i];
            continue;
        }
        g->ext[w] = g->ext[i];
        for (int x = 0; x < 4; x++) g->s[x][w] = g->s[x][i];
        w++;
    }
    g->n = w;
}

static void span_of(const view* v, int64_t* lo, int64_t* hi)
{
    *lo = 0; *hi = 0;
    for (int i = 0; i < v->nmode; i++) {
        if (v->ext[i] == 0) { *lo = 0; *hi = -1; return; }
        int64_t d = (v->ext[i] - 1) * v->str[i];
        if (d < 0) *lo += d; else *hi += d;
    }
}

static bool known_type(int t) { return t >= T_F32 && t <= T_BF16; }

int tk_plan_problem(const struct tapp_plan* p, const struct tapp_exec* ex, const void* alpha,
                    const void* beta, const void* C, tk_problem* out)
{
    memset(out, 0, sizeof *out);
    view v[4];
    for (int x = 0; x < 4; x++) {
        const struct tapp_info* t = (const struct tapp_info*)p->info[x];
        if (!t) return out->rc = TAPP_AMD_ERR_INVALID;
        if (t->nmode > p->nidx[x]) return out->rc = TAPP_AMD_ERR_INVALID;
        v[x].type = t->type;
        v[x].nmode = t->nmode;
        v[x].ext = t->extents;
        v[x].str = t->strides;
        v[x].idx = p->idx[x];
        out->type[x] = t->type;
    }
    const view *A = &v[0], *B = &v[1], *Cv = &v[2], *D = &v[3];

    /* 1. validation chain (product.c:199-206) */
    int rc = check_occurrence(D, A, B, 4);
    if (!rc) rc = check_extents(A, B, D, 9, 1, 2);
    if (!rc) rc = check_extents(B, A, D, 10, 1, 3);
    if (!rc) rc = check_extents(D, A, B, 11, 2, 3);
    if (!rc) rc = check_structure(Cv, D);
    if (!rc) rc = check_aliasing(D);
    double br, bi;
    scalar_value(beta, D->type, &br, &bi);
    out->beta_zero = br == 0.0 && bi == 0.0;
    if (!rc && C == NULL && !out->beta_zero) rc = 12;
    if (!rc && !ex) rc = 33;
    if (!rc && !(ex->kind == 1 || ex->kind == 2 || ex->kind == 12)) rc = 33;
    if (rc) return out->rc = rc;

    /* 4. compute type and op semantics */
    for (int x = 0; x < 4; x++)
        if (!known_type(v[x].type)) return out->rc = TAPP_AMD_ERR_UNSUPPORTED;
    bool cplx = tk_type_complex(A->type) || tk_type_complex(B->type) || tk_type_complex(Cv->type) ||
                tk_type_complex(D->type);
    bool wide;
    out->op_A = p->op[0]; out->op_B = p->op[1]; out->op_C = p->op[2]; out->op_D = p->op[3];
    out->mode = TK_MODE_PLAIN;
    switch (p->prec) {
    case TAPP_DEFAULT_PREC:
        wide = D->type == T_F64 || D->type == T_C64;
        break;
    case TAPP_F32F32_ACCUM_F32:
        wide = false; out->op_C = TAPP_IDENTITY;
        break;
    case TAPP_F64F64_ACCUM_F64:
        wide = true; out->op_C = TAPP_IDENTITY; out->conjD_via_f32 = true;
        break;
    case TAPP_F16F16_ACCUM_F16:
    case TAPP_F16F16_ACCUM_F32:
    case TAPP_BF16BF16_ACCUM_F32:
        wide = false;
        out->mode = p->prec == TAPP_BF16BF16_ACCUM_F32 ? TK_MODE_BF16 : TK_MODE_F16;
        out->op_A = out->op_B = out->op_C = TAPP_IDENTITY;
        if (p->prec == TAPP_F16F16_ACCUM_F16) out->op_D = TAPP_IDENTITY;
        out->round_C = true;
        break;
    default:
        return out->rc = TAPP_AMD_ERR_UNSUPPORTED;
    }
    if (!tk_type_complex(A->type)) out->op_A = TAPP_IDENTITY;
    if (!tk_type_complex(B->type)) out->op_B = TAPP_IDENTITY;
    if (!tk_type_complex(Cv->type)) out->op_C = TAPP_IDENTITY;
    if (!tk_type_complex(D->type)) out->op_D = TAPP_IDENTITY;
    out->ctype = cplx ? (wide ? TK_C64 : TK_C32) : (wide ? TK_F64 : TK_F32);
    if (ex->split_bf16 && !wide && out->mode == TK_MODE_PLAIN &&
        (p->prec == TAPP_DEFAULT_PREC || p->prec == TAPP_F32F32_ACCUM_F32))
        out->mode = ex->split_bf16 == 3 ? TK_MODE_SPLIT_F16
                    : ex->split_bf16 == 2 ? TK_MODE_SPLIT6_BF16 : TK_MODE_SPLIT_BF16;

    /* alpha and beta are read as type_D (product.c:330-331) */
    double ar, ai;
    scalar_value(alpha, D->type, &ar, &ai);
    if (!cplx) { ai = 0; bi = 0; }
    if (out->mode == TK_MODE_BF16 || out->mode == TK_MODE_F16) {
        /* one rounding from the value read as type_D */
        ar = rnd16_d(ar, out->mode); ai = rnd16_d(ai, out->mode);
        br = rnd16_d(br, out->mode); bi = rnd16_d(bi, out->mode);
    } else if (!wide) {
        ar = (float)ar; ai = (float)ai; br = (float)br; bi = (float)bi;
    }
    /* the 16-bit sums of unary labels (alloc_val :1104-1230: _Float16 /
     * __bf16 under the 16-bit precisions, the operand's own type under
     * DEFAULT_PREC) */
    for (int x = 0; x < 2; x++) {
        if (out->mode == TK_MODE_F16 || out->mode == TK_MODE_BF16) out->sum_round[x] = out->mode;
        else if (p->prec == TAPP_DEFAULT_PREC && v[x].type == T_F16) out->sum_round[x] = TK_MODE_F16;
        else if (p->prec == TAPP_DEFAULT_PREC && v[x].type == T_BF16) out->sum_round[x] = TK_MODE_BF16;
    }
    out->alpha[0] = ar; out->alpha[1] = ai;
    out->beta[0] = br; out->beta[1] = bi;

    /* 2. fold repeated labels */
    cmodes ca, cb, cc, cd;
    if ((rc = compress(A, &ca)) || (rc = compress(B, &cb)) || (rc = compress(Cv, &cc)) || (rc = compress(D, &cd)))
        return out->rc = rc;

    /* 3. classify */
    for (int j = 0; j < cd.n; j++) {
        int ia = findc(&ca, cd.idx[j]), ib = findc(&cb, cd.idx[j]);
        int64_t sa = ia >= 0 ? ca.str[ia] : 0, sb = ib >= 0 ? cb.str[ib] : 0;
        int64_t sc = j < cc.n ? cc.str[j] : 0;
        if (ia >= 0 && ib >= 0) add_mode(&out->L, cd.ext[j], sa, sb, sc, cd.str[j]);
        else if (ia >= 0) add_mode(&out->M, cd.ext[j], sa, 0, sc, cd.str[j]);
        else add_mode(&out->N, cd.ext[j], 0, sb, sc, cd.str[j]);
    }
    for (int i = 0; i < ca.n; i++) {
        int ib = findc(&cb, ca.idx[i]), id = findc(&cd, ca.idx[i]);
        if (ib >= 0 && id < 0) add_mode(&out->K, ca.ext[i], ca.str[i], cb.str[ib], 0, 0);
        else if (ib < 0 && id < 0) add_mode(&out->UA, ca.ext[i], ca.str[i], 0, 0, 0);
    }
    for (int i = 0; i < cb.n; i++)
        if (findc(&ca, cb.idx[i]) < 0 && findc(&cd, cb.idx[i]) < 0)
            add_mode(&out->UB, cb.ext[i], 0, cb.str[i], 0, 0);

    /* any zero extent among D's modes: nothing is written; nor with
     * TAPP_F16F16_ACCUM_F16 and a complex D, whose assign_D stores nothing
     * (product.c:11096-11099) */
    out->empty = group_size(&out->L) == 0 || group_size(&out->M) == 0 || group_size(&out->N) == 0 ||
                 (p->prec == TAPP_F16F16_ACCUM_F16 && tk_type_complex(D->type));

    simplify_group(&out->L, true, out->pin);
    simplify_group(&out->M, true, out->pin);
    simplify_group(&out->N, true, out->pin);
    /* contracted / unary: extent-1 modes dropped, order otherwise kept */
    {
        int64_t nopin[4] = {0, 0, 0, 0};
        simplify_group(&out->K, false, nopin);
        simplify_group(&out->UA, false, nopin);
        simplify_group(&out->UB, false, nopin);
    }

    /* mode order inside groups (any order is exact for L, M, N; K's order
     * only changes FP rounding; unary groups keep the reference's order) */
    sort_group(&out->M, 0, 3);
    sort_group(&out->N, 1, 3);
    sort_group(&out->L, 3, 0);
    {
        bool unit_a = false, unit_b = false;
        for (int i = 0; i < out->K.n; i++) {
            unit_a |= llabs(out->K.s[0][i]) == 1;
            unit_b |= llabs(out->K.s[1][i]) == 1;
        }
        if (unit_a || !unit_b) sort_group(&out->K, 0, 1);
        else sort_group(&out->K, 1, 0);
    }

    out->nL = group_size(&out->L);
    out->nM = group_size(&out->M);
    out->nN = group_size(&out->N);
    out->nK = group_size(&out->K);
    out->nUA = group_size(&out->UA);
    out->nUB = group_size(&out->UB);
    if (out->nUA == 0 || out->nUB == 0) out->nK = 0;   /* every product term is 0 */

    for (int x = 0; x < 4; x++) span_of(&v[x], &out->lo[x], &out->hi[x]);
    return out->rc = 0;
}
