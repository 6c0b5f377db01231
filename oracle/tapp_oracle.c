/*
 * Oracle for TAPP_execute_product -- TEST INFRASTRUCTURE ONLY.
 *
 * A single-threaded C restatement of reference_implementation/src/product.c
 * of TAPPorg/reference-implementation.  Each function names the reference
 * lines it restates.  Loaded only by the test suite, the smoke check and the
 * CPU-baseline leg of the benchmark, to compare results against; the GPU
 * library neither links nor calls it.
 *
 * What pins it (every function has at least one check in tests/):
 *   - orc_check, orc_classify: compared code-for-code and array-for-array
 *     with the compiled reference (oracle/_ref) and its exported helpers on
 *     random small cases, plus hand-built negative cases per error code.
 *   - orc_execute, types F32/F64/C32/C64 with DEFAULT_PREC, F32F32_ACCUM_F32,
 *     F64F64_ACCUM_F64: bit-exact against the compiled reference on random
 *     cases (same per-element operation order), and against the known
 *     answers printed by the reference's demo/driver/exercise programs.
 *   - orc_execute, mixed storage types (loop_generic): bit-exact against
 *     the compiled reference on every combination of F32/F64/C32/C64 over
 *     A, B, C, D under DEFAULT_PREC, F32F32_ACCUM_F32 and F64F64_ACCUM_F64.
 *   - orc_execute, TAPP_F16F16_ACCUM_F32 / _F16 and DEFAULT_PREC with
 *     binary16 storage (loop_generic): bit-exact against the clang build of
 *     the reference with TAPP_REFERENCE_ENABLE_F16 on every scenario class,
 *     including unary sums, beta != 0 and mixed storage types.  loop_generic
 *     also gives the per-type loops' bits on uniform F32/F64/C32/C64 data.
 *   - orc_execute, BF16BF16_ACCUM_F32 and bfloat16 storage: the reference's
 *     BF16 code does not compile (product.c:1219, 1268, 1586, 3562), so the
 *     __bf16 evaluation is restated by analogy with _Float16 (same rules,
 *     8-bit significand) and is PARITY UNPINNED; what is pinned is the
 *     ORC_GPU_PRODUCTS form the GPU library computes (equal to the binary32
 *     loop on bfloat16-rounded inputs), the bfloat16 rounding known answers
 *     and the FP64 oracle with the 2 u_bf16 + gamma_K bound.
 *   - orc_execute_f64: pinned by agreement with orc_execute at F64F64
 *     precision (same operation order, bit-exact for F64 data) and by closed
 *     forms (rank-1 inputs).
 *
 * Deviations from the reference, all outside its defined behaviour:
 *   - element offsets are int64 (the reference's are 32-bit int,
 *     product.c:340-343, and wrap past 2^31 elements);
 *   - a NULL alpha/beta pointer is treated as the value 0 (the reference
 *     dereferences it).
 */
#include "tapp_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdbool.h>
#include <stdlib.h>
#include <string.h>

/* TAPP enum values (api/include/tapp/datatype.h:17-95, product.h:17-21). */
enum { T_F32 = 0, T_F64 = 1, T_C32 = 2, T_C64 = 3, T_F16 = 4, T_BF16 = 5 };
enum { P_DEFAULT = -1, P_F32 = 0, P_F64 = 1, P_F16F16_F16 = 4, P_F16F16_F32 = 5,
       P_BF16BF16_F32 = 6 };
enum { OP_ID = 0, OP_CONJ = 1 };

/* ------------------------------------------------------------------ */
/* scalar helpers                                                      */
/* ------------------------------------------------------------------ */

static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

float orc_round_bf16(float x)
{
    uint32_t u = f2u(x);
    if ((u & 0x7f800000u) == 0x7f800000u) {          /* inf / nan */
        if (u & 0x007fffffu) u |= 0x00400000u;        /* quiet the nan */
        return u2f(u & 0xffff0000u);
    }
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;                               /* round to nearest even */
    return u2f(u & 0xffff0000u);
}

/* binary16 <-> binary32 by bit manipulation (IEEE 754-2008 binary16). */
static float half_bits_to_float(uint16_t h)
{
    uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 0x1f, m = h & 0x3ff;
    if (e == 0) {
        if (m == 0) return u2f(s);
        float v = ldexpf((float)m, -24);               /* subnormal */
        return s ? -v : v;
    }
    if (e == 31) return u2f(s | 0x7f800000u | (m << 13));
    return u2f(s | ((e - 15 + 127) << 23) | (m << 13));
}

static uint16_t float_to_half_bits(float x)
{
    uint32_t u = f2u(x), s = (u >> 16) & 0x8000u;
    uint32_t a = u & 0x7fffffffu;
    if (a >= 0x7f800000u) return (uint16_t)(s | 0x7c00u | (a > 0x7f800000u ? 0x200u : 0));
    if (a >= 0x477ff000u) return (uint16_t)(s | 0x7c00u); /* rounds to >= 65520: inf */
    if (a < 0x38800000u) {                              /* result subnormal or zero */
        /* value = a * 2^-24 units; round-to-nearest-even of x / 2^-24 */
        float v = u2f(a);
        float q = v * 16777216.0f;                      /* exact scaling by 2^24 */
        float r = rintf(q);                              /* current mode: nearest even */
        return (uint16_t)(s | (uint32_t)r);
    }
    uint32_t e = ((a >> 23) - 127 + 15), m = a & 0x7fffffu;
    uint32_t hm = m >> 13, rem = m & 0x1fffu;
    uint32_t h = (e << 10) | hm;
    if (rem > 0x1000u || (rem == 0x1000u && (hm & 1u))) h += 1;  /* may carry into exponent */
    return (uint16_t)(s | h);
}

float orc_round_f16(float x) { return half_bits_to_float(float_to_half_bits(x)); }

static float bf16_bits_to_float(uint16_t b) { return u2f((uint32_t)b << 16); }

/* round-to-nearest-even of a binary64 value to p significant bits with
 * smallest normal exponent emin (subnormal spacing 2^(emin - p + 1)) and
 * overflow to infinity at or above `ovf`: one rounding from the exact
 * value, as a conversion from binary64 or binary32 does */
static double rnd_bits(double x, int p, int emin, double ovf)
{
    if (!isfinite(x) || x == 0.0) return x;
    double a = fabs(x);
    int e;
    frexp(a, &e);                    /* a = f 2^e, f in [0.5, 1): ulp exponent e - p */
    int q = e - p;
    if (q < emin - p + 1) q = emin - p + 1;
    double r = ldexp(rint(ldexp(a, -q)), q);
    if (r >= ovf) r = INFINITY;
    return copysign(r, x);
}


static bool is_complex_type(int t) { return t == T_C32 || t == T_C64; }

/* Reads element `off` of a tensor of storage type `type` as complex double.
 * Storage values are exactly representable in binary64. */
static double complex load_z(const void* base, int type, int64_t off)
{
    switch (type) {
    case T_F32: return ((const float*)base)[off];
    case T_F64: return ((const double*)base)[off];
    case T_C32: { const float* p = (const float*)base + 2 * off; return CMPLX(p[0], p[1]); }
    case T_C64: { const double* p = (const double*)base + 2 * off; return CMPLX(p[0], p[1]); }
    case T_F16: return half_bits_to_float(((const uint16_t*)base)[off]);
    case T_BF16: return bf16_bits_to_float(((const uint16_t*)base)[off]);
    default: return 0;
    }
}

/* is_equal(beta, type_D, 0.0f) of product.c:1675-1861: the scalar is read as
 * type_D and compared with zero. */
static bool scalar_is_zero(const void* s, int type)
{
    if (!s) return true;
    double complex z = load_z(s, type, 0);
    return creal(z) == 0.0 && cimag(z) == 0.0;
}

/* ------------------------------------------------------------------ */
/* validation: product.c:197-206 and helpers :757-967                  */
/* ------------------------------------------------------------------ */

/* check_idx_occurrence, product.c:757-784 */
static int chk_occurrence(const orc_tensor* o, const orc_tensor* a, const orc_tensor* b, int code)
{
    for (int i = 0; i < o->nmode; i++) {
        bool hit = false;
        for (int j = 0; j < a->nmode && !hit; j++) hit = o->idx[i] == a->idx[j];
        for (int j = 0; j < b->nmode && !hit; j++) hit = o->idx[i] == b->idx[j];
        if (!hit) return code;
    }
    return 0;
}

/* check_extents, product.c:786-813: per mode of x, scan x, then y, then z. */
static int chk_extents(const orc_tensor* x, const orc_tensor* y, const orc_tensor* z,
                       int code_xx, int code_xy, int code_xz)
{
    for (int i = 0; i < x->nmode; i++) {
        for (int j = 0; j < x->nmode; j++)
            if (x->idx[i] == x->idx[j] && x->extents[i] != x->extents[j]) return code_xx;
        for (int j = 0; j < y->nmode; j++)
            if (x->idx[i] == y->idx[j] && x->extents[i] != y->extents[j]) return code_xy;
        for (int j = 0; j < z->nmode; j++)
            if (x->idx[i] == z->idx[j] && x->extents[i] != z->extents[j]) return code_xz;
    }
    return 0;
}

/* check_same_structure(C, D), product.c:815-834 (strides not compared). */
static int chk_structure(const orc_tensor* c, const orc_tensor* d)
{
    if (c->nmode != d->nmode) return 5;
    for (int i = 0; i < d->nmode; i++) {
        if (d->idx[i] != c->idx[i]) return 6;
        if (d->extents[i] != c->extents[i]) return 7;
    }
    return 0;
}

/* check_self_aliasing, product.c:836-870.  The reference's merge sort moves
 * values through `int` temporaries (merge_strides, :905-918), so every
 * |stride| and extent it compares has been truncated to 32 bits; the order
 * is ascending stride, ties by ascending extent.  Equal (stride, extent)
 * pairs are interchangeable, so any sort on that key gives the same
 * sequence. */
typedef struct { int32_t s, e; } se_pair;
static int cmp_se(const void* a, const void* b)
{
    const se_pair* x = a; const se_pair* y = b;
    if (x->s != y->s) return x->s < y->s ? -1 : 1;
    if (x->e != y->e) return x->e < y->e ? -1 : 1;
    return 0;
}
static int chk_aliasing(const orc_tensor* d)
{
    int n = d->nmode;
    if (n <= 1) return 0;
    for (int i = 0; i < n; i++) if (d->strides[i] == 0) return 8;
    se_pair* v = malloc(sizeof(se_pair) * (size_t)n);
    for (int i = 0; i < n; i++) {
        int64_t s = d->strides[i] < 0 ? -d->strides[i] : d->strides[i];
        v[i].s = (int32_t)(uint32_t)(uint64_t)s;
        v[i].e = (int32_t)(uint32_t)(uint64_t)d->extents[i];
    }
    qsort(v, (size_t)n, sizeof(se_pair), cmp_se);
    int rc = 0;
    for (int i = 0; i + 1 < n; i++)
        if ((int64_t)v[i + 1].s < (int64_t)v[i].s * (int64_t)v[i].e) { rc = 8; break; }
    free(v);
    return rc;
}

int orc_check(const orc_tensor* A, const orc_tensor* B, const orc_tensor* C,
              const orc_tensor* D, const void* beta, const void* C_data,
              const int* exec_kind)
{
    int rc = chk_occurrence(D, A, B, 4);
    if (!rc) rc = chk_extents(A, B, D, 9, 1, 2);
    if (!rc) rc = chk_extents(B, A, D, 10, 1, 3);
    if (!rc) rc = chk_extents(D, A, B, 11, 2, 3);
    if (!rc) rc = chk_structure(C, D);
    if (!rc) rc = chk_aliasing(D);
    if (!rc && C_data == NULL && !scalar_is_zero(beta, D->type)) rc = 12;   /* :872-876 */
    if (!rc) {                                                              /* :878-885 */
        if (!exec_kind) rc = 33;
        else if (!(*exec_kind == 1 || *exec_kind == 2 || *exec_kind == 12)) rc = 33;
    }
    return rc;
}

/* ------------------------------------------------------------------ */
/* plan stage: product.c:265-310 and helpers :538-715, :969-1010      */
/* ------------------------------------------------------------------ */

typedef struct {
    int n;
    int64_t* idx; int64_t* ext; int64_t* str;
} modes;

/* compress_repeated_indices, product.c:969-1010: first occurrence keeps its
 * place and extent; its stride becomes the sum over all occurrences. */
static modes compress(const orc_tensor* t)
{
    modes m;
    size_t cap = t->nmode > 0 ? (size_t)t->nmode : 1;
    m.idx = malloc(cap * 8); m.ext = malloc(cap * 8); m.str = malloc(cap * 8);
    m.n = 0;
    for (int i = 0; i < t->nmode; i++) {
        bool seen = false;
        for (int j = 0; j < m.n; j++) seen |= m.idx[j] == t->idx[i];
        if (seen) continue;
        int64_t s = 0;
        for (int j = 0; j < t->nmode; j++) if (t->idx[j] == t->idx[i]) s += t->strides[j];
        m.idx[m.n] = t->idx[i]; m.ext[m.n] = t->extents[i]; m.str[m.n] = s; m.n++;
    }
    return m;
}
static void modes_free(modes* m) { free(m->idx); free(m->ext); free(m->str); }

static int find(const modes* m, int64_t label)
{
    int at = -1;
    for (int j = 0; j < m->n; j++) if (m->idx[j] == label) at = j;   /* last match, as :629-633 */
    return at;
}

/* calculate_size, product.c:692-700: the product is kept in a 32-bit int. */
static int64_t size32(const int64_t* ext, int n)
{
    int32_t s = 1;
    for (int i = 0; i < n; i++) s = (int32_t)(uint32_t)((uint64_t)(int64_t)s * (uint64_t)ext[i]);
    return s;
}

typedef struct {
    modes a, b, c, d;
    int nk, nua, nub;
    int64_t *k_ext, *k_sa, *k_sb;       /* contracted, idx_A order (:538-570, :643-657) */
    int64_t *ua_ext, *ua_s;             /* unary A (:588-619) */
    int64_t *ub_ext, *ub_s;             /* unary B */
    int64_t *fa, *fb;                   /* per D position stride of A/B or 0 (:621-641) */
    int64_t size_free, size_k, size_ua, size_ub;
} plan;

static void plan_build(const orc_tensor* A, const orc_tensor* B, const orc_tensor* C,
                       const orc_tensor* D, plan* p)
{
    p->a = compress(A); p->b = compress(B); p->c = compress(C); p->d = compress(D);
    int na = p->a.n, nb = p->b.n, nd = p->d.n;
    size_t cap = (size_t)(na + nb + nd + 1);
    p->k_ext = calloc(cap, 8); p->k_sa = calloc(cap, 8); p->k_sb = calloc(cap, 8);
    p->ua_ext = calloc(cap, 8); p->ua_s = calloc(cap, 8);
    p->ub_ext = calloc(cap, 8); p->ub_s = calloc(cap, 8);
    p->fa = calloc(cap, 8); p->fb = calloc(cap, 8);
    p->nk = p->nua = p->nub = 0;
    for (int i = 0; i < na; i++) {
        int64_t l = p->a.idx[i];
        int ib = find(&p->b, l), id = find(&p->d, l);
        if (ib >= 0 && id < 0) {
            p->k_ext[p->nk] = p->a.ext[i]; p->k_sa[p->nk] = p->a.str[i]; p->k_sb[p->nk] = p->b.str[ib];
            p->nk++;
        } else if (ib < 0 && id < 0) {
            p->ua_ext[p->nua] = p->a.ext[i]; p->ua_s[p->nua] = p->a.str[i]; p->nua++;
        }
    }
    for (int i = 0; i < nb; i++) {
        int64_t l = p->b.idx[i];
        if (find(&p->a, l) < 0 && find(&p->d, l) < 0) {
            p->ub_ext[p->nub] = p->b.ext[i]; p->ub_s[p->nub] = p->b.str[i]; p->nub++;
        }
    }
    for (int j = 0; j < nd; j++) {
        int ia = find(&p->a, p->d.idx[j]), ib = find(&p->b, p->d.idx[j]);
        p->fa[j] = ia >= 0 ? p->a.str[ia] : 0;
        p->fb[j] = ib >= 0 ? p->b.str[ib] : 0;
    }
    p->size_free = size32(p->d.ext, nd);
    p->size_k = size32(p->k_ext, p->nk);
    p->size_ua = size32(p->ua_ext, p->nua);
    p->size_ub = size32(p->ub_ext, p->nub);
}

static void plan_free(plan* p)
{
    modes_free(&p->a); modes_free(&p->b); modes_free(&p->c); modes_free(&p->d);
    free(p->k_ext); free(p->k_sa); free(p->k_sb); free(p->ua_ext); free(p->ua_s);
    free(p->ub_ext); free(p->ub_s); free(p->fa); free(p->fb);
}

int orc_classify(const orc_tensor* A, const orc_tensor* B, const orc_tensor* C,
                 const orc_tensor* D, orc_classes* out)
{
    plan p;
    plan_build(A, B, C, D, &p);
    memset(out, 0, sizeof *out);
    if (p.nk > 64 || p.nua > 64 || p.nub > 64 || p.d.n > 64) { plan_free(&p); return -1; }
    out->n_contracted = p.nk; out->n_unary_A = p.nua; out->n_unary_B = p.nub; out->nmode_D = p.d.n;
    int k = 0, ua = 0, ub = 0;
    for (int i = 0; i < p.a.n; i++) {
        int64_t l = p.a.idx[i];
        int ib = find(&p.b, l), id = find(&p.d, l);
        if (ib >= 0 && id < 0) out->contracted[k++] = l;
        else if (ib < 0 && id < 0) out->unary_A[ua++] = l;
    }
    for (int i = 0; i < p.b.n; i++)
        if (find(&p.a, p.b.idx[i]) < 0 && find(&p.d, p.b.idx[i]) < 0) out->unary_B[ub++] = p.b.idx[i];
    for (int i = 0; i < p.nk; i++) {
        out->ext_contracted[i] = p.k_ext[i];
        out->contracted_strides_A[i] = p.k_sa[i]; out->contracted_strides_B[i] = p.k_sb[i];
    }
    for (int i = 0; i < p.nua; i++) out->unary_strides_A[i] = p.ua_s[i];
    for (int i = 0; i < p.nub; i++) out->unary_strides_B[i] = p.ub_s[i];
    for (int j = 0; j < p.d.n; j++) {
        out->free_strides_A[j] = p.fa[j]; out->free_strides_B[j] = p.fb[j];
        out->strides_D[j] = p.d.str[j]; out->extents_D[j] = p.d.ext[j];
        out->strides_C[j] = j < p.c.n ? p.c.str[j] : 0;
    }
    plan_free(&p);
    return 0;
}

/* increment_coordinates, product.c:702-715: mixed-radix odometer, first
 * coordinate fastest, wrapping to zero after the last. */
static void odometer(int64_t* c, int n, const int64_t* ext)
{
    for (int k = 0; k < n; k++) {
        c[k] = (c[k] + 1) % ext[k];
        if (c[k] != 0) return;
    }
}

/* ------------------------------------------------------------------ */
/* loop nest: product.c:312-399, instantiated per compute type        */
/* ------------------------------------------------------------------ */

/* Per-call semantics resolved from precision and types (alloc_accum :1012,
 * alloc_val :1104, create_prec_scalar :1232, sum_unary_contractions :1863,
 * calculate_beta_C_prec :3156, calculate_op_D :10737). */
typedef struct {
    int op_A, op_B, op_C, op_D;
    int round;            /* 0 none, 1 binary16, 2 bfloat16: applied to alpha, beta, A, B, C */
    bool conjf_roundtrip; /* op_D conj goes through conjf (F64F64 prec, :10814) */
    const void *alpha, *beta;
    int type_alpha;       /* alpha and beta are read as type_D (:330-331) */
    const void *A, *B, *C; void* D;
    int tA, tB, tC, tD;
    bool beta_zero;
} semantics;

static double complex rnd(double complex z, int mode)
{
    if (mode == 1) return CMPLX(orc_round_f16((float)creal(z)), orc_round_f16((float)cimag(z)));
    if (mode == 2) return CMPLX(orc_round_bf16((float)creal(z)), orc_round_bf16((float)cimag(z)));
    return z;
}

static double complex opz(double complex z, int op) { return op == OP_CONJ ? conj(z) : z; }

static void store_z(void* base, int type, int64_t off, double complex v)
{
    switch (type) {
    case T_F32: ((float*)base)[off] = (float)creal(v); break;
    case T_F64: ((double*)base)[off] = creal(v); break;
    case T_C32: { float* p = (float*)base + 2 * off; p[0] = (float)creal(v); p[1] = (float)cimag(v); } break;
    case T_C64: { double* p = (double*)base + 2 * off; p[0] = creal(v); p[1] = cimag(v); } break;
    /* one rounding from the accumulator's value, as the conversion in
     * assign_D does (a binary32 accumulator converts exactly to binary64) */
    case T_F16: ((uint16_t*)base)[off] = float_to_half_bits((float)rnd_bits(creal(v), 11, -14, 65520.0)); break;
    case T_BF16:
        ((uint16_t*)base)[off] =
            (uint16_t)(f2u((float)rnd_bits(creal(v), 8, -126, ldexp(2.0 - ldexp(1.0, -8), 127))) >> 16);
        break;
    default: break;
    }
}

/* The loop nest for compute type T.  Loads convert storage -> T (the value
 * the reference's get_val/sum_unary_contractions would hold), then every
 * operation happens in T in the reference's order:
 *   accum = beta * op_C(C) | 0                      (:351-360)
 *   for each contracted coordinate (first fastest):  (:361-395)
 *     sum_A = 0; sum_A += op_A(A) over unary A      (:370-380)
 *     sum_B = 0; sum_B += op_B(B) over unary B      (:381-391)
 *       (the addend keeps its storage precision: C promotes the sum, adds
 *       and converts back, e.g. float sum += double value, :1903)
 *     accum += alpha * sum_A * sum_B                (:393, left to right)
 *   accum = op_D(accum); D = accum                  (:396-397)        */
#define ORC_LOOP(NAME, T, CONV)                                                        \
static void NAME(const plan* p, const semantics* s)                                    \
{                                                                                      \
    int nd = p->d.n;                                                                   \
    int64_t* cf = calloc((size_t)nd + 1, 8);                                           \
    int64_t* ck = calloc((size_t)p->nk + 1, 8);                                        \
    int64_t* cua = calloc((size_t)p->nua + 1, 8);                                      \
    int64_t* cub = calloc((size_t)p->nub + 1, 8);                                      \
    T alpha = CONV(rnd(s->alpha ? load_z(s->alpha, s->type_alpha, 0) : 0, s->round));  \
    T beta = CONV(rnd(s->beta ? load_z(s->beta, s->type_alpha, 0) : 0, s->round));     \
    for (int64_t i = 0; i < p->size_free; i++) {                                       \
        int64_t oa = 0, ob = 0, oc = 0, od = 0;                                        \
        for (int j = 0; j < nd; j++) {                                                 \
            oa += cf[j] * p->fa[j]; ob += cf[j] * p->fb[j];                            \
            oc += cf[j] * p->c.str[j]; od += cf[j] * p->d.str[j];                      \
        }                                                                              \
        T acc;                                                                         \
        if (!s->beta_zero)                                                             \
            acc = beta * CONV(opz(rnd(load_z(s->C, s->tC, oc), s->round), s->op_C));   \
        else                                                                           \
            acc = 0;                                                                   \
        for (int64_t j = 0; j < p->size_k; j++) {                                      \
            int64_t ka = oa, kb = ob;                                                  \
            for (int k = 0; k < p->nk; k++) { ka += ck[k] * p->k_sa[k]; kb += ck[k] * p->k_sb[k]; } \
            T sa = 0, sb = 0;                                                          \
            for (int64_t u = 0; u < p->size_ua; u++) {                                 \
                int64_t x = ka;                                                        \
                for (int l = 0; l < p->nua; l++) x += cua[l] * p->ua_s[l];             \
                sa = CONV((double complex)sa + opz(rnd(load_z(s->A, s->tA, x), s->round), s->op_A)); \
                odometer(cua, p->nua, p->ua_ext);                                      \
            }                                                                          \
            for (int64_t u = 0; u < p->size_ub; u++) {                                 \
                int64_t x = kb;                                                        \
                for (int l = 0; l < p->nub; l++) x += cub[l] * p->ub_s[l];             \
                sb = CONV((double complex)sb + opz(rnd(load_z(s->B, s->tB, x), s->round), s->op_B)); \
                odometer(cub, p->nub, p->ub_ext);                                      \
            }                                                                          \
            if (s->round) { sa = CONV(rnd(sa, s->round)); sb = CONV(rnd(sb, s->round)); } \
            acc += alpha * sa * sb;                                                    \
            odometer(ck, p->nk, p->k_ext);                                             \
        }                                                                              \
        double complex out = acc;                                                      \
        if (s->op_D == OP_CONJ) {                                                      \
            if (s->conjf_roundtrip) out = conjf((float complex)out);                   \
            else out = conj(out);                                                      \
        }                                                                              \
        store_z(s->D, s->tD, od, out);                                                 \
        odometer(cf, nd, p->d.ext);                                                    \
    }                                                                                  \
    free(cf); free(ck); free(cua); free(cub);                                          \
}

#define CONV_F(z) ((float)creal(z))
#define CONV_D(z) (creal(z))
#define CONV_CF(z) ((float complex)(z))
#define CONV_CD(z) ((double complex)(z))

ORC_LOOP(loop_f32, float, CONV_F)
ORC_LOOP(loop_f64, double, CONV_D)
ORC_LOOP(loop_c32, float complex, CONV_CF)
ORC_LOOP(loop_c64, double complex, CONV_CD)

/* ------------------------------------------------------------------ */
/* the loop nest with binary16 / bfloat16 values (product.c:312-399)  */
/* ------------------------------------------------------------------ */

/* Which precision every value of the loop nest has.  The reference keeps
 * each quantity in a type chosen per precision branch: accum by alloc_accum
 * (:1012-1102), sum_A / sum_B / val_C by alloc_val (:1104-1230), alpha and
 * beta by create_prec_scalar (:1232-1490).  Those types reach _Float16
 * (TAPP_F16F16_* precisions, and DEFAULT_PREC with binary16 storage) and
 * __bf16.  The reference's F16 build is compiled by clang for x86-64, which
 * has no binary16 arithmetic: every _Float16 operation is evaluated in
 * binary32 and rounded to binary16 only at an assignment or when the
 * _Float16-typed result of an expression is converted to another type
 * (clang's default -ffloat16-excess-precision=standard).  Within one
 * expression clang may contract a*b + c into llvm.fmuladd, which baseline
 * x86-64 (no FMA) executes as a rounded multiply then a rounded add.  The
 * rules below restate exactly that evaluation, one rounding per
 * operation.  Every value is carried in binary64, which holds each of these
 * formats exactly; binary32 operations are computed in binary64 and rounded
 * once to binary32, which for +, -, * of binary32 operands gives the
 * binary32 result (53 >= 2 * 24 + 2). */
enum { K_H = 1, K_B = 2, K_F = 3, K_D = 4 };   /* _Float16, __bf16, float, double */

typedef struct { int k; bool c; } slot;      /* precision, complex */

static int kind_of_storage(int t)
{
    switch (t) {
    case T_F16: return K_H;
    case T_BF16: return K_B;
    case T_F32: case T_C32: return K_F;
    default: return K_D;
    }
}

/* the common (semantic) type of a binary operation */
static int kmax(int a, int b)
{
    if (a == b) return a;
    if (a == K_D || b == K_D) return K_D;
    return K_F;                      /* float with anything narrower; _Float16 with __bf16 */
}

/* the type an operation is evaluated in */
static int keval(int k) { return k == K_H || k == K_B ? K_F : k; }
static bool lowp(int k) { return k == K_H || k == K_B; }

static double rnd_k(double x, int k)
{
    switch (k) {
    case K_H: return rnd_bits(x, 11, -14, 65520.0);
    case K_B: return rnd_bits(x, 8, -126, ldexp(2.0 - ldexp(1.0, -8), 127));
    case K_F: return (double)(float)x;
    default: return x;
    }
}

static double complex rnd_z(double complex z, slot s)
{
    return CMPLX(rnd_k(creal(z), s.k), s.c ? rnd_k(cimag(z), s.k) : 0.0);
}

/* x * y evaluated in precision k (C's textbook complex product, one rounding
 * per operation; a real operand multiplies each part) */
static double complex mul_k(double complex x, bool xc, double complex y, bool yc, int k)
{
    double xr = creal(x), xi = cimag(x), yr = creal(y), yi = cimag(y);
    if (xc && yc)
        return CMPLX(rnd_k(rnd_k(xr * yr, k) - rnd_k(xi * yi, k), k), rnd_k(rnd_k(xi * yr, k) + rnd_k(xr * yi, k), k));
    if (xc) return CMPLX(rnd_k(xr * yr, k), rnd_k(xi * yr, k));
    if (yc) return CMPLX(rnd_k(xr * yr, k), rnd_k(xr * yi, k));
    return rnd_k(xr * yr, k);
}

/* dst op= (value v of semantic slot vs): C's compound assignment with the
 * excess-precision rule above.  A _Float16 / __bf16 right-hand side is
 * rounded to its own type first unless dst has that same type. */
static double complex assign_add(double complex dst, slot ds, double complex v, slot vs)
{
    if (lowp(vs.k) && vs.k != ds.k) v = rnd_z(v, vs);
    const int e = keval(kmax(ds.k, vs.k));
    double re = rnd_k(rnd_k(creal(dst) + creal(v), e), ds.k);
    double im = cimag(dst);
    if (ds.c && vs.c) im = rnd_k(rnd_k(cimag(dst) + cimag(v), e), ds.k);
    return CMPLX(re, ds.c ? im : 0.0);
}

static double complex assign_set(slot ds, double complex v, slot vs)
{
    if (lowp(vs.k) && vs.k != ds.k) v = rnd_z(v, vs);
    return CMPLX(rnd_k(creal(v), ds.k), ds.c && vs.c ? rnd_k(cimag(v), ds.k) : 0.0);
}

typedef struct {
    slot acc, sa, sb, vc, scal;   /* accum, sum_A, sum_B, val_C, alpha/beta */
    int conj_sums;                /* sum_unary_contractions applies op_A / op_B */
    bool no_store;                /* F16F16_ACCUM_F16 with complex D: assign_D stores nothing */
    bool gpu_products;            /* products and accumulation in binary32 (not the reference) */
} slots;

/* Slot precisions of one call (alloc_accum :1012, alloc_val :1104,
 * create_prec_scalar :1232, sum_unary_contractions :1863, assign_D
 * :10989-11118). */
static void choose_slots(int prec, const orc_tensor* A, const orc_tensor* B, const orc_tensor* C,
                         const orc_tensor* D, slots* q)
{
    int k_acc, k_val;
    memset(q, 0, sizeof *q);
    switch (prec) {
    case P_DEFAULT: k_acc = kind_of_storage(D->type); k_val = 0; break;
    case P_F32: case P_F16F16_F32: case P_BF16BF16_F32: k_acc = K_F; break;
    case P_F64: k_acc = K_D; break;
    default: k_acc = K_H; break;                 /* P_F16F16_F16 */
    }
    switch (prec) {
    case P_DEFAULT: k_val = 0; break;
    case P_F32: k_val = K_F; break;
    case P_F64: k_val = K_D; break;
    case P_BF16BF16_F32: k_val = K_B; break;
    default: k_val = K_H; break;
    }
    q->acc = (slot){k_acc, is_complex_type(D->type)};
    q->sa = (slot){k_val ? k_val : kind_of_storage(A->type), is_complex_type(A->type)};
    q->sb = (slot){k_val ? k_val : kind_of_storage(B->type), is_complex_type(B->type)};
    q->vc = (slot){k_val ? k_val : kind_of_storage(C->type), is_complex_type(C->type)};
    q->scal = (slot){k_val ? k_val : kind_of_storage(D->type), is_complex_type(D->type)};
    q->conj_sums = prec == P_DEFAULT || prec == P_F32 || prec == P_F64;
    q->no_store = prec == P_F16F16_F16 && is_complex_type(D->type);
}

static bool slots_lowp(const slots* q)
{
    return lowp(q->acc.k) || lowp(q->sa.k) || lowp(q->sb.k) || lowp(q->vc.k) || lowp(q->scal.k);
}

/* The loop nest of product.c:312-399 over slot precisions: the same
 * operation order as ORC_LOOP, with each value rounded as its slot and each
 * expression as the rules above. */
static void loop_generic(const plan* p, const semantics* s, const slots* q)
{
    int nd = p->d.n;
    int64_t* cf = calloc((size_t)nd + 1, 8);
    int64_t* ck = calloc((size_t)p->nk + 1, 8);
    int64_t* cua = calloc((size_t)p->nua + 1, 8);
    int64_t* cub = calloc((size_t)p->nub + 1, 8);
    slot acc_s = q->acc, sa_s = q->sa, sb_s = q->sb;
    if (q->gpu_products) { acc_s.k = K_F; }
    const double complex alpha = rnd_z(s->alpha ? load_z(s->alpha, s->type_alpha, 0) : 0, q->scal);
    const double complex beta = rnd_z(s->beta ? load_z(s->beta, s->type_alpha, 0) : 0, q->scal);
    const int kA = kind_of_storage(s->tA), kB = kind_of_storage(s->tB);
    /* alpha * sum_A * sum_B, left to right (:3431-10735) */
    const slot ab = {kmax(q->scal.k, sa_s.k), q->scal.c || sa_s.c};
    slot prod = {kmax(ab.k, sb_s.k), ab.c || sb_s.c};
    const int e_ab = keval(ab.k), e_prod = keval(prod.k);
    if (q->gpu_products) prod.k = K_F;
    const slot bc = {kmax(q->scal.k, q->vc.k), q->scal.c || q->vc.c};
    for (int64_t i = 0; i < p->size_free; i++) {
        int64_t oa = 0, ob = 0, oc = 0, od = 0;
        for (int j = 0; j < nd; j++) {
            oa += cf[j] * p->fa[j]; ob += cf[j] * p->fb[j];
            oc += cf[j] * p->c.str[j]; od += cf[j] * p->d.str[j];
        }
        double complex acc = 0;
        if (!s->beta_zero) {                                            /* :351-360 */
            double complex c = rnd_z(load_z(s->C, s->tC, oc), q->vc);   /* get_val :10830 */
            c = opz(c, s->op_C);
            slot bcs = bc;
            if (q->gpu_products) bcs.k = K_F;
            acc = assign_set(acc_s, mul_k(beta, q->scal.c, c, q->vc.c, keval(bc.k)), bcs);
        }
        for (int64_t j = 0; j < p->size_k; j++) {
            int64_t ka = oa, kb = ob;
            for (int k = 0; k < p->nk; k++) { ka += ck[k] * p->k_sa[k]; kb += ck[k] * p->k_sb[k]; }
            double complex sa = 0, sb = 0;                              /* zero_sum :1516 */
            for (int64_t u = 0; u < p->size_ua; u++) {                  /* :370-380, :1863 */
                int64_t x = ka;
                for (int l = 0; l < p->nua; l++) x += cua[l] * p->ua_s[l];
                double complex v = load_z(s->A, s->tA, x);
                if (q->conj_sums) v = opz(v, s->op_A);
                sa = assign_add(sa, sa_s, v, (slot){kA, is_complex_type(s->tA)});
                odometer(cua, p->nua, p->ua_ext);
            }
            for (int64_t u = 0; u < p->size_ub; u++) {                  /* :381-391 */
                int64_t x = kb;
                for (int l = 0; l < p->nub; l++) x += cub[l] * p->ub_s[l];
                double complex v = load_z(s->B, s->tB, x);
                if (q->conj_sums) v = opz(v, s->op_B);
                sb = assign_add(sb, sb_s, v, (slot){kB, is_complex_type(s->tB)});
                odometer(cub, p->nub, p->ub_ext);
            }
            double complex t = mul_k(alpha, q->scal.c, sa, sa_s.c, e_ab);
            t = mul_k(t, ab.c, sb, sb_s.c, e_prod);
            acc = assign_add(acc, acc_s, t, prod);                       /* :393 */
            odometer(ck, p->nk, p->k_ext);
        }
        if (s->op_D == OP_CONJ) {                                       /* :10737-10828 */
            if (s->conjf_roundtrip) acc = conjf((float complex)acc);
            else acc = conj(acc);
        }
        if (!q->no_store) store_z(s->D, s->tD, od, acc);                /* :10989-11118 */
        odometer(cf, nd, p->d.ext);
    }
    free(cf); free(ck); free(cua); free(cub);
}

static bool known_type(int t) { return t >= T_F32 && t <= T_BF16; }

/* Picks the compute type from the precision and the storage types, and the
 * element-op behaviour of each precision branch of the reference. */
static int resolve(const orc_tensor* A, const orc_tensor* B, const orc_tensor* C,
                   const orc_tensor* D, int prec, semantics* s, int* wide, int* cplx)
{
    if (!known_type(A->type) || !known_type(B->type) || !known_type(C->type) || !known_type(D->type))
        return ORC_UNSUPPORTED;
    *cplx = is_complex_type(A->type) || is_complex_type(B->type) ||
            is_complex_type(C->type) || is_complex_type(D->type);
    s->op_A = A->op; s->op_B = B->op; s->op_C = C->op; s->op_D = D->op;
    s->round = 0; s->conjf_roundtrip = false;
    switch (prec) {
    case P_DEFAULT:                               /* accum has type_D (:1016-1044) */
        *wide = D->type == T_F64 || D->type == T_C64;
        break;
    case P_F32:                                   /* :1045-1071; no op_C in _prec (:3156) */
        *wide = 0; s->op_C = OP_ID;
        break;
    case P_F64:                                   /* :1072-1092; op_D via conjf (:10814) */
        *wide = 1; s->op_C = OP_ID; s->conjf_roundtrip = true;
        break;
    case P_F16F16_F16:                            /* no conj anywhere (:1965-1969, :10823) */
    case P_F16F16_F32:
    case P_BF16BF16_F32:
        *wide = 0; s->round = prec == P_BF16BF16_F32 ? 2 : 1;
        s->op_A = OP_ID; s->op_B = OP_ID; s->op_C = OP_ID;
        if (prec == P_F16F16_F16) s->op_D = OP_ID;
        break;
    default:
        return ORC_UNSUPPORTED;
    }
    if (!is_complex_type(D->type)) s->op_D = OP_ID;   /* conj of a real accum: no-op */
    if (!is_complex_type(A->type)) s->op_A = OP_ID;
    if (!is_complex_type(B->type)) s->op_B = OP_ID;
    if (!is_complex_type(C->type)) s->op_C = OP_ID;
    return 0;
}

int orc_execute_ex(const orc_tensor* A, const orc_tensor* B, const orc_tensor* C,
                   const orc_tensor* D, int prec,
                   const void* alpha, const void* A_data, const void* B_data,
                   const void* beta, const void* C_data, void* D_data,
                   const int* exec_kind, int flags)
{
    int rc = orc_check(A, B, C, D, beta, C_data, exec_kind);
    if (rc) return rc;
    semantics s; int wide, cplx;
    rc = resolve(A, B, C, D, prec, &s, &wide, &cplx);
    if (rc) return rc;
    s.alpha = alpha; s.beta = beta; s.type_alpha = D->type;
    s.A = A_data; s.B = B_data; s.C = C_data; s.D = D_data;
    s.tA = A->type; s.tB = B->type; s.tC = C->type; s.tD = D->type;
    s.beta_zero = scalar_is_zero(beta, D->type);
    slots q;
    choose_slots(prec, A, B, C, D, &q);
    q.gpu_products = (flags & ORC_GPU_PRODUCTS) != 0;
    if (q.gpu_products) q.no_store = false;
    plan p;
    plan_build(A, B, C, D, &p);
    /* the per-type loops hold every value in one compute type, which is the
     * reference's only when all four storage types are equal and no value is
     * binary16 / bfloat16; otherwise each value keeps its own slot type */
    const bool uniform = A->type == B->type && A->type == C->type && A->type == D->type;
    if (!uniform || slots_lowp(&q) || (flags & (ORC_GENERIC | ORC_GPU_PRODUCTS))) loop_generic(&p, &s, &q);
    else if (cplx) { if (wide) loop_c64(&p, &s); else loop_c32(&p, &s); }
    else           { if (wide) loop_f64(&p, &s); else loop_f32(&p, &s); }
    plan_free(&p);
    return 0;
}

int orc_execute(const orc_tensor* A, const orc_tensor* B, const orc_tensor* C,
                const orc_tensor* D, int prec,
                const void* alpha, const void* A_data, const void* B_data,
                const void* beta, const void* C_data, void* D_data,
                const int* exec_kind)
{
    return orc_execute_ex(A, B, C, D, prec, alpha, A_data, B_data, beta, C_data, D_data, exec_kind, 0);
}

int orc_execute_f64(const orc_tensor* A, const orc_tensor* B, const orc_tensor* C,
                    const orc_tensor* D, int prec,
                    const void* alpha, const void* A_data, const void* B_data,
                    const void* beta, const void* C_data, void* D_data,
                    int round_inputs)
{
    int kind = 1;
    int rc = orc_check(A, B, C, D, beta, C_data, &kind);
    if (rc) return rc;
    semantics s; int wide, cplx;
    rc = resolve(A, B, C, D, prec, &s, &wide, &cplx);
    if (rc) return rc;
    s.round = round_inputs;
    s.conjf_roundtrip = false;
    s.alpha = alpha; s.beta = beta; s.type_alpha = D->type;
    s.A = A_data; s.B = B_data; s.C = C_data; s.D = D_data;
    s.tA = A->type; s.tB = B->type; s.tC = C->type; s.tD = D->type;
    s.beta_zero = scalar_is_zero(beta, D->type);
    plan p;
    plan_build(A, B, C, D, &p);
    if (cplx) loop_c64(&p, &s); else loop_f64(&p, &s);
    plan_free(&p);
    return 0;
}
