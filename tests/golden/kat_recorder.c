/*
 * Known-answer capture for the reference's deterministic programs
 * (test/demo.c, examples/driver/driver.c,
 * examples/exercise_contraction/answers/exercise_contraction_answers.c).
 *
 * Those programs print their results but assert nothing.  Linked against
 * this recorder instead of the reference library, they run unchanged: every
 * TAPP_* call is forwarded to the compiled reference (oracle/_ref, loaded
 * with RTLD_LOCAL), and each TAPP_execute_product call is written as one
 * JSON line -- layouts, labels, ops, precision, scalars, the input data over
 * each operand's address span, the return code and D after the call -- to
 * the file named by $TAPP_KAT_OUT.  make_golden.py turns these lines into
 * tests/golden/kat_reference_programs.json.
 */
#include <complex.h>
#include <dlfcn.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef intptr_t H;

static void* ref;
static FILE* out;

#define REF(name) ((fn_##name)resolve(#name))
static void* resolve(const char* name)
{
    if (!ref) {
        const char* path = getenv("TAPP_REF_LIB");
        ref = dlopen(path, RTLD_NOW | RTLD_LOCAL);
        if (!ref) { fprintf(stderr, "recorder: %s\n", dlerror()); exit(3); }
        out = fopen(getenv("TAPP_KAT_OUT"), "a");
        if (!out) { perror("recorder"); exit(3); }
    }
    void* f = dlsym(ref, name);
    if (!f) { fprintf(stderr, "recorder: missing %s\n", name); exit(3); }
    return f;
}

typedef int (*fn_TAPP_create_tensor_info)(H*, int, int, const int64_t*, const int64_t*);
typedef int (*fn_TAPP_destroy_tensor_info)(H);
typedef int (*fn_TAPP_create_tensor_product)(H*, H, int, H, const int64_t*, int, H, const int64_t*,
                                              int, H, const int64_t*, int, H, const int64_t*, int);
typedef int (*fn_TAPP_destroy_tensor_product)(H);
typedef int (*fn_TAPP_execute_product)(H, H, H*, const void*, const void*, const void*,
                                        const void*, const void*, void*);
typedef int (*fn_TAPP_create_executor)(H*);
typedef int (*fn_TAPP_destroy_executor)(H);
typedef int (*fn_TAPP_create_handle)(H*);
typedef int (*fn_TAPP_destroy_handle)(H);
typedef bool (*fn_TAPP_check_success)(int);
typedef size_t (*fn_TAPP_explain_error)(int, size_t, char*);

typedef struct { H ref; int type, nmode; int64_t ext[16], str[16]; } info;
typedef struct { H ref; int op[4]; info* t[4]; int64_t idx[4][16]; int prec; } plan;

int TAPP_create_tensor_info(H* h, int type, int nmode, const int64_t* e, const int64_t* s)
{
    info* i = calloc(1, sizeof *i);
    int rc = REF(TAPP_create_tensor_info)(&i->ref, type, nmode, e, s);
    i->type = type; i->nmode = nmode;
    for (int k = 0; k < nmode && k < 16; k++) { i->ext[k] = e[k]; i->str[k] = s[k]; }
    *h = (H)i;
    return rc;
}
int TAPP_destroy_tensor_info(H h) { info* i = (info*)h; int rc = REF(TAPP_destroy_tensor_info)(i->ref); free(i); return rc; }
int TAPP_create_executor(H* e) { return REF(TAPP_create_executor)(e); }
int TAPP_destroy_executor(H e) { return REF(TAPP_destroy_executor)(e); }
int TAPP_create_handle(H* h) { return REF(TAPP_create_handle)(h); }
int TAPP_destroy_handle(H h) { return REF(TAPP_destroy_handle)(h); }
bool TAPP_check_success(int e) { return REF(TAPP_check_success)(e); }
size_t TAPP_explain_error(int e, size_t n, char* m) { return REF(TAPP_explain_error)(e, n, m); }

int TAPP_create_tensor_product(H* p, H handle, int opA, H A, const int64_t* iA, int opB, H B,
                               const int64_t* iB, int opC, H C, const int64_t* iC, int opD, H D,
                               const int64_t* iD, int prec)
{
    plan* q = calloc(1, sizeof *q);
    info* t[4] = {(info*)A, (info*)B, (info*)C, (info*)D};
    const int64_t* ix[4] = {iA, iB, iC, iD};
    int ops[4] = {opA, opB, opC, opD};
    for (int k = 0; k < 4; k++) {
        q->t[k] = t[k]; q->op[k] = ops[k];
        for (int m = 0; m < t[k]->nmode && m < 16; m++) q->idx[k][m] = ix[k][m];
    }
    q->prec = prec;
    int rc = REF(TAPP_create_tensor_product)(&q->ref, handle, opA, t[0]->ref, iA, opB, t[1]->ref, iB,
                                             opC, t[2]->ref, iC, opD, t[3]->ref, iD, prec);
    *p = (H)q;
    return rc;
}
int TAPP_destroy_tensor_product(H p) { plan* q = (plan*)p; int rc = REF(TAPP_destroy_tensor_product)(q->ref); free(q); return rc; }

static int width(int type) { return type == 0 ? 1 : type == 1 ? 1 : 2; }
static bool is_double(int type) { return type == 1 || type == 3; }

static void span(const info* t, int64_t* lo, int64_t* hi)
{
    *lo = 0; *hi = 0;
    for (int k = 0; k < t->nmode; k++) {
        if (t->ext[k] == 0) { *lo = 0; *hi = -1; return; }
        int64_t d = (t->ext[k] - 1) * t->str[k];
        if (d < 0) *lo += d; else *hi += d;
    }
}

static void dump_values(const void* base, int type, int64_t lo, int64_t hi)
{
    fprintf(out, "[");
    for (int64_t o = lo; o <= hi; o++) {
        for (int w = 0; w < width(type); w++) {
            double v = is_double(type) ? ((const double*)base)[o * width(type) + w]
                                       : (double)((const float*)base)[o * width(type) + w];
            fprintf(out, "%s%.17g", (o == lo && w == 0) ? "" : ",", v);
        }
    }
    fprintf(out, "]");
}

static void dump_tensor(const char* name, const plan* q, int k, const void* data)
{
    const info* t = q->t[k];
    int64_t lo, hi;
    span(t, &lo, &hi);
    fprintf(out, "\"%s\":{\"type\":%d,\"nmode\":%d,\"op\":%d,\"extents\":[", name, t->type, t->nmode, q->op[k]);
    for (int m = 0; m < t->nmode; m++) fprintf(out, "%s%lld", m ? "," : "", (long long)t->ext[m]);
    fprintf(out, "],\"strides\":[");
    for (int m = 0; m < t->nmode; m++) fprintf(out, "%s%lld", m ? "," : "", (long long)t->str[m]);
    fprintf(out, "],\"idx\":[");
    for (int m = 0; m < t->nmode; m++) fprintf(out, "%s%lld", m ? "," : "", (long long)q->idx[k][m]);
    fprintf(out, "],\"offset\":%lld,\"data\":", (long long)lo);
    if (data) dump_values(data, t->type, lo, hi); else fprintf(out, "null");
    fprintf(out, "}");
}

int TAPP_execute_product(H p, H exec, H* status, const void* alpha, const void* A, const void* B,
                         const void* beta, const void* C, void* D)
{
    plan* q = (plan*)p;
    int tD = q->t[3]->type;
    fprintf(out, "{\"prec\":%d,\"alpha\":", q->prec);
    dump_values(alpha, tD, 0, 0);
    fprintf(out, ",\"beta\":");
    dump_values(beta, tD, 0, 0);
    fprintf(out, ",");
    dump_tensor("A", q, 0, A); fprintf(out, ",");
    dump_tensor("B", q, 1, B); fprintf(out, ",");
    dump_tensor("C", q, 2, C); fprintf(out, ",");
    dump_tensor("D_in", q, 3, D); fprintf(out, ",");
    int rc = REF(TAPP_execute_product)(q->ref, exec, status, alpha, A, B, beta, C, D);
    fprintf(out, "\"rc\":%d,", rc);
    dump_tensor("D_out", q, 3, D);
    fprintf(out, "}\n");
    fflush(out);
    return rc;
}
