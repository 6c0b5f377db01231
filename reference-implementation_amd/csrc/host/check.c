/*
 * Self-checking execution: executor kind 12 (reference product.c:229-249,
 * 424-448, executor.c:13).
 *
 * The reference's kind 12 computes each product twice, with TBLIS into a
 * copy E_ of D and with its own loops into D, compares the two results
 * element by element (compare_tensors_, tblis_bind.cpp:577-640) and returns
 * 137 when they differ, after printing diagnostics.  Here both results come
 * from the GPU, from two independent kernel families: the one the product
 * would run (MFMA GETT, split-K or one lane per output) and, as the check,
 * the one-lane-per-output FMA kernel -- or the MFMA GETT when the product
 * itself would run the former.  The check runs first, into a device copy E
 * of D's address span, so that both runs read the caller's original C even
 * when C and D are the same array; the product then runs into D as usual and
 * a device kernel compares D with E over the span (elements of the span
 * that D's layout does not cover are equal copies).  The printout lists up
 * to TK_COMPARE_KEEP differing positions with both values.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

static void print_value(int type, const void* base, int64_t i)
{
    switch (type) {
    case TAPP_F32: printf("%.9g", (double)((const float*)base)[i]); break;
    case TAPP_F64: printf("%.17g", ((const double*)base)[i]); break;
    case TAPP_C32:
        printf("(%.9g,%.9g)", (double)((const float*)base)[2 * i], (double)((const float*)base)[2 * i + 1]);
        break;
    case TAPP_C64: printf("(%.17g,%.17g)", ((const double*)base)[2 * i], ((const double*)base)[2 * i + 1]); break;
    case TAPP_F16:
    case TAPP_BF16: printf("0x%04x", (unsigned)((const uint16_t*)base)[i]); break;
    default: printf("?"); break;
    }
}

int tk_execute_checked(struct tapp_plan* p, struct tapp_exec* ex, const void* alpha, const void* A,
                       const void* B, const void* beta, const void* C, void* D)
{
    tk_problem P;
    int rc = tk_plan_problem(p, ex, alpha, beta, C, &P);
    if (rc) return rc;
    if (P.empty) return 0;
    if (!ex->have_stream) return TAPP_AMD_ERR_DEVICE;
    if (!A || !B || !D || (!P.beta_zero && !C)) return TAPP_AMD_ERR_INVALID;
    if (hipSetDevice(ex->device) != hipSuccess) return TAPP_AMD_ERR_DEVICE;

    /* D's span [lo, hi] contains element (0, ..., 0) at offset 0 */
    const int type = P.type[3];
    const int64_t es = (int64_t)tk_type_size(type);
    const int64_t n = P.hi[3] - P.lo[3] + 1;
    if (n <= 0) return tk_execute(p, ex, alpha, A, B, beta, C, D);
    const size_t span = (size_t)(n * es);
    const size_t span_al = (span + 255) & ~(size_t)255;
    const bool d_dev = tk_is_device_ptr(D);

    /* E | copy of a host D | count | first[] */
    char* buf = NULL;
    const size_t bytes = span_al * (d_dev ? 1 : 2) + 256 + sizeof(int64_t) * TK_COMPARE_KEEP;
    if (hipMalloc((void**)&buf, bytes) != hipSuccess) return TAPP_AMD_ERR_NOMEM;
    char* E = buf;
    char* Dcopy = d_dev ? NULL : buf + span_al;
    unsigned long long* count = (unsigned long long*)(buf + span_al * (d_dev ? 1 : 2));
    int64_t* first = (int64_t*)((char*)count + 256);
    const char* Dspan = (const char*)D + P.lo[3] * es;

    /* the check: the other kernel family, no tuning or timing */
    const int kern = tk_kernel_for(&P, ex->force_kernel);
    struct tapp_exec chk = *ex;
    chk.force_kernel = kern == TK_KERNEL_SIMPLE && P.nK > 0 ? TK_KERNEL_MFMA : TK_KERNEL_SIMPLE;
    chk.autotune = 0;
    chk.timing = 0;

    if (hipMemcpyAsync(E, Dspan, span, hipMemcpyDefault, ex->stream) != hipSuccess) {
        rc = TAPP_AMD_ERR_DEVICE;
        goto out;
    }
    rc = tk_execute_on(p, &chk, alpha, A, B, beta, C, E - P.lo[3] * es, true);
    /* the check run may have grown the shared workspace */
    ex->ws = chk.ws;
    ex->ws_size = chk.ws_size;
    if (rc) goto out;

    /* the product itself, as the caller's executor would run it */
    rc = tk_execute_on(p, ex, alpha, A, B, beta, C, D, true);
    if (rc) goto out;

    const void* x = Dspan;
    if (!d_dev) {
        if (hipMemcpyAsync(Dcopy, Dspan, span, hipMemcpyHostToDevice, ex->stream) != hipSuccess) {
            rc = TAPP_AMD_ERR_DEVICE;
            goto out;
        }
        x = Dcopy;
    }
    tk_compare cmp = {type, x, E, n, count, first};
    unsigned long long bad = 0;
    int64_t where[TK_COMPARE_KEEP];
    if (tk_launch_compare(&cmp, ex->stream) != 0 ||
        hipMemcpyAsync(&bad, count, sizeof bad, hipMemcpyDeviceToHost, ex->stream) != hipSuccess ||
        hipMemcpyAsync(where, first, sizeof where, hipMemcpyDeviceToHost, ex->stream) != hipSuccess ||
        hipStreamSynchronize(ex->stream) != hipSuccess) {
        rc = TAPP_AMD_ERR_DEVICE;
        goto out;
    }
    if (bad) {
        const int shown = bad < TK_COMPARE_KEEP ? (int)bad : TK_COMPARE_KEEP;
        char* vals = (char*)malloc((size_t)(2 * shown) * (size_t)es);
        printf("TAPP self-check (executor kind 12): %llu of %lld elements of D differ from the check run\n", bad,
               (long long)n);
        if (vals) {
            for (int i = 0; i < shown; i++) {
                hipMemcpy(vals + (2 * i) * es, (const char*)x + where[i] * es, (size_t)es, hipMemcpyDefault);
                hipMemcpy(vals + (2 * i + 1) * es, E + where[i] * es, (size_t)es, hipMemcpyDeviceToHost);
            }
            for (int i = 0; i < shown; i++) {
                printf("  D offset %lld: ", (long long)(where[i] + P.lo[3]));
                print_value(type, vals, 2 * i);
                printf(" - ");
                print_value(type, vals, 2 * i + 1);
                printf("\n");
            }
            free(vals);
        }
        fflush(stdout);
        rc = 137;
    }
out:
    hipStreamSynchronize(ex->stream);
    hipFree(buf);
    return rc;
}
