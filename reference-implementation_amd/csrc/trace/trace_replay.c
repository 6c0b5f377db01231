/*
 * Open-loop load generator for the serving loop (used by bench.py and the
 * tests, not part of the TAPP API).  Request i is submitted at
 * t0 + arrival_us[i] -- the submitting thread spins until then, so arrivals
 * follow the given (e.g. Poisson) schedule however the server keeps up --
 * and every status is waited on and released at the end.  The same trace
 * can be run one TAPP_execute_product call at a time instead, back to back
 * on one executor, as the baseline the serving loop is compared with.
 */
#define _POSIX_C_SOURCE 200809L
#include <stdlib.h>
#include <time.h>

#include "tapp_amd.h"

static double now_us(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec * 1e6 + (double)t.tv_nsec * 1e-3;
}

/* Returns the number of submits that failed; rc_out[i] receives each
 * request's code (submit or completion), late_us the total lateness of
 * submits behind their schedule. */
__attribute__((visibility("default"))) int tapp_trace_replay(TAPP_AMD_server server, int n,
                                                             const TAPP_tensor_product* plans,
                                                             const void* const* alpha, const void* const* A,
                                                             const void* const* B, const void* const* beta,
                                                             const void* const* C, void* const* D,
                                                             const double* arrival_us, int* rc_out,
                                                             double* late_us)
{
    TAPP_status* st = calloc((size_t)(n > 0 ? n : 1), sizeof(TAPP_status));
    if (!st) return -1;
    int failed = 0;
    double late = 0;
    const double t0 = now_us();
    for (int i = 0; i < n; i++) {
        const double due = t0 + arrival_us[i];
        double t = now_us();
        while (t < due) t = now_us();
        late += t - due;
        rc_out[i] = TAPP_AMD_server_submit(server, plans[i], alpha[i], A[i], B[i], beta[i], C ? C[i] : NULL, D[i],
                                           &st[i]);
        if (rc_out[i]) { failed++; st[i] = 0; }
    }
    for (int i = 0; i < n; i++) {
        if (!st[i]) continue;
        TAPP_error r = 0;
        TAPP_AMD_status_wait(server, st[i], &r);
        rc_out[i] = r;
        TAPP_destroy_status(st[i]);
    }
    free(st);
    if (late_us) *late_us = late;
    return failed;
}

/* Request i as one TAPP_execute_product call on `exec`, one after another
 * (each returns when its D is written); arrivals are not waited for, so
 * this is the most a one-call-at-a-time client gets through.  Returns the
 * number of failed calls; *elapsed_us receives the wall time. */
__attribute__((visibility("default"))) int tapp_trace_sequential(TAPP_executor exec, int n,
                                                                 const TAPP_tensor_product* plans,
                                                                 const void* const* alpha, const void* const* A,
                                                                 const void* const* B, const void* const* beta,
                                                                 const void* const* C, void* const* D, int* rc_out,
                                                                 double* elapsed_us)
{
    int failed = 0;
    const double t0 = now_us();
    for (int i = 0; i < n; i++) {
        rc_out[i] = TAPP_execute_product(plans[i], exec, NULL, alpha[i], A[i], B[i], beta[i], C ? C[i] : NULL, D[i]);
        if (rc_out[i]) failed++;
    }
    if (elapsed_us) *elapsed_us = now_us() - t0;
    return failed;
}
