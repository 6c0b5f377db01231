/*
 * Tensor products: plan objects and execution on the GPU.
 *
 * TAPP_create_tensor_product / TAPP_destroy_tensor_product keep the
 * reference's semantics (product.c:66-128): no validation, layouts held by
 * handle, label arrays copied with the layouts' current mode counts.
 *
 * TAPP_execute_product (reference product.c:130-465):
 *   plan stage (plan.c) -> offset tables (cached per plan and layout) ->
 *   operands on the device (device pointers used in place, host pointers
 *   staged over their address span) -> optional pre-pass (unary reduction,
 *   storage conversion) -> GETT kernel -> optional scatter to D's storage
 *   type -> D staged back if it is host memory.  The call returns when D is
 *   written, as the reference's does.
 *
 * Threads: the executor's context (stream, workspace) is locked for the
 * call and the plan while its tables are built and read, so concurrent calls
 * are safe -- on one executor, including the shared default context of bare
 * int executors, they take turns.
 */
#include <stdlib.h>
#include <string.h>

#include "internal.h"

TAPP_error TAPP_create_tensor_product(TAPP_tensor_product* plan, TAPP_handle handle, TAPP_element_op op_A,
                                      TAPP_tensor_info A, const int64_t* idx_A, TAPP_element_op op_B,
                                      TAPP_tensor_info B, const int64_t* idx_B, TAPP_element_op op_C,
                                      TAPP_tensor_info C, const int64_t* idx_C, TAPP_element_op op_D,
                                      TAPP_tensor_info D, const int64_t* idx_D, TAPP_prectype prec)
{
    struct tapp_plan* p = calloc(1, sizeof *p);
    if (!p) return TAPP_AMD_ERR_NOMEM;
    p->handle = handle;
    TAPP_tensor_info infos[4] = {A, B, C, D};
    const int64_t* idx[4] = {idx_A, idx_B, idx_C, idx_D};
    TAPP_element_op ops[4] = {op_A, op_B, op_C, op_D};
    for (int x = 0; x < 4; x++) {
        p->info[x] = infos[x];
        p->op[x] = ops[x];
        int n = infos[x] ? ((struct tapp_info*)infos[x])->nmode : 0;
        p->nidx[x] = n;
        p->idx[x] = malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
        if (!p->idx[x]) {
            for (int y = 0; y < x; y++) free(p->idx[y]);
            free(p);
            return TAPP_AMD_ERR_NOMEM;
        }
        if (n > 0) memcpy(p->idx[x], idx[x], sizeof(int64_t) * (size_t)n);
    }
    p->prec = prec;
    tk_mutex_init_recursive(&p->mu);
    *plan = (TAPP_tensor_product)p;
    return 0;
}

TAPP_error TAPP_destroy_tensor_product(TAPP_tensor_product plan)
{
    struct tapp_plan* p = (struct tapp_plan*)plan;
    if (!p) return 0;
    for (int x = 0; x < 4; x++) free(p->idx[x]);
    tk_plan_cache_free(p->cache);
    tk_plan_cache_free(p->hcache);
    tk_batch_cache_free(p->batch);
    pthread_mutex_destroy(&p->mu);
    free(p);
    return 0;
}

/* Resolves the executor handle the way the reference's validity rule sees
 * it: NULL -> no executor (code 33 at the end of validation); one of this
 * library's executors -> itself; any other pointer -> its int is the kind,
 * and a valid kind runs on the process-wide default context. */
static struct tapp_exec* resolve_exec(TAPP_executor exec, struct tapp_exec* shadow)
{
    if (!exec) return NULL;
    struct tapp_exec* e = tk_exec_lookup(exec);
    if (e) return e;
    int kind = *(const int*)exec;
    if (!(kind == 1 || kind == 2 || kind == 12)) {
        memset(shadow, 0, sizeof *shadow);
        shadow->kind = kind;
        return shadow;
    }
    return tk_default_exec();
}

TAPP_error TAPP_execute_product(TAPP_tensor_product plan, TAPP_executor exec, TAPP_status* status,
                                const void* alpha, const void* A, const void* B, const void* beta,
                                const void* C, void* D)
{
    (void)status;
    struct tapp_plan* p = (struct tapp_plan*)plan;
    if (!p) return TAPP_AMD_ERR_INVALID;
    struct tapp_exec shadow;
    struct tapp_exec* ex = resolve_exec(exec, &shadow);
    /* kind 12: every product computed twice and compared (check.c); the
     * kind is the int the executor points at, for this library's executors
     * their first field */
    int rc;
    tk_exec_lock(ex);
    if (ex && *(const int*)exec == 12) rc = tk_execute_checked(p, ex, alpha, A, B, beta, C, D);
    else rc = tk_execute(p, ex, alpha, A, B, beta, C, D);
    tk_exec_unlock(ex);
    return rc;
}

TAPP_error TAPP_AMD_execute_product_noalloc(TAPP_tensor_product plan, TAPP_executor exec, TAPP_status* status,
                                            const void* alpha, const void* A, const void* B, const void* beta,
                                            const void* C, void* D)
{
    (void)status;
    struct tapp_plan* p = (struct tapp_plan*)plan;
    if (!p) return TAPP_AMD_ERR_INVALID;
    /* validation first, with the reference's codes; a caller's bare int
     * executor is only read (resolving it would create the default context) */
    struct tapp_exec shadow;
    memset(&shadow, 0, sizeof shadow);
    struct tapp_exec* ex = exec ? tk_exec_lookup(exec) : NULL;
    if (exec && !ex) shadow.kind = *(const int*)exec;
    tk_problem P;
    int rc = tk_plan_problem(p, ex ? ex : (exec ? &shadow : NULL), alpha, beta, C, &P);
    if (rc) return rc;
    if (!ex) return TAPP_AMD_ERR_INVALID;
    if (ex->kind == 12) return TAPP_AMD_ERR_UNSUPPORTED;
    tk_exec_lock(ex);
    rc = tk_execute_noalloc(p, ex, alpha, A, B, beta, C, D);
    tk_exec_unlock(ex);
    return rc;
}

TAPP_error TAPP_AMD_plan_info(TAPP_tensor_product plan, TAPP_executor exec, const void* beta, const void* C,
                              int64_t sizes[6], int* compute_type)
{
    struct tapp_plan* p = (struct tapp_plan*)plan;
    if (!p) return TAPP_AMD_ERR_INVALID;
    struct tapp_exec shadow;
    struct tapp_exec* ex = resolve_exec(exec, &shadow);
    tk_problem P;
    float one = 1.0f;
    (void)one;
    int rc = tk_plan_problem(p, ex, NULL, beta, C, &P);
    if (rc) return rc;
    if (sizes) {
        sizes[0] = P.nL; sizes[1] = P.nM; sizes[2] = P.nN; sizes[3] = P.nK; sizes[4] = P.nUA; sizes[5] = P.nUB;
    }
    if (compute_type) *compute_type = P.ctype;
    return 0;
}
