/*
 * Host-layer internals (C11).  Object layouts and the plan stage.
 */
#ifndef TAPP_AMD_INTERNAL_H_
#define TAPP_AMD_INTERNAL_H_

#include <pthread.h>
#include <stdbool.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#include "tapp.h"
#include "tapp_amd.h"
#include "../tk.h"

/* --- objects ------------------------------------------------------------ */

/* Layout descriptor (reference: struct tensor_info, ref_impl.h:13-19). */
struct tapp_info {
    TAPP_datatype type;
    int nmode;
    int64_t* extents;
    int64_t* strides;
};

struct tapp_plan_cache;
struct tapp_batch_cache;

/* Product plan (reference: struct plan, ref_impl.h:21-37).  Layouts are held
 * by handle, so later TAPP_set_* calls on an info change the plan, as in the
 * reference (product.c:86-107). */
struct tapp_plan {
    TAPP_handle handle;
    TAPP_element_op op[4];
    TAPP_tensor_info info[4];
    int64_t* idx[4];
    int nidx[4];               /* label count copied at creation */
    TAPP_prectype prec;
    struct tapp_plan_cache* cache;
    struct tapp_plan_cache* hcache;   /* host-only tables of the allocation-free path */
    struct tapp_batch_cache* batch;   /* the last TAPP_execute_batched_product call, prepared */
    /* recursive; held while a call builds or reads the cache and enqueues
     * work that reads it, so calls on one plan from several threads are safe */
    pthread_mutex_t mu;
};

#define TAPP_AMD_EXEC_MAGIC 0x54415050u /* "TAPP" */

/* Executor = device context.  The first field is the int kind tag the
 * reference's executor points at (executor.c:9-18), so the reference's
 * validity rule (product.c:878-885) reads the same value. */
struct tapp_exec {
    int kind;                  /* 1 */
    uint32_t magic;
    int device;
    int cus;                   /* compute units of the device */
    hipStream_t stream;
    bool own_stream;
    bool have_stream;          /* a stream is bound (the null stream, 0, is a valid binding) */
    /* recursive; held by every API call that uses the context (stream,
     * workspace, events, last-call fields), so threads sharing an executor
     * -- or the process-wide default context of bare int executors -- take
     * turns instead of racing on the workspace */
    pthread_mutex_t mu;
    bool mu_init;
    hipEvent_t ev0, ev1;
    /* device workspace: grown on demand, reused across calls */
    void* ws;
    size_t ws_size;
    /* attributes */
    int split_bf16;
    int force_kernel;
    int tiles;                 /* TAPP_AMD_ATTR_TILES: 0 by shape, 1 large, 2 small, 3 tall */
    int autotune;              /* 0 rules, 1 measure unknown shape classes, 2 cache only */
    int last_reused;           /* the last batched call reused its plan's prepared batch */
    int timing;
    int last_kernel;
    char last_kernel_name[96];  /* kernel template of the last execute */
    double last_device_us;
};

/* Handle: the device current at creation and the table of resident tensors
 * (resident.c). */
struct tapp_handle_obj {
    uint32_t magic;
    int device;
    void* res;                 /* struct resident[cap] */
    int nres, cap;
};

/* --- plan stage ---------------------------------------------------------- */

#define TK_MAXG 64   /* max modes per group after compression */

typedef struct {
    int n;
    int64_t ext[TK_MAXG];
    int64_t s[4][TK_MAXG];     /* strides of A, B, C, D (0 where absent) */
} tk_group;

/* Everything TAPP_execute_product derives from the plan before touching
 * data. */
typedef struct {
    int rc;                    /* validation result, 0 = ok */
    int type[4];               /* storage types A, B, C, D */
    int ctype;                 /* compute type TK_* */
    int mode;                  /* TK_MODE_* */
    int op_A, op_B, op_C, op_D;
    bool conjD_via_f32;
    bool round_C;
    /* sums of A's and B's unary labels held in a 16-bit format, rounded at
     * every addition (0, TK_MODE_F16 or TK_MODE_BF16): the TAPP_F16F16_* and
     * BF16BF16 precisions, and DEFAULT_PREC with 16-bit storage */
    int sum_round[2];
    bool beta_zero;
    double alpha[2], beta[2];
    tk_group L, M, N, K, UA, UB;
    int64_t nL, nM, nN, nK, nUA, nUB;
    int64_t pin[4];            /* element offset added for pinned zero-stride D modes */
    int64_t lo[4], hi[4];      /* element span of A, B, C, D; hi < lo when empty */
    bool empty;                /* no D element is written */
} tk_problem;

int tk_plan_problem(const struct tapp_plan* p, const struct tapp_exec* ex, const void* alpha,
                    const void* beta, const void* C, tk_problem* out);

/* storage helpers */
size_t tk_type_size(int type);
bool tk_type_complex(int type);

/* operand residency (resident.c) */
bool tk_is_device_ptr(const void* p);
/* Device address to read operand p at, whose accessed bytes are
 * [p + lo_bytes, p + hi_bytes): p if it is device memory, the resident copy if
 * the handle holds that range, NULL for host memory to be staged. */
const void* tk_resolve_operand(TAPP_handle handle, const void* p, int64_t lo_bytes, int64_t hi_bytes);
void tk_handle_register(struct tapp_handle_obj* h);
void tk_handle_unregister(struct tapp_handle_obj* h);

/* cached autotuner (tune.c) */
uint64_t tk_tune_key(const int64_t* fields, int n);
int64_t tk_tune_bucket(int64_t x);
bool tk_tune_get(int device, uint64_t key, int* value);
void tk_tune_put(int device, uint64_t key, int value, const char* desc);
double tk_tune_time(struct tapp_exec* ex, int (*launch)(void* ctx), void* ctx, int reps);

/* locks (executor.c): recursive mutex set-up; executor locks are no-ops on
 * contexts without one (the stand-in of an invalid bare int executor) */
void tk_mutex_init_recursive(pthread_mutex_t* m);
void tk_exec_lock(struct tapp_exec* e);
void tk_exec_unlock(struct tapp_exec* e);

/* registry of live executors / handles */
struct tapp_exec* tk_exec_lookup(TAPP_executor e);
struct tapp_exec* tk_default_exec(void);
void tk_exec_register(struct tapp_exec* e);
void tk_exec_unregister(struct tapp_exec* e);

/* execution */
int tk_execute(struct tapp_plan* p, struct tapp_exec* ex, const void* alpha, const void* A,
               const void* B, const void* beta, const void* C, void* D);
int tk_execute_on(struct tapp_plan* p, struct tapp_exec* ex, const void* alpha, const void* A,
                  const void* B, const void* beta, const void* C, void* D, bool sync);
/* Synchronous execute whose device memory (tables, pre-pass buffers,
 * split-K partials) is the device scratch array (TK_SCRATCH_BYTES); device or
 * resident operands only.  Calls on one device take turns. */
int tk_execute_noalloc(struct tapp_plan* p, struct tapp_exec* ex, const void* alpha, const void* A,
                       const void* B, const void* beta, const void* C, void* D);
int tk_kernel_for(const tk_problem* P, int force);
/* self-checking execution (executor kind 12, check.c) */
int tk_execute_checked(struct tapp_plan* p, struct tapp_exec* ex, const void* alpha, const void* A,
                       const void* B, const void* beta, const void* C, void* D);
void tk_plan_cache_free(struct tapp_plan_cache* c);
/* Device base and layout key of the plan's current offset tables (NULL
 * before the first build). */
const void* tk_plan_tables_id(const struct tapp_plan* plan, uint64_t* key);
void tk_batch_cache_free(struct tapp_batch_cache* b);
/* Device memory a call of `p` would allocate through `ex` (asynchronous, as
 * the serving loop runs it): *work = its workspace, *tables = the plan's
 * offset tables when not yet on ex's device; *devops = every operand is on
 * the device or resident. */
int tk_call_bytes(struct tapp_plan* p, const struct tapp_exec* ex, const tk_problem* P, const void* A,
                  const void* B, const void* C, void* D, size_t* work, size_t* tables, bool* devops);
/* The plan's tables for the current layout are not yet on ex's device. */
bool tk_plan_needs_tables(struct tapp_plan* p, const struct tapp_exec* ex);
/* Grows the executor's workspace to exactly `bytes` (no headroom) when it is
 * smaller; frees it. */
int tk_exec_reserve_ws(struct tapp_exec* ex, size_t bytes);
void tk_exec_release_ws(struct tapp_exec* ex);
/* The problem-level conditions of the persistent launch (compute type,
 * precision form, unary labels, storage types). */
bool tk_group_may_run(const tk_problem* P);
int tk_problem_tables(struct tapp_plan* plan, const tk_problem* P, int device, const int64_t** mA,
                      const int64_t** kA, const int64_t** lA, const int64_t** nB, const int64_t** kB,
                      const int64_t** lB, const int64_t** mD, const int64_t** nD, const int64_t** lD,
                      const int64_t** mC, const int64_t** nC, const int64_t** lC);
void tk_problem_host_k(struct tapp_plan* plan, const int64_t** kA, const int64_t** kB);
void tk_problem_host_mn(struct tapp_plan* plan, const int64_t** mA, const int64_t** nB);
/* force_cls: -1 by rule (thin tiles when the smaller side is <= 16), or
 * TK_GROUP_THIN / TK_GROUP_MFMA (thin only where the rule allows it) */
bool tk_group_describe(struct tapp_plan* plan, struct tapp_exec* ex, const tk_problem* P, const void* A,
                       const void* B, const void* C, void* D, int force_cls, tk_prob* out, int* rc);

#endif /* TAPP_AMD_INTERNAL_H_ */
