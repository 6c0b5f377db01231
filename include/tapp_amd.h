/*
 * MI355X extensions to the TAPP C API.  Everything here is additive: a
 * program written against the reference headers (api/include/tapp/ and tapp.h)
 * compiles and runs unchanged against this library.
 */
#ifndef TAPP_AMD_H_
#define TAPP_AMD_H_

#include <stdint.h>

#include "tapp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes beyond the reference's (which are 0-15, 33 and 137,
 * error.c:15-76 and product.c:197-222, 463). */
enum {
    TAPP_AMD_ERR_UNSUPPORTED = 101, /* datatype / precision combination not supported */
    TAPP_AMD_ERR_DEVICE = 102,      /* HIP runtime error */
    TAPP_AMD_ERR_NOMEM = 103,       /* device or host memory exhausted */
    TAPP_AMD_ERR_INVALID = 104,     /* invalid handle, NULL argument, label count mismatch */
    TAPP_AMD_ERR_KEY = 105,         /* unknown attribute key */
    TAPP_AMD_ERR_BUDGET = 106,      /* request larger than the serving memory budget */
    TAPP_AMD_ERR_BUSY = 107,        /* serving queue full */
    TAPP_AMD_ERR_SCRATCH = 108      /* product needs more than the allocation-free entry point's scratch */
};

/* Threads and ordering.
 *   - TAPP_execute_product may be called from several threads at once, with
 *     the same or different plans and executors.  An executor is one device
 *     context (stream + workspace): calls sharing it -- including every call
 *     made with a bare int executor, which share one process-wide context --
 *     run one at a time.  Use one executor per thread for concurrency.
 *   - Each executor launches on its own non-blocking stream, which does not
 *     wait for work queued on other streams (the null stream included).
 *     Device operands must be complete when the call is made: synchronize
 *     the producing stream first, or bind that stream with
 *     TAPP_AMD_ATTR_STREAM (the null stream, 0, is a valid binding), after
 *     which the call's kernels run in order after the work queued on it.
 *     The call still returns only when D is written. */

/* Attribute keys for TAPP_attr_set / TAPP_attr_get / TAPP_attr_clear on an
 * executor (pass the TAPP_executor as the TAPP_attr).  Values are passed by
 * pointer to an int unless stated. */
enum {
    /* FP32 products with TAPP_DEFAULT_PREC or TAPP_F32F32_ACCUM_F32 run on
     * 16-bit MFMA with FP32 accumulation:
     * 1 = split-BF16 (x = hi + lo in bfloat16, three products: FP32-level
     *     accuracy at twice FP32's speed);
     * 2 = three-piece split-BF16 (x = h + m + l, six products, FP32 sums of
     *     128 k added with compensation: far below FP32's own error, at
     *     about FP32 speed);
     * 3 = split-F16 (x = hi + lo in binary16, three products, FP32 sums of
     *     128 k added into a total: below FP32's own error at about twice
     *     its speed).  binary16's range is checked: the kernel reports
     *     max |A| and max |B|, and an operand outside [2^-3, 2^15) is rerun
     *     scaled by a power of two, or on the FP32 kernel for NaN, inf and
     *     magnitudes below 2^-100.  A call that cannot rerun (C overlapping
     *     D with beta != 0, requests of the serving loop) runs value 2.
     * Products small enough for the one-lane-per-output kernel, and complex
     * products, run in FP32 with value 3.  Opt-in; default 0. */
    TAPP_AMD_ATTR_SPLIT_BF16 = 0x1001,
    /* Force a kernel family: 0 auto, 1 one-lane-per-output FMA, 2 MFMA,
     * 3 split-K, 5 small-tile MFMA (FP32; others run 2).  Default 0. */
    TAPP_AMD_ATTR_KERNEL = 0x1002,
    /* Read-only: kernel family used by the last execute (int). */
    TAPP_AMD_ATTR_LAST_KERNEL = 0x1003,
    /* Read-only: device time in microseconds of the last execute's kernels,
     * measured with HIP events on the executor's stream (double). */
    TAPP_AMD_ATTR_LAST_DEVICE_US = 0x1004,
    /* The HIP stream (hipStream_t) the executor launches on; set to bind an
     * external stream (0 = the null stream), get to read it.  The executor
     * waits for its own stream before switching.  Value: pointer to a
     * void*. */
    TAPP_AMD_ATTR_STREAM = 0x1005,
    /* 1: time every execute with HIP events (needed for LAST_DEVICE_US). */
    TAPP_AMD_ATTR_TIMING = 0x1006,
    /* int: 0 = built-in kernel rules (default); 1 = autotune: measure the
     * kernel configurations of a shape class the first time it runs (single
     * products: kernel family, MFMA tile size, split-K; product groups: thin
     * FMA or MFMA tiles) and keep the fastest in an on-disk cache
     * ($TAPP_AMD_TUNE_CACHE, else ~/.cache/tapp_amd/tune-<arch>.txt);
     * 2 = use cached choices, never measure.  Measuring runs the product
     * several times, so it is skipped when C and D overlap with beta != 0
     * and for requests of the serving loop. */
    TAPP_AMD_ATTR_AUTOTUNE = 0x1007,
    /* int: 1 = self-checking executor, the reference's executor kind 12
     * (product.c:229-249, 424-448): TAPP_execute_product computes each
     * product twice with independent kernel families (the check into a
     * device copy of D), compares the results element by element with the
     * reference's rule (tblis_bind.cpp:577-640: |(a - b) / max(a, b)| >
     * 5e-5; binary64 complex also needs an absolute difference > 1e-12),
     * prints up to 16 differing positions to stdout and returns 137 when any
     * differ.  Setting it writes 12 (or 1) into the executor's kind, so
     * writing 12 into the int the executor points at, as a reference caller
     * does, has the same effect.  Default 0. */
    TAPP_AMD_ATTR_SELF_CHECK = 0x1008,
    /* Read-only: 1 when the last TAPP_execute_batched_product on this
     * executor reused the plan's prepared batch (descriptors patched, graph
     * replayed), 0 when it built the batch (int). */
    TAPP_AMD_ATTR_LAST_REUSED = 0x1009,
    /* Read-only: name of the kernel template the last execute launched
     * (const char*, owned by the executor, valid until its next call). */
    TAPP_AMD_ATTR_LAST_KERNEL_NAME = 0x100A,
    /* int: MFMA tile size of single products.  0 = by shape (default);
     * 1 = large: FP32 split forms run 256 x 128 tiles (one workgroup per CU)
     * where both operands load along a contiguous direction; F64 and complex
     * run their large tiles; 2 = small: 128 x 128 split-form tiles, the
     * small F64 / complex tiles; 3 = tall: F64 and complex run tiles of the
     * small tiles' columns and twice their rows (FP32 as 0).  Every size
     * gives the same bits. */
    TAPP_AMD_ATTR_TILES = 0x100B
};

/* Kernel family identifiers reported by TAPP_AMD_ATTR_LAST_KERNEL. */
enum {
    TAPP_AMD_KERNEL_NONE = 0,
    TAPP_AMD_KERNEL_SIMPLE = 1,
    TAPP_AMD_KERNEL_MFMA = 2,
    TAPP_AMD_KERNEL_SPLITK = 3,
    TAPP_AMD_KERNEL_GROUP = 4,      /* persistent launch of a product group */
    TAPP_AMD_KERNEL_MFMA_SMALL = 5  /* FP32 16 x 16 MFMA tiles, K split over a workgroup's
                                       waves: few outputs, moderate K, one launch */
};

/* --- product groups ---------------------------------------------------------
 * Many independent products, each with its own plan, operands and scalars,
 * run as one persistent launch (a work queue of tiles over all products,
 * captured in a hipGraph).  The reference declares only the one-plan
 * TAPP_execute_batched_product (product.h:66-75) and defines nothing.
 *
 * create: validates every product exactly as TAPP_execute_product would and
 *   returns the first failing code; reads alpha[i]/beta[i] now (the values
 *   are copied: the caller's scalars may be freed after create); keeps the
 *   operand pointers.  Layouts and data pointers must not change while the
 *   group exists (the group replays the same addresses), and the plans and
 *   their tensor infos must outlive the group: its launch reads the plans'
 *   device offset tables, and products run one by one execute their plan.
 * execute: runs every product; returns when all D are written.
 * Products outside the persistent kernel's scope (non-FP32 compute, unary
 * labels, storage conversion, host pointers) run one by one inside execute. */
typedef intptr_t TAPP_AMD_product_group;

TAPP_EXPORT TAPP_error TAPP_AMD_create_product_group(TAPP_AMD_product_group* group, TAPP_executor exec,
                                                     int count, const TAPP_tensor_product* plans,
                                                     const void* const* alpha, const void* const* A,
                                                     const void* const* B, const void* const* beta,
                                                     const void* const* C, void* const* D);
TAPP_EXPORT TAPP_error TAPP_AMD_execute_product_group(TAPP_AMD_product_group group);
TAPP_EXPORT TAPP_error TAPP_AMD_destroy_product_group(TAPP_AMD_product_group group);
/* info[0] products in the persistent launch, info[1] products run one by
 * one, info[2] tiles of the persistent launch. */
TAPP_EXPORT TAPP_error TAPP_AMD_group_info(TAPP_AMD_product_group group, int64_t info[3]);

/* --- serving -----------------------------------------------------------------
 * A request loop with continuous batching on one GPU.  A dispatcher thread
 * owns `slots` HIP streams; whenever one is idle it admits queued requests
 * (FIFO) while the device memory the server holds fits `hbm_budget_bytes`,
 * launches them as one batch (small FP32 products with D in device memory
 * in one persistent launch, their host operands that are not resident --
 * up to 4 MiB per request -- copied into the stream's workspace first; the
 * rest on the general path) and keeps admitting while batches run on other
 * streams.  Held memory: each stream's workspace (sized to the larger of
 * the persistent launch's staged host operands and the largest general-path
 * request of its batch -- staged host operands, pre-pass and conversion
 * buffers, split-K partial sums -- and kept between batches), and the
 * offset tables that batches in flight build for their plans (afterwards
 * the plan keeps them, outside the budget); the persistent launch's
 * descriptor buffers are not counted.
 * When a request does not fit, idle streams' workspaces are freed first;
 * if it still does not fit it waits (an admission stall).  Device operands
 * and resident tensors cost nothing.  A heavy request (above 2^31 flop, or
 * a workspace above 4 MiB)
 * is passed over while every other stream holds one, so that one stream
 * stays free for the small requests queued behind it.  The process needs a
 * hardware queue per stream it uses (HIP's GPU_MAX_HW_QUEUES, default 4,
 * read when the runtime starts): at least `slots` + 2, or slot streams
 * share in-order queues and small requests wait behind large kernels.
 * Each request returns a TAPP_status (status.h:9-17) that
 * TAPP_AMD_status_wait waits on and TAPP_destroy_status releases.  Operands
 * must stay valid, and plans unchanged, until the request completes; the
 * plans of queued requests must not be executed elsewhere concurrently.
 * Requests in flight together run in no particular order (a batch's
 * persistent launch first, its requests completing when it ends): submit a
 * request that reads another's D only after that one has completed. */
typedef intptr_t TAPP_AMD_server;

enum { TAPP_AMD_SERVER_PAUSED = 1 };   /* create paused: queue and admission only */

TAPP_EXPORT TAPP_error TAPP_AMD_create_server(TAPP_AMD_server* server, TAPP_executor exec, size_t hbm_budget_bytes,
                                              int slots, int max_batch, int flags);
/* Validates like TAPP_execute_product; TAPP_AMD_ERR_BUDGET when the request's
 * own workspace and tables exceed the budget.  alpha/beta are copied. */
TAPP_EXPORT TAPP_error TAPP_AMD_server_submit(TAPP_AMD_server server, TAPP_tensor_product plan, const void* alpha,
                                              const void* A, const void* B, const void* beta, const void* C,
                                              void* D, TAPP_status* status);
TAPP_EXPORT TAPP_error TAPP_AMD_status_wait(TAPP_AMD_server server, TAPP_status status, TAPP_error* result);
/* Removes a still-queued request; TAPP_AMD_ERR_BUSY once it was admitted. */
TAPP_EXPORT TAPP_error TAPP_AMD_server_cancel(TAPP_AMD_server server, TAPP_status status);
TAPP_EXPORT TAPP_error TAPP_AMD_server_set_paused(TAPP_AMD_server server, int paused);
/* Waits until nothing is queued (or the server is paused) and no batch runs. */
TAPP_EXPORT TAPP_error TAPP_AMD_server_drain(TAPP_AMD_server server);
/* stats: completed, p50 us, p99 us, mean us, requests that waited for memory, max in flight,
 * batches, products run in persistent launches, GFLOP/s, queued, device
 * bytes held now, most device bytes held, workspace reclaims.  Latency =
 * submit to observed completion. */
TAPP_EXPORT TAPP_error TAPP_AMD_server_stats(TAPP_AMD_server server, double stats[13], int reset);
TAPP_EXPORT TAPP_error TAPP_AMD_destroy_server(TAPP_AMD_server server);

/* --- resident tensors ----------------------------------------------------
 * Handle-scoped device copies of host buffers (handle.h:9-24 leaves the
 * handle's contents to the implementation).  After
 * TAPP_AMD_make_resident(handle, host, bytes), every execute of a plan created
 * with `handle` -- TAPP_execute_product, product groups, the serving loop --
 * that passes, as A, B or C, a host pointer whose accessed bytes lie inside
 * [host, host + bytes) reads the device copy instead of staging over PCIe.
 * The copy is taken now and refreshed only by TAPP_AMD_update_resident, which
 * first waits for every kernel already queued on the device (so none reads a
 * half-updated copy; requests queued but not yet launched read the new
 * contents); D is
 * always written through the caller's pointer.  Ranges of one handle must not
 * overlap.  Release (or destroy the handle) only after the requests and
 * product groups that read a copy have completed or been destroyed.
 * Errors: TAPP_AMD_ERR_INVALID (not this library's handle, overlapping or
 * unknown range, device pointer), TAPP_AMD_ERR_NOMEM, TAPP_AMD_ERR_DEVICE. */
TAPP_EXPORT TAPP_error TAPP_AMD_make_resident(TAPP_handle handle, const void* host, size_t bytes);
TAPP_EXPORT TAPP_error TAPP_AMD_update_resident(TAPP_handle handle, const void* host);
TAPP_EXPORT TAPP_error TAPP_AMD_release_resident(TAPP_handle handle, const void* host);
/* info[0] resident buffers, info[1] their bytes. */
TAPP_EXPORT TAPP_error TAPP_AMD_resident_info(TAPP_handle handle, int64_t info[2]);

/* --- allocation-free execution ---------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): TAPP_execute_product with the same arguments, codes and
 * results, whose device memory -- the plan's offset tables, pre-pass buffers,
 * split-K partial sums -- is one static device array of 64 MiB per GPU
 * (allocated when the library's code object loads, held until the process
 * ends).  It calls no hipMalloc / hipFree; host memory for a plan's table
 * contents is allocated on the first call with a given layout.
 *   - A, B, C and D must be device memory (or host ranges resident on the
 *     plan's handle): other host operands return TAPP_AMD_ERR_INVALID.
 *   - A product whose buffers exceed the scratch first drops split-K; if it
 *     still does not fit, TAPP_AMD_ERR_SCRATCH (108), and
 *     TAPP_AMD_last_error_message gives the size it needed.
 *   - exec must come from TAPP_create_executor (a bare int executor would
 *     need the default context created: TAPP_AMD_ERR_INVALID); the
 *     self-checking kind 12 is not available here (TAPP_AMD_ERR_UNSUPPORTED).
 *     The autotuner uses cached choices only.
 *   - Every call on one GPU uses that GPU's one array: calls from several
 *     threads take turns (a per-device lock held until the call's kernels
 *     have finished); the array keeps the last call's contents.
 * Synchronous: D is written on return. */
TAPP_EXPORT TAPP_error TAPP_AMD_execute_product_noalloc(TAPP_tensor_product plan, TAPP_executor exec,
                                                        TAPP_status* status, const void* alpha, const void* A,
                                                        const void* B, const void* beta, const void* C, void* D);

/* Text of the last device error of the calling thread (which HIP call
 * failed and why), with TAPP_explain_error's maxlen contract. */
TAPP_EXPORT size_t TAPP_AMD_last_error_message(char* message, size_t maxlen);

/* Version string of this library. */
TAPP_EXPORT const char* TAPP_AMD_version(void);

/* Plan-stage introspection for tests: after validation, classification and
 * stride folding, reports group sizes (L Hadamard, M, N, K contracted, UA
 * and UB unary) and the chosen compute type (0 F32, 1 F64, 2 C32, 3 C64).
 * Returns the validation code TAPP_execute_product would return. */
TAPP_EXPORT TAPP_error TAPP_AMD_plan_info(TAPP_tensor_product plan, TAPP_executor exec,
                                          const void* beta, const void* C,
                                          int64_t sizes[6], int* compute_type);

#ifdef __cplusplus
}
#endif

#endif /* TAPP_AMD_H_ */
