/*
 * Serving loop: continuous batching of TAPP_execute_product requests on one
 * GPU.
 *
 *   client threads --submit--> FIFO queue --dispatcher thread--> stream slots
 *
 * The dispatcher owns `slots` device contexts (stream + workspace each).
 * Whenever a slot is idle it admits requests from the head of the queue while
 * the device memory they need fits the HBM budget, launches them as one batch on
 * that slot's stream and goes on: requests arriving while batches run are
 * admitted as soon as a slot and budget free up, and batches on different
 * slots overlap on the device.  Inside a batch, FP32 products that fit the
 * persistent kernel (small and thin products, D in device memory) run as
 * one persistent group launch -- a work queue over all their tiles, mixed
 * shape classes in one kernel.  Their host operands that are not resident
 * (up to heavy_bytes per request) are first copied into the batch's
 * workspace on the slot's stream.  Every other request (large GETTs,
 * complex, FP64, split-BF16, unary labels, host D) runs through the
 * TAPP_execute_product path on the same stream without waiting.  The
 * persistent launch goes first and has its own HIP event, so its (small)
 * requests complete when it ends, not behind the batch's large products;
 * a second event marks the end of the batch.
 *
 * Heavy requests -- above the persistent launch's flop limit, or with a
 * workspace above heavy_bytes -- hold a stream for a long time.  While
 * every other slot holds one, admission passes over them (they keep their
 * place in the queue) and takes the small requests behind them, so small
 * requests never wait for a stream behind heavy ones.
 *
 * Device memory is accounted as the server holds it:
 *   held = the slots' workspaces (each sized to the larger of its batch's
 *          largest general-path request -- staged host operands, pre-pass
 *          and conversion buffers, split-K partial sums -- and the sum of
 *          the host operands its persistent launch stages: the general path
 *          runs after that launch on the same stream, so both reuse one
 *          workspace; kept between batches)
 *        + offset tables that the batches in flight build for their plans
 *          (they stay with the plan afterwards, outside the budget).
 * The persistent launch's descriptor buffers (232 bytes per request of a
 * batch, per slot) are not counted.
 * A batch is admitted while held + its growth fits the budget; when the
 * first request does not fit, the workspaces of idle slots are freed and
 * the check repeated, and if it still does not fit the request waits (a
 * stall).  Requests that run in the persistent launch need workspace only
 * for their staged host operands; device operands and resident tensors
 * cost nothing.  A request whose own
 * workspace and tables exceed the whole budget is rejected at submit
 * (TAPP_AMD_ERR_BUDGET).
 *
 * A request's completion is a TAPP_status object (status.h:9-17: "status
 * objects are created by execution functions"), waited on with
 * TAPP_AMD_status_wait and released with TAPP_destroy_status.
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "internal.h"

#define SRV_MAX_SLOTS 16
#define STATUS_MAGIC 0x53544154u

typedef struct srv_request {
    uint32_t magic;
    int refs;                   /* client + server */
    int state;                  /* 0 queued, 1 running, 2 done, 3 cancelled */
    int rc;
    struct tapp_plan* plan;
    unsigned char alpha[16], beta[16];
    const void *A, *B, *C;
    void* D;
    size_t work;                /* workspace of a general-path run (0 for the persistent launch) */
    size_t stage;               /* host operands the persistent launch stages (256-aligned spans) */
    size_t tables;              /* offset tables the plan still has to build on the device */
    double t_submit, t_start, t_done;
    int in_group;               /* in the batch: 1 persistent launch, 0 general path, -1 nothing to launch */
    bool heavy;                 /* above the persistent launch's flop limit, or work above heavy_bytes */
    bool stalled;               /* waited at the head of the queue for memory at least once */
    struct srv_request* next;   /* queue / batch list */
} srv_request;

typedef struct {
    struct tapp_exec ex;        /* private context: stream, workspace, events */
    bool busy;
    srv_request* batch;         /* requests of the batch in flight */
    size_t batch_work;          /* workspace the batch was admitted with */
    size_t batch_tables;        /* tables its requests build (held while it runs) */
    hipEvent_t done;            /* end of the batch */
    hipEvent_t done_group;      /* end of the batch's persistent launch */
    bool group_pending;         /* its requests not yet completed */
    bool failed;                /* the device reported an error while the batch ran */
    bool heavy;                 /* the batch holds a heavy request */
    /* descriptor staging for the persistent launch */
    tk_prob* hprobs;
    int64_t* hend;
    size_t cap;
    void* dmem;
    size_t dcap;
} srv_slot;

struct tapp_server {
    uint32_t magic;
    int device;
    int nslots, max_batch;
    size_t budget;
    size_t tables_held;         /* tables of the batches in flight */
    int64_t large_flop;         /* products above this run on the general GETT path */
    size_t heavy_bytes;         /* a workspace above this makes a request heavy */
    int cus;
    srv_slot slot[SRV_MAX_SLOTS];
    pthread_t thread;
    bool running, stop, paused;
    pthread_mutex_t mu;
    pthread_cond_t cv_work, cv_done;
    srv_request *head, *tail;
    int queued;
    /* statistics */
    int64_t completed, stalls, batches, group_products, max_inflight, inflight, reclaims;
    size_t max_held, held_now;
    double* lat;                /* completed latencies, microseconds */
    size_t nlat, caplat;
    double flop_done, t_first, t_last;
};

static double now_us(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec * 1e6 + (double)t.tv_nsec * 1e-3;
}


static void release(srv_request* r)
{
    if (__atomic_sub_fetch(&r->refs, 1, __ATOMIC_ACQ_REL) == 0) { r->magic = 0; free(r); }
}

/* --- execution of one batch on one slot ----------------------------------- */

static int64_t product_flop(const tk_problem* P)
{
    return 2 * P->nL * P->nM * P->nN * (P->nK > 0 ? P->nK : 1);
}

static int ensure_desc(srv_slot* s, size_t n)
{
    if (n <= s->cap) return 0;
    size_t cap = n < 64 ? 64 : n * 2;
    tk_prob* hp = NULL;
    int64_t* he = NULL;
    if (hipHostMalloc((void**)&hp, sizeof(tk_prob) * cap, 0) != hipSuccess ||
        hipHostMalloc((void**)&he, sizeof(int64_t) * cap, 0) != hipSuccess) {
        (void)hipGetLastError();
        if (hp) hipHostFree(hp);
        return TAPP_AMD_ERR_NOMEM;
    }
    if (s->hprobs) { hipHostFree(s->hprobs); hipHostFree(s->hend); }
    s->hprobs = hp; s->hend = he; s->cap = cap;
    size_t bytes = (sizeof(tk_prob) + sizeof(int64_t)) * cap + 512;
    if (s->dmem) hipFree(s->dmem);
    s->dmem = NULL;
    if (hipMalloc(&s->dmem, bytes) != hipSuccess) { (void)hipGetLastError(); s->cap = 0; return TAPP_AMD_ERR_NOMEM; }
    s->dcap = bytes;
    return 0;
}

#define SRV_ALIGN256(x) (((x) + 255) & ~(size_t)255)

/* Bytes the persistent launch stages for a product it can run: the spans
 * of A, B and C (C only when beta is not 0) that are host memory and not
 * resident on the plan's handle, each 256-aligned.  -1 when the product
 * cannot join the launch (the kernel does not run it, or D is host memory:
 * the launch writes D in place). */
static int64_t group_stage_bytes(struct tapp_plan* p, const tk_problem* P, const void* A, const void* B,
                                 const void* C, void* D)
{
    if (!tk_group_may_run(P) || !tk_is_device_ptr(D) || !A || !B) return -1;
    const void* ops[3] = {A, B, P->beta_zero ? NULL : C};
    int64_t total = 0;
    for (int x = 0; x < 3; x++) {
        if (!ops[x] || P->hi[x] < P->lo[x]) continue;
        if (tk_resolve_operand(p->handle, ops[x], P->lo[x] * 4, (P->hi[x] + 1) * 4)) continue;
        total += (int64_t)SRV_ALIGN256((size_t)(P->hi[x] - P->lo[x] + 1) * 4);
    }
    return total;
}

/* Copies the host operands of a request joining the persistent launch into
 * the batch's workspace at *off (advanced past them) on the slot's stream
 * and points ops[] (A, B, C) at the copies.  False when they no longer fit
 * what the batch was admitted with (the plan's layouts changed since
 * submit) or a copy fails: the request then takes the general path. */
static bool stage_group_operands(struct tapp_exec* ex, struct tapp_plan* p, const tk_problem* P, size_t limit,
                                 size_t* off, const void* ops[3])
{
    for (int x = 0; x < 3; x++) {
        if (!ops[x] || (x == 2 && P->beta_zero) || P->hi[x] < P->lo[x]) continue;
        if (tk_resolve_operand(p->handle, ops[x], P->lo[x] * 4, (P->hi[x] + 1) * 4)) continue;
        const size_t bytes = (size_t)(P->hi[x] - P->lo[x] + 1) * 4;
        if (*off + SRV_ALIGN256(bytes) > limit) return false;
        char* buf = (char*)ex->ws + *off;
        if (hipMemcpyAsync(buf, (const char*)ops[x] + P->lo[x] * 4, bytes, hipMemcpyHostToDevice, ex->stream) !=
            hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        ops[x] = buf - P->lo[x] * 4;
        *off += SRV_ALIGN256(bytes);
    }
    return true;
}

static void launch_batch(struct tapp_server* sv, srv_slot* s)
{
    struct tapp_exec* ex = &s->ex;
    int n = 0;
    int64_t total = 0;
    size_t count = 0;
    s->group_pending = false;
    for (srv_request* r = s->batch; r; r = r->next) count++;
    /* the workspace the batch was admitted with, in one allocation */
    if (s->batch_work && tk_exec_reserve_ws(ex, s->batch_work)) {
        for (srv_request* r = s->batch; r; r = r->next) { r->rc = TAPP_AMD_ERR_NOMEM; r->in_group = -1; }
        hipEventRecord(s->done, ex->stream);
        return;
    }
    if (ensure_desc(s, count)) {
        for (srv_request* r = s->batch; r; r = r->next) { r->rc = TAPP_AMD_ERR_NOMEM; r->in_group = -1; }
        hipEventRecord(s->done, ex->stream);
        return;
    }
    /* which requests join the persistent launch; their host operands are
     * staged first, into the front of the workspace */
    size_t staged = 0;
    for (srv_request* r = s->batch; r; r = r->next) {
        r->t_start = now_us();
        r->in_group = -1;
        tk_problem P;
        int rc = tk_plan_problem(r->plan, ex, r->alpha, r->beta, r->C, &P);
        if (rc) { r->rc = rc; continue; }
        if (P.empty) { r->rc = 0; continue; }
        sv->flop_done += (double)product_flop(&P);
        tk_prob d;
        const void* ops[3] = {r->A, r->B, r->C};
        const bool fits = product_flop(&P) <= sv->large_flop &&
                          (!r->stage || stage_group_operands(ex, r->plan, &P, s->batch_work, &staged, ops));
        if (fits && tk_group_describe(r->plan, ex, &P, ops[0], ops[1], ops[2], r->D, -1, &d, &rc)) {
            s->hprobs[n] = d;
            total += d.L * d.tiles_s * d.tiles_g;
            s->hend[n] = total;
            n++;
            r->rc = 0;
            r->in_group = 1;
        } else if (rc) {
            r->rc = rc;
        } else {
            r->in_group = 0;
        }
    }
    /* the persistent launch first, with its own event */
    if (n > 0) {
        tk_prob* dprobs = (tk_prob*)s->dmem;
        int64_t* dend = (int64_t*)((char*)s->dmem + sizeof(tk_prob) * s->cap);
        unsigned long long* counter =
            (unsigned long long*)((char*)s->dmem + ((sizeof(tk_prob) + sizeof(int64_t)) * s->cap + 255) / 256 * 256);
        int e = (int)hipMemcpyAsync(dprobs, s->hprobs, sizeof(tk_prob) * (size_t)n, hipMemcpyHostToDevice, ex->stream);
        if (!e) e = (int)hipMemcpyAsync(dend, s->hend, sizeof(int64_t) * (size_t)n, hipMemcpyHostToDevice, ex->stream);
        int64_t nb = (int64_t)sv->cus * TK_GROUP_WG_PER_CU;
        if (!e) e = tk_launch_group(dprobs, dend, NULL, n, total, counter, (int)(total < nb ? total : nb), ex->stream);
        if (e)
            for (srv_request* r = s->batch; r; r = r->next)
                if (r->in_group == 1) r->rc = TAPP_AMD_ERR_DEVICE;
        sv->group_products += n;
        if (hipEventRecord(s->done_group, ex->stream) == hipSuccess) s->group_pending = true;
    }
    /* then the general path */
    for (srv_request* r = s->batch; r; r = r->next)
        if (r->in_group == 0) r->rc = tk_execute_on(r->plan, ex, r->alpha, r->A, r->B, r->beta, r->C, r->D, false);
    hipEventRecord(s->done, ex->stream);
}

/* --- dispatcher ----------------------------------------------------------- */

/* Whether a batch event has completed.  hipEventQuery answers success or
 * not-ready; any other answer is an error that says nothing about the
 * batch's progress, so instead of polling on (the batch would never
 * complete) the slot's stream is waited for, and the batch fails if that
 * reports an error too. */
static bool event_done(srv_slot* s, hipEvent_t e)
{
    const hipError_t q = hipEventQuery(e);
    if (q == hipSuccess) return true;
    if (q == hipErrorNotReady) return false;
    (void)hipGetLastError();
    const hipError_t w = hipStreamSynchronize(s->ex.stream);
    if (w != hipSuccess) {
        (void)hipGetLastError();
        s->failed = true;
    }
    fprintf(stderr, "tapp_amd serving loop: hipEventQuery returned %d (%s); stream wait returned %d\n", (int)q,
            hipGetErrorString(q), (int)w);
    return true;
}

/* Device memory the server holds against its budget now (dispatcher
 * thread: the only one that changes the workspaces). */
static size_t held_bytes(const struct tapp_server* sv)
{
    size_t h = sv->tables_held;
    for (int i = 0; i < sv->nslots; i++) h += sv->slot[i].ex.ws_size;
    return h;
}

/* Frees the workspaces of idle slots (all of them, `keep` included when its
 * workspace is larger than `keep_need`). */
static void reclaim(struct tapp_server* sv, int keep, size_t keep_need)
{
    for (int i = 0; i < sv->nslots; i++) {
        srv_slot* o = &sv->slot[i];
        if (o->busy || !o->ex.ws || (i == keep && o->ex.ws_size <= keep_need)) continue;
        tk_exec_release_ws(&o->ex);
        sv->reclaims++;
    }
}

static bool plan_in(const srv_request* list, const srv_request* stop, const struct tapp_plan* p)
{
    for (const srv_request* q = list; q && q != stop; q = q->next)
        if (q->plan == p) return true;
    return false;
}

/* Completes the batch's requests that are done: with group_only, those of
 * the persistent launch; otherwise all that remain, and the slot frees. */
static void finish_requests(struct tapp_server* sv, srv_slot* s, bool group_only)
{
    double t = now_us();
    srv_request** link = &s->batch;
    while (*link) {
        srv_request* r = *link;
        if (group_only && r->in_group != 1) { link = &r->next; continue; }
        *link = r->next;
        r->t_done = t;
        if (s->failed && r->rc == 0) r->rc = TAPP_AMD_ERR_DEVICE;
        r->state = 2;
        if (sv->nlat == sv->caplat) {
            size_t cap = sv->caplat ? sv->caplat * 2 : 4096;
            double* l = realloc(sv->lat, sizeof(double) * cap);
            if (l) { sv->lat = l; sv->caplat = cap; }
        }
        if (sv->nlat < sv->caplat) sv->lat[sv->nlat++] = r->t_done - r->t_submit;
        sv->completed++;
        sv->inflight--;
        sv->t_last = t;
        r->next = NULL;
        release(r);
    }
    if (group_only) return;
    sv->tables_held -= s->batch_tables;
    s->batch = NULL;
    s->batch_tables = 0;
    s->batch_work = 0;
    s->busy = false;
    s->group_pending = false;
    s->failed = false;
}

static void* dispatcher(void* arg)
{
    struct tapp_server* sv = arg;
    hipSetDevice(sv->device);
    pthread_mutex_lock(&sv->mu);
    for (;;) {
        bool progressed = false;
        /* completions */
        for (int i = 0; i < sv->nslots; i++) {
            srv_slot* s = &sv->slot[i];
            if (s->busy && s->group_pending && event_done(s, s->done_group)) {
                finish_requests(sv, s, true);
                s->group_pending = false;
                progressed = true;
            }
            if (s->busy && event_done(s, s->done)) {
                finish_requests(sv, s, false);
                progressed = true;
            }
        }
        if (progressed) pthread_cond_broadcast(&sv->cv_done);
        /* admissions */
        if (!sv->paused) {
            for (int i = 0; i < sv->nslots && sv->head; i++) {
                srv_slot* s = &sv->slot[i];
                if (s->busy) continue;
                /* heavy requests are passed over while every other slot
                 * holds one, so that one stream stays free for small ones */
                int heavy_busy = 0;
                for (int j = 0; j < sv->nslots; j++) heavy_busy += sv->slot[j].busy && sv->slot[j].heavy;
                const bool allow_heavy = heavy_busy < sv->nslots - 1 || sv->nslots == 1;
                srv_request *first = NULL, *last = NULL, *prev = NULL;
                srv_request** link = &sv->head;
                int taken = 0;
                size_t need = 0, stg = 0, tabs = 0;
                bool heavy = false;
                size_t held = held_bytes(sv);
                while (*link && taken < sv->max_batch) {
                    srv_request* r = *link;
                    if (r->state == 3) {          /* cancelled while queued */
                        *link = r->next;
                        if (sv->tail == r) sv->tail = prev;
                        sv->queued--;
                        r->next = NULL;
                        release(r);
                        continue;
                    }
                    if (r->heavy && !allow_heavy) {
                        prev = r;
                        link = &r->next;
                        continue;
                    }
                    /* tables count once per plan and only while not yet built */
                    const size_t t = (r->tables && !plan_in(first, NULL, r->plan) && tk_plan_needs_tables(r->plan, &s->ex))
                                         ? r->tables : 0;
                    const size_t n2 = r->work > need ? r->work : need, s2 = stg + r->stage;
                    const size_t w2 = n2 > s2 ? n2 : s2;
                    size_t grow = (w2 > s->ex.ws_size ? w2 - s->ex.ws_size : 0) + tabs + t;
                    if (held + grow > sv->budget && taken == 0) {
                        reclaim(sv, i, w2);
                        held = held_bytes(sv);
                        grow = (w2 > s->ex.ws_size ? w2 - s->ex.ws_size : 0) + t;
                    }
                    if (held + grow > sv->budget) {
                        bool any = false;
                        for (int j = 0; j < sv->nslots; j++) any |= sv->slot[j].busy;
                        if (taken == 0 && !any) {
                            /* nothing runs and nothing more can be freed: it will
                             * never fit (its plan's layout grew since submit) */
                            *link = r->next;
                            if (sv->tail == r) sv->tail = prev;
                            sv->queued--;
                            r->next = NULL;
                            r->rc = TAPP_AMD_ERR_BUDGET;
                            r->state = 2;
                            release(r);
                            pthread_cond_broadcast(&sv->cv_done);
                            continue;
                        }
                        if (taken == 0 && !r->stalled) { r->stalled = true; sv->stalls++; }
                        break;
                    }
                    *link = r->next;
                    if (sv->tail == r) sv->tail = prev;
                    sv->queued--;
                    r->next = NULL;
                    r->state = 1;
                    if (last) last->next = r; else first = r;
                    last = r;
                    need = n2;
                    stg = s2;
                    tabs += t;
                    heavy |= r->heavy;
                    taken++;
                }
                if (!taken) break;
                s->batch = first;
                s->heavy = heavy;
                s->batch_work = need > stg ? need : stg;
                s->batch_tables = tabs;
                sv->tables_held += tabs;
                sv->inflight += taken;
                if (sv->inflight > sv->max_inflight) sv->max_inflight = sv->inflight;
                sv->batches++;
                s->busy = true;
                pthread_mutex_unlock(&sv->mu);
                launch_batch(sv, s);
                pthread_mutex_lock(&sv->mu);
                progressed = true;
            }
        }
        sv->held_now = held_bytes(sv);
        if (sv->held_now > sv->max_held) sv->max_held = sv->held_now;
        bool any_busy = false;
        for (int i = 0; i < sv->nslots; i++) any_busy |= sv->slot[i].busy;
        if (sv->stop && !sv->head && !any_busy) break;
        if (!progressed) {
            if (any_busy) {
                /* poll completions at a short interval */
                pthread_mutex_unlock(&sv->mu);
                struct timespec ts = {0, 5000};
                nanosleep(&ts, NULL);
                pthread_mutex_lock(&sv->mu);
            } else {
                pthread_cond_wait(&sv->cv_work, &sv->mu);
            }
        }
    }
    pthread_mutex_unlock(&sv->mu);
    return NULL;
}

/* --- public API ------------------------------------------------------------- */

TAPP_error TAPP_AMD_create_server(TAPP_AMD_server* server, TAPP_executor exec, size_t hbm_budget_bytes, int slots,
                                  int max_batch, int flags)
{
    if (!server || slots <= 0 || slots > SRV_MAX_SLOTS || max_batch <= 0) return TAPP_AMD_ERR_INVALID;
    struct tapp_exec* base = tk_exec_lookup(exec);
    if (!base) return 33;
    struct tapp_server* sv = calloc(1, sizeof *sv);
    if (!sv) return TAPP_AMD_ERR_NOMEM;
    sv->magic = TAPP_AMD_EXEC_MAGIC;
    sv->device = base->device;
    sv->nslots = slots;
    sv->max_batch = max_batch;
    sv->budget = hbm_budget_bytes;
    sv->large_flop = (int64_t)1 << 31;
    sv->heavy_bytes = (size_t)4 << 20;
    sv->paused = (flags & TAPP_AMD_SERVER_PAUSED) != 0;
    if (!base->have_stream || hipDeviceGetAttribute(&sv->cus, hipDeviceAttributeMultiprocessorCount, base->device) !=
                             hipSuccess || sv->cus <= 0) {
        (void)hipGetLastError();
        sv->cus = 256;
    }
    pthread_mutex_init(&sv->mu, NULL);
    pthread_cond_init(&sv->cv_work, NULL);
    pthread_cond_init(&sv->cv_done, NULL);
    for (int i = 0; i < slots; i++) {
        struct tapp_exec* e = &sv->slot[i].ex;
        e->kind = 1;
        e->magic = TAPP_AMD_EXEC_MAGIC;
        e->device = base->device;
        e->cus = base->cus;
        e->split_bf16 = base->split_bf16;
        e->force_kernel = base->force_kernel;
        if (base->have_stream) {
            if (hipSetDevice(base->device) != hipSuccess ||
                hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&sv->slot[i].done, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&sv->slot[i].done_group, hipEventDisableTiming) != hipSuccess ||
                hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess) {
                (void)hipGetLastError();
                TAPP_AMD_destroy_server((TAPP_AMD_server)sv);
                return TAPP_AMD_ERR_DEVICE;
            }
            e->own_stream = true;
            e->have_stream = true;
        }
    }
    if (!base->have_stream) sv->paused = true;   /* no device: queue-only (admission logic) */
    if (pthread_create(&sv->thread, NULL, dispatcher, sv) != 0) {
        TAPP_AMD_destroy_server((TAPP_AMD_server)sv);
        return TAPP_AMD_ERR_NOMEM;
    }
    sv->running = true;
    *server = (TAPP_AMD_server)sv;
    return 0;
}

/* Device memory a request needs: the workspace of a general-path run
 * (staged host operands, pre-pass and conversion buffers, split-K
 * partials) -- none when it will run in the persistent launch, which
 * needs only its staged host operands (*stage) -- and the offset tables
 * its plan has yet to build on the device. */
static int footprint(struct tapp_server* sv, struct tapp_plan* p, const void* alpha, const void* A, const void* B,
                     const void* beta, const void* C, void* D, size_t* work, size_t* stage, size_t* tables,
                     int64_t* flop)
{
    const struct tapp_exec* ex = &sv->slot[0].ex;   /* the slots' attributes */
    tk_problem P;
    int rc = tk_plan_problem(p, ex, alpha, beta, C, &P);
    if (rc) return rc;
    *work = *stage = *tables = 0;
    *flop = P.empty ? 0 : product_flop(&P);
    if (P.empty) return 0;
    bool devops = false;
    if ((rc = tk_call_bytes(p, ex, &P, A, B, C, D, work, tables, &devops))) return rc;
    if (*flop > sv->large_flop) return 0;
    if (devops && tk_group_may_run(&P)) {
        *work = 0;
    } else {
        const int64_t sb = group_stage_bytes(p, &P, A, B, C, D);
        if (sb > 0 && (size_t)sb <= sv->heavy_bytes) { *work = 0; *stage = (size_t)sb; }
    }
    return 0;
}

TAPP_error TAPP_AMD_server_submit(TAPP_AMD_server server, TAPP_tensor_product plan, const void* alpha,
                                  const void* A, const void* B, const void* beta, const void* C, void* D,
                                  TAPP_status* status)
{
    struct tapp_server* sv = (struct tapp_server*)server;
    if (!sv || sv->magic != TAPP_AMD_EXEC_MAGIC || !plan || !status) return TAPP_AMD_ERR_INVALID;
    struct tapp_plan* p = (struct tapp_plan*)plan;
    size_t work = 0, stage = 0, tables = 0;
    int64_t fl = 0;
    int rc = footprint(sv, p, alpha, A, B, beta, C, D, &work, &stage, &tables, &fl);
    if (rc) return rc;
    if (work + stage + tables > sv->budget) return TAPP_AMD_ERR_BUDGET;
    srv_request* r = calloc(1, sizeof *r);
    if (!r) return TAPP_AMD_ERR_NOMEM;
    r->magic = STATUS_MAGIC;
    r->refs = 2;
    r->plan = p;
    const struct tapp_info* dinfo = (const struct tapp_info*)p->info[3];
    size_t ss = tk_type_size(dinfo->type);
    if (alpha && ss) memcpy(r->alpha, alpha, ss);
    if (beta && ss) memcpy(r->beta, beta, ss);
    r->A = A; r->B = B; r->C = C; r->D = D;
    r->work = work;
    r->stage = stage;
    r->tables = tables;
    r->heavy = fl > sv->large_flop || work > sv->heavy_bytes;
    r->t_submit = now_us();
    pthread_mutex_lock(&sv->mu);
    if (!sv->t_first) sv->t_first = r->t_submit;
    if (sv->tail) sv->tail->next = r; else sv->head = r;
    sv->tail = r;
    sv->queued++;
    pthread_cond_signal(&sv->cv_work);
    pthread_mutex_unlock(&sv->mu);
    *status = (TAPP_status)r;
    return 0;
}

/* Waits for the request behind `status`; *result receives its TAPP code. */
TAPP_error TAPP_AMD_status_wait(TAPP_AMD_server server, TAPP_status status, TAPP_error* result)
{
    struct tapp_server* sv = (struct tapp_server*)server;
    srv_request* r = (srv_request*)status;
    if (!sv || !r || r->magic != STATUS_MAGIC) return TAPP_AMD_ERR_INVALID;
    pthread_mutex_lock(&sv->mu);
    while (r->state < 2) pthread_cond_wait(&sv->cv_done, &sv->mu);
    if (result) *result = r->state == 3 ? TAPP_AMD_ERR_INVALID : r->rc;
    pthread_mutex_unlock(&sv->mu);
    return 0;
}

/* Cancels a request still in the queue; returns TAPP_AMD_ERR_BUSY when it has
 * already been admitted. */
TAPP_error TAPP_AMD_server_cancel(TAPP_AMD_server server, TAPP_status status)
{
    struct tapp_server* sv = (struct tapp_server*)server;
    srv_request* r = (srv_request*)status;
    if (!sv || !r || r->magic != STATUS_MAGIC) return TAPP_AMD_ERR_INVALID;
    pthread_mutex_lock(&sv->mu);
    int rc = 0;
    if (r->state == 0) {
        r->state = 3;
        pthread_cond_broadcast(&sv->cv_done);
    } else if (r->state == 1) {
        rc = TAPP_AMD_ERR_BUSY;
    }
    pthread_mutex_unlock(&sv->mu);
    return rc;
}

TAPP_error TAPP_destroy_status(TAPP_status status)
{
    srv_request* r = (srv_request*)status;
    if (!r) return 0;
    if (r->magic != STATUS_MAGIC) return TAPP_AMD_ERR_INVALID;
    /* the server's reference is dropped when the request completes or its
     * cancellation is observed by the dispatcher; callers destroy statuses of
     * finished or cancelled requests */
    release(r);
    return 0;
}

TAPP_error TAPP_AMD_server_set_paused(TAPP_AMD_server server, int paused)
{
    struct tapp_server* sv = (struct tapp_server*)server;
    if (!sv || sv->magic != TAPP_AMD_EXEC_MAGIC) return TAPP_AMD_ERR_INVALID;
    pthread_mutex_lock(&sv->mu);
    bool has_device = sv->slot[0].ex.stream != NULL;
    sv->paused = paused || !has_device;
    pthread_cond_signal(&sv->cv_work);
    pthread_mutex_unlock(&sv->mu);
    return 0;
}

TAPP_error TAPP_AMD_server_drain(TAPP_AMD_server server)
{
    struct tapp_server* sv = (struct tapp_server*)server;
    if (!sv || sv->magic != TAPP_AMD_EXEC_MAGIC) return TAPP_AMD_ERR_INVALID;
    pthread_mutex_lock(&sv->mu);
    for (;;) {
        bool busy = false;
        for (int i = 0; i < sv->nslots; i++) busy |= sv->slot[i].busy;
        if (!busy && (!sv->head || sv->paused)) break;
        pthread_cond_wait(&sv->cv_done, &sv->mu);
    }
    pthread_mutex_unlock(&sv->mu);
    return 0;
}

static int cmp_d(const void* a, const void* b)
{
    double x = *(const double*)a, y = *(const double*)b;
    return x < y ? -1 : x > y;
}

/* stats: [0] completed, [1] p50 us, [2] p99 us, [3] mean us, [4] stalls,
 * [5] max in flight, [6] batches, [7] products in persistent launches,
 * [8] GFLOP/s over first submit .. last completion, [9] queued now,
 * [4] counts requests that waited for memory at least once, [10] device
 * bytes held now, [11] most device bytes held, [12] workspace reclaims.  reset != 0 clears the counters (not the bytes held now). */
TAPP_error TAPP_AMD_server_stats(TAPP_AMD_server server, double stats[13], int reset)
{
    struct tapp_server* sv = (struct tapp_server*)server;
    if (!sv || sv->magic != TAPP_AMD_EXEC_MAGIC || !stats) return TAPP_AMD_ERR_INVALID;
    pthread_mutex_lock(&sv->mu);
    memset(stats, 0, sizeof(double) * 13);
    stats[0] = (double)sv->completed;
    if (sv->nlat) {
        double* c = malloc(sizeof(double) * sv->nlat);
        if (c) {
            memcpy(c, sv->lat, sizeof(double) * sv->nlat);
            qsort(c, sv->nlat, sizeof(double), cmp_d);
            stats[1] = c[(size_t)floor(0.50 * (double)(sv->nlat - 1))];
            stats[2] = c[(size_t)floor(0.99 * (double)(sv->nlat - 1))];
            double s = 0;
            for (size_t i = 0; i < sv->nlat; i++) s += c[i];
            stats[3] = s / (double)sv->nlat;
            free(c);
        }
    }
    stats[4] = (double)sv->stalls;
    stats[5] = (double)sv->max_inflight;
    stats[6] = (double)sv->batches;
    stats[7] = (double)sv->group_products;
    if (sv->t_last > sv->t_first) stats[8] = sv->flop_done / ((sv->t_last - sv->t_first) * 1e-6) / 1e9;
    stats[9] = (double)sv->queued;
    stats[10] = (double)sv->held_now;
    stats[11] = (double)sv->max_held;
    stats[12] = (double)sv->reclaims;
    if (reset) {
        sv->completed = sv->stalls = sv->batches = sv->group_products = sv->max_inflight = sv->reclaims = 0;
        sv->max_held = sv->held_now;
        sv->nlat = 0;
        sv->flop_done = 0;
        sv->t_first = sv->t_last = 0;
    }
    pthread_mutex_unlock(&sv->mu);
    return 0;
}

TAPP_error TAPP_AMD_destroy_server(TAPP_AMD_server server)
{
    struct tapp_server* sv = (struct tapp_server*)server;
    if (!sv) return 0;
    if (sv->running) {
        pthread_mutex_lock(&sv->mu);
        sv->stop = true;
        if (sv->paused) {
            /* drop what never ran */
            while (sv->head) {
                srv_request* r = sv->head;
                sv->head = r->next;
                r->state = 3;
                r->next = NULL;
                release(r);
            }
            sv->tail = NULL;
        }
        pthread_cond_signal(&sv->cv_work);
        pthread_cond_broadcast(&sv->cv_done);
        pthread_mutex_unlock(&sv->mu);
        pthread_join(sv->thread, NULL);
    }
    for (int i = 0; i < sv->nslots; i++) {
        srv_slot* s = &sv->slot[i];
        if (s->ex.stream) hipStreamSynchronize(s->ex.stream);
        if (s->ex.ws) hipFree(s->ex.ws);
        if (s->dmem) hipFree(s->dmem);
        if (s->hprobs) { hipHostFree(s->hprobs); hipHostFree(s->hend); }
        if (s->done) hipEventDestroy(s->done);
        if (s->done_group) hipEventDestroy(s->done_group);
        if (s->ex.ev0) hipEventDestroy(s->ex.ev0);
        if (s->ex.ev1) hipEventDestroy(s->ex.ev1);
        if (s->ex.own_stream && s->ex.stream) hipStreamDestroy(s->ex.stream);
    }
    free(sv->lat);
    pthread_mutex_destroy(&sv->mu);
    pthread_cond_destroy(&sv->cv_work);
    pthread_cond_destroy(&sv->cv_done);
    sv->magic = 0;
    free(sv);
    return 0;
}
