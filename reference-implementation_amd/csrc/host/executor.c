/*
 * Executor = device context: a HIP stream on the device current at creation,
 * a device workspace grown on demand, timing events and attributes.
 *
 * The reference's executor is a malloc'd int holding the kind 1
 * (executor.c:9-18) and TAPP_execute_product accepts any executor whose int
 * is 1, 2 or 12 (product.c:878-885).  The first field here is that int, and
 * a live-executor registry lets TAPP_execute_product tell this library's
 * executors from a caller's bare int: the latter runs on a process-wide
 * default context.
 */
#define _GNU_SOURCE   /* PTHREAD_MUTEX_RECURSIVE */
#include <pthread.h>
#include <stdlib.h>

#include "internal.h"

#define MAX_EXEC 4096
static struct tapp_exec* live[MAX_EXEC];
static int nlive;
static pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
static struct tapp_exec* default_exec;

void tk_exec_register(struct tapp_exec* e)
{
    pthread_mutex_lock(&mu);
    if (nlive < MAX_EXEC) live[nlive++] = e;
    pthread_mutex_unlock(&mu);
}

void tk_exec_unregister(struct tapp_exec* e)
{
    pthread_mutex_lock(&mu);
    for (int i = 0; i < nlive; i++)
        if (live[i] == e) { live[i] = live[--nlive]; break; }
    pthread_mutex_unlock(&mu);
}

struct tapp_exec* tk_exec_lookup(TAPP_executor h)
{
    struct tapp_exec* e = (struct tapp_exec*)h;
    struct tapp_exec* found = NULL;
    pthread_mutex_lock(&mu);
    for (int i = 0; i < nlive; i++)
        if (live[i] == e) { found = e; break; }
    pthread_mutex_unlock(&mu);
    return found;
}

void tk_mutex_init_recursive(pthread_mutex_t* m)
{
    pthread_mutexattr_t a;
    pthread_mutexattr_init(&a);
    pthread_mutexattr_settype(&a, PTHREAD_MUTEX_RECURSIVE);
    pthread_mutex_init(m, &a);
    pthread_mutexattr_destroy(&a);
}

void tk_exec_lock(struct tapp_exec* e)
{
    if (e && e->mu_init) pthread_mutex_lock(&e->mu);
}

void tk_exec_unlock(struct tapp_exec* e)
{
    if (e && e->mu_init) pthread_mutex_unlock(&e->mu);
}

static int exec_init(struct tapp_exec* e)
{
    e->kind = 1;
    e->magic = TAPP_AMD_EXEC_MAGIC;
    tk_mutex_init_recursive(&e->mu);
    e->mu_init = true;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return TAPP_AMD_ERR_DEVICE;
    e->device = dev;
    if (hipDeviceGetAttribute(&e->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || e->cus <= 0)
        e->cus = 256;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) return TAPP_AMD_ERR_DEVICE;
    e->own_stream = true;
    e->have_stream = true;
    if (hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess) return TAPP_AMD_ERR_DEVICE;
    return 0;
}

struct tapp_exec* tk_default_exec(void)
{
    pthread_mutex_lock(&mu);
    if (!default_exec) {
        struct tapp_exec* e = calloc(1, sizeof *e);
        if (e && exec_init(e) == 0) default_exec = e;
        else free(e);
    }
    struct tapp_exec* d = default_exec;
    pthread_mutex_unlock(&mu);
    return d;
}

TAPP_error TAPP_create_executor(TAPP_executor* exec)
{
    struct tapp_exec* e = calloc(1, sizeof *e);
    if (!e) return TAPP_AMD_ERR_NOMEM;
    int rc = exec_init(e);
    if (rc) {
        /* No usable device: keep the reference's contract (an executor whose
         * int is 1 is returned with code 0); execution reports the device
         * error. */
        e->kind = 1;
        e->stream = NULL;
        e->own_stream = false;
        e->have_stream = false;
    }
    tk_exec_register(e);
    *exec = (TAPP_executor)e;
    return 0;
}

TAPP_error TAPP_destroy_executor(TAPP_executor exec)
{
    struct tapp_exec* e = tk_exec_lookup(exec);
    if (!e) return 0;
    tk_exec_unregister(e);
    /* waits for a call another thread may still be making on it */
    tk_exec_lock(e);
    if (e->have_stream) hipStreamSynchronize(e->stream);
    if (e->ws) hipFree(e->ws);
    if (e->ev0) hipEventDestroy(e->ev0);
    if (e->ev1) hipEventDestroy(e->ev1);
    if (e->own_stream && e->stream) hipStreamDestroy(e->stream);
    tk_exec_unlock(e);
    if (e->mu_init) pthread_mutex_destroy(&e->mu);
    free(e);
    return 0;
}
