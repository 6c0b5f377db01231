/*
 * Resident tensors: handle-scoped device copies of host buffers.
 *
 * The reference's handle is an opaque library context (handle.h:9-24; a bare
 * malloc in handle.c:9-17) that plans keep (product.h:33-47).  Here a handle
 * also owns a table of resident buffers: TAPP_AMD_make_resident copies a host
 * buffer into HBM once, and every later execute of a plan created with that
 * handle that passes a host pointer whose accessed bytes lie inside the
 * buffer reads the device copy, so a serving caller that keeps reusing an
 * operand pays the PCIe transfer once instead of on every call.  A, B and C
 * are redirected; D is always written through the caller's pointer.
 *
 * tk_resolve_operand is the one place that decides where an operand lives:
 * device memory is used in place, a resident host range is translated to its
 * copy, and anything else is host memory that the caller stages.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "internal.h"

struct resident {
    const char* host;
    size_t bytes;
    char* dev;
};

#define MAX_HANDLES 4096
static struct tapp_handle_obj* live[MAX_HANDLES];
static int nlive;
static pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;   /* registry and every table */

bool tk_is_device_ptr(const void* p)
{
    if (!p) return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    (void)hipGetLastError();   /* an unregistered host pointer reports an error: not sticky */
    if (e != hipSuccess) return false;
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

void tk_handle_register(struct tapp_handle_obj* h)
{
    pthread_mutex_lock(&mu);
    if (nlive < MAX_HANDLES) live[nlive++] = h;
    pthread_mutex_unlock(&mu);
}

static struct tapp_handle_obj* find_locked(TAPP_handle handle)
{
    for (int i = 0; i < nlive; i++)
        if (live[i] == (struct tapp_handle_obj*)handle) return live[i];
    return NULL;
}

/* Unregisters one of this library's handles and frees its resident copies;
 * any other pointer is left untouched. */
void tk_handle_unregister(struct tapp_handle_obj* h)
{
    bool found = false;
    pthread_mutex_lock(&mu);
    for (int i = 0; i < nlive; i++)
        if (live[i] == h) { live[i] = live[--nlive]; found = true; break; }
    pthread_mutex_unlock(&mu);
    if (!found) return;
    struct resident* r = (struct resident*)h->res;
    if (r) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(h->device);
        for (int i = 0; i < h->nres; i++) (void)hipFree(r[i].dev);
        (void)hipSetDevice(cur);
        (void)hipGetLastError();
        free(r);
    }
    h->res = NULL;
    h->nres = 0;
}

const void* tk_resolve_operand(TAPP_handle handle, const void* p, int64_t lo_bytes, int64_t hi_bytes)
{
    if (!p) return NULL;
    if (tk_is_device_ptr(p)) return p;
    const void* out = NULL;
    pthread_mutex_lock(&mu);
    struct tapp_handle_obj* h = find_locked(handle);
    if (h && h->nres) {
        const char* lo = (const char*)p + lo_bytes;
        const char* hi = (const char*)p + hi_bytes;
        const struct resident* r = (const struct resident*)h->res;
        for (int i = 0; i < h->nres; i++)
            if (lo >= r[i].host && hi <= r[i].host + r[i].bytes) {
                out = r[i].dev + ((const char*)p - r[i].host);
                break;
            }
    }
    pthread_mutex_unlock(&mu);
    return out;
}

static int copy_in(int device, char* dev, const void* host, size_t bytes)
{
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) cur = device;
    if (hipSetDevice(device) != hipSuccess) { (void)hipGetLastError(); return TAPP_AMD_ERR_DEVICE; }
    hipError_t e = hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice);
    (void)hipSetDevice(cur);
    if (e != hipSuccess) { (void)hipGetLastError(); return TAPP_AMD_ERR_DEVICE; }
    return 0;
}

TAPP_error TAPP_AMD_make_resident(TAPP_handle handle, const void* host, size_t bytes)
{
    if (!host || bytes == 0 || tk_is_device_ptr(host)) return TAPP_AMD_ERR_INVALID;
    pthread_mutex_lock(&mu);
    struct tapp_handle_obj* h = find_locked(handle);
    int rc = h ? 0 : TAPP_AMD_ERR_INVALID;
    const char* lo = (const char*)host;
    struct resident* r = h ? (struct resident*)h->res : NULL;
    for (int i = 0; !rc && i < h->nres; i++)       /* ranges of one handle do not overlap */
        if (lo < r[i].host + r[i].bytes && r[i].host < lo + bytes) rc = TAPP_AMD_ERR_INVALID;
    if (!rc && h->nres == h->cap) {
        int cap = h->cap ? 2 * h->cap : 8;
        struct resident* n = realloc(r, sizeof *n * (size_t)cap);
        if (!n) rc = TAPP_AMD_ERR_NOMEM;
        else { h->res = n; h->cap = cap; r = n; }
    }
    char* dev = NULL;
    if (!rc) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        if (hipSetDevice(h->device) != hipSuccess || hipMalloc((void**)&dev, bytes) != hipSuccess) {
            (void)hipGetLastError();
            rc = TAPP_AMD_ERR_NOMEM;
            dev = NULL;
        }
        (void)hipSetDevice(cur);
    }
    if (!rc && (rc = copy_in(h->device, dev, host, bytes))) { (void)hipFree(dev); dev = NULL; }
    if (!rc) r[h->nres++] = (struct resident){lo, bytes, dev};
    pthread_mutex_unlock(&mu);
    return rc;
}

TAPP_error TAPP_AMD_update_resident(TAPP_handle handle, const void* host)
{
    pthread_mutex_lock(&mu);
    struct tapp_handle_obj* h = find_locked(handle);
    int rc = TAPP_AMD_ERR_INVALID;
    if (h) {
        struct resident* r = (struct resident*)h->res;
        for (int i = 0; i < h->nres; i++)
            if (r[i].host == (const char*)host) {
                /* kernels already queued on any stream (executors, serving
                 * slots) may still read the copy: let them finish first, as
                 * release does, so none sees a half-updated operand */
                if (hipSetDevice(h->device) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
                    (void)hipGetLastError();
                    rc = TAPP_AMD_ERR_DEVICE;
                    break;
                }
                rc = copy_in(h->device, r[i].dev, host, r[i].bytes);
                break;
            }
    }
    pthread_mutex_unlock(&mu);
    return rc;
}

TAPP_error TAPP_AMD_release_resident(TAPP_handle handle, const void* host)
{
    pthread_mutex_lock(&mu);
    struct tapp_handle_obj* h = find_locked(handle);
    int rc = TAPP_AMD_ERR_INVALID;
    if (h) {
        struct resident* r = (struct resident*)h->res;
        for (int i = 0; i < h->nres; i++)
            if (r[i].host == (const char*)host) {
                int cur = 0;
                (void)hipGetDevice(&cur);
                (void)hipSetDevice(h->device);
                (void)hipDeviceSynchronize();   /* reads of the copy may still be queued */
                (void)hipFree(r[i].dev);
                (void)hipSetDevice(cur);
                (void)hipGetLastError();
                r[i] = r[--h->nres];
                rc = 0;
                break;
            }
    }
    pthread_mutex_unlock(&mu);
    return rc;
}

TAPP_error TAPP_AMD_resident_info(TAPP_handle handle, int64_t info[2])
{
    if (!info) return TAPP_AMD_ERR_INVALID;
    pthread_mutex_lock(&mu);
    struct tapp_handle_obj* h = find_locked(handle);
    if (h) {
        const struct resident* r = (const struct resident*)h->res;
        int64_t bytes = 0;
        for (int i = 0; i < h->nres; i++) bytes += (int64_t)r[i].bytes;
        info[0] = h->nres;
        info[1] = bytes;
    }
    pthread_mutex_unlock(&mu);
    return h ? 0 : TAPP_AMD_ERR_INVALID;
}
