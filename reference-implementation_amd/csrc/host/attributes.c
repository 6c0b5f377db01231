/*
 * TAPP_attr_set / _get / _clear (attributes.h:14-18, declared but never
 * defined by the reference).  The attribute object is an executor; keys are
 * listed in tapp_amd.h.
 */
#include "internal.h"

static struct tapp_exec* exec_of(TAPP_attr attr)
{
    return tk_exec_lookup((TAPP_executor)attr);
}

TAPP_error TAPP_attr_set(TAPP_attr attr, TAPP_key key, void* value)
{
    struct tapp_exec* e = exec_of(attr);
    if (!e || !value) return TAPP_AMD_ERR_INVALID;
    switch (key) {
    case TAPP_AMD_ATTR_SPLIT_BF16: {
        int v = *(int*)value;
        if (v < 0 || v > 3) return TAPP_AMD_ERR_INVALID;
        e->split_bf16 = v;
        return 0;
    }
    case TAPP_AMD_ATTR_KERNEL: {
        int k = *(int*)value;
        if (k < 0 || k > 5 || k == 4) return TAPP_AMD_ERR_INVALID;
        e->force_kernel = k;
        return 0;
    }
    case TAPP_AMD_ATTR_TIMING: e->timing = *(int*)value != 0; return 0;
    case TAPP_AMD_ATTR_TILES: {
        int t = *(int*)value;
        if (t < 0 || t > 3) return TAPP_AMD_ERR_INVALID;
        e->tiles = t;
        return 0;
    }
    case TAPP_AMD_ATTR_AUTOTUNE: {
        int a = *(int*)value;
        if (a < 0 || a > 2) return TAPP_AMD_ERR_INVALID;
        e->autotune = a;
        return 0;
    }
    case TAPP_AMD_ATTR_SELF_CHECK: e->kind = *(int*)value ? 12 : 1; return 0;
    case TAPP_AMD_ATTR_STREAM: {
        /* any stream, the null stream (0) included: later calls run ordered
         * after the work already queued on it */
        hipStream_t s = *(hipStream_t*)value;
        tk_exec_lock(e);
        if (e->have_stream) hipStreamSynchronize(e->stream);
        if (e->own_stream && e->stream) hipStreamDestroy(e->stream);
        e->stream = s;
        e->own_stream = false;
        e->have_stream = true;
        tk_exec_unlock(e);
        return 0;
    }
    default: return TAPP_AMD_ERR_KEY;
    }
}

TAPP_error TAPP_attr_get(TAPP_attr attr, TAPP_key key, void** value)
{
    struct tapp_exec* e = exec_of(attr);
    if (!e || !value) return TAPP_AMD_ERR_INVALID;
    switch (key) {
    case TAPP_AMD_ATTR_SPLIT_BF16: *(int*)value = e->split_bf16; return 0;
    case TAPP_AMD_ATTR_KERNEL: *(int*)value = e->force_kernel; return 0;
    case TAPP_AMD_ATTR_TIMING: *(int*)value = e->timing; return 0;
    case TAPP_AMD_ATTR_TILES: *(int*)value = e->tiles; return 0;
    case TAPP_AMD_ATTR_AUTOTUNE: *(int*)value = e->autotune; return 0;
    case TAPP_AMD_ATTR_LAST_KERNEL: *(int*)value = e->last_kernel; return 0;
    case TAPP_AMD_ATTR_LAST_DEVICE_US: *(double*)value = e->last_device_us; return 0;
    case TAPP_AMD_ATTR_STREAM: *(hipStream_t*)value = e->stream; return 0;
    case TAPP_AMD_ATTR_SELF_CHECK: *(int*)value = e->kind == 12; return 0;
    case TAPP_AMD_ATTR_LAST_REUSED: *(int*)value = e->last_reused; return 0;
    case TAPP_AMD_ATTR_LAST_KERNEL_NAME: *(const char**)value = e->last_kernel_name; return 0;
    default: return TAPP_AMD_ERR_KEY;
    }
}

TAPP_error TAPP_attr_clear(TAPP_attr attr, TAPP_key key)
{
    struct tapp_exec* e = exec_of(attr);
    if (!e) return TAPP_AMD_ERR_INVALID;
    switch (key) {
    case TAPP_AMD_ATTR_SPLIT_BF16: e->split_bf16 = 0; return 0;
    case TAPP_AMD_ATTR_KERNEL: e->force_kernel = 0; return 0;
    case TAPP_AMD_ATTR_TIMING: e->timing = 0; return 0;
    case TAPP_AMD_ATTR_TILES: e->tiles = 0; return 0;
    case TAPP_AMD_ATTR_AUTOTUNE: e->autotune = 0; return 0;
    case TAPP_AMD_ATTR_SELF_CHECK: e->kind = 1; return 0;
    default: return TAPP_AMD_ERR_KEY;
    }
}
