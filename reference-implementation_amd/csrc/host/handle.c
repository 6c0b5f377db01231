/*
 * Library handle (reference: handle.c:9-17, a bare malloc).  Here it records
 * the HIP device current at creation and owns the handle's resident tensors
 * (resident.c).  Plans keep the handle by value; execution needs it only to
 * find resident operands, so a plan whose handle is not one of this
 * library's still executes, staging host operands as usual.
 */
#include <stdlib.h>

#include "internal.h"

TAPP_error TAPP_create_handle(TAPP_handle* handle)
{
    struct tapp_handle_obj* h = calloc(1, sizeof *h);
    if (!h) return TAPP_AMD_ERR_NOMEM;
    h->magic = TAPP_AMD_EXEC_MAGIC;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    h->device = dev;
    tk_handle_register(h);
    *handle = (TAPP_handle)h;
    return 0;
}

TAPP_error TAPP_destroy_handle(TAPP_handle handle)
{
    struct tapp_handle_obj* h = (struct tapp_handle_obj*)handle;
    if (!h) return 0;
    tk_handle_unregister(h);
    free(h);
    return 0;
}

const char* TAPP_AMD_version(void) { return "tapp-mi355x 0.1.0 (gfx950)"; }
