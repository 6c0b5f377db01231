/*
 * Tensor layout descriptors.  Behaviour of
 * reference_implementation/src/tensor.c:11-102: create copies extents and
 * strides and returns 14 for a negative nmode and 15 for a negative extent;
 * set_extents returns 15 for a negative extent; set_nmodes resizes the arrays
 * (their contents past the old length are unspecified until set).
 */
#include <stdlib.h>
#include <string.h>

#include "internal.h"

static int64_t* dup64(const int64_t* src, int n)
{
    int64_t* d = malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    if (d && n > 0) memcpy(d, src, sizeof(int64_t) * (size_t)n);
    return d;
}

TAPP_error TAPP_create_tensor_info(TAPP_tensor_info* info, TAPP_datatype type, int nmode,
                                   const int64_t* extents, const int64_t* strides)
{
    if (nmode < 0) return 14;
    for (int i = 0; i < nmode; i++)
        if (extents[i] < 0) return 15;
    struct tapp_info* t = malloc(sizeof *t);
    if (!t) return TAPP_AMD_ERR_NOMEM;
    t->type = type;
    t->nmode = nmode;
    t->extents = dup64(extents, nmode);
    t->strides = dup64(strides, nmode);
    if (!t->extents || !t->strides) {
        free(t->extents); free(t->strides); free(t);
        return TAPP_AMD_ERR_NOMEM;
    }
    *info = (TAPP_tensor_info)t;
    return 0;
}

TAPP_error TAPP_destroy_tensor_info(TAPP_tensor_info info)
{
    struct tapp_info* t = (struct tapp_info*)info;
    if (!t) return 0;
    free(t->extents);
    free(t->strides);
    free(t);
    return 0;
}

int TAPP_get_nmodes(TAPP_tensor_info info) { return ((struct tapp_info*)info)->nmode; }

TAPP_error TAPP_set_nmodes(TAPP_tensor_info info, int nmodes)
{
    if (nmodes < 0) return 14;
    struct tapp_info* t = (struct tapp_info*)info;
    size_t bytes = sizeof(int64_t) * (size_t)(nmodes > 0 ? nmodes : 1);
    int64_t* e = realloc(t->extents, bytes);
    if (!e) return TAPP_AMD_ERR_NOMEM;
    t->extents = e;
    int64_t* s = realloc(t->strides, bytes);
    if (!s) return TAPP_AMD_ERR_NOMEM;
    t->strides = s;
    for (int i = t->nmode; i < nmodes; i++) { t->extents[i] = 0; t->strides[i] = 0; }
    t->nmode = nmodes;
    return 0;
}

void TAPP_get_extents(TAPP_tensor_info info, int64_t* extents)
{
    struct tapp_info* t = (struct tapp_info*)info;
    if (t->nmode > 0) memcpy(extents, t->extents, sizeof(int64_t) * (size_t)t->nmode);
}

TAPP_error TAPP_set_extents(TAPP_tensor_info info, const int64_t* extents)
{
    struct tapp_info* t = (struct tapp_info*)info;
    for (int i = 0; i < t->nmode; i++)
        if (extents[i] < 0) return 15;
    if (t->nmode > 0) memcpy(t->extents, extents, sizeof(int64_t) * (size_t)t->nmode);
    return 0;
}

void TAPP_get_strides(TAPP_tensor_info info, int64_t* strides)
{
    struct tapp_info* t = (struct tapp_info*)info;
    if (t->nmode > 0) memcpy(strides, t->strides, sizeof(int64_t) * (size_t)t->nmode);
}

TAPP_error TAPP_set_strides(TAPP_tensor_info info, const int64_t* strides)
{
    struct tapp_info* t = (struct tapp_info*)info;
    if (t->nmode > 0) memcpy(t->strides, strides, sizeof(int64_t) * (size_t)t->nmode);
    return 0;
}
