/*
 * Error codes and messages.  TAPP_check_success and TAPP_explain_error keep
 * the reference's contract (error.c:10-76): success iff 0; the message for
 * codes 0-14 is the reference's text; maxlen == 0 returns the full length;
 * otherwise at most maxlen-1 characters plus a NUL are written and the count
 * written is returned.  Codes the reference returns but does not describe
 * (15 from tensor.c, 33 from product.c:878-885, 137 from product.c:463) and
 * this library's own codes get their own text instead of the reference's
 * uninitialised pointer (error.c:66-69).
 */
#include <string.h>

#include "internal.h"

bool TAPP_check_success(TAPP_error error) { return error == 0; }

static const char* message_for(TAPP_error e)
{
    switch (e) {
    case 0: return "Success.";
    case 1: return "The extents for the indices shared between tensor A and B does not match.";
    case 2: return "The extents for the indices shared between tensor A and D does not match.";
    case 3: return "The extents for the indices shared between tensor B and D does not match.";
    case 4: return "Tensor D has indices not shared with tensor A or B.";
    case 5: return "The tensors C and D have different amount of dimensions.";
    case 6: return "The indices of tensor C and D does not line up.";
    case 7: return "The extents for the indices shared between tensor C and D does not match.";
    case 8: return "Aliasing found within tensor D.";
    case 9: return "An idx in tensor A has two different extents.";
    case 10: return "An idx in tensor B has two different extents.";
    case 11: return "An idx in tensor D has two different extents.";
    case 12: return "C should not be NULL while beta is not zero.";
    case 13: return "Nmode can not be negative.";
    case 14: return "Extents can not be negative.";
    case 15: return "Extents can not be negative (tensor info).";
    case 33: return "Invalid or missing executor.";
    case 137: return "Self-check mismatch between executors.";
    case TAPP_AMD_ERR_UNSUPPORTED: return "Datatype or precision combination not supported.";
    case TAPP_AMD_ERR_DEVICE: return "HIP runtime error.";
    case TAPP_AMD_ERR_NOMEM: return "Out of memory.";
    case TAPP_AMD_ERR_INVALID: return "Invalid argument or handle.";
    case TAPP_AMD_ERR_KEY: return "Unknown attribute key.";
    case TAPP_AMD_ERR_BUDGET: return "Request exceeds the memory budget.";
    case TAPP_AMD_ERR_BUSY: return "Queue full.";
    case TAPP_AMD_ERR_SCRATCH: return "Product needs more device memory than the allocation-free scratch holds.";
    default: return "Unknown error code.";
    }
}

size_t TAPP_explain_error(TAPP_error error, size_t maxlen, char* message)
{
    const char* m = message_for(error);
    size_t len = strlen(m);
    if (maxlen == 0) return len;
    size_t n = maxlen - 1 < len ? maxlen - 1 : len;
    memcpy(message, m, n);
    message[n] = '\0';
    return n;
}
