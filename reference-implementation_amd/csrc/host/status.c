/*
 * Status objects (status.h:9-17).  TAPP_execute_product is synchronous and
 * writes no status, like the reference; TAPP_destroy_status therefore only
 * has to accept 0 and objects it may be handed.
 */
#include "internal.h"

TAPP_error TAPP_destroy_status(TAPP_status status)
{
    (void)status;
    return 0;
}
