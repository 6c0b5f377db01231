/* Status objects (api/include/tapp/status.h:9-17): created by execution
 * functions.  The reference declares but never defines TAPP_destroy_status. */
#ifndef TAPP_STATUS_H_
#define TAPP_STATUS_H_

#include <stdint.h>

#include "util.h"
#include "error.h"

typedef intptr_t TAPP_status;

TAPP_EXPORT TAPP_error TAPP_destroy_status(TAPP_status status);

#endif /* TAPP_STATUS_H_ */
