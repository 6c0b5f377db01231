/* Executor (api/include/tapp/executor.h:9-18).  Here it is the device
 * context: HIP stream, device workspace and plan cache.  Its first field is
 * an int kind tag, as in reference_implementation/src/executor.c:9-18. */
#ifndef TAPP_EXECUTOR_H_
#define TAPP_EXECUTOR_H_

#include <stdint.h>

#include "util.h"
#include "error.h"

typedef intptr_t TAPP_executor;

TAPP_EXPORT TAPP_error TAPP_create_executor(TAPP_executor* exec);
TAPP_EXPORT TAPP_error TAPP_destroy_executor(TAPP_executor exec);

#endif /* TAPP_EXECUTOR_H_ */
