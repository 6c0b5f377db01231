/* Library handle (api/include/tapp/handle.h:9-24).  Here it owns the HIP
 * device binding and the resident-tensor table. */
#ifndef TAPP_HANDLE_H_
#define TAPP_HANDLE_H_

#include <stdint.h>

#include "util.h"
#include "error.h"

typedef intptr_t TAPP_handle;

TAPP_EXPORT TAPP_error TAPP_create_handle(TAPP_handle* handle);
TAPP_EXPORT TAPP_error TAPP_destroy_handle(TAPP_handle handle);

#endif /* TAPP_HANDLE_H_ */
