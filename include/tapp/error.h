/* Error codes: an int, 0 is success (api/include/tapp/error.h:9-26). */
#ifndef TAPP_ERROR_H_
#define TAPP_ERROR_H_

#include <stddef.h>
#include <stdbool.h>

#include "util.h"

typedef int TAPP_error;

/* true iff error == 0 (reference_implementation/src/error.c:10-12). */
TAPP_EXPORT bool TAPP_check_success(TAPP_error error);

/* Copies at most maxlen-1 characters of the message for `error` into
 * `message` and NUL-terminates it; returns the number written.  With
 * maxlen == 0 nothing is written and the full message length is returned
 * (error.h:14-26, error.c:15-76). */
TAPP_EXPORT size_t TAPP_explain_error(TAPP_error error, size_t maxlen, char* message);

#endif /* TAPP_ERROR_H_ */
