/* Key/value attributes (api/include/tapp/attributes.h:9-18).  The reference
 * declares these but never defines them; this library defines them on
 * executor handles (keys in tapp_amd.h). */
#ifndef TAPP_ATTRIBUTES_H_
#define TAPP_ATTRIBUTES_H_

#include <stdint.h>

#include "util.h"
#include "error.h"

typedef intptr_t TAPP_attr;
typedef int TAPP_key;

TAPP_EXPORT TAPP_error TAPP_attr_set(TAPP_attr attr, TAPP_key key, void* value);
TAPP_EXPORT TAPP_error TAPP_attr_get(TAPP_attr attr, TAPP_key key, void** value);
TAPP_EXPORT TAPP_error TAPP_attr_clear(TAPP_attr attr, TAPP_key key);

#endif /* TAPP_ATTRIBUTES_H_ */
