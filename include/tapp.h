/* TAPP C API umbrella header (api/include/tapp.h:1-21). */
#ifndef TAPP_TAPP_H_
#define TAPP_TAPP_H_

#ifdef __cplusplus
extern "C" {
#endif

#include "tapp/error.h"
#include "tapp/attributes.h"
#include "tapp/datatype.h"
#include "tapp/handle.h"
#include "tapp/executor.h"
#include "tapp/status.h"
#include "tapp/tensor.h"
#include "tapp/product.h"

#ifdef __cplusplus
}
#endif

#endif /* TAPP_TAPP_H_ */
