/* Tensor products (api/include/tapp/product.h:15-75):
 *   D <- op_D( alpha * sum op_A(A) op_B(B) + beta * op_C(C) )
 * with einsum-style int64 labels per mode. */
#ifndef TAPP_PRODUCT_H_
#define TAPP_PRODUCT_H_

#include <stdint.h>

#include "util.h"
#include "error.h"
#include "handle.h"
#include "executor.h"
#include "datatype.h"
#include "status.h"
#include "tensor.h"

typedef int TAPP_element_op;

enum
{
    TAPP_IDENTITY = 0,
    TAPP_CONJUGATE = 1,
};

#define TAPP_IN_PLACE NULL

typedef intptr_t TAPP_tensor_product;

/* Records the four layouts (by handle, as product.c:66-117 does), copies the
 * label arrays, and the precision.  Never fails. */
TAPP_EXPORT TAPP_error TAPP_create_tensor_product(TAPP_tensor_product* plan,
                                                  TAPP_handle handle,
                                                  TAPP_element_op op_A,
                                                  TAPP_tensor_info A,
                                                  const int64_t* idx_A,
                                                  TAPP_element_op op_B,
                                                  TAPP_tensor_info B,
                                                  const int64_t* idx_B,
                                                  TAPP_element_op op_C,
                                                  TAPP_tensor_info C,
                                                  const int64_t* idx_C,
                                                  TAPP_element_op op_D,
                                                  TAPP_tensor_info D,
                                                  const int64_t* idx_D,
                                                  TAPP_prectype prec);

TAPP_EXPORT TAPP_error TAPP_destroy_tensor_product(TAPP_tensor_product plan);

/* Validates (codes of product.c:197-222) and executes the product on the
 * GPU.  A, B, C, D may be host or device pointers. */
TAPP_EXPORT TAPP_error TAPP_execute_product(TAPP_tensor_product plan,
                                            TAPP_executor exec,
                                            TAPP_status* status,
                                            const void* alpha,
                                            const void* A,
                                            const void* B,
                                            const void* beta,
                                            const void* C,
                                            void* D);

/* The same plan over num_batches operand sets with one alpha and beta
 * (product.h:66-75; declared but not defined by the reference). */
TAPP_EXPORT TAPP_error TAPP_execute_batched_product(TAPP_tensor_product plan,
                                                    TAPP_executor exec,
                                                    TAPP_status* status,
                                                    int num_batches,
                                                    const void* alpha,
                                                    const void** A,
                                                    const void** B,
                                                    const void* beta,
                                                    const void** C,
                                                    void** D);

#endif /* TAPP_PRODUCT_H_ */
