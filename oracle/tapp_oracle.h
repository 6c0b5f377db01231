/*
 * Oracle for TAPP_execute_product: a plain-C, single-threaded restatement of
 * the reference's algorithm (reference_implementation/src/product.c).
 *
 * TEST INFRASTRUCTURE ONLY.  The GPU library's results are compared with
 * it: only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark load it.  The product library never links, loads or calls it.
 *
 * Symbols use the prefix orc_ so the oracle can sit in one process beside
 * both the product library and the compiled reference, which all export
 * TAPP_*.
 */
#ifndef TAPP_ORACLE_H_
#define TAPP_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One operand as the reference sees it inside TAPP_execute_product after
 * unpacking the plan (product.c:140-195): storage type, layout, labels and
 * element op. */
typedef struct orc_tensor {
    int type;               /* TAPP_datatype */
    int nmode;
    const int64_t* extents;
    const int64_t* strides; /* in elements, may be negative or zero */
    const int64_t* idx;     /* labels */
    int op;                 /* TAPP_element_op */
} orc_tensor;

/* Validation chain of product.c:197-206, same order, same codes.
 * exec_kind == NULL stands for a null executor handle; otherwise *exec_kind
 * is the int the executor handle points at. */
int orc_check(const orc_tensor* A, const orc_tensor* B, const orc_tensor* C,
              const orc_tensor* D, const void* beta, const void* C_data,
              const int* exec_kind);

/* Full restatement of TAPP_execute_product (product.c:130-465) for the
 * brute-force executor: validation, plan stage and loop nest.  Returns the
 * reference's error code, 0 on success, or ORC_UNSUPPORTED for a
 * type/precision combination the reference library does not define. */
int orc_execute(const orc_tensor* A, const orc_tensor* B, const orc_tensor* C,
                const orc_tensor* D, int prec,
                const void* alpha, const void* A_data, const void* B_data,
                const void* beta, const void* C_data, void* D_data,
                const int* exec_kind);

/* orc_execute with options:
 *   ORC_GENERIC: run the slot-precision loop (the one binary16 / bfloat16
 *     values need) for every precision; on binary32 / binary64 data it
 *     must give the same bits as the per-type loops, which tests check.
 *   ORC_GPU_PRODUCTS: the TAPP_F16F16_* / TAPP_BF16BF16_ACCUM_F32 and
 *     binary16-storage semantics of the GPU library, not of the reference:
 *     alpha, beta, C and the unary sums are rounded as the reference does
 *     (inputs rounded to the 16-bit format, sums rounded at each addition),
 *     but alpha * sum_A * sum_B is not rounded to 16 bits and the
 *     accumulator is binary32.  For products without unary labels this is
 *     the binary32 loop on inputs rounded to the 16-bit format. */
#define ORC_GENERIC 1
#define ORC_GPU_PRODUCTS 2
int orc_execute_ex(const orc_tensor* A, const orc_tensor* B, const orc_tensor* C,
                   const orc_tensor* D, int prec,
                   const void* alpha, const void* A_data, const void* B_data,
                   const void* beta, const void* C_data, void* D_data,
                   const int* exec_kind, int flags);

/* Same as orc_execute but computing every operation in binary64 (complex
 * binary64 when any operand is complex) whatever the precision, after
 * applying the input rounding `round_inputs` (0 none, 1 binary16, 2
 * bfloat16, round-to-nearest-even) to alpha, beta, A, B and C.  This is the
 * FP64 accuracy oracle for the low-precision and split forms. */
int orc_execute_f64(const orc_tensor* A, const orc_tensor* B, const orc_tensor* C,
                    const orc_tensor* D, int prec,
                    const void* alpha, const void* A_data, const void* B_data,
                    const void* beta, const void* C_data, void* D_data,
                    int round_inputs);

/* Index classification after compress_repeated_indices (product.c:265-310),
 * exposed so tests can compare it with the reference's own helpers. */
typedef struct orc_classes {
    int n_contracted, n_unary_A, n_unary_B, nmode_D;
    int64_t contracted[64], unary_A[64], unary_B[64];
    int64_t ext_contracted[64];
    int64_t free_strides_A[64], free_strides_B[64];
    int64_t contracted_strides_A[64], contracted_strides_B[64];
    int64_t unary_strides_A[64], unary_strides_B[64];
    int64_t strides_D[64], strides_C[64], extents_D[64];
} orc_classes;

int orc_classify(const orc_tensor* A, const orc_tensor* B, const orc_tensor* C,
                 const orc_tensor* D, orc_classes* out);

/* Rounding helpers (round-to-nearest-even), exposed for tests. */
float orc_round_bf16(float x);
float orc_round_f16(float x);

#define ORC_UNSUPPORTED 101

#ifdef __cplusplus
}
#endif

#endif /* TAPP_ORACLE_H_ */
