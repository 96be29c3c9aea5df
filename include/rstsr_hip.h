/*
 * rstsr_hip.h — C ABI of the MI355X (gfx950) device backend for RSTSR.
 *
 * Each entry point replaces one method of the rstsr-core device traits that
 * DeviceFaer / DeviceOpenBLAS implement; the citation beside it names the
 * reference trait (file:line, paths relative to the rstsr repository root).
 * All pointers are plain device or host addresses, sizes are in bytes or
 * elements as stated, layouts are (ndim, shape[], stride[], offset) in
 * ELEMENTS exactly as rstsr's Layout<IxD> (strides may be negative or zero).
 *
 * Every function returns 0 on success and a nonzero rstsr_status_t on failure;
 * rstsr_hip_last_error() returns the calling thread's last message. Work is
 * enqueued on the context's stream; functions that hand a value to the host
 * (*_all, memcpy_d2h) synchronise that stream before returning.
 */
#ifndef RSTSR_HIP_H
#define RSTSR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error kinds, numbered as rstsr-common/src/error.rs:15-48 variants used here. */
typedef enum {
    RSTSR_OK = 0,
    RSTSR_ERR_INVALID_LAYOUT = 1,
    RSTSR_ERR_INVALID_VALUE = 2,
    RSTSR_ERR_VALUE_OUT_OF_RANGE = 3,
    RSTSR_ERR_DEVICE_MISMATCH = 4,
    RSTSR_ERR_DEVICE = 5,
    RSTSR_ERR_MEMORY = 6,
    RSTSR_ERR_UNIMPLEMENTED = 7,
    RSTSR_ERR_INDEX = 8,
    RSTSR_ERR_RUNTIME = 9
} rstsr_status_t;

/* Element types (Rust: f32, f64, i32, i64, Complex<f32>, Complex<f64>, usize, bool). */
typedef enum {
    RSTSR_F32 = 0,
    RSTSR_F64 = 1,
    RSTSR_I32 = 2,
    RSTSR_I64 = 3,
    RSTSR_C64 = 4,
    RSTSR_C128 = 5,
    RSTSR_U64 = 6,
    RSTSR_BOOL = 7
} rstsr_dtype_t;

/* Binary operators (rstsr-core/src/operators/ops/op_ternary_arithmetic.rs:3-56; the device impls
 * of feature_rayon/auto_impl/op_ternary_arithmetic.rs:3-14), operands of one dtype:
 *  - ADD, SUB, MUL, DIV, REM: f32, f64, i32, i64, c64, c128. Integers wrap; integer / and %
 *    truncate toward zero; float % is fmod; complex / and % are num-complex 0.4's formulas.
 *    Rust panics on an integer zero divisor and on MIN / -1 (MIN % -1); the device gives 0,
 *    the wrapped MIN and 0 for those instead.
 *  - BITAND, BITOR, BITXOR: i32, i64, bool.  SHL, SHR: i32, i64, the shift amount taken
 *    modulo the bit width (Rust release semantics), SHR arithmetic.
 * Other (operator, dtype) pairs have no impl in the reference and fail with UnImplemented. */
typedef enum {
    RSTSR_OP_ADD = 0,
    RSTSR_OP_SUB = 1,
    RSTSR_OP_MUL = 2,
    RSTSR_OP_DIV = 3,
    RSTSR_OP_REM = 4,
    RSTSR_OP_BITAND = 5,
    RSTSR_OP_BITOR = 6,
    RSTSR_OP_BITXOR = 7,
    RSTSR_OP_SHL = 8,
    RSTSR_OP_SHR = 9
} rstsr_binop_t;

/* Reductions (rstsr-core/src/operators/reduction.rs:3-95). */
typedef enum {
    RSTSR_RED_SUM = 0,
    RSTSR_RED_MAX = 1,
    RSTSR_RED_MIN = 2,
    RSTSR_RED_ARGMAX = 3,
    RSTSR_RED_ARGMIN = 4,
    /* feature_rayon/auto_impl/reduction.rs: prod :131-165, mean :167-205, var :207-259,
     * std :261-315, l2_norm :317-354. var/std/l2_norm of complex arrays are real. */
    RSTSR_RED_PROD = 5,
    RSTSR_RED_MEAN = 6,
    RSTSR_RED_VAR = 7,
    RSTSR_RED_STD = 8,
    RSTSR_RED_L2NORM = 9,
    /* all / any of a bool array -> bool (auto_impl/reduction.rs:466-519); count_nonzero of any
     * dtype -> usize (:521-551). RSTSR_RED_SUM of a bool array is OpSumBoolAPI (:678-715): the
     * number of true elements as a usize. */
    RSTSR_RED_ALL = 10,
    RSTSR_RED_ANY = 11,
    RSTSR_RED_COUNT_NONZERO = 12
} rstsr_redop_t;

/* Unary elementwise operations: the 35 traits of rstsr-core/src/operators/ops/op_binary_common.rs:3-39
 * (OpAcosAPI ... OpSignAPI, in that order), then OpNegAPI and OpNotAPI
 * (operators/ops/op_binary_arithmetic.rs:90-109). Output dtypes follow the traits' TOut:
 *  - acos .. trunc except the classifiers, square and signbit: T::FloatType
 *    (rstsr-dtype-traits DTypeIntoFloatAPI: i32/i64 -> f64, floats and complex unchanged);
 *    ceil, expm1, floor, log1p, round and trunc need a real float (num::Float bound);
 *  - is_finite, is_inf, is_nan (floats, complex) and signbit (integers, floats): bool;
 *  - square, sign, neg: T; not: integers and bool;
 *  - abs, imag, real: ExtNum::AbsOut (complex -> its real type, otherwise T).
 * log1p has a trait but no implementation in the reference devices; it is provided here
 * for real floats. */
typedef enum {
    RSTSR_UN_ACOS = 0,
    RSTSR_UN_ACOSH = 1,
    RSTSR_UN_ASIN = 2,
    RSTSR_UN_ASINH = 3,
    RSTSR_UN_ATAN = 4,
    RSTSR_UN_ATANH = 5,
    RSTSR_UN_CEIL = 6,
    RSTSR_UN_CONJ = 7,
    RSTSR_UN_COS = 8,
    RSTSR_UN_COSH = 9,
    RSTSR_UN_EXP = 10,
    RSTSR_UN_EXPM1 = 11,
    RSTSR_UN_FLOOR = 12,
    RSTSR_UN_INV = 13,
    RSTSR_UN_IS_FINITE = 14,
    RSTSR_UN_IS_INF = 15,
    RSTSR_UN_IS_NAN = 16,
    RSTSR_UN_LOG = 17,
    RSTSR_UN_LOG1P = 18,
    RSTSR_UN_LOG2 = 19,
    RSTSR_UN_LOG10 = 20,
    RSTSR_UN_RECIPROCAL = 21,
    RSTSR_UN_ROUND = 22,
    RSTSR_UN_SIGNBIT = 23,
    RSTSR_UN_SIN = 24,
    RSTSR_UN_SINH = 25,
    RSTSR_UN_SQUARE = 26,
    RSTSR_UN_SQRT = 27,
    RSTSR_UN_TAN = 28,
    RSTSR_UN_TANH = 29,
    RSTSR_UN_TRUNC = 30,
    RSTSR_UN_ABS = 31,
    RSTSR_UN_IMAG = 32,
    RSTSR_UN_REAL = 33,
    RSTSR_UN_SIGN = 34,
    RSTSR_UN_NEG = 35,
    RSTSR_UN_NOT = 36
} rstsr_unop_t;

/* GEMM arithmetic: plain FP64 MFMA, or Ozaki-II split over INT8 MFMA with CRT (f64; c128 with
 * real alpha/beta as three real products when beta = 0 and C is one matrix, else four). */
typedef enum { RSTSR_GEMM_PLAIN = 0, RSTSR_GEMM_OZAKI_INT8 = 1 } rstsr_gemm_mode_t;

/* Layout<IxD>: rstsr-common/src/layout/layoutbase.rs:15-23 (shape: usize, stride: isize, offset: usize). */
typedef struct {
    size_t ndim;
    const size_t* shape;
    const ptrdiff_t* stride;
    size_t offset;
} rstsr_layout_t;

typedef struct rstsr_hip_ctx rstsr_hip_ctx_t;

/* ---- device (DeviceBaseAPI, rstsr-core/src/storage/device.rs:3-7) ---- */
/* Create a context bound to one GPU and one stream (DeviceFaer::new analogue, device_faer/device.rs:12-60). */
int rstsr_hip_ctx_create(int device_id, rstsr_hip_ctx_t** out);
int rstsr_hip_ctx_destroy(rstsr_hip_ctx_t* ctx);
/* same_device (storage/device.rs:4): 1 if both contexts drive the same GPU. */
int rstsr_hip_same_device(const rstsr_hip_ctx_t* a, const rstsr_hip_ctx_t* b);
/* Thread-local message for the last failing call on this thread. */
const char* rstsr_hip_last_error(void);
int rstsr_hip_synchronize(rstsr_hip_ctx_t* ctx);
/* hipStream_t of the context, for callers that enqueue their own work on it. */
int rstsr_hip_get_stream(rstsr_hip_ctx_t* ctx, void** stream);
/* No counterpart in the reference (a CPU device has no queue): order ctx's stream after the
 * work already queued on `other` (an event on other's stream, waited on by ctx's stream;
 * during a capture on ctx, a host wait for other's stream). Call it before ctx's work reads
 * or writes a buffer that another context on the same device last touched. */
int rstsr_hip_stream_wait(rstsr_hip_ctx_t* ctx, rstsr_hip_ctx_t* other);
int rstsr_hip_device_count(int* count);

/* ---- storage (DeviceCreationAnyAPI storage/creation.rs:3-34, DeviceStorageAPI storage/device.rs:23-57) ---- */
/* uninit_impl / empty_impl: nbytes of uninitialised device memory (a block freed earlier on
 * this context is reused when one of the same size class is cached). */
int rstsr_hip_malloc(rstsr_hip_ctx_t* ctx, size_t nbytes, void** ptr);
/* Drop of Raw; the library frees only memory this call names. A buffer freed on the context
 * that allocated it is kept in that context's cache for reuse (ordered by its stream), up to
 * RSTSR_HIP_CACHE_BYTES (default: a quarter of the device memory); a copy between contexts
 * finishes before rstsr_hip_memcpy_d2d returns, and with several contexts on one device a
 * free waits for the device. */
int rstsr_hip_free(rstsr_hip_ctx_t* ctx, void* ptr);
/* No counterpart in the reference (its host allocator has no cache): return every cached
 * block of the context to the runtime, e.g. before another library allocates device memory;
 * and the bytes the cache holds. */
int rstsr_hip_empty_cache(rstsr_hip_ctx_t* ctx);
int rstsr_hip_cached_bytes(rstsr_hip_ctx_t* ctx, size_t* bytes);
/* outof_cpu_vec / from_cpu_vec (host -> device). */
int rstsr_hip_memcpy_h2d(rstsr_hip_ctx_t* ctx, void* dst, const void* src, size_t nbytes);
/* to_cpu_vec / into_cpu_vec / get_index (device -> host, synchronous). */
int rstsr_hip_memcpy_d2h(rstsr_hip_ctx_t* ctx, void* dst, const void* src, size_t nbytes);
/* Host transfers above run pageable copies of 8 MiB and more through two pinned 32 MiB chunks
 * (the host copy of one overlaps the DMA of the other); pinned host memory goes by direct DMA.
 * The _mode variants (no counterpart in the reference; for measurement) force a path:
 * 0 automatic, 1 the HIP runtime's own copy, 2 the staged copy. */
int rstsr_hip_memcpy_h2d_mode(rstsr_hip_ctx_t* ctx, void* dst, const void* src, size_t nbytes, int mode);
int rstsr_hip_memcpy_d2h_mode(rstsr_hip_ctx_t* ctx, void* dst, const void* src, size_t nbytes, int mode);
/* Raw: Clone (device -> device, stream ordered). */
int rstsr_hip_memcpy_d2d(rstsr_hip_ctx_t* ctx, void* dst, const void* src, size_t nbytes);

/* ---- assignment (rstsr-core/src/operators/assignment.rs:5-53) ---- */
/* OpAssignArbitaryAPI::assign_arbitary(_uninit): same element COUNT, any shapes, copied in
 * logical row-major order (cpu_rayon/assignment.rs:14-93); dtype cast c <- a. */
int rstsr_hip_assign_arbitary(rstsr_hip_ctx_t* ctx, rstsr_dtype_t tc, void* c, const rstsr_layout_t* lc,
                              rstsr_dtype_t ta, const void* a, const rstsr_layout_t* la);
/* OpAssignAPI::assign: same shape (cpu_rayon/assignment.rs:95-171); dtype cast c <- a. */
int rstsr_hip_assign(rstsr_hip_ctx_t* ctx, rstsr_dtype_t tc, void* c, const rstsr_layout_t* lc, rstsr_dtype_t ta,
                     const void* a, const rstsr_layout_t* la);
/* OpAssignAPI::fill (cpu_rayon/assignment.rs:181-225); *value is one host element of dtype t. */
int rstsr_hip_fill(rstsr_hip_ctx_t* ctx, rstsr_dtype_t t, void* c, const rstsr_layout_t* lc, const void* value);

/* ---- elementwise (operators/ops/op_ternary_arithmetic.rs:3-56, op_binary_arithmetic.rs:16-43) ---- */
/* c = a (op) b; lc, la, lb already broadcast to one shape (tensor/operators/op_binary_arithmetic.rs:197-208). */
int rstsr_hip_op_mutc_refa_refb(rstsr_hip_ctx_t* ctx, rstsr_binop_t op, rstsr_dtype_t t, void* c,
                                const rstsr_layout_t* lc, const void* a, const rstsr_layout_t* la, const void* b,
                                const rstsr_layout_t* lb);
/* c = a (op) scalar b (host value). */
int rstsr_hip_op_mutc_refa_numb(rstsr_hip_ctx_t* ctx, rstsr_binop_t op, rstsr_dtype_t t, void* c,
                                const rstsr_layout_t* lc, const void* a, const rstsr_layout_t* la, const void* b);
/* c = scalar a (op) b. */
int rstsr_hip_op_mutc_numa_refb(rstsr_hip_ctx_t* ctx, rstsr_binop_t op, rstsr_dtype_t t, void* c,
                                const rstsr_layout_t* lc, const void* a, const void* b, const rstsr_layout_t* lb);
/* c = a (op) b with one operand complex and the other its real type: the reference's
 * OpAddAPI/OpSubAPI/OpMulAPI/OpDivAPI<TA, TB, TC> with TA != TB (auto_impl/op_ternary_arithmetic.rs:3-14
 * over num-complex 0.4's `Complex<T> op T` and `T op Complex<T>`): tc is c64 or c128, (ta, tb) is
 * (tc, its real type) or (the real type, tc). Either of la / lb may be NULL: that operand is one
 * host value (op_mutc_refa_numb / op_mutc_numa_refb). Layouts already broadcast to lc's shape. */
int rstsr_hip_op_mutc_mixed(rstsr_hip_ctx_t* ctx, rstsr_binop_t op, rstsr_dtype_t tc, void* c,
                            const rstsr_layout_t* lc, rstsr_dtype_t ta, const void* a, const rstsr_layout_t* la,
                            rstsr_dtype_t tb, const void* b, const rstsr_layout_t* lb);
/* c (op)= b  (OpAddAssignAPI etc.). */
int rstsr_hip_op_muta_refb(rstsr_hip_ctx_t* ctx, rstsr_binop_t op, rstsr_dtype_t t, void* c,
                           const rstsr_layout_t* lc, const void* b, const rstsr_layout_t* lb);
/* c (op)= scalar b. */
int rstsr_hip_op_muta_numb(rstsr_hip_ctx_t* ctx, rstsr_binop_t op, rstsr_dtype_t t, void* c,
                           const rstsr_layout_t* lc, const void* b);

/* ---- reductions (operators/reduction.rs:3-95; auto_impl/reduction.rs:7-464) ---- */
/* sum_all / max_all / min_all: one host value of dtype t (sum of f32 is accumulated in f64). */
int rstsr_hip_reduce_all(rstsr_hip_ctx_t* ctx, rstsr_redop_t op, rstsr_dtype_t t, const void* a,
                         const rstsr_layout_t* la, void* host_out);
/* argmax_all / argmin_all: raveled row-major index into la's shape (reduction.rs:577). */
int rstsr_hip_argreduce_all(rstsr_hip_ctx_t* ctx, rstsr_redop_t op, rstsr_dtype_t t, const void* a,
                            const rstsr_layout_t* la, uint64_t* host_out);
/* Output layout of a reduction over `axes` (reduce_axes_cpu_rayon, cpu_rayon/reduction.rs:153 =
 * layout_for_array_copy(rest, K)); shape_out/stride_out need room for la->ndim entries. */
int rstsr_hip_reduce_out_layout(const rstsr_layout_t* la, size_t naxes, const ptrdiff_t* axes, size_t* ndim_out,
                                size_t* shape_out, ptrdiff_t* stride_out);
/* sum_axes / max_axes / min_axes / ... into caller-allocated `out` laid out as `lo`; the output
 * element is T, except T::Real for var/std/l2_norm of complex, bool for all/any and u64 (usize)
 * for count_nonzero and the sum of a bool array. */
int rstsr_hip_reduce_axes(rstsr_hip_ctx_t* ctx, rstsr_redop_t op, rstsr_dtype_t t, const void* a,
                          const rstsr_layout_t* la, size_t naxes, const ptrdiff_t* axes, void* out,
                          const rstsr_layout_t* lo);
/* No counterpart in the reference: rstsr_hip_reduce_axes (same arguments and results) for
 * callers that cannot allocate device memory (code under hipGraph stream capture, tight
 * serving loops, harnesses that forbid hipMalloc). It calls no hipMalloc/hipFree: the
 * partials of split reductions live in a fixed 8 MiB device array per GPU, resolved when
 * the first context on that GPU is created; a plan needing more fails with
 * RSTSR_ERR_VALUE_OUT_OF_RANGE before anything is enqueued. Every call on one GPU shares
 * that array, so calls that overlap on one GPU (other threads, other contexts' streams)
 * race; calls ordered on one stream do not. */
int rstsr_hip_reduce_axes_noalloc(rstsr_hip_ctx_t* ctx, rstsr_redop_t op, rstsr_dtype_t t, const void* a,
                                  const rstsr_layout_t* la, size_t naxes, const ptrdiff_t* axes, void* out,
                                  const rstsr_layout_t* lo);
/* argmax_axes / argmin_axes: u64 indices raveled row-major over the reduced axes in the given
 * order (cpu_serial/reduction.rs:553-580), into `out` laid out as `lo`. */
int rstsr_hip_argreduce_axes(rstsr_hip_ctx_t* ctx, rstsr_redop_t op, rstsr_dtype_t t, const void* a,
                             const rstsr_layout_t* la, size_t naxes, const ptrdiff_t* axes, uint64_t* out,
                             const rstsr_layout_t* lo);
/* OpAllCloseAPI (operators/reduction.rs:73-94; isclose.rs:92-106): |a-b| <= atol + rtol*|b|, NaN never close. */
int rstsr_hip_allclose(rstsr_hip_ctx_t* ctx, rstsr_dtype_t t, const void* a, const rstsr_layout_t* la,
                       const void* b, const rstsr_layout_t* lb, double rtol, double atol, int equal_nan,
                       int* host_out);

/* ---- unary elementwise (op_binary_common.rs:41-55, op_binary_arithmetic.rs:95-109) ---- */
/* Output dtype of `op` on input dtype `ta` (the trait's TOut); RSTSR_ERR_UNIMPLEMENTED when the
 * reference has no implementation for that input type. */
int rstsr_hip_unary_out_dtype(rstsr_unop_t op, rstsr_dtype_t ta, rstsr_dtype_t* tc);
/* OpXxxAPI::op_muta_refb: c = op(a) elementwise, layouts of one shape (the tensor API passes
 * lc = layout_for_array_copy(la, K), tensor/operators/op_unary_common.rs:150). tc must equal
 * rstsr_hip_unary_out_dtype(op, ta). c may alias a only with an identical layout. */
int rstsr_hip_unary(rstsr_hip_ctx_t* ctx, rstsr_unop_t op, rstsr_dtype_t tc, void* c, const rstsr_layout_t* lc,
                    rstsr_dtype_t ta, const void* a, const rstsr_layout_t* la);
/* OpXxxAPI::op_muta: c = op(c) in place; only ops whose output dtype equals t (the owned-tensor
 * path of op_unary_common.rs:175-190). */
int rstsr_hip_unary_inplace(rstsr_hip_ctx_t* ctx, rstsr_unop_t op, rstsr_dtype_t t, void* c,
                            const rstsr_layout_t* lc);

/* ---- matmul (DeviceMatMulAPI, operators/matmul.rs:7-25; device_faer/matmul.rs:204-239) ---- */
/* c = alpha * a % b + beta * c with rstsr's row-major matmul rules 1-7 (layout/matmul.rs:9-29).
 * beta == 0 never reads c. alpha/beta point to host scalars of dtype t. mode selects the arithmetic
 * for f64/c128 (other dtypes ignore it). */
int rstsr_hip_matmul(rstsr_hip_ctx_t* ctx, rstsr_dtype_t t, void* c, const rstsr_layout_t* lc, const void* a,
                     const rstsr_layout_t* la, const void* b, const rstsr_layout_t* lb, const void* alpha,
                     const void* beta, rstsr_gemm_mode_t mode);
/* 2-D GEMM on strided matrices: C[m,n] = alpha*sum_k A[m,k]B[k,n] + beta*C, each operand given by
 * (row stride, col stride) in elements (GEMMDriverAPI, rstsr-blas-traits/src/blas3/gemm.rs:4-20). */
int rstsr_hip_gemm(rstsr_hip_ctx_t* ctx, rstsr_dtype_t t, rstsr_gemm_mode_t mode, size_t m, size_t n, size_t k,
                   const void* alpha, const void* a, ptrdiff_t rsa, ptrdiff_t csa, const void* b, ptrdiff_t rsb,
                   ptrdiff_t csb, const void* beta, void* c, ptrdiff_t rsc, ptrdiff_t csc);

/* One problem of a grouped GEMM (uneven shapes, each operand by row/column strides). */
typedef struct {
    size_t m, n, k;
    const void* a;
    ptrdiff_t rsa, csa;
    const void* b;
    ptrdiff_t rsb, csb;
    void* c;
    ptrdiff_t rsc, csc;
} rstsr_gemm_desc_t;
/* Grouped GEMM: C_g = alpha A_g B_g + beta C_g for count uneven problems whose outputs do not
 * overlap. Plain f64/c128 run as ONE persistent launch whose tile shape the autotuner picks;
 * the f64 split mode runs each of its stages (norms, residue split, one INT8 launch over all
 * moduli, CRT) as one launch over the whole group; other dtypes run one call per problem. No
 * trait in the reference takes a list of problems; the closest is DeviceFaer's batched loop
 * over matmul tasks (device_faer/matmul.rs:151-199). The descriptor array is read before the
 * call returns (the caller may free it); the products run on the context stream without the
 * host waiting. Not available during graph capture (UnImplemented). */
int rstsr_hip_gemm_grouped(rstsr_hip_ctx_t* ctx, rstsr_dtype_t t, rstsr_gemm_mode_t mode, size_t count,
                           const rstsr_gemm_desc_t* descs, const void* alpha, const void* beta);

/* ---- BLAS-style drivers (cblas flag values of rstsr-common/src/flags.rs:62-213: order 101 C / 102 F,
 * trans 111 N / 112 T / 113 C (conjugate transpose) / 114 CN (conjugate only), uplo 121 U / 122 L,
 * diag 131 N (non-unit) / 132 U (unit), side 141 L / 142 R). Dtypes f32, f64, c64, c128; mode as
 * for rstsr_hip_gemm. ---- */
/* GEMMDriverAPI::driver_gemm (rstsr-blas-traits/src/blas3/gemm.rs:4-20):
 * C (m x n) = alpha op(A) op(B) + beta C, leading dimensions as in BLAS (at least the stored row
 * length for order C, column length for order F). Conjugated complex operands are copied
 * conjugated into context scratch first. */
int rstsr_hip_driver_gemm(rstsr_hip_ctx_t* ctx, rstsr_dtype_t t, rstsr_gemm_mode_t mode, int order, int transa,
                          int transb, size_t m, size_t n, size_t k, const void* alpha, const void* a, size_t lda,
                          const void* b, size_t ldb, const void* beta, void* c, size_t ldc);
/* SYHEMMDriverAPI::driver_syhemm (blas3/syhemm.rs:4-19): C = alpha A B + beta C (side L, A m x m) or
 * alpha B A + beta C (side R, A n x n) with A symmetric (hermi 0) or Hermitian (hermi 1); only
 * the uplo triangle of A is read (and only the real part of a Hermitian diagonal). */
int rstsr_hip_driver_syhemm(rstsr_hip_ctx_t* ctx, rstsr_dtype_t t, rstsr_gemm_mode_t mode, int hermi, int order,
                            int side, int uplo, size_t m, size_t n, const void* alpha, const void* a, size_t lda,
                            const void* b, size_t ldb, const void* beta, void* c, size_t ldc);
/* TRSMDriverAPI::driver_trsm (blas3/trsm.rs:4-19): solves op(A) X = alpha B (side L, A m x m) or
 * X op(A) = alpha B (side R, A n x n) for triangular A, B (m x n) overwritten by X; only the uplo
 * triangle of A is read, and with diag U not its diagonal (taken as one). alpha = 0 sets B to
 * zero without reading it. The trailing updates of the blocked solve run on the GEMM of `mode`. */
int rstsr_hip_driver_trsm(rstsr_hip_ctx_t* ctx, rstsr_dtype_t t, rstsr_gemm_mode_t mode, int order, int side,
                          int uplo, int transa, int diag, size_t m, size_t n, const void* alpha, const void* a,
                          size_t lda, void* b, size_t ldb);

/* ---- hipGraph capture of the context stream (no counterpart in the reference) ----
 * Everything enqueued between begin and end is recorded, not run; entry points that hand
 * values to the host (*_all, memcpy_d2h, allclose) cannot be captured and fail. */
int rstsr_hip_graph_begin(rstsr_hip_ctx_t* ctx);
/* Ends capture and returns an executable graph (hipGraphExec_t) in *graph_exec. Buffers freed
 * during the capture (temporaries the graph still uses) are held by the graph. */
int rstsr_hip_graph_end(rstsr_hip_ctx_t* ctx, void** graph_exec);
int rstsr_hip_graph_launch(rstsr_hip_ctx_t* ctx, void* graph_exec);
/* Waits for running launches when the graph holds buffers, then frees them with the graph. */
int rstsr_hip_graph_destroy(void* graph_exec);

/* ---- timing on the context stream (no counterpart in the reference; used by benchmarks) ---- */
int rstsr_hip_event_record(rstsr_hip_ctx_t* ctx, int slot);
/* Milliseconds between two recorded slots (synchronises on the later one). */
int rstsr_hip_event_elapsed(rstsr_hip_ctx_t* ctx, int slot_begin, int slot_end, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* RSTSR_HIP_H */
