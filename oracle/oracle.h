/*
 * oracle.h -- CPU restatement of clBool's boolean SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker (or as the timed CPU
 * baseline), never as a fallback for the GPU path.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the clBool source tree).  All indices are uint32_t, as in
 * the reference (src/core/matrix_base.hpp:12-16).
 *
 * How each function is pinned (checked by tests/test_oracle.py):
 *   orc_fr_*                  against the reference's own fast_random.h, built
 *                             into oracle/_ref, and a committed fixture
 *   orc_generate_coo_pairs    through orc_fr_* plus std::sort/unique semantics
 *                             (checked against a numpy restatement)
 *   orc_spgemm_*              dense boolean brute force, I*A = A, A*0, and the
 *                             README example's hand-derived answer
 *   orc_workload              brute force sum over dense rows
 *   orc_coo_add, orc_kron     set semantics on dense masks
 *   orc_exclusive_scan        closed form (numpy cumsum)
 *   orc_csr_from_pairs        the reference's check_test_utils.cpp:21-39 vector
 *   orc_transpose             dense transpose; transpose twice is the identity
 *   orc_submatrix             dense slicing of the brute-force matrix
 *   orc_reduce                dense row-any
 *   orc_spgemm_digest,        a numpy restatement of splitmix64 over the
 *   orc_dcsr_row_digest       Gustavson product's rows; a dropped or changed
 *                             column and an unsorted row are caught
 *   orc_mlp_forward           a numpy evaluation of relu(relu(x W1' + b1)
 *                             W2' + b2) w3 + b3 with the documented layout (a
 *                             planted missing ReLU is caught)
 *   orc_mlp_loss_grad         finite differences of the loss
 *   orc_adam_step             closed-form first step
 */
#ifndef CLBOOL_ORACLE_H
#define CLBOOL_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- xorshf96 restatement: libs/utils/libutils/fast_random.h:8-40 ------ */
typedef struct { uint64_t x, y, z; } orc_fastrandom;
void orc_fr_reset(orc_fastrandom *r, uint64_t seed);
int32_t orc_fr_next(orc_fastrandom *r, int32_t min, int32_t max);
void orc_fr_stream(uint64_t seed, int32_t min, int32_t max, uint32_t n, int32_t *out);

/* ---- matrices owned by the oracle (free with orc_free) ----------------- */
typedef struct {
    uint32_t nzr;      /* number of stored rows (DCSR) or nrows (CSR)        */
    uint64_t nnz;
    uint32_t *rpt;     /* nzr + 1                                            */
    uint32_t *rows;    /* nzr (NULL for CSR)                                 */
    uint32_t *cols;    /* nnz                                                */
} orc_mat;
void orc_free(orc_mat *m);

typedef struct {
    uint64_t n;
    uint32_t *rows;
    uint32_t *cols;
} orc_pairs;
void orc_free_pairs(orc_pairs *p);

/* generate_coo_pairs_cpu: src/coo/coo_utils.cpp:204-218 with the fill of
 * :135-151 (FastRandom seeded by the requested count, draws in [0, max-1],
 * then sort and unique). */
orc_pairs *orc_generate_coo_pairs(uint32_t pseudo_nnz, uint32_t max_size);

/* generate_coo_cpu: src/coo/coo_utils.cpp:220-239 with the fill of :153-160
 * (draws in [0, max] inclusive). */
orc_pairs *orc_generate_coo(uint32_t pseudo_nnz, uint32_t max_size);

/* coo_pairs_to_dcsr_cpu: src/coo/coo_utils.cpp:241-266 */
orc_mat *orc_pairs_to_dcsr(const uint32_t *rows, const uint32_t *cols, uint64_t n);

/* matrix_multiplication_cpu: src/coo/coo_utils.cpp:470-514, restated
 * faithfully, including the linear std::find over B's rows (:487).  This is
 * the reference test suite's host SpGEMM and the timed CPU baseline. */
orc_mat *orc_spgemm_reference(const uint32_t *a_rpt, const uint32_t *a_rows, const uint32_t *a_cols, uint32_t a_nzr,
                              const uint32_t *b_rpt, const uint32_t *b_rows, const uint32_t *b_cols, uint32_t b_nzr);

/* The same product in canonical DCSR by Gustavson's algorithm with a dense
 * marker (OpenMP over rows).  Output equals orc_spgemm_reference for
 * well-formed inputs; used for parity at sizes the linear find cannot reach.
 * If row_sel != NULL only the A rows row_sel[0..n_sel) (indices into A's
 * stored rows, ascending) are multiplied. */
orc_mat *orc_spgemm_gustavson(const uint32_t *a_rpt, const uint32_t *a_rows, const uint32_t *a_cols, uint32_t a_nzr,
                              const uint32_t *b_rpt, const uint32_t *b_rows, const uint32_t *b_cols, uint32_t b_nzr,
                              uint32_t b_nrows, uint32_t b_ncols,
                              const uint32_t *row_sel, uint32_t n_sel, int threads);

/* Dense brute force: C[i][j] = OR_k A[i][k] AND B[k][j] over byte masks
 * (row-major, a is m x k, b is k x n, c is m x n).  Pins the SpGEMM. */
/* Row digests: count and sum of splitmix64(col) per row (checker only). */
void orc_spgemm_digest(const uint32_t *a_rpt, const uint32_t *a_cols, uint32_t r0, uint32_t r1,
                       const uint32_t *b_rpt, const uint32_t *b_rows, const uint32_t *b_cols, uint32_t b_nzr,
                       uint32_t b_ncols, int threads, uint64_t *cnt, uint64_t *hash);
uint64_t orc_dcsr_row_digest(const uint32_t *rpt, const uint32_t *cols, uint64_t base, uint32_t r0, uint32_t r1,
                             int threads, uint64_t *hash);
void orc_dense_bool_mm(const uint8_t *a, const uint8_t *b, uint8_t *c, uint32_t m, uint32_t k, uint32_t n);

/* get_workload: src/coo/coo_utils.cpp:439-456 (and count_workload.cl:30-53):
 * out[i] = sum over k in A row i of |B row k|, missing rows count 0. */
void orc_workload(const uint32_t *a_rpt, const uint32_t *a_cols, uint32_t a_nzr,
                  const uint32_t *b_rpt, const uint32_t *b_rows, uint32_t b_nzr, uint64_t *out);

/* matrix_addition_cpu: src/coo/coo_utils.cpp:316-324 (std::merge then
 * std::unique of (row, col) pairs). */
orc_pairs *orc_coo_add(const uint32_t *ar, const uint32_t *ac, uint64_t an,
                       const uint32_t *br, const uint32_t *bc, uint64_t bn);

/* kronecker_product_cpu: src/coo/coo_utils.cpp:327-342 (all pair products,
 * row = ar*b_nrows + br, col = ac*b_ncols + bc, then sort). */
orc_pairs *orc_kron(const uint32_t *ar, const uint32_t *ac, uint64_t an,
                    const uint32_t *br, const uint32_t *bc, uint64_t bn,
                    uint32_t b_nrows, uint32_t b_ncols);

/* Exclusive scan in place; returns the total.  test/src/procedures.cpp:24-31
 * and src/common/cl_operations.cpp:4-88 semantics. */
uint64_t orc_exclusive_scan_u32(uint32_t *data, uint64_t n);

/* csr_cpu_from_pairs: src/common/matrices_conversions.cpp:158-177 */
orc_mat *orc_csr_from_pairs(const uint32_t *rows, const uint32_t *cols, uint64_t n, uint32_t nrows);

/* dcsr_to_csr: src/common/matrices_conversions.cpp:247-271 (CSR rpt of
 * nrows + 1 built from DCSR rows and lengths; cols shared). */
/* matrix_coo_cpu::transpose + coo_to_dcsr_cpu (src/core/cpu_matrices.hpp:
 * 89-103, src/coo/coo_utils.cpp:268-294). */
orc_mat *orc_transpose(const uint32_t *rpt, const uint32_t *rows, const uint32_t *cols, uint32_t nzr);
/* submatrix_cpu: src/common/cpu_matrix_operations.cpp:6-52 */
orc_mat *orc_submatrix(const uint32_t *rpt, const uint32_t *rows, const uint32_t *cols, uint32_t nzr, uint32_t i,
                       uint32_t j, uint32_t nrows, uint32_t ncols);
/* reduce: src/common/cpu_matrix_operations.cpp:63-76 */
orc_mat *orc_reduce(const uint32_t *rows, uint32_t nzr);
void orc_dcsr_to_csr_rpt(const uint32_t *rpt, const uint32_t *rows, uint32_t nzr, uint32_t nrows, uint32_t *csr_rpt);

/* csr_to_dcsr: src/common/matrices_conversions.cpp:211-244 (drop empty rows).
 * Returns an orc_mat whose cols is NULL (cols are shared with the CSR). */
orc_mat *orc_csr_to_dcsr(const uint32_t *csr_rpt, uint32_t nrows);

/* matrix_coo(..., sorted, noDuplicates): src/core/matrix_coo.cpp:17-47 with
 * coo::sort_arrays (lexicographic (row, col) sort) and reduce_duplicates2
 * (:117-170, keep the first of equal pairs). */
orc_pairs *orc_coo_build(const uint32_t *rows, const uint32_t *cols, uint64_t n, int sorted, int no_duplicates);

/* ---- MLP restatement for the learned row-size predictor --------------- */
/* Layout (row-major fp32): W1[H1][F], b1[H1], W2[H2][H1], b2[H2], w3[H2], b3.
 * Activation ReLU; output y = w3.h2 + b3 (a prediction of log2(1 + nnz)).
 * Loss = mean over the batch of 0.5*(y - t)^2.  Computed in double. */
typedef struct { uint32_t F, H1, H2; } orc_mlp_dims;
size_t orc_mlp_param_count(orc_mlp_dims d);
void orc_mlp_forward(orc_mlp_dims d, const float *params, const float *x, uint32_t n, float *y);
double orc_mlp_loss_grad(orc_mlp_dims d, const float *params, const float *x, const float *t, uint32_t n,
                         double *grad);
void orc_adam_step(size_t n, float *p, float *m, float *v, const float *g, float lr, float b1, float b2, float eps,
                   uint32_t step);

#ifdef __cplusplus
}
#endif
#endif
