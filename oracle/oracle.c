/*
 * oracle.c -- CPU restatement of clBool's boolean SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The product library never links
 * or calls this file; it is the checker for the HIP kernels and the timed CPU
 * baseline.  Citations are file:line in the clBool source tree.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* small growable u32 vector                                                 */
typedef struct { uint32_t *d; uint64_t n, cap; } vec32;

static void v_push(vec32 *v, uint32_t x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 64;
        v->d = (uint32_t *)realloc(v->d, v->cap * sizeof(uint32_t));
    }
    v->d[v->n++] = x;
}

static orc_mat *mat_new(void) { return (orc_mat *)calloc(1, sizeof(orc_mat)); }

void orc_free(orc_mat *m) {
    if (!m) return;
    free(m->rpt);
    free(m->rows);
    free(m->cols);
    free(m);
}

void orc_free_pairs(orc_pairs *p) {
    if (!p) return;
    free(p->rows);
    free(p->cols);
    free(p);
}

/* ------------------------------------------------------------------------ */
/* FastRandom (xorshf96, 64-bit state as on LP64 Linux):
 * libs/utils/libutils/fast_random.h:8-32.                                   */
void orc_fr_reset(orc_fastrandom *r, uint64_t seed) {
    /* fast_random.h:15-19 */
    r->x = seed;
    r->y = 362436069ull;
    r->z = 521288629ull;
}

int32_t orc_fr_next(orc_fastrandom *r, int32_t min, int32_t max) {
    /* fast_random.h:22-32: three xorshifts of x, rotate (x, y, z), then
     * min + (unsigned)(z % ((unsigned long)max - min + 1)). */
    r->x ^= r->x << 16;
    r->x ^= r->x >> 5;
    r->x ^= r->x << 1;
    uint64_t t = r->x;
    r->x = r->y;
    r->y = r->z;
    r->z = t ^ r->x ^ r->y;
    uint64_t range = (uint64_t)(int64_t)max - (uint64_t)(int64_t)min + 1ull;
    uint32_t off = (uint32_t)(r->z % range);
    return (int32_t)((uint32_t)min + off);
}

void orc_fr_stream(uint64_t seed, int32_t min, int32_t max, uint32_t n, int32_t *out) {
    orc_fastrandom r;
    orc_fr_reset(&r, seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = orc_fr_next(&r, min, max);
}

/* ------------------------------------------------------------------------ */
/* LSD radix sort of 64-bit keys (used to sort (row, col) pairs packed as
 * row << 32 | col, which is the lexicographic std::pair order).            */
static void radix_sort_u64(uint64_t *keys, uint64_t n) {
    if (n < 2) return;
    uint64_t *tmp = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *src = keys, *dst = tmp;
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = pass * 16;
        uint64_t *cnt = (uint64_t *)calloc(65537, sizeof(uint64_t));
        for (uint64_t i = 0; i < n; ++i) cnt[((src[i] >> shift) & 0xFFFF) + 1]++;
        for (int b = 0; b < 65536; ++b) cnt[b + 1] += cnt[b];
        for (uint64_t i = 0; i < n; ++i) dst[cnt[(src[i] >> shift) & 0xFFFF]++] = src[i];
        free(cnt);
        uint64_t *s = src; src = dst; dst = s;
    }
    /* four passes: result is back in keys */
    free(tmp);
}

static void radix_sort_u32(uint32_t *keys, uint64_t n, uint32_t *tmp) {
    if (n < 2) return;
    if (n <= 32) {
        for (uint64_t i = 1; i < n; ++i) {
            uint32_t x = keys[i];
            uint64_t j = i;
            while (j > 0 && keys[j - 1] > x) { keys[j] = keys[j - 1]; --j; }
            keys[j] = x;
        }
        return;
    }
    uint32_t *src = keys, *dst = tmp;
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = pass * 8;
        uint64_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        for (uint64_t i = 0; i < n; ++i) cnt[((src[i] >> shift) & 0xFF) + 1]++;
        for (int b = 0; b < 256; ++b) cnt[b + 1] += cnt[b];
        for (uint64_t i = 0; i < n; ++i) dst[cnt[(src[i] >> shift) & 0xFF]++] = src[i];
        uint32_t *s = src; src = dst; dst = s;
    }
}

static orc_pairs *pairs_from_keys(const uint64_t *keys, uint64_t n) {
    orc_pairs *p = (orc_pairs *)calloc(1, sizeof(orc_pairs));
    p->n = n;
    p->rows = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    p->cols = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    for (uint64_t i = 0; i < n; ++i) {
        p->rows[i] = (uint32_t)(keys[i] >> 32);
        p->cols[i] = (uint32_t)keys[i];
    }
    return p;
}

static uint64_t unique_u64(uint64_t *k, uint64_t n) {
    if (n == 0) return 0;
    uint64_t w = 1;
    for (uint64_t i = 1; i < n; ++i)
        if (k[i] != k[w - 1]) k[w++] = k[i];
    return w;
}

/* generate_coo_pairs_cpu: coo_utils.cpp:204-218; fill_random_matrix :135-151 */
orc_pairs *orc_generate_coo_pairs(uint32_t pseudo_nnz, uint32_t max_size) {
    if (max_size == 0 && pseudo_nnz != 0) return NULL; /* :136-140 throws */
    uint64_t *keys = (uint64_t *)malloc(((uint64_t)pseudo_nnz + 1) * sizeof(uint64_t));
    orc_fastrandom r;
    orc_fr_reset(&r, pseudo_nnz); /* :146 FastRandom r(nnz) */
    for (uint32_t i = 0; i < pseudo_nnz; ++i) {
        uint32_t row = (uint32_t)orc_fr_next(&r, 0, (int32_t)(max_size - 1)); /* :148 */
        uint32_t col = (uint32_t)orc_fr_next(&r, 0, (int32_t)(max_size - 1)); /* :149 */
        keys[i] = ((uint64_t)row << 32) | col;
    }
    radix_sort_u64(keys, pseudo_nnz);               /* :213 std::sort */
    uint64_t n = unique_u64(keys, pseudo_nnz);       /* :215 unique    */
    orc_pairs *p = pairs_from_keys(keys, n);
    free(keys);
    return p;
}

/* generate_coo_cpu: coo_utils.cpp:220-239; fill_random_matrix_pairs :153-160 */
orc_pairs *orc_generate_coo(uint32_t pseudo_nnz, uint32_t max_size) {
    uint64_t *keys = (uint64_t *)malloc(((uint64_t)pseudo_nnz + 1) * sizeof(uint64_t));
    orc_fastrandom r;
    orc_fr_reset(&r, pseudo_nnz);
    for (uint32_t i = 0; i < pseudo_nnz; ++i) {
        uint32_t row = (uint32_t)orc_fr_next(&r, 0, (int32_t)max_size); /* inclusive, :157 */
        uint32_t col = (uint32_t)orc_fr_next(&r, 0, (int32_t)max_size); /* :158 */
        keys[i] = ((uint64_t)row << 32) | col;
    }
    radix_sort_u64(keys, pseudo_nnz);
    uint64_t n = unique_u64(keys, pseudo_nnz);
    orc_pairs *p = pairs_from_keys(keys, n);
    free(keys);
    return p;
}

/* coo_pairs_to_dcsr_cpu: coo_utils.cpp:241-266 */
orc_mat *orc_pairs_to_dcsr(const uint32_t *rows, const uint32_t *cols, uint64_t n) {
    orc_mat *m = mat_new();
    if (n == 0) return m; /* :242-244 empty matrix_dcsr_cpu */
    vec32 rpt = {0}, rc = {0};
    uint32_t cur = rows[0];
    v_push(&rc, cur);
    v_push(&rpt, 0);
    for (uint64_t i = 0; i < n; ++i) {
        if (rows[i] != cur) {
            cur = rows[i];
            v_push(&rc, cur);
            v_push(&rpt, (uint32_t)i);
        }
    }
    v_push(&rpt, (uint32_t)n);
    m->nzr = (uint32_t)rc.n;
    m->nnz = n;
    m->rpt = rpt.d;
    m->rows = rc.d;
    m->cols = (uint32_t *)malloc(n * sizeof(uint32_t));
    memcpy(m->cols, cols, n * sizeof(uint32_t));
    return m;
}

/* ------------------------------------------------------------------------ */
/* matrix_multiplication_cpu: coo_utils.cpp:470-514                          */
orc_mat *orc_spgemm_reference(const uint32_t *a_rpt, const uint32_t *a_rows, const uint32_t *a_cols, uint32_t a_nzr,
                              const uint32_t *b_rpt, const uint32_t *b_rows, const uint32_t *b_cols, uint32_t b_nzr) {
    vec32 c_cols = {0}, c_rpt = {0}, c_rows = {0}, cur = {0};
    uint32_t current_pointer = 0;
    uint32_t *tmp = NULL;
    uint64_t tmp_cap = 0;
    for (uint32_t i = 0; i < a_nzr; ++i) {
        uint32_t start = a_rpt[i], end = a_rpt[i + 1];
        int is_row = 0;
        for (uint32_t j = start; j < end; ++j) {
            /* :487 linear std::find over B's rows */
            uint32_t pos = b_nzr;
            for (uint32_t q = 0; q < b_nzr; ++q)
                if (b_rows[q] == a_cols[j]) { pos = q; break; }
            if (pos != b_nzr) {
                if (!is_row) { /* :489-493 row recorded on the first hit */
                    v_push(&c_rpt, current_pointer);
                    v_push(&c_rows, a_rows[i]);
                    is_row = 1;
                }
                for (uint32_t k = b_rpt[pos]; k < b_rpt[pos + 1]; ++k) v_push(&cur, b_cols[k]);
            }
        }
        if (cur.n == 0) continue; /* :502 */
        if (cur.n > tmp_cap) {
            tmp_cap = cur.n;
            tmp = (uint32_t *)realloc(tmp, tmp_cap * sizeof(uint32_t));
        }
        radix_sort_u32(cur.d, cur.n, tmp); /* :504 sort */
        uint64_t w = 1;                    /* :505 unique */
        for (uint64_t q = 1; q < cur.n; ++q)
            if (cur.d[q] != cur.d[w - 1]) cur.d[w++] = cur.d[q];
        for (uint64_t q = 0; q < w; ++q) v_push(&c_cols, cur.d[q]);
        current_pointer += (uint32_t)w;
        cur.n = 0;
    }
    v_push(&c_rpt, current_pointer); /* :512 */
    free(tmp);
    free(cur.d);
    orc_mat *m = mat_new();
    m->nzr = (uint32_t)c_rows.n;
    m->nnz = c_cols.n;
    m->rpt = c_rpt.d;
    m->rows = c_rows.d;
    m->cols = c_cols.d;
    return m;
}

/* Gustavson with a stamped dense marker.  Same output as the restatement
 * above for well-formed DCSR inputs (sorted unique rows, nonempty rows).    */
static uint32_t b_lookup(const uint32_t *b_rows, uint32_t b_nzr, const uint32_t *dense_map, uint64_t map_n,
                         uint32_t key) {
    if (dense_map) return key < map_n ? dense_map[key] : UINT32_MAX;
    uint32_t lo = 0, hi = b_nzr;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (b_rows[mid] < key) lo = mid + 1; else hi = mid;
    }
    return (lo < b_nzr && b_rows[lo] == key) ? lo : UINT32_MAX;
}

orc_mat *orc_spgemm_gustavson(const uint32_t *a_rpt, const uint32_t *a_rows, const uint32_t *a_cols, uint32_t a_nzr,
                              const uint32_t *b_rpt, const uint32_t *b_rows, const uint32_t *b_cols, uint32_t b_nzr,
                              uint32_t b_nrows, uint32_t b_ncols,
                              const uint32_t *row_sel, uint32_t n_sel, int threads) {
    (void)b_nrows;
    orc_mat *m = mat_new();
    uint32_t n_rows = row_sel ? n_sel : a_nzr;
    if (n_rows == 0 || b_nzr == 0) {
        m->rpt = (uint32_t *)calloc(1, sizeof(uint32_t));
        return m;
    }
    /* marker width: the largest stored column + 1 (robust to cols >= ncols) */
    uint32_t maxc = b_ncols ? b_ncols - 1 : 0;
    for (uint64_t q = 0; q < b_rpt[b_nzr]; ++q) if (b_cols[q] > maxc) maxc = b_cols[q];
    uint64_t width = (uint64_t)maxc + 1;
    uint64_t map_n = (uint64_t)b_rows[b_nzr - 1] + 1;
    uint32_t *dense_map = NULL;
    if (map_n <= (1ull << 27)) {
        dense_map = (uint32_t *)malloc(map_n * sizeof(uint32_t));
        memset(dense_map, 0xFF, map_n * sizeof(uint32_t));
        for (uint32_t q = 0; q < b_nzr; ++q) dense_map[b_rows[q]] = q;
    }
    uint64_t *cnt = (uint64_t *)calloc((uint64_t)n_rows + 1, sizeof(uint64_t));
    if (threads <= 0) threads = 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads)
#endif
    {
        uint32_t *mark = (uint32_t *)calloc(width, sizeof(uint32_t));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t s = 0; s < (int64_t)n_rows; ++s) {
            uint32_t i = row_sel ? row_sel[s] : (uint32_t)s;
            uint32_t stamp = (uint32_t)s + 1;
            uint64_t c = 0;
            for (uint32_t j = a_rpt[i]; j < a_rpt[i + 1]; ++j) {
                uint32_t pos = b_lookup(b_rows, b_nzr, dense_map, map_n, a_cols[j]);
                if (pos == UINT32_MAX) continue;
                for (uint32_t k = b_rpt[pos]; k < b_rpt[pos + 1]; ++k) {
                    uint32_t col = b_cols[k];
                    if (mark[col] != stamp) { mark[col] = stamp; ++c; }
                }
            }
            cnt[s] = c;
        }
        free(mark);
    }
    /* offsets */
    uint64_t nnz = 0, nzr = 0;
    for (uint32_t s = 0; s < n_rows; ++s) { uint64_t c = cnt[s]; cnt[s] = nnz; nnz += c; nzr += c ? 1 : 0; }
    cnt[n_rows] = nnz;
    m->nnz = nnz;
    m->nzr = (uint32_t)nzr;
    m->cols = (uint32_t *)malloc((nnz ? nnz : 1) * sizeof(uint32_t));
    m->rpt = (uint32_t *)malloc((nzr + 1) * sizeof(uint32_t));
    m->rows = (uint32_t *)malloc((nzr ? nzr : 1) * sizeof(uint32_t));
#ifdef _OPENMP
#pragma omp parallel num_threads(threads)
#endif
    {
        uint32_t *mark = (uint32_t *)calloc(width, sizeof(uint32_t));
        uint32_t *tmp = NULL;
        uint64_t tmp_cap = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t s = 0; s < (int64_t)n_rows; ++s) {
            uint64_t len = cnt[s + 1] - cnt[s];
            if (!len) continue;
            uint32_t i = row_sel ? row_sel[s] : (uint32_t)s;
            uint32_t stamp = (uint32_t)s + 1;
            uint32_t *out = m->cols + cnt[s];
            uint64_t c = 0;
            for (uint32_t j = a_rpt[i]; j < a_rpt[i + 1]; ++j) {
                uint32_t pos = b_lookup(b_rows, b_nzr, dense_map, map_n, a_cols[j]);
                if (pos == UINT32_MAX) continue;
                for (uint32_t k = b_rpt[pos]; k < b_rpt[pos + 1]; ++k) {
                    uint32_t col = b_cols[k];
                    if (mark[col] != stamp) { mark[col] = stamp; out[c++] = col; }
                }
            }
            if (len > tmp_cap) { tmp_cap = len; tmp = (uint32_t *)realloc(tmp, tmp_cap * sizeof(uint32_t)); }
            radix_sort_u32(out, len, tmp);
        }
        free(tmp);
        free(mark);
    }
    uint32_t w = 0;
    for (uint32_t s = 0; s < n_rows; ++s) {
        if (cnt[s + 1] == cnt[s]) continue;
        uint32_t i = row_sel ? row_sel[s] : s;
        m->rpt[w] = (uint32_t)cnt[s];
        m->rows[w] = a_rows[i];
        ++w;
    }
    m->rpt[w] = (uint32_t)nnz;
    free(cnt);
    free(dense_map);
    return m;
}

/* ------------------------------------------------------------------------ */
/* Row digests (checker for products too large to compare column by column;
 * no reference counterpart).  A row's digest is its column count and the
 * sum, modulo 2^64, of splitmix64(col) over its columns (Vigna's splitmix64
 * finaliser): equal column sets give equal digests in any order, and a
 * changed, missing or extra column changes the sum unless two 64-bit mixes
 * cancel.  Sortedness and uniqueness are checked separately on the GPU side
 * (orc_dcsr_row_digest).                                                     */
static inline uint64_t mix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Digests of the rows of C = A * B for stored rows [r0, r1) of A (Gustavson
 * with a stamped dense marker, as orc_spgemm_gustavson, but nothing is
 * materialised): cnt[i - r0] = nnz of C's row a_rows[i], hash[i - r0] its
 * digest sum.  Rows with no output get (0, 0).                              */
void orc_spgemm_digest(const uint32_t *a_rpt, const uint32_t *a_cols, uint32_t r0, uint32_t r1,
                       const uint32_t *b_rpt, const uint32_t *b_rows, const uint32_t *b_cols, uint32_t b_nzr,
                       uint32_t b_ncols, int threads, uint64_t *cnt, uint64_t *hash) {
    if (r1 <= r0) return;
    if (b_nzr == 0) {
        memset(cnt, 0, (uint64_t)(r1 - r0) * sizeof(uint64_t));
        memset(hash, 0, (uint64_t)(r1 - r0) * sizeof(uint64_t));
        return;
    }
    uint32_t maxc = b_ncols ? b_ncols - 1 : 0;
    for (uint64_t q = 0; q < b_rpt[b_nzr]; ++q) if (b_cols[q] > maxc) maxc = b_cols[q];
    uint64_t width = (uint64_t)maxc + 1;
    uint64_t map_n = (uint64_t)b_rows[b_nzr - 1] + 1;
    uint32_t *dense_map = NULL;
    if (map_n <= (1ull << 27)) {
        dense_map = (uint32_t *)malloc(map_n * sizeof(uint32_t));
        memset(dense_map, 0xFF, map_n * sizeof(uint32_t));
        for (uint32_t q = 0; q < b_nzr; ++q) dense_map[b_rows[q]] = q;
    }
    if (threads <= 0) threads = 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads)
#endif
    {
        uint32_t *mark = (uint32_t *)calloc(width, sizeof(uint32_t));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (int64_t i = (int64_t)r0; i < (int64_t)r1; ++i) {
            uint32_t stamp = (uint32_t)(i - r0) + 1;
            uint64_t c = 0, h = 0;
            for (uint32_t j = a_rpt[i]; j < a_rpt[i + 1]; ++j) {
                uint32_t pos = b_lookup(b_rows, b_nzr, dense_map, map_n, a_cols[j]);
                if (pos == UINT32_MAX) continue;
                for (uint32_t k = b_rpt[pos]; k < b_rpt[pos + 1]; ++k) {
                    uint32_t col = b_cols[k];
                    if (mark[col] != stamp) { mark[col] = stamp; ++c; h += mix64(col); }
                }
            }
            cnt[i - r0] = c;
            hash[i - r0] = h;
        }
        free(mark);
    }
    free(dense_map);
}

/* Digests of stored rows [r0, r1) of a DCSR matrix whose columns
 * cols[rpt[r0] - base ...] are given (a chunk read back from the device):
 * hash[r - r0] as above.  Returns the number of rows whose columns are not
 * strictly ascending (0 for a canonical matrix).                             */
uint64_t orc_dcsr_row_digest(const uint32_t *rpt, const uint32_t *cols, uint64_t base, uint32_t r0, uint32_t r1,
                             int threads, uint64_t *hash) {
    uint64_t bad = 0;
    if (threads <= 0) threads = 1;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(dynamic, 64) reduction(+ : bad)
#endif
    for (int64_t r = (int64_t)r0; r < (int64_t)r1; ++r) {
        uint64_t h = 0;
        int ok = 1;
        for (uint64_t q = rpt[r]; q < rpt[r + 1]; ++q) {
            uint32_t c = cols[q - base];
            if (q > rpt[r] && cols[q - 1 - base] >= c) ok = 0;
            h += mix64(c);
        }
        hash[r - r0] = h;
        bad += ok ? 0 : 1;
    }
    return bad;
}

void orc_dense_bool_mm(const uint8_t *a, const uint8_t *b, uint8_t *c, uint32_t m, uint32_t k, uint32_t n) {
    for (uint32_t i = 0; i < m; ++i)
        for (uint32_t j = 0; j < n; ++j) {
            uint8_t v = 0;
            for (uint32_t q = 0; q < k && !v; ++q) v = (uint8_t)(a[(uint64_t)i * k + q] & b[(uint64_t)q * n + j]);
            c[(uint64_t)i * n + j] = v ? 1 : 0;
        }
}

/* get_workload: coo_utils.cpp:439-456 */
void orc_workload(const uint32_t *a_rpt, const uint32_t *a_cols, uint32_t a_nzr,
                  const uint32_t *b_rpt, const uint32_t *b_rows, uint32_t b_nzr, uint64_t *out) {
    for (uint32_t i = 0; i < a_nzr; ++i) {
        uint64_t w = 0;
        for (uint32_t j = a_rpt[i]; j < a_rpt[i + 1]; ++j) {
            uint32_t pos = b_lookup(b_rows, b_nzr, NULL, 0, a_cols[j]);
            if (pos != UINT32_MAX) w += b_rpt[pos + 1] - b_rpt[pos];
        }
        out[i] = w;
    }
}

/* matrix_addition_cpu: coo_utils.cpp:316-324 */
orc_pairs *orc_coo_add(const uint32_t *ar, const uint32_t *ac, uint64_t an,
                       const uint32_t *br, const uint32_t *bc, uint64_t bn) {
    uint64_t *k = (uint64_t *)malloc((an + bn + 1) * sizeof(uint64_t));
    uint64_t i = 0, j = 0, w = 0;
    /* std::merge: on ties the element of the first range goes first */
    while (i < an && j < bn) {
        uint64_t x = ((uint64_t)ar[i] << 32) | ac[i];
        uint64_t y = ((uint64_t)br[j] << 32) | bc[j];
        if (y < x) { k[w++] = y; ++j; } else { k[w++] = x; ++i; }
    }
    while (i < an) { k[w++] = ((uint64_t)ar[i] << 32) | ac[i]; ++i; }
    while (j < bn) { k[w++] = ((uint64_t)br[j] << 32) | bc[j]; ++j; }
    uint64_t n = unique_u64(k, w);
    orc_pairs *p = pairs_from_keys(k, n);
    free(k);
    return p;
}

/* kronecker_product_cpu: coo_utils.cpp:327-342 (uint32 arithmetic wraps) */
orc_pairs *orc_kron(const uint32_t *ar, const uint32_t *ac, uint64_t an,
                    const uint32_t *br, const uint32_t *bc, uint64_t bn,
                    uint32_t b_nrows, uint32_t b_ncols) {
    uint64_t n = an * bn;
    uint64_t *k = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    uint64_t w = 0;
    for (uint64_t i = 0; i < an; ++i)
        for (uint64_t j = 0; j < bn; ++j) {
            uint32_t r = ar[i] * b_nrows + br[j];
            uint32_t c = ac[i] * b_ncols + bc[j];
            k[w++] = ((uint64_t)r << 32) | c;
        }
    radix_sort_u64(k, n);
    orc_pairs *p = pairs_from_keys(k, n);
    free(k);
    return p;
}

uint64_t orc_exclusive_scan_u32(uint32_t *data, uint64_t n) {
    uint64_t acc = 0;
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t v = data[i];
        data[i] = (uint32_t)acc;
        acc += v;
    }
    return acc;
}

/* csr_cpu_from_pairs: matrices_conversions.cpp:158-177 */
orc_mat *orc_csr_from_pairs(const uint32_t *rows, const uint32_t *cols, uint64_t n, uint32_t nrows) {
    orc_mat *m = mat_new();
    m->nzr = nrows;
    m->nnz = n;
    m->rpt = (uint32_t *)calloc((uint64_t)nrows + 1, sizeof(uint32_t));
    m->cols = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    m->rpt[nrows] = (uint32_t)n;
    uint64_t ptr = 0, j = 0;
    for (uint32_t i = 0; i < nrows; ++i) {
        m->rpt[i] = (uint32_t)ptr;
        while (j < n && rows[j] == i) {
            m->cols[j] = cols[j];
            ++ptr;
            ++j;
        }
    }
    return m;
}

/* dcsr_to_csr: matrices_conversions.cpp:247-271 with dcsr_to_csr_set_size
 * (conversions.cl:8-18) and the scan. */
void orc_dcsr_to_csr_rpt(const uint32_t *rpt, const uint32_t *rows, uint32_t nzr, uint32_t nrows, uint32_t *csr_rpt) {
    memset(csr_rpt, 0, ((uint64_t)nrows + 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < nzr; ++i)
        if (rows[i] <= nrows) csr_rpt[rows[i]] = rpt[i + 1] - rpt[i];
    orc_exclusive_scan_u32(csr_rpt, (uint64_t)nrows + 1);
}

/* csr_to_dcsr: matrices_conversions.cpp:211-244 with
 * prepare_for_shift_empty_rows and set_positions_pointers_and_rows_csr
 * (set_positions.cl:69-85). */
orc_mat *orc_csr_to_dcsr(const uint32_t *csr_rpt, uint32_t nrows) {
    orc_mat *m = mat_new();
    uint32_t nzr = 0;
    for (uint32_t i = 0; i < nrows; ++i) nzr += csr_rpt[i] != csr_rpt[i + 1];
    m->nzr = nzr;
    m->nnz = csr_rpt[nrows];
    m->rpt = (uint32_t *)malloc(((uint64_t)nzr + 1) * sizeof(uint32_t));
    m->rows = (uint32_t *)malloc((nzr ? nzr : 1) * sizeof(uint32_t));
    uint32_t w = 0;
    for (uint32_t i = 0; i < nrows; ++i)
        if (csr_rpt[i] != csr_rpt[i + 1]) {
            m->rpt[w] = csr_rpt[i];
            m->rows[w] = i;
            ++w;
        }
    m->rpt[nzr] = csr_rpt[nrows];
    return m;
}

/* matrix_coo constructor: matrix_coo.cpp:17-47 */
orc_pairs *orc_coo_build(const uint32_t *rows, const uint32_t *cols, uint64_t n, int sorted, int no_duplicates) {
    uint64_t *k = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; ++i) k[i] = ((uint64_t)rows[i] << 32) | cols[i];
    if (!sorted) radix_sort_u64(k, n);           /* :40-42 coo::sort_arrays  */
    uint64_t w = no_duplicates ? n : unique_u64(k, n); /* :44-46 reduce_duplicates2 */
    orc_pairs *p = pairs_from_keys(k, w);
    free(k);
    return p;
}

/* ------------------------------------------------------------------------ */
/* Transpose as the reference's test builds it (test/src/operations.cpp:
 * 209-238): matrix_coo_cpu::transpose swaps rows and cols and sorts the pairs
 * (src/core/cpu_matrices.hpp:89-103), then coo_to_dcsr_cpu
 * (src/coo/coo_utils.cpp:268-294).                                          */
orc_mat *orc_transpose(const uint32_t *rpt, const uint32_t *rows, const uint32_t *cols, uint32_t nzr) {
    const uint64_t n = nzr ? rpt[nzr] : 0;
    if (n == 0) return mat_new();
    uint64_t *k = (uint64_t *)malloc(n * sizeof(uint64_t));
    for (uint32_t i = 0; i < nzr; ++i)
        for (uint32_t q = rpt[i]; q < rpt[i + 1]; ++q) k[q] = ((uint64_t)cols[q] << 32) | rows[i];
    radix_sort_u64(k, n);
    orc_pairs *p = pairs_from_keys(k, n);
    free(k);
    orc_mat *m = orc_pairs_to_dcsr(p->rows, p->cols, p->n);
    orc_free_pairs(p);
    return m;
}

/* submatrix_cpu: src/common/cpu_matrix_operations.cpp:6-52 (bounds taken in
 * 64 bits; the reference's i + nrows and j + ncols are uint32 sums). */
static uint32_t lower_bound32(const uint32_t *a, uint32_t n, uint64_t v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if ((uint64_t)a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

orc_mat *orc_submatrix(const uint32_t *rpt, const uint32_t *rows, const uint32_t *cols, uint32_t nzr, uint32_t i,
                       uint32_t j, uint32_t nrows, uint32_t ncols) {
    const uint32_t rb = lower_bound32(rows, nzr, i);
    const uint32_t re = lower_bound32(rows, nzr, (uint64_t)i + nrows);
    orc_mat *m = mat_new();
    if (rb == re) return m; /* :17-20 */
    vec32 r = {0}, w = {0}, c = {0};
    uint32_t count = 0;
    for (uint32_t idx = rb; idx < re; ++idx) {
        const uint32_t s = rpt[idx], len = rpt[idx + 1] - rpt[idx];
        const uint32_t lo = lower_bound32(cols + s, len, j);
        const uint32_t hi = lower_bound32(cols + s, len, (uint64_t)j + ncols);
        if (hi == lo) continue; /* :36 */
        for (uint32_t q = lo; q < hi; ++q) v_push(&c, cols[s + q] - j);
        v_push(&r, count);
        v_push(&w, rows[idx] - i);
        count += hi - lo;
    }
    v_push(&r, count); /* :46 */
    m->nzr = (uint32_t)w.n;
    m->nnz = c.n;
    m->rpt = r.d;
    m->rows = w.d;
    m->cols = c.d;
    return m;
}

/* reduce: src/common/cpu_matrix_operations.cpp:63-76 -- column 0 in every
 * stored row. */
orc_mat *orc_reduce(const uint32_t *rows, uint32_t nzr) {
    orc_mat *m = mat_new();
    m->nzr = nzr;
    m->nnz = nzr;
    m->rpt = (uint32_t *)malloc(((uint64_t)nzr + 1) * sizeof(uint32_t));
    m->rows = (uint32_t *)malloc((nzr ? nzr : 1) * sizeof(uint32_t));
    m->cols = (uint32_t *)calloc(nzr ? nzr : 1, sizeof(uint32_t));
    for (uint32_t q = 0; q <= nzr; ++q) m->rpt[q] = q;
    if (nzr) memcpy(m->rows, rows, nzr * sizeof(uint32_t));
    return m;
}

/* ------------------------------------------------------------------------ */
/* MLP restatement (not in the reference: the learned row-size predictor).  */
size_t orc_mlp_param_count(orc_mlp_dims d) {
    return (size_t)d.H1 * d.F + d.H1 + (size_t)d.H2 * d.H1 + d.H2 + d.H2 + 1;
}

static void mlp_offsets(orc_mlp_dims d, size_t *o) {
    o[0] = 0;                                 /* W1 */
    o[1] = o[0] + (size_t)d.H1 * d.F;         /* b1 */
    o[2] = o[1] + d.H1;                       /* W2 */
    o[3] = o[2] + (size_t)d.H2 * d.H1;        /* b2 */
    o[4] = o[3] + d.H2;                       /* w3 */
    o[5] = o[4] + d.H2;                       /* b3 */
}

void orc_mlp_forward(orc_mlp_dims d, const float *p, const float *x, uint32_t n, float *y) {
    size_t o[6];
    mlp_offsets(d, o);
    double h1[256], h2[256];
    for (uint32_t s = 0; s < n; ++s) {
        const float *xs = x + (size_t)s * d.F;
        for (uint32_t j = 0; j < d.H1; ++j) {
            double a = p[o[1] + j];
            for (uint32_t f = 0; f < d.F; ++f) a += (double)p[o[0] + (size_t)j * d.F + f] * xs[f];
            h1[j] = a > 0 ? a : 0;
        }
        for (uint32_t j = 0; j < d.H2; ++j) {
            double a = p[o[3] + j];
            for (uint32_t q = 0; q < d.H1; ++q) a += (double)p[o[2] + (size_t)j * d.H1 + q] * h1[q];
            h2[j] = a > 0 ? a : 0;
        }
        double out = p[o[5]];
        for (uint32_t j = 0; j < d.H2; ++j) out += (double)p[o[4] + j] * h2[j];
        y[s] = (float)out;
    }
}

double orc_mlp_loss_grad(orc_mlp_dims d, const float *p, const float *x, const float *t, uint32_t n, double *g) {
    size_t o[6];
    mlp_offsets(d, o);
    size_t np = orc_mlp_param_count(d);
    memset(g, 0, np * sizeof(double));
    double loss = 0;
    double a1[256], h1[256], a2[256], h2[256], d2[256], d1[256];
    for (uint32_t s = 0; s < n; ++s) {
        const float *xs = x + (size_t)s * d.F;
        for (uint32_t j = 0; j < d.H1; ++j) {
            double a = p[o[1] + j];
            for (uint32_t f = 0; f < d.F; ++f) a += (double)p[o[0] + (size_t)j * d.F + f] * xs[f];
            a1[j] = a;
            h1[j] = a > 0 ? a : 0;
        }
        for (uint32_t j = 0; j < d.H2; ++j) {
            double a = p[o[3] + j];
            for (uint32_t q = 0; q < d.H1; ++q) a += (double)p[o[2] + (size_t)j * d.H1 + q] * h1[q];
            a2[j] = a;
            h2[j] = a > 0 ? a : 0;
        }
        double y = p[o[5]];
        for (uint32_t j = 0; j < d.H2; ++j) y += (double)p[o[4] + j] * h2[j];
        double e = y - (double)t[s];
        loss += 0.5 * e * e;
        double dy = e / n;
        g[o[5]] += dy;
        for (uint32_t j = 0; j < d.H2; ++j) {
            g[o[4] + j] += dy * h2[j];
            d2[j] = a2[j] > 0 ? dy * p[o[4] + j] : 0;
        }
        for (uint32_t q = 0; q < d.H1; ++q) d1[q] = 0;
        for (uint32_t j = 0; j < d.H2; ++j) {
            g[o[3] + j] += d2[j];
            for (uint32_t q = 0; q < d.H1; ++q) {
                g[o[2] + (size_t)j * d.H1 + q] += d2[j] * h1[q];
                d1[q] += d2[j] * p[o[2] + (size_t)j * d.H1 + q];
            }
        }
        for (uint32_t q = 0; q < d.H1; ++q) {
            double dq = a1[q] > 0 ? d1[q] : 0;
            g[o[1] + q] += dq;
            for (uint32_t f = 0; f < d.F; ++f) g[o[0] + (size_t)q * d.F + f] += dq * xs[f];
        }
    }
    return loss / n;
}

/* Adam (Kingma & Ba 2015, Algorithm 1) with bias correction. */
void orc_adam_step(size_t n, float *p, float *m, float *v, const float *g, float lr, float b1, float b2, float eps,
                   uint32_t step) {
    double c1 = 1.0 - pow((double)b1, (double)step);
    double c2 = 1.0 - pow((double)b2, (double)step);
    for (size_t i = 0; i < n; ++i) {
        double mi = (double)b1 * m[i] + (1.0 - b1) * g[i];
        double vi = (double)b2 * v[i] + (1.0 - b2) * (double)g[i] * g[i];
        m[i] = (float)mi;
        v[i] = (float)vi;
        double mh = mi / c1, vh = vi / c2;
        p[i] = (float)((double)p[i] - (double)lr * mh / (sqrt(vh) + eps));
    }
}
