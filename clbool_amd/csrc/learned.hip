// learned.hip -- learned row-size predictor and the one-pass SpGEMM.
//
// Not a restatement of any reference code: clBool sizes every hash table from
// the symbolic count pass (hash.cpp:122-201).  Here a 5 -> 32 -> 32 -> 1 MLP
// predicts each row's output size from per-row features; the prediction
// sizes the row's LDS table and picks its bin, and a single numeric pass
// writes every row into a temporary buffer at predicted offsets.  A row
// with more distinct columns than its predicted capacity is re-run exactly
// (capacity min(P, ncols), which cannot overflow), so every row's result is
// exact whatever the prediction.  The final DCSR is compacted from the
// temporary buffers with exact offsets.  The only precedent in the reference
// is the unused overflow-and-record-failed-rows scheme of hash_large.cl:103-131.
//
// The MLP is trained on the GPU by the kernels below (fused forward and
// backward per block of samples with LDS-staged activations, a deterministic
// reduction of per-block gradients, Adam) on exact row sizes harvested from
// the plain symbolic pass.
#include "spgemm.hpp"

#include <chrono>
#include <cmath>
#include <random>

namespace clb {
namespace spg {

constexpr int NF = 5, NH1 = 32, NH2 = 32;
constexpr int OW1 = 0, OB1 = OW1 + NH1 * NF, OW2 = OB1 + NH1, OB2 = OW2 + NH2 * NH1, OW3 = OB2 + NH2,
              OB3 = OW3 + NH2;
constexpr int NPARAM = OB3 + 1;
static_assert(NPARAM == CLB_MLP_PARAMS, "parameter layout");
constexpr float FSCALE = 1.0f / 32.0f;  // features and target: log2(1 + x) / 32
constexpr int TRAIN_WG = 128;
constexpr u32 CAP_PAD = 8;

struct Learned {
    DevArray<float> params, m, v, grad, gpart, lpart;
    DevArray<float> X, Y;  // harvested samples
    DevArray<u32> idx;
    u64 n = 0, cap = 0;
    u32 step = 0;
    bool ready = false;
};

static Learned *state(clb_controls_s *ctl) {
    if (!ctl->learned) ctl->learned = new Learned();
    return static_cast<Learned *>(ctl->learned);
}

// ---------------------------------------------------------------------------
__device__ __forceinline__ float flog(u32 x) { return log2f(1.0f + (float)x) * FSCALE; }

// Feature vectors from the per-row statistics of k_row_stats: log2(1 + x)/32
// of products, A-row length, min and max B-row length, and column span (last
// column of any of the row's B rows minus the first, plus one).
__global__ void k_features(Mats g, u32 nrows, const u32 *__restrict__ prod, const u32 *__restrict__ mn,
                           const u32 *__restrict__ mx, const u32 *__restrict__ fc, const u32 *__restrict__ lc,
                           float *__restrict__ feat) {
    const u32 row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nrows) return;
    const u32 i = g.row0 + row;
    const u32 p = prod[row];
    const u32 span = p ? lc[row] - fc[row] + 1 : 0;
    float *f = feat + (u64)row * NF;
    f[0] = flog(p);
    f[1] = flog(g.a_rpt[i + 1] - g.a_rpt[i]);
    f[2] = flog(p ? mn[row] : 0);
    f[3] = flog(mx[row]);
    f[4] = flog(span);
}

// MLP forward of one sample with parameters in LDS.
__device__ __forceinline__ float mlp_fwd(const float *__restrict__ w, const float *x, float *a1, float *a2) {
#pragma unroll
    for (int j = 0; j < NH1; ++j) {
        float a = w[OB1 + j];
#pragma unroll
        for (int f = 0; f < NF; ++f) a = fmaf(w[OW1 + j * NF + f], x[f], a);
        a1[j] = a;
    }
#pragma unroll
    for (int j = 0; j < NH2; ++j) {
        float a = w[OB2 + j];
#pragma unroll
        for (int q = 0; q < NH1; ++q) a = fmaf(w[OW2 + j * NH1 + q], fmaxf(a1[q], 0.0f), a);
        a2[j] = a;
    }
    float y = w[OB3];
#pragma unroll
    for (int j = 0; j < NH2; ++j) y = fmaf(w[OW3 + j], fmaxf(a2[j], 0.0f), y);
    return y;
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ params, const float *__restrict__ x,
                                                     u32 n, float *__restrict__ y) {
    __shared__ float w[NPARAM];
    for (int p = threadIdx.x; p < NPARAM; p += 256) w[p] = params[p];
    __syncthreads();
    const u32 s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    float xs[NF], a1[NH1], a2[NH2];
#pragma unroll
    for (int f = 0; f < NF; ++f) xs[f] = x[(u64)s * NF + f];
    y[s] = mlp_fwd(w, xs, a1, a2);
}

// Capacity of every row from the prediction: ceil(pred * (1 + margin)) +
// CAP_PAD, clamped to the exact upper bound min(P, ncols); tiny rows (the
// wave kernel holds any of them) get the bound.  Rows the bin rule sends to
// a bitmap keep the predicted capacity when bitmap_pred is set (a bitmap row
// over its capacity is dumped whole or re-run exactly), else the bound.
__global__ __launch_bounds__(256) void k_predict_caps(const float *__restrict__ params,
                                                      const float *__restrict__ feat, const u32 *__restrict__ prod,
                                                      u32 n, BinCfg c, float margin, u32 bitmap_pred,
                                                      u32 *__restrict__ cap, float *__restrict__ pred) {
    __shared__ float w[NPARAM];
    for (int p = threadIdx.x; p < NPARAM; p += 256) w[p] = params[p];
    __syncthreads();
    const u32 s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const u32 P = prod[s];
    const u32 ub = P < c.ncols ? P : c.ncols;
    if (P == 0) {
        cap[s] = 0;
        if (pred) pred[s] = 0;
        return;
    }
    float xs[NF], a1[NH1], a2[NH2];
#pragma unroll
    for (int f = 0; f < NF; ++f) xs[f] = feat[(u64)s * NF + f];
    const float y = mlp_fwd(w, xs, a1, a2);
    float est = exp2f(fminf(fmaxf(y, 0.0f), 1.0f) * 32.0f) - 1.0f;
    if (pred) pred[s] = est;
    float want = ceilf(est * (1.0f + margin)) + (float)CAP_PAD;
    u32 k = want >= (float)ub ? ub : (want < 1.0f ? 1u : (u32)want);
    if (P <= c.tiny_max) k = ub;  // the wave kernel holds any tiny row
    if (!bitmap_pred && classify(P, k, c) >= BIN_BITMAP0) k = ub;
    cap[s] = k;
}

// Capacities of the one pass without a predictor: min(P, ncols) per row.
__global__ void k_bound_caps(const u32 *__restrict__ prod, u32 n, u32 ncols, u32 *__restrict__ cap) {
    const u32 s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const u32 P = prod[s];
    cap[s] = P < ncols ? P : ncols;
}

// Fused forward and backward over TRAIN_WG samples per block.  Writes the
// block's gradient partial sums to gpart[block][NPARAM] and its loss share.
__global__ __launch_bounds__(TRAIN_WG) void k_mlp_grad(const float *__restrict__ params,
                                                       const float *__restrict__ X, const float *__restrict__ Y,
                                                       const u32 *__restrict__ idx, u32 n, float inv_batch,
                                                       float *__restrict__ gpart, float *__restrict__ lpart) {
    __shared__ float w[NPARAM];
    __shared__ float sx[TRAIN_WG][NF];
    __shared__ float sh1[TRAIN_WG][NH1 + 1];
    __shared__ float sd1[TRAIN_WG][NH1 + 1];
    __shared__ float sh2[TRAIN_WG][NH2 + 1];
    __shared__ float sd2[TRAIN_WG][NH2 + 1];
    __shared__ float sdy[TRAIN_WG];
    __shared__ float red[TRAIN_WG / 64 + 1];
    const int t = threadIdx.x;
    for (int p = t; p < NPARAM; p += TRAIN_WG) w[p] = params[p];
    __syncthreads();
    const u32 s = blockIdx.x * TRAIN_WG + t;
    const bool valid = s < n;
    float xs[NF], a1[NH1], a2[NH2];
    const u32 src = valid ? (idx ? idx[s] : s) : 0;
#pragma unroll
    for (int f = 0; f < NF; ++f) xs[f] = valid ? X[(u64)src * NF + f] : 0.0f;
    const float y = mlp_fwd(w, xs, a1, a2);
    const float e = valid ? y - Y[src] : 0.0f;
    const float dy = e * inv_batch;
    // backward through the layers
    float d2[NH2];
#pragma unroll
    for (int j = 0; j < NH2; ++j) d2[j] = a2[j] > 0.0f ? dy * w[OW3 + j] : 0.0f;
#pragma unroll
    for (int q = 0; q < NH1; ++q) {
        float d = 0.0f;
#pragma unroll
        for (int j = 0; j < NH2; ++j) d = fmaf(d2[j], w[OW2 + j * NH1 + q], d);
        sd1[t][q] = a1[q] > 0.0f ? d : 0.0f;
        sh1[t][q] = fmaxf(a1[q], 0.0f);
    }
#pragma unroll
    for (int j = 0; j < NH2; ++j) {
        sd2[t][j] = d2[j];
        sh2[t][j] = fmaxf(a2[j], 0.0f);
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) sx[t][f] = xs[f];
    sdy[t] = dy;
    const float l = dev::block_sum<float, TRAIN_WG>(0.5f * e * e * inv_batch, red);
    if (t == 0) lpart[blockIdx.x] = l;
    __syncthreads();
    // parameter-parallel sums over the block's samples
    float *gp = gpart + (u64)blockIdx.x * NPARAM;
    for (int p = t; p < NPARAM; p += TRAIN_WG) {
        float acc = 0.0f;
        if (p < OB1) {
            const int j = p / NF, f = p % NF;
            for (int q = 0; q < TRAIN_WG; ++q) acc = fmaf(sd1[q][j], sx[q][f], acc);
        } else if (p < OW2) {
            const int j = p - OB1;
            for (int q = 0; q < TRAIN_WG; ++q) acc += sd1[q][j];
        } else if (p < OB2) {
            const int j = (p - OW2) / NH1, k = (p - OW2) % NH1;
            for (int q = 0; q < TRAIN_WG; ++q) acc = fmaf(sd2[q][j], sh1[q][k], acc);
        } else if (p < OW3) {
            const int j = p - OB2;
            for (int q = 0; q < TRAIN_WG; ++q) acc += sd2[q][j];
        } else if (p < OB3) {
            const int j = p - OW3;
            for (int q = 0; q < TRAIN_WG; ++q) acc = fmaf(sdy[q], sh2[q][j], acc);
        } else {
            for (int q = 0; q < TRAIN_WG; ++q) acc += sdy[q];
        }
        gp[p] = acc;
    }
}

// grad[p] = sum over blocks of gpart[b][p] (fixed order: deterministic).
__global__ void k_grad_reduce(const float *__restrict__ gpart, u32 nblocks, float *__restrict__ grad,
                              const float *__restrict__ lpart, float *__restrict__ loss) {
    const u32 p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < (u32)NPARAM) {
        float acc = 0.0f;
        for (u32 b = 0; b < nblocks; ++b) acc += gpart[(u64)b * NPARAM + p];
        grad[p] = acc;
    }
    if (p == 0 && loss) {
        float l = 0.0f;
        for (u32 b = 0; b < nblocks; ++b) l += lpart[b];
        *loss = l;
    }
}

// Adam (Kingma & Ba 2015, Algorithm 1) with bias correction.
__global__ void k_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, u32 n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_harvest(const float *__restrict__ feat, const u32 *__restrict__ prod,
                          const u32 *__restrict__ nnz_row, u32 nrows, u32 stride, float *__restrict__ X,
                          float *__restrict__ Y, u64 base) {
    const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
    const u64 r = (u64)k * stride;
    if (r >= nrows) return;
    for (int f = 0; f < NF; ++f) X[(base + k) * NF + f] = feat[r * NF + f];
    Y[base + k] = log2f(1.0f + (float)nnz_row[r]) * FSCALE;
    (void)prod;
}

// key2[i] = min(P, ncols) for rows whose one-pass output overflowed, else 0.
__global__ void k_failed_keys(const u32 *__restrict__ nnz_row, const u32 *__restrict__ prod, u32 n, u32 ncols,
                              u32 *__restrict__ key2) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u32 P = prod[i];
    key2[i] = nnz_row[i] == FAILED ? (P < ncols ? P : ncols) : 0u;
}

// Copy each row from the temporary buffers to its exact place and accumulate
// the prediction error |log2(1+pred) - log2(1+n)|.  A workgroup takes
// CR_ROWS consecutive rows: each wave copies one of them, except rows of at
// least CR_LONG columns, which the whole workgroup copies in turn; the error
// is reduced in LDS and added with one atomic per workgroup.
constexpr u32 CR_WG = 256, CR_ROWS = CR_WG / 64, CR_LONG = 2048;
__global__ __launch_bounds__(CR_WG) void k_compact_rows(const u32 *__restrict__ nnz_row, const u64 *__restrict__ off,
                                                        u32 nrows, const u32 *__restrict__ temp,
                                                        const u64 *__restrict__ toff, const u32 *__restrict__ side,
                                                        const u64 *__restrict__ soff, const u32 *__restrict__ key2,
                                                        const u32 *__restrict__ isdense, const float *__restrict__ pred,
                                                        u32 *__restrict__ cols, double *__restrict__ err) {
    __shared__ double red[CR_ROWS];
    const u32 wv = dev::wave_id(), l = (u32)dev::lane();
    const u32 r0 = blockIdx.x * CR_ROWS;
    auto copy = [&](const u32 *src, u32 *dst, u32 n, u32 t, u32 step) {
        for (; t + 3 * step < n; t += 4 * step) {
            const u32 a = src[t], b = src[t + step], c = src[t + 2 * step], d = src[t + 3 * step];
            dst[t] = a;
            dst[t + step] = b;
            dst[t + 2 * step] = c;
            dst[t + 3 * step] = d;
        }
        for (; t < n; t += step) dst[t] = src[t];
    };
    for (u32 j = 0; j < CR_ROWS; ++j) {
        const u32 r = r0 + j;
        if (r >= nrows) break;  // uniform
        const u32 n = nnz_row[r];
        if (!n || (isdense && isdense[r])) continue;  // dense bitmap rows are emitted separately
        const u32 *src = (key2 && key2[r]) ? side + soff[r] : temp + toff[r];
        u32 *dst = cols + off[r];
        if (n >= CR_LONG) copy(src, dst, n, threadIdx.x, CR_WG);
        else if (wv == j) copy(src, dst, n, l, 64u);
    }
    if (pred && err) {
        double e = 0.0;
        const u32 r = r0 + wv;
        if (l == 0 && r < nrows) {
            const u32 n = nnz_row[r];
            if (n) e = fabs(log2(1.0 + (double)pred[r]) - log2(1.0 + (double)n));
        }
        if (l == 0) red[wv] = e;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (u32 k = 0; k < CR_ROWS; ++k) t += red[k];
            if (t != 0.0) atomicAdd(err, t);
        }
    }
}

// The same copy cut by output position instead of by row: workgroup c writes
// output columns [c * CC_TILE, (c + 1) * CC_TILE), so every workgroup moves
// the same number of columns however the row sizes vary, and short rows do
// not leave most of a workgroup idle.  The rows that overlap the tile are
// found by binary search over the exact offsets (off[r] <= q < off[r + 1];
// empty rows have off[r] == off[r + 1] and are never the last row with
// off[r] <= q); their start and source are staged in LDS when at most
// CC_ROWS rows overlap, and each thread finds its column's row by a binary
// search there (or over off in HBM otherwise).
constexpr u32 CC_WG = 256, CC_ITEMS = 16, CC_TILE = CC_WG * CC_ITEMS, CC_ROWS = 512;
__device__ __forceinline__ u32 last_le(const u64 *__restrict__ a, u32 lo, u32 hi, u64 v) {
    // last index i in [lo, hi) with a[i] <= v (a[lo] <= v)
    while (hi - lo > 1) {
        const u32 mid = lo + (hi - lo) / 2;
        if (a[mid] <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}
__global__ __launch_bounds__(CC_WG) void k_compact_tiles(const u32 *__restrict__ nnz_row,
                                                         const u64 *__restrict__ off, u32 nrows,
                                                         const u32 *__restrict__ temp, const u64 *__restrict__ toff,
                                                         const u32 *__restrict__ side, const u64 *__restrict__ soff,
                                                         const u32 *__restrict__ key2,
                                                         const u32 *__restrict__ isdense, u32 *__restrict__ cols) {
    __shared__ u64 s_off[CC_ROWS + 1];
    __shared__ u64 s_src[CC_ROWS];  // bit 63: from side; ~0: skip
    __shared__ u32 s_r[2];
    const u64 total = off[nrows];
    const u64 q0 = (u64)blockIdx.x * CC_TILE;
    const u64 q1 = q0 + CC_TILE < total ? q0 + CC_TILE : total;
    if (threadIdx.x == 0) {
        s_r[0] = last_le(off, 0, nrows, q0);
        s_r[1] = last_le(off, 0, nrows, q1 - 1);
    }
    __syncthreads();
    const u32 r0 = s_r[0], r1 = s_r[1];
    const u32 m = r1 - r0 + 1;
    auto source = [&](u32 r) -> u64 {
        if (!nnz_row[r] || (isdense && isdense[r])) return ~0ull;
        return (key2 && key2[r]) ? (soff[r] | (1ull << 63)) : toff[r];
    };
    const bool staged = m <= CC_ROWS;
    if (staged) {
        for (u32 j = threadIdx.x; j < m; j += CC_WG) {
            s_off[j] = off[r0 + j];
            s_src[j] = source(r0 + j);
        }
        __syncthreads();
    }
    for (u32 k = 0; k < CC_ITEMS; ++k) {
        const u64 q = q0 + (u64)k * CC_WG + threadIdx.x;
        if (q >= q1) break;
        u64 start, src;
        if (staged) {
            u32 lo = 0, hi = m;
            while (hi - lo > 1) {
                const u32 mid = (lo + hi) / 2;
                if (s_off[mid] <= q) lo = mid;
                else hi = mid;
            }
            start = s_off[lo];
            src = s_src[lo];
        } else {
            const u32 r = last_le(off, r0, r1 + 1, q);
            start = off[r];
            src = source(r);
        }
        if (src == ~0ull) continue;
        const u64 d = q - start;
        cols[q] = (src >> 63) ? side[(src & ~(1ull << 63)) + d] : temp[src + d];
    }
}

// Mean-error numerator sum over rows of |log2(1 + pred) - log2(1 + n)|
// (rows with output), one atomic per workgroup.
__global__ __launch_bounds__(256) void k_pred_error(const u32 *__restrict__ nnz_row, const float *__restrict__ pred,
                                                    u32 nrows, double *__restrict__ err) {
    __shared__ double red[256 / 64 + 1];
    const u32 r = blockIdx.x * 256 + threadIdx.x;
    double e = 0.0;
    if (r < nrows) {
        const u32 n = nnz_row[r];
        if (n) e = fabs(log2(1.0 + (double)pred[r]) - log2(1.0 + (double)n));
    }
    const double s = dev::block_sum<double, 256>(e, red);
    if (threadIdx.x == 0 && s != 0.0) atomicAdd(err, s);
}

void compact_rows(hipStream_t st, u32 nrows, u64 nnz, const u32 *nnz_row, const u64 *off, const u32 *temp,
                  const u64 *toff, const u32 *side, const u64 *soff, const u32 *key2, const u32 *isdense,
                  const float *pred, u32 *cols, double *err, bool by_tiles) {
    if (!nrows) return;
    if (!by_tiles) {
        k_compact_rows<<<div_up(nrows, CR_ROWS), CR_WG, 0, st>>>(nnz_row, off, nrows, temp, toff, side, soff, key2,
                                                                  isdense, pred, cols, err);
        CLB_LAUNCHED("compact_rows");
        return;
    }
    if (pred && err) {
        k_pred_error<<<div_up(nrows, 256), 256, 0, st>>>(nnz_row, pred, nrows, err);
        CLB_LAUNCHED("pred_error");
    }
    if (nnz) {
        k_compact_tiles<<<div_up(nnz, CC_TILE), CC_WG, 0, st>>>(nnz_row, off, nrows, temp, toff, side, soff, key2,
                                                                isdense, cols);
        CLB_LAUNCHED("compact_tiles");
    }
}

// ---------------------------------------------------------------------------
static void prepare_features(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, u32 rb, u32 nrows, Setup &s,
                             DevArray<float> &feat) {
    // brange, Mats and cfg as in the plain path, but the count kernel is the
    // feature kernel (it also writes s.prod and the product total)
    prepare(ctl, a, b, rb, nrows, s, false);
    feat.reset(ctl, (u64)nrows * NF);
    DevArray<u32> mn(ctl, nrows), mx(ctl, nrows), fc(ctl, nrows), lc(ctl, nrows);
    fill_u32(ctl, s.prod.ptr, nrows, 0u);
    fill_u32(ctl, mn.ptr, nrows, 0xFFFFFFFFu);
    fill_u32(ctl, mx.ptr, nrows, 0u);
    fill_u32(ctl, fc.ptr, nrows, 0xFFFFFFFFu);
    fill_u32(ctl, lc.ptr, nrows, 0u);
    k_row_stats<true><<<div_up(a.nnz, RS_WG), RS_WG, 0, ctl->stream>>>(s.g, nrows, s.prod.ptr, s.scal.ptr, mn.ptr,
                                                                       mx.ptr, fc.ptr, lc.ptr);
    CLB_LAUNCHED("row_stats");
    k_features<<<div_up(nrows, 256), 256, 0, ctl->stream>>>(s.g, nrows, s.prod.ptr, mn.ptr, mx.ptr, fc.ptr, lc.ptr,
                                                             feat.ptr);
    CLB_LAUNCHED("features");
}

static BinCfg learned_bincfg(const clb_controls_s *ctl, u32 ncols) {
    BinCfg c = make_bincfg(ctl, ncols);
    // the predicted size decides, and rows with at least onepass_bitmap_min
    // products take a bitmap whatever their size (0: the size alone)
    const u32 pmin = ctl->tuning.reserved[CLB_TUNE_ONEPASS_BITMAP_MIN];
    c.bitmap_min = pmin ? pmin : 0xFFFFFFFFu;
    return c;
}

static void ensure_ready(clb_controls_s *ctl) {
    Learned *L = state(ctl);
    if (!L->ready)
        raise(INVALID_VALUE,
              "the learned row-size predictor has no parameters: train it (clb_predictor_train) or load them "
              "(clb_predictor_set_params)");
}

// Capacities (and predictions) of rows [0, s.nrows) of a prepared block.
static void learned_caps(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, Setup &s, bool bound_only,
                         u32 *cap, float *pred) {
    if (bound_only) {
        prepare(ctl, a, b, s.row0, s.nrows, s, true);
        k_bound_caps<<<div_up(s.nrows, 256), 256, 0, ctl->stream>>>(s.P, s.nrows, b.ncols, cap);
        CLB_LAUNCHED("bound_caps");
    } else {
        ensure_ready(ctl);
        DevArray<float> feat;
        prepare_features(ctl, a, b, s.row0, s.nrows, s, feat);
        const BinCfg c = learned_bincfg(ctl, b.ncols);
        const float margin = ctl->tuning.learned_margin_pct / 100.0f;
        const u32 bitmap_pred = ctl->tuning.reserved[CLB_TUNE_BITMAP_PRED_CAPS] ? 1u : 0u;
        k_predict_caps<<<div_up(s.nrows, 256), 256, 0, ctl->stream>>>(state(ctl)->params.ptr, feat.ptr, s.P,
                                                                       s.nrows, c, margin, bitmap_pred, cap, pred);
        CLB_LAUNCHED("predict_caps");
    }
}

void plan_sizes_learned(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, Setup &s, u32 *size,
                        float *pred, bool bound_only) {
    learned_caps(ctl, a, b, s, bound_only, size, bound_only ? nullptr : pred);
    s.cfg = learned_bincfg(ctl, b.ncols);
}

// One block of the one pass, alive until its columns are written.  Its last
// phase (the compaction of the temporary slots and the dense-row emission)
// runs on the controls' tail streams, so a pipelined plan can start the next
// block's passes before it ends; every buffer it reads stays here until then.
struct OnePass {
    clb_controls_s *ctl = nullptr;
    Setup s;
    LaunchTimer lt;
    DevArray<u32> cap_own, nnz_row, perm, key2, perm2, side, pos, temp;
    DevArray<float> pred_own;
    DevArray<u64> toff, soff, offs, gpre;
    DevArray<double> err;
    DumpBufs dump;
    BinHost h{};
    clb_dcsr out{};
    clb_spgemm_stats st{};
    hipEvent_t joined = nullptr, done = nullptr;
    // the end of the block's row kernels on each side stream (launch_bins
    // with a deferred join); post makes the controls' stream wait for them
    hipEvent_t bins[clb_controls_s::NSIDE] = {};
    bool tail = false, timing = false, replan = false;
    // kept from launch for post
    clb_dcsr a{}, b{};
    PlanRows prc{};
    bool planned = false, posted = false, empty = false;
    const float *pred = nullptr;  // predicted sizes (compaction's error statistic)
    std::chrono::steady_clock::time_point t0;
    ~OnePass() {
        for (hipEvent_t &e : bins)
            if (e) {
                (void)hipEventSynchronize(e);
                (void)hipEventDestroy(e);
            }
        if (tail)
            for (int k = 0; k < clb_controls_s::NTAIL; ++k) (void)hipStreamSynchronize(ctl->tail[k]);
        if (joined) (void)hipEventDestroy(joined);
        if (done) (void)hipEventDestroy(done);
        clb_dcsr_release(&out);  // nothing left when the result was handed over
    }
};

void OnePassDeleter::operator()(OnePass *p) const { delete p; }

OnePassPtr one_pass_launch(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, u32 rb, u32 re, bool timing,
                           bool bound_only, const PlanRows *pr) {
    OnePassPtr P(new OnePass());
    OnePass &p = *P;
    p.ctl = ctl;
    p.timing = timing;
    p.lt.on = timing;
    p.t0 = std::chrono::steady_clock::now();
    p.out.nrows = a.nrows;
    p.out.ncols = b.ncols;
    p.a = a;
    p.b = b;
    if (pr) {
        p.prc = *pr;
        p.planned = true;
    }
    clb_spgemm_stats &st = p.st;
    Setup &s = p.s;
    const u32 nrows = re - rb;
    p.nnz_row.reset(ctl, nrows);
    u32 *cap = nullptr;
    const float *pred = nullptr;
    if (pr) {
        // planned block: capacities (and predictions) of all rows were
        // computed at planning time
        prepare(ctl, a, b, rb, nrows, s, false, pr->shared);
        s.P = pr->prod;
        cap = pr->size;
        pred = pr->pred;
    } else {
        s.nrows = nrows;
        s.row0 = rb;
        p.cap_own.reset(ctl, nrows);
        if (!bound_only) p.pred_own.reset(ctl, nrows);
        learned_caps(ctl, a, b, s, bound_only, p.cap_own.ptr, p.pred_own.ptr);
        cap = p.cap_own.ptr;
        pred = p.pred_own.ptr;
    }
    s.cfg = learned_bincfg(ctl, b.ncols);
    BinBuf bb;
    bin_rows(ctl, s, cap, p.nnz_row.ptr, p.perm, bb);
    p.toff.reset(ctl, (u64)nrows + 1);
    exclusive_scan_u32_to_u64(ctl, cap, p.toff.ptr, nrows, s.scal.ptr + 4);
    read_bins(ctl, bb, p.h);
    u64 sc[8];
    readback(ctl, sc, s.scal.ptr, sizeof(sc));
    st.products = pr ? pr->products : sc[0];
    st.rows = nrows;
    st.n_bins = NBINS;
    st.capacity = sc[4];
    p.pred = pred;
    if (st.products == 0) {
        p.empty = true;
        return P;
    }
    p.temp.reset(ctl, sc[4] ? sc[4] : 1);
    u32 *fails = reinterpret_cast<u32 *>(s.scal.ptr + 5);
    Out o{p.nnz_row.ptr, p.toff.ptr, p.temp.ptr, cap, fails, Dump{}, ONE_PASS_PROBES};
    setup_dump(ctl, s, p.h, p.dump, o, true, p.perm.ptr, cap);
    ensure_splits(ctl, s, p.h);
    for (hipEvent_t &e : p.bins) CLB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming), EVENT_WAITING_ERROR);
    launch_bins(ctl, s, p.perm.ptr, p.h.off, ONE, o, &p.lt, CLB_KP_ONEPASS, p.bins);
    return P;
}

void one_pass_post(OnePass &p) {
    if (p.posted) return;
    p.posted = true;
    if (p.empty) return;
    clb_controls_s *ctl = p.ctl;
    clb_spgemm_stats &st = p.st;
    Setup &s = p.s;
    const clb_dcsr &a = p.a, &b = p.b;
    const PlanRows *pr = p.planned ? &p.prc : nullptr;
    const u32 nrows = s.nrows;
    const bool timing = p.timing;
    const float *pred = p.pred;
    join_deferred(ctl, p.bins);
    u32 *fails = reinterpret_cast<u32 *>(s.scal.ptr + 5);
    u32 nfail = 0;
    readback(ctl, &nfail, fails, 4);
    st.fallback_rows = nfail;
    if (timing) {  // per-bin statistics of the one-pass launch (for its roofline)
        u64 outs[NBINS];
        sum_by_bin(ctl, p.perm.ptr, p.h, nrows, p.nnz_row.ptr, outs);
        for (u32 k = 0; k < NBINS; ++k) {
            st.bin_rows[k] = p.h.off[k + 1] - p.h.off[k];
            st.bin_products[k] = p.h.products[k];
            st.bin_a_nnz[k] = p.h.a_nnz[k];
            st.bin_out_nnz[k] = outs[k];
        }
    }
    // exact re-run of the rows that overflowed their predicted capacity
    if (nfail) {
        p.key2.reset(ctl, nrows);
        k_failed_keys<<<div_up(nrows, 256), 256, 0, ctl->stream>>>(p.nnz_row.ptr, s.P, nrows, b.ncols,
                                                                    p.key2.ptr);
        CLB_LAUNCHED("failed_keys");
        const BinCfg keep = s.cfg;
        s.cfg = make_bincfg(ctl, b.ncols);  // exact sizing as in the plain symbolic pass
        BinBuf bb2;
        BinHost h2;
        bin_rows(ctl, s, p.key2.ptr, nullptr, p.perm2, bb2);
        p.soff.reset(ctl, (u64)nrows + 1);
        exclusive_scan_u32_to_u64(ctl, p.key2.ptr, p.soff.ptr, nrows, s.scal.ptr + 6);
        read_bins(ctl, bb2, h2);
        u64 side_total = 0;
        readback(ctl, &side_total, s.scal.ptr + 6, 8);
        p.side.reset(ctl, side_total ? side_total : 1);
        u32 *fails2 = reinterpret_cast<u32 *>(s.scal.ptr + 7);
        Out o2{p.nnz_row.ptr, p.soff.ptr, p.side.ptr, p.key2.ptr, fails2, Dump{}, 0};  // exact: full probing
        ensure_splits(ctl, s, h2);
        // on the controls' stream: the side streams may already hold the
        // next block's row kernels, which this block's re-run must not wait for
        launch_bins(ctl, s, p.perm2.ptr, h2.off, ONE, o2, &p.lt, CLB_KP_FALLBACK, nullptr, true);
        s.cfg = keep;
        u32 nfail2 = 0;
        readback(ctl, &nfail2, fails2, 4);
        if (nfail2) raise(RUN_KERNEL_FAILURE, "exact re-run overflowed its upper-bound capacity");
    }
    // exact offsets
    scan_sizes(ctl, s, p.nnz_row.ptr, p.offs, p.pos);
    u64 sc2[3];
    readback(ctl, sc2, s.scal.ptr, sizeof(sc2));
    const u64 nnz = sc2[1];
    const u32 nzr = static_cast<u32>(sc2[2]);
    if (pr && nnz > pr->bound) {
        // the block's result exceeds the plan's bound (rows grew past their
        // capacities): hand the exact row sizes back for a finer cut
        CLB_HIP(hipMemcpyAsync(pr->size, p.nnz_row.ptr, 4ull * nrows, hipMemcpyDeviceToDevice, ctl->stream),
                COPY_BUFFER_ERROR);
        sync(ctl);
        if (pr->replan) *pr->replan = true;
        p.replan = true;
        return;
    }
    if (nnz > 0xFFFFFFFFull)
        raise(INVALID_VALUE, "result has more nonzeros than uint32 indices allow; multiply by row blocks");
    st.nnz = nnz;
    st.nzr = nzr;
    if (nnz == 0) return;
    p.err.reset(ctl, 1);
    fill_u64(ctl, reinterpret_cast<u64 *>(p.err.ptr), 1, 0ull);
    clb_buffer_s *cols = buffer_new(ctl, 4ull * nnz);
    p.out.cols = cols;  // owned by p.out from here
    const u32 ndense = emit_prepare(ctl, p.dump, p.gpre, timing ? &st : nullptr);
    assemble(ctl, s, a, p.nnz_row.ptr, p.offs.ptr, p.pos.ptr, nnz, nzr, cols, p.out);
    // the tail: compaction of the temporary slots and emission of the dumped
    // rows write disjoint rows, side by side on the two tail streams, after
    // everything above on the controls' stream
    CLB_HIP(hipEventCreateWithFlags(&p.joined, hipEventDisableTiming), EVENT_WAITING_ERROR);
    CLB_HIP(hipEventCreateWithFlags(&p.done, hipEventDisableTiming), EVENT_WAITING_ERROR);
    CLB_HIP(hipEventRecord(ctl->ev_fork, ctl->stream), EVENT_WAITING_ERROR);
    p.tail = true;
    hipStream_t t0s = ctl->tail[0], t1s = ctl->tail[1];
    CLB_HIP(hipStreamWaitEvent(t0s, ctl->ev_fork, 0), EVENT_WAITING_ERROR);
    CLB_HIP(hipStreamWaitEvent(t1s, ctl->ev_fork, 0), EVENT_WAITING_ERROR);
    p.lt.begin(t0s);
    compact_rows(t0s, nrows, nnz, p.nnz_row.ptr, p.offs.ptr, p.temp.ptr, p.toff.ptr, p.side.ptr, p.soff.ptr,
                 nfail ? p.key2.ptr : nullptr, p.dump.on ? p.dump.isdense.ptr : nullptr, pred, u32p(cols),
                 p.err.ptr, ctl->tuning.reserved[CLB_TUNE_COMPACT_TILES] != 0);
    p.lt.end((int)CLB_KF_COMPACT, t0s);
    emit_launch(t1s, ndense, p.perm.ptr, p.h, p.dump, p.gpre.ptr, p.offs.ptr, u32p(cols), &p.lt, CLB_KP_ONEPASS);
    CLB_HIP(hipEventRecord(p.joined, t0s), EVENT_WAITING_ERROR);
    CLB_HIP(hipStreamWaitEvent(t1s, p.joined, 0), EVENT_WAITING_ERROR);
    CLB_HIP(hipEventRecord(p.done, t1s), EVENT_WAITING_ERROR);
}

OnePassPtr one_pass_start(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, u32 rb, u32 re, bool timing,
                          bool bound_only, const PlanRows *pr) {
    OnePassPtr p = one_pass_launch(ctl, a, b, rb, re, timing, bound_only, pr);
    one_pass_post(*p);
    return p;
}

bool one_pass_replan(const OnePass &p) { return p.replan; }

u64 one_pass_bytes(const OnePass &p) {
    u64 b = p.temp.cap + p.dump.store.cap + p.dump.cnt.cap + p.side.cap + p.toff.cap + p.offs.cap +
            p.soff.cap + p.nnz_row.cap + p.perm.cap + p.pos.cap;
    if (p.out.cols) b += p.out.cols->cap;
    return b;
}

void one_pass_finish(OnePass &p, clb_dcsr &out, clb_spgemm_stats &st) {
    clb_controls_s *ctl = p.ctl;
    one_pass_post(p);
    if (p.done) CLB_HIP(hipEventSynchronize(p.done), EVENT_WAITING_ERROR);
    p.tail = false;
    if (p.err.ptr && p.st.nzr) {
        double e = 0;
        readback(ctl, &e, p.err.ptr, 8);
        p.st.pred_abs_log2_err = e / p.st.nzr;
    }
    p.lt.collect(p.st);
    if (p.timing)
        p.st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - p.t0).count();
    out = p.out;
    p.out = clb_dcsr{};  // handed over
    st = p.st;
}

void spgemm_learned(clb_controls_s *ctl, const clb_dcsr &a, const clb_dcsr &b, u32 rb, u32 re, bool timing,
                    clb_dcsr &out, clb_spgemm_stats &st, bool bound_only, const PlanRows *pr) {
    OnePassPtr p = one_pass_start(ctl, a, b, rb, re, timing, bound_only, pr);
    clb_dcsr r{};
    one_pass_finish(*p, r, st);
    out = r;
}

// ---------------------------------------------------------------------------
static void init_params(clb_controls_s *ctl, Learned *L, u32 seed) {
    std::mt19937 rng(seed);
    std::vector<float> p(NPARAM, 0.0f);
    auto he = [&](int fan_in) { return std::normal_distribution<float>(0.0f, std::sqrt(2.0f / fan_in)); };
    auto d1 = he(NF), d2 = he(NH1), d3 = he(NH2);
    for (int k = 0; k < NH1 * NF; ++k) p[OW1 + k] = d1(rng);
    for (int k = 0; k < NH2 * NH1; ++k) p[OW2 + k] = d2(rng);
    for (int k = 0; k < NH2; ++k) p[OW3 + k] = d3(rng) * 0.5f;
    for (int k = 0; k < NH1; ++k) p[OB1 + k] = 0.01f;
    for (int k = 0; k < NH2; ++k) p[OB2 + k] = 0.01f;
    p[OB3] = 0.3f;
    L->params.reset(ctl, NPARAM);
    L->m.reset(ctl, NPARAM);
    L->v.reset(ctl, NPARAM);
    L->grad.reset(ctl, NPARAM);
    CLB_HIP(hipMemcpyAsync(L->params.ptr, p.data(), NPARAM * 4, hipMemcpyHostToDevice, ctl->stream),
            WRITE_BUFFER_ERROR);
    fill_u32(ctl, reinterpret_cast<u32 *>(L->m.ptr), NPARAM, 0);
    fill_u32(ctl, reinterpret_cast<u32 *>(L->v.ptr), NPARAM, 0);
    L->step = 0;
    sync(ctl);
}

// One Adam step on a batch [b0, b0 + nb) of the (shuffled) samples.
static void train_step(clb_controls_s *ctl, Learned *L, const float *X, const float *Y, const u32 *idx, u32 nb,
                       float lr, float *loss_dev) {
    const u32 blocks = div_up(nb, TRAIN_WG);
    if (L->gpart.n < (u64)blocks * NPARAM) L->gpart.reset(ctl, (u64)blocks * NPARAM);
    if (L->lpart.n < blocks) L->lpart.reset(ctl, blocks);
    k_mlp_grad<<<blocks, TRAIN_WG, 0, ctl->stream>>>(L->params.ptr, X, Y, idx, nb, 1.0f / nb, L->gpart.ptr,
                                                     L->lpart.ptr);
    CLB_LAUNCHED("mlp_grad");
    k_grad_reduce<<<div_up(NPARAM, 256), 256, 0, ctl->stream>>>(L->gpart.ptr, blocks, L->grad.ptr, L->lpart.ptr,
                                                                 loss_dev);
    CLB_LAUNCHED("grad_reduce");
    L->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)L->step), c2 = 1.0f - std::pow(b2, (float)L->step);
    k_adam<<<div_up(NPARAM, 256), 256, 0, ctl->stream>>>(L->params.ptr, L->m.ptr, L->v.ptr, L->grad.ptr, NPARAM, lr,
                                                         b1, b2, 1e-8f, c1, c2);
    CLB_LAUNCHED("adam");
}

}  // namespace spg

void learned_destroy_state(clb_controls_s *ctl) {
    if (ctl->learned) {
        delete static_cast<spg::Learned *>(ctl->learned);
        ctl->learned = nullptr;
    }
}

}  // namespace clb

using namespace clb;
using namespace clb::spg;

extern "C" {

clb_status clb_predictor_reset(clb_controls ctl, uint32_t seed) {
    return run_guarded([&] {
        Learned *L = state(ctl);
        init_params(ctl, L, seed);
        L->n = 0;
        L->ready = true;
    });
}

clb_status clb_predictor_harvest(clb_controls ctl, const clb_dcsr *a, const clb_dcsr *b, uint32_t max_rows,
                                 uint64_t *total_samples) {
    return run_guarded([&] {
        check_dims(*a, *b);
        Learned *L = state(ctl);
        if (!L->ready) {
            init_params(ctl, L, 1);
            L->ready = true;
        }
        if (a->nnz && b->nnz && a->nzr && max_rows) {
            const u32 nrows = a->nzr;
            Setup s;
            DevArray<float> feat;
            prepare_features(ctl, *a, *b, 0, nrows, s, feat);
            DevArray<u32> nnz_row(ctl, nrows), perm;
            BinBuf bb;
            BinHost h;
            bin_rows(ctl, s, nullptr, nnz_row.ptr, perm, bb);
            read_bins(ctl, bb, h);
            Out o{nnz_row.ptr, nullptr, nullptr, nullptr, reinterpret_cast<u32 *>(s.scal.ptr + 3), Dump{}, 0};
            ensure_splits(ctl, s, h);
            launch_bins(ctl, s, perm.ptr, h.off, SYM, o, nullptr, -1);
            const u32 stride = nrows > max_rows ? div_up(nrows, max_rows) : 1;
            const u32 take = div_up(nrows, stride);
            if (L->n + take > L->cap) {  // grow the sample store
                const u64 ncap = std::max<u64>(L->cap * 2, L->n + take + 4096);
                DevArray<float> X2(ctl, ncap * NF), Y2(ctl, ncap);
                if (L->n) {
                    CLB_HIP(hipMemcpyAsync(X2.ptr, L->X.ptr, L->n * NF * 4, hipMemcpyDeviceToDevice, ctl->stream),
                            COPY_BUFFER_ERROR);
                    CLB_HIP(hipMemcpyAsync(Y2.ptr, L->Y.ptr, L->n * 4, hipMemcpyDeviceToDevice, ctl->stream),
                            COPY_BUFFER_ERROR);
                }
                sync(ctl);
                std::swap(L->X.ptr, X2.ptr);
                std::swap(L->X.cap, X2.cap);
                std::swap(L->X.n, X2.n);
                std::swap(L->Y.ptr, Y2.ptr);
                std::swap(L->Y.cap, Y2.cap);
                std::swap(L->Y.n, Y2.n);
                L->cap = ncap;
            }
            k_harvest<<<div_up(take, 256), 256, 0, ctl->stream>>>(feat.ptr, s.prod.ptr, nnz_row.ptr, nrows, stride,
                                                                   L->X.ptr, L->Y.ptr, L->n);
            CLB_LAUNCHED("harvest");
            sync(ctl);
            L->n += take;
        }
        if (total_samples) *total_samples = L->n;
    });
}

clb_status clb_predictor_train(clb_controls ctl, const clb_train_opts *opts, clb_train_stats *stats) {
    return run_guarded([&] {
        Learned *L = state(ctl);
        if (!L->ready || L->n == 0) raise(INVALID_VALUE, "no harvested samples to train on");
        const u32 epochs = opts ? opts->epochs : 20;
        const u32 batch = opts && opts->batch ? opts->batch : 4096;
        const float lr = opts && opts->lr > 0 ? opts->lr : 3e-3f;
        std::mt19937 rng(opts ? opts->seed : 1);
        auto t0 = std::chrono::steady_clock::now();
        const u32 n = static_cast<u32>(L->n);
        std::vector<u32> order(n);
        for (u32 k = 0; k < n; ++k) order[k] = k;
        L->idx.reset(ctl, n);
        DevArray<float> losses(ctl, 1);
        double first = 0, last = 0;
        u32 steps = 0;
        std::vector<float> ep_loss;
        DevArray<float> lbuf(ctl, div_up(n, batch) + 1);
        for (u32 e = 0; e < epochs; ++e) {
            std::shuffle(order.begin(), order.end(), rng);
            CLB_HIP(hipMemcpyAsync(L->idx.ptr, order.data(), 4ull * n, hipMemcpyHostToDevice, ctl->stream),
                    WRITE_BUFFER_ERROR);
            u32 nb_steps = 0;
            for (u32 b0 = 0; b0 < n; b0 += batch) {
                const u32 nb = std::min(batch, n - b0);
                train_step(ctl, L, L->X.ptr, L->Y.ptr, L->idx.ptr + b0, nb, lr, lbuf.ptr + nb_steps);
                ++nb_steps;
                ++steps;
            }
            std::vector<float> lh(nb_steps);
            CLB_HIP(hipMemcpyAsync(lh.data(), lbuf.ptr, 4ull * nb_steps, hipMemcpyDeviceToHost, ctl->stream),
                    READ_BUFFER_ERROR);
            sync(ctl);
            double m = 0;
            for (float x : lh) m += x;
            m /= nb_steps;
            if (e == 0) first = m;
            last = m;
        }
        if (stats) {
            stats->loss_first = first;
            stats->loss_last = last;
            stats->samples = L->n;
            stats->steps = steps;
            stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
    });
}

clb_status clb_predictor_get_params(clb_controls ctl, float *params, uint32_t n) {
    return run_guarded([&] {
        ensure_ready(ctl);
        if (n != NPARAM) raise(INVALID_ARGUMENT, "expected CLB_MLP_PARAMS parameters");
        readback(ctl, params, state(ctl)->params.ptr, NPARAM * 4);
    });
}

clb_status clb_predictor_set_params(clb_controls ctl, const float *params, uint32_t n) {
    return run_guarded([&] {
        if (n != NPARAM) raise(INVALID_ARGUMENT, "expected CLB_MLP_PARAMS parameters");
        Learned *L = state(ctl);
        if (!L->params.ptr) init_params(ctl, L, 1);
        CLB_HIP(hipMemcpyAsync(L->params.ptr, params, NPARAM * 4, hipMemcpyHostToDevice, ctl->stream),
                WRITE_BUFFER_ERROR);
        fill_u32(ctl, reinterpret_cast<u32 *>(L->m.ptr), NPARAM, 0);
        fill_u32(ctl, reinterpret_cast<u32 *>(L->v.ptr), NPARAM, 0);
        L->step = 0;
        sync(ctl);
        L->ready = true;
    });
}

clb_status clb_predictor_predict(clb_controls ctl, const clb_dcsr *a, const clb_dcsr *b, float *pred_nnz) {
    return run_guarded([&] {
        check_dims(*a, *b);
        ensure_ready(ctl);
        if (!a->nzr) return;
        if (!a->nnz || !b->nnz) {
            std::memset(pred_nnz, 0, 4ull * a->nzr);
            return;
        }
        Setup s;
        DevArray<float> feat, pred(ctl, a->nzr);
        DevArray<u32> cap(ctl, a->nzr);
        prepare_features(ctl, *a, *b, 0, a->nzr, s, feat);
        const BinCfg c = learned_bincfg(ctl, b->ncols);
        k_predict_caps<<<div_up(a->nzr, 256), 256, 0, ctl->stream>>>(state(ctl)->params.ptr, feat.ptr, s.prod.ptr,
                                                                      a->nzr, c, 0.0f, 1u, cap.ptr, pred.ptr);
        CLB_LAUNCHED("predict_caps");
        readback(ctl, pred_nnz, pred.ptr, 4ull * a->nzr);
    });
}

clb_status clb_mlp_forward(clb_controls ctl, const float *params, const float *x, uint32_t n, float *y) {
    return run_guarded([&] {
        if (!n) return;
        DevArray<float> p(ctl, NPARAM), dx(ctl, (u64)n * NF), dy(ctl, n);
        CLB_HIP(hipMemcpyAsync(p.ptr, params, NPARAM * 4, hipMemcpyHostToDevice, ctl->stream), WRITE_BUFFER_ERROR);
        CLB_HIP(hipMemcpyAsync(dx.ptr, x, 4ull * n * NF, hipMemcpyHostToDevice, ctl->stream), WRITE_BUFFER_ERROR);
        k_mlp_forward<<<div_up(n, 256), 256, 0, ctl->stream>>>(p.ptr, dx.ptr, n, dy.ptr);
        CLB_LAUNCHED("mlp_forward");
        readback(ctl, y, dy.ptr, 4ull * n);
    });
}

clb_status clb_mlp_loss_grad(clb_controls ctl, const float *params, const float *x, const float *t, uint32_t n,
                             float *grad, float *loss) {
    return run_guarded([&] {
        if (!n) raise(INVALID_ARGUMENT, "empty batch");
        Learned L;
        L.params.reset(ctl, NPARAM);
        L.grad.reset(ctl, NPARAM);
        DevArray<float> dx(ctl, (u64)n * NF), dt(ctl, n), dl(ctl, 1);
        CLB_HIP(hipMemcpyAsync(L.params.ptr, params, NPARAM * 4, hipMemcpyHostToDevice, ctl->stream),
                WRITE_BUFFER_ERROR);
        CLB_HIP(hipMemcpyAsync(dx.ptr, x, 4ull * n * NF, hipMemcpyHostToDevice, ctl->stream), WRITE_BUFFER_ERROR);
        CLB_HIP(hipMemcpyAsync(dt.ptr, t, 4ull * n, hipMemcpyHostToDevice, ctl->stream), WRITE_BUFFER_ERROR);
        const u32 blocks = div_up(n, TRAIN_WG);
        L.gpart.reset(ctl, (u64)blocks * NPARAM);
        L.lpart.reset(ctl, blocks);
        k_mlp_grad<<<blocks, TRAIN_WG, 0, ctl->stream>>>(L.params.ptr, dx.ptr, dt.ptr, nullptr, n, 1.0f / n,
                                                         L.gpart.ptr, L.lpart.ptr);
        CLB_LAUNCHED("mlp_grad");
        k_grad_reduce<<<div_up(NPARAM, 256), 256, 0, ctl->stream>>>(L.gpart.ptr, blocks, L.grad.ptr, L.lpart.ptr,
                                                                     dl.ptr);
        CLB_LAUNCHED("grad_reduce");
        readback(ctl, grad, L.grad.ptr, NPARAM * 4);
        readback(ctl, loss, dl.ptr, 4);
    });
}

clb_status clb_mlp_adam(clb_controls ctl, float *params, float *m, float *v, const float *grad, float lr, float b1,
                        float b2, float eps, uint32_t step) {
    return run_guarded([&] {
        if (!step) raise(INVALID_ARGUMENT, "Adam steps count from 1");
        DevArray<float> p(ctl, NPARAM), dm(ctl, NPARAM), dv(ctl, NPARAM), g(ctl, NPARAM);
        CLB_HIP(hipMemcpyAsync(p.ptr, params, NPARAM * 4, hipMemcpyHostToDevice, ctl->stream), WRITE_BUFFER_ERROR);
        CLB_HIP(hipMemcpyAsync(dm.ptr, m, NPARAM * 4, hipMemcpyHostToDevice, ctl->stream), WRITE_BUFFER_ERROR);
        CLB_HIP(hipMemcpyAsync(dv.ptr, v, NPARAM * 4, hipMemcpyHostToDevice, ctl->stream), WRITE_BUFFER_ERROR);
        CLB_HIP(hipMemcpyAsync(g.ptr, grad, NPARAM * 4, hipMemcpyHostToDevice, ctl->stream), WRITE_BUFFER_ERROR);
        const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
        k_adam<<<div_up(NPARAM, 256), 256, 0, ctl->stream>>>(p.ptr, dm.ptr, dv.ptr, g.ptr, NPARAM, lr, b1, b2, eps,
                                                             c1, c2);
        CLB_LAUNCHED("adam");
        readback(ctl, params, p.ptr, NPARAM * 4);
        readback(ctl, m, dm.ptr, NPARAM * 4);
        readback(ctl, v, dv.ptr, NPARAM * 4);
    });
}

}  // extern "C"
