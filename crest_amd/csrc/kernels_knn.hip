// Exact kNN (knn.rs:48-151,335-369) on MFMA: a distance screen keeps, per query, the
// best L candidates seen by each of its two lanes, then an FP64 rerank picks the k
// nearest and reports the reference's FP32 Euclidean distances.
//
// The screen works on centred rows x' = x - mean (distances are translation
// invariant; centring keeps the GEMM expansion well conditioned).  Distances come
// straight out of the MFMA: operands are augmented with the squared norms so that
// A_c . B_q = |c|^2 + |q|^2 - 2 c.q (the GEMM expansion of knn.rs:126):
//   FP32       A_c = [x_c | |c|^2, 1 | 0..]          B_q = [-2 x_q | 1, |q|^2 | 0..]
//   split-BF16 A_c = [h | h | l | nh, nl, 1, 1 | 0..]  B_q = [-2h | -2l | -2h | 1, 1, qh, ql | 0..]
//   FP16       A_c = [s x_c | nh, nl, 1, 1 | 0..]      B_q = [-2 s x_q | 1, 1, qh, ql | 0..]
// with x = h + l (bf16 hi/lo) so that A.B = -2(hh + hl + lh) + norms, and for FP16
// a power-of-two prescale s that brings the largest squared norm to at most 2^14
// (exact, and ranking-neutral), norms split into fp16 hi + lo.  Padding candidates
// carry |c|^2 = +inf and never enter a list.
//
// Operands are stored "fragment-major": 16-byte units of row i, chunk j at
// ((i/64)*KC + j)*64 + i%64 (16-bit types; FP32 splits each chunk into two float2
// halves), so one 64-row tile is one contiguous block that a workgroup copies into
// LDS and each MFMA fragment read from LDS is a contiguous 512-byte run per half-wave.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

#include <rocprim/rocprim.hpp>

#include "common.hpp"

namespace crest {

typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;
typedef __attribute__((ext_vector_type(4))) uint32_t u32x4;

constexpr int kTC = 64;    // candidates per LDS tile
constexpr int kQW = 64;    // queries per wave (two 32-column MFMA groups)
constexpr int kPadRows = 512;  // n is padded to a multiple of the largest workgroup's queries
constexpr int kPrecFP32 = 1, kPrecBF16X3 = 2, kPrecFP16 = 3;

__device__ inline uint16_t bf16_rn(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf16_f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Instantiated operand widths (16-byte chunks per row).
static const int kBf16Chunks[] = {4, 8, 12, 16, 20, 24, 32, 52};  // KC = 2*KSTEPS
static const int kF32Chunks[] = {2, 4, 8, 14, 16, 24, 34};         // KC (even), 4 floats each
static const int kF16Chunks[] = {4, 8, 16, 24, 32};                 // KC = 2*KSTEPS

int knn_kpad(int prec, int p) {
    if (prec == kPrecBF16X3 || prec == kPrecFP16) {
        const int need = prec == kPrecFP16 ? (p + 4 + 7) / 8 : (3 * p + 4 + 7) / 8;
        if (prec == kPrecFP16) {
            for (int kc : kF16Chunks)
                if (kc >= need) return kc * 8;
            return -1;
        }
        for (int kc : kBf16Chunks)
            if (kc >= need) return kc * 8;
        return -1;
    }
    const int need = (p + 2 + 3) / 4;
    for (int kc : kF32Chunks)
        if (kc >= need) return kc * 4;
    return -1;
}

// ---------------------------------------------------------------- prep
// 1. Column means and standard deviations (deterministic two-stage FP64 sums).
// 2. A spatial order: a Morton key of the (up to) four highest-variance columns,
//    quantised to +-4 sd, and a stable radix sort of the keys.  perm[i] is the row at
//    sorted position i.  Nearby rows end up in nearby positions, so a query's
//    neighbours sit in the candidate stages its workgroup visits first.
// 3. The augmented operands of the rows in sorted order, centred (x' = x - mean):
//    squared norms Σ x'² in f32, summed sequentially as knn.rs:346 does, written one
//    16-byte chunk at a time so that a wave (64 consecutive rows) stores 1 KB runs.
constexpr int kColParts = 512;
constexpr int kKeyCols = 4;

struct KnnWork {
    double *part;    // [kColParts * p]
    float *mean;     // [p]
    float *sd;       // [p]
    int32_t *cols;   // [kKeyCols] key columns (-1 = unused)
    uint32_t *maxbits;
    uint32_t *keys_c, *perm_c;   // [nc_pad] sorted keys and permutation of the candidate rows
    uint32_t *keys_q, *perm_q;   // [nq_pad] the same for separate query rows
    uint32_t *keys_in, *vals_in; // unsorted keys and row ids
    void *temp;
    size_t temp_bytes;
};

static size_t knn_sort_temp(int64_t n) {
    size_t b = 0;
    (void)rocprim::radix_sort_pairs((void *)nullptr, b, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                    (uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)n, 0, 32);
    return b;
}

// nq == 0: the queries are the candidates (kNN graph).
static KnnWork knn_work(void *w, int64_t nc, int64_t nq, int p, size_t *total) {
    const int64_t nc_pad = round_up(nc, kPadRows), nq_pad = round_up(nq, kPadRows);
    const int64_t nmax = nc_pad > nq_pad ? nc_pad : nq_pad;
    uint8_t *b = (uint8_t *)w;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        uint8_t *r = b ? b + o : nullptr;
        o = round_up((int64_t)(o + bytes), 256);
        return (void *)r;
    };
    KnnWork k;
    k.part = (double *)take((size_t)kColParts * p * sizeof(double));
    k.mean = (float *)take((size_t)p * sizeof(float));
    k.sd = (float *)take((size_t)p * sizeof(float));
    k.cols = (int32_t *)take(kKeyCols * sizeof(int32_t));
    k.maxbits = (uint32_t *)take(sizeof(uint32_t));
    k.keys_c = (uint32_t *)take((size_t)nc_pad * 4);
    k.perm_c = (uint32_t *)take((size_t)nc_pad * 4);
    k.keys_q = nq ? (uint32_t *)take((size_t)nq_pad * 4) : k.keys_c;
    k.perm_q = nq ? (uint32_t *)take((size_t)nq_pad * 4) : k.perm_c;
    k.keys_in = (uint32_t *)take((size_t)nmax * 4);
    k.vals_in = (uint32_t *)take((size_t)nmax * 4);
    const size_t t1 = knn_sort_temp(nc), t2 = nq ? knn_sort_temp(nq) : 0;
    k.temp_bytes = t1 > t2 ? t1 : t2;
    k.temp = take(k.temp_bytes);
    if (total) *total = o;
    return k;
}

size_t knn_work_bytes(int64_t nc, int64_t nq, int p) {
    size_t t = 0;
    (void)knn_work(nullptr, nc, nq, p, &t);
    return t;
}

// Block b sums rows [b*rows_per, (b+1)*rows_per) of x (mode 0) or of (x - mean)^2
// (mode 1): thread t takes column t % p of every (256/p)-th row, and the row groups
// are combined in order through LDS.
// With cell_off (a batch of subsets), blockIdx.y = subset y: rows cell_off[y] ..
// cell_off[y+1]-1 split into gridDim.x parts, its means at mean + y p, its parts at
// part + y gridDim.x p.
__global__ __launch_bounds__(256) void k_knn_colsum(const float *x, const float *mean, int mode, int64_t n, int p,
                                                    int64_t rows_per, double *part, const int64_t *cell_off) {
    __shared__ double red[256];
    if (cell_off) {
        const int y = blockIdx.y;
        x += cell_off[y] * p;
        n = cell_off[y + 1] - cell_off[y];
        rows_per = (n + gridDim.x - 1) / gridDim.x;
        part += (int64_t)y * gridDim.x * p;
        mean += (int64_t)y * p;
    }
    const int g = 256 / p, t = threadIdx.x, f = t % p, j = t / p;
    const int64_t r0 = blockIdx.x * rows_per, r1 = r0 + rows_per < n ? r0 + rows_per : n;
    double acc = 0.0;
    if (j < g) {
        const double mu = mode ? (double)mean[f] : 0.0;
        for (int64_t r = r0 + j; r < r1; r += g) {
            const double v = (double)x[r * p + f] - mu;
            acc += mode ? v * v : v;
        }
    }
    red[t] = acc;
    __syncthreads();
    if (t < p) {
        double s = 0.0;
        for (int jj = 0; jj < g; ++jj) s += red[jj * p + t];
        part[(int64_t)blockIdx.x * p + t] = s;
    }
}

// mode 0: mean[f] = Σ/n and the running maximum norm is reset; mode 1: sd[f] and the
// key columns (largest variance first, ties to the lower column).
// With cell_off, block y finalises subset y (its own mean, sd, key columns, maximum norm).
__global__ void k_knn_colfinal(const double *part, int nparts, int64_t n, int p, int mode, float *mean, float *sd,
                               int32_t *cols, uint32_t *maxbits, const int64_t *cell_off) {
    __shared__ float var[256];
    if (cell_off) {
        const int y = blockIdx.x;
        part += (int64_t)y * nparts * p;
        n = cell_off[y + 1] - cell_off[y];
        mean += (int64_t)y * p;
        sd += (int64_t)y * p;
        cols += (int64_t)y * kKeyCols;
        maxbits += y;
    }
    for (int f = threadIdx.x; f < p; f += blockDim.x) {
        double s = 0.0;
        for (int b = 0; b < nparts; ++b) s += part[(int64_t)b * p + f];
        if (mode == 0) {
            mean[f] = (float)(s / (double)n);
        } else {
            const float v = (float)(s / (double)n);
            var[f] = v;
            sd[f] = sqrtf(v);
        }
    }
    if (mode == 0) {
        if (threadIdx.x == 0) *maxbits = 0u;
        return;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t used[4] = {0, 0, 0, 0};
        for (int c = 0; c < kKeyCols; ++c) {
            int best = -1;
            for (int f = 0; f < p; ++f)
                if (!((used[f >> 6] >> (f & 63)) & 1) && (best < 0 || var[f] > var[best])) best = f;
            cols[c] = best;
            if (best >= 0) used[best >> 6] |= 1ull << (best & 63);
        }
    }
}

__device__ inline uint32_t morton_key(const float *row, const float *mean, const float *sd, const int32_t *cols) {
    int m = 0;
    uint32_t q[kKeyCols];
    for (int c = 0; c < kKeyCols; ++c)
        if (cols[c] >= 0) ++m;
    const int bits = 32 / m < 10 ? 32 / m : 10;
    for (int c = 0; c < m; ++c) {
        const int f = cols[c];
        const float s = sd[f];
        const float z = s > 0.0f ? (row[f] - mean[f]) / s : 0.0f;
        float v = floorf(z * (float)(1 << (bits - 3)) + (float)(1 << (bits - 1)));
        v = v < 0.0f ? 0.0f : v;  // NaN rows end at 0 as well
        const float top = (float)((1 << bits) - 1);
        q[c] = (uint32_t)(v < top ? v : top);
    }
    uint32_t key = 0;
    for (int b = bits - 1; b >= 0; --b)
        for (int c = 0; c < m; ++c) key = (key << 1) | ((q[c] >> b) & 1u);
    return key;
}

__global__ __launch_bounds__(256) void k_knn_keys(const float *x, int64_t n, int p, const float *mean,
                                                  const float *sd, const int32_t *cols, uint32_t *keys,
                                                  uint32_t *vals) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = morton_key(x + i * p, mean, sd, cols);
    vals[i] = (uint32_t)i;
}

// Batch of subsets (blockIdx.y = subset y): key (y << 32) | Morton key of the row
// under subset y's statistics, so one sort orders the subsets and, within each, the
// rows spatially.
__global__ __launch_bounds__(256) void k_knn_keys_seg(const float *x, const int64_t *cell_off, int p,
                                                      const float *mean, const float *sd, const int32_t *cols,
                                                      uint64_t *keys, uint32_t *vals) {
    const int y = blockIdx.y;
    const int64_t i = cell_off[y] + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= cell_off[y + 1]) return;
    keys[i] = ((uint64_t)y << 32) | morton_key(x + i * p, mean + (int64_t)y * p, sd + (int64_t)y * p,
                                               cols + (int64_t)y * kKeyCols);
    vals[i] = (uint32_t)i;
}

__device__ inline float centred_norm(const float *r, const float *mean, int p, float s) {
    float nn = 0.0f;
    for (int f = 0; f < p; ++f) {
        const float v = (r[f] - mean[f]) * s;
        const float sq = v * v;
        nn = nn + sq;
    }
    return nn;
}

__global__ __launch_bounds__(256) void k_knn_maxnorm(const float *x, int64_t n, int p, const float *mean,
                                                     uint32_t *maxbits, const int64_t *cell_off) {
    if (cell_off) {  // blockIdx.y = subset
        const int y = blockIdx.y;
        x += cell_off[y] * p;
        n = cell_off[y + 1] - cell_off[y];
        mean += (int64_t)y * p;
        maxbits += y;
    }
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    float nn = i < n ? centred_norm(x + i * p, mean, p, 1.0f) : 0.0f;
    for (int o = 32; o > 0; o >>= 1) nn = fmaxf(nn, __shfl_xor(nn, o, 64));
    if ((threadIdx.x & 63) == 0 && nn > 0.0f) atomicMax(maxbits, __float_as_uint(nn));
}

// Power-of-two s with s^2 * max|x'|^2 <= 2^14, so every |s x'| <= 128 and every
// fp16 norm stays far below the fp16 maximum (65504).
__device__ inline float fp16_prescale(const uint32_t *maxbits) {
    const float mx = __uint_as_float(*maxbits);
    if (!(mx > 0.0f) || !(mx <= 3.0e38f)) return 1.0f;
    int ex;
    (void)frexpf(mx, &ex);  // mx < 2^ex
    int e = 14 - ex;
    e = e >= 0 ? e / 2 : -((1 - e) / 2);  // floor(e / 2)
    e = e < -60 ? -60 : (e > 60 ? 60 : e);
    return ldexpf(1.0f, e);
}

// Operand row resolution.  Plain: row i (sorted position) holds data row perm[i],
// rows i >= n are padding.  Batch (sg.cell_off set, blockIdx.y = subset y): subset y
// owns operand rows row0[y] .. row0[y+1]-1, of which the first n_y hold its rows in
// sorted order (perm[cell_off[y] + local]); perm_pad records the data row of every
// operand row (0xFFFFFFFF for padding).
struct PrepSeg {
    const int64_t *cell_off;
    const int64_t *row0;
    uint32_t *perm_pad;
};
__device__ inline bool prep_row(const PrepSeg &sg, const uint32_t *perm, int64_t n, int64_t n_pad, int p,
                                int64_t &i, bool &real, int64_t &src, const float *&mean, const uint32_t *&maxbits) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (sg.cell_off) {
        const int y = blockIdx.y;
        if (t >= sg.row0[y + 1] - sg.row0[y]) return false;
        i = sg.row0[y] + t;
        real = t < sg.cell_off[y + 1] - sg.cell_off[y];
        src = real ? (int64_t)perm[sg.cell_off[y] + t] : 0;
        mean += (int64_t)y * p;
        if (maxbits) maxbits += y;
        sg.perm_pad[i] = real ? (uint32_t)src : 0xFFFFFFFFu;
        return true;
    }
    if (t >= n_pad) return false;
    i = t;
    real = t < n;
    src = real ? (int64_t)perm[t] : 0;
    return true;
}

// Element k of the split-BF16 operands of a centred row r' (values r - mean) with
// norm nn.
__device__ inline void bf16_elem(const float *r, const float *mean, int p, float nn, int k, uint16_t &a,
                                 uint16_t &b) {
    a = 0;
    b = 0;
    if (k < 3 * p) {
        const int seg = k / p, f = k - seg * p;
        const float v = r[f] - mean[f];
        const uint16_t h = bf16_rn(v);
        const uint16_t l = bf16_rn(v - bf16_f(h));
        a = seg == 2 ? l : h;
        b = bf16_rn(-2.0f * bf16_f(seg == 1 ? l : h));
    } else if (k < 3 * p + 4) {
        const uint16_t nh = bf16_rn(nn);
        const uint16_t nl = bf16_rn(nn - bf16_f(nh));
        const int t = k - 3 * p;
        a = t == 0 ? nh : (t == 1 ? nl : (uint16_t)0x3F80);
        b = t == 2 ? nh : (t == 3 ? nl : (uint16_t)0x3F80);
    }
}

// Operand row i (sorted position) holds data row perm[i]; rows i >= n are padding.
__global__ void k_knn_prep_bf16(const float *x, const uint32_t *perm, const float *mean, int64_t n, int64_t n_pad,
                                int p, int kc, uint4 *xa, uint4 *xb, PrepSeg sg) {
    int64_t i, src;
    bool real;
    const uint32_t *nomax = nullptr;
    if (!prep_row(sg, perm, n, n_pad, p, i, real, src, mean, nomax)) return;
    const float *r = x + src * p;
    const float nn = real ? centred_norm(r, mean, p, 1.0f) : INFINITY;
    for (int j = 0; j < kc; ++j) {
        uint16_t ea[8], eb[8];
        for (int e = 0; e < 8; ++e) {
            const int k = 8 * j + e;
            if (real) {
                bf16_elem(r, mean, p, nn, k, ea[e], eb[e]);
            } else {  // padding candidate: +inf norm, never a neighbour
                const int t = k - 3 * p;
                ea[e] = t == 0 ? (uint16_t)0x7F80 : ((t == 2 || t == 3) ? (uint16_t)0x3F80 : (uint16_t)0);
                eb[e] = 0;
            }
        }
        uint4 va, vb;
        __builtin_memcpy(&va, ea, 16);
        __builtin_memcpy(&vb, eb, 16);
        const int64_t u = ((i / 64) * kc + j) * 64 + (i % 64);
        if (xa) xa[u] = va;
        if (xb) xb[u] = vb;
    }
}

// FP16 operands of a centred, prescaled row; fp16 conversions round to nearest even.
__global__ void k_knn_prep_f16(const float *x, const uint32_t *perm, const float *mean, const uint32_t *maxbits,
                               int64_t n, int64_t n_pad, int p, int kc, uint4 *xa, uint4 *xb, PrepSeg sg) {
    int64_t i, src;
    bool real;
    if (!prep_row(sg, perm, n, n_pad, p, i, real, src, mean, maxbits)) return;
    const float s = fp16_prescale(maxbits);
    const float *r = x + src * p;
    const float nn = real ? centred_norm(r, mean, p, s) : INFINITY;
    const _Float16 nh = (_Float16)nn;
    const _Float16 nl = real ? (_Float16)(nn - (float)nh) : (_Float16)0.0f;
    const _Float16 one = (_Float16)1.0f, zero = (_Float16)0.0f;
    for (int j = 0; j < kc; ++j) {
        _Float16 ea[8], eb[8];
        for (int e = 0; e < 8; ++e) {
            const int k = 8 * j + e;
            ea[e] = eb[e] = zero;
            if (k < p) {
                if (real) {
                    const _Float16 h = (_Float16)((r[k] - mean[k]) * s);
                    ea[e] = h;
                    eb[e] = (_Float16)(-2.0f * (float)h);
                }
            } else if (k < p + 4) {
                const int t = k - p;
                ea[e] = t == 0 ? nh : (t == 1 ? nl : one);
                eb[e] = real ? (t == 2 ? nh : (t == 3 ? nl : one)) : zero;
            }
        }
        uint4 va, vb;
        __builtin_memcpy(&va, ea, 16);
        __builtin_memcpy(&vb, eb, 16);
        const int64_t u = ((i / 64) * kc + j) * 64 + (i % 64);
        if (xa) xa[u] = va;
        if (xb) xb[u] = vb;
    }
}

// The same operands for 64 consecutive sorted positions per workgroup (no subsets):
// their rows, in permuted order, are first read into LDS with the threads walking
// along each row (a thread per row reading its row element by element touches 64
// rows' lines per load), then each row's norm and the 16-byte units are built from
// LDS with exactly the arithmetic of k_knn_prep_f16, so the operands are identical.
__global__ __launch_bounds__(256) void k_knn_prep_f16_tile(const float *x, const uint32_t *perm, const float *mean,
                                                           const uint32_t *maxbits, int64_t n, int p, int kc,
                                                           uint4 *xa, uint4 *xb) {
    extern __shared__ float tile[];  // [64][p + 1]
    __shared__ float nnv[64];
    const int ld = p + 1;
    const int64_t i0 = (int64_t)blockIdx.x * 64;
    const float s = fp16_prescale(maxbits);
    for (int e = threadIdx.x; e < 64 * p; e += blockDim.x) {
        const int r = e / p, f = e - r * p;
        const int64_t i = i0 + r;
        tile[r * ld + f] = i < n ? x[(int64_t)perm[i] * p + f] : 0.0f;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int r = threadIdx.x;
        float nn = 0.0f;
        for (int f = 0; f < p; ++f) {
            const float v = (tile[r * ld + f] - mean[f]) * s;
            const float sq = v * v;
            nn = nn + sq;
        }
        nnv[r] = i0 + r < n ? nn : INFINITY;
    }
    __syncthreads();
    const _Float16 one = (_Float16)1.0f, zero = (_Float16)0.0f;
    for (int q = threadIdx.x; q < 64 * kc; q += blockDim.x) {
        const int r = q & 63, j = q >> 6;
        const bool real = i0 + r < n;
        const float nn = nnv[r];
        const _Float16 nh = (_Float16)nn;
        const _Float16 nl = real ? (_Float16)(nn - (float)nh) : (_Float16)0.0f;
        _Float16 ea[8], eb[8];
        for (int e = 0; e < 8; ++e) {
            const int k = 8 * j + e;
            ea[e] = eb[e] = zero;
            if (k < p) {
                if (real) {
                    const _Float16 h = (_Float16)((tile[r * ld + k] - mean[k]) * s);
                    ea[e] = h;
                    eb[e] = (_Float16)(-2.0f * (float)h);
                }
            } else if (k < p + 4) {
                const int t = k - p;
                ea[e] = t == 0 ? nh : (t == 1 ? nl : one);
                eb[e] = real ? (t == 2 ? nh : (t == 3 ? nl : one)) : zero;
            }
        }
        uint4 va, vb;
        __builtin_memcpy(&va, ea, 16);
        __builtin_memcpy(&vb, eb, 16);
        const int64_t u = ((int64_t)blockIdx.x * kc + j) * 64 + r;
        if (xa) xa[u] = va;
        if (xb) xb[u] = vb;
    }
}

// The largest centred squared norm of rows [0, n) (natural order), 64 rows per
// workgroup read along the rows into LDS; each row's norm in k_knn_maxnorm's order.
__global__ __launch_bounds__(256) void k_knn_maxnorm_tile(const float *x, int64_t n, int p, const float *mean,
                                                          uint32_t *maxbits) {
    extern __shared__ float tile[];  // [64][p + 1]
    const int ld = p + 1;
    const int64_t i0 = (int64_t)blockIdx.x * 64;
    const int64_t rows = n - i0 < 64 ? n - i0 : 64;
    for (int64_t e = threadIdx.x; e < rows * p; e += blockDim.x) {
        const int r = (int)(e / p), f = (int)(e - (int64_t)r * p);
        tile[r * ld + f] = x[(i0 + r) * p + f];
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int r = threadIdx.x;
        float nn = 0.0f;
        if (r < rows)
            for (int f = 0; f < p; ++f) {
                const float v = (tile[r * ld + f] - mean[f]) * 1.0f;
                const float sq = v * v;
                nn = nn + sq;
            }
        for (int o = 32; o > 0; o >>= 1) nn = fmaxf(nn, __shfl_xor(nn, o, 64));
        if (threadIdx.x == 0 && nn > 0.0f) atomicMax(maxbits, __float_as_uint(nn));
    }
}

__device__ inline void f32_elem(const float *r, const float *mean, int p, float nn, int k, float &a, float &b) {
    a = 0.0f;
    b = 0.0f;
    if (k < p) {
        const float v = r[k] - mean[k];
        a = v;
        b = -2.0f * v;
    } else if (k == p) {
        a = nn;
        b = 1.0f;
    } else if (k == p + 1) {
        a = 1.0f;
        b = nn;
    }
}

__global__ void k_knn_prep_f32(const float *x, const uint32_t *perm, const float *mean, int64_t n, int64_t n_pad,
                               int p, int kc, float2 *xa, float2 *xb, PrepSeg sg) {
    int64_t i, src;
    bool real;
    const uint32_t *nomax = nullptr;
    if (!prep_row(sg, perm, n, n_pad, p, i, real, src, mean, nomax)) return;
    const float *r = x + src * p;
    const float nn = real ? centred_norm(r, mean, p, 1.0f) : INFINITY;
    for (int j = 0; j < kc; ++j)
        for (int h = 0; h < 2; ++h) {
            float a0, a1, b0, b1;
            if (real) {
                f32_elem(r, mean, p, nn, 4 * j + 2 * h, a0, b0);
                f32_elem(r, mean, p, nn, 4 * j + 2 * h + 1, a1, b1);
            } else {
                a0 = (4 * j + 2 * h == p) ? INFINITY : 0.0f;
                a1 = (4 * j + 2 * h + 1 == p) ? INFINITY : 0.0f;
                b0 = b1 = 0.0f;
            }
            const int64_t u = (((i / 64) * kc + j) * 2 + h) * 64 + (i % 64);
            if (xa) xa[u] = make_float2(a0, a1);
            if (xb) xb[u] = make_float2(b0, b1);
        }
}

int64_t knn_rows_padded(int64_t n) { return round_up(n, kPadRows); }

static void sort_rows(const KnnWork &w, const float *x, int64_t n, int p, uint32_t *keys_out, uint32_t *perm,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_knn_keys, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, x, n, p, w.mean, w.sd, w.cols,
                       w.keys_in, w.vals_in);
    size_t tb = w.temp_bytes;
    if (rocprim::radix_sort_pairs(w.temp, tb, w.keys_in, keys_out, w.vals_in, perm, (size_t)n, 0, 32, s) !=
        hipSuccess)
        throw std::runtime_error("rocprim radix_sort_pairs failed");
}

static void write_operands(int prec, const float *x, const uint32_t *perm, const KnnWork &w, int64_t n, int p,
                           int kpad, void *xa, void *xb, hipStream_t s) {
    const int64_t n_pad = knn_rows_padded(n);
    const int g = (int)ceil_div(n_pad, 256);
    if (prec == kPrecFP16)
        hipLaunchKernelGGL(k_knn_prep_f16_tile, dim3((unsigned)(n_pad / 64)), dim3(256),
                           (size_t)64 * (p + 1) * sizeof(float), s, x, perm, w.mean, w.maxbits, n, p, kpad / 8,
                           (uint4 *)xa, (uint4 *)xb);
    else if (prec == kPrecBF16X3)
        hipLaunchKernelGGL(k_knn_prep_bf16, dim3(g), dim3(256), 0, s, x, perm, w.mean, n, n_pad, p, kpad / 8,
                           (uint4 *)xa, (uint4 *)xb, PrepSeg{});
    else
        hipLaunchKernelGGL(k_knn_prep_f32, dim3(g), dim3(256), 0, s, x, perm, w.mean, n, n_pad, p, kpad / 4,
                           (float2 *)xa, (float2 *)xb, PrepSeg{});
}

KnnOrder launch_knn_prep(int prec, const float *xc, int64_t nc, const float *xq, int64_t nq, int p, int kpad,
                         void *work, void *xa, void *xb, hipStream_t s) {
    const bool graph = xq == nullptr;
    const KnnWork w = knn_work(work, nc, graph ? 0 : nq, p, nullptr);
    const int64_t rows_per = ceil_div(nc, kColParts);
    const int nparts = (int)ceil_div(nc, rows_per);
    const int64_t *no = nullptr;
    hipLaunchKernelGGL(k_knn_colsum, dim3(nparts), dim3(256), 0, s, xc, w.mean, 0, nc, p, rows_per, w.part, no);
    hipLaunchKernelGGL(k_knn_colfinal, dim3(1), dim3(256), 0, s, w.part, nparts, nc, p, 0, w.mean, w.sd, w.cols,
                       w.maxbits, no);
    hipLaunchKernelGGL(k_knn_colsum, dim3(nparts), dim3(256), 0, s, xc, w.mean, 1, nc, p, rows_per, w.part, no);
    hipLaunchKernelGGL(k_knn_colfinal, dim3(1), dim3(256), 0, s, w.part, nparts, nc, p, 1, w.mean, w.sd, w.cols,
                       w.maxbits, no);
    sort_rows(w, xc, nc, p, w.keys_c, w.perm_c, s);
    if (!graph) sort_rows(w, xq, nq, p, w.keys_q, w.perm_q, s);
    if (prec == kPrecFP16) {
        const size_t lds = (size_t)64 * (p + 1) * sizeof(float);
        hipLaunchKernelGGL(k_knn_maxnorm_tile, dim3((unsigned)ceil_div(nc, 64)), dim3(256), lds, s, xc, nc, p, w.mean,
                           w.maxbits);
        if (!graph)
            hipLaunchKernelGGL(k_knn_maxnorm_tile, dim3((unsigned)ceil_div(nq, 64)), dim3(256), lds, s, xq, nq, p,
                               w.mean, w.maxbits);
    }
    if (graph) {
        write_operands(prec, xc, w.perm_c, w, nc, p, kpad, xa, xb, s);
    } else {
        write_operands(prec, xc, w.perm_c, w, nc, p, kpad, xa, nullptr, s);
        write_operands(prec, xq, w.perm_q, w, nq, p, kpad, nullptr, xb, s);
    }
    KnnOrder o;
    o.perm_c = w.perm_c;
    o.perm_q = w.perm_q;
    o.keys_c = graph ? nullptr : w.keys_c;
    o.keys_q = graph ? nullptr : w.keys_q;
    o.mean = w.mean;
    o.maxbits = prec == kPrecFP16 ? w.maxbits : nullptr;
    return o;
}

// ---------------------------------------------------------------- batch prep
// The same three steps for a batch of independent subsets of the rows of x (subset y
// = rows cell_off[y] .. cell_off[y+1]-1): statistics, key columns and FP16 prescale
// per subset; one stable sort of 64-bit keys (subset, Morton key) orders the subsets
// and each subset spatially; subset y's operands occupy rows row0[y] .. row0[y+1]-1
// (its rows padded to kPadRows), and perm_pad maps every operand row back to its data
// row (0xFFFFFFFF for padding).
constexpr int kBatchColParts = 64;

struct KnnBatchWork {
    double *part;
    float *mean, *sd;
    int32_t *cols;
    uint32_t *maxbits;
    uint64_t *keys_in, *keys_out;
    uint32_t *vals_in, *perm, *perm_pad;
    void *temp;
    size_t temp_bytes;
};

static KnnBatchWork knn_batch_work(void *w, int64_t n, int64_t n_pad, int n_sub, int p, size_t *total) {
    uint8_t *b = (uint8_t *)w;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        uint8_t *r = b ? b + o : nullptr;
        o = round_up((int64_t)(o + bytes), 256);
        return (void *)r;
    };
    KnnBatchWork k;
    k.part = (double *)take((size_t)n_sub * kBatchColParts * p * sizeof(double));
    k.mean = (float *)take((size_t)n_sub * p * sizeof(float));
    k.sd = (float *)take((size_t)n_sub * p * sizeof(float));
    k.cols = (int32_t *)take((size_t)n_sub * kKeyCols * sizeof(int32_t));
    k.maxbits = (uint32_t *)take((size_t)n_sub * sizeof(uint32_t));
    k.keys_in = (uint64_t *)take((size_t)n * 8);
    k.keys_out = (uint64_t *)take((size_t)n * 8);
    k.vals_in = (uint32_t *)take((size_t)n * 4);
    k.perm = (uint32_t *)take((size_t)n * 4);
    k.perm_pad = (uint32_t *)take((size_t)n_pad * 4);
    size_t tb = 0;
    (void)rocprim::radix_sort_pairs((void *)nullptr, tb, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                    (uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)n, 0, 64);
    k.temp_bytes = tb;
    k.temp = take(tb);
    if (total) *total = o;
    return k;
}

const float *knn_batch_means(void *work, int64_t n, int64_t n_pad, int n_sub, int p) {
    return knn_batch_work(work, n, n_pad, n_sub, p, nullptr).mean;
}

size_t knn_batch_work_bytes(int64_t n, int64_t n_pad, int n_sub, int p) {
    size_t t = 0;
    (void)knn_batch_work(nullptr, n, n_pad, n_sub, p, &t);
    return t;
}

KnnOrder launch_knn_prep_batch(int prec, const float *x, int64_t n, const int64_t *cell_off, int n_sub,
                               int64_t max_rows, const int64_t *row0, int64_t n_pad, int64_t max_pad, int p, int kpad,
                               void *work, void *xa, void *xb, hipStream_t s) {
    const KnnBatchWork w = knn_batch_work(work, n, n_pad, n_sub, p, nullptr);
    const dim3 gp(kBatchColParts, n_sub), grow((unsigned)ceil_div(max_rows, 256), n_sub);
    hipLaunchKernelGGL(k_knn_colsum, gp, dim3(256), 0, s, x, w.mean, 0, (int64_t)0, p, (int64_t)0, w.part, cell_off);
    hipLaunchKernelGGL(k_knn_colfinal, dim3(n_sub), dim3(256), 0, s, w.part, kBatchColParts, (int64_t)0, p, 0, w.mean,
                       w.sd, w.cols, w.maxbits, cell_off);
    hipLaunchKernelGGL(k_knn_colsum, gp, dim3(256), 0, s, x, w.mean, 1, (int64_t)0, p, (int64_t)0, w.part, cell_off);
    hipLaunchKernelGGL(k_knn_colfinal, dim3(n_sub), dim3(256), 0, s, w.part, kBatchColParts, (int64_t)0, p, 1, w.mean,
                       w.sd, w.cols, w.maxbits, cell_off);
    hipLaunchKernelGGL(k_knn_keys_seg, grow, dim3(256), 0, s, x, cell_off, p, w.mean, w.sd, w.cols, w.keys_in,
                       w.vals_in);
    int end_bit = 33;
    while (end_bit < 64 && ((int64_t)1 << (end_bit - 32)) < n_sub) ++end_bit;
    size_t tb = w.temp_bytes;
    if (rocprim::radix_sort_pairs(w.temp, tb, w.keys_in, w.keys_out, w.vals_in, w.perm, (size_t)n, 0, end_bit, s) !=
        hipSuccess)
        throw std::runtime_error("rocprim radix_sort_pairs failed");
    if (prec == kPrecFP16)
        hipLaunchKernelGGL(k_knn_maxnorm, grow, dim3(256), 0, s, x, (int64_t)0, p, w.mean, w.maxbits, cell_off);
    const PrepSeg sg{cell_off, row0, w.perm_pad};
    const dim3 gpad((unsigned)ceil_div(max_pad, 256), n_sub);
    if (prec == kPrecFP16)
        hipLaunchKernelGGL(k_knn_prep_f16, gpad, dim3(256), 0, s, x, w.perm, w.mean, w.maxbits, (int64_t)0, (int64_t)0,
                           p, kpad / 8, (uint4 *)xa, (uint4 *)xb, sg);
    else if (prec == kPrecBF16X3)
        hipLaunchKernelGGL(k_knn_prep_bf16, gpad, dim3(256), 0, s, x, w.perm, w.mean, (int64_t)0, (int64_t)0, p,
                           kpad / 8, (uint4 *)xa, (uint4 *)xb, sg);
    else
        hipLaunchKernelGGL(k_knn_prep_f32, gpad, dim3(256), 0, s, x, w.perm, w.mean, (int64_t)0, (int64_t)0, p,
                           kpad / 4, (float2 *)xa, (float2 *)xb, sg);
    KnnOrder o;
    o.perm_c = w.perm_pad;
    o.perm_q = w.perm_pad;
    o.keys_c = nullptr;
    o.keys_q = nullptr;
    o.mean = w.mean;
    o.maxbits = prec == kPrecFP16 ? w.maxbits : nullptr;
    return o;
}

// ---------------------------------------------------------------- tile staging
// N values carried in registers across a loop iteration: a recursive struct instead
// of an array, so no indexing ever asks for a stack slot.  Element m of thread t is
// unit t + m*STRIDE of the tile.
template <class T, int N> struct Regs {
    T head;
    Regs<T, N - 1> tail;
};
template <class T> struct Regs<T, 0> {};

template <class T, int N, int STRIDE> __device__ inline void regs_load(Regs<T, N> &r, const T *src, int i) {
    r.head = src[i];
    if constexpr (N > 1) regs_load<T, N - 1, STRIDE>(r.tail, src, i + STRIDE);
}
template <class T, int N, int STRIDE> __device__ inline void regs_store(const Regs<T, N> &r, T *dst, int i) {
    dst[i] = r.head;
    if constexpr (N > 1) regs_store<T, N - 1, STRIDE>(r.tail, dst, i + STRIDE);
}

// ---------------------------------------------------------------- top-L lists
// Sorted ascending; a new entry goes after equal ones, so among equal distances the
// entry offered first stays ahead, as push_topk does (knn.rs:68-87).  With the
// comparisons c_j = d < ld[j] taken on the old list (monotone in j), slot j keeps
// its entry, takes d, or takes slot j-1's entry; the distances need only
// ld[j] = min(ld[j], max(d, ld[j-1])), and no value is carried from slot to slot.
// Entries are +0.0 or positive (or +inf), so their bit patterns order as unsigned
// integers and the min/max run on the integer unit without NaN canonicalisation.
// The median of three unsigned integers in one instruction (v_med3_u32).
__device__ inline uint32_t med3_u32(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

template <int L>
__device__ __attribute__((always_inline)) inline void topk_insert(float (&ld)[L], uint32_t (&li)[L], float d, uint32_t id) {
    const uint32_t db = __float_as_uint(d);
    bool c[L];
#pragma unroll
    for (int j = 0; j < L; ++j) c[j] = db < __float_as_uint(ld[j]);
#pragma unroll
    for (int j = L - 1; j > 0; --j) {
        li[j] = c[j - 1] ? li[j - 1] : (c[j] ? id : li[j]);
        // ld[j-1] <= ld[j]: min(ld[j], max(d, ld[j-1])) is the median of the three
        ld[j] = __uint_as_float(med3_u32(db, __float_as_uint(ld[j - 1]), __float_as_uint(ld[j])));
    }
    li[0] = c[0] ? id : li[0];
    ld[0] = __uint_as_float(min(__float_as_uint(ld[0]), db));
}

// Register r of lane l holds the pair (candidate cb + 8(r/4) + 4(l/32) + r%4, query
// of column l%32).  The registers that hold an entry below some lane's threshold
// are visited in a wave-uniform loop, so a[r] is a uniform register index; each
// visit inserts into the lists of the lanes whose value is still below their
// threshold.
template <int L>
__device__ __attribute__((always_inline)) inline void topk_update(float (&ld)[L], uint32_t (&li)[L], const f32x16 &a, uint32_t cb, int half) {
    // the minimum on the bit patterns as signed integers: non-negative floats order as
    // their patterns, negative ones fall below any threshold (which is >= 0), so
    // "some value < thr" is decided without canonicalising the MFMA outputs
    int32_t m = __float_as_int(a[0]);
#pragma unroll
    for (int r = 1; r < 16; ++r) m = min(m, __float_as_int(a[r]));
    const float thr = ld[L - 1];
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(m < __float_as_int(thr)) == 0, 1)) return;
    uint32_t u = 0;
#pragma unroll
    for (int r = 0; r < 16; ++r) u |= (__builtin_amdgcn_ballot_w64(a[r] < thr) != 0 ? 1u : 0u) << r;
    while (u) {
        const int r = __builtin_ctz(u);
        u &= u - 1;
        const float v = a[r];
        if (v < ld[L - 1]) topk_insert<L>(ld, li, v > 0.0f ? v : 0.0f, cb + 8 * (r >> 2) + 4 * half + (r & 3));
    }
}

// Insertion buffer slots per lane and list in the 16-bit screens (measured at 1M x 50,
// FP16, k = 15: 129.1 ms with direct insertions, 122.9 ms with buffers of 4); 0 (direct
// insertions) for the shapes whose registers already spill without the buffers.
template <int KC, int L> constexpr int screen_buf() {
    return ((L == 16 && (KC == 12 || KC == 16)) || KC == 52 || (L == 64 && KC >= 16)) ? 0 : 4;
}

// Deferred insertions: a lane's entries below its list's threshold are appended to
// a per-lane buffer of B slots in LDS (slot s of the lane at buf[64 s]), and the
// buffers are merged into the lists together once some lane's buffer is full.  A
// list insertion costs the whole wave's issue time whichever lanes need it, and
// lanes insert at different fragments, so merging every lane's buffer in one
// wave-uniform pass turns about one insertion pass per inserted entry into one per
// B entries of the fullest lane.  The threshold of the append test is the list's at
// the last merge (entries that stop qualifying are dropped by the merge).
template <int L, int B>
__device__ __attribute__((always_inline)) inline void topk_flush(float (&ld)[L], uint32_t (&li)[L], const uint2 *buf, uint32_t &cnt) {
    // slots in a uniform loop up to the fullest lane's count
    uint32_t s = 0;
    while (__builtin_amdgcn_ballot_w64(s < cnt) != 0) {
        if (s < cnt) {
            const uint2 e = buf[64 * s];
            if (__uint_as_float(e.x) < ld[L - 1]) topk_insert<L>(ld, li, __uint_as_float(e.x), e.y);
        }
        ++s;
    }
    cnt = 0;
}

template <int L, int B>
__device__ __attribute__((always_inline)) inline void topk_update_buf(float (&ld)[L], uint32_t (&li)[L], const f32x16 &a, uint32_t cb, int half,
                                       uint2 *buf, uint32_t &cnt) {
    int32_t m = __float_as_int(a[0]);
#pragma unroll
    for (int r = 1; r < 16; ++r) m = min(m, __float_as_int(a[r]));
    // the slow path is rare after the first stages: laid out away from the MFMA stream
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(m < __float_as_int(ld[L - 1])) == 0, 1)) return;
    const float thr = ld[L - 1];
    uint32_t u = 0;
#pragma unroll
    for (int r = 0; r < 16; ++r) u |= (__builtin_amdgcn_ballot_w64(a[r] < thr) != 0 ? 1u : 0u) << r;
    while (u) {
        const int r = __builtin_ctz(u);
        u &= u - 1;
        const float v = a[r];
        if (v < ld[L - 1]) {
            buf[64 * cnt] = make_uint2(__float_as_uint(v > 0.0f ? v : 0.0f), cb + 8 * (r >> 2) + 4 * half + (r & 3));
            ++cnt;
        }
        if (__builtin_amdgcn_ballot_w64(cnt == (uint32_t)B) != 0) topk_flush<L, B>(ld, li, buf, cnt);
    }
}

template <int L>
__device__ inline void self_mask(f32x16 &a, uint32_t cb, int half, uint32_t q) {
#pragma unroll
    for (int r = 0; r < 16; ++r)
        if (cb + 8 * (r >> 2) + 4 * half + (r & 3) == q) a[r] = INFINITY;
}

template <int L>
__device__ inline void write_lists(const float (&ld)[L], const uint32_t (&li)[L], int64_t q, int half, float *cand_d,
                                   uint32_t *cand_i) {
    float *od = cand_d + q * (2 * L) + half * L;
    uint32_t *oi = cand_i + q * (2 * L) + half * L;
#pragma unroll
    for (int j = 0; j < L; ++j) {
        od[j] = ld[j];
        oi[j] = li[j];
    }
}

// ---------------------------------------------------------------- screen, 16-bit operands
// One MFMA K-step of 16: v_mfma_f32_32x32x16_{bf16,f16}.  TY 0 = bf16, 1 = fp16.
template <int TY> __device__ inline f32x16 mfma16(u32x4 a, u32x4 b, f32x16 c) {
    if constexpr (TY == 0)
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                       0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0,
                                                      0, 0);
}

// Workgroup of NW waves, 64 queries per wave.  Two waves per SIMD fit when the
// B fragments, both accumulator pairs and the lists leave room in 256 VGPRs
// (checked against the compiler's register report: no spills).  Four-wave
// workgroups throughout: eight-wave ones halve the stage traffic per query but make
// twice as many waves meet at every stage barrier, and with the A operands read a
// fragment ahead four waves measured faster (1M x 50 FP16: 131.3 against 134.6 ms).
// The software pipeline (two accumulator pairs, below) is used where it keeps two
// waves per SIMD, or where one wave per SIMD is all that fits anyway.
template <int KC, int L> constexpr bool screen_pipe() { return (L <= 16 && KC <= 16) || !(L <= 16 && KC <= 20); }
template <int KC, int L> constexpr int screen_waves() {
    return screen_pipe<KC, L>() ? ((L <= 16 && KC <= 16) ? 2 : 1) : 2;
}
template <int TY, int KC, int L> constexpr int screen_nw() { return 4; }  // waves per workgroup
// 64-candidate blocks per LDS stage: as many (up to four) as keep the
// double-buffered stages within 64 KB of LDS.
template <int KC> constexpr int screen_tb() { return KC <= 8 ? 4 : KC <= 16 ? 2 : 1; }

// A stage holds TB blocks of 64 candidates = 2*TB fragments of 32.  Pipelined,
// fragments alternate between two accumulator pairs, and each fragment's list
// update is issued after the MFMAs of the next one, so the matrix core works while
// the wave's VALU merges the previous results.
template <int TY, int KC, int L, int NW, int TB, int BUF>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(screen_waves<KC, L>()))) void k_knn_screen16(
    const u32x4 *__restrict__ xa, const u32x4 *__restrict__ xb, int64_t nc, int64_t nq, const uint32_t *keys_c,
    const uint32_t *keys_q, int self, float *cand_d, uint32_t *cand_i, KnnSeg sg) {
    constexpr int KS = KC / 2;     // MFMA K-steps of 16
    constexpr int BU = KC * kTC;   // 16-byte units per 64-candidate block
    constexpr int TU = TB * BU;    // units per stage
    constexpr int NT = NW * 64;
    constexpr int NF = 2 * TB;     // fragments per stage
    static_assert(TU % NT == 0, "stage must split evenly over the workgroup");
    constexpr int PER = TU / NT;
    __shared__ u32x4 lds[2][TU];
    // insertion buffers (BUF > 0): [wave][list][slot][lane]
    __shared__ uint2 qbuf[BUF > 0 ? NW * 2 * BUF * 64 : 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, col = lane & 31, half = lane >> 5;
    uint2 *buf0 = qbuf + (BUF > 0 ? (wave * 2) * BUF * 64 + lane : 0);
    uint2 *buf1 = qbuf + (BUF > 0 ? (wave * 2 + 1) * BUF * 64 + lane : 0);
    uint32_t cnt0 = 0, cnt1 = 0;
    // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous range of query blocks, so the workgroups sharing an XCD's L2 have
    // neighbouring home stages and walk overlapping candidate streams.
    const int64_t nb = gridDim.x, xcd = blockIdx.x % 8, per = nb / 8, rem = nb % 8;
    int64_t wg = xcd * per + (xcd < rem ? xcd : rem) + blockIdx.x / 8;
    if (sg.order) {  // batch: query blocks in the order of the table (largest subsets first)
        constexpr int SB = kPadRows / NT;
        wg = (int64_t)sg.order[blockIdx.x / SB] * SB + blockIdx.x % SB;
    }
    const int64_t q0 = wg * NT + wave * kQW;
    const uint32_t qa = (uint32_t)(q0 + col), qb = (uint32_t)(q0 + 32 + col);
    const u32x4 *qbase = xb + (q0 / 64) * KC * 64;
    u32x4 b0[KS], b1[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        b0[s] = qbase[(2 * s + half) * 64 + col];
        b1[s] = qbase[(2 * s + half) * 64 + 32 + col];
    }
    float d0[L], d1[L];
    uint32_t i0[L], i1[L];
#pragma unroll
    for (int j = 0; j < L; ++j) {
        d0[j] = d1[j] = INFINITY;
        i0[j] = i1[j] = 0xFFFFFFFFu;
    }
    auto mma = [&](const u32x4 *stage, int f, f32x16 &c0, f32x16 &c1) {
        const u32x4 *base = stage + (f >> 1) * BU + (f & 1) * 32 + col;
        c0 = f32x16{};
        c1 = f32x16{};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const u32x4 a = base[(2 * s + half) * kTC];
            c0 = mfma16<TY>(a, b0[s], c0);
            c1 = mfma16<TY>(a, b1[s], c1);
        }
    };
    // the same, split: the A operands of fragment f from LDS, then its MFMAs, so that
    // the next fragment's operands are read while this one's MFMAs run
    auto lda = [&](const u32x4 *stage, int f, u32x4 (&a)[KS]) {
        const u32x4 *base = stage + (f >> 1) * BU + (f & 1) * 32 + col;
#pragma unroll
        for (int s = 0; s < KS; ++s) a[s] = base[(2 * s + half) * kTC];
    };
    auto mmaa = [&](const u32x4 (&a)[KS], f32x16 &c0, f32x16 &c1) {
        c0 = f32x16{};
        c1 = f32x16{};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            c0 = mfma16<TY>(a[s], b0[s], c0);
            c1 = mfma16<TY>(a[s], b1[s], c1);
        }
    };
    auto update = [&](f32x16 &c0, f32x16 &c1, uint32_t cb) __attribute__((always_inline)) {
        if (self && (cb == (uint32_t)q0 || cb == (uint32_t)q0 + 32)) {
            self_mask<L>(c0, cb, half, qa);
            self_mask<L>(c1, cb, half, qb);
        }
        if constexpr (BUF > 0) {
            // one vote for both query groups: in the common case (nothing below either
            // lane's threshold) both list updates are skipped by a single branch
            int32_t m0 = __float_as_int(c0[0]), m1 = __float_as_int(c1[0]);
#pragma unroll
            for (int r = 1; r < 16; ++r) {
                m0 = min(m0, __float_as_int(c0[r]));
                m1 = min(m1, __float_as_int(c1[r]));
            }
            if (__builtin_expect(__builtin_amdgcn_ballot_w64((m0 < __float_as_int(d0[L - 1])) |
                                                             (m1 < __float_as_int(d1[L - 1]))) == 0,
                                 1))
                return;
            topk_update_buf<L, BUF>(d0, i0, c0, cb, half, buf0, cnt0);
            topk_update_buf<L, BUF>(d1, i1, c1, cb, half, buf1, cnt1);
        } else {
            topk_update<L>(d0, i0, c0, cb, half);
            topk_update<L>(d1, i1, c1, cb, half);
        }
    };
    int64_t n_stages = (nc + kPadRows - 1) / kPadRows * (kPadRows / (kTC * TB));
    // Stages are visited outward from the one holding the middle of this
    // workgroup's queries (h, h+1, h-1, h+2, ...), so in the spatial order the
    // nearest candidates come first and the lists tighten early.  For a separate
    // query set, that is where the middle query's key falls among the sorted
    // candidate keys.  The workgroups that run together on one XCD (32 CUs, G at a
    // time, dispatched in order) share one home: the middle of their joint query
    // range, so they walk the same stages in step and read them through the L2 once.
    constexpr int64_t G = 32 * 4 * screen_waves<KC, L>() / NW;
    const int64_t jx = blockIdx.x / 8, cnt = per + (xcd < rem ? 1 : 0), g0 = (jx / G) * G;
    const int64_t gn = cnt - g0 < G ? cnt - g0 : G;
    int64_t mid = (wg - (jx - g0)) * NT + gn * NT / 2;
    if (keys_q) {
        mid = mid < nq ? mid : nq - 1;
        const uint32_t key = keys_q[mid];
        int64_t lo = 0, hi = nc;
        while (lo < hi) {
            const int64_t m2 = (lo + hi) >> 1;
            if (keys_c[m2] < key) lo = m2 + 1;
            else hi = m2;
        }
        mid = lo;
    }
    int64_t st0 = 0;  // first stage of the candidate range
    if (sg.order) {   // batch: the candidates are the rows of this block's subset
        const int sub = sg.blk_seg[wg * NT / kSegBlock];
        const int64_t r0 = sg.row0[sub], nv = sg.nrow[sub];
        st0 = r0 / (kTC * TB);
        n_stages = (nv + kTC * TB - 1) / (kTC * TB);
        mid = wg * NT + NT / 2;
        mid = mid < r0 ? r0 : (mid > r0 + nv - 1 ? r0 + nv - 1 : mid);
    }
    const int64_t home = mid / (kTC * TB);
    const int64_t h = home < st0 + n_stages ? home : st0 + n_stages - 1;
    const int64_t below = h - st0, above = st0 + n_stages - 1 - h, m = below < above ? below : above;
    auto stage_at = [&](int64_t k) -> int64_t {
        if (k <= 2 * m) return (k & 1) ? h + (k + 1) / 2 : h - k / 2;
        return above > below ? h + (k - m) : h - (k - m);
    };
    f32x16 x0, x1, y0, y1;  // even fragments in x, odd in y
    uint32_t cy = 0;        // first candidate of the fragment held in y
    // LDS holds visit t (being computed) and receives visit t+1 from registers loaded
    // one iteration earlier: after the MFMAs of each fragment the workgroup stores its
    // share of those registers (PER / NF units per thread) and reloads them with visit
    // t+2, so the stage traffic runs beside the matrix work instead of in a burst at
    // the barrier, and each load has a whole iteration to arrive.  (The final
    // iterations re-load the last stage; those copies are never read.)
    u32x4 pre[PER];
    auto move = [&](u32x4 *dst, const u32x4 *src, int f) __attribute__((always_inline)) {
#pragma unroll
        for (int m = f * PER / NF; m < (f + 1) * PER / NF; ++m) {
            dst[threadIdx.x + m * NT] = pre[m];
            pre[m] = src[threadIdx.x + m * NT];
        }
    };
    auto compute = [&](const u32x4 *stage, int64_t t, u32x4 *dst, const u32x4 *src) __attribute__((always_inline)) {
        const uint32_t c0 = (uint32_t)(stage_at(t) * kTC * TB);
        if constexpr (screen_pipe<KC, L>()) {
            u32x4 ae[KS], ao[KS];  // operands of the even / odd fragment
            lda(stage, 0, ae);
#pragma unroll
            for (int f = 0; f < NF; f += 2) {
                lda(stage, f + 1, ao);
                mmaa(ae, x0, x1);
                move(dst, src, f);
                if (f > 0 || t > 0) update(y0, y1, cy);
                if (f + 2 < NF) lda(stage, f + 2, ae);
                mmaa(ao, y0, y1);
                move(dst, src, f + 1);
                cy = c0 + (uint32_t)((f + 1) * 32);
                update(x0, x1, c0 + (uint32_t)(f * 32));
            }
        } else {
#pragma unroll
            for (int f = 0; f < NF; ++f) {
                mma(stage, f, x0, x1);
                move(dst, src, f);
                update(x0, x1, c0 + (uint32_t)(f * 32));
            }
        }
    };
    // visit 1 into registers, then visit 0 into LDS: the stores of visit 0 wait for
    // every load before them, so the loop starts with no load in flight and its waits
    // count only the loop's own loads (reloaded in unit order, one unit per fragment)
    {
        const u32x4 *second = xa + stage_at(n_stages > 1 ? 1 : 0) * TU;
#pragma unroll
        for (int m = 0; m < PER; ++m) pre[m] = second[threadIdx.x + m * NT];
    }
    {
        const u32x4 *first = xa + stage_at(0) * TU;
        for (int i = threadIdx.x; i < TU; i += NT) lds[0][i] = first[i];
    }
    __syncthreads();
    for (int64_t t = 0; t < n_stages; ++t) {
        const int64_t tn = t + 2 < n_stages ? t + 2 : n_stages - 1;
        compute(lds[t & 1], t, lds[(t & 1) ^ 1], xa + stage_at(tn) * TU);
        __syncthreads();
    }
    if constexpr (screen_pipe<KC, L>()) update(y0, y1, cy);
    if constexpr (BUF > 0) {
        topk_flush<L, BUF>(d0, i0, buf0, cnt0);
        topk_flush<L, BUF>(d1, i1, buf1, cnt1);
    }
    write_lists<L>(d0, i0, qa, half, cand_d, cand_i);
    write_lists<L>(d1, i1, qb, half, cand_d, cand_i);
}

// ---------------------------------------------------------------- screen, FP32
// v_mfma_f32_32x32x2_f32: per 16-byte chunk, lane half h supplies floats 2h, 2h+1 in
// two MFMAs, so the chunk's four K values are covered by the two half-waves.
template <int KC, int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((L <= 16 && KC <= 16) ? 2 : 1))) void k_knn_screen_f32(const float2 *__restrict__ xa, const float2 *__restrict__ xb,
                                                        int64_t nc_pad, int self, float *cand_d, uint32_t *cand_i,
                                                        KnnSeg sg) {
    constexpr int TU = KC * 2 * kTC;  // float2 units per tile
    static_assert(TU % 256 == 0, "tile must split evenly over 256 threads");
    constexpr int PER = TU / 256;
    __shared__ float2 lds[2][TU];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, col = lane & 31, half = lane >> 5;
    int64_t wg = blockIdx.x;
    if (sg.order) wg = (int64_t)sg.order[blockIdx.x / (kPadRows / 256)] * (kPadRows / 256) + blockIdx.x % (kPadRows / 256);
    const int64_t q0 = wg * 256 + wave * kQW;
    const uint32_t qa = (uint32_t)(q0 + col), qb = (uint32_t)(q0 + 32 + col);
    const float2 *qbase = xb + (q0 / 64) * KC * 2 * 64;
    float2 b0[KC], b1[KC];
#pragma unroll
    for (int j = 0; j < KC; ++j) {
        b0[j] = qbase[(2 * j + half) * 64 + col];
        b1[j] = qbase[(2 * j + half) * 64 + 32 + col];
    }
    float d0[L], d1[L];
    uint32_t i0[L], i1[L];
#pragma unroll
    for (int j = 0; j < L; ++j) {
        d0[j] = d1[j] = INFINITY;
        i0[j] = i1[j] = 0xFFFFFFFFu;
    }
    int64_t n_tiles = nc_pad / kTC, t0 = 0;
    if (sg.order) {  // batch: the candidates are the rows of this block's subset
        const int sub = sg.blk_seg[wg * 256 / kSegBlock];
        t0 = sg.row0[sub] / kTC;
        n_tiles = (sg.nrow[sub] + kTC - 1) / kTC;
    }
    xa += t0 * TU;
    // same loop-carried prefetch as the 16-bit screen
    for (int i = threadIdx.x; i < TU; i += 256) lds[0][i] = xa[i];
    Regs<float2, PER> pre;
    regs_load<float2, PER, 256>(pre, xa + (n_tiles > 1 ? 1 : 0) * TU, threadIdx.x);
    __syncthreads();
    for (int64_t t = 0; t < n_tiles; ++t) {
        const int buf = (int)(t & 1);
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            f32x16 a0 = {}, a1 = {};
#pragma unroll
            for (int j = 0; j < KC; ++j) {
                const float2 a = lds[buf][(2 * j + half) * kTC + mt * 32 + col];
                a0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b0[j].x, a0, 0, 0, 0);
                a0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b0[j].y, a0, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b1[j].x, a1, 0, 0, 0);
                a1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b1[j].y, a1, 0, 0, 0);
            }
            const uint32_t cb = (uint32_t)((t0 + t) * kTC + mt * 32);
            if (self && (cb == (uint32_t)q0 || cb == (uint32_t)q0 + 32)) {
                self_mask<L>(a0, cb, half, qa);
                self_mask<L>(a1, cb, half, qb);
            }
            topk_update<L>(d0, i0, a0, cb, half);
            topk_update<L>(d1, i1, a1, cb, half);
        }
        regs_store<float2, PER, 256>(pre, lds[buf ^ 1], threadIdx.x);
        const int64_t tn = t + 2 < n_tiles ? t + 2 : n_tiles - 1;
        regs_load<float2, PER, 256>(pre, xa + tn * TU, threadIdx.x);
        __syncthreads();
    }
    write_lists<L>(d0, i0, qa, half, cand_d, cand_i);
    write_lists<L>(d1, i1, qb, half, cand_d, cand_i);
}

template <int L>
static bool launch_screen_L(int prec, const void *xa, const void *xb, int64_t nc, int64_t nq, int kpad, int self,
                            const uint32_t *keys_c, const uint32_t *keys_q, float *cd, uint32_t *ci,
                            hipStream_t s, const KnnSeg &sg) {
    const int64_t nq_pad = knn_rows_padded(nq), nc_pad = knn_rows_padded(nc);
#define CREST_X16(TY, KCV)                                                                                   \
    case KCV: {                                                                                             \
        constexpr int NW = screen_nw<TY, KCV, L>();                                                         \
        hipLaunchKernelGGL((k_knn_screen16<TY, KCV, L, NW, screen_tb<KCV>(), screen_buf<KCV, L>()>),         \
                           dim3((unsigned)(nq_pad / (NW * 64))), dim3(NW * 64), 0, s, (const u32x4 *)xa,     \
                           (const u32x4 *)xb, nc, nq, keys_c, keys_q, self, cd, ci, sg);                    \
        return true;                                                                                        \
    }
#define CREST_F(KCV)                                                                                          \
    case KCV:                                                                                                 \
        hipLaunchKernelGGL((k_knn_screen_f32<KCV, L>), dim3((unsigned)(nq_pad / 256)), dim3(256), 0, s,      \
                           (const float2 *)xa, (const float2 *)xb, nc_pad, self, cd, ci, sg);                \
        return true;
    if (prec == kPrecBF16X3) {
        switch (kpad / 8) {
            CREST_X16(0, 4) CREST_X16(0, 8) CREST_X16(0, 12) CREST_X16(0, 16) CREST_X16(0, 20) CREST_X16(0, 24)
            CREST_X16(0, 32) CREST_X16(0, 52)
        }
    } else if (prec == kPrecFP16) {
        switch (kpad / 8) {
            CREST_X16(1, 4) CREST_X16(1, 8) CREST_X16(1, 16) CREST_X16(1, 24) CREST_X16(1, 32)
        }
    } else {
        switch (kpad / 4) { CREST_F(2) CREST_F(4) CREST_F(8) CREST_F(14) CREST_F(16) CREST_F(24) CREST_F(34) }
    }
#undef CREST_X16
#undef CREST_F
    return false;
}

// Screen of nq queries against nc candidates (self != 0: the same rows, each
// query's own row excluded) with lists of L per lane (L in {16, 32, 64}, L >= k);
// cand arrays hold 2L per query for knn_rows_padded(nq) rows.  Returns false when no
// kernel is instantiated for (prec, kpad, L).
bool launch_knn_screen(int prec, const void *xa, const void *xb, int64_t nc, int64_t nq, int kpad, int L, int self,
                       const KnnOrder &o, float *cand_d, uint32_t *cand_i, hipStream_t s) {
    const KnnSeg sg{};
    if (L == 16) return launch_screen_L<16>(prec, xa, xb, nc, nq, kpad, self, o.keys_c, o.keys_q, cand_d, cand_i, s, sg);
    if (L == 32) return launch_screen_L<32>(prec, xa, xb, nc, nq, kpad, self, o.keys_c, o.keys_q, cand_d, cand_i, s, sg);
    if (L == 64) return launch_screen_L<64>(prec, xa, xb, nc, nq, kpad, self, o.keys_c, o.keys_q, cand_d, cand_i, s, sg);
    return false;
}

// Batch: n_pad operand rows holding every subset's rows (subset y at row0[y], n_y
// real rows, padded to kPadRows); each query block screens only its own subset.
bool launch_knn_screen_batch(int prec, const void *xa, const void *xb, int64_t n_pad, int kpad, int L,
                             const KnnSeg &sg, float *cand_d, uint32_t *cand_i, hipStream_t s) {
    if (L == 16) return launch_screen_L<16>(prec, xa, xb, n_pad, n_pad, kpad, 1, nullptr, nullptr, cand_d, cand_i, s, sg);
    if (L == 32) return launch_screen_L<32>(prec, xa, xb, n_pad, n_pad, kpad, 1, nullptr, nullptr, cand_d, cand_i, s, sg);
    if (L == 64) return launch_screen_L<64>(prec, xa, xb, n_pad, n_pad, kpad, 1, nullptr, nullptr, cand_d, cand_i, s, sg);
    return false;
}

// ---------------------------------------------------------------- rerank
// knn.rs:48-66: 8 f32 partial sums, summed in order, then the tail; mul and add
// rounded separately.
__device__ inline float ref_sq_dist(const float *a, const float *b, int p) {
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const int m = p - p % 8;
    for (int i = 0; i < m; i += 8)
#pragma unroll
        for (int l = 0; l < 8; ++l) {
            const float t = a[i + l] - b[i + l];
            const float sq = t * t;
            acc[l] = acc[l] + sq;
        }
    float s = acc[0];
#pragma unroll
    for (int l = 1; l < 8; ++l) s = s + acc[l];
    for (int i = m; i < p; ++i) {
        const float t = a[i] - b[i];
        const float sq = t * t;
        s = s + sq;
    }
    return s;
}

// Screen error coefficients (the bound of KnnCert, in the operands' scaled units).
// K = kpad terms per dot product accumulate in FP32 on MFMA: at most 2 gamma_K of
// the sum of their magnitudes (allowing a non-IEEE accumulation order), which is at
// most (a + b)^2; the squared norms are FP32 sums of p squares (gamma_p), and the
// centring x - mean rounds once (2 u32 (a + b)^2 on the distance); together c2.
//   FP16: h = fp16(v) = v + e with |e| <= u|v| + eta (u = 2^-11, eta = 2^-25 the
//     subnormal half-spacing), products exact in FP32: |2 sum(h_c h_q - v_c v_q)|
//     <= 2(2u + u^2) a b + 2 eta (1 + u) sqrt(p) (a + b) + 2 p eta^2; the norm split
//     nh + nl loses at most u^2 nn + eta per norm.
//   split-BF16: v = h + l + r with |r| <= 2^-16 |v|; the missing l l, (h + l) r, r (h + l)
//     and r r terms total at most 3.01 * 2^-16 |v_c v_q| per coordinate (doubled); the
//     norm split loses at most 2^-16 nn.
//   FP32: operands are the centred FP32 values themselves.
// CREST_KNN_CERT_SCALE (tests) multiplies the bound, so that the recheck paths run on
// any data.
double knn_cert_scale() {
    const char *e = std::getenv("CREST_KNN_CERT_SCALE");
    const double v = e ? std::atof(e) : 1.0;
    return v > 0.0 ? v : 1.0;
}

static KnnCert knn_cert_coefs(int prec, int p, int kpad, double scale);

KnnCert knn_cert(int prec, int p, int kpad, const float *cand_d, uint32_t *count, uint32_t *list, int64_t cap) {
    KnnCert c = knn_cert_coefs(prec, p, kpad, knn_cert_scale());
    c.cand_d = cand_d;
    c.count = count;
    c.list = list;
    c.cap = cap;
    return c;
}

static KnnCert knn_cert_coefs(int prec, int p, int kpad, double scale) {
    KnnCert c{};
    const double u32 = std::ldexp(1.0, -24);
    c.c2 = (2.0 * kpad + p + 16) * u32 * 1.01;
    c.c4 = 1e-30;
    if (prec == kPrecFP16) {
        const double u = std::ldexp(1.0, -11), eta = std::ldexp(1.0, -25);
        c.c1 = 2.0 * (2.0 * u + u * u) * 1.001;
        c.c2 += u * u;
        c.c3 = 2.0 * eta * (1.0 + u) * std::sqrt((double)p) * 1.001;
        c.c4 += 2.0 * p * eta * eta + 2.0 * eta;
    } else if (prec == kPrecBF16X3) {
        const double u2 = std::ldexp(1.0, -16);
        c.c1 = 2.0 * 3.01 * u2;
        c.c2 += u2;
    }
    c.c1 *= scale;
    c.c2 *= scale;
    c.c3 *= scale;
    c.c4 *= scale;
    return c;
}

// The GEMM screen of the large-k path: key = fl(fl(|q|^2 + |c|^2) - 2 fl(q.c)) from
// the rows centred on the candidates' mean (FP32; a = |q'|, b = |c'|, s = 1): the squared norms are sequential FP32
// sums ((p + 1) u32 each), the SGEMM dot at most 2 gamma_p |q||c| <= gamma_p (a + b)^2 / 2
// in any accumulation order, and the two final roundings u32 (a + b)^2 each.
static double knn_gemm_c2(int p) { return (2.0 * p + 16) * std::ldexp(1.0, -24) * 1.01; }

KnnCert knn_cert_gemm(int p, uint32_t *count) {
    KnnCert c{};
    c.count = count;
    const double k = knn_cert_scale();
    c.c2 = knn_gemm_c2(p) * k;
    c.c4 = 1e-30 * k;
    return c;
}

// The certificate of query position i (one wave; dk = the k-th smallest reranked
// FP64 squared distance, INFINITY when fewer than k candidates were valid).
__device__ inline void knn_certify(const KnnCert &ct, const float *xq, int p, const float *mean,
                                   const uint32_t *maxbits, int64_t i, int n_cand, double dk, uint32_t row,
                                   int lane) {
    double part = 0.0;
    for (int f = lane; f < p; f += 64) {
        const double v = (double)xq[f] - (double)mean[f];
        part += v * v;
    }
    const double qq = wave_sum_f64(part);
    if (lane != 0) return;
    const float *cd = ct.cand_d + i * (int64_t)n_cand;
    const int L = n_cand / 2;
    const double t = fmin((double)cd[L - 1], (double)cd[n_cand - 1]);
    // a half-list that is not full holds every candidate of its half
    bool ok = !(t < (double)INFINITY);
    if (!ok && dk < (double)INFINITY) {
        const double s = maxbits ? (double)fp16_prescale(maxbits) : 1.0;
        const double a = s * sqrt(qq) * (1.0 + 1e-6);
        const double b = a + s * sqrt(dk) * (1.0 + 1e-6);
        const double e = ct.c1 * a * b + ct.c2 * (a + b) * (a + b) + ct.c3 * (a + b) + ct.c4;
        ok = t - e > s * s * dk * (1.0 + 1e-12);
    }
    if (!ok) {
        const uint32_t at = atomicAdd(ct.count, 1u);
        if ((int64_t)at < ct.cap) {
            ct.list[at] = row;
            if (ct.pos) ct.pos[at] = (uint32_t)i;
        }
    }
}

__device__ inline double wave_max_f64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// One wave per query.  Candidates are ranked by their exact FP64 squared distance
// (ties by index); the k best are reported with the reference's FP32 distance and
// ordered by (distance, index) as knn.rs:356-367 does.
template <int NCL>
__global__ __launch_bounds__(256) void k_knn_rerank(const float *__restrict__ xqs, int64_t nq,
                                                    const float *__restrict__ x, int64_t n, int p,
                                                    const uint32_t *__restrict__ perm_q,
                                                    const uint32_t *__restrict__ perm, int self,
                                                    const uint32_t *__restrict__ cand_i, int n_cand, int k,
                                                    uint32_t *idx_out, float *dist_out, const int32_t *blk_seg,
                                                    const int64_t *cell_off, KnnCert ct, const uint32_t *qmap,
                                                    const float *mean, const uint32_t *maxbits) {
    const int lane = threadIdx.x & 63;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;  // sorted query position
    if (i >= nq) return;
    const int64_t q = perm_q[i];
    if (q == 0xFFFFFFFFu) return;  // padding row of a batch
    // batch: indices are reported within the query's subset
    const int y = blk_seg ? blk_seg[i / kSegBlock] : 0;
    const uint32_t ibase = blk_seg ? (uint32_t)cell_off[y] : 0u;
    // a recheck answers query qmap[q]: its output row, and the row a graph excludes
    const uint32_t qo = qmap ? qmap[q] : (uint32_t)q;
    const float *xq = xqs + q * (int64_t)p;
    const uint32_t *ci = cand_i + i * (int64_t)n_cand;
    double dd[NCL];
    uint32_t cc[NCL];
    bool vv[NCL];
#pragma unroll
    for (int s = 0; s < NCL; ++s) {
        dd[s] = INFINITY;
        cc[s] = 0xFFFFFFFFu;
        vv[s] = false;
    }
    for (int j = 0; j < n_cand; ++j) {
        const uint32_t cpos = ci[j];
        bool valid = (int64_t)cpos < n && !(self && !qmap && (int64_t)cpos == i);
        const uint32_t c = valid ? perm[cpos] : 0xFFFFFFFFu;
        valid = valid && c != 0xFFFFFFFFu && !(self && qmap && c == qo);
        double part = 0.0;
        if (valid) {
            const float *xc = x + (int64_t)c * p;
            for (int f = lane; f < p; f += 64) {
                const double df = (double)xq[f] - (double)xc[f];
                part += df * df;
            }
        }
        const double sum = wave_sum_f64(part);
        if ((j & 63) == lane) {
            dd[j >> 6] = valid ? sum : INFINITY;
            cc[j >> 6] = c;
            vv[j >> 6] = valid;
        }
    }
    // rank by (d64, index) among valid candidates
    int rank[NCL];
#pragma unroll
    for (int s = 0; s < NCL; ++s) rank[s] = 0;
    for (int j = 0; j < n_cand; ++j) {
        const double dj = __shfl(dd[j >> 6], j & 63, 64);
        const uint32_t cj = __shfl(cc[j >> 6], j & 63, 64);
        const bool vj = __shfl((int)vv[j >> 6], j & 63, 64) != 0;
#pragma unroll
        for (int s = 0; s < NCL; ++s)
            if (vj && (dj < dd[s] || (dj == dd[s] && cj < cc[s]))) ++rank[s];
    }
    bool sel[NCL];
    float df[NCL];
#pragma unroll
    for (int s = 0; s < NCL; ++s) {
        sel[s] = vv[s] && rank[s] < k && (s * 64 + lane) < n_cand;
        df[s] = sel[s] ? sqrtf(ref_sq_dist(xq, x + (int64_t)cc[s] * p, p)) : INFINITY;
    }
    int rank2[NCL];
#pragma unroll
    for (int s = 0; s < NCL; ++s) rank2[s] = 0;
    for (int j = 0; j < n_cand; ++j) {
        const float fj = __shfl(df[j >> 6], j & 63, 64);
        const uint32_t cj = __shfl(cc[j >> 6], j & 63, 64);
        const bool sj = __shfl((int)sel[j >> 6], j & 63, 64) != 0;
#pragma unroll
        for (int s = 0; s < NCL; ++s)
            if (sj && (fj < df[s] || (fj == df[s] && cj < cc[s]))) ++rank2[s];
    }
#pragma unroll
    for (int s = 0; s < NCL; ++s)
        if (sel[s]) {
            idx_out[(int64_t)qo * k + rank2[s]] = cc[s] - ibase;
            dist_out[(int64_t)qo * k + rank2[s]] = df[s];
        }
    if (ct.count) {
        double dk = -(double)INFINITY;
        int ns = 0;
#pragma unroll
        for (int s = 0; s < NCL; ++s) {
            dk = sel[s] ? fmax(dk, dd[s]) : dk;
            ns += __popcll(__builtin_amdgcn_ballot_w64(sel[s]));
        }
        dk = ns < k ? (double)INFINITY : wave_max_f64(dk);
        knn_certify(ct, xq, p, mean + (int64_t)y * p, maxbits ? maxbits + y : nullptr, i, n_cand, dk, qo, lane);
    }
}

// The same selection with LPC lanes per candidate (64/LPC candidate slots, one wave
// per query).  Lane `sub` of a candidate's group reads features sub*W .. sub*W+W-1
// (W = 8/LPC) of every 8-block, so a group loads whole 32-byte runs of the row and
// holds exactly the reference's FP32 partial sums sub*W .. sub*W+W-1; the FP32
// distance is then summed in the reference's order after a gather of the 8 partial
// sums.  The FP64 distance of the same features is summed by a butterfly over the
// group, which leaves the same value in every lane of it.
template <int LPC>
__global__ __launch_bounds__(256) void k_knn_rerank_grp(const float *__restrict__ xqs, int64_t nq,
                                                        const float *__restrict__ x, int64_t n, int p,
                                                        const uint32_t *__restrict__ perm_q,
                                                        const uint32_t *__restrict__ perm, int self,
                                                        const uint32_t *__restrict__ cand_i, int n_cand, int k,
                                                        uint32_t *idx_out, float *dist_out, const int32_t *blk_seg,
                                                        const int64_t *cell_off, KnnCert ct, const uint32_t *qmap,
                                                        const float *mean, const uint32_t *maxbits) {
    constexpr int W = 8 / LPC;
    constexpr int S = 64 / LPC;
    const int lane = threadIdx.x & 63;
    const int slot = lane / LPC, sub = lane % LPC;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;  // sorted query position
    if (i >= nq) return;  // uniform across the wave
    const int64_t q = perm_q[i];
    if (q == 0xFFFFFFFFu) return;  // padding row of a batch
    const int y = blk_seg ? blk_seg[i / kSegBlock] : 0;
    const uint32_t ibase = blk_seg ? (uint32_t)cell_off[y] : 0u;
    const uint32_t qo = qmap ? qmap[q] : (uint32_t)q;
    const float *xq = xqs + q * (int64_t)p;
    const bool has = slot < n_cand;
    const uint32_t cpos = has ? cand_i[i * (int64_t)n_cand + slot] : 0xFFFFFFFFu;
    bool valid = has && (int64_t)cpos < n && !(self && !qmap && (int64_t)cpos == i);
    const uint32_t c = valid ? perm[cpos] : 0xFFFFFFFFu;
    valid = valid && c != 0xFFFFFFFFu && !(self && qmap && c == qo);
    const float *xc = x + (int64_t)(valid ? c : 0u) * p;
    const int m = p - p % 8;
    float acc[W];
#pragma unroll
    for (int w = 0; w < W; ++w) acc[w] = 0.f;
    double part = 0.0;
    if (valid) {
        for (int b = sub * W; b < m; b += 8)
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const float a = xq[b + w], e = xc[b + w];
                const float t = a - e;
                const float sq = t * t;
                acc[w] = acc[w] + sq;
                const double dd = (double)a - (double)e;
                part += dd * dd;
            }
        for (int f = m + sub; f < p; f += LPC) {
            const double dd = (double)xq[f] - (double)xc[f];
            part += dd * dd;
        }
    }
#pragma unroll
    for (int o = 1; o < LPC; o <<= 1) part += __shfl_xor(part, o, 64);
    // acc[0] + acc[1] + ... + acc[7] in order, then the tail (knn.rs:48-66)
    const int base = lane - sub;
    float s32 = 0.f;
#pragma unroll
    for (int j = 0; j < LPC; ++j)
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const float v = __shfl(acc[w], base + j, 64);
            s32 = (j == 0 && w == 0) ? v : s32 + v;
        }
    if (valid)
        for (int f = m; f < p; ++f) {
            const float t = xq[f] - xc[f];
            const float sq = t * t;
            s32 = s32 + sq;
        }
    const double d = valid ? part : (double)INFINITY;
    const int ns = n_cand < S ? n_cand : S;
    int rank = 0;
    for (int s = 0; s < ns; ++s) {
        const double ds = __shfl(d, s * LPC, 64);
        const uint32_t cs = __shfl(c, s * LPC, 64);
        const int vs = __shfl((int)valid, s * LPC, 64);
        if (vs && (ds < d || (ds == d && cs < c))) ++rank;
    }
    const bool sel = valid && rank < k;
    const float dfl = sel ? sqrtf(s32) : INFINITY;
    int rank2 = 0;
    for (int s = 0; s < ns; ++s) {
        const float fs = __shfl(dfl, s * LPC, 64);
        const uint32_t cs = __shfl(c, s * LPC, 64);
        const int ss = __shfl((int)sel, s * LPC, 64);
        if (ss && (fs < dfl || (fs == dfl && cs < c))) ++rank2;
    }
    if (sel && sub == 0) {
        idx_out[(int64_t)qo * k + rank2] = c - ibase;
        dist_out[(int64_t)qo * k + rank2] = dfl;
    }
    if (ct.count) {
        const bool one = sel && sub == 0;
        const int ns = __popcll(__builtin_amdgcn_ballot_w64(one));
        const double dk = ns < k ? (double)INFINITY : wave_max_f64(one ? d : -(double)INFINITY);
        knn_certify(ct, xq, p, mean + (int64_t)y * p, maxbits ? maxbits + y : nullptr, i, n_cand, dk, qo, lane);
    }
}

void launch_knn_rerank(const float *xq, int64_t nq, const float *xc, int64_t nc, int p, const KnnOrder &o, int self,
                       const uint32_t *cand_i, int n_cand, int k, uint32_t *idx_out, float *dist_out,
                       hipStream_t s, const int32_t *blk_seg, const int64_t *cell_off, const KnnCert *cert,
                       const uint32_t *qmap) {
    const int g = (int)ceil_div(nq, 4);  // one wave per query, four per workgroup
    const KnnCert ct = cert ? *cert : KnnCert{};
    const float *mean = o.mean;
    const uint32_t *maxbits = o.maxbits;
    if (n_cand <= 16)
        hipLaunchKernelGGL(k_knn_rerank_grp<4>, dim3(g), dim3(256), 0, s, xq, nq, xc, nc, p, o.perm_q, o.perm_c, self,
                           cand_i, n_cand, k, idx_out, dist_out, blk_seg, cell_off, ct, qmap, mean,
                           maxbits);
    else if (n_cand <= 32)
        hipLaunchKernelGGL(k_knn_rerank_grp<2>, dim3(g), dim3(256), 0, s, xq, nq, xc, nc, p, o.perm_q, o.perm_c, self,
                           cand_i, n_cand, k, idx_out, dist_out, blk_seg, cell_off, ct, qmap, mean,
                           maxbits);
    else if (n_cand <= 64)
        hipLaunchKernelGGL(k_knn_rerank<1>, dim3(g), dim3(256), 0, s, xq, nq, xc, nc, p, o.perm_q, o.perm_c, self,
                           cand_i, n_cand, k, idx_out, dist_out, blk_seg, cell_off, ct, qmap, mean,
                           maxbits);
    else
        hipLaunchKernelGGL(k_knn_rerank<2>, dim3(g), dim3(256), 0, s, xq, nq, xc, nc, p, o.perm_q, o.perm_c, self,
                           cand_i, n_cand, k, idx_out, dist_out, blk_seg, cell_off, ct, qmap, mean,
                           maxbits);
}

// out[r] = x[rows[r]] (p floats each): the query rows of a recheck.
__global__ void k_gather_rows(const float *x, int p, const uint32_t *rows, int64_t m, float *out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * p) return;
    const int64_t r = e / p;
    out[e] = x[(int64_t)rows[r] * p + (e - r * p)];
}

void launch_gather_rows(const float *x, int p, const uint32_t *rows, int64_t m, float *out, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)ceil_div(m * p, 256)), dim3(256), 0, s, x, p, rows, m, out);
}

// Row r's k1 nearest (ti/td, in output order) written as output row rows[r]: with
// `self`, the entry of row rows[r] itself is dropped and the first k others kept.
__global__ void k_knn_place(const uint32_t *ti, const float *td, int64_t m, int k1, int k, const uint32_t *rows,
                            int self, uint32_t *idx_out, float *dist_out, const uint32_t *ibase) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const uint32_t q = rows[r];
    const uint32_t base = ibase ? ibase[r] : 0u;  // indices reported within the query's subset
    int o = 0;
    for (int j = 0; j < k1 && o < k; ++j) {
        const uint32_t c = ti[r * k1 + j];
        if (self && c == q) continue;
        idx_out[(int64_t)q * k + o] = c - base;
        dist_out[(int64_t)q * k + o] = td[r * k1 + j];
        ++o;
    }
}

void launch_knn_place(const uint32_t *ti, const float *td, int64_t m, int k1, int k, const uint32_t *rows, int self,
                      uint32_t *idx_out, float *dist_out, hipStream_t s, const uint32_t *ibase) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_knn_place, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, ti, td, m, k1, k, rows, self,
                       idx_out, dist_out, ibase);
}

// Threshold scan of a recheck, on rows centred on the candidates' mean: the key of a
// pair is the GEMM screen's, fl(fl(|q'|^2 + |c'|^2) - 2 fl(q'.c')) with the dot an FP32
// fma chain (so it is within knn_cert_gemm's bound of |q - c|^2), and every candidate
// whose key is at most thr[q] is appended to query q's list (the query's own row
// excluded when self_rows is given).  One workgroup per (64-query block, candidate
// slice); the query rows sit in LDS (read as broadcasts), each thread takes every
// 256th candidate of the slice and keeps the 64 dot products in registers.  The LDS
// rows take 64 * 4 * (p rounded up to 8) bytes: at most 64 KiB for p <= 256.
constexpr int kScanQ = 64;
constexpr int64_t kMaxGridY = 65535;  // the largest gridDim.y
__global__ __launch_bounds__(256) void k_knn_thr_scan(const float *__restrict__ xcc, const float *__restrict__ cn,
                                                      int64_t nc, int p, const float *__restrict__ xqc,
                                                      const float *__restrict__ qn, int64_t m,
                                                      const float *__restrict__ thr,
                                                      const uint32_t *__restrict__ self_rows, int64_t slice,
                                                      uint32_t *cnt, uint32_t *cand, int cap) {
    extern __shared__ float sq[];  // [kScanQ][pp], rows zero-padded to pp = p rounded up to 8
    __shared__ float sthr[kScanQ], sqn[kScanQ];
    __shared__ uint32_t sself[kScanQ];
    const int pp = (p + 7) / 8 * 8;
    const int64_t q0 = (int64_t)blockIdx.y * kScanQ;
    const int nqb = (int)(m - q0 < kScanQ ? m - q0 : kScanQ);
    for (int e = threadIdx.x; e < kScanQ * pp; e += blockDim.x) {
        const int j = e / pp, f = e - j * pp;
        sq[e] = (j < nqb && f < p) ? xqc[(q0 + j) * p + f] : 0.0f;
    }
    if (threadIdx.x < kScanQ) {
        const int j = threadIdx.x;
        sthr[j] = j < nqb ? thr[q0 + j] : -1.0f;  // keys are >= 0: nothing passes a negative threshold
        sqn[j] = j < nqb ? qn[q0 + j] : 0.0f;
        sself[j] = (j < nqb && self_rows) ? self_rows[q0 + j] : 0xFFFFFFFFu;
    }
    __syncthreads();
    const int64_t c0 = (int64_t)blockIdx.x * slice, c1 = c0 + slice < nc ? c0 + slice : nc;
    for (int64_t c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
        const float *cr = xcc + c * p;
        float acc[kScanQ];
#pragma unroll
        for (int j = 0; j < kScanQ; ++j) acc[j] = 0.0f;
        for (int f0 = 0; f0 < pp; f0 += 8) {
            float cv[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) cv[t] = f0 + t < p ? cr[f0 + t] : 0.0f;
#pragma unroll
            for (int j = 0; j < kScanQ; ++j)
#pragma unroll
                for (int t = 0; t < 8; ++t) acc[j] = fmaf(sq[j * pp + f0 + t], cv[t], acc[j]);
        }
        const float cnc = cn[c];
#pragma unroll
        for (int j = 0; j < kScanQ; ++j) {
            const float sn = sqn[j] + cnc;
            const float t2 = 2.0f * acc[j];
            const float key = sn - t2;
            if (!(key <= sthr[j]) || (uint32_t)c == sself[j]) continue;
            const uint32_t at = atomicAdd(&cnt[q0 + j], 1u);
            if ((int64_t)at < cap) cand[(q0 + j) * cap + at] = (uint32_t)c;
        }
    }
}

// The same threshold test on a block of FP32 dot products from the GEMM (dots[r][c] =
// q'_(q0+r) . c', rows padded to ld, a multiple of 4): four candidates per thread.
__global__ __launch_bounds__(256) void k_knn_thr_filter(const float *__restrict__ dots, int64_t ld, int64_t nc,
                                                        const float *__restrict__ cn, const float *__restrict__ qn,
                                                        const float *__restrict__ thr,
                                                        const uint32_t *__restrict__ self_rows, int64_t q0,
                                                        uint32_t *cnt, uint32_t *cand, int cap) {
    const int64_t r = blockIdx.y, q = q0 + r;
    const int64_t c = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (c >= nc) return;
    const float4 d = *reinterpret_cast<const float4 *>(dots + r * ld + c);
    const float dv[4] = {d.x, d.y, d.z, d.w};
    const float t = thr[q], qv = qn[q];
    const uint32_t self = self_rows ? self_rows[q] : 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int64_t cj = c + j;
        if (cj >= nc) break;
        const float sn = qv + cn[cj];
        const float t2 = 2.0f * dv[j];
        const float key = sn - t2;
        if (!(key <= t) || (uint32_t)cj == self) continue;
        const uint32_t at = atomicAdd(&cnt[q], 1u);
        if ((int64_t)at < cap) cand[q * cap + at] = (uint32_t)cj;
    }
}

void launch_knn_thr_filter(const float *dots, int64_t ld, int64_t nc, const float *cn, const float *qn,
                           const float *thr, const uint32_t *rows, int self, int64_t q0, int64_t nrows, uint32_t *cnt,
                           uint32_t *cand, int cap, hipStream_t s) {
    // gridDim.y is at most 65,535: rows in launches of that many
    for (int64_t r0 = 0; r0 < nrows; r0 += kMaxGridY) {
        const int64_t nr = std::min<int64_t>(kMaxGridY, nrows - r0);
        hipLaunchKernelGGL(k_knn_thr_filter, dim3((unsigned)ceil_div(nc, 1024), (unsigned)nr), dim3(256), 0, s,
                           dots + r0 * ld, ld, nc, cn, qn, thr, self ? rows : nullptr, q0 + r0, cnt, cand, cap);
    }
}

void launch_knn_scan_thresholds(const float *dist, const uint32_t *rows, const float *xqc, int64_t m, int k, int p,
                                float *thr, hipStream_t s);

// The threshold test on FP32 MFMA (v_mfma_f32_32x32x2_f32), rows centred as above.
// Workgroup = 4 waves x 32 queries; each wave keeps its 32 query rows as B operands in
// registers (lane l: query l % 32, K-step s: element 2s + l / 32; KS = 16, 26, 32 or
// 48 K-steps, so p <= 96 with two LDS tile buffers of 2 KS x 65 floats).  The
// workgroup streams its candidate slice in tiles of 64 rows through two LDS buffers
// laid out [k][65] (A operand of lane l: candidate l % 32 of a 32-row half, element
// 2s + l / 32); register r of lane l then holds the dot product of candidate
// 8(r/4) + 4(l/32) + r%4 of the half with query l % 32, and the key
// fl(fl(|q'|^2 + |c'|^2) - 2 dot) is tested against that query's threshold.  Rows past
// nc are zero with norm +inf (never appended); query rows past m have threshold -1.
// With a block table (a batch of subsets): blockIdx.y names block y of the table —
// queries q0 .. q0 + nq - 1 (at most 128, one subset's) against candidates c0 .. c1 - 1
// of that subset in slices of `slice` (workgroups past the subset's last slice exit).
template <int KS>  // MFMA K-steps: features 0 .. 2 KS - 1 (zero past p)
__global__ __launch_bounds__(256) void k_knn_thr_mfma(const float *__restrict__ xcc, const float *__restrict__ cn,
                                                      int64_t nc, int p, const float *__restrict__ xqc,
                                                      const float *__restrict__ qn, int64_t m,
                                                      const float *__restrict__ thr,
                                                      const uint32_t *__restrict__ self_rows, int64_t slice,
                                                      uint32_t *cnt, uint32_t *cand, int cap,
                                                      const KnnScanBlock *__restrict__ blocks) {
    constexpr int LD = 65, PP = 2 * KS, PER = PP / 4;  // 64 * PP floats per tile over 256 threads
    __shared__ float ta[2][PP * LD];
    __shared__ float tn[2][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, col = lane & 31, half = lane >> 5;
    int64_t q = (int64_t)blockIdx.y * 128 + wave * 32 + col, cb = 0, ce = nc;
    bool qreal = q < m;
    if (blocks) {
        const KnnScanBlock bk = blocks[blockIdx.y];
        if ((int64_t)blockIdx.x * bk.slice >= bk.c1 - bk.c0) return;  // the whole workgroup
        q = bk.q0 + wave * 32 + col;
        qreal = wave * 32 + col < bk.nq;
        cb = bk.c0;
        ce = bk.c1;
        slice = bk.slice;
    }
    float b[KS];
#pragma unroll
    for (int st = 0; st < KS; ++st) {
        const int f = 2 * st + half;
        b[st] = (qreal && f < p) ? xqc[q * p + f] : 0.0f;
    }
    const float qt = qreal ? thr[q] : -1.0f, qv = qreal ? qn[q] : 0.0f;
    const uint32_t qself = (qreal && self_rows) ? self_rows[q] : 0xFFFFFFFFu;
    const int64_t c0 = cb + (int64_t)blockIdx.x * slice;
    const int64_t c1 = c0 + slice < ce ? c0 + slice : ce;
    const int64_t nt = (c1 - c0 + 63) / 64;
    const float *slice_base = xcc + c0 * p;
    const int64_t slice_len = c1 - c0;
    // element e of a tile: row e / PP, feature e % PP, stored at [feature][row]
    float pre[PER];
    auto load = [&](int64_t t) {
        const float *tb = slice_base + t * 64 * p;
        const int64_t rows_left = slice_len - t * 64;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = i * 256 + threadIdx.x, r = e / PP, f = e - r * PP;
            pre[i] = (r < rows_left && f < p) ? tb[r * p + f] : 0.0f;
        }
    };
    auto store = [&](int buf, int64_t t) {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = i * 256 + threadIdx.x, r = e / PP, f = e - r * PP;
            ta[buf][f * LD + r] = pre[i];
        }
        if (threadIdx.x < 64) {
            const int64_t c = c0 + t * 64 + threadIdx.x;
            tn[buf][threadIdx.x] = c < c1 ? cn[c] : INFINITY;
        }
    };
    if (nt > 0) {
        load(0);
        store(0, 0);
        if (nt > 1) load(1);
    }
    __syncthreads();
    for (int64_t t = 0; t < nt; ++t) {
        const int buf = (int)(t & 1);
        f32x16 acc0 = {}, acc1 = {};
#pragma unroll
        for (int st = 0; st < KS; ++st) {
            const float *row = &ta[buf][(2 * st + half) * LD + col];
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(row[0], b[st], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(row[32], b[st], acc1, 0, 0, 0);
        }
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = mt * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
                const float sn = qv + tn[buf][i];
                const float t2 = 2.0f * (mt ? acc1[r] : acc0[r]);
                const float key = sn - t2;
                const uint32_t c = (uint32_t)(c0 + t * 64 + i);
                if (key <= qt && c != qself) {
                    const uint32_t at = atomicAdd(&cnt[q], 1u);
                    if ((int64_t)at < cap) cand[q * cap + at] = c;
                }
            }
        if (t + 1 < nt) store(buf ^ 1, t + 1);
        if (t + 2 < nt) load(t + 2);
        __syncthreads();
    }
}

static void launch_thr_mfma_grid(dim3 grid, const float *xcc, const float *cn, int64_t nc, int p, const float *xqc,
                                 const float *qn, int64_t m, const float *thr, const uint32_t *sr, int64_t slice,
                                 uint32_t *cnt, uint32_t *cand, int cap, const KnnScanBlock *blocks, hipStream_t s) {
    if (p <= 32)
        hipLaunchKernelGGL(k_knn_thr_mfma<16>, grid, dim3(256), 0, s, xcc, cn, nc, p, xqc, qn, m, thr, sr, slice, cnt,
                           cand, cap, blocks);
    else if (p <= 52)  // 50 principal components: 26 K-steps, no all-zero step
        hipLaunchKernelGGL(k_knn_thr_mfma<26>, grid, dim3(256), 0, s, xcc, cn, nc, p, xqc, qn, m, thr, sr, slice, cnt,
                           cand, cap, blocks);
    else if (p <= 64)
        hipLaunchKernelGGL(k_knn_thr_mfma<32>, grid, dim3(256), 0, s, xcc, cn, nc, p, xqc, qn, m, thr, sr, slice, cnt,
                           cand, cap, blocks);
    else
        hipLaunchKernelGGL(k_knn_thr_mfma<48>, grid, dim3(256), 0, s, xcc, cn, nc, p, xqc, qn, m, thr, sr, slice, cnt,
                           cand, cap, blocks);
}

void launch_knn_thr_mfma(const float *xcc, const float *cn, int64_t nc, int p, const float *xqc, const float *qn,
                         int64_t m, const float *thr, const uint32_t *rows, int self, uint32_t *cnt, uint32_t *cand,
                         int cap, hipStream_t s) {
    if (m <= 0) return;
    // about 2,048 workgroups in all: slices of at least 4,096 candidates, whole tiles
    const int64_t qblocks = ceil_div(m, 128);
    const int64_t want = std::max<int64_t>(1, 2048 / qblocks);
    const int64_t slice = round_up(std::max<int64_t>(4096, ceil_div(nc, want)), 64);
    for (int64_t b0 = 0; b0 < qblocks; b0 += kMaxGridY) {  // gridDim.y <= 65,535
        const int64_t q0 = b0 * 128, mq = std::min<int64_t>(m - q0, kMaxGridY * 128);
        const dim3 grid((unsigned)ceil_div(nc, slice), (unsigned)ceil_div(mq, 128));
        launch_thr_mfma_grid(grid, xcc, cn, nc, p, xqc + q0 * p, qn + q0, mq, thr + q0, self ? rows + q0 : nullptr,
                             slice, cnt + q0, cand + q0 * cap, cap, nullptr, s);
    }
}

void launch_knn_thr_mfma_blocks(const float *xcc, const float *cn, int p, const float *xqc, const float *qn,
                                const float *thr, const uint32_t *rows, const KnnScanBlock *blocks, int n_blocks,
                                int64_t max_slices, uint32_t *cnt, uint32_t *cand, int cap, hipStream_t s) {
    for (int64_t b0 = 0; b0 < n_blocks; b0 += kMaxGridY) {  // gridDim.y <= 65,535
        const dim3 grid((unsigned)max_slices, (unsigned)std::min<int64_t>(kMaxGridY, n_blocks - b0));
        launch_thr_mfma_grid(grid, xcc, cn, 0, p, xqc, qn, 0, thr, rows, 0, cnt, cand, cap, blocks + b0, s);
    }
}

// thr[r]: T = the square of output row rows[r]'s current k-th distance (the
// reference's f32 formula, relative error at most (p/8 + 10) u32 after the square root
// is undone) widened by (4p + 64) u32, so T >= the true k-th squared distance; then
// T + c2 (2a + sqrt(T))^2 (knn_cert_gemm's bound at a = |q'|, b <= a + sqrt(T)) rounded
// up: every candidate at most as far as the true k-th neighbour has a key below it.
__global__ void k_knn_scan_thr(const float *dist, const uint32_t *rows, const float *xqc, int64_t m, int k, int p,
                               double c2, float *thr) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const double dk = (double)dist[(int64_t)rows[r] * k + (k - 1)];
    const double T = dk * dk * (1.0 + (4.0 * p + 64.0) * 0x1p-24) + 1e-37;
    double qq = 0.0;
    for (int f = 0; f < p; ++f) qq += (double)xqc[r * p + f] * (double)xqc[r * p + f];
    const double a = sqrt(qq) * (1.0 + 1e-6);
    const double w = 2.0 * a + sqrt(T) * (1.0 + 1e-6);
    const double t = T + c2 * w * w;
    float tf = (float)t;
    if ((double)tf < t) tf = nextafterf(tf, INFINITY);
    thr[r] = tf;
}

void launch_knn_scan_thresholds(const float *dist, const uint32_t *rows, const float *xqc, int64_t m, int k, int p,
                                float *thr, hipStream_t s) {
    if (m <= 0) return;
    // the key's own bound (never the test scale of the certificate)
    hipLaunchKernelGGL(k_knn_scan_thr, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, dist, rows, xqc, m, k, p,
                       knn_gemm_c2(p), thr);
}

void launch_knn_thr_scan(const float *xcc, const float *cn, int64_t nc, int p, const float *xqc, const float *qn,
                         int64_t m, const float *dist, const uint32_t *rows, int k, int self, float *thr, uint32_t *cnt,
                         uint32_t *cand, int cap, hipStream_t s) {
    if (m <= 0) return;
    launch_knn_scan_thresholds(dist, rows, xqc, m, k, p, thr, s);
    // about 2,048 workgroups in all: slices of at least 4,096 candidates; query blocks
    // in launches of at most 65,535 (the largest gridDim.y)
    const int64_t qblocks = ceil_div(m, kScanQ);
    const int64_t want = std::max<int64_t>(1, 2048 / qblocks);
    const int64_t slice = std::max<int64_t>(4096, ceil_div(nc, want));
    const size_t lds = (size_t)kScanQ * ((p + 7) / 8 * 8) * sizeof(float);
    for (int64_t b0 = 0; b0 < qblocks; b0 += kMaxGridY) {
        const int64_t q0 = b0 * kScanQ, mq = std::min<int64_t>(m - q0, kMaxGridY * kScanQ);
        hipLaunchKernelGGL(k_knn_thr_scan, dim3((unsigned)ceil_div(nc, slice), (unsigned)ceil_div(mq, kScanQ)),
                           dim3(256), lds, s, xcc, cn, nc, p, xqc + q0 * p, qn + q0, mq, thr + q0,
                           self ? rows + q0 : nullptr, slice, cnt + q0, cand + q0 * cap, cap);
    }
}

// Thresholds of the FP16 threshold test: T as for the FP32 test (the square of the
// query's current k-th distance, widened by (4p + 64) u32, at least the true k-th
// squared distance), then s^2 T + E16(a, a + s sqrt(T)) with a = s|q - mean| (the
// screen's centring and prescale) and E16 the FP16 screen's bound (c1 .. c4, without
// any test scale), rounded up: every candidate at most as far as the true k-th
// neighbour has a screen value below it.
__global__ void k_knn_scan_thr16(const float *dist, const uint32_t *rows, const float *xq_rows, const float *mean,
                                 const uint32_t *maxbits, int64_t m, int k, int p, double c1, double c2, double c3,
                                 double c4, float *thr) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= m) return;
    const double dk = (double)dist[(int64_t)rows[r] * k + (k - 1)];
    const double T = dk * dk * (1.0 + (4.0 * p + 64.0) * 0x1p-24) + 1e-37;
    const float *xr = xq_rows + r * p;
    double qq = 0.0;
    for (int f = 0; f < p; ++f) {
        const double v = (double)xr[f] - (double)mean[f];
        qq += v * v;
    }
    const double s = maxbits ? (double)fp16_prescale(maxbits) : 1.0;
    const double a = s * sqrt(qq) * (1.0 + 1e-6);
    const double b = a + s * sqrt(T) * (1.0 + 1e-6);
    const double t = s * s * T * (1.0 + 1e-12) + c1 * a * b + c2 * (a + b) * (a + b) + c3 * (a + b) + c4;
    float tf = (float)t;
    if ((double)tf < t) tf = nextafterf(tf, INFINITY);
    thr[r] = tf;
}

// The test itself: workgroup = 4 waves x 32 queries; a wave keeps its queries' FP16
// operands (the screen's B fragments, read at their screen positions) in registers and
// reads each 64-candidate block's A fragments straight from the fragment-major
// operands (v_mfma_f32_32x32x16_f16, the screen's own arithmetic); register r of lane
// l holds candidate 8(r/4) + 4(l/32) + r%4 of a 32-row half against query l % 32.
// Padding candidates carry an infinite norm (never below a threshold); queries past m
// have threshold -inf.
template <int KC>
__global__ __launch_bounds__(256) void k_knn_thr_mfma16(const u32x4 *__restrict__ xa, const u32x4 *__restrict__ xb,
                                                        int64_t n_tiles, const uint32_t *__restrict__ pos,
                                                        const uint32_t *__restrict__ perm_c, int64_t m,
                                                        const float *__restrict__ thr, int self, int64_t slice_tiles,
                                                        uint32_t *cnt, uint32_t *cand, int cap) {
    constexpr int KS = KC / 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, col = lane & 31, half = lane >> 5;
    const int64_t r = (int64_t)blockIdx.y * 128 + wave * 32 + col;
    const bool real = r < m;
    const uint32_t qi = real ? pos[r] : 0u;
    u32x4 b[KS];
#pragma unroll
    for (int st = 0; st < KS; ++st) {
        const u32x4 z = {0u, 0u, 0u, 0u};
        b[st] = real ? xb[((int64_t)(qi / 64) * KC + 2 * st + half) * 64 + qi % 64] : z;
    }
    const float t = real ? thr[r] : -INFINITY;
    const int64_t t0 = (int64_t)blockIdx.x * slice_tiles;
    const int64_t t1 = t0 + slice_tiles < n_tiles ? t0 + slice_tiles : n_tiles;
    for (int64_t tile = t0; tile < t1; ++tile) {
        const u32x4 *base = xa + tile * KC * 64 + col;
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            u32x4 a[KS];
#pragma unroll
            for (int st = 0; st < KS; ++st) a[st] = base[(2 * st + half) * 64 + mt * 32];
            f32x16 acc = {};
#pragma unroll
            for (int st = 0; st < KS; ++st) acc = mfma16<1>(a[st], b[st], acc);
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const uint32_t cp = (uint32_t)(tile * 64 + mt * 32 + 8 * (e >> 2) + 4 * half + (e & 3));
                if (acc[e] <= t && !(self && cp == qi)) {
                    const uint32_t at = atomicAdd(&cnt[r], 1u);
                    if ((int64_t)at < cap) cand[r * cap + at] = perm_c[cp];
                }
            }
        }
    }
}

void launch_knn_thr16(const void *xa, const void *xb, int kpad, int64_t nc, const KnnOrder &o, const float *xq_rows,
                      const float *dist, const uint32_t *rows, const uint32_t *pos, int64_t m, int k, int p, int self,
                      float *thr, uint32_t *cnt, uint32_t *cand, int cap, hipStream_t s) {
    if (m <= 0) return;
    const KnnCert c = knn_cert_coefs(kPrecFP16, p, kpad, 1.0);
    hipLaunchKernelGGL(k_knn_scan_thr16, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, dist, rows, xq_rows, o.mean,
                       o.maxbits, m, k, p, c.c1, c.c2, c.c3, c.c4, thr);
    const int64_t n_tiles = knn_rows_padded(nc) / 64;
    const int64_t qblocks = ceil_div(m, 128);
    const int64_t want = std::max<int64_t>(1, 2048 / qblocks);
    const int64_t slice = std::max<int64_t>(1, ceil_div(n_tiles, want));
    const u32x4 *A = (const u32x4 *)xa, *B = (const u32x4 *)xb;
    for (int64_t b0 = 0; b0 < qblocks; b0 += kMaxGridY) {  // gridDim.y <= 65,535
        const int64_t q0 = b0 * 128, mq = std::min<int64_t>(m - q0, kMaxGridY * 128);
        const dim3 grid((unsigned)ceil_div(n_tiles, slice), (unsigned)ceil_div(mq, 128));
        switch (kpad / 8) {
#define CREST_T16(KCV)                                                                                        \
    case KCV:                                                                                                 \
        hipLaunchKernelGGL(k_knn_thr_mfma16<KCV>, grid, dim3(256), 0, s, A, B, n_tiles, pos + q0, o.perm_c, mq, \
                           thr + q0, self, slice, cnt + q0, cand + q0 * cap, cap);                            \
        break;
            CREST_T16(4)
            CREST_T16(8)
            CREST_T16(16)
            CREST_T16(24)
            CREST_T16(32)
#undef CREST_T16
            default: throw std::runtime_error("no FP16 threshold kernel for this width");
        }
    }
}

// ---------------------------------------------------------------- large k
// k beyond the register lists (Scrublet asks for k ~ 1.5 sqrt(n), doublets.py:
// 131-137).  The dot products of a block of queries with every candidate come from
// one FP32 GEMM; a workgroup per query then radix-selects the k2 smallest screened
// keys of its row, max(|q|^2 + |c|^2 - 2 q.c, 0) as knn.rs:126 forms it (the query
// itself excluded in a graph), and a second workgroup per query ranks those k2
// exactly (FP64 squared distance, ties by index), keeps k and orders them by the
// reference's f32 distance and index (knn.rs:356-367).

// knn.rs:346: sum of squares of a row in f32, in column order.
__global__ void k_knn_sqnorm(const float *__restrict__ x, int64_t n, int p, float *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *r = x + i * p;
    float s = 0.f;
    for (int f = 0; f < p; ++f) {
        const float sq = r[f] * r[f];
        s = s + sq;
    }
    out[i] = s;
}

// Centring for the GEMM screen (any width): column sums of row parts in FP64 (thread =
// column), summed in part order, and the centred copy x - mean.
__global__ void k_col_sum_any(const float *x, int64_t n, int p, int64_t rows_per, double *part) {
    const int f = blockIdx.y * 256 + threadIdx.x;
    if (f >= p) return;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per, r1 = r0 + rows_per < n ? r0 + rows_per : n;
    double acc = 0.0;
    for (int64_t r = r0; r < r1; ++r) acc += (double)x[r * p + f];
    part[(int64_t)blockIdx.x * p + f] = acc;
}

__global__ void k_col_mean_any(const double *part, int nparts, int64_t n, int p, float *mean) {
    const int f = blockIdx.x * 256 + threadIdx.x;
    if (f >= p) return;
    double s = 0.0;
    for (int b = 0; b < nparts; ++b) s += part[(int64_t)b * p + f];
    mean[f] = (float)(s / (double)n);
}

__global__ void k_centre_rows(const float *x, int64_t n, int p, const float *mean, float *out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * p) return;
    out[e] = x[e] - mean[e % p];
}

// Every row of a batch centred on its own subset's mean (subset y = rows cell_off[y] ..
// cell_off[y + 1] - 1, means [n_sub][p]).
__global__ void k_centre_rows_seg(const float *x, const int64_t *cell_off, int n_sub, int p, const float *means,
                                  float *out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= cell_off[n_sub] * p) return;
    const int64_t r = e / p;
    int lo = 0, hi = n_sub;  // the last y with cell_off[y] <= r
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (cell_off[mid] <= r) lo = mid; else hi = mid;
    }
    out[e] = x[e] - means[(int64_t)lo * p + (e - r * p)];
}

void launch_knn_centre_seg(const float *x, int64_t n, const int64_t *cell_off, int n_sub, int p, const float *means,
                           float *out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_centre_rows_seg, dim3((unsigned)ceil_div(n * p, 256)), dim3(256), 0, s, x, cell_off, n_sub,
                       p, means, out);
}

int knn_centre_parts(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(4096, (n + 255) / 256)); }

void launch_knn_centre(const float *x, int64_t n, int p, float *mean, bool compute_mean, double *part, float *out,
                       hipStream_t s) {
    if (n <= 0) return;
    const unsigned cy = (unsigned)ceil_div(p, 256);
    if (compute_mean) {
        const int parts = knn_centre_parts(n);
        const int64_t rows_per = ceil_div(n, parts);
        hipLaunchKernelGGL(k_col_sum_any, dim3((unsigned)parts, cy), dim3(256), 0, s, x, n, p, rows_per, part);
        hipLaunchKernelGGL(k_col_mean_any, dim3(cy), dim3(256), 0, s, part, parts, n, p, mean);
    }
    hipLaunchKernelGGL(k_centre_rows, dim3((unsigned)ceil_div(n * p, 256)), dim3(256), 0, s, x, n, p, mean, out);
}

// Non-negative f32 distances order like their bit patterns.
__device__ inline uint32_t screen_key(float dot, float qn, float cn) {
    const float s = qn + cn;
    const float t = 2.0f * dot;
    const float d = s - t;
    return __float_as_uint(d > 0.f ? d : 0.f);
}

constexpr int kSelBins = 2048;
constexpr int kSelUnroll = 4;
constexpr int kSelGroups = 4096;  // >= the largest k2

struct SelShared {
    uint32_t hist[kSelBins];
    uint32_t bin, before;
    uint32_t cnt[2][4];
    uint32_t pcnt[8][4];
};

// The k-th smallest (1-based) of `count` keys key_at(i): three radix rounds over
// key bits 31-21, 20-10 and 9-0.  Returns it and sets need = k - #(keys below it).
// Called by all 256 threads of the workgroup.
template <class KeyAt>
__device__ uint32_t radix_kth(int64_t count, KeyAt key_at, uint32_t k, uint32_t &need_out, SelShared &sh) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t prefix = 0, need = k;
    for (int round = 0; round < 3; ++round) {
        const int shift = round == 0 ? 21 : round == 1 ? 10 : 0;
        const uint32_t hmask = round == 0 ? 0u : round == 1 ? 0xFFE00000u : 0xFFFFFC00u;  // bits already fixed
        const uint32_t dmask = round == 2 ? 0x3FFu : 0x7FFu;
        for (int i = tid; i < kSelBins; i += 256) sh.hist[i] = 0;
        __syncthreads();
        for (int64_t i0 = 0; i0 < count; i0 += 256 * kSelUnroll) {
            uint32_t kk[kSelUnroll];
#pragma unroll
            for (int u = 0; u < kSelUnroll; ++u) {
                const int64_t i = i0 + u * 256 + tid;
                kk[u] = i < count ? key_at(i) : 0xFFFFFFFFu;
            }
#pragma unroll
            for (int u = 0; u < kSelUnroll; ++u) {
                const bool act = i0 + u * 256 + tid < count && (kk[u] & hmask) == prefix;
                const uint32_t bin = (kk[u] >> shift) & dmask;
                uint64_t m = __ballot(act);
                while (m) {  // one LDS add per distinct bin of the wave
                    const int leader = __builtin_ctzll(m);
                    const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)bin, leader);
                    const uint64_t same = __ballot(act && bin == b);
                    if (lane == leader) atomicAdd(&sh.hist[b], (uint32_t)__popcll(same));
                    m &= ~same;
                }
            }
        }
        __syncthreads();
        if (wave == 0) {  // the bin holding the need-th key: 32 bins per lane, then a wave scan
            constexpr int per = kSelBins / 64;
            uint32_t sum = 0;
            for (int b = 0; b < per; ++b) sum += sh.hist[lane * per + b];
            uint32_t inc = sum;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t v = __shfl_up(inc, o, 64);
                if (lane >= o) inc += v;
            }
            const uint64_t hit = __ballot(inc >= need);
            if (lane == __builtin_ctzll(hit)) {
                uint32_t acc = inc - sum;
                int b = lane * per;
                while (acc + sh.hist[b] < need) acc += sh.hist[b++];
                sh.bin = (uint32_t)b;
                sh.before = acc;
            }
        }
        __syncthreads();
        prefix |= sh.bin << shift;
        need -= sh.before;
        __syncthreads();
    }
    need_out = need;
    return prefix;
}

// Among `count` entries in candidate order (key key_at(i), candidate idx_at(i)),
// write to out[0..k2) every candidate whose key is below T, the k2-th smallest key,
// then the first ones whose key equals T.  Called by all 256 threads.
template <class KeyAt, class IdxAt>
__device__ void select_ordered(int64_t count, KeyAt key_at, IdxAt idx_at, int k2, uint32_t *out, SelShared &sh) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t need = 0;
    const uint32_t T = radix_kth(count, key_at, (uint32_t)k2, need, sh);
    const uint32_t n_lt = (uint32_t)k2 - need;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t base_lt = 0, base_eq = 0;
    for (int64_t i0 = 0; i0 < count; i0 += 256) {
        const int64_t i = i0 + tid;
        const uint32_t kk = i < count ? key_at(i) : 0xFFFFFFFFu;
        const bool lt = i < count && kk < T, eq = i < count && kk == T;
        const uint64_t blt = __ballot(lt), beq = __ballot(eq);
        if (lane == 0) {
            sh.cnt[0][wave] = (uint32_t)__popcll(blt);
            sh.cnt[1][wave] = (uint32_t)__popcll(beq);
        }
        __syncthreads();
        uint32_t off_lt = base_lt, off_eq = base_eq;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            if (w < wave) {
                off_lt += sh.cnt[0][w];
                off_eq += sh.cnt[1][w];
            }
            base_lt += sh.cnt[0][w];
            base_eq += sh.cnt[1][w];
        }
        __syncthreads();
        if (lt) out[off_lt + __popcll(blt & below)] = idx_at(i);
        if (eq) {
            const uint32_t pos = off_eq + (uint32_t)__popcll(beq & below);
            if (pos < need) out[n_lt + pos] = idx_at(i);
        }
    }
}

// One workgroup (4 waves) per query row of `dots` (row r = query q0 + r at dots +
// r ld, ld a multiple of 4, cn readable up to ld): the k2 candidates of smallest
// screened key, in candidate order, into cand[r].  Pass 1 takes the minimum key of
// each of G = 4096 >= k2 groups (a fixed map of c mod G) and U, the k2-th smallest
// of those minima: k2 distinct candidates have keys <= U, so the k2-th smallest key
// of the row is <= U, while few others are (about k2 (1 + k2/G)).  Pass 2 copies the
// candidates with key <= U in candidate order into comp[r] (capacity cap); the
// selection runs on that copy, or on the whole row when it overflowed.  Rows of at
// most 2G candidates are selected directly.  Both passes read four candidates per
// lane per load.
__global__ __launch_bounds__(256) void k_knn_select_rows(const float *__restrict__ dots, int64_t ld, int64_t nc,
                                                         const float *__restrict__ qn_all,
                                                         const float *__restrict__ cn, int64_t q0, int self, int k2,
                                                         int cap, uint32_t *__restrict__ comp_key,
                                                         uint32_t *__restrict__ comp_idx, uint32_t *__restrict__ cand) {
    __shared__ SelShared sh;
    __shared__ uint32_t gmin[kSelGroups];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t r = blockIdx.x;
    const int64_t q = q0 + r;
    const float *row = dots + r * ld;
    const float qn = qn_all[q];
    auto key = [&](int64_t c) -> uint32_t {
        if (self && c == q) return 0x7F800000u;  // the query itself: +inf
        return screen_key(row[c], qn, cn[c]);
    };
    // keys of candidates c .. c+3 (c a multiple of 4 below ld); past nc: all ones
    auto key4 = [&](int64_t c, uint32_t *kk) {
        const float4 d = *reinterpret_cast<const float4 *>(row + c);
        const float4 n4 = *reinterpret_cast<const float4 *>(cn + c);
        const float dv[4] = {d.x, d.y, d.z, d.w}, nv[4] = {n4.x, n4.y, n4.z, n4.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t cj = c + j;
            kk[j] = cj >= nc ? 0xFFFFFFFFu : (self && cj == q) ? 0x7F800000u : screen_key(dv[j], qn, nv[j]);
        }
    };
    uint32_t *out = cand + r * (int64_t)k2;
    if (nc <= 2 * kSelGroups) {
        select_ordered(
            nc, [&](int64_t i) { return key(i); }, [&](int64_t i) { return (uint32_t)i; }, k2, out, sh);
        return;
    }
    // pass 1: group minima (group of c: (c/4) mod 1024 + 1024 (c mod 4)), then their k2-th smallest
    constexpr int V = 4;  // float4 loads per lane per step
    for (int i = tid; i < kSelGroups; i += 256) gmin[i] = 0xFFFFFFFFu;
    __syncthreads();
    for (int64_t c0 = 0; c0 < nc; c0 += 1024 * V) {
        uint32_t kk[V][4];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int64_t c = c0 + (int64_t)(v * 256 + tid) * 4;
            if (c < nc) key4(c, kk[v]);
        }
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int64_t c = c0 + (int64_t)(v * 256 + tid) * 4;
            if (c < nc) {
                const uint32_t g = (uint32_t)((c >> 2) & 1023);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (kk[v][j] != 0xFFFFFFFFu) atomicMin(&gmin[g + 1024 * j], kk[v][j]);
            }
        }
    }
    __syncthreads();
    uint32_t need_g = 0;
    const uint32_t U = radix_kth((int64_t)kSelGroups, [&](int64_t i) { return gmin[i]; }, (uint32_t)k2, need_g, sh);
    // pass 2: ordered copy of the keys <= U; order within a step: v, wave, lane, j
    uint32_t *ck = comp_key + r * (int64_t)cap;
    uint32_t *cix = comp_idx + r * (int64_t)cap;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t m = 0;
    for (int64_t c0 = 0; c0 < nc && m <= (uint32_t)cap; c0 += 1024 * V) {
        uint32_t kk[V][4];
        bool any = false;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int64_t c = c0 + (int64_t)(v * 256 + tid) * 4;
            if (c < nc)
                key4(c, kk[v]);
            else
                kk[v][0] = kk[v][1] = kk[v][2] = kk[v][3] = 0xFFFFFFFFu;
#pragma unroll
            for (int j = 0; j < 4; ++j) any |= kk[v][j] <= U;  // U < all ones
        }
        if (!__syncthreads_or(any)) continue;
        uint32_t pre[V];  // hits of the lanes below, over all four slots
#pragma unroll
        for (int v = 0; v < V; ++v) {
            uint32_t tot = 0;
            pre[v] = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint64_t b = __ballot(kk[v][j] <= U);
                pre[v] += (uint32_t)__popcll(b & below);
                tot += (uint32_t)__popcll(b);
            }
            if (lane == 0) sh.pcnt[v][wave] = tot;
        }
        __syncthreads();
        uint32_t off[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            off[v] = m;
            for (int w = 0; w < 4; ++w) {
                if (w < wave) off[v] += sh.pcnt[v][w];
                m += sh.pcnt[v][w];
            }
        }
        __syncthreads();
#pragma unroll
        for (int v = 0; v < V; ++v) {
            uint32_t pos = off[v] + pre[v];
            const int64_t c = c0 + (int64_t)(v * 256 + tid) * 4;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (kk[v][j] <= U) {
                    if (pos < (uint32_t)cap) {
                        ck[pos] = kk[v][j];
                        cix[pos] = (uint32_t)(c + j);
                    }
                    ++pos;
                }
        }
    }
    __syncthreads();  // the copy is complete and visible to the workgroup
    if (m <= (uint32_t)cap)
        select_ordered(
            (int64_t)m, [&](int64_t i) { return ck[i]; }, [&](int64_t i) { return cix[i]; }, k2, out, sh);
    else
        select_ordered(
            nc, [&](int64_t i) { return key(i); }, [&](int64_t i) { return (uint32_t)i; }, k2, out, sh);
}

// Bitonic sort of n2 (a power of two) (distance, index) pairs in LDS, ascending.
__device__ void lds_bitonic(double *sd, uint32_t *si, int n2) {
    for (int size = 2; size <= n2; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < n2 / 2; t += blockDim.x) {
                const int lo = 2 * t - (t & (stride - 1));
                const int hi = lo + stride;
                const bool up = (lo & size) == 0;
                const double da = sd[lo], db = sd[hi];
                const uint32_t ia = si[lo], ib = si[hi];
                const bool gt = da > db || (da == db && ia > ib);
                if (gt == up) {
                    sd[lo] = db;
                    sd[hi] = da;
                    si[lo] = ib;
                    si[hi] = ia;
                }
            }
            __syncthreads();
        }
}

// One workgroup per query: FP64 squared distances of its k2 candidates, sorted by
// (distance, index); the first k get the reference's f32 distance and are sorted by
// (that distance, index) for output.  LDS: n2 doubles + n2 indices, n2 = 2^m >= k2.
__global__ __launch_bounds__(256) void k_knn_rerank_sort(const float *__restrict__ xq_all,
                                                         const float *__restrict__ xc, int p,
                                                         const uint32_t *__restrict__ cand, int k2, int n2, int k,
                                                         int64_t q0, uint32_t *idx_out, float *dist_out,
                                                         const float *__restrict__ tau, KnnCert ct,
                                                         const float *__restrict__ xq_scr) {
    extern __shared__ double lds_d[];
    double *sd = lds_d;
    uint32_t *si = reinterpret_cast<uint32_t *>(lds_d + n2);
    const int64_t q = q0 + blockIdx.x;
    const float *xq = xq_all + q * (int64_t)p;
    const uint32_t *ci = cand + (int64_t)blockIdx.x * k2;
    for (int j = threadIdx.x; j < n2; j += blockDim.x) {
        double d = INFINITY;
        uint32_t c = 0xFFFFFFFFu;
        if (j < k2 && ci[j] != 0xFFFFFFFFu) {  // an empty slot of a threshold scan's list
            c = ci[j];
            const float *xr = xc + (int64_t)c * p;
            double acc = 0.0;
            for (int f = 0; f < p; ++f) {
                const double dd = (double)xq[f] - (double)xr[f];
                acc += dd * dd;
            }
            d = acc;
        }
        sd[j] = d;
        si[j] = c;
    }
    __syncthreads();
    lds_bitonic(sd, si, n2);
    const double dk = sd[k - 1];  // the k-th smallest FP64 squared distance (INFINITY: fewer valid)
    __syncthreads();
    int m = 1;
    while (m < k) m <<= 1;
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
        float f = INFINITY;
        uint32_t c = 0xFFFFFFFFu;
        if (j < k) {
            c = si[j];
            f = sqrtf(ref_sq_dist(xq, xc + (int64_t)c * p, p));
        }
        sd[j] = (double)f;
        si[j] = c;
    }
    __syncthreads();
    lds_bitonic(sd, si, m);
    for (int j = threadIdx.x; j < k; j += blockDim.x) {
        idx_out[q * k + j] = si[j];
        dist_out[q * k + j] = (float)sd[j];
    }
    // certificate (KnnCert with s = 1, a = |q| of the screen's (centred) row; every
    // candidate the select dropped has a key >= tau, the largest kept key)
    if (ct.count && threadIdx.x == 0) {
        const double t = (double)tau[blockIdx.x];
        bool ok = !(t < (double)INFINITY);
        if (!ok && dk < (double)INFINITY) {
            const float *xs = xq_scr + q * (int64_t)p;
            double qq = 0.0;
            for (int f = 0; f < p; ++f) qq += (double)xs[f] * (double)xs[f];
            const double a = sqrt(qq) * (1.0 + 1e-6);
            const double b = a + sqrt(dk) * (1.0 + 1e-6);
            const double e = ct.c1 * a * b + ct.c2 * (a + b) * (a + b) + ct.c3 * (a + b) + ct.c4;
            ok = t - e > dk * (1.0 + 1e-12);
        }
        if (!ok) atomicAdd(ct.count, 1u);
    }
}

// tau[r] = the largest screen key among query row r's k2 selected candidates (every
// candidate the select dropped has a key at least this).
__global__ void k_knn_cand_max(const float *__restrict__ dots, int64_t ld, const float *__restrict__ qn_all,
                               const float *__restrict__ cn, int64_t q0, int64_t rows, const uint32_t *cand, int k2,
                               float *tau) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const float qn = qn_all[q0 + r];
    uint32_t mx = 0;
    for (int j = 0; j < k2; ++j) {
        const uint32_t c = cand[r * k2 + j];
        const uint32_t key = screen_key(dots[r * ld + c], qn, cn[c]);
        mx = key > mx ? key : mx;
    }
    tau[r] = __uint_as_float(mx);
}

void launch_knn_cand_max(const float *dots, int64_t ld, const float *qn, const float *cn, int64_t q0, int64_t rows,
                         const uint32_t *cand, int k2, float *tau, hipStream_t s) {
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_knn_cand_max, dim3((unsigned)ceil_div(rows, 256)), dim3(256), 0, s, dots, ld, qn, cn, q0,
                       rows, cand, k2, tau);
}

void launch_knn_sqnorm(const float *x, int64_t n, int p, float *out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_knn_sqnorm, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, x, n, p, out);
}

void launch_knn_select_rows(const float *dots, int64_t ld, int64_t nc, const float *qn, const float *cn,
                            int64_t q0, int64_t rows, int self, int k2, int cap, uint32_t *comp_key,
                            uint32_t *comp_idx, uint32_t *cand, hipStream_t s) {
    if (rows <= 0) return;
    hipLaunchKernelGGL(k_knn_select_rows, dim3((unsigned)rows), dim3(256), 0, s, dots, ld, nc, qn, cn, q0, self, k2,
                       cap, comp_key, comp_idx, cand);
}

void launch_knn_rerank_sort(const float *xq, const float *xc, int p, const uint32_t *cand, int k2, int k,
                            int64_t q0, int64_t rows, uint32_t *idx_out, float *dist_out, hipStream_t s,
                            const float *tau, const KnnCert *cert, const float *xq_scr) {
    if (rows <= 0) return;
    int n2 = 1;
    while (n2 < k2) n2 <<= 1;
    const size_t lds = (size_t)n2 * (sizeof(double) + sizeof(uint32_t));
    const KnnCert ct = (cert && tau && xq_scr) ? *cert : KnnCert{};
    hipLaunchKernelGGL(k_knn_rerank_sort, dim3((unsigned)rows), dim3(256), lds, s, xq, xc, p, cand, k2, n2, k, q0,
                       idx_out, dist_out, tau, ct, xq_scr);
}

}  // namespace crest
