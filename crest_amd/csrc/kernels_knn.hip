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
// Column means (deterministic two-stage FP64 sum), then the centred rows' squared
// norms Σ x'² in f32, summed sequentially as knn.rs:346 does, and the augmented
// operands, one 16-byte chunk at a time so that a wave (64 consecutive rows) stores
// 1 KB runs.  The work buffer holds kColParts*p partial sums, the p means and the
// largest centred norm (as its bit pattern; norms are >= 0, so integer order is
// float order).
constexpr int kColParts = 512;

size_t knn_work_bytes(int p) { return (size_t)kColParts * p * sizeof(double) + (size_t)p * sizeof(float) + 64; }

struct KnnWork {
    double *part;
    float *mean;
    uint32_t *maxbits;
};
static KnnWork knn_work(void *w, int p) {
    KnnWork k;
    k.part = (double *)w;
    k.mean = (float *)(k.part + (size_t)kColParts * p);
    k.maxbits = (uint32_t *)(k.mean + p);
    return k;
}

// Block b sums rows [b*rows_per, (b+1)*rows_per): thread t takes column t % p of
// every (256/p)-th row, and the row groups are combined in order through LDS.
__global__ __launch_bounds__(256) void k_knn_colsum(const float *x, int64_t n, int p, int64_t rows_per,
                                                    double *part) {
    __shared__ double red[256];
    const int g = 256 / p, t = threadIdx.x, f = t % p, j = t / p;
    const int64_t r0 = blockIdx.x * rows_per, r1 = r0 + rows_per < n ? r0 + rows_per : n;
    double acc = 0.0;
    if (j < g)
        for (int64_t r = r0 + j; r < r1; r += g) acc += (double)x[r * p + f];
    red[t] = acc;
    __syncthreads();
    if (t < p) {
        double s = 0.0;
        for (int jj = 0; jj < g; ++jj) s += red[jj * p + t];
        part[(int64_t)blockIdx.x * p + t] = s;
    }
}

__global__ void k_knn_colmean(const double *part, int nparts, int64_t n, int p, float *mean, uint32_t *maxbits) {
    for (int f = threadIdx.x; f < p; f += blockDim.x) {
        double s = 0.0;
        for (int b = 0; b < nparts; ++b) s += part[(int64_t)b * p + f];
        mean[f] = (float)(s / (double)n);
    }
    if (threadIdx.x == 0) *maxbits = 0u;
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
                                                     uint32_t *maxbits) {
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

__global__ void k_knn_prep_bf16(const float *x, const float *mean, int64_t n, int64_t n_pad, int p, int kc,
                                uint4 *xa, uint4 *xb) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n_pad) return;
    const bool real = i < n;
    const float *r = x + (real ? i : 0) * (int64_t)p;
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
        xa[u] = va;
        xb[u] = vb;
    }
}

// FP16 operands of a centred, prescaled row; fp16 conversions round to nearest even.
__global__ void k_knn_prep_f16(const float *x, const float *mean, const uint32_t *maxbits, int64_t n, int64_t n_pad,
                               int p, int kc, uint4 *xa, uint4 *xb) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n_pad) return;
    const bool real = i < n;
    const float s = fp16_prescale(maxbits);
    const float *r = x + (real ? i : 0) * (int64_t)p;
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
        xa[u] = va;
        xb[u] = vb;
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

__global__ void k_knn_prep_f32(const float *x, const float *mean, int64_t n, int64_t n_pad, int p, int kc,
                               float2 *xa, float2 *xb) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n_pad) return;
    const bool real = i < n;
    const float *r = x + (real ? i : 0) * (int64_t)p;
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
            xa[u] = make_float2(a0, a1);
            xb[u] = make_float2(b0, b1);
        }
}

int64_t knn_rows_padded(int64_t n) { return round_up(n, kPadRows); }

void launch_knn_prep(int prec, const float *x, int64_t n, int p, int kpad, void *work, void *xa, void *xb,
                     hipStream_t s) {
    const KnnWork w = knn_work(work, p);
    const int64_t n_pad = knn_rows_padded(n);
    const int64_t rows_per = ceil_div(n, kColParts);
    const int nparts = (int)ceil_div(n, rows_per);
    hipLaunchKernelGGL(k_knn_colsum, dim3(nparts), dim3(256), 0, s, x, n, p, rows_per, w.part);
    hipLaunchKernelGGL(k_knn_colmean, dim3(1), dim3(256), 0, s, w.part, nparts, n, p, w.mean, w.maxbits);
    const int g = (int)ceil_div(n_pad, 256);
    if (prec == kPrecFP16) {
        hipLaunchKernelGGL(k_knn_maxnorm, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, x, n, p, w.mean,
                           w.maxbits);
        hipLaunchKernelGGL(k_knn_prep_f16, dim3(g), dim3(256), 0, s, x, w.mean, w.maxbits, n, n_pad, p, kpad / 8,
                           (uint4 *)xa, (uint4 *)xb);
    } else if (prec == kPrecBF16X3) {
        hipLaunchKernelGGL(k_knn_prep_bf16, dim3(g), dim3(256), 0, s, x, w.mean, n, n_pad, p, kpad / 8, (uint4 *)xa,
                           (uint4 *)xb);
    } else {
        hipLaunchKernelGGL(k_knn_prep_f32, dim3(g), dim3(256), 0, s, x, w.mean, n, n_pad, p, kpad / 4, (float2 *)xa,
                           (float2 *)xb);
    }
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
// earlier (lower-index) candidate stays ahead, as push_topk does (knn.rs:68-87).
template <int L>
__device__ inline void topk_insert(float (&ld)[L], uint32_t (&li)[L], float d, uint32_t id) {
#pragma unroll
    for (int j = 0; j < L; ++j) {
        const bool sw = d < ld[j];
        const float td = sw ? d : ld[j];
        const uint32_t ti = sw ? id : li[j];
        d = sw ? ld[j] : d;
        id = sw ? li[j] : id;
        ld[j] = td;
        li[j] = ti;
    }
}

// Register r of lane l holds the pair (candidate cb + 8(r/4) + 4(l/32) + r%4, query
// of column l%32); candidates are offered in increasing index order.
template <int L>
__device__ inline void topk_update(float (&ld)[L], uint32_t (&li)[L], const f32x16 &a, uint32_t cb, int half) {
    float m = a[0];
#pragma unroll
    for (int r = 1; r < 16; ++r) m = fminf(m, a[r]);
    if (m < ld[L - 1]) {
        uint32_t mask = 0;
#pragma unroll
        for (int r = 0; r < 16; ++r) mask |= (a[r] < ld[L - 1]) ? (1u << r) : 0u;
        while (mask) {
            const int r = __builtin_ctz(mask);
            mask &= mask - 1;
            float v = a[0];
#pragma unroll
            for (int j = 1; j < 16; ++j) v = (r == j) ? a[j] : v;
            if (v < ld[L - 1]) topk_insert<L>(ld, li, fmaxf(v, 0.0f), cb + 8 * (r >> 2) + 4 * half + (r & 3));
        }
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
// B fragments and lists leave room in 256 VGPRs (checked against the compiler's
// register report: no spills).  Eight-wave workgroups (half the tile traffic per
// query) are used where that register budget holds.
template <int KC, int L> constexpr int screen_waves() { return (L <= 16 && KC <= 20) ? 2 : 1; }
template <int TY, int KC, int L> constexpr int screen_nw() { return TY == 1 && L <= 16 && KC % 8 == 0 && KC <= 24 ? 8 : 4; }

template <int TY, int KC, int L, int NW>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(screen_waves<KC, L>()))) void k_knn_screen16(
    const u32x4 *__restrict__ xa, const u32x4 *__restrict__ xb, int64_t n_pad, float *cand_d, uint32_t *cand_i) {
    constexpr int KS = KC / 2;    // MFMA K-steps of 16
    constexpr int TU = KC * kTC;  // 16-byte units per tile
    constexpr int NT = NW * 64;
    static_assert(TU % NT == 0, "tile must split evenly over the workgroup");
    constexpr int PER = TU / NT;
    __shared__ u32x4 lds[2][TU];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, col = lane & 31, half = lane >> 5;
    const int64_t q0 = (int64_t)blockIdx.x * NT + wave * kQW;
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
    const int64_t n_tiles = n_pad / kTC;
    // Pipeline: LDS holds tile t (being computed) and receives tile t+1 from
    // registers loaded one iteration earlier; the loads for tile t+2 are issued after
    // that store, so their latency overlaps the whole next iteration.  (The final
    // iterations re-load the last tile; those copies are never read.)
    for (int i = threadIdx.x; i < TU; i += NT) lds[0][i] = xa[i];
    Regs<u32x4, PER> pre;
    regs_load<u32x4, PER, NT>(pre, xa + (n_tiles > 1 ? 1 : 0) * TU, threadIdx.x);
    __syncthreads();
    for (int64_t t = 0; t < n_tiles; ++t) {
        const int buf = (int)(t & 1);
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            f32x16 a0 = {}, a1 = {};
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const u32x4 a = lds[buf][(2 * s + half) * kTC + mt * 32 + col];
                a0 = mfma16<TY>(a, b0[s], a0);
                a1 = mfma16<TY>(a, b1[s], a1);
            }
            const uint32_t cb = (uint32_t)(t * kTC + mt * 32);
            if (cb == (uint32_t)q0 || cb == (uint32_t)q0 + 32) {
                self_mask<L>(a0, cb, half, qa);
                self_mask<L>(a1, cb, half, qb);
            }
            topk_update<L>(d0, i0, a0, cb, half);
            topk_update<L>(d1, i1, a1, cb, half);
        }
        regs_store<u32x4, PER, NT>(pre, lds[buf ^ 1], threadIdx.x);
        const int64_t tn = t + 2 < n_tiles ? t + 2 : n_tiles - 1;
        regs_load<u32x4, PER, NT>(pre, xa + tn * TU, threadIdx.x);
        __syncthreads();
    }
    write_lists<L>(d0, i0, qa, half, cand_d, cand_i);
    write_lists<L>(d1, i1, qb, half, cand_d, cand_i);
}

// ---------------------------------------------------------------- screen, FP32
// v_mfma_f32_32x32x2_f32: per 16-byte chunk, lane half h supplies floats 2h, 2h+1 in
// two MFMAs, so the chunk's four K values are covered by the two half-waves.
template <int KC, int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((L <= 16 && KC <= 16) ? 2 : 1))) void k_knn_screen_f32(const float2 *__restrict__ xa, const float2 *__restrict__ xb,
                                                        int64_t n_pad, float *cand_d, uint32_t *cand_i) {
    constexpr int TU = KC * 2 * kTC;  // float2 units per tile
    static_assert(TU % 256 == 0, "tile must split evenly over 256 threads");
    constexpr int PER = TU / 256;
    __shared__ float2 lds[2][TU];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, col = lane & 31, half = lane >> 5;
    const int64_t q0 = (int64_t)blockIdx.x * 256 + wave * kQW;
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
    const int64_t n_tiles = n_pad / kTC;
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
            const uint32_t cb = (uint32_t)(t * kTC + mt * 32);
            if (cb == (uint32_t)q0 || cb == (uint32_t)q0 + 32) {
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
static bool launch_screen_L(int prec, const void *xa, const void *xb, int64_t n_pad, int kpad, float *cd, uint32_t *ci,
                            hipStream_t s) {
#define CREST_X16(TY, KCV)                                                                                       \
    case KCV: {                                                                                                 \
        constexpr int NW = screen_nw<TY, KCV, L>();                                                             \
        hipLaunchKernelGGL((k_knn_screen16<TY, KCV, L, NW>), dim3((unsigned)(n_pad / (NW * 64))), dim3(NW * 64), \
                           0, s, (const u32x4 *)xa, (const u32x4 *)xb, n_pad, cd, ci);                          \
        return true;                                                                                            \
    }
#define CREST_F(KCV)                                                                                              \
    case KCV:                                                                                                     \
        hipLaunchKernelGGL((k_knn_screen_f32<KCV, L>), dim3((unsigned)(n_pad / 256)), dim3(256), 0, s,          \
                           (const float2 *)xa, (const float2 *)xb, n_pad, cd, ci);                               \
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

// Screen with lists of L per lane (L in {16, 32, 64}, L >= k); cand arrays hold 2L per
// query for knn_rows_padded(n) rows.  Returns false when no kernel is instantiated for
// (prec, kpad, L).
bool launch_knn_screen(int prec, const void *xa, const void *xb, int64_t n, int kpad, int L, float *cand_d,
                       uint32_t *cand_i, hipStream_t s) {
    const int64_t n_pad = knn_rows_padded(n);
    if (L == 16) return launch_screen_L<16>(prec, xa, xb, n_pad, kpad, cand_d, cand_i, s);
    if (L == 32) return launch_screen_L<32>(prec, xa, xb, n_pad, kpad, cand_d, cand_i, s);
    if (L == 64) return launch_screen_L<64>(prec, xa, xb, n_pad, kpad, cand_d, cand_i, s);
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

// One wave per query.  Candidates are ranked by their exact FP64 squared distance
// (ties by index); the k best are reported with the reference's FP32 distance and
// ordered by (distance, index) as knn.rs:356-367 does.
template <int NCL>
__global__ __launch_bounds__(256) void k_knn_rerank(const float *__restrict__ x, int64_t n, int p,
                                                    const uint32_t *__restrict__ cand_i, int n_cand, int k,
                                                    uint32_t *idx_out, float *dist_out) {
    const int lane = threadIdx.x & 63;
    const int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    if (q >= n) return;
    const float *xq = x + q * (int64_t)p;
    const uint32_t *ci = cand_i + q * (int64_t)n_cand;
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
        const uint32_t c = ci[j];
        const bool valid = (int64_t)c < n && (int64_t)c != q;
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
            idx_out[q * k + rank2[s]] = cc[s];
            dist_out[q * k + rank2[s]] = df[s];
        }
}

void launch_knn_rerank(const float *x, int64_t n, int p, const float *cand_d, const uint32_t *cand_i, int n_cand,
                       int k, uint32_t *idx_out, float *dist_out, hipStream_t s) {
    (void)cand_d;
    const int g = (int)ceil_div(n, 4);
    if (n_cand <= 64)
        hipLaunchKernelGGL(k_knn_rerank<1>, dim3(g), dim3(256), 0, s, x, n, p, cand_i, n_cand, k, idx_out, dist_out);
    else
        hipLaunchKernelGGL(k_knn_rerank<2>, dim3(g), dim3(256), 0, s, x, n, p, cand_i, n_cand, k, idx_out, dist_out);
}

}  // namespace crest
