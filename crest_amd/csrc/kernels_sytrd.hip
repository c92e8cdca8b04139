// Householder tridiagonalization of a batch of symmetric matrices, one workgroup per
// matrix: A = Q T Q^T with Q = H(0) H(1) ... H(n-2), in the layout of LAPACK's dsytrd
// with uplo = lower — D (diagonal of T), E (off-diagonal), tau, and the reflector
// entries v(i+2:n) in A(i+2:n, i) (v(i+1) = 1 implicit, A(i+1, i) = E(i)).  The upper
// triangle is never read.  This is what the top-p eigensolver and its back-transform
// consume (kernels.rs:324-345 take the covariance's top eigenpairs).
//
// The blocked scheme of LAPACK's dlatrd/dsytrd: within a panel of NB columns, column i
// is first brought up to date with the panel's earlier reflectors (A -= V W^T + W V^T
// restricted to that column), its reflector is generated (dlarfg), and
// w = tau (A22 v - V (W^T v) - W (V^T v)), w -= tau/2 (w^T v) v is formed from the
// trailing matrix as it was at the start of the panel; after the panel, one symmetric
// rank-2NB update brings the trailing lower triangle up to date.
//
// A matrix of order 2,000 is 32 MB, so a batch streams from HBM: the symmetric
// matrix-vector product reads each lower-triangle element once (64 x 64 tiles, one per
// wave at a time: the row part accumulates per lane, the transposed part is summed
// across lanes by a halving exchange), so the whole reduction reads ~n^3/6 doubles per
// matrix, half of what a full-storage product reads.  The partial products of the eight
// waves go to per-wave LDS vectors and are summed in wave order, so the result does
// not depend on scheduling.
#include <cmath>

#include "common.hpp"

namespace crest {

namespace {

constexpr int kSyNB = 32;       // panel width
constexpr int kSyWaves = 8;     // 512 threads
constexpr int kSyThreads = kSyWaves * 64;
constexpr int kSyRU = 4;        // rows per thread: row r belongs to thread r % 512
constexpr int kSyMaxN = kSyThreads * kSyRU;
constexpr int kSyStage = 66;    // LDS stride of a staged column (16-byte aligned, 4 banks apart)

__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ inline double wsum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline double readlane_f64(double v, int l) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over the 64 lanes of t[c] for 16 entries at once: after the step of width h each
// lane keeps the half of its values selected by lane bit h and adds its partner's copy
// of that half; after four steps lane l holds entry l & 15 summed over its 16-lane
// group, and two more exchanges add the other three groups.
template <int H> __device__ inline void halve(double *t, int lane) {
    const bool up = (lane & H) != 0;
#pragma unroll
    for (int j = 0; j < H; ++j) {
        const double send = up ? t[j] : t[j + H];
        const double keep = up ? t[j + H] : t[j];
        t[j] = keep + __shfl_xor(send, H, 64);
    }
}
__device__ inline double transpose_sum16(double *t, int lane) {
    halve<8>(t, lane);
    halve<4>(t, lane);
    halve<2>(t, lane);
    halve<1>(t, lane);
    double v = t[0] + __shfl_xor(t[0], 16, 64);
    return v + __shfl_xor(v, 32, 64);
}

}  // namespace

size_t sytrd_wg_lds(int n) {
    const int ldn = (n + 63) & ~63;
    const int ypart = kSyWaves * (ldn > 16 * kSyStage ? ldn : 16 * kSyStage);
    return (size_t)(ypart + ldn + 64 + kSyWaves * 64 + kSyWaves + 3 * kSyNB + 8) * sizeof(double);
}
bool sytrd_wg_supported(int n) { return n >= 1 && n <= kSyMaxN && sytrd_wg_lds(n) <= 160 * 1024; }

// grid (n_mat), 512 threads, LDS sytrd_wg_lds(n).  Wg: scratch [n_mat][NB][n].
__global__ __launch_bounds__(kSyThreads) void k_sytrd_lower(double *A, int64_t sA, int lda, double *D, int64_t sD,
                                                            double *E, int64_t sE, double *tau, int64_t sT,
                                                            double *Wg, int n) {
    extern __shared__ double lds[];
    const int ldn = (n + 63) & ~63;
    const int ylen = ldn > 16 * kSyStage ? ldn : 16 * kSyStage;
    double *ypart = lds;                        // [8][ylen]: symv partials, or syr2k staging
    // [ldn + 64] current reflector, by row; zero beyond n (the last column tile of the
    // product reads up to 63 entries past the matrix)
    double *vl = ypart + kSyWaves * ylen;
    double *red64 = vl + ldn + 64;              // [64] (sized [8][64])
    double *red = red64 + kSyWaves * 64;        // [8]
    double *rowW = red + kSyWaves;              // [NB] W(i, :) of the next column's row
    double *rowV = rowW + kSyNB;                // [NB] V(i, :)
    double *panE = rowV + kSyNB;                // [NB] E of the panel's columns
    double *sc = panE + kSyNB;                  // [8] broadcast scalars

    const int mat = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    A += mat * sA;
    D += mat * sD;
    E += mat * sE;
    tau += mat * sT;
    Wg += (int64_t)mat * kSyNB * n;
    for (int r = tid; r < ldn + 64; r += kSyThreads) vl[r] = 0.0;
    lds_barrier();

    for (int j0 = 0; j0 < n - 1; j0 += kSyNB) {
        const int pw = n - 1 - j0 < kSyNB ? n - 1 - j0 : kSyNB;
        for (int jj = 0; jj < pw; ++jj) {
            const int i = j0 + jj, i1 = i + 1;
            // (a) column i of the thread's rows brought up to date with the panel
            double colv[kSyRU], vr[kSyRU];
            double ss = 0.0;
#pragma unroll
            for (int u = 0; u < kSyRU; ++u) {
                const int r = tid + kSyThreads * u;
                colv[u] = 0.0;
                if (r >= i && r < n) {
                    double a = A[r + (int64_t)i * lda];
                    for (int q = 0; q < jj; ++q)
                        a -= A[r + (int64_t)(j0 + q) * lda] * rowW[q] + Wg[(int64_t)q * n + r] * rowV[q];
                    colv[u] = a;
                    if (r >= i + 2) ss += a * a;
                    if (r == i1) sc[0] = a;
                    if (r == i) {
                        A[r + (int64_t)i * lda] = a;
                        D[i] = a;
                    }
                }
            }
            ss = wsum(ss);
            if (lane == 0) red[wave] = ss;
            lds_barrier();
            // (b) the reflector (dlarfg): beta = -sign(alpha) ||(alpha, x)||,
            // tau = (beta - alpha) / beta, v = x / (alpha - beta)
            double xn2 = 0.0;
#pragma unroll
            for (int w = 0; w < kSyWaves; ++w) xn2 += red[w];
            const double alpha = sc[0];
            double beta = alpha, tv = 0.0, scale = 1.0;
            if (xn2 > 0.0) {
                beta = -copysign(hypot(alpha, sqrt(xn2)), alpha);
                tv = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
#pragma unroll
            for (int u = 0; u < kSyRU; ++u) {
                const int r = tid + kSyThreads * u;
                vr[u] = 0.0;
                if (r >= i1 && r < n) {
                    const double v = r == i1 ? 1.0 : colv[u] * scale;
                    vr[u] = v;
                    vl[r] = v;
                    A[r + (int64_t)i * lda] = v;
                }
            }
            if (tid == 0) {
                E[i] = beta;
                tau[i] = tv;
                panE[jj] = beta;
            }
            lds_barrier();
            // (c1) W^T v and V^T v over the panel's earlier columns, one wave per dot
            // product: red64[q] = W(:, q)^T v, red64[32 + q] = V(:, q)^T v over rows i1..n-1
            for (int k = wave; k < 2 * jj; k += kSyWaves) {
                const bool isw = k < jj;
                const int q = isw ? k : k - jj;
                const double *src = isw ? Wg + (int64_t)q * n : A + (int64_t)(j0 + q) * lda;
                double acc = 0.0;
                for (int r = i1 + lane; r < n; r += 64) acc += src[r] * vl[r];
                acc = wsum(acc);
                if (lane == 0) red64[(isw ? 0 : kSyNB) + q] = acc;
            }
            // (c2) y = A22 v over the lower triangle of rows/columns i1..n-1, in tiles of
            // 64 rows x 32 columns (block row R, column block C <= 2R + 1)
            {
                double *yp = ypart + wave * ylen;
                for (int r = i1 + lane; r < n; r += 64) yp[r] = 0.0;
                const int T = (n - i1 + 63) >> 6;
                int t = 0;
#pragma unroll 1
                for (int R = 0; R < T; ++R)
#pragma unroll 1
                    for (int C = 0; C <= 2 * R + 1; ++C, ++t) {
                        if ((t & (kSyWaves - 1)) != wave) continue;
                        const int r = i1 + 64 * R + lane, cb = i1 + 32 * C;
                        if (cb >= n) continue;
                        // column offset from the block row's first row: >= 0 only on the
                        // two blocks that hold the diagonal; entries above it are not read
                        const int off = 32 * C - 64 * R;
                        const bool rv = r < n;
                        // every load in bounds (row and column clamped to n - 1; entries
                        // above the diagonal are read and discarded), so they issue together
                        const double *arow = A + (rv ? r : n - 1);
                        double a[32];
#pragma unroll
                        for (int c = 0; c < 32; ++c) {
                            const int col = cb + c < n ? cb + c : n - 1;
                            a[c] = arow[(int64_t)col * lda];
                        }
#pragma unroll
                        for (int c = 0; c < 32; ++c) a[c] = (rv && cb + c < n && off + c <= lane) ? a[c] : 0.0;
                        double acc = 0.0;
#pragma unroll
                        for (int c = 0; c < 32; ++c) acc += a[c] * vl[cb + c];
                        const double vrow = rv ? vl[r] : 0.0;
#pragma unroll
                        for (int c = 0; c < 32; ++c) a[c] = off + c < lane ? a[c] * vrow : 0.0;
                        const double ts0 = transpose_sum16(a, lane);
                        const double ts1 = transpose_sum16(a + 16, lane);
                        if (rv) yp[r] += acc;
                        if (lane < 16 && cb + lane < n) yp[cb + lane] += ts0;
                        if (lane < 16 && cb + 16 + lane < n) yp[cb + 16 + lane] += ts1;
                    }
            }
            lds_barrier();
            // (c3) w = tau (y - V (W^T v) - W (V^T v)), then w -= tau/2 (w^T v) v
            const double s64 = jj > 0 ? red64[lane] : 0.0;
            double wl[kSyRU], part = 0.0;
#pragma unroll
            for (int u = 0; u < kSyRU; ++u) {
                const int r = tid + kSyThreads * u;
                wl[u] = 0.0;
                if (r >= i1 && r < n) {
                    double y = 0.0;
#pragma unroll
                    for (int w = 0; w < kSyWaves; ++w) y += ypart[w * ylen + r];
                    for (int q = 0; q < jj; ++q)
                        y -= A[r + (int64_t)(j0 + q) * lda] * readlane_f64(s64, q) +
                             Wg[(int64_t)q * n + r] * readlane_f64(s64, kSyNB + q);
                    wl[u] = tv * y;
                    part += wl[u] * vr[u];
                }
            }
            part = wsum(part);
            if (lane == 0) red[wave] = part;
            lds_barrier();
            double wv = 0.0;
#pragma unroll
            for (int w = 0; w < kSyWaves; ++w) wv += red[w];
            const double a2 = -0.5 * tv * wv;
#pragma unroll
            for (int u = 0; u < kSyRU; ++u) {
                const int r = tid + kSyThreads * u;
                if (r >= i1 && r < n) {
                    const double w = wl[u] + a2 * vr[u];
                    Wg[(int64_t)jj * n + r] = w;
                    if (r == i1 && jj + 1 < pw) {  // row i1's panel entries for the next column
                        for (int q = 0; q < jj; ++q) {
                            rowW[q] = Wg[(int64_t)q * n + r];
                            rowV[q] = A[r + (int64_t)(j0 + q) * lda];
                        }
                        rowW[jj] = w;
                        rowV[jj] = 1.0;
                    }
                }
            }
            // a full barrier: the next column's dot products read this column's W and V
            // entries written by other threads
            __syncthreads();
        }
        // (e) trailing rows/columns t0..n-1: A -= V W^T + W V^T (lower triangle)
        __syncthreads();
        const int t0 = j0 + pw;
        if (t0 < n) {
            double *stg = ypart + wave * ylen;  // [16 columns][kSyStage]: V(c, 0:32), W(c, 0:32)
            const int TB = (n - t0 + 63) >> 6;
            for (int R = wave; R < TB; R += kSyWaves) {
                const int r = t0 + 64 * R + lane;
                const bool rv = r < n;
                double vrw[kSyNB], wrw[kSyNB];
                const int rc = rv ? r : n - 1;
#pragma unroll
                for (int q = 0; q < kSyNB; ++q) {
                    const int qc = q < pw ? q : 0;
                    vrw[q] = A[rc + (int64_t)(j0 + qc) * lda];
                    wrw[q] = Wg[(int64_t)qc * n + rc];
                }
#pragma unroll
                for (int q = 0; q < kSyNB; ++q) {
                    vrw[q] = (rv && q < pw) ? vrw[q] : 0.0;
                    wrw[q] = (rv && q < pw) ? wrw[q] : 0.0;
                }
                const int cend = t0 + 64 * R + 64 < n ? t0 + 64 * R + 64 : n;
                for (int c0 = t0; c0 < cend; c0 += 16) {
                    // lane (c, qg) stages column c0 + (lane & 15), entries q = (lane >> 4) + 4 k
#pragma unroll
                    for (int k = 0; k < 16; ++k) {
                        const int c = lane & 15, q = (lane >> 4) + 4 * k, col = c0 + c;
                        double val = 0.0;
                        if (col < n) {
                            if (q < kSyNB) {
                                if (q < pw) val = A[col + (int64_t)(j0 + q) * lda];
                            } else if (q - kSyNB < pw) {
                                val = Wg[(int64_t)(q - kSyNB) * n + col];
                            }
                        }
                        stg[c * kSyStage + q] = val;
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
#pragma unroll 2
                    for (int c = 0; c < 16; ++c) {
                        const int col = c0 + c;
                        if (rv && col <= r && col < n) {
                            const double *sv = stg + c * kSyStage;
                            double acc = 0.0;
#pragma unroll
                            for (int q = 0; q < kSyNB; ++q) acc += vrw[q] * sv[kSyNB + q] + wrw[q] * sv[q];
                            A[r + (int64_t)col * lda] -= acc;
                        }
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
                }
            }
        }
        __syncthreads();
        // the subdiagonal back into A, as LAPACK leaves it
        if (tid < pw) A[(j0 + tid + 1) + (int64_t)(j0 + tid) * lda] = panE[tid];
        __syncthreads();
    }
    if (tid == 0) D[n - 1] = A[(n - 1) + (int64_t)(n - 1) * lda];
}

void sytrd_set_attrs() {
    hipFuncSetAttribute((const void *)k_sytrd_lower, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

void launch_sytrd_lower(double *A, int64_t sA, int lda, double *D, int64_t sD, double *E, int64_t sE, double *tau,
                        int64_t sT, double *W, int n, int n_mat, hipStream_t s) {
    if (n_mat <= 0 || n <= 0) return;
    hipLaunchKernelGGL(k_sytrd_lower, dim3(n_mat), dim3(kSyThreads), sytrd_wg_lds(n), s, A, sA, lda, D, sD, E, sE,
                       tau, sT, W, n);
}

size_t sytrd_wg_scratch(int n) { return (size_t)kSyNB * n; }

}  // namespace crest
