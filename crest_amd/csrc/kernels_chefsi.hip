// Kernels of the Chebyshev-filtered subspace iteration for the p largest eigenpairs of
// the d x d covariance (the FP64 eigensolve behind pca_from_gram, kernels.rs:311-348).
// A block of p + 16 vectors is filtered with Chebyshev polynomials of the covariance
// that damp the spectrum below the block's smallest Ritz value, orthonormalised by
// CholeskyQR2 and rotated to Ritz vectors; converged leading pairs are locked and
// deflated out of the operator.  The products with the covariance are rocBLAS dgemm
// calls split over K into partial products; these kernels sum the partials where they
// are consumed and do the small dense pieces.  (Filtered subspace iteration: Zhou &
// Saad, SIAM J. Matrix Anal. Appl. 29 (2007) 954-971; CholeskyQR2: Fukaya,
// Nakatsukasa, Yanagisawa & Yamamoto, 2014; parallel cyclic Jacobi: Golub & Van Loan,
// Matrix Computations, 4th ed., 8.5.)
#include "common.hpp"

namespace crest {

constexpr int kChMaxK = 80;  // largest block (the small dense kernels work in LDS)
constexpr int kChThreads = 1024;

// A deterministic start block (rows x cols, leading dimension ld): values in [-1, 1)
// from the splitmix64 hash of (seed, row, column); rows rows .. ld - 1 are zero.
__global__ void k_ch_rand(double *v, int rows, int ld, int cols, uint64_t seed) {
    const int64_t n = (int64_t)ld * cols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e % ld, c = e / ld;
        double out = 0.0;
        if (r < rows) {
            uint64_t z = seed + 0x9e3779b97f4a7c15ull * (uint64_t)(c * rows + r + 1);
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            z ^= z >> 31;
            out = (double)(z >> 11) * (1.0 / 4503599627370496.0) - 1.0;
        }
        v[e] = out;
    }
}

// out[i] = sum over s < S of part[s * stride + i]
__global__ void k_ch_sum(double *out, const double *part, int S, int64_t stride, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v = part[i];
        for (int s = 1; s < S; ++s) v += part[s * stride + i];
        out[i] = v;
    }
}

// One step of the three-term recurrence: y_next = a1 (w - c y) - a2 y_prev, where
// w = sum of S partial products of the operator with y (stride apart); y_prev ==
// nullptr for the first step (a2 unused).
__global__ void k_ch_combine(double *y_next, const double *w, int S, int64_t stride, const double *y,
                             const double *y_prev, int64_t n, double a1, double c, double a2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double wi = w[i];
        for (int s = 1; s < S; ++s) wi += w[s * stride + i];
        double v = a1 * (wi - c * y[i]);
        if (y_prev) v -= a2 * y_prev[i];
        y_next[i] = v;
    }
}

typedef double f64x4_ch __attribute__((ext_vector_type(4)));

// Partial products of the symmetric dp x dp operator A (column-major) with the block
// Y (dp x n, leading dimension dp): out[s] (dp x n, column-major) = A Y over rows
// k0 .. k0 + kc - 1 of Y, k0 = s kc (kc a multiple of 32).  Grid (dp / 64, S,
// ceil(n / 32)): a workgroup (256 threads) takes rows r0 .. r0 + 63 and columns
// c0 .. c0 + 31, so that a block of 72 columns still gives three workgroups per CU to
// hide the loads.  Stages of 32 k: the workgroup loads A(r0 .. r0 + 63, k) -- stored at
// row dp + k by symmetry, so each row is a contiguous run -- and Y(k, c0 .. c0 + 31)
// with 16-byte loads, the next stage's into registers while the current one
// multiplies out of LDS, where both sit k-major so that the MFMA operand reads are
// consecutive.  Wave w takes rows r0 + 16 w .. + 15 on v_mfma_f64_16x16x4f64 (lane l
// supplies P[l & 15][l >> 4] = A(row, k) and Q[l >> 4][l & 15] = Y(k, col), and holds
// D[4 v + (l >> 4)][l & 15] in register v).
// One workgroup's tile: rows r0 .. r0 + 63, columns c0 .. c0 + 31 of A Y over k0 .. k0 +
// kc - 1, into o (column-major, dp rows).
__device__ inline void ch_gemm_tile(const double *__restrict__ A, const double *__restrict__ Y, int dp, int n, int r0,
                                    int k0, int kc, int c0, double *__restrict__ o) {
    constexpr int KS = 32, NC = 32, LA = 64 + 4, LY = NC + 4;  // k per stage, columns; LDS row lengths
    __shared__ double As[2][KS * LA], Ys[2][KS * LY];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, li = lane & 15, lk = lane >> 4;
    // this thread's share of a stage: A row ar, k offsets 8 ap .. 8 ap + 7; Y column
    // c0 + yc, k offsets 8 yp .. 8 yp + 7 for the first 128 threads... four double2 each
    const int ar = tid >> 2, ap = tid & 3;
    const double2 *ga = reinterpret_cast<const double2 *>(A + (int64_t)(r0 + ar) * dp + k0 + 8 * ap);
    const int yc = (tid & 127) >> 2, yp = tid & 3, yh = tid >> 7;  // yh: which half of the 8 k
    const bool yok = c0 + yc < n;
    const double2 *gy = reinterpret_cast<const double2 *>(Y + (int64_t)(yok ? c0 + yc : 0) * dp + k0 + 8 * yp + 4 * yh);
    double2 ra[4], ry[2];
    auto fetch = [&](int kb) __attribute__((always_inline)) {
#pragma unroll
        for (int u = 0; u < 4; ++u) ra[u] = ga[kb / 2 + u];
#pragma unroll
        for (int u = 0; u < 2; ++u) ry[u] = yok ? gy[kb / 2 + u] : make_double2(0.0, 0.0);
    };
    auto stash = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            As[buf][(8 * ap + 2 * u) * LA + ar] = ra[u].x;
            As[buf][(8 * ap + 2 * u + 1) * LA + ar] = ra[u].y;
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            Ys[buf][(8 * yp + 4 * yh + 2 * u) * LY + yc] = ry[u].x;
            Ys[buf][(8 * yp + 4 * yh + 2 * u + 1) * LY + yc] = ry[u].y;
        }
    };
    f64x4_ch acc[2] = {(f64x4_ch){0.0, 0.0, 0.0, 0.0}, (f64x4_ch){0.0, 0.0, 0.0, 0.0}};
    fetch(0);
    stash(0);
    __syncthreads();
    int cur = 0;
    for (int kb = 0; kb < kc; kb += KS) {
        const bool more = kb + KS < kc;
        if (more) fetch(kb + KS);
        const double *as = As[cur], *ys = Ys[cur];
#pragma unroll
        for (int st = 0; st < KS / 4; ++st) {
            const int kr = 4 * st + lk;
            const double a = as[kr * LA + 16 * w + li];
            acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, ys[kr * LY + li], acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, ys[kr * LY + 16 + li], acc[1], 0, 0, 0);
        }
        if (more) {
            stash(cur ^ 1);
            __syncthreads();
            cur ^= 1;
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int col = c0 + 16 * t + li;
        if (col < n)
#pragma unroll
            for (int v = 0; v < 4; ++v) o[(int64_t)col * dp + r0 + 16 * w + 4 * v + lk] = acc[t][v];
    }
}

// The same product for a whole block of up to 16 NG columns per workgroup (rows r0 ..
// r0 + 63, all K = dp): each row panel of A is read once for all the block's columns.
// Stage loads: Y(k0 .. k0 + 31, c) for the 16 NG columns as 16 NG x 16 double2, thread
// t taking double2 u 256 + t (16 threads per column, its 256 contiguous bytes).
template <int NG>
__device__ inline void ch_gemm_panel(const double *__restrict__ A, const double *__restrict__ Y, int dp, int n,
                                     int r0, double *__restrict__ o, int k0 = 0, int kc = -1) {
    if (kc < 0) kc = dp;
    A += k0;
    Y += k0;
    constexpr int KS = 32, NC = 16 * NG, LA = 64 + 4, LY = NC + 4;
    __shared__ double As[2][KS * LA], Ys[2][KS * LY];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, li = lane & 15, lk = lane >> 4;
    const int ar = tid >> 2, ap = tid & 3;
    const double2 *ga = reinterpret_cast<const double2 *>(A + (int64_t)(r0 + ar) * dp + 8 * ap);
    const double2 *gy[NG];
    bool yok[NG];
#pragma unroll
    for (int u = 0; u < NG; ++u) {
        const int q = u * 256 + tid, yc = q >> 4, kp = q & 15;
        yok[u] = yc < n;
        gy[u] = reinterpret_cast<const double2 *>(Y + (int64_t)(yok[u] ? yc : 0) * dp + 2 * kp);
    }
    double2 ra[4], ry[NG];
    auto fetch = [&](int kb) __attribute__((always_inline)) {
#pragma unroll
        for (int u = 0; u < 4; ++u) ra[u] = ga[kb / 2 + u];
#pragma unroll
        for (int u = 0; u < NG; ++u) ry[u] = yok[u] ? gy[u][kb / 2] : make_double2(0.0, 0.0);
    };
    auto stash = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            As[buf][(8 * ap + 2 * u) * LA + ar] = ra[u].x;
            As[buf][(8 * ap + 2 * u + 1) * LA + ar] = ra[u].y;
        }
#pragma unroll
        for (int u = 0; u < NG; ++u) {
            const int q = u * 256 + tid, yc = q >> 4, kp = q & 15;
            Ys[buf][(2 * kp) * LY + yc] = ry[u].x;
            Ys[buf][(2 * kp + 1) * LY + yc] = ry[u].y;
        }
    };
    f64x4_ch acc[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) acc[g] = (f64x4_ch){0.0, 0.0, 0.0, 0.0};
    fetch(0);
    stash(0);
    __syncthreads();
    int cur = 0;
    for (int kb = 0; kb < kc; kb += KS) {
        const bool more = kb + KS < kc;
        if (more) fetch(kb + KS);
        const double *as = As[cur], *ys = Ys[cur];
#pragma unroll
        for (int st = 0; st < KS / 4; ++st) {
            const int kr = 4 * st + lk;
            const double a = as[kr * LA + 16 * w + li];
#pragma unroll
            for (int g = 0; g < NG; ++g)
                acc[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, ys[kr * LY + 16 * g + li], acc[g], 0, 0, 0);
        }
        if (more) {
            stash(cur ^ 1);
            __syncthreads();
            cur ^= 1;
        }
    }
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int col = 16 * g + li;
        if (col < n)
#pragma unroll
            for (int v = 0; v < 4; ++v) o[(int64_t)col * dp + r0 + 16 * w + 4 * v + lk] = acc[g][v];
    }
}

// One matrix: one workgroup per (64-row panel, K part) over all n <= 16 NG columns;
// part s writes its partial product to out + s dp n.
template <int NG>
__global__ __launch_bounds__(256) void k_ch_gemm_panel(const double *__restrict__ A, const double *__restrict__ Y,
                                                       int dp, int n, int kc, double *__restrict__ out) {
    ch_gemm_panel<NG>(A, Y, dp, n, blockIdx.x * 64, out + (int64_t)blockIdx.y * dp * n, blockIdx.y * kc, kc);
}

__global__ __launch_bounds__(256) void k_ch_gemm(const double *__restrict__ A, const double *__restrict__ Y, int dp,
                                                 int n, int kc, double *__restrict__ out) {
    ch_gemm_tile(A, Y, dp, n, blockIdx.x * 64, blockIdx.y * kc, kc, blockIdx.z * 32, out + (int64_t)blockIdx.y * dp * n);
}

// 1 / x and 1 / sqrt(x) in FP64 from the hardware estimates and one Newton step
// (a few ulp): the rotations and pivots need c^2 + s^2 = 1 and r^2 = a to rounding,
// not correctly rounded quotients.
__device__ inline double ch_rcp(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return r * (2.0 - x * r);
}
__device__ inline double ch_rsq(double x) {
    const double r = __builtin_amdgcn_rsq(x);
    return r * (1.5 - 0.5 * x * r * r);
}

constexpr int kChCholThreads = 256;

// Rinv = R^-1 for the Cholesky factor of the k x k Gram G = Y^T Y, the sum of S
// partials (column-major, k * k apart).  With D = diag(G)^-1/2 the factor of the
// column-normalised Gn = D G D = Rt^T Rt is taken, so R = Rt D^-1 and Rinv = D Rt^-1:
// columns of very different lengths (the filter amplifies each direction by its own
// factor) cost no accuracy.  When Gn is within 1e-9 of the identity everywhere (the
// second pass of CholeskyQR2), Rt^-1 = I - E_upper - diag(E) / 2 with E = Gn - I, to
// within |E|^2 (second-order) terms, is taken without factoring.  Otherwise a pivot
// below 1e-13 (the block lost rank) sets *bad = 1 and leaves Rinv unwritten.  One
// barrier per column of the right-looking factorization and per row of the inversion
// (Rt^-1 from the last row up); lanes run along the contiguous index and each thread
// loads a batch of operands before it stores.  LDS 2 k^2 doubles.
// G is read with leading dimension ld (partials pstride apart) and Rinv written with
// leading dimension ld, zero in rows and columns k .. ld - 1 (ld = k for one matrix).
__device__ inline void ch_chol_rinv(const double *Gp, int S, int k, int64_t pstride, int ld, double *Rinv,
                                    double *bad) {
    extern __shared__ double lds[];
    double *L = lds;          // [k][k] column-major: Schur complements; then X = Rt^-1 (row-major, upper)
    double *F = lds + k * k;  // [k][k] column-major: the lower factor Rt^T
    __shared__ double dsc[kChMaxK], rdg[kChMaxK];  // column scales; pivots of Rt
    __shared__ int far_from_identity;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) far_from_identity = 0;
    for (int e = tid; e < k * k; e += blockDim.x) {
        const int64_t ge = (int64_t)(e / k) * ld + e % k;
        double g = Gp[ge];
        for (int s = 1; s < S; ++s) g += Gp[s * pstride + ge];
        L[e] = g;
    }
    for (int e = tid; e < ld * ld; e += blockDim.x)  // the padding of Rinv
        if (e / ld >= k || e % ld >= k) Rinv[e] = 0.0;
    __syncthreads();
    for (int i = tid; i < k; i += blockDim.x) {
        const double g = L[i * k + i];
        dsc[i] = g > 0.0 ? ch_rsq(g) : 0.0;
    }
    __syncthreads();
    for (int e = tid; e < k * k; e += blockDim.x) {
        const int c = e / k, r = e % k;
        const double g = L[e] * dsc[r] * dsc[c];
        L[e] = g;
        if (fabs(g - (r == c ? 1.0 : 0.0)) > 1e-9) far_from_identity = 1;
    }
    __syncthreads();
    if (!far_from_identity) {
        for (int e = tid; e < k * k; e += blockDim.x) {
            const int c = e / k, i = e % k;
            const double x = i < c ? -L[e] : i == c ? 1.0 - 0.5 * (L[e] - 1.0) : 0.0;
            Rinv[(int64_t)c * ld + i] = dsc[i] * x;
        }
        return;
    }
    for (int j = 0; j < k; ++j) {
        const double piv = L[j * k + j];
        if (!(piv > 1e-13)) {  // uniform: every thread read the same value after the barrier
            if (tid == 0) *bad = 1.0;
            return;
        }
        const double rinv = ch_rsq(piv), r = piv * rinv;
        if (tid == 0) rdg[j] = r;
        for (int i = j + tid; i < k; i += blockDim.x) F[j * k + i] = (i == j) ? r : L[j * k + i] * rinv;
        // a_il -= (a_ij / r)(a_lj / r) for j < l <= i: lanes over rows i (consecutive in
        // the column-major L, so conflict-free), wave w over columns l = j + 1 + w + 4 u,
        // five columns per batch of loads
        const int base = j + 1;
        for (int h = 0; h < 2; ++h) {
            const int i = base + lane + 64 * h;
            if (base + 64 * h >= k) break;
            const bool iok = i < k;
            const double li = iok ? L[j * k + i] * rinv : 0.0;
            for (int l0 = base + wave; l0 < k; l0 += 20) {
                double ll[5], cur[5];
#pragma unroll
                for (int u = 0; u < 5; ++u) {
                    const int l = l0 + 4 * u;
                    ll[u] = l < k ? L[j * k + l] * rinv : 0.0;
                    cur[u] = (iok && l <= i) ? L[l * k + i] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 5; ++u) {
                    const int l = l0 + 4 * u;
                    if (iok && l <= i) L[l * k + i] = cur[u] - li * ll[u];
                }
            }
        }
        __syncthreads();
    }
    // Rt X = I with Rt[i][l] = F[i * k + l]: X (row-major in L) starts as I; at step i
    // row i is final up to the division by Rt[i][i], and rows l < i take its share.
    // Finished rows go to F's upper triangle (F[c k + i], c >= i), which no later step
    // reads; Rinv is written once at the end (a global store inside the loop would make
    // every barrier wait for it).
    for (int e = tid; e < k * k; e += blockDim.x) L[e] = (e / k == e % k) ? 1.0 : 0.0;
    __syncthreads();
    for (int i = k - 1; i >= 0; --i) {
        const double di = ch_rcp(rdg[i]);
        // lanes over columns c >= i (consecutive in the row-major X), wave w over rows
        // l = w + 4 u < i, five rows per batch of loads
        for (int h = 0; h < 2; ++h) {
            const int c = i + lane + 64 * h;
            if (i + 64 * h >= k) break;
            const bool cok = c < k;
            const double xi = cok ? L[i * k + c] * di : 0.0;
            if (wave == 0 && cok) F[c * k + i] = xi;  // row i of Rt^-1 into F's upper triangle
            for (int l0 = wave; l0 < i; l0 += 20) {
                double fl[5], cur[5];
#pragma unroll
                for (int u = 0; u < 5; ++u) {
                    const int l = l0 + 4 * u;
                    fl[u] = l < i ? F[l * k + i] : 0.0;
                    cur[u] = (cok && l < i) ? L[l * k + c] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 5; ++u) {
                    const int l = l0 + 4 * u;
                    if (cok && l < i) L[l * k + c] = cur[u] - fl[u] * xi;
                }
            }
        }
        __syncthreads();
    }
    for (int e = tid; e < k * k; e += blockDim.x) {
        const int c = e / k, i = e % k;
        Rinv[(int64_t)c * ld + i] = i <= c ? dsc[i] * F[e] : 0.0;
    }
}

__global__ __launch_bounds__(kChCholThreads) void k_ch_chol_rinv(const double *Gp, int S, int k, double *Rinv,
                                                                double *bad) {
    ch_chol_rinv(Gp, S, k, (int64_t)k * k, k, Rinv, bad);
}

// Eigenpairs of the symmetric n x n matrix H = sym(sum of S partials, column-major,
// n * n apart) by the two-sided cyclic Jacobi method in round-robin order: each round
// rotates m / 2 disjoint pairs (m = n rounded up to even; an odd n gets a zero
// row/column whose rotations are skipped).  A round has two phases: one thread per
// pair computes its rotation, then one thread per 2 x 2 block pair (a <= b) rotates
// the block and writes it and its transpose, and the columns of V rotate; pairs whose
// rotation is the identity are skipped.  A pair is rotated while |h_pq| > 2e-16
// sqrt(|h_pp h_qq|) and > 1e-18 ||H||_F.  At most max_sweeps sweeps run; with strict
// set, not converging within them sets *bad = 1.  Output: theta (descending when desc,
// else ascending) and Z [n x n] column-major, column j the eigenvector of theta[j].
// One workgroup; LDS 2 m (m + 1) doubles.
// H is read with leading dimension ld and Z written with leading dimension ld, zero in
// rows and columns n .. ld - 1 (ld = n for one matrix).
__device__ inline void ch_jacobi(const double *Hp, int S, int n, int ld_h, int max_sweeps, int strict, int desc,
                                 double *theta, double *Z, double *bad) {
    extern __shared__ double lds[];
    const int m = n + (n & 1), ld = m + 1, half = m / 2, m1 = m - 1;
    double *A = lds;           // [m][ld], row-major
    double *V = lds + m * ld;  // [m][ld], row-major: V[i][j]
    __shared__ double rc[kChMaxK / 2], rs[kChMaxK / 2], rt[kChMaxK / 2];
    __shared__ int rp[kChMaxK / 2], rq[kChMaxK / 2];
    __shared__ double red[kChThreads / 64];
    __shared__ int todo;
    const int tid = threadIdx.x;
    const int64_t nn = (int64_t)n * n;
    double fro = 0.0;
    for (int e = tid; e < m * m; e += blockDim.x) {
        const int i = e / m, j = e % m;
        double h = 0.0;
        if (i < n && j < n) {
            double a = Hp[(int64_t)j * ld_h + i], b = Hp[(int64_t)i * ld_h + j];
            for (int s = 1; s < S; ++s) {
                a += Hp[s * nn + (int64_t)j * ld_h + i];
                b += Hp[s * nn + (int64_t)i * ld_h + j];
            }
            h = 0.5 * (a + b);
        }
        A[i * ld + j] = h;
        V[i * ld + j] = (i == j) ? 1.0 : 0.0;
        fro += h * h;
    }
    for (int o = 32; o > 0; o >>= 1) fro += __shfl_xor(fro, o);
    if ((tid & 63) == 0) red[tid >> 6] = fro;
    __syncthreads();
    fro = 0.0;
    for (int w = 0; w < (int)(blockDim.x / 64); ++w) fro += red[w];
    const double abs_tol = 1e-18 * sqrt(fro);
    // roles: block (ba, bb), ba <= bb, of the upper triangle of pair slots; V column
    // pair vb on rows vr0, vr0 + vstep, ...
    int ba = -1, bb = -1;
    if (tid < half * (half + 1) / 2) {
        int rem = tid, a = 0;
        while (rem >= half - a) {
            rem -= half - a;
            ++a;
        }
        ba = a;
        bb = a + rem;
    }
    const int vstep = (int)blockDim.x / half;
    const int vb = tid % half, vr0 = tid / half;
    const bool vrole = vr0 < vstep;
    bool converged = false;
    for (int sweep = 0; sweep < max_sweeps; ++sweep) {
        if (tid == 0) todo = 0;
        __syncthreads();
        for (int e = tid; e < m * m; e += blockDim.x) {
            const int i = e / m, j = e % m;
            if (i < j) {
                const double apq = fabs(A[i * ld + j]);
                if (apq > abs_tol && apq > 2e-16 * sqrt(fabs(A[i * ld + i] * A[j * ld + j]))) todo = 1;
            }
        }
        __syncthreads();
        if (!todo) {
            converged = true;
            break;
        }
        for (int r = 0; r < m1; ++r) {
            if (tid < half) {  // slot tid of round r: players tid and m - 1 - tid, rotated by r
                int u = tid == 0 ? 0 : tid - 1 + r, w = m - 2 - tid + r;
                if (u >= m1) u -= m1;
                if (w >= m1) w -= m1;
                int p = tid == 0 ? 0 : 1 + u, q = 1 + w;
                if (p > q) {
                    const int x = p;
                    p = q;
                    q = x;
                }
                const double app = A[p * ld + p], aqq = A[q * ld + q], apq = A[p * ld + q];
                double c = 1.0, s = 0.0, t = 0.0;
                if (fabs(apq) > abs_tol && fabs(apq) > 2e-16 * sqrt(fabs(app * aqq))) {
                    const double tau = (aqq - app) * ch_rcp(2.0 * apq);
                    if (fabs(tau) > 1e150) {
                        t = 0.5 * ch_rcp(tau);
                    } else {
                        const double h2 = 1.0 + tau * tau;
                        t = (tau >= 0.0 ? 1.0 : -1.0) * ch_rcp(fabs(tau) + h2 * ch_rsq(h2));
                    }
                    c = ch_rsq(1.0 + t * t);
                    s = t * c;
                }
                rp[tid] = p;
                rq[tid] = q;
                rc[tid] = c;
                rs[tid] = s;
                rt[tid] = t;
            }
            __syncthreads();
            if (ba >= 0) {
                const int pa = rp[ba], qa = rq[ba];
                const double ta = rt[ba];
                if (ba == bb) {
                    if (ta != 0.0) {
                        const double apq = A[pa * ld + qa];
                        A[pa * ld + pa] -= ta * apq;
                        A[qa * ld + qa] += ta * apq;
                        A[pa * ld + qa] = 0.0;
                        A[qa * ld + pa] = 0.0;
                    }
                } else {
                    const int pb = rp[bb], qb = rq[bb];
                    const double tb = rt[bb];
                    if (ta != 0.0 || tb != 0.0) {
                        const double ca = rc[ba], sa = rs[ba], cb = rc[bb], sb = rs[bb];
                        const double b00 = A[pa * ld + pb], b01 = A[pa * ld + qb], b10 = A[qa * ld + pb],
                                     b11 = A[qa * ld + qb];
                        const double r00 = ca * b00 - sa * b10, r01 = ca * b01 - sa * b11;
                        const double r10 = sa * b00 + ca * b10, r11 = sa * b01 + ca * b11;
                        const double n00 = r00 * cb - r01 * sb, n01 = r00 * sb + r01 * cb;
                        const double n10 = r10 * cb - r11 * sb, n11 = r10 * sb + r11 * cb;
                        A[pa * ld + pb] = n00;
                        A[pa * ld + qb] = n01;
                        A[qa * ld + pb] = n10;
                        A[qa * ld + qb] = n11;
                        A[pb * ld + pa] = n00;
                        A[qb * ld + pa] = n01;
                        A[pb * ld + qa] = n10;
                        A[qb * ld + qa] = n11;
                    }
                }
            }
            if (vrole && rt[vb] != 0.0) {
                const int p = rp[vb], q = rq[vb];
                const double c = rc[vb], s = rs[vb];
                for (int i = vr0; i < m; i += vstep) {
                    const double x = V[i * ld + p], y = V[i * ld + q];
                    V[i * ld + p] = x * c - y * s;
                    V[i * ld + q] = x * s + y * c;
                }
            }
            __syncthreads();
        }
    }
    if (strict && !converged && tid == 0) *bad = 1.0;
    for (int i = tid; i < n; i += blockDim.x) {  // order of the diagonal (ties by index)
        const double di = A[i * ld + i];
        int rank = 0;
        for (int j = 0; j < n; ++j) {
            const double dj = A[j * ld + j];
            rank += (desc ? dj > di : dj < di) || (dj == di && j < i);
        }
        theta[rank] = di;
        for (int r = 0; r < ld_h; ++r) Z[(int64_t)rank * ld_h + r] = r < n ? V[r * ld + i] : 0.0;
    }
    for (int e = tid; e < (ld_h - n) * ld_h; e += blockDim.x) Z[(int64_t)n * ld_h + e] = 0.0;
}

__global__ __launch_bounds__(kChThreads) void k_ch_jacobi(const double *Hp, int S, int n, int max_sweeps, int strict,
                                                         int desc, double *theta, double *Z, double *bad) {
    ch_jacobi(Hp, S, n, n, max_sweeps, strict, desc, theta, Z, bad);
}

// out[j] = || cv[:, j] - theta[j] v[:, j] ||_2 over d rows (leading dimension ld), one
// workgroup per column.
__global__ __launch_bounds__(256) void k_ch_resid(const double *cv, const double *v, const double *theta, int d,
                                                 int ld, double *out) {
    __shared__ double part[256];
    const int j = blockIdx.x;
    const double t = theta[j];
    double s = 0.0;
    for (int i = threadIdx.x; i < d; i += blockDim.x) {
        const double r = cv[(int64_t)j * ld + i] - t * v[(int64_t)j * ld + i];
        s += r * r;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[j] = sqrt(part[0]);
}

// Locking: dst[:, j] = src[:, j], dst_val[j] = src_val[j] and dst_scaled[:, j] =
// dst[:, j] * dst_val[j] for j < n (the deflation's left factor); columns of ld rows.
__global__ void k_ch_lock(const double *src, const double *src_val, int ld, int n, double *dst, double *dst_val,
                          double *dst_scaled) {
    const int64_t total = (int64_t)ld * n;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(e / ld), i = (int)(e % ld);
        const double t = src_val[j], v = src[(int64_t)j * ld + i];
        dst[e] = v;
        dst_scaled[e] = v * t;
        if (i == 0) dst_val[j] = t;
    }
}

// ---------------------------------------------------------------- batches
// The same iteration for many matrices side by side (chefsi_topp_batch): matrix z's
// operator at A + z dp^2, its blocks at buf + z dp K (K = the block size, column-major,
// dp rows), its small matrices at G + z K^2, its statistics at stat + z (2K + 2).  The
// per-matrix state of the current iteration is a ChbMat: ka active columns (0: the
// matrix takes no part), the filter degree m, the first active column voff of the
// Ritz block after locking, the locked count L, the columns J locked at the start of
// this iteration, and the filter's centre c and half-width e.

// rows r0.. x columns c0.. of A_z Y_z[:, yoff ..] for the matrices whose product is
// live at this step (step < m; step < 0: any active matrix)
// (one workgroup per 64-row panel of each matrix, all its active columns)
template <int NG>
__global__ __launch_bounds__(256) void k_chb_gemm(const double *__restrict__ A, const double *__restrict__ Y,
                                                  int dp, int K, const ChbMat *P, int step, int use_voff,
                                                  double *__restrict__ out) {
    const int z = blockIdx.y;
    const ChbMat m = P[z];
    const int n = (m.ka > 0 && (step < 0 || step < m.m)) ? m.ka : 0;
    if (n == 0) return;
    const int64_t dk = (int64_t)dp * K;
    ch_gemm_panel<NG>(A + (int64_t)z * dp * dp, Y + z * dk + (use_voff ? (int64_t)m.voff * dp : 0), dp, n,
                      blockIdx.x * 64, out + z * dk);
}

// The three-term recurrence for every matrix: step 0: y1 = (w - c y) / e (w = A y,
// both at column voff); step j >= 1 while j < m: y_next = 2 (w - c y) / e - y_prev
// (y_prev at column voff when prev_voff); j >= m: y_next = y (the matrix's filter is
// done; its result moves along with the buffer rotation).  Columns at or past ka are
// set to 0.
__global__ void k_chb_combine(double *y_next, const double *w, const double *y, const double *y_prev, int dp, int K,
                              int nm, const ChbMat *P, int step, int prev_voff) {
    const int64_t dk = (int64_t)dp * K, total = dk * nm;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(e / dk);
        const int64_t r = e % dk;
        const int j = (int)(r / dp);
        const ChbMat m = P[z];
        double v = 0.0;
        if (j < m.ka) {
            const int64_t base = z * dk;
            if (step == 0) {
                const int64_t o = base + r + (int64_t)m.voff * dp;
                v = (w[o] - m.c * y[o]) * (1.0 / m.e);
            } else if (step < m.m) {
                const double yp = y_prev[base + r + (prev_voff ? (int64_t)m.voff * dp : 0)];
                v = (2.0 / m.e) * (w[base + r] - m.c * y[base + r]) - yp;
            } else {
                v = y[base + r];
            }
        }
        y_next[e] = v;
    }
}

// the columns of an active matrix at or past ka set to 0 (so that no stale value meets
// a padded factor in a product)
__global__ void k_chb_zero_tail(double *y, int dp, int K, int nm, const ChbMat *P) {
    const int64_t dk = (int64_t)dp * K, total = dk * nm;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(e / dk);
        if ((int)((e % dk) / dp) >= P[z].ka) y[e] = 0.0;
    }
}

__global__ __launch_bounds__(kChCholThreads) void k_chb_chol_rinv(const double *G, int K, const ChbMat *P,
                                                                 double *Rinv, double *stat) {
    const int z = blockIdx.x, ka = P[z].ka;
    if (ka == 0) return;
    const int64_t kk = (int64_t)K * K;
    ch_chol_rinv(G + z * kk, 1, ka, kk, K, Rinv + z * kk, stat + (int64_t)z * (2 * K + 2) + 2 * K);
}

__global__ __launch_bounds__(kChThreads) void k_chb_jacobi(const double *H, int K, const ChbMat *P, int max_sweeps,
                                                          int strict, int desc, double *Z, double *stat) {
    const int z = blockIdx.x, ka = P[z].ka;
    if (ka == 0) return;
    const int64_t kk = (int64_t)K * K;
    double *st = stat + (int64_t)z * (2 * K + 2);
    ch_jacobi(H + z * kk, 1, ka, K, max_sweeps, strict, desc, st, Z + z * kk, st + 2 * K + 1);
}

// residual norms || cv[:, j] - theta_j v[:, j] || over d rows, j < ka
__global__ __launch_bounds__(256) void k_chb_resid(const double *cv, const double *v, int d, int dp, int K,
                                                  const ChbMat *P, double *stat) {
    __shared__ double part[256];
    const int j = blockIdx.x, z = blockIdx.y;
    if (j >= P[z].ka) return;
    double *st = stat + (int64_t)z * (2 * K + 2);
    const int64_t o = (int64_t)z * dp * K + (int64_t)j * dp;
    const double t = st[j];
    double s = 0.0;
    for (int i = threadIdx.x; i < d; i += blockDim.x) {
        const double r = cv[o + i] - t * v[o + i];
        s += r * r;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) st[K + j] = sqrt(part[0]);
}

// Locking of the first J columns of the Ritz block v (values in stat): QL[:, L + j] =
// v[:, j], thL[L + j] = theta_j; the deflation's factors wl[:, j] = theta_j v[:, j] and
// ql[:, j] = v[:, j] for j < J, zero elsewhere (so A -= wl ql^T over all K columns
// removes exactly the J new directions).
__global__ void k_chb_lock(const double *v, const double *stat, int dp, int K, int p, int nm, const ChbMat *P,
                           double *QL, double *thL, double *wl, double *ql) {
    const int64_t dk = (int64_t)dp * K, total = dk * nm;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(e / dk);
        const int64_t r = e % dk;
        const int j = (int)(r / dp), i = (int)(r % dp);
        const ChbMat m = P[z];
        double a = 0.0, b = 0.0;
        if (j < m.J) {
            const double t = stat[(int64_t)z * (2 * K + 2) + j], x = v[e];
            a = x * t;
            b = x;
            QL[(int64_t)z * dp * p + (int64_t)(m.L + j) * dp + i] = x;
            if (i == 0) thL[(int64_t)z * p + m.L + j] = t;
        }
        wl[e] = a;
        ql[e] = b;
    }
}

// A_z = C_z (d x d, column-major) padded with zeros to dp x dp, for the matrices with
// ka > 0
__global__ void k_chb_pad(const double *C, int d, int dp, int nm, const ChbMat *P, double *A) {
    const int64_t dd = (int64_t)dp * dp, total = dd * nm;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(e / dd);
        if (P[z].ka == 0) continue;
        const int64_t r = e % dd;
        const int c = (int)(r / dp), i = (int)(r % dp);
        A[e] = (c < d && i < d) ? C[(int64_t)z * d * d + (int64_t)c * d + i] : 0.0;
    }
}

// dst[:, j] = src[:, off + j] for j < ka of the matrices with ka > 0 (off = voff when
// use_voff; src leading dimension per matrix: src_cols columns of dp rows)
__global__ void k_chb_copy(double *dst, const double *src, int src_cols, int dp, int K, int nm, const ChbMat *P,
                           int use_voff) {
    const int64_t dk = (int64_t)dp * K, total = dk * nm;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(e / dk);
        const int64_t r = e % dk;
        const int j = (int)(r / dp);
        const ChbMat m = P[z];
        if (m.ka == 0 || j >= m.ka) continue;
        dst[e] = src[(int64_t)z * dp * src_cols + (int64_t)(j + (use_voff ? m.voff : 0)) * dp + r % dp];
    }
}

// the final p vectors (d rows) and values of each matrix, compact for the selection
__global__ void k_chb_final(const double *v, const double *stat, int d, int dp, int K, int p, int nm, double *vout,
                            double *evals) {
    const int64_t dpp = (int64_t)d * p, total = dpp * nm;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(e / dpp);
        const int64_t r = e % dpp;
        const int j = (int)(r / d), i = (int)(r % d);
        vout[e] = v[(int64_t)z * dp * K + (int64_t)j * dp + i];
        if (i == 0) evals[(int64_t)z * p + j] = stat[(int64_t)z * (2 * K + 2) + j];
    }
}

// the deterministic start block of chefsi_topp, the same for every matrix
__global__ void k_chb_rand(double *v, int rows, int ld, int cols, int nm, uint64_t seed) {
    const int64_t n = (int64_t)ld * cols, total = n * nm;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = e % n, r = q % ld, c = q / ld;
        double out = 0.0;
        if (r < rows) {
            uint64_t x = seed + 0x9e3779b97f4a7c15ull * (uint64_t)(c * rows + r + 1);
            x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
            x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
            x ^= x >> 31;
            out = (double)(x >> 11) * (1.0 / 4503599627370496.0) - 1.0;
        }
        v[e] = out;
    }
}

int chefsi_max_block() { return kChMaxK; }
static size_t ch_chol_lds(int k) { return (size_t)2 * k * k * sizeof(double); }
static size_t ch_jacobi_lds(int n) {
    const int m = n + (n & 1);
    return (size_t)2 * m * (m + 1) * sizeof(double);
}
void chefsi_set_attrs() {
    hipFuncSetAttribute((const void *)k_ch_chol_rinv, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)ch_chol_lds(kChMaxK));
    hipFuncSetAttribute((const void *)k_ch_jacobi, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)ch_jacobi_lds(kChMaxK));
}
static int ch_grid(int64_t n) { return (int)std::min<int64_t>(std::max<int64_t>(ceil_div(n, 256), 1), 4096); }
void launch_ch_rand(double *v, int rows, int ld, int cols, uint64_t seed, hipStream_t s) {
    hipLaunchKernelGGL(k_ch_rand, dim3(ch_grid((int64_t)ld * cols)), dim3(256), 0, s, v, rows, ld, cols, seed);
}
void launch_ch_sum(double *out, const double *part, int S, int64_t stride, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_ch_sum, dim3(ch_grid(n)), dim3(256), 0, s, out, part, S, stride, n);
}
void launch_ch_combine(double *y_next, const double *w, int S, int64_t stride, const double *y, const double *y_prev,
                       int64_t n, double a1, double c, double a2, hipStream_t s) {
    hipLaunchKernelGGL(k_ch_combine, dim3(ch_grid(n)), dim3(256), 0, s, y_next, w, S, stride, y, y_prev, n, a1, c,
                       a2);
}
void launch_ch_chol_rinv(const double *Gp, int S, int k, double *Rinv, double *bad, hipStream_t s) {
    hipLaunchKernelGGL(k_ch_chol_rinv, dim3(1), dim3(kChCholThreads), ch_chol_lds(k), s, Gp, S, k, Rinv, bad);
}
void launch_ch_jacobi(const double *Hp, int S, int n, int max_sweeps, bool strict, bool desc, double *theta,
                      double *Z, double *bad, hipStream_t s) {
    hipLaunchKernelGGL(k_ch_jacobi, dim3(1), dim3(kChThreads), ch_jacobi_lds(n), s, Hp, S, n, max_sweeps, (int)strict,
                       (int)desc, theta, Z, bad);
}
void launch_ch_gemm(const double *A, const double *Y, int dp, int n, int S, double *out, hipStream_t s) {
    const dim3 g(dp / 64, S), b(256);
    const int kc = dp / S;
    switch ((n + 15) / 16) {
    case 1: hipLaunchKernelGGL(k_ch_gemm_panel<1>, g, b, 0, s, A, Y, dp, n, kc, out); break;
    case 2: hipLaunchKernelGGL(k_ch_gemm_panel<2>, g, b, 0, s, A, Y, dp, n, kc, out); break;
    case 3: hipLaunchKernelGGL(k_ch_gemm_panel<3>, g, b, 0, s, A, Y, dp, n, kc, out); break;
    case 4: hipLaunchKernelGGL(k_ch_gemm_panel<4>, g, b, 0, s, A, Y, dp, n, kc, out); break;
    case 5: hipLaunchKernelGGL(k_ch_gemm_panel<5>, g, b, 0, s, A, Y, dp, n, kc, out); break;
    default: hipLaunchKernelGGL(k_ch_gemm, dim3(dp / 64, S, (n + 31) / 32), b, 0, s, A, Y, dp, n, kc, out); break;
    }
}
void launch_ch_resid(const double *cv, const double *v, const double *theta, int d, int ld, int k, double *out,
                     hipStream_t s) {
    hipLaunchKernelGGL(k_ch_resid, dim3(k), dim3(256), 0, s, cv, v, theta, d, ld, out);
}
static int chb_grid(int64_t n) { return (int)std::min<int64_t>(std::max<int64_t>(ceil_div(n, 256), 1), 16384); }
void chefsi_batch_set_attrs() {
    hipFuncSetAttribute((const void *)k_chb_chol_rinv, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)ch_chol_lds(kChMaxK));
    hipFuncSetAttribute((const void *)k_chb_jacobi, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)ch_jacobi_lds(kChMaxK));
}
void launch_chb_gemm(const double *A, const double *Y, int dp, int K, int nm, const ChbMat *P, int step, bool use_voff,
                     double *out, hipStream_t s) {
    const dim3 g(dp / 64, nm), b(256);
    switch ((K + 15) / 16) {
    case 1: hipLaunchKernelGGL(k_chb_gemm<1>, g, b, 0, s, A, Y, dp, K, P, step, (int)use_voff, out); break;
    case 2: hipLaunchKernelGGL(k_chb_gemm<2>, g, b, 0, s, A, Y, dp, K, P, step, (int)use_voff, out); break;
    case 3: hipLaunchKernelGGL(k_chb_gemm<3>, g, b, 0, s, A, Y, dp, K, P, step, (int)use_voff, out); break;
    case 4: hipLaunchKernelGGL(k_chb_gemm<4>, g, b, 0, s, A, Y, dp, K, P, step, (int)use_voff, out); break;
    default: hipLaunchKernelGGL(k_chb_gemm<5>, g, b, 0, s, A, Y, dp, K, P, step, (int)use_voff, out); break;
    }
}
void launch_chb_combine(double *y_next, const double *w, const double *y, const double *y_prev, int dp, int K, int nm,
                        const ChbMat *P, int step, bool prev_voff, hipStream_t s) {
    hipLaunchKernelGGL(k_chb_combine, dim3(chb_grid((int64_t)dp * K * nm)), dim3(256), 0, s, y_next, w, y, y_prev, dp,
                       K, nm, P, step, (int)prev_voff);
}
void launch_chb_zero_tail(double *y, int dp, int K, int nm, const ChbMat *P, hipStream_t s) {
    hipLaunchKernelGGL(k_chb_zero_tail, dim3(chb_grid((int64_t)dp * K * nm)), dim3(256), 0, s, y, dp, K, nm, P);
}
void launch_chb_chol_rinv(const double *G, int K, int nm, const ChbMat *P, double *Rinv, double *stat, hipStream_t s) {
    hipLaunchKernelGGL(k_chb_chol_rinv, dim3(nm), dim3(kChCholThreads), ch_chol_lds(K), s, G, K, P, Rinv, stat);
}
void launch_chb_jacobi(const double *H, int K, int nm, const ChbMat *P, int max_sweeps, bool strict, bool desc,
                       double *Z, double *stat, hipStream_t s) {
    hipLaunchKernelGGL(k_chb_jacobi, dim3(nm), dim3(kChThreads), ch_jacobi_lds(K), s, H, K, P, max_sweeps,
                       (int)strict, (int)desc, Z, stat);
}
void launch_chb_resid(const double *cv, const double *v, int d, int dp, int K, int nm, const ChbMat *P, double *stat,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_chb_resid, dim3(K, nm), dim3(256), 0, s, cv, v, d, dp, K, P, stat);
}
void launch_chb_lock(const double *v, const double *stat, int dp, int K, int p, int nm, const ChbMat *P, double *QL,
                     double *thL, double *wl, double *ql, hipStream_t s) {
    hipLaunchKernelGGL(k_chb_lock, dim3(chb_grid((int64_t)dp * K * nm)), dim3(256), 0, s, v, stat, dp, K, p, nm, P, QL,
                       thL, wl, ql);
}
void launch_chb_pad(const double *C, int d, int dp, int nm, const ChbMat *P, double *A, hipStream_t s) {
    hipLaunchKernelGGL(k_chb_pad, dim3(chb_grid((int64_t)dp * dp * nm)), dim3(256), 0, s, C, d, dp, nm, P, A);
}
void launch_chb_copy(double *dst, const double *src, int src_cols, int dp, int K, int nm, const ChbMat *P,
                     bool use_voff, hipStream_t s) {
    hipLaunchKernelGGL(k_chb_copy, dim3(chb_grid((int64_t)dp * K * nm)), dim3(256), 0, s, dst, src, src_cols, dp, K,
                       nm, P, (int)use_voff);
}
void launch_chb_final(const double *v, const double *stat, int d, int dp, int K, int p, int nm, double *vout,
                      double *evals, hipStream_t s) {
    hipLaunchKernelGGL(k_chb_final, dim3(chb_grid((int64_t)d * p * nm)), dim3(256), 0, s, v, stat, d, dp, K, p, nm, vout,
                       evals);
}
void launch_chb_rand(double *v, int rows, int ld, int cols, int nm, uint64_t seed, hipStream_t s) {
    hipLaunchKernelGGL(k_chb_rand, dim3(chb_grid((int64_t)ld * cols * nm)), dim3(256), 0, s, v, rows, ld, cols, nm,
                       seed);
}
void launch_ch_lock(const double *src, const double *src_val, int ld, int n, double *dst, double *dst_val,
                    double *dst_scaled, hipStream_t s) {
    hipLaunchKernelGGL(k_ch_lock, dim3(ch_grid((int64_t)ld * n)), dim3(256), 0, s, src, src_val, ld, n, dst, dst_val,
                       dst_scaled);
}

}  // namespace crest
