// Top-p eigenpairs of a symmetric tridiagonal T (the output of a Householder
// tridiagonalization A = Q T Q^T) and their back-transformation to eigenvectors of A.
// Only the p largest eigenpairs are computed, which is all tl.pca reads
// (kernels.rs:324-345 keep n_comps of the d eigenpairs).
//
//  1. k_tri_bisect: eigenvalue m (ascending index) by multisection on Sturm counts:
//     each lane of a wave counts the eigenvalues below one of 64 points of the current
//     interval, so every round shrinks it 65-fold; stops at the bisection accuracy
//     ulp * ||T|| absolute (relative for large eigenvalues).
//  2. k_tri_vectors: inverse iteration (T - lambda I) x = b with a partially pivoted
//     LU of the tridiagonal, one workgroup per cluster of close eigenvalues (gap below
//     1e-3 ||T||): the cluster's vectors iterate together, one lane each with its own
//     shift (coincident eigenvalues pushed apart by 10 ulp ||T||), and after every
//     round they are orthogonalised in order (Gram-Schmidt, twice) and normalised;
//     rounds stop once every residual 1/||x|| has been at rounding level twice.
//  3. k_tri_back: X = Q Y with Q = H(0) H(1) ... H(n-2), the reflectors of a lower
//     tridiagonalization (H(i) = I - tau_i v v^T, v = [0..0, 1, A(i+2:n, i)]), applied
//     to a block of columns held in registers, last reflector first.
#include <cmath>

#include "common.hpp"

namespace crest {

constexpr double kUlp = 2.220446049250313e-16;
constexpr double kSafeMin = 2.2250738585072014e-308;

__device__ inline double wave_max_f64(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wave_min_f64(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// N wave sums at once, level by level, so the lane exchanges of one level are all in
// flight together instead of N dependent chains.
template <int N> __device__ inline void wave_sum_n(double (&v)[N]) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        double t[N];
#pragma unroll
        for (int c = 0; c < N; ++c) t[c] = __shfl_xor(v[c], o, 64);
#pragma unroll
        for (int c = 0; c < N; ++c) v[c] += t[c];
    }
}

// Number of eigenvalues of T below x (diagonal d, squared off-diagonal e2): the
// signs of the pivots of the LDL^T factorisation of T - x I, tiny pivots replaced by
// -pivmin so that the count stays monotone in x.
__device__ inline int sturm_count(const double *d, const double *e2, int n, double x, double pivmin) {
    double q = d[0] - x;
    if (fabs(q) < pivmin) q = -pivmin;
    int c = q < 0.0 ? 1 : 0;
    for (int i = 1; i < n; ++i) {
        const double t = e2[i - 1] / q;
        q = (d[i] - x) - t;
        if (fabs(q) < pivmin) q = -pivmin;
        c += q < 0.0 ? 1 : 0;
    }
    return c;
}

// Gershgorin interval [gl, gu], ||T||_inf and the pivot floor, identical in every lane.
__device__ inline void tri_bounds(const double *d, const double *e, int n, double &gl, double &gu, double &tnorm,
                                  double &pivmin) {
    const int lane = threadIdx.x & 63;
    double lo = INFINITY, hi = -INFINITY, nrm = 0.0, e2max = 0.0;
    for (int i = lane; i < n; i += 64) {
        const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0);
        lo = fmin(lo, d[i] - r);
        hi = fmax(hi, d[i] + r);
        nrm = fmax(nrm, fabs(d[i]) + r);
        if (i + 1 < n) e2max = fmax(e2max, e[i] * e[i]);
    }
    gl = wave_min_f64(lo);
    gu = wave_max_f64(hi);
    tnorm = wave_max_f64(nrm);
    pivmin = kSafeMin * fmax(1.0, wave_max_f64(e2max));
}

// grid (ceil(p/4), n_mat), 256 threads; dynamic LDS 2n doubles.  Wave w of block b
// finds eigenvalue n - p + 4b + w (ascending) of matrix blockIdx.y; W[j] is the j-th
// of the p largest in ascending order.
__global__ __launch_bounds__(256) void k_tri_bisect(const double *D, int64_t sD, const double *E, int64_t sE, int n,
                                                    int p, double *W, int64_t sW) {
    extern __shared__ double sm[];
    double *d = sm, *e2 = sm + n;
    const int mat = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    D += mat * sD;
    E += mat * sE;
    W += mat * sW;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        d[i] = D[i];
        e2[i] = i + 1 < n ? E[i] * E[i] : 0.0;
    }
    __syncthreads();
    const int j = blockIdx.x * 4 + wave;
    if (j >= p) return;
    double gl, gu, tnorm, pivmin;
    tri_bounds(D, E, n, gl, gu, tnorm, pivmin);
    const double fudge = 2.0 * kUlp * tnorm * n + 2.0 * pivmin;
    double lo = gl - fudge, hi = gu + fudge;
    const int m = n - p + j;  // the count below lambda_m is at most m
    const double atol = kUlp * fmax(fabs(lo), fabs(hi));
    for (int round = 0; round < 64; ++round) {
        const double tol = fmax(fmax(atol, 2.0 * pivmin), 2.0 * kUlp * fmax(fabs(lo), fabs(hi)));
        if (hi - lo <= tol) break;
        const double x = lo + (hi - lo) * ((double)(lane + 1) / 65.0);
        const int c = sturm_count(d, e2, n, x, pivmin);
        const uint64_t above = ~__ballot(c <= m);  // points above lambda_m
        const int f = above ? __builtin_ctzll(above) : 64;  // the first of them
        const double nlo = f > 0 ? __shfl(x, f - 1, 64) : lo;
        const double nhi = f < 64 ? __shfl(x, f, 64) : hi;
        lo = nlo;
        hi = nhi;
    }
    if (lane == 0) W[j] = 0.5 * (lo + hi);
}

// Cluster bounds of the p eigenvalues w[0..p) (ascending): j starts a cluster when
// w[j] - w[j-1] > ortol.  Returns false when cluster `c` does not exist.
__device__ inline bool cluster_of(const double *w, int p, double ortol, int c, int &c0, int &c1) {
    int k = -1;
    for (int j = 0; j < p; ++j) {
        if (j == 0 || w[j] - w[j - 1] > ortol) {
            ++k;
            if (k == c) c0 = j;
            if (k == c + 1) {
                c1 = j;
                return true;
            }
        }
    }
    if (k == c) {
        c1 = p;
        return true;
    }
    return false;
}

// Counter-based start vector in [-1, 1].
__device__ inline double start_value(int i, int j) {
    uint32_t h = (uint32_t)i * 0x9E3779B1u ^ ((uint32_t)j + 0x7F4A7C15u) * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return (double)h * (2.0 / 4294967295.0) - 1.0;
}

// Chunk of the steps of one lane's sequential recurrences: operands of CH steps are
// loaded into registers before any of them is used, so their latency overlaps.
constexpr int kTriChunk = 16;

// grid (p, n_mat), NT threads (64 or 256); dynamic LDS tri_vectors_lds(n).
// Workgroup c computes the vectors of cluster c (if it exists) of matrix blockIdx.y:
// lane t of wave 0 iterates vector c0 + t of the cluster (64 at a time), all with
// their own shifts at once, and after every round the cluster's vectors are
// orthogonalised in order (Gram-Schmidt, twice) and normalised by all the waves,
// rows spread over the threads.  Per matrix, the LU factors (fac:
// 3 arrays [n][p]: reciprocal pivots, U's first superdiagonal, multipliers; swp
// [n][p] bytes: row swaps) and the vectors Y [n][p] are stored row-interleaved, so
// the lanes' accesses of one step are contiguous.  info[mat] is set when a vector's
// solve overflows or never reaches rounding-level residual.
template <int NT>
__global__ __launch_bounds__(NT) void k_tri_vectors(const double *D, int64_t sD, const double *E, int64_t sE, int n,
                                                    int p, const double *W, int64_t sW, double *Y, int64_t sY,
                                                    double *fac, int64_t sF, uint8_t *swpb, int64_t sB, int *info) {
    extern __shared__ double sm[];
    double *d = sm, *e = sm + n;
    __shared__ double inv_s[64];
    __shared__ double gs_red[2][NT / 64][16];
    const int mat = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // sum over the waves of 16 values already summed within each wave, in wave order;
    // the two LDS slots alternate, so a slot is rewritten only after the barrier of the
    // reduction in between, which every thread passes after reading it
    int red_i = 0;
    auto block_sum16 = [&](double (&v)[16]) {
        if constexpr (NT == 64) return;  // one wave: its sums are the block's
        double *slot = &gs_red[red_i & 1][0][0];
        ++red_i;
        if (lane == 0)
#pragma unroll
            for (int q = 0; q < 16; ++q) slot[wave * 16 + q] = v[q];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            double a = slot[q];
#pragma unroll
            for (int w = 1; w < NT / 64; ++w) a += slot[16 * w + q];
            v[q] = a;
        }
    };
    D += mat * sD;
    E += mat * sE;
    W += mat * sW;
    Y += mat * sY;
    fac += mat * sF;
    swpb += mat * sB;
    double *rinv = fac, *u1 = fac + (int64_t)n * p, *mul = fac + 2 * (int64_t)n * p;
    double gl, gu, tnorm, pivmin;
    tri_bounds(D, E, n, gl, gu, tnorm, pivmin);
    int c0 = 0, c1 = 0;
    if (!cluster_of(W, p, 1e-3 * tnorm, blockIdx.x, c0, c1)) return;
    const double tiny = kUlp * tnorm > 0.0 ? kUlp * tnorm : kSafeMin;
    // couplings at rounding level split T (a perturbation no larger than rounding), so
    // no pivot of the factorisation is one of them
    for (int i = tid; i < n; i += NT) {
        d[i] = D[i];
        e[i] = i + 1 < n && fabs(E[i]) > tiny ? E[i] : 0.0;
    }
    __syncthreads();
    const double pertol = 10.0 * kUlp * tnorm;
    const double restol = 8.0 * n * kUlp * tnorm;
    bool failed = false;
    for (int g0 = c0; g0 < c1; g0 += 64) {
        const int g1 = c1 - g0 < 64 ? c1 : g0 + 64;
        const int j = g0 + lane;
        const bool act = wave == 0 && j < g1;
        if (act) {
            // shift: eigenvalues closer than pertol to the previous (shifted) one of the
            // cluster are moved up by pertol, so no two solves share a shift
            double lam = W[c0];
            for (int jj = c0 + 1; jj <= j; ++jj) {
                const double w = W[jj];
                lam = w - lam < pertol ? lam + pertol : w;
            }
            // LU with partial pivoting of T - lam I; U's second superdiagonal is e[k+1]
            // exactly where rows k, k+1 were swapped
            double r0 = d[0] - lam, r1 = n > 1 ? e[0] : 0.0;
            for (int k = 0; k + 1 < n; ++k) {
                const double s0 = e[k], s1 = d[k + 1] - lam, s2 = k + 2 < n ? e[k + 1] : 0.0;
                const int64_t o = (int64_t)k * p + j;
                if (fabs(s0) > fabs(r0)) {
                    const double m = r0 / s0;
                    rinv[o] = 1.0 / s0;
                    u1[o] = s1;
                    swpb[o] = 1;
                    mul[o] = m;
                    const double t = m * s1;
                    r0 = r1 - t;
                    r1 = -(m * s2);
                } else {
                    if (r0 == 0.0) r0 = tiny;
                    const double m = s0 / r0;
                    rinv[o] = 1.0 / r0;
                    u1[o] = r1;
                    swpb[o] = 0;
                    mul[o] = m;
                    const double t = m * r1;
                    r0 = s1 - t;
                    r1 = s2;
                }
            }
            if (r0 == 0.0) r0 = tiny;
            rinv[(int64_t)(n - 1) * p + j] = 1.0 / r0;
            for (int k = 0; k < n; ++k) Y[(int64_t)k * p + j] = start_value(k, j);
        }
        int conv = 0;
        for (int it = 0; it < 8; ++it) {
            if (act) {
                // forward: the row operations of the factorisation applied to x
                double cur = Y[j];
                int k0 = 0;
                for (; k0 + kTriChunk < n; k0 += kTriChunk) {
                    double nx[kTriChunk], mm[kTriChunk];
                    uint8_t sw[kTriChunk];
#pragma unroll
                    for (int u = 0; u < kTriChunk; ++u) {
                        const int64_t o = (int64_t)(k0 + u) * p + j;
                        nx[u] = Y[o + p];
                        mm[u] = mul[o];
                        sw[u] = swpb[o];
                    }
#pragma unroll
                    for (int u = 0; u < kTriChunk; ++u) {
                        const int64_t o = (int64_t)(k0 + u) * p + j;
                        if (sw[u]) {
                            Y[o] = nx[u];
                            const double t = mm[u] * nx[u];
                            cur = cur - t;
                        } else {
                            Y[o] = cur;
                            const double t = mm[u] * cur;
                            cur = nx[u] - t;
                        }
                    }
                }
                for (int k = k0; k + 1 < n; ++k) {
                    const int64_t o = (int64_t)k * p + j;
                    const double nxt = Y[o + p], m = mul[o];
                    if (swpb[o]) {
                        Y[o] = nxt;
                        const double t = m * nxt;
                        cur = cur - t;
                    } else {
                        Y[o] = cur;
                        const double t = m * cur;
                        cur = nxt - t;
                    }
                }
                // back substitution with U
                double xk1 = cur * rinv[(int64_t)(n - 1) * p + j], xk2 = 0.0;
                Y[(int64_t)(n - 1) * p + j] = xk1;
                int k = n - 2;
                for (; k - kTriChunk + 1 >= 0; k -= kTriChunk) {
                    double xv[kTriChunk], uv[kTriChunk], rv[kTriChunk], ev[kTriChunk];
                    uint8_t sw[kTriChunk];
#pragma unroll
                    for (int u = 0; u < kTriChunk; ++u) {
                        const int64_t o = (int64_t)(k - u) * p + j;
                        xv[u] = Y[o];
                        uv[u] = u1[o];
                        rv[u] = rinv[o];
                        sw[u] = swpb[o];
                        ev[u] = e[k - u + 1];
                    }
#pragma unroll
                    for (int u = 0; u < kTriChunk; ++u) {
                        const double a = uv[u] * xk1;
                        const double b = sw[u] ? ev[u] * xk2 : 0.0;
                        const double v = ((xv[u] - a) - b) * rv[u];
                        Y[(int64_t)(k - u) * p + j] = v;
                        xk2 = xk1;
                        xk1 = v;
                    }
                }
                for (; k >= 0; --k) {
                    const int64_t o = (int64_t)k * p + j;
                    const double a = u1[o] * xk1;
                    const double b = swpb[o] ? e[k + 1] * xk2 : 0.0;
                    const double v = ((Y[o] - a) - b) * rinv[o];
                    Y[o] = v;
                    xk2 = xk1;
                    xk1 = v;
                }
            }
            __threadfence_block();
            __syncthreads();
            // Gram-Schmidt over the cluster in order, twice: column t loses its components
            // along columns c0..t-1, taken 16 at a time so that each lane reads 16
            // consecutive values of its rows; rows spread over the lanes
            for (int t = g0; t < g1; ++t) {
                for (int pass = 0; pass < 2; ++pass)
                    for (int s0 = c0; s0 < t; s0 += 16) {
                        const int ns = t - s0 < 16 ? t - s0 : 16;
                        double acc[16];
#pragma unroll
                        for (int q = 0; q < 16; ++q) acc[q] = 0.0;
#pragma unroll 4
                        for (int r = tid; r < n; r += NT) {
                            const double *row = Y + (int64_t)r * p;
                            const double xt = row[t];
#pragma unroll
                            for (int q = 0; q < 16; ++q)
                                if (q < ns) acc[q] += row[s0 + q] * xt;
                        }
                        wave_sum_n<16>(acc);
                        block_sum16(acc);
#pragma unroll 4
                        for (int r = tid; r < n; r += NT) {
                            double *row = Y + (int64_t)r * p;
                            double v = row[t];
#pragma unroll
                            for (int q = 0; q < 16; ++q)
                                if (q < ns) v -= acc[q] * row[s0 + q];
                            row[t] = v;
                        }
                    }
                double nx = 0.0;
                for (int r = tid; r < n; r += NT) {
                    const double v = Y[(int64_t)r * p + t];
                    nx += v * v;
                }
                {
                    double one[16] = {wave_sum(nx)};
                    block_sum16(one);
                    nx = sqrt(one[0]);
                }
                const bool ok = nx > 0.0 && isfinite(nx);
                const double inv = ok ? 1.0 / nx : 0.0;
                for (int r = tid; r < n; r += NT) Y[(int64_t)r * p + t] *= inv;
                if (tid == 0) inv_s[t - g0] = ok ? inv : INFINITY;
                __threadfence_block();
                __syncthreads();
            }
            bool all = true, bad = false;
            for (int t = 0; t < g1 - g0; ++t) {
                all = all && inv_s[t] <= restol;
                bad = bad || !(inv_s[t] < INFINITY);
            }
            if (bad) {
                failed = true;
                break;
            }
            if (all && ++conv == 2) break;
            __syncthreads();
        }
        if (conv == 0) failed = true;
    }
    if (failed && tid == 0) atomicOr(info + mat, 1);
}

// grid (ceil(p/CB), n_mat), 256 threads.
// X[:, c] = Q Y[:, c] for the CB columns of this block (Y [n][p] row-major, X column-major).
// Thread t owns rows t + 256 u (u < RT, n <= 256 RT) of the CB columns, in registers;
// per reflector the CB dot products are reduced over the four waves through a
// double-buffered LDS slot (one barrier per reflector, which waits for the LDS writes
// only, so loads in flight stay in flight), and the thread's entries of the
// reflectors two steps ahead are loaded while the current one is applied.
template <int CB, int RT>
__global__ __launch_bounds__(256) void k_tri_back(const double *A, int64_t sA, int lda, const double *tau,
                                                  int64_t sTau, int n, int p, const double *Y, int64_t sY, double *X,
                                                  int64_t sX) {
    __shared__ double red[2][4][CB];
    const int mat = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    A += mat * sA;
    tau += mat * sTau;
    Y += mat * sY;
    X += mat * sX;
    const int cb0 = blockIdx.x * CB;
    const int ncol = p - cb0 < CB ? p - cb0 : CB;
    double y[CB][RT];
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
        for (int u = 0; u < RT; ++u) {
            const int r = tid + 256 * u;
            y[c][u] = (r < n && c < ncol) ? Y[(int64_t)r * p + cb0 + c] : 0.0;
        }
    // v entries of reflector i for the thread's rows: 1 at row i+1, A(r, i) below, 0 above
    auto load_v = [&](int i, double (&v)[RT], double &t) {
        const double *col = A + (int64_t)i * lda;
#pragma unroll
        for (int u = 0; u < RT; ++u) {
            const int r = tid + 256 * u;
            v[u] = (r < n && r > i + 1) ? col[r] : (r == i + 1 ? 1.0 : 0.0);
        }
        t = tau[i];
    };
    // reflector i applied to the block: dot products, their sum over the four waves
    // through a double-buffered LDS slot (one barrier that waits for LDS only), update
    auto apply = [&](int i, const double (&v)[RT], double t) __attribute__((always_inline)) {
        double *slot = &red[i & 1][0][0];
        double part[CB];
#pragma unroll
        for (int c = 0; c < CB; ++c) {
            double a0 = 0.0, a1 = 0.0;
#pragma unroll
            for (int u = 0; u < RT; u += 2) {
                a0 += v[u] * y[c][u];
                if (u + 1 < RT) a1 += v[u + 1] * y[c][u + 1];
            }
            part[c] = a0 + a1;
        }
        wave_sum_n<CB>(part);
        if (lane == 0)
#pragma unroll
            for (int c = 0; c < CB; ++c) slot[wave * CB + c] = part[c];
        // s_waitcnt lgkmcnt(0) (vmcnt and expcnt left at their maxima), then the barrier
        __builtin_amdgcn_s_waitcnt(0xC07F);
        __builtin_amdgcn_s_barrier();
#pragma unroll
        for (int c = 0; c < CB; ++c) {
            const double w = t * (((slot[c] + slot[CB + c]) + slot[2 * CB + c]) + slot[3 * CB + c]);
#pragma unroll
            for (int u = 0; u < RT; ++u) y[c][u] -= w * v[u];
        }
    };
    // four register sets in a ring: the entries of the reflector three steps ahead are
    // loaded while the current one is applied
    double v0[RT], v1[RT], v2[RT], v3[RT], t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
    if (n >= 2) load_v(n - 2, v0, t0);
    if (n >= 3) load_v(n - 3, v1, t1);
    if (n >= 4) load_v(n - 4, v2, t2);
    for (int i = n - 2; i >= 0; i -= 4) {
        if (i >= 3) load_v(i - 3, v3, t3);
        apply(i, v0, t0);
        if (i < 1) break;
        if (i >= 4) load_v(i - 4, v0, t0);
        apply(i - 1, v1, t1);
        if (i < 2) break;
        if (i >= 5) load_v(i - 5, v1, t1);
        apply(i - 2, v2, t2);
        if (i < 3) break;
        if (i >= 6) load_v(i - 6, v2, t2);
        apply(i - 3, v3, t3);
    }
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
        for (int u = 0; u < RT; ++u) {
            const int r = tid + 256 * u;
            if (r < n && c < ncol) X[(int64_t)(cb0 + c) * n + r] = y[c][u];
        }
}

// The same X = Q Y with the reflectors taken four at a time: one reduction per group
// gives the four blocks of dot products v_a^T y and the six products v_a^T v_b, and
// the group's four updates follow from the recurrence
//   w_a = tau_a (v_a^T y - sum_{b<a} w_b v_a^T v_b),  y -= w_0 v_0 + ... + w_3 v_3
// (v_0 = reflector i, applied first, down to v_3 = reflector i-3), so there is one
// barrier per four reflectors.  The next group's entries load during the current
// group's update.  grid (ceil(p/CB), n_mat), 256 threads, n <= 256 RT.
template <int CB, int RT>
__global__ __launch_bounds__(256) void k_tri_back4(const double *A, int64_t sA, int lda, const double *tau,
                                                   int64_t sTau, int n, int p, const double *Y, int64_t sY, double *X,
                                                   int64_t sX) {
    constexpr int NV = 4 * CB + 6;  // dot products and reflector products per group
    __shared__ double red[2][4][NV];
    const int mat = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    A += mat * sA;
    tau += mat * sTau;
    Y += mat * sY;
    X += mat * sX;
    const int cb0 = blockIdx.x * CB;
    const int ncol = p - cb0 < CB ? p - cb0 : CB;
    double y[CB][RT];
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
        for (int u = 0; u < RT; ++u) {
            const int r = tid + 256 * u;
            y[c][u] = (r < n && c < ncol) ? Y[(int64_t)r * p + cb0 + c] : 0.0;
        }
    // v entries of reflector i for the thread's rows (1 at row i+1, A(r, i) below, 0
    // above); reflector i < 0 is the identity (v = 0, tau = 0)
    auto load_v = [&](int i, double (&v)[RT], double &t) __attribute__((always_inline)) {
        const int ic = i >= 0 ? i : 0;
        const double *col = A + (int64_t)ic * lda;
#pragma unroll
        for (int u = 0; u < RT; ++u) {
            const int r = tid + 256 * u;
            v[u] = (i >= 0 && r < n && r > i + 1) ? col[r] : ((i >= 0 && r == i + 1) ? 1.0 : 0.0);
        }
        t = i >= 0 ? tau[ic] : 0.0;
    };
    double va[4][RT], vb[4][RT], ta[4], tb[4];
    auto group = [&](int i, int g, double (&v)[4][RT], const double (&t)[4], double (&vn)[4][RT], double (&tn)[4])
                     __attribute__((always_inline)) {
        double part[NV];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < CB; ++c) {
                double s0 = 0.0;
#pragma unroll
                for (int u = 0; u < RT; ++u) s0 += v[a][u] * y[c][u];
                part[a * CB + c] = s0;
            }
        int k = 4 * CB;
#pragma unroll
        for (int a = 1; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < a; ++b) {
                double s0 = 0.0;
#pragma unroll
                for (int u = 0; u < RT; ++u) s0 += v[a][u] * v[b][u];
                part[k++] = s0;
            }
        wave_sum_n<NV>(part);
        double *slot = &red[g & 1][0][0];
        if (lane == 0)
#pragma unroll
            for (int j = 0; j < NV; ++j) slot[wave * NV + j] = part[j];
        // the next group's entries load while this group is reduced and applied
#pragma unroll
        for (int a = 0; a < 4; ++a) load_v(i - 4 - a, vn[a], tn[a]);
        // s_waitcnt lgkmcnt(0) (vmcnt and expcnt left at their maxima), then the barrier
        __builtin_amdgcn_s_waitcnt(0xC07F);
        __builtin_amdgcn_s_barrier();
        double gm[6];
#pragma unroll
        for (int j = 0; j < 6; ++j)
            gm[j] = ((slot[4 * CB + j] + slot[NV + 4 * CB + j]) + slot[2 * NV + 4 * CB + j]) + slot[3 * NV + 4 * CB + j];
        // gm: v1.v0, v2.v0, v2.v1, v3.v0, v3.v1, v3.v2
#pragma unroll
        for (int c = 0; c < CB; ++c) {
            double sv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a)
                sv[a] = ((slot[a * CB + c] + slot[NV + a * CB + c]) + slot[2 * NV + a * CB + c]) +
                        slot[3 * NV + a * CB + c];
            const double w0 = t[0] * sv[0];
            const double w1 = t[1] * (sv[1] - w0 * gm[0]);
            const double w2 = t[2] * ((sv[2] - w0 * gm[1]) - w1 * gm[2]);
            const double w3 = t[3] * (((sv[3] - w0 * gm[3]) - w1 * gm[4]) - w2 * gm[5]);
#pragma unroll
            for (int u = 0; u < RT; ++u) {
                y[c][u] -= w0 * v[0][u];
                y[c][u] -= w1 * v[1][u];
                y[c][u] -= w2 * v[2][u];
                y[c][u] -= w3 * v[3][u];
            }
        }
    };
#pragma unroll
    for (int a = 0; a < 4; ++a) load_v(n - 2 - a, va[a], ta[a]);
    int g = 0;
    for (int i = n - 2; i >= 0; i -= 8, g += 2) {
        group(i, g, va, ta, vb, tb);
        if (i - 4 < 0) break;
        group(i - 4, g + 1, vb, tb, va, ta);
    }
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
        for (int u = 0; u < RT; ++u) {
            const int r = tid + 256 * u;
            if (r < n && c < ncol) X[(int64_t)(cb0 + c) * n + r] = y[c][u];
        }
}

// Scratch of launch_tri_topp per matrix: 3 n p doubles (fac) and n p bytes (swp).
// LDS limits of the three kernels for order n (0 when the order is too large).
size_t tri_vectors_lds(int n) { return (size_t)2 * n * sizeof(double); }
bool tri_topp_supported(int n) { return n <= 4096 && tri_vectors_lds(n) <= 150 * 1024; }

void tri_set_attrs() {
    hipFuncSetAttribute((const void *)k_tri_bisect, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
    hipFuncSetAttribute((const void *)k_tri_vectors<64>, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
    hipFuncSetAttribute((const void *)k_tri_vectors<256>, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
}

// Top-p eigenpairs of n_mat tridiagonalised matrices: A (reflectors, lower, lda,
// stride sA), tau (stride sTau), D/E (diagonal / off-diagonal, strides sD/sE).  W
// [n_mat][p] eigenvalues ascending, Y scratch and X output [n_mat][n x p] column-major
// eigenvectors of A (ascending), info [n_mat] zeroed by the caller.
void launch_tri_topp(const double *A, int64_t sA, int lda, const double *tau, int64_t sTau, const double *D,
                     int64_t sD, const double *E, int64_t sE, int n, int p, int n_mat, double *W, double *Y,
                     double *X, double *fac, uint8_t *swp, int *info, hipStream_t s) {
    const int64_t sW = p, sY = (int64_t)n * p;
    hipLaunchKernelGGL(k_tri_bisect, dim3((p + 3) / 4, n_mat), dim3(256), (size_t)2 * n * sizeof(double), s, D, sD, E,
                       sE, n, p, W, sW);
    // one matrix: four waves share each cluster's Gram-Schmidt sweeps (47.0 -> 44.0 ms
    // at n = 2,000, p = 50); a batch keeps one wave per workgroup, whose grid already
    // fills the GPU (256 threads measured 5 ms slower over 256 subsets)
    if ((int64_t)n_mat * p < 256)
        hipLaunchKernelGGL(k_tri_vectors<256>, dim3(p, n_mat), dim3(256), tri_vectors_lds(n), s, D, sD, E, sE, n, p,
                           (const double *)W, sW, Y, sY, fac, 3 * sY, swp, sY, info);
    else
        hipLaunchKernelGGL(k_tri_vectors<64>, dim3(p, n_mat), dim3(64), tri_vectors_lds(n), s, D, sD, E, sE, n, p,
                           (const double *)W, sW, Y, sY, fac, 3 * sY, swp, sY, info);
    // columns per block: the thread's rows of CB columns stay in registers.  Fewer
    // blocks than CUs (one matrix): four reflectors per barrier; a batch fills the GPU
    // with one reflector per barrier at twice the occupancy.
    if (n <= 2048 && (int64_t)n_mat * ((p + 3) / 4) < 256) {
        // one column per block while the blocks fit on the CUs: the per-group reduction
        // shrinks to 4 + 6 values (the n = 2,000, p = 50 eigensolve: 42.7 ms, against
        // 44.2 with four columns per block and 43.1 with two)
        if ((int64_t)n_mat * p <= 256)
            hipLaunchKernelGGL((k_tri_back4<1, 8>), dim3(p, n_mat), dim3(256), 0, s, A, sA, lda, tau, sTau, n, p, Y,
                               sY, X, sY);
        else
            hipLaunchKernelGGL((k_tri_back4<4, 8>), dim3((p + 3) / 4, n_mat), dim3(256), 0, s, A, sA, lda, tau, sTau,
                               n, p, Y, sY, X, sY);
    } else if (n <= 2048)
        hipLaunchKernelGGL((k_tri_back<8, 8>), dim3((p + 7) / 8, n_mat), dim3(256), 0, s, A, sA, lda, tau, sTau, n,
                           p, Y, sY, X, sY);
    else
        hipLaunchKernelGGL((k_tri_back<4, 16>), dim3((p + 3) / 4, n_mat), dim3(256), 0, s, A, sA, lda, tau, sTau, n,
                           p, Y, sY, X, sY);
}

}  // namespace crest
