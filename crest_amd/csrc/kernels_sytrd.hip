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
// matrix-vector product reads each lower-triangle element once (64-row x 32-column
// tiles; a wave walks down one 32-column block, the row part accumulating per lane,
// the transposed part per lane and column, summed across lanes once per block by a
// halving exchange), so the whole reduction reads ~n^3/6 doubles per matrix, half of
// what a full-storage product reads.  The partial products of the eight waves go to
// per-wave LDS vectors and are summed in wave order, so the result does not depend on
// scheduling.  The trailing update after each panel runs on FP64 MFMA.  Loads go
// through buffer descriptors with the column offset in a scalar register, so a tile's
// 32 loads share one address register.
#include <cmath>

#include "common.hpp"

namespace crest {

namespace {

constexpr int kSyNB = 32;       // panel width
constexpr int kSyWaves = 8;     // 512 threads
constexpr int kSyThreads = kSyWaves * 64;
constexpr int kSyRU = 4;        // rows per thread: row r belongs to thread r % 512
constexpr int kSyMaxN = kSyThreads * kSyRU;
typedef double f64x4 __attribute__((ext_vector_type(4)));

__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Loads through buffer descriptors: a uniform byte offset in a scalar register plus one
// 32-bit vector offset per lane, so a tile's loads share one address register (a matrix
// of order <= 2,048 spans at most 32 MB).  Offsets past the descriptor's range read 0.
typedef __amdgpu_buffer_rsrc_t BufRsrc;
__device__ inline BufRsrc buf_rsrc(const double *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}
__device__ inline double bload(BufRsrc r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ inline void bstore(BufRsrc r, uint32_t voff, uint32_t soff, double v) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v), r, voff, soff, 0);
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

// a[u] -= V(r_u, q) cw(q) + W(r_u, q) cv(q) for q = 0 .. jj-1 in order, for the thread's
// rows r_u = tid + 512 u (clamped to n - 1; the caller discards rows it does not own).
// The loads of four panel columns are issued together.
template <class FW, class FV>
__device__ inline void panel_correct(double (&a)[kSyRU], BufRsrc ra, int lda, BufRsrc rw, int n, int j0, int jj,
                                     int tid, FW cw, FV cv) {
    for (int q0 = 0; q0 < jj; q0 += 4) {
        double vv[kSyRU][4], ww[kSyRU][4];
#pragma unroll
        for (int u = 0; u < kSyRU; ++u) {
            const int r = tid + kSyThreads * u;
            const uint32_t rc = (uint32_t)(r < n ? r : n - 1);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int q = q0 + k < jj ? q0 + k : jj - 1;
                vv[u][k] = bload(ra, rc * 8u, (uint32_t)((j0 + q) * lda) * 8u);
                ww[u][k] = bload(rw, rc * 8u, (uint32_t)(q * n) * 8u);
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (q0 + k < jj) {
                const double x = cw(q0 + k), y = cv(q0 + k);
#pragma unroll
                for (int u = 0; u < kSyRU; ++u) a[u] -= vv[u][k] * x + ww[u][k] * y;
            }
    }
}

}  // namespace

size_t sytrd_wg_lds(int n) {
    const int ldn = (n + 63) & ~63;
    return (size_t)(kSyWaves * ldn + ldn + 64 + kSyWaves * 64 + kSyWaves + 3 * kSyNB + 8) * sizeof(double);
}
bool sytrd_wg_supported(int n) { return n >= 1 && n <= kSyMaxN && sytrd_wg_lds(n) <= 160 * 1024; }

// grid (n_mat), 512 threads, LDS sytrd_wg_lds(n).  Wg: scratch [n_mat][NB][n].
__global__ __launch_bounds__(kSyThreads) void k_sytrd_lower(double *A, int64_t sA, int lda, double *D, int64_t sD,
                                                            double *E, int64_t sE, double *tau, int64_t sT,
                                                            double *Wg, int n) {
    extern __shared__ double lds[];
    const int ldn = (n + 63) & ~63;
    const int ylen = ldn;
    double *ypart = lds;                        // [8][ylen]: the waves' partial products
    // [ldn + 64] current reflector, by row; zero beyond n (the last column tile of the
    // product reads up to 63 entries past the matrix)
    double *vl = ypart + kSyWaves * ylen;
    double *red64 = vl + ldn + 64;              // [64] (sized [8][64])
    double *red = red64 + kSyWaves * 64;        // [8]
    double *rowW = red + kSyWaves;              // [NB] W(i, :) of the next column's row
    double *rowV = rowW + kSyNB;                // [NB] V(i, :)
    double *panE = rowV + kSyNB;                // [NB] E of the panel's columns
    double *sc = panE + kSyNB;                  // [8] broadcast scalars

    // the wave index through readfirstlane, so that values derived from it (work item
    // indices, column offsets) are known to be uniform and live in scalar registers
    const int mat = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    A += mat * sA;
    D += mat * sD;
    E += mat * sE;
    tau += mat * sT;
    Wg += (int64_t)mat * kSyNB * n;
    const BufRsrc ra = buf_rsrc(A, (uint32_t)n * (uint32_t)lda * 8u);
    const BufRsrc rw = buf_rsrc(Wg, (uint32_t)kSyNB * (uint32_t)n * 8u);
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
                colv[u] = bload(ra, (uint32_t)(r < n ? r : n - 1) * 8u, (uint32_t)(i * lda) * 8u);
            }
            panel_correct(colv, ra, lda, rw, n, j0, jj, tid, [&](int q) { return rowW[q]; },
                          [&](int q) { return rowV[q]; });
#pragma unroll
            for (int u = 0; u < kSyRU; ++u) {
                const int r = tid + kSyThreads * u;
                if (r >= i && r < n) {
                    const double a = colv[u];
                    if (r >= i + 2) ss += a * a;
                    if (r == i1) sc[0] = a;
                    if (r == i) {
                        A[r + (int64_t)i * lda] = a;
                        D[i] = a;
                    }
                } else {
                    colv[u] = 0.0;
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
            // (c1 + c2) y = A22 v over the lower triangle of rows/columns i1..n-1, by
            // column blocks of 32 (block C = columns i1 + 32C .., round-robin over the
            // waves): the wave walks down the block in 64-row tiles (block rows R >= C/2),
            // each lane adding its row's part to the wave's LDS vector and accumulating its
            // share of the transposed part per column; one cross-lane sum per column block.
            // Then W^T v and V^T v over the panel's earlier columns, in 64-row x 32-column
            // items (block row R; W(r, 0:32) or V(r, 0:32)) continuing the round-robin,
            // summed into the wave's own red64 row (k < 32: W(:, k)^T v, 32 + q: V(:, q)^T v).
            {
                double *yp = ypart + wave * ylen;
                for (int r = i1 + lane; r < n; r += 64) yp[r] = 0.0;
                red64[wave * 64 + lane] = 0.0;
                const int T = (n - i1 + 63) >> 6;
                const int NC = (n - i1 + 31) >> 5;  // column blocks
                auto load = [&](int R, int C, double (&x)[32]) __attribute__((always_inline)) {
                    const int r = i1 + 64 * R + lane;
                    const uint32_t rc = (uint32_t)(r < n ? r : n - 1);
                    const int cb = i1 + 32 * C;
#pragma unroll
                    for (int c = 0; c < 32; ++c) {
                        const int col = cb + c < n ? cb + c : n - 1;
                        x[c] = bload(ra, rc * 8u, (uint32_t)(col * lda) * 8u);
                    }
                };
                double tacc[32];
                auto tile = [&](int R, int C, double (&x)[32]) __attribute__((always_inline)) {
                    const int r = i1 + 64 * R + lane, cb = i1 + 32 * C;
                    const bool rv = r < n;
                    const double vrow = rv ? vl[r] : 0.0;
                    // column offset from the block row's first row: >= 0 only on the block
                    // row that holds the diagonal; entries above it are dropped
                    const int off = 32 * C - 64 * R;
                    double acc = 0.0;
                    // masks as 0/1 factors (every entry read is finite: the matrix is
                    // fully written and never modified above the diagonal)
#pragma unroll
                    for (int c = 0; c < 32; ++c) {
                        const double xv = x[c] * ((rv && cb + c < n && off + c <= lane) ? 1.0 : 0.0);
                        acc += xv * vl[cb + c];
                        tacc[c] += xv * (off + c < lane ? vrow : 0.0);
                    }
                    if (rv) yp[r] += acc;
                };
#pragma unroll 1
                for (int C = wave; C < NC; C += kSyWaves) {
#pragma unroll
                    for (int c = 0; c < 32; ++c) tacc[c] = 0.0;
#pragma unroll 1
                    for (int R = C >> 1; R < T; ++R) {
                        double x[32];
                        load(R, C, x);
                        tile(R, C, x);
                    }
                    const int cb = i1 + 32 * C;
                    const double ts0 = transpose_sum16(tacc, lane);
                    const double ts1 = transpose_sum16(tacc + 16, lane);
                    if (lane < 16 && cb + lane < n) yp[cb + lane] += ts0;
                    if (lane < 16 && cb + 16 + lane < n) yp[cb + 16 + lane] += ts1;
                }
                // panel items (R, g), numbered after the column blocks
                if (jj > 0)
                    for (int it = (wave - NC % kSyWaves + kSyWaves) % kSyWaves; it < 2 * T; it += kSyWaves) {
                        const int R = it >> 1, g = it & 1;
                        const int r = i1 + 64 * R + lane;
                        const uint32_t rc = (uint32_t)(r < n ? r : n - 1);
                        const bool rv = r < n;
                        double x[32];
                        if (g == 0) {
#pragma unroll
                            for (int c = 0; c < 32; ++c) x[c] = bload(rw, rc * 8u, (uint32_t)((c < jj ? c : 0) * n) * 8u);
                        } else {
#pragma unroll
                            for (int c = 0; c < 32; ++c)
                                x[c] = bload(ra, rc * 8u, (uint32_t)((j0 + (c < jj ? c : 0)) * lda) * 8u);
                        }
                        const double vrow = rv ? vl[r] : 0.0;
#pragma unroll
                        for (int c = 0; c < 32; ++c) x[c] *= (rv && c < jj) ? vrow : 0.0;
                        const double ts0 = transpose_sum16(x, lane);
                        const double ts1 = transpose_sum16(x + 16, lane);
                        if (lane < 16) {
                            red64[wave * 64 + g * kSyNB + lane] += ts0;
                            red64[wave * 64 + g * kSyNB + 16 + lane] += ts1;
                        }
                    }
            }
            lds_barrier();
            // (c3) w = tau (y - V (W^T v) - W (V^T v)), then w -= tau/2 (w^T v) v
            double s64 = 0.0;
            if (jj > 0)
#pragma unroll
                for (int w = 0; w < kSyWaves; ++w) s64 += red64[w * 64 + lane];
            double wl[kSyRU], part = 0.0;
#pragma unroll
            for (int u = 0; u < kSyRU; ++u) {
                const int r = tid + kSyThreads * u, rc = r < n ? r : n - 1;
                double y = 0.0;
#pragma unroll
                for (int w = 0; w < kSyWaves; ++w) y += ypart[w * ylen + rc];
                wl[u] = y;
            }
            panel_correct(wl, ra, lda, rw, n, j0, jj, tid, [&](int q) { return readlane_f64(s64, q); },
                          [&](int q) { return readlane_f64(s64, kSyNB + q); });
#pragma unroll
            for (int u = 0; u < kSyRU; ++u) {
                const int r = tid + kSyThreads * u;
                if (r >= i1 && r < n) {
                    wl[u] = tv * wl[u];
                    part += wl[u] * vr[u];
                } else {
                    wl[u] = 0.0;
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
                        rowW[jj] = w;
                        rowV[jj] = 1.0;
                    }
                }
            }
            // entries q < jj of row i1 (written in earlier columns, so visible after the
            // full barriers since), one per thread
            if (tid < jj && jj + 1 < pw) {
                rowW[tid] = Wg[(int64_t)tid * n + i1];
                rowV[tid] = A[i1 + (int64_t)(j0 + tid) * lda];
            }
            // a full barrier: the next column's dot products read this column's W and V
            // entries written by other threads
            __syncthreads();
        }
        // (e) trailing rows/columns t0..n-1: A -= V W^T + W V^T (lower triangle), on FP64
        // MFMA in 16 x 16 tiles.  Tile (rows r0.., columns c0..): D = P Q with K = 64,
        // P[i][k] = W(c0 + i, k) for k < 32 and V(c0 + i, k - 32) above, Q[k][j] =
        // V(r0 + j, k) for k < 32 and W(r0 + j, k - 32) above, so D[i][j] is the update of
        // A(r0 + j, c0 + i).  v_mfma_f64_16x16x4f64 at step s takes P[l & 15][4s + (l >> 4)]
        // and Q[4s + (l >> 4)][l & 15] from lane l and leaves D[4v + (l >> 4)][l & 15] in its
        // register v: one lane's outputs are consecutive rows of four columns.  A wave
        // takes 16-row strips round-robin; the strip's Q stays in registers and the next
        // tile's P and A entries load while the current tile is computed.
        __syncthreads();
        const int t0 = j0 + pw;
        if (t0 < n) {
            const int S = (n - t0 + 15) >> 4, li = lane & 15, lk = lane >> 4;
            for (int sR = wave; sR < S; sR += kSyWaves) {
                const int row = t0 + 16 * sR + li, rowc = row < n ? row : n - 1;
                const uint32_t span = (uint32_t)(j0 * lda) * 8u;  // byte offset of the panel
                double qb[16];
#pragma unroll
                for (int s2 = 0; s2 < 8; ++s2) {
                    const int q = 4 * s2 + lk, qc = q < pw ? q : 0;
                    qb[s2] = bload(ra, (uint32_t)(qc * lda + rowc) * 8u, span);
                    qb[8 + s2] = bload(rw, (uint32_t)(qc * n + rowc) * 8u, 0u);
                }
#pragma unroll
                for (int s2 = 0; s2 < 8; ++s2) {
                    const bool ok = row < n && 4 * s2 + lk < pw;
                    qb[s2] = ok ? qb[s2] : 0.0;
                    qb[8 + s2] = ok ? qb[8 + s2] : 0.0;
                }
                auto load = [&](int cB, double (&pa)[16], double (&av)[4]) __attribute__((always_inline)) {
                    const int c0 = t0 + 16 * cB, colp = c0 + li, colpc = colp < n ? colp : n - 1;
#pragma unroll
                    for (int s2 = 0; s2 < 8; ++s2) {
                        const int q = 4 * s2 + lk, qc = q < pw ? q : 0;
                        pa[s2] = bload(rw, (uint32_t)(qc * n + colpc) * 8u, 0u);
                        pa[8 + s2] = bload(ra, (uint32_t)(qc * lda + colpc) * 8u, span);
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int col = c0 + 4 * v + lk;
                        av[v] = bload(ra, (uint32_t)((col < n ? col : n - 1) * lda + rowc) * 8u, 0u);
                    }
                };
                auto tile = [&](int cB, double (&pa)[16], const double (&av)[4]) __attribute__((always_inline)) {
                    const int c0 = t0 + 16 * cB, colp = c0 + li;
#pragma unroll
                    for (int s2 = 0; s2 < 8; ++s2) {
                        const bool ok = colp < n && 4 * s2 + lk < pw;
                        pa[s2] = ok ? pa[s2] : 0.0;
                        pa[8 + s2] = ok ? pa[8 + s2] : 0.0;
                    }
                    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int s2 = 0; s2 < 16; ++s2) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[s2], qb[s2], acc, 0, 0, 0);
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int col = c0 + 4 * v + lk;
                        if (row < n && col <= row) bstore(ra, (uint32_t)(col * lda + row) * 8u, 0u, av[v] - acc[v]);
                    }
                };
                double pa0[16], av0[4], pa1[16], av1[4];
                load(0, pa0, av0);
                for (int cB = 0; cB <= sR; cB += 2) {
                    if (cB + 1 <= sR) load(cB + 1, pa1, av1);
                    tile(cB, pa0, av0);
                    if (cB + 1 > sR) break;
                    if (cB + 2 <= sR) load(cB + 2, pa0, av0);
                    tile(cB + 1, pa1, av1);
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
    (void)hipFuncSetAttribute((const void *)k_sytrd_lower, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

void launch_sytrd_lower(double *A, int64_t sA, int lda, double *D, int64_t sD, double *E, int64_t sE, double *tau,
                        int64_t sT, double *W, int n, int n_mat, hipStream_t s) {
    if (n_mat <= 0 || n <= 0) return;
    hipLaunchKernelGGL(k_sytrd_lower, dim3(n_mat), dim3(kSyThreads), sytrd_wg_lds(n), s, A, sA, lda, D, sD, E, sE,
                       tau, sT, W, n);
}

size_t sytrd_wg_scratch(int n) { return (size_t)kSyNB * n; }

}  // namespace crest
