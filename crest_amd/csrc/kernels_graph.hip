// UMAP fuzzy simplicial set of the kNN graph (umap/graph.rs:29-107; py.rs:72-99): per
// cell, (rho, sigma) by the 64-step bisection of smooth_knn_dist and the directed
// membership strengths keyed by the unordered pair; a radix sort of the keys; the
// fuzzy union w = a + b - ab of the (at most two) entries of a pair; and the edges
// with source < target in key order.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "common.hpp"

namespace crest {

// mean over all distances (graph.rs:66): per-block FP64 partial sums, then one block.
__global__ void k_sum_partial(const float *d, int64_t m, double *part) {
    __shared__ double s[256];
    double a = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
        a += (double)d[i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}
__global__ void k_sum_final(const double *part, int np, double *out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double a = 0.0;
        for (int i = 0; i < np; ++i) a += part[i];
        *out = a;
    }
}

// One thread per cell: smooth_knn_dist (local_connectivity = 1) then the directed
// strengths; invalid entries (self, missing neighbour) get key ~0.
__global__ void k_membership(const uint32_t *idx, const float *dist, int64_t n, int k, int n_neighbors,
                             const double *sum_all, uint64_t *keys, float *w) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *d = dist + i * k;
    const uint32_t *ix = idx + i * k;
    const double mean_all = *sum_all / (double)(n * (int64_t)(k + 1));
    const double target = log2((double)n_neighbors);
    double rho = 0.0;
    for (int t = 0; t < k; ++t)
        if (d[t] > 0.0f) {
            rho = (double)d[t];
            break;
        }
    double lo = 0.0, hi = INFINITY, mid = 1.0;
    for (int it = 0; it < 64; ++it) {
        double psum = 0.0;
        for (int t = 0; t < k; ++t) {
            const double v = (double)d[t] - rho;
            psum += v > 0.0 ? exp(-v / mid) : 1.0;
        }
        if (fabs(psum - target) < 1e-5) break;
        if (psum > target) {
            hi = mid;
            mid = (lo + hi) / 2.0;
        } else {
            lo = mid;
            if (hi == INFINITY)
                mid *= 2.0;
            else
                mid = (lo + hi) / 2.0;
        }
    }
    double s = 0.0;
    for (int t = 0; t < k; ++t) s += (double)d[t];
    const double mean_i = s / (double)(k + 1);
    const double fl = rho > 0.0 ? 1e-3 * mean_i : 1e-3 * mean_all;
    const double sigma = mid > fl ? mid : fl;
    for (int t = 0; t < k; ++t) {
        const uint32_t j = ix[t];
        const int64_t o = i * k + t;
        if (j == 0xFFFFFFFFu || (int64_t)j == i) {
            keys[o] = ~0ull;
            w[o] = 0.0f;
            continue;
        }
        const double v = (double)d[t] - rho;
        const double ww = (v <= 0.0 || sigma == 0.0) ? 1.0 : exp(-v / sigma);
        const uint64_t a = (uint32_t)i < j ? (uint64_t)i : (uint64_t)j;
        const uint64_t b = (uint32_t)i < j ? (uint64_t)j : (uint64_t)i;
        keys[o] = (a << 32) | b;
        w[o] = (float)ww;
    }
}

// Fuzzy union at the head of each run of equal keys; flag[p] = 1 for kept edges.
__global__ void k_union(const uint64_t *keys, const float *w, int64_t m, float *wu, uint32_t *flag) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= m) return;
    const uint64_t key = keys[p];
    const bool head = key != ~0ull && (p == 0 || keys[p - 1] != key);
    float u = 0.0f;
    if (head) {
        u = w[p];
        for (int64_t q = p + 1; q < m && keys[q] == key; ++q) {
            const float b = w[q];
            const float s = u + b;
            const float pr = u * b;
            u = s - pr;
        }
    }
    wu[p] = u;
    flag[p] = head && u > 0.0f ? 1u : 0u;
}

__global__ void k_emit(const uint64_t *keys, const float *wu, const uint32_t *flag, const uint32_t *pos, int64_t m,
                       uint32_t *rows, uint32_t *cols, float *wout) {
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= m || !flag[p]) return;
    const uint32_t o = pos[p];
    rows[o] = (uint32_t)(keys[p] >> 32);
    cols[o] = (uint32_t)(keys[p] & 0xFFFFFFFFull);
    wout[o] = wu[p];
}

__global__ void k_count(const uint32_t *flag, const uint32_t *pos, int64_t m, uint32_t *count) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *count = m > 0 ? pos[m - 1] + flag[m - 1] : 0u;
}

size_t conn_temp_bytes(int64_t m) {
    size_t a = 0, b = 0;
    (void)rocprim::radix_sort_pairs((void *)nullptr, a, (uint64_t *)nullptr, (uint64_t *)nullptr, (float *)nullptr,
                              (float *)nullptr, (size_t)m, 0, 64);
    (void)rocprim::exclusive_scan((void *)nullptr, b, (uint32_t *)nullptr, (uint32_t *)nullptr, 0u, (size_t)m,
                            rocprim::plus<uint32_t>());
    return a > b ? a : b;
}

// Buffers: keys/keys2 u64[m], w/w2/wu f32[m], flag/pos u32[m], part f64[256] + 1,
// temp of conn_temp_bytes(m).  Returns the edge count through *count (device u32).
void launch_connectivities(const uint32_t *idx, const float *dist, int64_t n, int k, int n_neighbors,
                           uint64_t *keys, uint64_t *keys2, float *w, float *w2, float *wu, uint32_t *flag,
                           uint32_t *pos, double *part, void *temp, size_t temp_bytes, uint32_t *rows,
                           uint32_t *cols, float *wout, uint32_t *count, hipStream_t s) {
    const int64_t m = n * k;
    hipLaunchKernelGGL(k_sum_partial, dim3(256), dim3(256), 0, s, dist, m, part);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(64), 0, s, part, 256, part + 256);
    hipLaunchKernelGGL(k_membership, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, idx, dist, n, k, n_neighbors,
                       part + 256, keys, w);
    size_t tb = temp_bytes;
    (void)rocprim::radix_sort_pairs(temp, tb, keys, keys2, w, w2, (size_t)m, 0, 64, s);
    const unsigned g = (unsigned)ceil_div(m, 256);
    hipLaunchKernelGGL(k_union, dim3(g), dim3(256), 0, s, keys2, w2, m, wu, flag);
    tb = temp_bytes;
    (void)rocprim::exclusive_scan(temp, tb, flag, pos, 0u, (size_t)m, rocprim::plus<uint32_t>(), s);
    hipLaunchKernelGGL(k_emit, dim3(g), dim3(256), 0, s, keys2, wu, flag, pos, m, rows, cols, wout);
    // count = pos[m-1] + flag[m-1]
    hipLaunchKernelGGL(k_count, dim3(1), dim3(64), 0, s, flag, pos, m, count);
}

}  // namespace crest
