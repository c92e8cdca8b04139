// Highly variable genes, flavour "seurat", from the per-gene statistics of the chunk
// kernels (crest/pp.py:201-292 with _mean_var pp.py:123-126 and _pd_cut_codes
// pp.py:129-136: scanpy's algorithm).  One workgroup does the whole selection on the
// device, so gene statistics + HVG are one device pass with no host round trip.
//
//   mean_log, var_log = moments of x;  mean, var = moments of expm1(x)  (ddof 1, >= 0)
//   mean == 0 -> 1e-12;  disp = var / mean;  disp == 0 -> NaN;  disp = log(disp)
//   mean = log1p(mean);  bins: 20 equal-width bins over [min, max] of mean, the lowest
//   edge lowered by 0.1% of the range, right-closed (pandas.cut); per bin the mean
//   and ddof-1 standard deviation of the non-NaN dispersions, a bin with one gene
//   normalised to 1; disp_norm = (disp - bin mean) / bin sd;  selection: the genes
//   whose disp_norm is >= the n_top-th largest (ties all kept), or the mean /
//   dispersion window when n_top is not given.
//
// The bin sums follow numpy's summation order (0 + pairwise sum: eight accumulators
// over blocks of up to 128 values, blocks split in halves at multiples of eight), so
// the bin means and deviations are numpy's whenever the dispersions are.  log and
// log1p are the device's FP64 functions (within an ulp of numpy's); the index set is
// checked against the oracle's.  The n_top-th largest value is found exactly by a
// radix select over order-preserving 64-bit keys.
#include "common.hpp"

namespace crest {

constexpr int kHvgThreads = 1024;
constexpr int kHvgWaves = kHvgThreads / kWave;
constexpr int kHvgMaxBins = 256;

// numpy's pairwise sum of f(i), i in [0, n), with res starting from 0.
template <class F> __device__ double np_pairwise(F f, int64_t lo, int64_t n) {
    if (n < 8) {
        double r = 0.0;
        for (int64_t i = 0; i < n; ++i) r = r + f(lo + i);
        return r;
    }
    if (n <= 128) {
        double r[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) r[j] = f(lo + j);
        int64_t i = 8;
        for (; i < n - (n % 8); i += 8) {
#pragma unroll
            for (int j = 0; j < 8; ++j) r[j] = r[j] + f(lo + i + j);
        }
        double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) res = res + f(lo + i);
        return res;
    }
    int64_t n2 = n / 2;
    n2 -= n2 % 8;
    return np_pairwise(f, lo, n2) + np_pairwise(f, lo + n2, n - n2);
}
template <class F> __device__ double np_sum(F f, int64_t n) { return 0.0 + np_pairwise(f, 0, n); }

// order-preserving key of a double (ascending keys <=> ascending values; no NaN)
__device__ inline uint64_t f64_key(double v) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ inline double key_f64(uint64_t k) {
    const uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)u);
}

struct HvgParams {
    int64_t n_vars;
    double n;          // cells
    int n_bins;
    int64_t n_top;     // < 0: the mean / dispersion window instead
    double min_mean, max_mean, min_disp, max_disp;
};

// _mean_var (pp.py:123-126): mean = s1/n; var = max((s2 - n mean^2)/(n - 1), 0)
__device__ inline void mean_var(double s1, double s2, double n, double &m, double &v) {
    m = s1 / n;
    v = n > 1.0 ? (s2 - n * m * m) / (n - 1.0) : 0.0;
    v = (v != v) ? v : (v > 0.0 ? v : 0.0);  // np.maximum(var, 0): NaN stays NaN
}

__device__ inline double block_min_max(double v, bool want_max, double *red) {
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, kWave);
        v = want_max ? (w > v ? w : v) : (w < v ? w : v);
    }
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int w = 1; w < kHvgWaves; ++w) r = want_max ? (red[w] > r ? red[w] : r) : (red[w] < r ? red[w] : r);
        red[kHvgWaves] = r;
    }
    __syncthreads();
    const double r = red[kHvgWaves];
    __syncthreads();
    return r;
}

// stats [n_vars, 5]: Σx, Σx², Σexpm1(x), Σexpm1(x)², nnz.  work: codes int32 [n_vars],
// compact double [n_vars].  Outputs [n_vars]: means (log1p of the expm1 mean),
// dispersions, dispersions_norm, hv (0/1), mean_log, std_log.
__global__ __launch_bounds__(kHvgThreads) void k_hvg_seurat(const double *stats, HvgParams P, int32_t *codes,
                                                            double *compact, double *means, double *disp,
                                                            double *disp_norm, uint8_t *hv, double *mean_log,
                                                            double *std_log) {
    __shared__ double red[kHvgWaves + 1];
    __shared__ double edges[kHvgMaxBins + 1];
    __shared__ double bavg[kHvgMaxBins], bdev[kHvgMaxBins];
    __shared__ int bcount[kHvgMaxBins], boff[kHvgMaxBins], bbase[kHvgMaxBins];
    __shared__ int wcnt[kHvgWaves][kHvgMaxBins];
    __shared__ unsigned int hist[256];
    __shared__ uint64_t sel_prefix;
    __shared__ int64_t sel_need;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const int64_t G = P.n_vars;
    const int nb = P.n_bins;

    // per-gene moments, dispersion, log1p mean
    double lo = __longlong_as_double(0x7ff0000000000000ll), hi = -lo;
    for (int64_t g = tid; g < G; g += kHvgThreads) {
        const double *s = stats + g * 5;
        double ml, vl, m, v;
        mean_var(s[0], s[1], P.n, ml, vl);
        mean_var(s[2], s[3], P.n, m, v);
        if (m == 0.0) m = 1e-12;
        double dp = v / m;
        if (dp == 0.0) dp = __longlong_as_double(0x7ff8000000000000ll);
        dp = log(dp);
        const double mm = log1p(m);
        double sl = sqrt(vl);
        if (sl == 0.0) sl = 1.0;
        means[g] = mm;
        disp[g] = dp;
        mean_log[g] = ml;
        std_log[g] = sl;
        lo = mm < lo ? mm : lo;
        hi = mm > hi ? mm : hi;
    }
    lo = block_min_max(lo, false, red);
    hi = block_min_max(hi, true, red);
    // numpy.linspace(lo, hi, nb + 1): edge i = i * ((hi - lo) / nb) + lo, the last = hi;
    // then the lowest edge moves down by 0.1% of the range
    if (tid == 0) {
        const double step = (hi - lo) / (double)nb;
        for (int i = 0; i < nb; ++i) edges[i] = (double)i * step + lo;
        edges[nb] = hi;
        edges[0] = edges[0] - (hi - lo) * 0.001;
    }
    for (int b = tid; b < nb; b += kHvgThreads) {
        bcount[b] = 0;
        bbase[b] = 0;
    }
    __syncthreads();
    // bin codes: (number of edges < mean) - 1, clipped to [0, nb - 1]; a single value
    // for every gene when the range is empty
    for (int64_t g = tid; g < G; g += kHvgThreads) {
        int c = 0;
        if (lo != hi) {
            const double x = means[g];
            int below = 0;
            for (int i = 0; i <= nb; ++i) below += edges[i] < x ? 1 : 0;
            c = below - 1;
            c = c < 0 ? 0 : (c > nb - 1 ? nb - 1 : c);
        }
        codes[g] = c;
        if (disp[g] == disp[g]) atomicAdd(&bcount[c], 1);
    }
    __syncthreads();
    if (tid == 0) {
        int o = 0;
        for (int b = 0; b < nb; ++b) {
            boff[b] = o;
            o += bcount[b];
        }
    }
    __syncthreads();
    // each bin's non-NaN dispersions, in gene order, into compact[boff[b] ...]
    for (int64_t t0 = 0; t0 < G; t0 += kHvgThreads) {
        const int64_t g = t0 + tid;
        const bool valid = g < G;
        const int b = valid ? codes[g] : -1;
        const double v = valid ? disp[g] : 0.0;
        const bool keep = valid && v == v;
        int rank = 0;
        for (int bb = 0; bb < nb; ++bb) {
            const uint64_t m = __ballot(keep && b == bb);
            if (lane == 0) wcnt[wave][bb] = __popcll(m);
            if (keep && b == bb) rank = __popcll(m & ((1ull << lane) - 1ull));
        }
        __syncthreads();
        if (keep) {
            int off = bbase[b];
            for (int w = 0; w < wave; ++w) off += wcnt[w][b];
            compact[boff[b] + off + rank] = v;
        }
        __syncthreads();
        for (int bb = tid; bb < nb; bb += kHvgThreads) {
            int s = 0;
            for (int w = 0; w < kHvgWaves; ++w) s += wcnt[w][bb];
            bbase[bb] += s;
        }
        __syncthreads();
    }
    // per bin: mean and ddof-1 deviation in numpy's order; one-gene bins normalise to 1
    for (int b = tid; b < nb; b += kHvgThreads) {
        const int c = bcount[b];
        const double *seg = compact + boff[b];
        const double nan = __longlong_as_double(0x7ff8000000000000ll);
        double avg = nan, dev = nan;
        if (c > 0) avg = np_sum([&](int64_t i) { return seg[i]; }, c) / (double)c;
        if (c > 1) {
            const double a = avg;
            const double ss = np_sum(
                [&](int64_t i) {
                    const double x = seg[i] - a;
                    return x * x;
                },
                c);
            dev = sqrt(ss / (double)(c - 1));
        }
        if (dev != dev) {
            dev = avg;
            avg = 0.0;
        }
        bavg[b] = avg;
        bdev[b] = dev;
    }
    __syncthreads();
    int64_t n_ok = 0;
    for (int64_t g = tid; g < G; g += kHvgThreads) {
        const int c = codes[g];
        const double dn = (disp[g] - bavg[c]) / bdev[c];
        disp_norm[g] = dn;
        n_ok += dn == dn ? 1 : 0;
    }
    // count of non-NaN normalised dispersions
    double cnt = (double)n_ok;
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, kWave);
    if (lane == 0) red[wave] = cnt;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < kHvgWaves; ++w) t += red[w];
        red[kHvgWaves] = t;
    }
    __syncthreads();
    const int64_t n_valid = (int64_t)red[kHvgWaves];
    __syncthreads();

    if (P.n_top < 0) {
        // the mean / dispersion window on nan_to_num(disp_norm)
        for (int64_t g = tid; g < G; g += kHvgThreads) {
            double dn = disp_norm[g];
            if (dn != dn) dn = 0.0;
            if (dn == __longlong_as_double(0x7ff0000000000000ll)) dn = 1.7976931348623157e308;
            if (dn == -__longlong_as_double(0x7ff0000000000000ll)) dn = -1.7976931348623157e308;
            const double mm = means[g];
            hv[g] = (mm > P.min_mean && mm < P.max_mean && dn > P.min_disp && dn < P.max_disp) ? 1 : 0;
        }
        return;
    }
    int64_t k = P.n_top < n_valid ? P.n_top : n_valid;
    k = k < G ? k : G;
    if (k <= 0) {
        for (int64_t g = tid; g < G; g += kHvgThreads) hv[g] = 0;
        return;
    }
    // radix select (MSB first, 8 bits per pass) of the k-th largest key
    if (tid == 0) {
        sel_prefix = 0;
        sel_need = k;
    }
    for (int pass = 0; pass < 8; ++pass) {
        const int shift = 56 - 8 * pass;
        for (int i = tid; i < 256; i += kHvgThreads) hist[i] = 0;
        __syncthreads();
        const uint64_t pre = sel_prefix;
        const uint64_t mask_hi = pass == 0 ? 0ull : (~0ull << (shift + 8));
        for (int64_t g = tid; g < G; g += kHvgThreads) {
            const double dn = disp_norm[g];
            if (dn != dn) continue;
            const uint64_t key = f64_key(dn);
            if ((key & mask_hi) != pre) continue;
            atomicAdd(&hist[(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            int64_t need = sel_need;
            int d = 255;
            for (; d > 0; --d) {
                if ((int64_t)hist[d] >= need) break;
                need -= hist[d];
            }
            sel_prefix = pre | ((uint64_t)d << shift);
            sel_need = need;
        }
        __syncthreads();
    }
    const double cutoff = key_f64(sel_prefix);
    for (int64_t g = tid; g < G; g += kHvgThreads) {
        double dn = disp_norm[g];
        if (dn != dn) dn = -__longlong_as_double(0x7ff0000000000000ll);  // nan_to_num(nan = -inf)
        hv[g] = dn >= cutoff ? 1 : 0;
    }
}

bool hvg_bins_supported(int n_bins) { return n_bins >= 1 && n_bins <= kHvgMaxBins; }

void launch_hvg_seurat(const double *stats, int64_t n_vars, double n_cells, int n_bins, int64_t n_top,
                       double min_mean, double max_mean, double min_disp, double max_disp, int32_t *codes,
                       double *compact, double *means, double *disp, double *disp_norm, uint8_t *hv,
                       double *mean_log, double *std_log, hipStream_t s) {
    HvgParams P{n_vars, n_cells, n_bins, n_top, min_mean, max_mean, min_disp, max_disp};
    hipLaunchKernelGGL(k_hvg_seurat, dim3(1), dim3(kHvgThreads), 0, s, stats, P, codes, compact, means, disp,
                       disp_norm, hv, mean_log, std_log);
}

}  // namespace crest
