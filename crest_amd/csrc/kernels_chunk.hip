// Streaming chunk kernels: per-gene statistics, CSR -> dense scaled tile (with the
// column sums of the scaled matrix), the FP64 sparse projection and materialize.
//
// One wavefront owns one cell (span of CSR entries) at a time: the 64 lanes
// stride over the cell's entries, so reads of `genes`/`values` are coalesced, and
// the per-cell total used by normalize_total is a wave reduction.
#include "common.hpp"

namespace crest {

// ---------------------------------------------------------------- fills
__global__ void k_fill_f64(double *p, int64_t n, double v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f32(float *p, int64_t n, float v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
void launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t s) {
    if (n <= 0) return;
    int g = (int)std::min<int64_t>(ceil_div(n, 256), 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(256), 0, s, p, n, v);
}
void launch_fill_f64(double *p, int64_t n, double v, hipStream_t s) {
    if (n <= 0) return;
    int g = (int)std::min<int64_t>(ceil_div(n, 256), 4096);
    hipLaunchKernelGGL(k_fill_f64, dim3(g), dim3(256), 0, s, p, n, v);
}
void launch_fill_f32(float *p, int64_t n, float v, hipStream_t s) {
    if (n <= 0) return;
    int g = (int)std::min<int64_t>(ceil_div(n, 256), 4096);
    hipLaunchKernelGGL(k_fill_f32, dim3(g), dim3(256), 0, s, p, n, v);
}

// Sum over the kept genes of one cell (kernels.rs:69-76), identical in every lane.
__device__ inline double cell_total(const DevChunk &c, int64_t lo, int64_t hi, int lane) {
    double t = 0.0;
    for (int64_t e = lo + lane; e < hi; e += kWave)
        if (chunk_var(c, c.genes[e]) >= 0) t += (double)c.values[e];
    return wave_sum_f64(t);
}

// ---------------------------------------------------------------- gene_stats
// kernels.rs:141-160.  Each workgroup owns a contiguous block of cells and
// accumulates a window of vars in LDS (FP64), then adds it to `out` once.
// With `ranges`, workgroup b takes row range ranges[b] of segment segs[ranges[b].seg]
// and accumulates into that segment's subset's block of `out` ([subset][n_vars][5]).
__global__ __launch_bounds__(1024) void k_gene_stats(DevChunk c, double *out, int n_vars, const SegDesc *segs,
                                                    const RowRange *ranges) {
    __shared__ double acc[5 * kStatsWindow];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int64_t base = c.base;
    int64_t r0, r1;
    if (ranges) {
        const RowRange w = ranges[blockIdx.x];
        const SegDesc sd = segs[w.seg];
        c = seg_rows(c, sd.cell0 + w.r0, sd.cell0 + w.r1);
        out += (int64_t)sd.subset * n_vars * 5;
        r0 = 0;
        r1 = c.n_spans;
    } else {
        const int64_t per = (c.n_spans + gridDim.x - 1) / gridDim.x;
        r0 = (int64_t)blockIdx.x * per;
        r1 = r0 + per < c.n_spans ? r0 + per : c.n_spans;
    }
    for (int w0 = 0; w0 < n_vars; w0 += kStatsWindow) {
        for (int i = threadIdx.x; i < 5 * kStatsWindow; i += blockDim.x) acc[i] = 0.0;
        __syncthreads();
        for (int64_t r = r0 + wave; r < r1; r += blockDim.x / kWave) {
            if (chunk_row(c, r) < 0) continue;
            const int64_t lo = c.span_off[r] - base, hi = c.span_off[r + 1] - base;
            const double ns = norm_factor(cell_total(c, lo, hi, lane), c.target_sum);
            for (int64_t e = lo + lane; e < hi; e += kWave) {
                const int32_t m = chunk_var(c, c.genes[e]);
                if (m >= w0 && m < w0 + kStatsWindow) {
                    const double x = (double)lazy_transform(c.values[e], ns, c.target_sum, c.log1p);
                    const double ex = expm1(x);
                    double *a = acc + (m - w0);
                    atomicAdd(a, x);
                    atomicAdd(a + kStatsWindow, x * x);
                    atomicAdd(a + 2 * kStatsWindow, ex);
                    atomicAdd(a + 3 * kStatsWindow, ex * ex);
                    atomicAdd(a + 4 * kStatsWindow, 1.0);
                }
            }
        }
        __syncthreads();
        const int wn = n_vars - w0 < kStatsWindow ? n_vars - w0 : kStatsWindow;
        for (int i = threadIdx.x; i < wn; i += blockDim.x)
            for (int s = 0; s < 5; ++s) {
                const double v = acc[s * kStatsWindow + i];
                if (v != 0.0) atomicAdd(&out[(int64_t)(w0 + i) * 5 + s], v);
            }
        __syncthreads();
    }
}

// One chunk: 1,024-thread workgroups (16 waves, one cell each at a time) so that the
// two workgroups the 80 KB accumulator leaves room for per CU keep 32 waves' loads in
// flight; at most two workgroups per CU are launched.
void launch_gene_stats(const DevChunk &c, double *out, int n_vars, hipStream_t s) {
    if (c.n_spans <= 0 || n_vars <= 0) return;
    int g = (int)std::min<int64_t>(ceil_div(c.n_spans, 64), 512);
    hipLaunchKernelGGL(k_gene_stats, dim3(g), dim3(1024), 0, s, c, out, n_vars, (const SegDesc *)nullptr,
                       (const RowRange *)nullptr);
}
void launch_gene_stats_ranges(const DevChunk &c, const SegDesc *segs, const RowRange *ranges, int64_t n_ranges,
                              double *out, int n_vars, hipStream_t s) {
    if (n_ranges <= 0 || n_vars <= 0) return;
    hipLaunchKernelGGL(k_gene_stats, dim3((unsigned)n_ranges), dim3(1024), 0, s, c, out, n_vars, segs, ranges);
}

// ---------------------------------------------------------------- qc
// kernels.rs:162-195 with one wave per cell: for every kept entry whose (transformed)
// value is non-zero, the row gets its total, its count of such genes and its total
// over flagged genes, and the gene its cell count and total.  FP64 lane partial sums
// and atomics: exact for integer counts, whose FP64 sums below 2^53 do not depend on
// the order of addition (the reference's own order depends on its thread count).
__global__ __launch_bounds__(256) void k_qc(DevChunk c, const uint8_t *flag, double *row_total, double *row_ngenes,
                                            double *row_flag, uint32_t *gene_ncells, double *gene_total) {
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int nwave = blockDim.x / kWave;
    const int64_t base = c.base;
    for (int64_t r = (int64_t)blockIdx.x * nwave + wave; r < c.n_spans; r += (int64_t)gridDim.x * nwave) {
        const int64_t row = chunk_row(c, r);
        if (row < 0) continue;
        const int64_t lo = c.span_off[r] - base, hi = c.span_off[r + 1] - base;
        const double ns = norm_factor(cell_total(c, lo, hi, lane), c.target_sum);
        double t = 0.0, f = 0.0, k = 0.0;
        for (int64_t e = lo + lane; e < hi; e += kWave) {
            const int32_t m = chunk_var(c, c.genes[e]);
            if (m < 0) continue;
            const float x = lazy_transform(c.values[e], ns, c.target_sum, c.log1p);
            if (x == 0.0f) continue;
            t += (double)x;
            k += 1.0;
            if (flag[m]) f += (double)x;
            atomicAdd(&gene_ncells[m], 1u);
            atomicAdd(&gene_total[m], (double)x);
        }
        t = wave_sum_f64(t);
        f = wave_sum_f64(f);
        k = wave_sum_f64(k);
        if (lane == 0) {
            atomicAdd(&row_total[row], t);
            atomicAdd(&row_ngenes[row], k);
            atomicAdd(&row_flag[row], f);
        }
    }
}

void launch_qc(const DevChunk &c, const uint8_t *flag, double *row_total, double *row_ngenes, double *row_flag,
               uint32_t *gene_ncells, double *gene_total, hipStream_t s) {
    if (c.n_spans <= 0) return;
    const int g = (int)std::min<int64_t>(ceil_div(c.n_spans, 4), 4096);
    hipLaunchKernelGGL(k_qc, dim3(g), dim3(256), 0, s, c, flag, row_total, row_ngenes, row_flag, gene_ncells,
                       gene_total);
}

// ---------------------------------------------------------------- densify
// Dense scaled tile S of a chunk (the sparse S of Z = 1 b^T + S, kernels.rs:197-219),
// written in the blocked layout tile[(group*d_pad + h)*8 + r] (row = 8*group + r), so
// that one 16-byte load gives an MFMA lane its K-run of consecutive rows.  Row i of
// the tile is span i of the chunk; filtered cells and padding rows are zero.
// colsum[h] += s over every non-zero (kernels.rs:258-260).
__device__ inline uint16_t bf16_bits_rn(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf16_to_float(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

template <int FMT> struct TileElem;
template <> struct TileElem<TILE_F64> { typedef double T; static constexpr int DW = 1024; };
// Column window per pass: small enough that three workgroups fit a CU's LDS, so
// enough tile stores are in flight (the tile is written at close to HBM speed).
template <> struct TileElem<TILE_F32> { typedef float T; static constexpr int DW = 1024; };
template <> struct TileElem<TILE_BF16X2> { typedef uint16_t T; static constexpr int DW = 1024; };
template <> struct TileElem<TILE_F16X2> { typedef uint16_t T; static constexpr int DW = 1024; };

// With `ranges`, workgroup b densifies row groups [r0/8, r1/8) of segment
// segs[ranges[b].seg] into tile groups grp0 + g, with that subset's scale parameters
// (mean/sd + subset d), column sums (colsum + subset d_pad) and Gram diagonal
// (gram + subset d_pad^2).
template <int FMT>
__global__ __launch_bounds__(256) void k_densify(DevChunk c, DevScale sc, void *tile_a, void *tile_b,
                                                 int64_t n_groups, int d_pad, double *colsum, double *gram,
                                                 const SegDesc *segs, const RowRange *ranges) {
    typedef typename TileElem<FMT>::T T;
    constexpr int DW = TileElem<FMT>::DW;
    constexpr bool X2 = FMT == TILE_BF16X2 || FMT == TILE_F16X2;  // hi/lo pairs
    constexpr int NST = X2 ? 2 : 1;
    constexpr bool LL = X2;  // also Σ l² per column, the Gram diagonal's L^T L term
    extern __shared__ __align__(16) unsigned char lds_raw[];
    double *cs = reinterpret_cast<double *>(lds_raw);              // [DW]
    double *cl = cs + DW;                                          // [DW] when LL
    T *stage = reinterpret_cast<T *>(lds_raw + (LL ? 2 : 1) * DW * sizeof(double)); // [NST][DW][8]
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int nwave = blockDim.x / kWave;
    const int64_t base = c.base;
    int64_t g_begin = blockIdx.x, g_end = n_groups, g_step = gridDim.x, grp0 = 0;
    if (ranges) {
        const RowRange w = ranges[blockIdx.x];
        const SegDesc sd = segs[w.seg];
        c = seg_rows(c, sd.cell0, sd.cell0 + sd.n);
        g_begin = w.r0 / kRowGroup;
        g_end = w.r1 / kRowGroup;
        g_step = 1;
        grp0 = sd.grp0;
        colsum += (int64_t)sd.subset * d_pad;
        gram += (int64_t)sd.subset * d_pad * d_pad;
        sc.mean += (int64_t)sd.subset * sc.d;
        sc.sd += (int64_t)sd.subset * sc.d;
    }
    for (int h0 = 0; h0 < d_pad; h0 += DW) {
        const int dw = d_pad - h0 < DW ? d_pad - h0 : DW;
        for (int i = threadIdx.x; i < DW; i += blockDim.x) {
            cs[i] = 0.0;
            if constexpr (LL) cl[i] = 0.0;
        }
        for (int64_t g = g_begin; g < g_end; g += g_step) {
            // zero the stage (vectorised)
            uint4 *st4 = reinterpret_cast<uint4 *>(stage);
            const int n4 = (int)((int64_t)NST * DW * 8 * sizeof(T) / 16);
            for (int i = threadIdx.x; i < n4; i += blockDim.x) st4[i] = make_uint4(0, 0, 0, 0);
            __syncthreads();
            for (int r = wave; r < kRowGroup; r += nwave) {
                const int64_t span = g * kRowGroup + r;
                if (span >= c.n_spans) break;
                if (chunk_row(c, span) < 0) continue;
                const int64_t lo = c.span_off[span] - base, hi = c.span_off[span + 1] - base;
                const double ns = norm_factor(cell_total(c, lo, hi, lane), c.target_sum);
                for (int64_t e = lo + lane; e < hi; e += kWave) {
                    const int32_t m = chunk_var(c, c.genes[e]);
                    if (m < 0) continue;
                    const int32_t h = sc.sub[m];
                    if (h < h0 || h >= h0 + dw) continue;
                    const float v = lazy_transform(c.values[e], ns, c.target_sum, c.log1p);
                    const double s = sc.scale ? scale_value(v, sc.mean[h], sc.sd[h], sc.max_value, 1) : (double)v;
                    const int off = (h - h0) * kRowGroup + r;
                    if constexpr (FMT == TILE_F64) {
                        stage[off] = s;
                    } else if constexpr (FMT == TILE_F32) {
                        stage[off] = (float)s;
                    } else if constexpr (FMT == TILE_BF16X2) {
                        const uint16_t hb = bf16_bits_rn((float)s);
                        const uint16_t lb = bf16_bits_rn((float)(s - (double)bf16_to_float(hb)));
                        stage[off] = hb;
                        stage[DW * kRowGroup + off] = lb;
                        const double lf = (double)bf16_to_float(lb);
                        atomicAdd(&cl[h - h0], lf * lf);
                    } else {  // split-FP16: the residual from the FP64 value, both halves RNE
                        const _Float16 hh = (_Float16)(float)s;
                        const _Float16 ll = (_Float16)(float)(s - (double)(float)hh);
                        stage[off] = __builtin_bit_cast(uint16_t, hh);
                        stage[DW * kRowGroup + off] = __builtin_bit_cast(uint16_t, ll);
                        const double lf = (double)(float)ll;
                        atomicAdd(&cl[h - h0], lf * lf);
                    }
                    atomicAdd(&cs[h - h0], s);
                }
            }
            __syncthreads();
            // stage -> tile (contiguous run of dw*8 elements per stage)
            for (int t = 0; t < NST; ++t) {
                T *dst = reinterpret_cast<T *>(t == 0 ? tile_a : tile_b) + ((grp0 + g) * d_pad + h0) * kRowGroup;
                const uint4 *src = reinterpret_cast<const uint4 *>(stage + (int64_t)t * DW * kRowGroup);
                uint4 *dst4 = reinterpret_cast<uint4 *>(dst);
                const int m4 = (int)((int64_t)dw * kRowGroup * sizeof(T) / 16);
                for (int i = threadIdx.x; i < m4; i += blockDim.x) dst4[i] = src[i];
            }
            __syncthreads();
        }
        for (int i = threadIdx.x; i < dw; i += blockDim.x) {
            if (cs[i] != 0.0) atomicAdd(&colsum[h0 + i], cs[i]);
            if constexpr (LL)
                if (cl[i] != 0.0) atomicAdd(&gram[(int64_t)(h0 + i) * d_pad + h0 + i], cl[i]);
        }
        __syncthreads();
    }
}

void launch_densify(const DevChunk &c, const DevScale &sc, int fmt, void *tile_a, void *tile_b,
                    int64_t rows_pad, int d_pad, double *colsum, double *gram, hipStream_t s) {
    const int64_t n_groups = rows_pad / kRowGroup;
    if (n_groups <= 0) return;
    size_t lds;
    int g;
    switch (fmt) {
    case TILE_F64:
        lds = TileElem<TILE_F64>::DW * (sizeof(double) + 8 * sizeof(double));
        g = (int)std::min<int64_t>(n_groups, 512);
        hipFuncSetAttribute((const void *)k_densify<TILE_F64>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(k_densify<TILE_F64>, dim3(g), dim3(256), lds, s, c, sc, tile_a, tile_b, n_groups, d_pad, colsum,
                           gram, (const SegDesc *)nullptr, (const RowRange *)nullptr);
        break;
    case TILE_F32:
        lds = TileElem<TILE_F32>::DW * (sizeof(double) + 8 * sizeof(float));
        g = (int)std::min<int64_t>(n_groups, 2048);
        hipFuncSetAttribute((const void *)k_densify<TILE_F32>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(k_densify<TILE_F32>, dim3(g), dim3(256), lds, s, c, sc, tile_a, tile_b, n_groups, d_pad, colsum,
                           gram, (const SegDesc *)nullptr, (const RowRange *)nullptr);
        break;
    default:
        lds = TileElem<TILE_BF16X2>::DW * (2 * sizeof(double) + 16 * sizeof(uint16_t));
        g = (int)std::min<int64_t>(n_groups, 2048);
        hipFuncSetAttribute((const void *)k_densify<TILE_BF16X2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (fmt == TILE_F16X2) {
            hipFuncSetAttribute((const void *)k_densify<TILE_F16X2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds);
            hipLaunchKernelGGL(k_densify<TILE_F16X2>, dim3(g), dim3(256), lds, s, c, sc, tile_a, tile_b, n_groups, d_pad,
                               colsum, gram, (const SegDesc *)nullptr, (const RowRange *)nullptr);
        } else {
            hipLaunchKernelGGL(k_densify<TILE_BF16X2>, dim3(g), dim3(256), lds, s, c, sc, tile_a, tile_b, n_groups, d_pad,
                               colsum, gram, (const SegDesc *)nullptr, (const RowRange *)nullptr);
        }
        break;
    }
}

// The dynamic LDS limits of every k_densify form (set once before any graph capture).
void densify_set_attrs() {
    hipFuncSetAttribute((const void *)k_densify<TILE_F64>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)(TileElem<TILE_F64>::DW * (sizeof(double) + 8 * sizeof(double))));
    hipFuncSetAttribute((const void *)k_densify<TILE_F32>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)(TileElem<TILE_F32>::DW * (sizeof(double) + 8 * sizeof(float))));
    const int l16 = (int)(TileElem<TILE_BF16X2>::DW * (2 * sizeof(double) + 16 * sizeof(uint16_t)));
    hipFuncSetAttribute((const void *)k_densify<TILE_BF16X2>, hipFuncAttributeMaxDynamicSharedMemorySize, l16);
    hipFuncSetAttribute((const void *)k_densify<TILE_F16X2>, hipFuncAttributeMaxDynamicSharedMemorySize, l16);
}

void launch_densify_ranges(const DevChunk &c, const DevScale &sc, int fmt, void *tile_a, void *tile_b, int d_pad,
                           const SegDesc *segs, const RowRange *ranges, int64_t n_ranges, double *colsum,
                           double *gram, hipStream_t s) {
    if (n_ranges <= 0) return;
    const dim3 g((unsigned)n_ranges), b(256);
    size_t lds;
    switch (fmt) {
    case TILE_F64:
        lds = TileElem<TILE_F64>::DW * (sizeof(double) + 8 * sizeof(double));
        hipLaunchKernelGGL(k_densify<TILE_F64>, g, b, lds, s, c, sc, tile_a, tile_b, (int64_t)0, d_pad, colsum, gram,
                           segs, ranges);
        break;
    case TILE_F32:
        lds = TileElem<TILE_F32>::DW * (sizeof(double) + 8 * sizeof(float));
        hipLaunchKernelGGL(k_densify<TILE_F32>, g, b, lds, s, c, sc, tile_a, tile_b, (int64_t)0, d_pad, colsum, gram,
                           segs, ranges);
        break;
    case TILE_F16X2:
        lds = TileElem<TILE_F16X2>::DW * (2 * sizeof(double) + 16 * sizeof(uint16_t));
        hipLaunchKernelGGL(k_densify<TILE_F16X2>, g, b, lds, s, c, sc, tile_a, tile_b, (int64_t)0, d_pad, colsum, gram,
                           segs, ranges);
        break;
    default:
        lds = TileElem<TILE_BF16X2>::DW * (2 * sizeof(double) + 16 * sizeof(uint16_t));
        hipLaunchKernelGGL(k_densify<TILE_BF16X2>, g, b, lds, s, c, sc, tile_a, tile_b, (int64_t)0, d_pad, colsum,
                           gram, segs, ranges);
        break;
    }
}

// ---------------------------------------------------------------- dense centred rows
// Row r < m of segment e = blockIdx.y (one subset's cells): the FP64 scaled values of
// the cell minus the subset's column means colsum / n, every HVG column written (rows
// n..m-1 zero).  One wave per row, the row staged in LDS (zeros become -mean).
__global__ __launch_bounds__(256) void k_dense_centred(DevChunk c, DevScale sc, const SegDesc *segs, int m,
                                                       const double *colsum, int ld_c, double *out) {
    extern __shared__ double rowbuf[];  // [4][d]
    const int e = blockIdx.y, lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const SegDesc sd = segs[e];
    const int d = sc.d;
    const int64_t r = (int64_t)blockIdx.x * 4 + wave;
    if (r >= m) return;
    double *row = out + ((int64_t)e * m + r) * d;
    if (r >= sd.n) {
        for (int h = lane; h < d; h += kWave) row[h] = 0.0;
        return;
    }
    double *buf = rowbuf + (int64_t)wave * d;
    const double *cs = colsum + (int64_t)sd.subset * ld_c;
    const double n = (double)sd.n;
    for (int h = lane; h < d; h += kWave) buf[h] = -(cs[h] / n);
    __builtin_amdgcn_wave_barrier();
    const DevChunk cc = seg_rows(c, sd.cell0 + r, sd.cell0 + r + 1);
    if (chunk_row(cc, 0) >= 0) {
        const int64_t lo = cc.span_off[0] - cc.base, hi = cc.span_off[1] - cc.base;
        const double ns = norm_factor(cell_total(cc, lo, hi, lane), cc.target_sum);
        const double *mean = sc.mean + (int64_t)sd.subset * d, *sdv = sc.sd + (int64_t)sd.subset * d;
        for (int64_t q = lo + lane; q < hi; q += kWave) {
            const int32_t v = chunk_var(cc, cc.genes[q]);
            if (v < 0) continue;
            const int32_t h = sc.sub[v];
            if (h < 0) continue;
            const float x = lazy_transform(cc.values[q], ns, cc.target_sum, cc.log1p);
            const double s = sc.scale ? scale_value(x, mean[h], sdv[h], sc.max_value, 1) : (double)x;
            buf[h] = s - cs[h] / n;
        }
    }
    __builtin_amdgcn_wave_barrier();
    for (int h = lane; h < d; h += kWave) row[h] = buf[h];
}

void launch_dense_centred(const DevChunk &c, const DevScale &sc, const SegDesc *segs, int n_seg, int m,
                          const double *colsum, int ld_c, double *out, hipStream_t s) {
    if (n_seg <= 0) return;
    const size_t lds = (size_t)4 * sc.d * sizeof(double);
    hipFuncSetAttribute((const void *)k_dense_centred, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k_dense_centred, dim3((unsigned)ceil_div(m, 4), n_seg), dim3(256), lds, s, c, sc, segs, m,
                       colsum, ld_c, out);
}

// ---------------------------------------------------------------- project
// kernels.rs:350-371: per kept cell, acc[j] = Σ_e s_e * V[h_e, j] in FP64 over the
// cell's HVG entries in chunk order (multiply and add rounded separately, as the
// reference does), then out[row, j] += (float)acc[j] in FP32.  Lane j owns column
// j (and j+64, ...); the (h, s) of 64 entries at a time are computed one per lane
// and broadcast with shuffles.
// With `ranges`, workgroup b projects row range ranges[b] of its segment with that
// subset's loadings (V + subset d k) and scale parameters.
// entries whose V rows are in flight together (1M x 2k: 4: 7.15 ms, 8: 6.85, 16: 6.88)
constexpr int kProjFlight = 8;
template <int NC>
__global__ __launch_bounds__(256) void k_project(DevChunk c, DevScale sc, const double *V, int k, float *out,
                                                  int64_t row0, const SegDesc *segs, const RowRange *ranges) {
    const int lane = threadIdx.x & (kWave - 1);
    int64_t wave_id = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    int64_t n_waves = (int64_t)gridDim.x * blockDim.x / kWave;
    if (ranges) {
        const RowRange w = ranges[blockIdx.x];
        const SegDesc sd = segs[w.seg];
        c = seg_rows(c, sd.cell0 + w.r0, sd.cell0 + w.r1);
        V += (int64_t)sd.subset * sc.d * k;
        sc.mean += (int64_t)sd.subset * sc.d;
        sc.sd += (int64_t)sd.subset * sc.d;
        wave_id = threadIdx.x / kWave;
        n_waves = blockDim.x / kWave;
    }
    const int64_t base = c.base;
    for (int64_t span = wave_id; span < c.n_spans; span += n_waves) {
        const int64_t row = chunk_row(c, span);
        if (row < 0) continue;
        const int64_t lo = c.span_off[span] - base, hi = c.span_off[span + 1] - base;
        const double ns = norm_factor(cell_total(c, lo, hi, lane), c.target_sum);
        double acc[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) acc[q] = 0.0;
        for (int64_t e0 = lo; e0 < hi; e0 += kWave) {
            const int64_t e = e0 + lane;
            int32_t h = -1;
            double s = 0.0;
            if (e < hi) {
                const int32_t m = chunk_var(c, c.genes[e]);
                if (m >= 0) {
                    h = sc.sub[m];
                    if (h >= 0) {
                        const float v = lazy_transform(c.values[e], ns, c.target_sum, c.log1p);
                        s = sc.scale ? scale_value(v, sc.mean[h], sc.sd[h], sc.max_value, 1) : (double)v;
                    }
                }
            }
            // kProjFlight entries at a time: their V rows are loaded before any of
            // their multiply-adds.  An entry's HVG index and scaled value are read from its
            // lane into scalar registers (wave-uniform), so the V row address is
            // scalar and the lanes only load and multiply-add; lanes past k load
            // column k - 1 and are never stored.  An entry outside the Gram is given
            // the value 0, so it adds a signed zero, which leaves the sum unchanged
            // (the sum starts at +0.0 and, rounding to nearest, never becomes -0.0),
            // so the result equals skipping it.
            // (cnt is wave-uniform: the span is the wave's)
            const int cnt = __builtin_amdgcn_readfirstlane(hi - e0 < kWave ? (int)(hi - e0) : kWave);
            const int s_lo = (int)(__double_as_longlong(s) & 0xffffffffll), s_hi = (int)(__double_as_longlong(s) >> 32);
            for (int j0 = 0; j0 < cnt; j0 += kProjFlight) {
                double sj[kProjFlight], vj[kProjFlight][NC];
#pragma unroll
                for (int u = 0; u < kProjFlight; ++u) {
                    const int j = j0 + u < cnt ? j0 + u : cnt - 1;
                    const int32_t hr = __builtin_amdgcn_readlane(h, j);
                    const bool use = j0 + u < cnt && hr >= 0;
                    const double sr = __longlong_as_double(((long long)__builtin_amdgcn_readlane(s_hi, j) << 32) |
                                                           (unsigned int)__builtin_amdgcn_readlane(s_lo, j));
                    sj[u] = use ? sr : 0.0;
                    const double *vr = V + (int64_t)(use ? hr : 0) * k;
#pragma unroll
                    for (int q = 0; q < NC; ++q) {
                        const int col = lane + q * kWave;
                        vj[u][q] = vr[col < k ? col : k - 1];
                    }
                }
#pragma unroll
                for (int u = 0; u < kProjFlight; ++u)
#pragma unroll
                    for (int q = 0; q < NC; ++q) {
                        const double prod = sj[u] * vj[u][q];
                        acc[q] = acc[q] + prod;
                    }
            }
        }
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const int col = lane + q * kWave;
            if (col < k) {
                float *o = out + (row - row0) * (int64_t)k + col;
                *o = *o + (float)acc[q];
            }
        }
    }
}

void launch_project(const DevChunk &c, const DevScale &sc, const double *V, int k, float *out, int64_t row0,
                    hipStream_t s) {
    if (c.n_spans <= 0 || k <= 0) return;
    int g = (int)std::min<int64_t>(ceil_div(c.n_spans, 4), 8192);
    const SegDesc *ns = nullptr;
    const RowRange *nr = nullptr;
    if (k <= 64)
        hipLaunchKernelGGL(k_project<1>, dim3(g), dim3(256), 0, s, c, sc, V, k, out, row0, ns, nr);
    else if (k <= 128)
        hipLaunchKernelGGL(k_project<2>, dim3(g), dim3(256), 0, s, c, sc, V, k, out, row0, ns, nr);
    else
        hipLaunchKernelGGL(k_project<4>, dim3(g), dim3(256), 0, s, c, sc, V, k, out, row0, ns, nr);
}
void launch_project_ranges(const DevChunk &c, const DevScale &sc, const double *V, int k, const SegDesc *segs,
                           const RowRange *ranges, int64_t n_ranges, float *out, hipStream_t s) {
    if (n_ranges <= 0 || k <= 0) return;
    const dim3 g((unsigned)n_ranges), b(256);
    if (k <= 64)
        hipLaunchKernelGGL(k_project<1>, g, b, 0, s, c, sc, V, k, out, (int64_t)0, segs, ranges);
    else if (k <= 128)
        hipLaunchKernelGGL(k_project<2>, g, b, 0, s, c, sc, V, k, out, (int64_t)0, segs, ranges);
    else
        hipLaunchKernelGGL(k_project<4>, g, b, 0, s, c, sc, V, k, out, (int64_t)0, segs, ranges);
}

// ---------------------------------------------------------------- materialize
// kernels.rs:122-139: kept entries of kept cells, transformed, in chunk order.
__global__ __launch_bounds__(256) void k_mat_count(DevChunk c, int64_t *kept) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wave_id = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t n_waves = (int64_t)gridDim.x * blockDim.x / kWave;
    const int64_t base = c.base;
    for (int64_t span = wave_id; span < c.n_spans; span += n_waves) {
        int64_t cnt = 0;
        if (chunk_row(c, span) >= 0) {
            const int64_t lo = c.span_off[span] - base, hi = c.span_off[span + 1] - base;
            for (int64_t e = lo + lane; e < hi; e += kWave) cnt += chunk_var(c, c.genes[e]) >= 0;
            for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, kWave);
        }
        if (lane == 0) kept[span] = cnt;
    }
}

__global__ __launch_bounds__(256) void k_mat_write(DevChunk c, const int64_t *out_off, uint32_t *rows, uint32_t *vars, float *vals) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wave_id = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int64_t n_waves = (int64_t)gridDim.x * blockDim.x / kWave;
    const int64_t base = c.base;
    for (int64_t span = wave_id; span < c.n_spans; span += n_waves) {
        const int64_t row = chunk_row(c, span);
        if (row < 0) continue;
        const int64_t lo = c.span_off[span] - base, hi = c.span_off[span + 1] - base;
        const double ns = norm_factor(cell_total(c, lo, hi, lane), c.target_sum);
        int64_t pos = out_off[span];
        for (int64_t e0 = lo; e0 < hi; e0 += kWave) {
            const int64_t e = e0 + lane;
            const int32_t m = e < hi ? chunk_var(c, c.genes[e]) : -1;
            const bool keep = m >= 0;
            const uint64_t mask = __ballot(keep);
            const int before = __popcll(mask & ((1ull << lane) - 1ull));
            if (keep) {
                rows[pos + before] = (uint32_t)row;
                vars[pos + before] = (uint32_t)m;
                vals[pos + before] = lazy_transform(c.values[e], ns, c.target_sum, c.log1p);
            }
            pos += __popcll(mask);
        }
    }
}

void launch_materialize_count(const DevChunk &c, int64_t *kept, hipStream_t s) {
    if (c.n_spans <= 0) return;
    int g = (int)std::min<int64_t>(ceil_div(c.n_spans, 4), 8192);
    hipLaunchKernelGGL(k_mat_count, dim3(g), dim3(256), 0, s, c, kept);
}
void launch_materialize(const DevChunk &c, const int64_t *out_off, uint32_t *rows, uint32_t *vars, float *vals, hipStream_t s) {
    if (c.n_spans <= 0) return;
    int g = (int)std::min<int64_t>(ceil_div(c.n_spans, 4), 8192);
    hipLaunchKernelGGL(k_mat_write, dim3(g), dim3(256), 0, s, c, out_off, rows, vars, vals);
}

}  // namespace crest
