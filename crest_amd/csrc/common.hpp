// Shared definitions of the gfx950 library: the per-entry arithmetic of the lazy
// transform (identical on host and device), launch helpers and kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace crest {

constexpr int kWave = 64;
constexpr int kRowGroup = 8;      // rows interleaved per (row-group, gene) unit of a tile
constexpr int kRowGroupSparse = 64;  // rows per group of the sparse FP64 Gram's packed entries
constexpr int kGramBlock = 128;   // Gram output block (per workgroup)
constexpr int kStatsWindow = 2048;  // vars accumulated in LDS per gene_stats window

// Element formats of a densified tile.
enum TileFmt : int { TILE_F64 = 0, TILE_F32 = 1, TILE_BF16X2 = 2, TILE_F16X2 = 3 };

#define CREST_LOG1PF_FN __host__ __device__ inline
#include "log1pf_glibc.h"

// normalize_total + log1p of one entry (kernels.rs:77-82).  `ns` is the
// per-cell factor target_sum/total (1 when total == 0).  log1p is f32 with the
// rounding of the reference's f32::ln_1p (glibc log1pf), identical on host and device.
__host__ __device__ inline float lazy_transform(float x, double ns, double target_sum, int log1p) {
    if (!(target_sum > 0.0 || log1p)) return x;
    float y = target_sum > 0.0 ? (float)((double)x * ns) : x;
    return log1p ? crest_log1pf_glibc(y) : y;
}

__host__ __device__ inline double norm_factor(double total, double target_sum) {
    return (target_sum > 0.0 && total > 0.0) ? target_sum / total : 1.0;
}

// scanpy-style scale with clipping, sparse form (kernels.rs:209-218):
// s = clip((x - mu)/sd) - clip(-mu/sd); the identity when scaling is off.
__host__ __device__ inline double clip_mv(double z, double mv) {
    if (mv > 0.0) {
        if (z < -mv) return -mv;
        if (z > mv) return mv;
    }
    return z;
}
__host__ __device__ inline double scale_value(float v, double mu, double sd, double mv, int scale) {
    if (!scale) return (double)v;
    return clip_mv(((double)v - mu) / sd, mv) - clip_mv(-mu / sd, mv);
}

// Chunk geometry on the device: spans of entries per cell.
struct DevChunk {
    const int64_t *span_off;   // [n_spans+1] absolute offsets; entries are genes[off-base]
    const uint32_t *span_cell; // [n_spans] store cell id, or nullptr -> first_cell + i
    uint32_t first_cell;
    int64_t n_spans;
    const uint32_t *genes;
    const float *values;
    const int64_t *cell_map;   // store cell -> row
    int64_t n_cell_map;
    const int32_t *gene_var;   // store gene -> var (-1 filtered)
    int64_t n_gene_map;
    double target_sum;
    int log1p;
    int64_t base;              // absolute offset of genes[0] / values[0]
    int64_t nnz;               // entries of genes / values (span_off[n_spans] - base for a whole chunk)
};

__device__ inline int64_t chunk_row(const DevChunk &c, int64_t i) {
    uint64_t cell = c.span_cell ? (uint64_t)c.span_cell[i] : (uint64_t)c.first_cell + (uint64_t)i;
    if (cell >= (uint64_t)c.n_cell_map) return -1;
    return c.cell_map[cell];
}
__device__ inline int32_t chunk_var(const DevChunk &c, uint32_t g) {
    return (int64_t)g < c.n_gene_map ? c.gene_var[g] : -1;
}

__device__ inline double wave_sum_f64(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Scale parameters on the device.
struct DevScale {
    const int32_t *sub;  // var -> hvg index (-1 = not in the Gram)
    const double *mean;  // [d]
    const double *sd;    // [d]
    double max_value;
    int scale;
    int d;
};

// Batches of independent cell subsets (many subsets in one launch).  A segment is a
// contiguous run of cells [cell0, cell0 + n) of one subset; in a densified tile its
// rows start at row group grp0 (rows padded to 16 with zeros).  A row range is a
// piece [r0, r1) of a segment's rows handled by one workgroup.
struct SegDesc {
    int64_t cell0, n, grp0;
    int32_t subset, pad_;
};
struct RowRange {
    int64_t r0, r1;
    int32_t seg, pad_;
};
// One Gram block over row groups [g0, g1) of a tile into partial slot `slot`
// (blk < 0: no work).  A reduce item sums slots [slot0, slot0 + n_slots) of block
// blk into the Gram of `subset`.
struct GramItem {
    int32_t blk, subset;
    int64_t g0, g1, slot;  // slot < 0: add the block into the Gram of `subset` directly
};
struct RedItem {
    int32_t blk, subset;
    int64_t slot0;
    int32_t n_slots, pad_;
};

// Rows [a, b) of a chunk as a chunk of their own (span offsets stay absolute).
__host__ __device__ inline DevChunk seg_rows(DevChunk c, int64_t a, int64_t b) {
    c.span_off += a;
    if (c.span_cell)
        c.span_cell += a;
    else
        c.first_cell += (uint32_t)a;
    c.n_spans = b - a;
    return c;
}

// ---- kernel launchers (each defined next to its kernel) ----------------------
void launch_gene_stats_ranges(const DevChunk &c, const SegDesc *segs, const RowRange *ranges, int64_t n_ranges,
                              double *out, int n_vars, hipStream_t s);
void densify_set_attrs();
bool gram_dma_enabled();
void launch_densify_ranges(const DevChunk &c, const DevScale &sc, int fmt, void *tile_a, void *tile_b, int d_pad,
                           const SegDesc *segs, const RowRange *ranges, int64_t n_ranges, double *colsum,
                           double *gram, hipStream_t s);
void launch_project_ranges(const DevChunk &c, const DevScale &sc, const double *V, int k, const SegDesc *segs,
                           const RowRange *ranges, int64_t n_ranges, float *out, hipStream_t s);
void launch_add_diag(const double *src, int d_pad, double *gram, hipStream_t s);
void launch_gram_items(int fmt, const void *tile_a, const void *tile_b, int d_pad, const GramItem *items,
                       int64_t n_items, int grid, double *partial, double *gram, hipStream_t s);
int gram_items_occupancy(int fmt);
void launch_gram_reduce_items(const double *partial, const RedItem *items, int64_t n_items, int d_pad,
                              double *gram, hipStream_t s);
void launch_cov_batch(const double *gram, const double *colsum, int d, int ld_g, const int64_t *dev_cell_off,
                      int n_sub, double *cov, double *total, hipStream_t s, const int32_t *ids = nullptr);
void launch_eig_select_batch(const double *evecs, const double *evals, int d, int k, int n_sub, double *variance,
                             double *loadings, hipStream_t s, int ncols = 0, const int32_t *ids = nullptr);
// The n x n dual of small subsets: dense centred scaled rows, the covariance trace from
// the Gram diagonal, and the eigenpairs of C from those of Zc Zc^T.
void launch_dense_centred(const DevChunk &c, const DevScale &sc, const SegDesc *segs, int n_seg, int m,
                          const double *colsum, int ld_c, double *out, hipStream_t s);
void launch_trace_gram(const double *gram, const double *colsum, int d, int ld_g, const int64_t *dev_cell_off,
                       const int32_t *ids, int n_ids, double *total, hipStream_t s);
void launch_dual_finish(double *V, const double *W, int d, int p, const int64_t *dev_cell_off, const int32_t *ids,
                        int n_ids, double *evals, hipStream_t s);
void launch_shift_fill_batch(const double *colsum, int ld_c, const double *V, int d, int k,
                             const int64_t *dev_cell_off, int n_sub, int64_t max_rows, float *x, float *shift_f,
                             hipStream_t s);
void launch_gene_stats(const DevChunk &c, double *out, int n_vars, hipStream_t s);
void launch_qc(const DevChunk &c, const uint8_t *flag, double *row_total, double *row_ngenes, double *row_flag,
               uint32_t *gene_ncells, double *gene_total, hipStream_t s);
void launch_densify(const DevChunk &c, const DevScale &sc, int fmt, void *tile_a, void *tile_b,
                    int64_t rows_pad, int d_pad, double *colsum, double *gram, hipStream_t s);
void launch_materialize_count(const DevChunk &c, int64_t *row_kept, hipStream_t s);
void launch_materialize(const DevChunk &c, const int64_t *out_off, uint32_t *rows, uint32_t *vars,
                        float *vals, hipStream_t s);
void launch_project(const DevChunk &c, const DevScale &sc, const double *V, int k, float *out,
                    int64_t row0, hipStream_t s);
void launch_gram(int fmt, const void *tile_a, const void *tile_b, int64_t rows_pad, int d_pad,
                 int n_split, double *partial, hipStream_t s);
void launch_gram_reduce(const double *partial, int n_split, int d_pad, double *gram, hipStream_t s);
void launch_cov(const double *gram, const double *colsum, int d, int d_pad, double n, double *cov,
                double *total, hipStream_t s);
void launch_eig_select(const double *evecs, const double *evals, int d, int k, double *variance,
                       double *loadings, hipStream_t s, int ncols = 0);
// Top-p eigenpairs after a lower Householder tridiagonalization (kernels_eig.hip).
bool tri_topp_supported(int n);
// kernels_sytrd.hip: lower Householder tridiagonalization (LAPACK dsytrd layout), one
// workgroup per matrix; W scratch sytrd_wg_scratch(n) doubles per matrix.
size_t sytrd_wg_lds(int n);
bool sytrd_wg_supported(int n);
size_t sytrd_wg_scratch(int n);
void sytrd_set_attrs();
void launch_sytrd_lower(double *A, int64_t sA, int lda, double *D, int64_t sD, double *E, int64_t sE, double *tau,
                        int64_t sT, double *W, int n, int n_mat, hipStream_t s);
void tri_set_attrs();
void launch_tri_topp(const double *A, int64_t sA, int lda, const double *tau, int64_t sTau, const double *D,
                     int64_t sD, const double *E, int64_t sE, int n, int p, int n_mat, double *W, double *Y,
                     double *X, double *fac, uint8_t *swp, int *info, hipStream_t s);
// kernels_chefsi.hip: the pieces of the Chebyshev-filtered subspace iteration around
// its dgemm calls (blocks of at most chefsi_max_block() vectors).
int chefsi_max_block();
void chefsi_set_attrs();
void launch_ch_rand(double *v, int rows, int ld, int cols, uint64_t seed, hipStream_t s);
void launch_ch_sum(double *out, const double *part, int S, int64_t stride, int64_t n, hipStream_t s);
void launch_ch_combine(double *y_next, const double *w, int S, int64_t stride, const double *y, const double *y_prev,
                       int64_t n, double a1, double c, double a2, hipStream_t s);
void launch_ch_chol_rinv(const double *Gp, int S, int k, double *Rinv, double *bad, hipStream_t s);
void launch_ch_jacobi(const double *Hp, int S, int n, int max_sweeps, bool strict, bool desc, double *theta,
                      double *Z, double *bad, hipStream_t s);
void launch_ch_gemm(const double *A, const double *Y, int dp, int n, int S, double *out, hipStream_t s);
void launch_ch_resid(const double *cv, const double *v, const double *theta, int d, int ld, int k, double *out,
                     hipStream_t s);
void launch_ch_lock(const double *src, const double *src_val, int ld, int n, double *dst, double *dst_val,
                    double *dst_scaled, hipStream_t s);
// batches (chefsi_topp_batch): per-matrix state of one iteration
struct ChbMat {
    int32_t ka, m, voff, L, J, pad_;
    double c, e;
};
void chefsi_batch_set_attrs();
void launch_chb_gemm(const double *A, const double *Y, int dp, int K, int nm, const ChbMat *P, int step, bool use_voff,
                     double *out, hipStream_t s);
void launch_chb_combine(double *y_next, const double *w, const double *y, const double *y_prev, int dp, int K, int nm,
                        const ChbMat *P, int step, bool prev_voff, hipStream_t s);
void launch_chb_zero_tail(double *y, int dp, int K, int nm, const ChbMat *P, hipStream_t s);
void launch_chb_chol_rinv(const double *G, int K, int nm, const ChbMat *P, double *Rinv, double *stat, hipStream_t s);
void launch_chb_jacobi(const double *H, int K, int nm, const ChbMat *P, int max_sweeps, bool strict, bool desc,
                       double *Z, double *stat, hipStream_t s);
void launch_chb_resid(const double *cv, const double *v, int d, int dp, int K, int nm, const ChbMat *P, double *stat,
                      hipStream_t s);
void launch_chb_lock(const double *v, const double *stat, int dp, int K, int p, int nm, const ChbMat *P, double *QL,
                     double *thL, double *wl, double *ql, hipStream_t s);
void launch_chb_pad(const double *C, int d, int dp, int nm, const ChbMat *P, double *A, hipStream_t s);
void launch_chb_copy(double *dst, const double *src, int src_cols, int dp, int K, int nm, const ChbMat *P,
                     bool use_voff, hipStream_t s);
void launch_chb_final(const double *v, const double *stat, int d, int dp, int K, int p, int nm, double *vout,
                      double *evals, hipStream_t s);
void launch_chb_rand(double *v, int rows, int ld, int cols, int nm, uint64_t seed, hipStream_t s);
void launch_shift_fill(const double *colsum, const double *V, int d, int k, double n, int64_t n_rows,
                       float *x, float *shift_scratch, hipStream_t s);
// kNN screen operands: rows are padded to knn_rows_padded(n) and stored in a
// spatial order; `work` holds knn_work_bytes(nc, nq, p) bytes (column statistics,
// sort keys, the FP16 prescale and the permutations; nq = 0 when the queries are the
// candidates).  Candidate position i of the screen is row perm_c[i], query
// position j is query row perm_q[j]; keys_c / keys_q are the sorted keys of a
// separate query set (null for a kNN graph).
struct KnnOrder {
    const uint32_t *perm_c, *perm_q, *keys_c, *keys_q;
    const float *mean;        // centring of the operands (per subset in a batch: + y*p)
    const uint32_t *maxbits;  // source of the FP16 prescale (per subset in a batch: + y)
};
// Certificate of a screen's candidate lists.  Every candidate a half-list rejected
// or evicted has a screen value at least the list's final L-th value t_h; the screen
// value of a pair differs from s^2 |c - q|^2 (s the operands' prescale) by at most
//   E(a, b) = c1 a b + c2 (a + b)^2 + c3 (a + b) + c4,   a = s|q'|, b = s|c'|
// (centred norms), and a candidate closer than the k-th reranked distance D_k has
// |c'| <= |q'| + sqrt(D_k).  So when min(t_0, t_1) - E(a, a + s sqrt(D_k)) > s^2 D_k,
// no rejected candidate is as close as the k-th: the reported set is the exact one.
// Queries for which this fails are counted and listed (their output rows).
struct KnnCert {
    const float *cand_d;  // the screen's kept values [query position][n_cand], halves ascending
    double c1, c2, c3, c4;
    uint32_t *count;      // null: no certificate
    uint32_t *list;       // the first `cap` uncertified query rows
    int64_t cap;
    uint32_t *pos;        // optional: their query positions in the screen's order
};
KnnCert knn_cert(int prec, int p, int kpad, const float *cand_d, uint32_t *count, uint32_t *list, int64_t cap);
double knn_cert_scale();
// A batch of independent subsets in one screen: subset y's operand rows are
// row0[y] .. row0[y] + nrow[y] - 1 (then padding up to a multiple of 512); blk_seg
// gives the subset of every kSegBlock-row block of operand rows, and order lists the
// 512-row query blocks in the order their workgroups are dispatched.  All null for a
// single set of rows.
constexpr int kSegBlock = 256;
struct KnnSeg {
    const int64_t *row0;
    const int64_t *nrow;
    const int32_t *blk_seg;
    const int32_t *order;
};
int64_t knn_rows_padded(int64_t n);
size_t knn_work_bytes(int64_t nc, int64_t nq, int p);
KnnOrder launch_knn_prep(int prec, const float *xc, int64_t nc, const float *xq, int64_t nq, int p, int kpad,
                         void *work, void *xa, void *xb, hipStream_t s);
int knn_kpad(int prec, int p);
bool launch_knn_screen(int prec, const void *xa, const void *xb, int64_t nc, int64_t nq, int kpad, int L,
                       int self, const KnnOrder &o, float *cand_d, uint32_t *cand_i, hipStream_t s);
// qmap (a recheck of a subset of the queries): query row r of xq answers query
// qmap[r] — its output row, and with `self` the candidate row it must not report.
void launch_knn_rerank(const float *xq, int64_t nq, const float *xc, int64_t nc, int p, const KnnOrder &o,
                       int self, const uint32_t *cand_i, int n_cand, int k, uint32_t *idx_out, float *dist_out,
                       hipStream_t s, const int32_t *blk_seg = nullptr, const int64_t *cell_off = nullptr,
                       const KnnCert *cert = nullptr, const uint32_t *qmap = nullptr);
void launch_gather_rows(const float *x, int p, const uint32_t *rows, int64_t m, float *out, hipStream_t s);
void launch_knn_place(const uint32_t *ti, const float *td, int64_t m, int k1, int k, const uint32_t *rows, int self,
                      uint32_t *idx_out, float *dist_out, hipStream_t s, const uint32_t *ibase = nullptr);
// Batch forms (subset y = rows cell_off[y] .. cell_off[y+1]-1 of x, operand rows
// row0[y] .. of n_pad in all; max_rows / max_pad are the largest n_y and padded n_y).
size_t knn_batch_work_bytes(int64_t n, int64_t n_pad, int n_sub, int p);
KnnOrder launch_knn_prep_batch(int prec, const float *x, int64_t n, const int64_t *cell_off, int n_sub,
                               int64_t max_rows, const int64_t *row0, int64_t n_pad, int64_t max_pad, int p,
                               int kpad, void *work, void *xa, void *xb, hipStream_t s);
bool launch_knn_screen_batch(int prec, const void *xa, const void *xb, int64_t n_pad, int kpad, int L,
                             const KnnSeg &sg, float *cand_d, uint32_t *cand_i, hipStream_t s);
// Large-k kNN: f32 row norms; per query row of an FP32 dot-product block (row
// stride ld, a multiple of 4; cn readable up to ld entries), the indices of the k2
// smallest screened keys; exact rank of those k2 and the k
// nearest in the reference's output order (k2 <= kKnnLargeMaxCand).
constexpr int kKnnLargeMaxCand = 4096;
void launch_knn_sqnorm(const float *x, int64_t n, int p, float *out, hipStream_t s);
void launch_knn_select_rows(const float *dots, int64_t ld, int64_t nc, const float *qn, const float *cn,
                            int64_t q0, int64_t rows, int self, int k2, int cap, uint32_t *comp_key,
                            uint32_t *comp_idx, uint32_t *cand, hipStream_t s);
// With tau and cert: the certificate of the GEMM screen (knn_cert_gemm) counted per row.
void launch_knn_rerank_sort(const float *xq, const float *xc, int p, const uint32_t *cand, int k2, int k,
                            int64_t q0, int64_t rows, uint32_t *idx_out, float *dist_out, hipStream_t s,
                            const float *tau = nullptr, const KnnCert *cert = nullptr,
                            const float *xq_scr = nullptr);
// Threshold scan of a recheck (m gathered query rows answering output rows rows[];
// xcc / xqc centred on the candidates' mean, cn / qn their FP32 squared norms):
// candidates within the widened current k-th distance appended (at most cap per row,
// cnt counts all; cand pre-filled with 0xFFFFFFFF, cnt zeroed).
void launch_knn_thr_scan(const float *xcc, const float *cn, int64_t nc, int p, const float *xqc, const float *qn,
                         int64_t m, const float *dist, const uint32_t *rows, int k, int self, float *thr, uint32_t *cnt,
                         uint32_t *cand, int cap, hipStream_t s);
// The thresholds alone (thr[r] for the m gathered rows), and the threshold test on a
// block of GEMM dot products (rows q0 .. q0 + nrows - 1 of the gathered set).
void launch_knn_scan_thresholds(const float *dist, const uint32_t *rows, const float *xqc, int64_t m, int k, int p,
                                float *thr, hipStream_t s);
void launch_knn_thr_filter(const float *dots, int64_t ld, int64_t nc, const float *cn, const float *qn,
                           const float *thr, const uint32_t *rows, int self, int64_t q0, int64_t nrows, uint32_t *cnt,
                           uint32_t *cand, int cap, hipStream_t s);
// The threshold test on the FP16 screen's own operands (xa / xb, kpad halves per row,
// m queries at screen positions pos answering output rows rows[]): thresholds
// s^2 T + E16(a, a + s sqrt(T)) from the queries' current k-th distances, then every
// candidate whose FP16 screen value is at most its query's threshold appended as its
// data row (perm_c), the query itself excluded in a graph.
void launch_knn_thr16(const void *xa, const void *xb, int kpad, int64_t nc, const KnnOrder &o, const float *xq_rows,
                      const float *dist, const uint32_t *rows, const uint32_t *pos, int64_t m, int k, int p, int self,
                      float *thr, uint32_t *cnt, uint32_t *cand, int cap, hipStream_t s);
// One block of a batch's threshold test: queries q0 .. q0 + nq - 1 (nq <= 128) of one
// subset against its candidates c0 .. c1 - 1, in slices of `slice` rows.
struct KnnScanBlock {
    int64_t q0, c0, c1, slice;
    int32_t nq, pad_;
};
void launch_knn_thr_mfma_blocks(const float *xcc, const float *cn, int p, const float *xqc, const float *qn,
                                const float *thr, const uint32_t *rows, const KnnScanBlock *blocks, int n_blocks,
                                int64_t max_slices, uint32_t *cnt, uint32_t *cand, int cap, hipStream_t s);
void launch_knn_centre_seg(const float *x, int64_t n, const int64_t *cell_off, int n_sub, int p, const float *means,
                           float *out, hipStream_t s);
const float *knn_batch_means(void *work, int64_t n, int64_t n_pad, int n_sub, int p);
// The threshold test on FP32 MFMA (p <= 96; thresholds from launch_knn_scan_thresholds).
void launch_knn_thr_mfma(const float *xcc, const float *cn, int64_t nc, int p, const float *xqc, const float *qn,
                         int64_t m, const float *thr, const uint32_t *rows, int self, uint32_t *cnt, uint32_t *cand,
                         int cap, hipStream_t s);
// The GEMM screen's operands: x - mean (mean computed from x when compute_mean; part
// holds knn_centre_parts(n) * p doubles).
int knn_centre_parts(int64_t n);
void launch_knn_centre(const float *x, int64_t n, int p, float *mean, bool compute_mean, double *part, float *out,
                       hipStream_t s);
void launch_knn_cand_max(const float *dots, int64_t ld, const float *qn, const float *cn, int64_t q0, int64_t rows,
                         const uint32_t *cand, int k2, float *tau, hipStream_t s);
KnnCert knn_cert_gemm(int p, uint32_t *count);
size_t conn_temp_bytes(int64_t m);
void launch_connectivities(const uint32_t *idx, const float *dist, int64_t n, int k, int n_neighbors,
                           uint64_t *keys, uint64_t *keys2, float *w, float *w2, float *wu, uint32_t *flag,
                           uint32_t *pos, double *part, void *temp, size_t temp_bytes, uint32_t *rows,
                           uint32_t *cols, float *wout, uint32_t *count, hipStream_t s);
// kernels_hvg.hip: seurat HVG selection from [n_vars, 5] gene statistics (one
// workgroup); work: codes int32 [n_vars], compact double [n_vars].
bool hvg_bins_supported(int n_bins);
void launch_hvg_seurat(const double *stats, int64_t n_vars, double n_cells, int n_bins, int64_t n_top,
                       double min_mean, double max_mean, double min_disp, double max_disp, int32_t *codes,
                       double *compact, double *means, double *disp, double *disp_norm, uint8_t *hv,
                       double *mean_log, double *std_log, hipStream_t s);
void launch_fill_rows(const float *shift_f, int k, int64_t n_rows, float *x, hipStream_t s);
// kernels_sparse.hip: the FP64 Gram by sparse pair products (d_pad <= 4096): pack the
// scaled HVG entries of a chunk by 128-column bucket, then one workgroup per
// (Gram block, row split) into partial [blk][split][128 x 128] and colsum.
bool sparse_gram_supported(int d_pad);
size_t sparse_meta_words(int64_t n_rows, int d_pad);
size_t sparse_gram_lds();
void launch_sparse_pack(const DevChunk &c, const DevScale &sc, int d_pad, double *vals, uint8_t *cols, uint32_t *meta,
                        hipStream_t s);
// splits split0 .. split0 + n_sl - 1 of n_split (over the row groups) of the sparse Gram;
// launch_sparse_colsum then sums the column partials of all n_split splits
void launch_gram_sparse_part(const double *vals, const uint8_t *cols, const uint32_t *meta, const int64_t *span_off,
                             int64_t base, int64_t n_rows, int d_pad, int n_split, int split0, int n_sl, double *partial,
                             double *cs_partial, hipStream_t s);
void launch_sparse_colsum(const double *cs_partial, int d_pad, int n_split, double *colsum, hipStream_t s);
void launch_gram_sparse(const double *vals, const uint8_t *cols, const uint32_t *meta, const int64_t *span_off,
                        int64_t base, int64_t n_rows, int d_pad, int n_split, double *partial, double *cs_partial,
                        double *colsum, hipStream_t s);
// the batch forms: 64-row groups from a table (each within one subset, scale
// parameters of that subset), Gram items over group ranges (GramItem.g0/g1 index
// the group table)
void launch_sparse_pack_groups(const DevChunk &c, const DevScale &sc, int d_pad, const SegDesc *groups,
                               int64_t n_groups, double *vals, uint8_t *cols, uint32_t *meta, hipStream_t s);
void launch_gram_sparse_items(const double *vals, const uint8_t *cols, const uint32_t *meta, const int64_t *span_off,
                              int64_t base, const SegDesc *groups, const GramItem *items, int64_t n_items, int d_pad,
                              double *partial, double *gram, double *colsum, hipStream_t s);
void sparse_set_attrs();
void launch_fill_f64(double *p, int64_t n, double v, hipStream_t s);
void launch_fill_u32(uint32_t *p, int64_t n, uint32_t v, hipStream_t s);
void launch_fill_f32(float *p, int64_t n, float v, hipStream_t s);

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

}  // namespace crest
