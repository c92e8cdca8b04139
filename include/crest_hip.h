/*
 * crest_hip.h — C ABI of the MI355X (gfx950) implementation of CREST's streamed
 * counts -> scale -> PCA -> kNN path.
 *
 * Every operator entry point below replaces one function of the reference's PyO3
 * module `crest.crest` (src/py.rs) and keeps its argument meaning, accumulation
 * semantics (outputs marked "+=" are accumulated in place, as the reference does)
 * and error classes.  Pointers are host pointers unless the name says `dev`.
 * All functions are thread-safe (one mutex per device context) and return
 * CREST_OK (0) or a negative status; the message of the last failure on the
 * calling thread is returned by crest_last_error().
 *
 * Status classes map onto the reference's Python exceptions:
 *   CREST_E_VALUE   (-1)  -> ValueError   (py.rs:11-27,162-189,226-228,250-268,279-281,295-304)
 *   CREST_E_RUNTIME (-2)  -> RuntimeError (py.rs:396, kernels.rs:315-317)
 *   CREST_E_DEVICE  (-3)  -> RuntimeError (HIP / rocSOLVER failure; no reference counterpart)
 */
#ifndef CREST_HIP_H
#define CREST_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CREST_OK 0
#define CREST_E_VALUE (-1)
#define CREST_E_RUNTIME (-2)
#define CREST_E_DEVICE (-3)

/* Arithmetic of the dense contractions (Gram SYRK and the kNN distance screen).
 * The transform, gene statistics, centring, eigensolve, projection and the kNN
 * rerank are FP64 in every mode.  The screen only chooses candidates: the reported
 * neighbours and distances always come from the exact rerank.                    */
#define CREST_PREC_FP64 0   /* Gram on FP64 MFMA; kNN screen on FP32 MFMA          */
#define CREST_PREC_FP32 1   /* FP32 MFMA, FP32 accumulation flushed to FP64          */
#define CREST_PREC_BF16X3 2 /* split-BF16 hi+lo, three MFMA products (the Gram adds
                               the exact diagonal of the lo.lo term), FP32 accum.   */
#define CREST_PREC_FP16 3   /* Gram: split-FP16 hi+lo, three MFMA products plus the
                               exact diagonal of the lo.lo term, FP32 accumulation;
                               needs scaled values clipped at 0 < max_value <= 30000
                               (else CREST_E_VALUE).  kNN screen: FP16 operands
                               (power-of-two prescaled), one MFMA product, FP32
                               accumulation                                         */

/* One chunk of raw counts plus the recorded filters and lazy transform: the
 * 9-tuple `(genes, values, cell_map, gene_map, target_sum, log1p, cells, indptr,
 * first_cell)` of crest/core.py:265-278 and py.rs:162-189,195-211.
 * Exactly one of `cells` / `indptr` is non-NULL.                                  */
typedef struct crest_chunk {
    const uint32_t *genes;    /* [nnz] store gene id of each entry                  */
    const float *values;      /* [nnz] raw value of each entry                      */
    size_t nnz;
    const int64_t *cell_map;  /* [n_cell_map] store cell -> row (-1 = filtered)     */
    size_t n_cell_map;
    const int32_t *gene_map;  /* [n_gene_map] store gene -> var (-1 = filtered)     */
    size_t n_gene_map;
    double target_sum;        /* normalize_total target (<= 0: off)                 */
    int32_t log1p;            /* apply log1p after normalisation                    */
    const uint32_t *cells;    /* [nnz] per-entry store cell id, non-decreasing      */
    const int64_t *indptr;    /* [n_indptr] absolute CSR offsets of cells first..   */
    size_t n_indptr;
    uint32_t first_cell;
} crest_chunk;

/* Library / device state ------------------------------------------------------ */

/* Select the device used by this thread's calls (default 0). */
int crest_set_device(int32_t device);
/* Precision used by crest_gram_accumulate / crest_knn_graph (default FP64). */
int crest_set_precision(int32_t prec);
int32_t crest_get_precision(void);
/* Precision of the kNN distance screen alone (crest_knn_graph): CREST_PREC_FP64 /
 * FP32 / BF16X3 / FP16, or -1 (default) to follow crest_set_precision.  The getter
 * returns the precision in effect.  No reference counterpart (the reference's
 * screen is FP32).                                                               */
int crest_set_knn_precision(int32_t prec);
int32_t crest_get_knn_precision(void);
/* Message of the last failing call on this thread ("" if none). */
const char *crest_last_error(void);
/* Number of visible HIP devices (0 when none; never fails). */
int32_t crest_device_count(void);

/* Operator entry points (host buffers) ------------------------------------------ */

/* Replaces py.rs:221-232 `gene_stats(*ctx, out)` (kernels.rs:141-160).
 * out (+=) is [n_vars, 5] row-major f64: Σx, Σx², Σexpm1(x), Σexpm1(x)², nnz
 * over the transformed values of kept cells and genes; n_vars = max(gene_map)+1. */
int crest_gene_stats(const crest_chunk *chunk, double *out, size_t out_len);

/* Replaces py.rs:234-248 `qc(*ctx, flag, cell_total, cell_ngenes, cell_flag,
 * gene_ncells, gene_total)` (kernels.rs:162-195), the accumulation behind
 * crest.pp.calculate_qc_metrics (call with an un-transformed chunk).  For every kept
 * entry with a non-zero value: cell_total[row] (+=) Σx, cell_ngenes[row] (+=) count,
 * cell_flag[row] (+=) Σx over genes with flag[var] != 0; gene_ncells[var] (+=) count,
 * gene_total[var] (+=) Σx.  The row arrays hold at least n_rows = max(cell_map)+1
 * entries; flag and the gene arrays exactly n_vars.  Sums are FP64 and exact for
 * integer counts.                                                                  */
int crest_qc(const crest_chunk *chunk, const uint8_t *flag, size_t n_flag, double *cell_total, size_t n_cell_total,
             uint32_t *cell_ngenes, size_t n_cell_ngenes, double *cell_flag, size_t n_cell_flag,
             uint32_t *gene_ncells, size_t n_gene_ncells, double *gene_total, size_t n_gene_total);

/* Replaces py.rs:270-287 `gram_accumulate(*ctx, sub_map, gram, colsum, scale,
 * mean, std, max_value)` (kernels.rs:197-309).  gram (+=) is [d, d] row-major
 * f64, of which the upper triangle a <= b is written; colsum (+=) is [d].
 * sub_map is [n_sub_map] var -> 0..d-1 or -1; mean/std are [d] (may be NULL when
 * scale == 0).                                                                    */
int crest_gram_accumulate(const crest_chunk *chunk, const int32_t *sub_map, size_t n_sub_map,
                          double *gram, size_t gram_len, double *colsum, size_t d, int32_t scale,
                          const double *mean, size_t n_mean, const double *std, size_t n_std,
                          double max_value);

/* No counterpart in the reference: crest_gram_accumulate for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  The same densify + Gram kernels over a chunk whose arrays all
 * live in device memory, launched on `stream` (a hipStream_t; NULL = the default
 * stream) with no allocation, no memset and no synchronisation: the caller waits on
 * the stream.  dev_chunk->indptr is device memory with indptr[0] == 0 (genes/values
 * start at that entry), cells must be NULL; the arrays are not validated entry by
 * entry (sub_map values must be < d).  dev_gram (+=) is [d, d] row-major (upper
 * triangle written), dev_colsum (+=) [d].  Precision: crest_set_precision's.
 * Scratch: one 128 MiB device array per GPU (the d_pad x d_pad Gram, the partials
 * and the densified tiles), so d <= 2048 (larger d: CREST_E_VALUE).  Calls that
 * overlap on one GPU share that array and race; calls on one stream do not.      */
int crest_gram_accumulate_noalloc(const crest_chunk *dev_chunk, const int32_t *dev_sub_map, size_t n_sub_map,
                                  double *dev_gram, double *dev_colsum, size_t d, int32_t scale,
                                  const double *dev_mean, const double *dev_std, double max_value, void *stream);

/* Replaces py.rs:386-399 `pca_from_gram(gram, colsum, n, k)` (kernels.rs:311-348).
 * Writes variance [k'] (descending, floored at 0), loadings [d, k'] row-major with
 * the sklearn svd_flip sign rule, *total = trace of the covariance and *k_out = k'
 * = min(k, d).  The caller sizes variance/loadings for min(k, d).                */
int crest_pca_from_gram(const double *gram, size_t gram_len, const double *colsum, size_t d,
                        size_t n, size_t k, double *variance, double *loadings, double *total,
                        size_t *k_out);

/* Replaces py.rs:289-309 `project(*ctx, sub_map, loadings, out, scale, mean, std,
 * max_value)` (kernels.rs:350-371).  out (+=) is [n_rows, k] f32 with n_rows =
 * max(cell_map)+1; out_len must be >= n_rows*k.  loadings is [d, k] row-major.    */
int crest_project(const crest_chunk *chunk, const int32_t *sub_map, size_t n_sub_map,
                  const double *loadings, size_t d, size_t k, float *out, size_t out_len,
                  int32_t scale, const double *mean, size_t n_mean, const double *std,
                  size_t n_std, double max_value);

/* Replaces py.rs:52-70 `knn_graph(data, k, exact=None, seed=0)` (knn.rs:335-369).
 * data is [n, dim] row-major f32.  k' = min(k, n-1) (>= 1 when n > 1) is written to
 * *k_out; indices [n, k'] u32 and distances [n, k'] f32 (Euclidean, ascending, ties
 * by index, self excluded) must hold n*min(k, max(n-1,1)) entries.  The search is
 * brute force for every n (no IVF): `exact` (-1 = None, 0, 1) and `seed` are
 * accepted and ignored.  Any dim.  k' <= 64 uses the screen precision of
 * crest_set_knn_precision, and every query's set is certified: the screen's error
 * bound proves that no candidate it dropped is as close as the k-th reranked
 * neighbour, and the queries it cannot prove are screened again in FP32 and
 * certified again (crest_knn_certificate counts them).  64 < k' <= 3584 (Scrublet's
 * k, doublets.py:131-137), or rows wider than the register screens hold (134 columns
 * FP32/FP64, 137 split-BF16, 252 FP16), screen in FP32 on GEMM dot products, whose
 * sets are certified against that screen's bound (no recheck); larger k' is a
 * ValueError.                                                                     */
int crest_knn_graph(const float *data, size_t n, size_t dim, size_t k, int32_t exact,
                    uint64_t seed, uint32_t *indices, float *distances, size_t *k_out);

/* Replaces py.rs:635-653 `knn_query(reference, query, k, exact=None, seed=0)`
 * (knn.rs:485-546), the reference-to-query search behind crest.ingest.  reference
 * is [n_ref, dim] and query [n_query, dim], both row-major f32.  k' = min(k, n_ref)
 * (0 when n_query == 0) is written to *k_out; indices [n_query, k'] u32 and
 * distances [n_query, k'] f32 (Euclidean, ascending, ties by index) must hold
 * n_query*min(k, n_ref) entries.  Brute force with the certificate of
 * crest_knn_graph: `exact` and `seed` are accepted and ignored.                  */
int crest_knn_query(const float *reference, size_t n_ref, const float *query, size_t n_query, size_t dim,
                    size_t k, int32_t exact, uint64_t seed, uint32_t *indices, float *distances, size_t *k_out);

/* The certificate of the last crest_knn_graph / crest_knn_query / pipeline kNN on
 * this device (no reference counterpart: the reference searches in f32 throughout).
 * checked = queries whose screened set was certified or rechecked; rechecked = those
 * the screen's bound could not certify, screened again by the FP32 GEMM screen;
 * uncertified = those still not proven after that (a candidate within the FP32
 * screen's error of the k-th distance, i.e. a near tie).  k' > 64 and rows too wide
 * for the register screens search by the GEMM screen directly: rechecked = 0 and
 * uncertified = the queries its bound does not prove.                             */
int crest_knn_certificate(int64_t *checked, int64_t *rechecked, int64_t *uncertified);

/* The certificate's bound for a screen precision and row width (host only, no device
 * needed): coef[4] = (c1, c2, c3, c4) with |screen(q, c) - s^2 |c - q|^2| <=
 * c1 a b + c2 (a + b)^2 + c3 (a + b) + c4, a = s|q - mean|, b = s|c - mean| (s the
 * FP16 power-of-two prescale, 1 otherwise); *kpad = the operand width (-1: no
 * register screen at this width).  prec = -1: the FP32 GEMM screen of the large-k
 * path (rows centred the same way; *kpad = dim).  No reference counterpart.       */
int crest_knn_screen_bound(int32_t prec, size_t dim, double *coef, int32_t *kpad);

/* Replaces py.rs:72-99 `connectivities(indices, distances, n_neighbors)`
 * (umap/graph.rs:29-107, scanpy's fuzzy simplicial set).  indices/distances are
 * [n, k] as knn_graph returns them; writes the undirected edges with source <
 * target in (source, target) order into rows/cols/weights of capacity `cap`
 * (>= n*k suffices); *n_out is the edge count.                                    */
int crest_connectivities(const uint32_t *indices, const float *distances, size_t n, size_t k,
                         size_t n_neighbors, uint32_t *rows, uint32_t *cols, float *weights,
                         size_t cap, size_t *n_out);

/* Replaces py.rs:213-219 `materialize(*ctx)` (kernels.rs:122-139) for parity
 * checks of the lazy transform.  Writes the kept, transformed entries of the chunk
 * in chunk order: rows/vars u32 and values f32, each of capacity `cap` (>= nnz);
 * *n_out is the count written.                                                    */
int crest_materialize(const crest_chunk *chunk, uint32_t *rows, uint32_t *vars, float *values,
                      size_t cap, size_t *n_out);

/* Sessions (no reference counterpart; the explicit device context of one tl.pca or
 * highly_variable_genes run).  The reference keeps its accumulators in numpy arrays
 * between per-chunk native calls (crest/tl.py:24-91, crest/pp.py:115-120); a session
 * keeps them in device memory instead: chunks of any chunk source (the 9-tuple of
 * crest/core.py:265-278, host memory) are pushed one at a time through two pinned
 * staging slots and a copy stream, so the upload of one chunk overlaps the kernels of
 * the previous one, and nothing is downloaded until a result is asked for.  A call
 * returns once its chunk is staged; kernels run asynchronously.  The cell_map and
 * gene_map arrays are read on their first use in each pass (statistics, Gram,
 * projection) and again only when another array (pointer or length) is passed; do
 * not modify them within a pass.
 * Calls on one session must not overlap; sessions are independent.               */
typedef struct crest_session crest_session;

/* A session over n_vars kept genes (max(gene_map)+1 of its chunks) and n_rows kept
 * cells (max(cell_map)+1); gene statistics start at zero.                         */
int crest_session_create(size_t n_vars, size_t n_rows, crest_session **out);
int crest_session_destroy(crest_session *s);

/* Gene statistics of one chunk accumulated on the device (kernels.rs:141-160).     */
int crest_session_push_stats(crest_session *s, const crest_chunk *chunk);
/* The accumulated [n_vars, 5] statistics (Σx, Σx², Σexpm1(x), Σexpm1(x)², nnz).  */
int crest_session_stats(crest_session *s, double *out, size_t out_len);

/* Seurat-flavour highly variable genes (crest/pp.py:201-292) from the session's
 * statistics of n_rows cells, on the device.  n_top < 0: the mean / dispersion
 * window instead of the top n_top.  Each output has n_vars entries: means (log1p of
 * the expm1 mean), dispersions (log; NaN where zero), dispersions_norm,
 * highly_variable (0/1), mean_log, std_log (0 -> 1).  1 <= n_bins <= 256.      */
int crest_session_hvg_seurat(crest_session *s, int64_t n_top, int32_t n_bins, double min_mean, double max_mean,
                             double min_disp, double max_disp, double *means, double *dispersions,
                             double *dispersions_norm, uint8_t *highly_variable, double *mean_log,
                             double *std_log);

/* Start the Gram pass of tl.pca (crest/tl.py:61-71): sub_map [n_vars] -> HVG index
 * 0..d-1 or -1; with scale, mean/std [d] given, or both NULL to derive them on the
 * device from the session's statistics (tl.py:30-34).  Precision: crest_set_precision's
 * (FP64 by default).  Resets the session's Gram and colsum.                       */
int crest_session_begin_gram(crest_session *s, const int32_t *sub_map, size_t n_sub_map, size_t d, int32_t scale,
                             const double *mean, const double *std, double max_value);
/* gram_accumulate of one chunk (kernels.rs:221-309) into the session's device Gram. */
int crest_session_push_gram(crest_session *s, const crest_chunk *chunk);
/* pca_from_gram (kernels.rs:311-348) on the session's Gram and colsum with n = n_rows
 * and k components (tl.py:72-73): variance [k], loadings [d, k] row-major, *total,
 * colsum [d] (for the shift, tl.py:76); mean/std [d] (may be NULL) receive the scale
 * parameters in use.                                                                */
int crest_session_solve(crest_session *s, size_t k, double *variance, double *loadings, double *total,
                        double *colsum, double *mean, double *std);
/* Start the projection pass: every score row set to prefill [k] (tl.py:77-78).      */
int crest_session_begin_project(crest_session *s, const float *prefill, size_t k);
/* project of one chunk (kernels.rs:350-371) into the session's device scores.       */
int crest_session_push_project(crest_session *s, const crest_chunk *chunk);
/* The scores [n_rows, k] (f32).                                                     */
int crest_session_x_pca(crest_session *s, float *out, size_t out_len);

/* Device-resident pipeline (no reference counterpart: the fused form of tl.pca +
 * pp.neighbors over a CSR already resident in HBM) -------------------------------- */

typedef struct crest_pipeline_args {
    /* CSR of raw counts in device memory: cells x genes */
    const int64_t *dev_indptr; /* [n_cells+1], indptr[0] == 0                       */
    const uint32_t *dev_genes; /* [nnz]                                             */
    const float *dev_values;   /* [nnz]                                             */
    size_t n_cells;
    size_t n_genes;
    /* host: store gene -> HVG index 0..d-1, or -1 (kept, not HVG), length n_genes  */
    const int32_t *hvg_of_gene;
    size_t d;
    double target_sum; /* normalize_total target (<= 0: off)                       */
    int32_t log1p;
    int32_t scale;      /* pp.scale recorded                                         */
    double max_value;   /* clip (<= 0: off)                                          */
    size_t n_comps;     /* requested PCs                                             */
    size_t k;           /* neighbours per cell (knn_graph k)                         */
    int32_t prec_gram;  /* CREST_PREC_* for the Gram (FP16: split-FP16)              */
    int32_t prec_knn;   /* CREST_PREC_FP64/FP32 (FP32 screen), BF16X3 or FP16         */
    size_t chunk_rows;  /* cells per densified tile (0 = default)                    */
    /* outputs in device memory */
    float *dev_x_pca;        /* [n_cells, k_pcs]                                     */
    uint32_t *dev_knn_idx;   /* [n_cells, k']                                        */
    float *dev_knn_dist;     /* [n_cells, k']                                        */
} crest_pipeline_args;

typedef struct crest_pipeline_result {
    size_t k_pcs;          /* min(n_comps, d, n-1)                                   */
    size_t k_nn;           /* k'                                                     */
    double total_variance;
    /* per-stage device times in ms (HIP events on the pipeline stream)            */
    double ms_stats, ms_gram, ms_eig, ms_project, ms_knn_prep, ms_knn_screen, ms_knn_rerank;
    double ms_total;
} crest_pipeline_result;

/* Size the host buffers: variance [n_comps], loadings [d, n_comps] (row-major, k_pcs
 * columns used).  Runs gene_stats -> mean/std -> Gram -> eigensolve -> project ->
 * kNN on the current device's pipeline stream and waits for it to finish.         */
int crest_pipeline_run(const crest_pipeline_args *args, double *variance, double *loadings,
                       crest_pipeline_result *result);

/* The same pipeline with the CSR of raw counts in host memory (the out-of-core
 * form; no reference counterpart): args->dev_indptr/dev_genes/dev_values are
 * ignored.  Chunks of stream_rows cells (0 = 32,768) are uploaded through two
 * device slots on a copy stream while the previous chunk is computed, three passes
 * (statistics, Gram, projection), so device memory holds two chunks, the Gram and
 * the kNN operands, never the whole matrix.  host_pinned != 0 says the three arrays
 * come from crest_host_alloc and are uploaded straight from there; otherwise each
 * chunk is first copied into a pinned staging buffer.                             */
int crest_pipeline_run_streamed(const crest_pipeline_args *args, const int64_t *host_indptr,
                                const uint32_t *host_genes, const float *host_values, int32_t host_pinned,
                                size_t stream_rows, double *variance, double *loadings,
                                crest_pipeline_result *result);

/* A batch of independent cell subsets (no reference counterpart: the reference runs
 * tl.pca + pp.neighbors once per subset, crest/tl.py:37-91 and crest/pp.py:296-315).
 * Subset i is cells subset_off[i] .. subset_off[i+1]-1 of the device CSR in `pipe`;
 * each gets its own gene statistics and mean/std (over its own cells), its own
 * Gram, covariance, eigensolve and projection, and its own exact kNN graph among its
 * own cells — exactly what crest_pipeline_run computes for that subset alone.
 * Every subset needs at least max(n_comps, k) + 1 cells (smaller ones go through
 * crest_pipeline_run); all share hvg_of_gene, the transform and the precisions.
 * Outputs: pipe.dev_x_pca [n_cells, k_pcs] and pipe.dev_knn_idx / dev_knn_dist
 * [n_cells, k] by cell (kNN indices counted from the subset's first cell);
 * variance [n_subsets, n_comps], loadings [n_subsets, d, n_comps] and total
 * [n_subsets] in host memory.
 * The stages run as one launch each over all subsets (the Gram as a persistent
 * kernel over a work list of (subset, row range, block) items), except the FP64
 * eigensolve, which runs batched over eig_batch subsets per call.               */
typedef struct crest_batch_args {
    crest_pipeline_args pipe;
    size_t n_subsets;
    const int64_t *subset_off; /* host [n_subsets+1], non-decreasing, 0 .. n_cells      */
    size_t tile_rows;   /* rows of one densified batch tile (0 = 524,288)                 */
    size_t gram_rows;   /* largest row range of one Gram work item (0 = autotuned)        */
    size_t eig_batch;   /* subsets per batched eigensolve call (0 = 256)                  */
    int32_t use_graph;  /* capture the launches before and after the eigensolve into two
                           HIP graphs, replayed by later calls with identical arguments  */
} crest_batch_args;

typedef struct crest_batch_result {
    size_t k_pcs;      /* min(n_comps, d)                                             */
    size_t k_nn;       /* k                                                           */
    double ms_stats, ms_gram, ms_eig, ms_project, ms_knn_prep, ms_knn_screen, ms_knn_rerank;
    double ms_total;   /* with use_graph: ms_stats+ms_gram is the first graph, ms_eig the
                          eigensolve, ms_project the second graph (stage split not timed) */
    size_t tile_rows, gram_rows; /* shapes used (gram_rows from the autotuner if 0)   */
    int32_t graph_replayed;      /* 1 when captured graphs were replayed              */
    int32_t gram_tuned;          /* 1 when this call ran the gram_rows autotuner      */
} crest_batch_result;

int crest_pipeline_run_batched(const crest_batch_args *args, double *variance, double *loadings, double *total,
                               crest_batch_result *result);

/* Pinned (page-locked) host memory for crest_pipeline_run_streamed's inputs. */
int crest_host_alloc(size_t bytes, void **ptr);
int crest_host_free(void *ptr);

/* Release every cached device buffer of the current device. */
int crest_release(void);

/* Device memory for callers of crest_pipeline_run (no reference counterpart):
 * allocation on the current device, synchronous copies, and a wait for the
 * library's stream.                                                               */
int crest_dev_alloc(size_t bytes, void **ptr);
int crest_dev_free(void *ptr);
int crest_memcpy_htod(void *dst, const void *src, size_t bytes);
int crest_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int crest_synchronize(void);
/* Wait for the work queued on `stream` (a hipStream_t; NULL = the default stream) of
 * the current device, e.g. after crest_gram_accumulate_noalloc.                   */
int crest_stream_synchronize(void *stream);
/* Device-to-device copy by a plain vectorised kernel (k_copy16): the reference point
 * for calibrating HBM byte counters.  bytes must be a multiple of 16.            */
int crest_dev_copy(void *dst, const void *src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* CREST_HIP_H */
