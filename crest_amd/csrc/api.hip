// C ABI of the gfx950 library (include/crest_hip.h): argument validation with the
// reference's error classes (py.rs), a per-device context (stream, rocBLAS handle,
// grow-only device buffers) and the orchestration of the kernels.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/crest_hip.h"
#include "common.hpp"

namespace crest {

// ---------------------------------------------------------------- errors
static thread_local std::string g_err;
static thread_local int g_device = 0;
static std::atomic<int> g_prec{CREST_PREC_FP64};
static std::atomic<int> g_knn_prec{-1};  // -1: follow g_prec

struct Fail {
    int code;
    std::string msg;
};
[[noreturn]] static void fail(int code, const std::string &msg) { throw Fail{code, msg}; }
static void value_error(const std::string &msg) { fail(CREST_E_VALUE, msg); }

#define HIPCHK(x)                                                                                      \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) fail(CREST_E_DEVICE, std::string(#x ": ") + hipGetErrorString(e_));     \
    } while (0)
#define BLASCHK(x)                                                                                    \
    do {                                                                                              \
        rocblas_status s_ = (x);                                                                      \
        if (s_ != rocblas_status_success)                                                             \
            fail(CREST_E_DEVICE, std::string(#x ": ") + rocblas_status_to_string(s_));               \
    } while (0)

template <class F> static int guarded(F &&f) {
    try {
        f();
        g_err.clear();
        return CREST_OK;
    } catch (const Fail &e) {
        g_err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        g_err = e.what();
        return CREST_E_DEVICE;
    }
}

// ---------------------------------------------------------------- context
struct Ctx {
    std::mutex mu;
    int dev = 0;
    hipStream_t st = nullptr;
    rocblas_handle blas = nullptr;
    // certificate of the last kNN: queries checked, rechecked in FP32, left uncertified
    int64_t knn_checked = 0, knn_rechecked = 0, knn_uncertified = 0;
    std::map<std::string, std::pair<void *, size_t>> bufs;
    std::map<std::string, std::pair<void *, size_t>> pinned;  // host staging (hipHostMalloc)
    hipEvent_t ev[12] = {};
    hipStream_t copy = nullptr;        // host -> device uploads of streamed chunks
    // a second compute stream (default priority) with its own rocBLAS handle, for work
    // that overlaps the main stream's (the batch's n x n duals)
    hipStream_t side = nullptr;
    rocblas_handle blas_side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    // the Gram's tile pipeline: tile buffer b densified on the side stream, then read by
    // the Gram on the main stream
    hipEvent_t tile_ready[2] = {}, tile_free[2] = {};
    hipEvent_t copied[2] = {}, consumed[2] = {};
    // batch of subsets: the two captured graphs and the arguments they were captured for
    std::string graph_key;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    // autotuner: shape key -> chosen Gram row range
    std::map<std::string, int64_t> tuned;
    // staging of pushed chunks (sessions): two pinned host buffers and their device
    // twins, grow-only for the process; a slot is reused once its consumer finished
    struct StageSlot {
        uint8_t *pin = nullptr, *dev = nullptr;
        size_t pin_bytes = 0, dev_bytes = 0;
        hipEvent_t copied = nullptr, consumed = nullptr;
        bool used = false;
    } stage[2];
    int stage_next = 0;

    void init(int d) {
        dev = d;
        HIPCHK(hipSetDevice(dev));
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipStreamCreateWithFlags(&copy, hipStreamNonBlocking));
        BLASCHK(rocblas_create_handle(&blas));
        BLASCHK(rocblas_set_stream(blas, st));
        HIPCHK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        BLASCHK(rocblas_create_handle(&blas_side));
        BLASCHK(rocblas_set_stream(blas_side, side));
        HIPCHK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&join, hipEventDisableTiming));
        for (int i = 0; i < 2; ++i) {
            HIPCHK(hipEventCreateWithFlags(&tile_ready[i], hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&tile_free[i], hipEventDisableTiming));
        }
        for (auto &e : ev) HIPCHK(hipEventCreate(&e));
        for (int i = 0; i < 2; ++i) {
            HIPCHK(hipEventCreateWithFlags(&copied[i], hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&consumed[i], hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&stage[i].copied, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&stage[i].consumed, hipEventDisableTiming));
        }
    }
    // Pinned host buffer `name` of at least `bytes` (grow-only).
    void *host(const std::string &name, size_t bytes) {
        bytes = std::max<size_t>(bytes, 256);
        auto &b = pinned[name];
        if (b.second < bytes) {
            if (b.first) HIPCHK(hipHostFree(b.first));
            b.first = nullptr;
            b.second = 0;
            HIPCHK(hipHostMalloc(&b.first, bytes, hipHostMallocDefault));
            b.second = bytes;
        }
        return b.first;
    }
    // Device buffer `name` of at least `bytes` (grow-only; contents undefined).
    template <class T> T *get(const std::string &name, size_t count) {
        size_t bytes = std::max<size_t>(count * sizeof(T), 256);
        auto &b = bufs[name];
        if (b.second < bytes) {
            if (b.first) HIPCHK(hipFree(b.first));
            b.first = nullptr;
            b.second = 0;
            HIPCHK(hipMalloc(&b.first, bytes));
            b.second = bytes;
        }
        return static_cast<T *>(b.first);
    }
    template <class T> T *upload(const std::string &name, const T *host, size_t count) {
        T *d = get<T>(name, count);
        if (count) HIPCHK(hipMemcpyAsync(d, host, count * sizeof(T), hipMemcpyHostToDevice, st));
        return d;
    }
    template <class T> void download(T *host, const T *dev, size_t count) {
        if (count) HIPCHK(hipMemcpyAsync(host, dev, count * sizeof(T), hipMemcpyDeviceToHost, st));
    }
    void sync() { HIPCHK(hipStreamSynchronize(st)); }
    void check_launch() { HIPCHK(hipGetLastError()); }
    void drop_graphs() {
        for (auto &g : graph)
            if (g) {
                hipGraphExecDestroy(g);
                g = nullptr;
            }
        graph_key.clear();
    }
    void release() {
        HIPCHK(hipStreamSynchronize(copy));
        HIPCHK(hipStreamSynchronize(side));
        drop_graphs();
        for (auto &kv : bufs)
            if (kv.second.first) hipFree(kv.second.first);
        bufs.clear();
        for (auto &kv : pinned)
            if (kv.second.first) hipHostFree(kv.second.first);
        pinned.clear();
        for (auto &sl : stage) {
            if (sl.pin) hipHostFree(sl.pin);
            if (sl.dev) hipFree(sl.dev);
            sl.pin = sl.dev = nullptr;
            sl.pin_bytes = sl.dev_bytes = 0;
            sl.used = false;
        }
    }
};

static std::mutex g_ctx_mu;
static Ctx *g_ctx[64] = {};

static Ctx &ctx() {
    const int d = g_device;
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (d < 0 || d >= 64) value_error("device index out of range");
    if (!g_ctx[d]) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) fail(CREST_E_DEVICE, "no HIP device is visible");
        if (d >= n) value_error("device index out of range");
        Ctx *c = new Ctx();
        c->init(d);
        g_ctx[d] = c;
    }
    HIPCHK(hipSetDevice(d));
    return *g_ctx[d];
}

// ---------------------------------------------------------------- chunk validation
// py.rs:162-189 (view) and :191-193 (n_var).
struct HostChunk {
    const crest_chunk *c;
    std::vector<int64_t> span_off;  // only for per-entry cell ids
    std::vector<uint32_t> span_cell;
    int64_t n_spans = 0;
    int64_t base = 0;
    int64_t n_rows = 0;  // max(cell_map) + 1
    int64_t n_vars = 0;  // max(gene_map) + 1
};

// max(map) + 1 over a host array (0 when empty or all negative)
template <class T> static int64_t map_extent(const T *m, size_t n) {
    int64_t mx = -1;
    for (size_t i = 0; i < n; ++i) mx = std::max<int64_t>(mx, (int64_t)m[i]);
    return mx + 1 > 0 ? mx + 1 : 0;
}

// Geometry of a chunk (spans, base) without the map extents.
static HostChunk check_chunk_geometry(const crest_chunk *c) {
    if (!c) value_error("chunk is NULL");
    HostChunk h;
    h.c = c;
    if ((c->nnz && (!c->genes || !c->values)) || (c->n_cell_map && !c->cell_map) || (c->n_gene_map && !c->gene_map))
        value_error("chunk arrays must not be NULL");
    if ((c->cells != nullptr) == (c->indptr != nullptr)) value_error("pass exactly one of cells / indptr");
    if (c->cells) {
        for (size_t i = 1; i < c->nnz; ++i)
            if (c->cells[i - 1] > c->cells[i]) value_error("chunk must be sorted by cell");
        size_t p = 0;
        while (p < c->nnz) {
            size_t q = p + 1;
            while (q < c->nnz && c->cells[q] == c->cells[p]) ++q;
            h.span_off.push_back((int64_t)p);
            h.span_cell.push_back(c->cells[p]);
            p = q;
        }
        h.span_off.push_back((int64_t)c->nnz);
        h.n_spans = (int64_t)h.span_cell.size();
        h.base = 0;
    } else {
        const int64_t *p = c->indptr;
        const size_t m = c->n_indptr;
        bool ok = m > 0;
        for (size_t i = 1; ok && i < m; ++i) ok = p[i - 1] <= p[i];
        if (ok) ok = (uint64_t)(p[m - 1] - p[0]) == (uint64_t)c->nnz;
        if (!ok) value_error("indptr must be non-decreasing and span the chunk");
        h.n_spans = (int64_t)m - 1;
        h.base = p[0];
    }
    return h;
}

static HostChunk check_chunk(const crest_chunk *c) {
    HostChunk h = check_chunk_geometry(c);
    h.n_rows = map_extent(c->cell_map, c->n_cell_map);
    h.n_vars = map_extent(c->gene_map, c->n_gene_map);
    return h;
}

// Row of each span on the host (for the rows a chunk touches).
static int64_t host_row(const HostChunk &h, int64_t i) {
    const crest_chunk *c = h.c;
    const uint64_t cell = c->cells ? (uint64_t)h.span_cell[i] : (uint64_t)c->first_cell + (uint64_t)i;
    if (cell >= c->n_cell_map) return -1;
    return c->cell_map[cell];
}

static DevChunk upload_chunk(Ctx &x, const HostChunk &h) {
    const crest_chunk *c = h.c;
    DevChunk d{};
    d.genes = x.upload("genes", c->genes, c->nnz);
    d.values = x.upload("values", c->values, c->nnz);
    d.cell_map = x.upload("cell_map", c->cell_map, c->n_cell_map);
    d.n_cell_map = (int64_t)c->n_cell_map;
    d.gene_var = x.upload("gene_map", c->gene_map, c->n_gene_map);
    d.n_gene_map = (int64_t)c->n_gene_map;
    d.target_sum = c->target_sum;
    d.log1p = c->log1p ? 1 : 0;
    d.n_spans = h.n_spans;
    d.base = h.base;
    d.nnz = (int64_t)c->nnz;
    if (c->cells) {
        d.span_off = x.upload("span_off", h.span_off.data(), h.span_off.size());
        d.span_cell = x.upload("span_cell", h.span_cell.data(), h.span_cell.size());
        d.first_cell = 0;
    } else {
        d.span_off = x.upload("span_off", c->indptr, c->n_indptr);
        d.span_cell = nullptr;
        d.first_cell = c->first_cell;
    }
    return d;
}

// Rows [r0, r1) of a device chunk (entry offsets stay absolute).
static DevChunk sub_chunk(const DevChunk &c, int64_t r0, int64_t r1) {
    DevChunk s = c;
    s.span_off = c.span_off + r0;
    if (c.span_cell)
        s.span_cell = c.span_cell + r0;
    else
        s.first_cell = c.first_cell + (uint32_t)r0;
    s.n_spans = r1 - r0;
    return s;
}

// py.rs:263-268
static void check_sub(const int32_t *sub, size_t n_sub, int64_t nv, size_t d) {
    if ((int64_t)n_sub != nv) value_error("sub_map must have length n_vars and values < d");
    for (size_t i = 0; i < n_sub; ++i)
        if (sub[i] >= (int64_t)d) value_error("sub_map must have length n_vars and values < d");
}
// py.rs:250-261
static void check_scale(int32_t scale, const double *mean, size_t n_mean, const double *sd, size_t n_sd, size_t d) {
    if (!scale) return;
    if (!mean || !sd) value_error("scale=True needs mean and std");
    if (n_mean != d || n_sd != d) value_error("mean/std must have length d");
}

static int tile_fmt(int prec) {
    return prec == CREST_PREC_FP32 ? TILE_F32 : prec == CREST_PREC_BF16X3 ? TILE_BF16X2 : prec == CREST_PREC_FP16 ? TILE_F16X2 : TILE_F64;
}
// Split-FP16 tiles hold |s| <= 65504: guaranteed when the scaled values are clipped,
// |s| = |clip(z) - clip(-mu/sd)| <= 2 max_value.
static void check_gram_prec(int prec, int scale, double max_value) {
    if (prec == CREST_PREC_FP16 && !(scale && max_value > 0.0 && max_value <= 30000.0))
        value_error("the FP16 Gram needs scaled values clipped at 0 < max_value <= 30000");
}
static size_t tile_elem_bytes(int fmt) { return fmt == TILE_F64 ? 8 : fmt == TILE_F32 ? 4 : 2; }

static int choose_split(int nblk, int64_t rows_pad, int fmt) {
    const int64_t ksteps = rows_pad / 16;
    // about 1,024 workgroups for the LDS-staged split-16-bit Gram (8 splits of the 136
    // blocks at d = 2,000: 22.4 ms for the 1M-cell Gram stage, against 23.6 with 16
    // splits and 24.8 with 4), about 2,048 for the other forms
    const bool x16 = (fmt == TILE_F16X2 || fmt == TILE_BF16X2) && gram_dma_enabled();
    const int target = x16 ? 1024 : 2048;
    int s = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(target, nblk), ksteps / 8));
    if (s >= 8) s = s / 8 * 8;
    return std::max(s, 1);
}

// Gram + colsum of the rows of a device chunk, accumulated into dev_gram/dev_colsum.
// With more than one tile, the tiles go through two buffers: tile t+1 is densified on
// the side stream while the Gram of tile t runs on the main stream (densify is bound by
// its tile writes, the Gram by L2 and MFMA, so the two share the CUs well).  Densify's
// exact lo^2 diagonal then accumulates into a buffer of its own, added to the Gram's
// diagonal at the end, since the split-K reduce adds into dev_gram without atomics.
// The FP64 Gram by sparse pair products (kernels_sparse.hip): the default for
// CREST_PREC_FP64 up to 4,096 HVGs; CREST_GRAM_SPARSE=0 selects the dense FP64 MFMA form.
static bool gram_sparse_enabled(int prec, int d_pad) {
    if (prec != CREST_PREC_FP64 || !sparse_gram_supported(d_pad)) return false;
    const char *e = std::getenv("CREST_GRAM_SPARSE");
    return !(e && std::string(e) == "0");
}

static void gram_rows_sparse(Ctx &x, const DevChunk &c, const DevScale &sc, int d_pad, double *dev_gram,
                             double *dev_colsum) {
    if (c.n_spans <= 0) return;
    const int nb = d_pad / kGramBlock, nblk = nb * (nb + 1) / 2;
    const int n_split = 16;
    const size_t ne = (size_t)std::max<int64_t>(c.nnz, 1);
    double *vals = x.get<double>("sp_vals", ne);
    uint8_t *cols = x.get<uint8_t>("sp_cols", ne);
    uint32_t *meta = x.get<uint32_t>("sp_meta", sparse_meta_words(c.n_spans, d_pad));
    double *partial = x.get<double>("gram_partial", (size_t)nblk * n_split * kGramBlock * kGramBlock);
    double *cs_partial = x.get<double>("sp_cs_partial", (size_t)nb * n_split * kGramBlock);
    const int64_t n_groups = ceil_div(c.n_spans, kRowGroupSparse);
    const int half = n_split / 2;
    const int64_t g_half = (int64_t)half * n_groups / n_split;  // the groups of splits 0 .. half-1
    if (g_half > 0 && g_half < n_groups) {
        // the rows of the second half are packed on the side stream while the pair
        // products of the first half run (the pack waits on its loads, the products
        // on issue and LDS atomics, so they share the CUs); each half's packed entries
        // sit at their own CSR offsets and its group records at its groups' places
        const int64_t r_half = g_half * kRowGroupSparse;
        launch_sparse_pack(seg_rows(c, 0, r_half), sc, d_pad, vals, cols, meta, x.st);
        HIPCHK(hipEventRecord(x.fork, x.st));
        HIPCHK(hipStreamWaitEvent(x.side, x.fork, 0));
        launch_sparse_pack(seg_rows(c, r_half, c.n_spans), sc, d_pad, vals, cols,
                           meta + sparse_meta_words(r_half, d_pad), x.side);
        HIPCHK(hipEventRecord(x.join, x.side));
        launch_gram_sparse_part(vals, cols, meta, c.span_off, c.base, c.n_spans, d_pad, n_split, 0, half, partial,
                                cs_partial, x.st);
        HIPCHK(hipStreamWaitEvent(x.st, x.join, 0));
        launch_gram_sparse_part(vals, cols, meta, c.span_off, c.base, c.n_spans, d_pad, n_split, half, n_split - half,
                                partial, cs_partial, x.st);
        launch_sparse_colsum(cs_partial, d_pad, n_split, dev_colsum, x.st);
    } else {
        launch_sparse_pack(c, sc, d_pad, vals, cols, meta, x.st);
        launch_gram_sparse(vals, cols, meta, c.span_off, c.base, c.n_spans, d_pad, n_split, partial, cs_partial,
                           dev_colsum, x.st);
    }
    launch_gram_reduce(partial, n_split, d_pad, dev_gram, x.st);
    x.check_launch();
}

static void gram_rows(Ctx &x, const DevChunk &c, const DevScale &sc, int prec, int d_pad, int64_t tile_rows,
                      double *dev_gram, double *dev_colsum) {
    if (gram_sparse_enabled(prec, d_pad)) return gram_rows_sparse(x, c, sc, d_pad, dev_gram, dev_colsum);
    const int fmt = tile_fmt(prec);
    const int nb = d_pad / kGramBlock, nblk = nb * (nb + 1) / 2;
    const int64_t n_tiles = ceil_div(c.n_spans, tile_rows);
    const bool pipe = n_tiles > 1;
    double *gdiag = dev_gram;
    if (pipe) {
        gdiag = x.get<double>("gram_diag_acc", (size_t)d_pad * d_pad);
        launch_fill_f64(gdiag, (int64_t)d_pad * d_pad, 0.0, x.st);
        HIPCHK(hipEventRecord(x.fork, x.st));
        HIPCHK(hipStreamWaitEvent(x.side, x.fork, 0));
    }
    for (int64_t t = 0; t < n_tiles; ++t) {
        const int64_t r0 = t * tile_rows, r1 = std::min<int64_t>(c.n_spans, r0 + tile_rows);
        const int64_t rows_pad = round_up(r1 - r0, 16);
        const size_t tile_elems = (size_t)rows_pad * d_pad;
        const int b = pipe ? (int)(t & 1) : 0;
        const std::string sfx = b ? "1" : "";
        void *ta = x.get<uint8_t>("tile_a" + sfx, tile_elems * tile_elem_bytes(fmt));
        void *tb = (fmt == TILE_BF16X2 || fmt == TILE_F16X2) ? x.get<uint8_t>("tile_b" + sfx, tile_elems * 2)
                                                             : nullptr;
        const int n_split = choose_split(nblk, rows_pad, fmt);
        double *partial = x.get<double>("gram_partial", (size_t)nblk * n_split * kGramBlock * kGramBlock);
        if (pipe) {
            if (t >= 2) HIPCHK(hipStreamWaitEvent(x.side, x.tile_free[b], 0));  // the Gram of tile t-2 is done
            launch_densify(sub_chunk(c, r0, r1), sc, fmt, ta, tb, rows_pad, d_pad, dev_colsum, gdiag, x.side);
            HIPCHK(hipEventRecord(x.tile_ready[b], x.side));
            HIPCHK(hipStreamWaitEvent(x.st, x.tile_ready[b], 0));
        } else {
            launch_densify(sub_chunk(c, r0, r1), sc, fmt, ta, tb, rows_pad, d_pad, dev_colsum, dev_gram, x.st);
        }
        launch_gram(fmt, ta, tb, rows_pad, d_pad, n_split, partial, x.st);
        launch_gram_reduce(partial, n_split, d_pad, dev_gram, x.st);
        if (pipe) HIPCHK(hipEventRecord(x.tile_free[b], x.st));
        x.check_launch();
    }
    if (pipe) launch_add_diag(gdiag, d_pad, dev_gram, x.st);  // after the last Gram, so after every densify
}

// Eigen-decomposition of the covariance from (gram [ld_g], colsum, n): variance and
// loadings (device), total (device scalar).
// The eigensolver.  "topp" (default): rocSOLVER's Householder tridiagonalization
// (dsytrd) and the top-p kernels of kernels_eig.hip (bisection, inverse iteration,
// back-transform of the p vectors); "dsyevd": rocSOLVER's full divide-and-conquer
// eigensolve (measured at d = 2,000, p = 50: 49 vs 51 ms for one matrix, 6.7 vs 10.2
// ms per matrix for 256 side by side).  CREST_EIGENSOLVER=dsyevd selects dsyevd; it
// also runs for orders beyond the top-p kernels' limits and for any matrix whose
// inverse iteration did not converge.
static bool use_topp(int d, int k, int n_mat) {
    (void)n_mat;
    if (!(k >= 1 && k <= d && tri_topp_supported(d))) return false;
    const char *e = std::getenv("CREST_EIGENSOLVER");
    return !(e && std::string(e) == "dsyevd");
}

// Block size of the filtered subspace iteration for the top p of order d: p + 16
// rounded up to a multiple of 8, at most chefsi_max_block() and d / 2; 0 when it does
// not apply.  It runs by default from order 512 (below that the tridiagonal path takes
// a few milliseconds); CREST_EIGENSOLVER=chefsi forces it wherever the block fits,
// =topp or =dsyevd switches it off.
static int chefsi_block(int d, int p) {
    const int k = (int)round_up(p + 16, 8);
    if (p < 1 || k > chefsi_max_block() || 2 * k > d) return 0;
    const char *e = std::getenv("CREST_EIGENSOLVER");
    const std::string mode = e ? e : "";
    if (mode == "chefsi") return k;
    return (mode.empty() && d >= 512) ? k : 0;
}

constexpr int kChebMaxDegree = 64;  // longest filter per iteration

// Degree of the next filter: the largest m <= kChebMaxDegree whose T_m at the top of the active
// block (x_top in the filter's variable) stays below kappa, so the filtered block
// is no worse conditioned than kappa from the spread of its directions' gains.
static int chefsi_degree(double x_top, double kappa) {
    if (!(x_top > 1.0 + 1e-12)) return kChebMaxDegree;
    const int m = (int)std::floor(std::acosh(kappa) / std::acosh(x_top));
    return std::max(1, std::min(kChebMaxDegree, m));
}

// The degree the wanted Ritz vectors still unlocked need: the filter multiplies a
// vector of value theta_j by T_m(x_j) (x_j = (theta_j - c) / e) against everything in
// the damped interval, so its residual falls by about that factor; the degree that
// brings every res_j to `target` (a hundredth of the locking threshold, so one more
// Rayleigh-Ritz locks them), or the longest filter when some x_j gives no bound.  Locking still tests
// the residuals themselves, so a degree that falls short costs one more iteration.
static int chefsi_needed(const double *theta, const double *res, int n_want, double c, double e, double target) {
    int m = 1;
    for (int j = 0; j < n_want; ++j) {
        const double x = (theta[j] - c) / e;
        if (!(x > 1.0 + 1e-12)) return kChebMaxDegree;
        const double need = res[j] / target;
        if (need > 1.0) m = std::max(m, (int)std::ceil(std::acosh(need) / std::acosh(x)));
    }
    return std::min(m, kChebMaxDegree);
}

// The p largest eigenpairs of the symmetric d x d matrix C (column-major, ld d) by
// Chebyshev-filtered subspace iteration, into dev_var / dev_load with the sign rule
// of launch_eig_select.  The work happens in a copy A of C padded with zeros to
// d_pad = d rounded up to 256 (the padding rows of every block stay zero).  Per
// iteration: filter the active block with T_m((A - c) / e), where [0, b] (b = the
// block's smallest Ritz value) is the damped interval; remove the locked directions;
// CholeskyQR2 with the columns in descending Ritz order (so each column is
// orthogonalised against those above it, whose contamination the filter amplified);
// Rayleigh-Ritz against A, the small eigenproblem by one-workgroup Jacobi limited to
// two sweeps (the basis it leaves is still orthonormal and spans the same subspace;
// the rotation it misses shows up in the residuals and is finished in a later
// iteration); residual norms.  Ritz pairs converged to 1e-12 of the largest Ritz value
// so far are locked from the top down and deflated out of A (A -= q theta q^T), which
// keeps the filter degree up as the top of the spectrum leaves.  A final
// Rayleigh-Ritz against C over the locked vectors, with Jacobi to convergence, gives
// the result.  Products with A run on k_ch_gemm split over K (kSplit partials summed
// by the kernel that consumes them), the Gram-like products as rocBLAS dgemm split the
// same way (their partials summed by a grid-wide kernel before the one-workgroup
// kernels read them).  Returns false -- with dev_var / dev_load unwritten -- when the block loses
// rank, the final Jacobi solve does not converge, the spectrum is too flat to filter
// (b below 1e-8 of the top) or 60 iterations do not converge; the caller then takes
// the tridiagonal path.  C is left as it was.
static bool chefsi_topp(Ctx &x, const double *C, int d, int p, double *dev_var, double *dev_load) {
    const int k = chefsi_block(d, p);
    if (!k) return false;
    const int dp = (int)round_up(d, 256);
    const int kSplit = 8;                       // K chunks of the d_pad x d_pad products (multiples of 32)
    const int kSkinny = std::min(32, dp / 64);  // K chunks of the k x k Gram products
    const int kSweeps = 2;                      // Jacobi sweeps of an intermediate Rayleigh-Ritz
    const int64_t dk = (int64_t)dp * k;
    double *A = x.get<double>("ch_a", (size_t)dp * dp);
    double *buf[4];
    for (int i = 0; i < 4; ++i) buf[i] = x.get<double>("ch_b" + std::to_string(i), (size_t)dk);
    double *W = x.get<double>("ch_w", (size_t)dk);
    double *WZ = x.get<double>("ch_wz", (size_t)dk);
    double *P = x.get<double>("ch_part", (size_t)kSplit * dk);         // partials of A Y
    double *Ps = x.get<double>("ch_part_s", (size_t)kSkinny * k * k);  // partials of Y^T Y, Y^T W, QL^T Y
    double *G = x.get<double>("ch_g", (size_t)k * k);  // summed Gram-like product
    double *Ri = x.get<double>("ch_ri", (size_t)k * k);
    double *Z = x.get<double>("ch_z", (size_t)k * k);
    double *QL = x.get<double>("ch_ql", (size_t)dp * p);
    double *thL = x.get<double>("ch_thl", p);
    double *T = x.get<double>("ch_t", (size_t)p * k);
    double *Vout = x.get<double>("ch_vout", (size_t)d * p);
    // stat: Ritz values [k] | residual norms [k] | block lost rank | Jacobi failed
    double *stat = x.get<double>("ch_stat", (size_t)2 * k + 2);
    double *hs = static_cast<double *>(x.host("ch_stat", (2 * (size_t)k + 2) * sizeof(double)));
    const bool trace = std::getenv("CREST_CHEFSI_TRACE") != nullptr;
    chefsi_set_attrs();
    const auto N = rocblas_operation_none, Tr = rocblas_operation_transpose;
    auto gemm = [&](rocblas_operation ta, rocblas_operation tb, int m, int n, int kk, double alpha, const double *a,
                    int lda, const double *b, int ldb, double beta, double *c, int ldc) {
        BLASCHK(rocblas_dgemm(x.blas, ta, tb, m, n, kk, &alpha, a, lda, b, ldb, &beta, c, ldc));
    };
    // partials of a^T b for a (dp x m), b (dp x n): Ps + s m n, s < kSkinny
    auto inner = [&](const double *a, int m, const double *b, int n) {
        const double one = 1.0, zero = 0.0;
        const int kc = dp / kSkinny;
        BLASCHK(rocblas_dgemm_strided_batched(x.blas, Tr, N, m, n, kc, &one, a, dp, kc, b, dp, kc, &zero, Ps, m,
                                              (int64_t)m * n, kSkinny));
    };
    auto other = [&](std::initializer_list<int> used) {
        for (int i = 0; i < 4; ++i)
            if (std::find(used.begin(), used.end(), i) == used.end()) return i;
        return -1;
    };
    // CholeskyQR2 of buf[y] (dp x ka), leaving buf[keep] alone; returns the buffer that
    // holds the result
    auto cholqr2 = [&](int y, int ka, int keep) {
        for (int pass = 0; pass < 2; ++pass) {
            const int o = other({y, keep});
            inner(buf[y], ka, buf[y], ka);
            launch_ch_sum(G, Ps, kSkinny, (int64_t)ka * ka, (int64_t)ka * ka, x.st);
            launch_ch_chol_rinv(G, 1, ka, Ri, stat + 2 * k, x.st);
            gemm(N, N, dp, ka, ka, 1.0, buf[y], dp, Ri, ka, 0.0, buf[o], dp);
            y = o;
        }
        return y;
    };
    // Rayleigh-Ritz of the orthonormal buf[y] (dp x ka) against op: Ritz vectors into
    // the returned buffer (not buf[keep]; descending unless final), op times them into
    // WZ, values and residual norms into stat, then downloaded into hs
    auto rayleigh_ritz = [&](const double *op, int y, int ka, int keep, bool final) {
        launch_ch_gemm(op, buf[y], dp, ka, kSplit, P, x.st);
        launch_ch_sum(W, P, kSplit, (int64_t)dp * ka, (int64_t)dp * ka, x.st);
        inner(buf[y], ka, W, ka);
        launch_ch_sum(G, Ps, kSkinny, (int64_t)ka * ka, (int64_t)ka * ka, x.st);
        launch_ch_jacobi(G, 1, ka, final ? 40 : kSweeps, final, !final, stat, Z, stat + 2 * k + 1, x.st);
        const int v = other({y, keep});
        gemm(N, N, dp, ka, ka, 1.0, buf[y], dp, Z, ka, 0.0, buf[v], dp);
        gemm(N, N, dp, ka, ka, 1.0, W, dp, Z, ka, 0.0, WZ, dp);
        launch_ch_resid(WZ, buf[v], stat, d, dp, ka, stat + k, x.st);
        x.check_launch();
        x.download(hs, stat, 2 * (size_t)k + 2);
        x.sync();
        return v;
    };
    auto give_up = [&](const char *why) {
        if (trace) std::fprintf(stderr, "[chefsi] fallback: %s\n", why);
        return false;
    };
    auto finite = [&](int ka) {
        for (int j = 0; j < ka; ++j)
            if (!std::isfinite(hs[j]) || !std::isfinite(hs[k + j])) return false;
        return true;
    };
    auto load_padded = [&](const double *src) {
        launch_fill_f64(A, (int64_t)dp * dp, 0.0, x.st);
        HIPCHK(hipMemcpy2DAsync(A, (size_t)dp * sizeof(double), src, (size_t)d * sizeof(double),
                                (size_t)d * sizeof(double), d, hipMemcpyDeviceToDevice, x.st));
    };

    load_padded(C);
    launch_fill_f64(stat, 2 * (int64_t)k + 2, 0.0, x.st);
    launch_ch_rand(buf[0], d, dp, k, 0x5eed0f5ca1edull + (uint64_t)d * 1000003u + (uint64_t)p, x.st);
    int ka = k, L = 0, iters = 0, products = 1;
    int v = rayleigh_ritz(A, cholqr2(0, ka, -1), ka, -1, false);
    if (hs[2 * k] != 0.0 || !finite(ka) || !(hs[0] > 0.0)) return give_up("start block");
    // the active block: ka Ritz vectors (descending) at vcol, A times them at wzcol, their
    // values and residual norms at hs[j0 ..] and hs[k + j0 ..]
    const double *vcol = buf[v], *wzcol = WZ;
    int j0 = 0;
    double top = 0.0, tol = 0.0;  // the largest Ritz value so far; the locking threshold
    std::vector<double> prev(2 * (size_t)k);
    double kappa = 1e3;  // the first filter acts on a random block: keep its gains close
    bool done = false;
    for (int it = 0, retries = 0; it < 60; ++it) {
        top = std::max(top, hs[0]);
        tol = 1e-12 * top;
        int J = 0;
        while (J < p - L && hs[k + J] <= tol) ++J;
        if (J) {  // lock the J largest (descending into QL) and deflate them out of A
            launch_ch_lock(vcol, stat, dp, J, QL + (int64_t)L * dp, thL + L, W, x.st);
            gemm(N, Tr, dp, dp, J, -1.0, W, dp, QL + (int64_t)L * dp, dp, 1.0, A, dp);
            L += J;
            ka -= J;
        }
        j0 = J;
        if (L >= p) {
            done = true;
            break;
        }
        vcol += (int64_t)J * dp;
        wzcol += (int64_t)J * dp;
        std::copy(hs, hs + 2 * k, prev.begin());
        const double a = 0.0, b = prev[j0 + ka - 1], theta_top = prev[j0];
        if (!(b > 1e-8 * top)) return give_up("flat spectrum");
        const double c = 0.5 * (a + b), e = 0.5 * (b - a);
        const int64_t n_el = (int64_t)dp * ka;
        for (;;) {
            const int m = std::min(chefsi_degree((theta_top - c) / e, kappa),
                                   chefsi_needed(&prev[j0], &prev[k + j0], std::min(p - L, ka), c, e, 1e-2 * tol));
            if (trace)
                std::fprintf(stderr, "[chefsi] it %d locked %d active %d degree %d top-resid %.3e b/top %.6e\n", it, L,
                             ka, m, prev[k + j0] / top, b / top);
            // y1 = (A y0 - c y0) / e with A y0 = wzcol; y_{j+1} = 2 (A y_j - c y_j) / e - y_{j-1};
            // the Ritz vectors (buf[v]) stay for a retry
            int y1 = other({v});
            const double *y0p = vcol;
            launch_ch_combine(buf[y1], wzcol, 1, 0, vcol, nullptr, n_el, 1.0 / e, c, 0.0, x.st);
            int y0 = -1;  // buffer of y0 once it is no longer the Ritz block
            for (int j = 1; j < m; ++j) {
                launch_ch_gemm(A, buf[y1], dp, ka, kSplit, P, x.st);
                const int y2 = other({y0, y1, v});
                launch_ch_combine(buf[y2], P, kSplit, n_el, buf[y1], y0p, n_el, 2.0 / e, c, 1.0, x.st);
                y0 = y1;
                y0p = buf[y1];
                y1 = y2;
            }
            if (L) {  // y -= QL (QL^T y)
                inner(QL, L, buf[y1], ka);
                launch_ch_sum(T, Ps, kSkinny, (int64_t)L * ka, (int64_t)L * ka, x.st);
                gemm(N, N, dp, ka, L, -1.0, QL, dp, T, L, 1.0, buf[y1], dp);
            }
            x.check_launch();
            const int nv = rayleigh_ritz(A, cholqr2(y1, ka, v), ka, v, false);
            ++iters;
            products += m;
            if (hs[2 * k] == 0.0 && finite(ka)) {
                v = nv;
                vcol = buf[v];
                wzcol = WZ;
                kappa = std::min(kappa * 100.0, 1e8);
                break;
            }
            // the filtered block lost rank: filter the same Ritz vectors with a lower degree
            if (++retries > 8 || kappa < 10.0) return give_up("rank loss");
            kappa = std::sqrt(kappa);
            launch_fill_f64(stat + 2 * k, 1, 0.0, x.st);
            launch_ch_gemm(A, vcol, dp, ka, kSplit, P, x.st);
            launch_ch_sum(WZ, P, kSplit, n_el, n_el, x.st);
            wzcol = WZ;
        }
    }
    if (!done) return give_up("no convergence");
    // final Rayleigh-Ritz of the locked vectors against C
    load_padded(C);
    HIPCHK(hipMemcpyAsync(buf[0], QL, (size_t)dp * p * sizeof(double), hipMemcpyDeviceToDevice, x.st));
    v = rayleigh_ritz(A, cholqr2(0, p, -1), p, -1, true);
    if (hs[2 * k] != 0.0 || hs[2 * k + 1] != 0.0 || !finite(p)) return give_up("final Rayleigh-Ritz");
    for (int j = 0; j < p; ++j)
        if (!(hs[k + j] <= 100.0 * tol)) return give_up("final residual");
    if (trace) std::fprintf(stderr, "[chefsi] converged: %d iterations, %d products with A\n", iters, products);
    HIPCHK(hipMemcpy2DAsync(Vout, (size_t)d * sizeof(double), buf[v], (size_t)dp * sizeof(double),
                            (size_t)d * sizeof(double), p, hipMemcpyDeviceToDevice, x.st));
    launch_eig_select(Vout, stat, d, p, dev_var, dev_load, x.st, p);
    x.check_launch();
    return true;
}

// The iteration of chefsi_topp for nm matrices of order d side by side (the covariances
// C, d x d column-major, d^2 apart): the same start block, filter, CholeskyQR2,
// Rayleigh-Ritz, locking and final solve per matrix, with every product of all the
// matrices in one launch (the big products on k_chb_gemm, the skinny ones as strided
// batched dgemm), one host round trip per iteration for all of them, and per-matrix
// decisions (degree, locking, retries) as in chefsi_topp.  A matrix that would make
// chefsi_topp give up -- rank loss after its retries, a flat spectrum, no convergence
// in 60 iterations, a failed final solve -- drops out of the batch; solved[z] tells
// which matrices got their variance and loadings (subset ids[z] of var / load).
// Buffers are named with `pre`.  C is left as it was.
static std::vector<char> chefsi_topp_batch(Ctx &x, const std::string &pre, const double *C, int nm, int d, int p,
                                           const int32_t *ids_dev, double *var, double *load) {
    std::vector<char> solved(nm, 0);
    const int K = chefsi_block(d, p);
    if (!K || nm <= 0) return solved;
    const int dp = (int)round_up(d, 256);
    const int64_t dk = (int64_t)dp * K, dd = (int64_t)dp * dp, kk = (int64_t)K * K, ns = 2 * (int64_t)K + 2;
    double *A = x.get<double>(pre + "chb_a", (size_t)nm * dd);
    double *buf[4];
    for (int i = 0; i < 4; ++i) buf[i] = x.get<double>(pre + "chb_b" + std::to_string(i), (size_t)nm * dk);
    double *W = x.get<double>(pre + "chb_w", (size_t)nm * dk);
    double *WZ = x.get<double>(pre + "chb_wz", (size_t)nm * dk);
    double *Wl = x.get<double>(pre + "chb_wl", (size_t)nm * dk);
    double *Qn = x.get<double>(pre + "chb_qn", (size_t)nm * dk);
    double *QL = x.get<double>(pre + "chb_ql", (size_t)nm * dp * p);
    double *thL = x.get<double>(pre + "chb_thl", (size_t)nm * p);
    double *T = x.get<double>(pre + "chb_t", (size_t)nm * p * K);
    double *G = x.get<double>(pre + "chb_g", (size_t)nm * kk);
    double *Ri = x.get<double>(pre + "chb_ri", (size_t)nm * kk);
    double *Z = x.get<double>(pre + "chb_z", (size_t)nm * kk);
    double *stat = x.get<double>(pre + "chb_stat", (size_t)nm * ns);
    double *Vout = x.get<double>(pre + "chb_vout", (size_t)nm * d * p);
    double *evals = x.get<double>(pre + "chb_evals", (size_t)nm * p);
    ChbMat *Pd = x.get<ChbMat>(pre + "chb_p", (size_t)nm);
    ChbMat *Ph = static_cast<ChbMat *>(x.host(pre + "chb_p", (size_t)nm * sizeof(ChbMat)));
    double *hs = static_cast<double *>(x.host(pre + "chb_stat", (size_t)nm * ns * sizeof(double)));
    const bool trace = std::getenv("CREST_CHEFSI_TRACE") != nullptr;
    chefsi_batch_set_attrs();
    const auto N = rocblas_operation_none, Tr = rocblas_operation_transpose;
    auto bgemm = [&](rocblas_operation ta, rocblas_operation tb, int m, int n, int kq, double alpha, const double *a,
                     int lda, int64_t sa, const double *b, int ldb, int64_t sb, double beta, double *c, int ldc,
                     int64_t sc) {
        BLASCHK(rocblas_dgemm_strided_batched(x.blas, ta, tb, m, n, kq, &alpha, a, lda, sa, b, ldb, sb, &beta, c, ldc,
                                              sc, nm));
    };
    auto other = [&](std::initializer_list<int> used) {
        for (int i = 0; i < 4; ++i)
            if (std::find(used.begin(), used.end(), i) == used.end()) return i;
        return -1;
    };
    auto upload = [&]() { HIPCHK(hipMemcpyAsync(Pd, Ph, (size_t)nm * sizeof(ChbMat), hipMemcpyHostToDevice, x.st)); };
    auto cholqr2 = [&](int y, int keep) {
        for (int pass = 0; pass < 2; ++pass) {
            const int o = other({y, keep});
            bgemm(Tr, N, K, K, dp, 1.0, buf[y], dp, dk, buf[y], dp, dk, 0.0, G, K, kk);
            launch_chb_chol_rinv(G, K, nm, Pd, Ri, stat, x.st);
            bgemm(N, N, dp, K, K, 1.0, buf[y], dp, dk, Ri, K, kk, 0.0, buf[o], dp, dk);
            y = o;
        }
        return y;
    };
    auto rayleigh_ritz = [&](int y, int keep, bool final) {
        launch_chb_gemm(A, buf[y], dp, K, nm, Pd, -1, false, W, x.st);
        bgemm(Tr, N, K, K, dp, 1.0, buf[y], dp, dk, W, dp, dk, 0.0, G, K, kk);
        launch_chb_jacobi(G, K, nm, Pd, final ? 40 : 2, final, !final, Z, stat, x.st);
        const int v = other({y, keep});
        bgemm(N, N, dp, K, K, 1.0, buf[y], dp, dk, Z, K, kk, 0.0, buf[v], dp, dk);
        bgemm(N, N, dp, K, K, 1.0, W, dp, dk, Z, K, kk, 0.0, WZ, dp, dk);
        launch_chb_resid(WZ, buf[v], d, dp, K, nm, Pd, stat, x.st);
        x.check_launch();
        x.download(hs, stat, (size_t)nm * ns);
        x.sync();
        return v;
    };
    auto finite = [&](int z, int ka) {
        const double *h = hs + z * ns;
        for (int j = 0; j < ka; ++j)
            if (!std::isfinite(h[j]) || !std::isfinite(h[K + j])) return false;
        return true;
    };
    enum { ACTIVE, DONE, OUT };
    struct Mat {
        int state = ACTIVE, ka = 0, L = 0, j0 = 0, retries = 0;
        bool retry = false;
        double top = 0.0, tol = 0.0, kappa = 1e3;
        std::vector<double> prev;
    };
    std::vector<Mat> M(nm);
    for (auto &m : M) {
        m.ka = K;
        m.prev.assign(2 * K, 0.0);
    }
    // start: every matrix's padded operator, the start block, zeroed buffers
    for (int z = 0; z < nm; ++z) Ph[z] = ChbMat{K, 0, 0, 0, 0, 0, 0.0, 1.0};
    upload();
    launch_chb_pad(C, d, dp, nm, Pd, A, x.st);
    for (double *b : {buf[1], buf[2], buf[3], W, WZ}) launch_fill_f64(b, (int64_t)nm * dk, 0.0, x.st);
    launch_fill_f64(QL, (int64_t)nm * dp * p, 0.0, x.st);
    launch_fill_f64(stat, (int64_t)nm * ns, 0.0, x.st);
    launch_chb_rand(buf[0], d, dp, K, nm, 0x5eed0f5ca1edull + (uint64_t)d * 1000003u + (uint64_t)p, x.st);
    int v = rayleigh_ritz(cholqr2(0, -1), -1, false);
    for (int z = 0; z < nm; ++z)
        if (hs[z * ns + 2 * K] != 0.0 || !finite(z, K) || !(hs[z * ns] > 0.0)) M[z].state = OUT;
    int iters = 0;
    for (int it = 0; it < 60; ++it) {
        // per matrix: locking from the last Rayleigh-Ritz, then the next filter
        bool any_lock = false, any_active = false, any_locked = false;
        int deg = 0;
        for (int z = 0; z < nm; ++z) {
            Mat &m = M[z];
            ChbMat &q = Ph[z];
            q = ChbMat{0, 0, 0, m.L, 0, 0, 0.0, 1.0};
            if (m.state != ACTIVE) continue;
            const double *h = hs + z * ns;
            if (!m.retry) {
                m.top = std::max(m.top, h[0]);
                m.tol = 1e-12 * m.top;
                int J = 0;
                while (J < p - m.L && h[K + J] <= m.tol) ++J;
                q.J = J;
                any_lock = any_lock || J > 0;
                m.L += J;
                m.ka -= J;
                m.j0 = J;
                std::copy(h, h + 2 * K, m.prev.begin());
                if (m.L >= p) {
                    m.state = DONE;
                    continue;
                }
            }
            const double b = m.prev[m.j0 + m.ka - 1], theta_top = m.prev[m.j0];
            if (!(b > 1e-8 * m.top)) {
                m.state = OUT;
                if (trace) std::fprintf(stderr, "[chefsi-batch] matrix %d: flat spectrum\n", z);
                continue;
            }
            const double c = 0.5 * b, e = 0.5 * b;
            q.ka = m.ka;
            q.m = std::min(chefsi_degree((theta_top - c) / e, m.kappa),
                           chefsi_needed(&m.prev[m.j0], &m.prev[K + m.j0], std::min(p - m.L, m.ka), c, e, 1e-2 * m.tol));
            q.voff = m.j0;
            q.c = c;
            q.e = e;
            deg = std::max(deg, q.m);
            any_active = true;
            any_locked = any_locked || m.L > 0;
        }
        upload();
        if (any_lock) {  // lock the converged leading columns and deflate them out of A
            launch_chb_lock(buf[v], stat, dp, K, p, nm, Pd, QL, thL, Wl, Qn, x.st);
            bgemm(N, Tr, dp, dp, K, -1.0, Wl, dp, dk, Qn, dp, dk, 1.0, A, dp, dd);
        }
        if (!any_active) break;
        if (trace)
            std::fprintf(stderr, "[chefsi-batch] it %d: max degree %d, active %d\n", it, deg,
                         (int)std::count_if(M.begin(), M.end(), [](const Mat &m) { return m.state == ACTIVE; }));
        launch_fill_f64(stat, (int64_t)nm * ns, 0.0, x.st);
        // the filter, deg steps for all; a matrix of lower degree carries its result
        int y1 = other({v});
        launch_chb_combine(buf[y1], WZ, buf[v], nullptr, dp, K, nm, Pd, 0, false, x.st);
        int y0 = -1;
        for (int j = 1; j < deg; ++j) {
            launch_chb_gemm(A, buf[y1], dp, K, nm, Pd, j, false, W, x.st);
            const int y2 = other({y0, y1, v});
            launch_chb_combine(buf[y2], W, buf[y1], j == 1 ? buf[v] : buf[y0], dp, K, nm, Pd, j, j == 1, x.st);
            y0 = y1;
            y1 = y2;
        }
        if (any_locked) {  // y -= QL (QL^T y)
            bgemm(Tr, N, p, K, dp, 1.0, QL, dp, (int64_t)dp * p, buf[y1], dp, dk, 0.0, T, p, (int64_t)p * K);
            bgemm(N, N, dp, K, p, -1.0, QL, dp, (int64_t)dp * p, T, p, (int64_t)p * K, 1.0, buf[y1], dp, dk);
        }
        x.check_launch();
        const int nv = rayleigh_ritz(cholqr2(y1, v), v, false);
        ++iters;
        // failures: filter the same Ritz vectors again at a lower degree (as chefsi_topp)
        bool any_retry = false;
        for (int z = 0; z < nm; ++z) {
            Mat &m = M[z];
            ChbMat &q = Ph[z];
            q = ChbMat{0, 0, 0, m.L, 0, 0, 0.0, 1.0};
            if (m.state != ACTIVE) continue;
            if (hs[z * ns + 2 * K] == 0.0 && finite(z, m.ka)) {
                m.kappa = std::min(m.kappa * 100.0, 1e8);
                m.retry = false;
                continue;
            }
            if (++m.retries > 8 || m.kappa < 10.0) {
                m.state = OUT;
                if (trace) std::fprintf(stderr, "[chefsi-batch] matrix %d: rank loss\n", z);
                continue;
            }
            m.kappa = std::sqrt(m.kappa);
            m.retry = true;
            // its Ritz block moves to the front of buf[nv]; A times it into WZ
            q.ka = m.ka;
            q.voff = m.j0;
            for (int i = 0; i < m.ka; ++i) {
                m.prev[i] = m.prev[m.j0 + i];
                m.prev[K + i] = m.prev[K + m.j0 + i];
            }
            m.j0 = 0;
            any_retry = true;
        }
        if (any_retry) {
            upload();
            launch_chb_copy(buf[nv], buf[v], K, dp, K, nm, Pd, true, x.st);
            x.check_launch();
            x.sync();  // the upload and the copy are done before the host changes Ph
            for (int z = 0; z < nm; ++z) Ph[z].voff = 0;
            upload();
            launch_chb_gemm(A, buf[nv], dp, K, nm, Pd, -1, false, WZ, x.st);
            x.check_launch();
            x.sync();
        }
        v = nv;
    }
    // the final Rayleigh-Ritz of each finished matrix's locked vectors against C
    x.sync();  // the last upload of Ph is done
    bool any_done = false;
    for (int z = 0; z < nm; ++z) {
        if (M[z].state == ACTIVE) M[z].state = OUT;
        Ph[z] = ChbMat{M[z].state == DONE ? p : 0, 0, 0, 0, 0, 0, 0.0, 1.0};
        any_done = any_done || M[z].state == DONE;
    }
    if (any_done) {
        upload();
        launch_chb_pad(C, d, dp, nm, Pd, A, x.st);
        launch_chb_copy(buf[0], QL, p, dp, K, nm, Pd, false, x.st);
        launch_chb_zero_tail(buf[0], dp, K, nm, Pd, x.st);
        launch_fill_f64(stat, (int64_t)nm * ns, 0.0, x.st);
        v = rayleigh_ritz(cholqr2(0, -1), -1, true);
        for (int z = 0; z < nm; ++z) {
            if (M[z].state != DONE) continue;
            const double *h = hs + z * ns;
            bool ok = h[2 * K] == 0.0 && h[2 * K + 1] == 0.0 && finite(z, p);
            for (int j = 0; ok && j < p; ++j) ok = h[K + j] <= 100.0 * M[z].tol;
            solved[z] = ok ? 1 : 0;
        }
        launch_chb_final(buf[v], stat, d, dp, K, p, nm, Vout, evals, x.st);
        launch_eig_select_batch(Vout, evals, d, p, nm, var, load, x.st, p, ids_dev);
        x.check_launch();
    }
    if (trace)
        std::fprintf(stderr, "[chefsi-batch] %d matrices: %d solved after %d iterations\n", nm,
                     (int)std::count(solved.begin(), solved.end(), 1), iters);
    return solved;
}

static void pca_device(Ctx &x, const double *dev_gram, int ld_g, const double *dev_colsum, int d, double n, int k,
                       double *dev_var, double *dev_load, double *dev_total) {
    double *cov = x.get<double>("cov", (size_t)d * d);
    double *w = x.get<double>("eig_w", d);
    double *e = x.get<double>("eig_e", d);
    rocblas_int *info = x.get<rocblas_int>("eig_info", 1);
    launch_cov(dev_gram, dev_colsum, d, ld_g, n, cov, dev_total, x.st);
    x.check_launch();
    if (chefsi_topp(x, cov, d, k, dev_var, dev_load)) return;
    if (use_topp(d, k, 1)) {
        double *tau = x.get<double>("eig_tau", d);
        double *W = x.get<double>("eig_topw", k);
        double *Y = x.get<double>("eig_y", (size_t)d * k);
        double *X = x.get<double>("eig_x", (size_t)d * k);
        int *tinfo = x.get<int>("eig_tinfo", 1);
        double *fac = x.get<double>("eig_fac", (size_t)3 * d * k);
        uint8_t *swp = x.get<uint8_t>("eig_swp", (size_t)d * k);
        tri_set_attrs();
        BLASCHK(rocsolver_dsytrd(x.blas, rocblas_fill_lower, d, cov, d, w, e, tau));
        launch_fill_u32(reinterpret_cast<uint32_t *>(tinfo), 1, 0u, x.st);
        launch_tri_topp(cov, 0, d, tau, 0, w, 0, e, 0, d, k, 1, W, Y, X, fac, swp, tinfo, x.st);
        launch_eig_select(X, W, d, k, dev_var, dev_load, x.st, k);
        x.check_launch();
        int hi = 0;
        x.download(&hi, tinfo, 1);
        x.sync();
        if (hi == 0) return;
        launch_cov(dev_gram, dev_colsum, d, ld_g, n, cov, dev_total, x.st);  // dsytrd overwrote it
        x.check_launch();
    }
    BLASCHK(rocsolver_dsyevd(x.blas, rocblas_evect_original, rocblas_fill_upper, d, cov, d, w, e, info));
    launch_eig_select(cov, w, d, k, dev_var, dev_load, x.st);
    x.check_launch();
    rocblas_int hinfo = 0;
    x.download(&hinfo, info, 1);
    x.sync();
    if (hinfo != 0) fail(CREST_E_RUNTIME, "eigensolver did not converge (info " + std::to_string(hinfo) + ")");
}

// k nearest rows of dev_xc [nc, p] for each row of dev_xq [nq, p] into
// dev_idx/dev_dist [nq, k]; dev_xq == nullptr gives the kNN graph of dev_xc (the
// queries are the candidates, each row's own entry excluded).
// k > 64 (knn_device): FP32 dot products of a block of queries (<= 2 GiB) with all candidates
// (rocBLAS) on rows centred on the candidates' mean, per-row radix selection of
// k2 = k + max(32, k/8) screened candidates
// (rounded up to 64, at most the candidate count), exact rank and output order.
// With `cert` (a device counter): each query's set certified against the GEMM screen's
// error bound (knn_cert_gemm), the queries not proven counted.
static void knn_device_large(Ctx &x, const float *dev_xc, int64_t nc, const float *dev_xq, int64_t nq, int p, int k,
                             uint32_t *dev_idx, float *dev_dist, hipEvent_t *ev3, uint32_t *cert = nullptr) {
    const bool graph = dev_xq == nullptr;
    if (graph) {
        dev_xq = dev_xc;
        nq = nc;
    }
    const int64_t n_valid = nc - (graph ? 1 : 0);
    int64_t k2 = ((int64_t)k + std::max(32, k / 8) + 63) / 64 * 64;
    k2 = std::min(k2, n_valid);
    if (k2 > kKnnLargeMaxCand || k > n_valid)
        value_error("k is too large for the exact GPU kNN (at most 3584 neighbours)");
    if (nc > INT32_MAX || p > INT32_MAX) value_error("too many points for the large-k GPU kNN");
    const int cap = (int)std::min<int64_t>(4 * k2, nc);  // per-row copy of the keys below the group bound
    const int64_t qb = std::max<int64_t>(
        1, std::min<int64_t>({nq, ((int64_t)1 << 29) / (nc + 3), ((int64_t)1 << 26) / cap}));  // dots <= 2 GiB
    const int64_t ld = (nc + 3) / 4 * 4;  // dot rows padded for four-candidate loads
    float *cn = x.get<float>("knnl_cn", (size_t)ld);
    float *qn = graph ? cn : x.get<float>("knnl_qn", (size_t)nq);
    float *dots = x.get<float>("knnl_dots", (size_t)qb * ld);
    uint32_t *cand = x.get<uint32_t>("knnl_cand", (size_t)qb * k2);
    uint32_t *comp_key = x.get<uint32_t>("knnl_comp_key", (size_t)qb * cap);
    uint32_t *comp_idx = x.get<uint32_t>("knnl_comp_idx", (size_t)qb * cap);
    float *tau = cert ? x.get<float>("knnl_tau", (size_t)qb) : nullptr;
    const KnnCert ct = knn_cert_gemm(p, cert);
    // the screen works on rows centred on the candidates' mean (its dot products then
    // carry no common offset); the rerank reads the original rows
    float *mean = x.get<float>("knnl_mean", (size_t)p);
    double *part = x.get<double>("knnl_part", (size_t)knn_centre_parts(nc) * p);
    float *sc = x.get<float>("knnl_xc", (size_t)nc * p);
    float *sq = graph ? sc : x.get<float>("knnl_xq", (size_t)nq * p);
    if (ev3) HIPCHK(hipEventRecord(ev3[0], x.st));
    launch_knn_centre(dev_xc, nc, p, mean, true, part, sc, x.st);
    if (!graph) launch_knn_centre(dev_xq, nq, p, mean, false, part, sq, x.st);
    launch_knn_sqnorm(sc, nc, p, cn, x.st);
    if (!graph) launch_knn_sqnorm(sq, nq, p, qn, x.st);
    if (ev3) HIPCHK(hipEventRecord(ev3[1], x.st));
    const float one = 1.0f, zero = 0.0f;
    for (int64_t q0 = 0; q0 < nq; q0 += qb) {
        const int64_t rows = std::min(qb, nq - q0);
        // dots[r][c] = x_c . q_(q0+r): column-major (nc x rows) = Xc^T (nc x p) * Q (p x rows)
        BLASCHK(rocblas_sgemm(x.blas, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)nc,
                              (rocblas_int)rows, p, &one, sc, p, sq + q0 * p, p, &zero, dots,
                              (rocblas_int)ld));
        launch_knn_select_rows(dots, ld, nc, qn, cn, q0, rows, graph ? 1 : 0, (int)k2, cap, comp_key, comp_idx, cand,
                               x.st);
        if (cert) launch_knn_cand_max(dots, ld, qn, cn, q0, rows, cand, (int)k2, tau, x.st);
        launch_knn_rerank_sort(dev_xq, dev_xc, p, cand, (int)k2, k, q0, rows, dev_idx, dev_dist, x.st, tau,
                               cert ? &ct : nullptr, sq);
    }
    if (ev3) {  // one timed span: screen and rank interleave per query block
        HIPCHK(hipEventRecord(ev3[2], x.st));
        HIPCHK(hipEventRecord(ev3[3], x.st));
    }
    x.check_launch();
}

static uint32_t read_count(Ctx &x, const uint32_t *dev) {
    uint32_t m = 0;
    HIPCHK(hipMemcpyAsync(&m, dev, sizeof(m), hipMemcpyDeviceToHost, x.st));
    HIPCHK(hipStreamSynchronize(x.st));
    return m;
}

// The GEMM recheck of m queries (rows listed on the device): their rows screened again
// against every candidate by the FP32 GEMM screen (k + max(32, k/8) candidates rounded
// up to 64, k + 1 in a graph so that the query's own row can be dropped), reranked in
// FP64 into their output rows, and certified with the GEMM screen's bound; returns how
// many are still not proven (a candidate within the FP32 screen's error of the k-th
// distance: a near tie).
static int64_t knn_recheck_gemm(Ctx &x, const float *dev_xc, int64_t nc, const float *dev_xq, int p, int k,
                                const uint32_t *rows, int64_t m, uint32_t *dev_idx, float *dev_dist) {
    const bool graph = dev_xq == nullptr;
    float *xsub = x.get<float>("knnr_xq", (size_t)m * p);
    launch_gather_rows(graph ? dev_xc : dev_xq, p, rows, m, xsub, x.st);
    const int k1 = graph ? (int)std::min<int64_t>(k + 1, nc) : k;
    uint32_t *ti = x.get<uint32_t>("knnr_ti", (size_t)m * k1);
    float *td = x.get<float>("knnr_td", (size_t)m * k1);
    uint32_t *cnt = x.get<uint32_t>("knnr_cnt", 1);
    launch_fill_u32(cnt, 1, 0u, x.st);
    knn_device_large(x, dev_xc, nc, xsub, m, p, k1, ti, td, nullptr, cnt);
    launch_knn_place(ti, td, m, k1, k, rows, graph ? 1 : 0, dev_idx, dev_dist, x.st);
    x.check_launch();
    return read_count(x, cnt);
}

// The queries a screen's certificate did not cover (m rows of the query set, listed
// on the device).  Their current k-th distance bounds the true one from above, so a
// threshold scan in FP32 over every candidate (rows centred on the candidates' mean,
// the GEMM screen's key; the threshold widened by the roundings of both distances and
// the key's bound) keeps every candidate that can be among the k nearest; an FP64
// rank of those (at most 256 per query)
// gives the exact k, written into the queries' output rows.  A query with more
// candidates than that under its threshold (many points at almost the same distance)
// takes the GEMM recheck.  Returns how many queries stay unproven.
// The threshold scan's products: FP32 MFMA (0, rows of at most 96 columns; after an
// FP16 screen the FP16 test on the screen's operands instead), rocBLAS dot products
// filtered in one pass (1, wider rows), or the VALU kernel (2); CREST_KNN_SCAN=gemm /
// valu picks the second / third, =fp32 the FP32 MFMA test after an FP16 screen too.
static int knn_scan_kind() {
    const char *e = std::getenv("CREST_KNN_SCAN");
    if (e && std::strcmp(e, "gemm") == 0) return 1;
    if (e && std::strcmp(e, "valu") == 0) return 2;
    if (e && std::strcmp(e, "fp32") == 0) return 3;  // the FP32 MFMA test also after an FP16 screen
    return 0;
}

// Buffers of a recheck of m queries against at most nc_max candidates (per-query arrays
// are indexed by the query's place in the recheck; the candidate arrays are reused by
// successive rechecks on the stream).
constexpr int kRecheckCap = 256;
struct RecheckBufs {
    float *xsub, *xsubc, *qn, *thr, *td, *xcc, *cn, *mean;
    uint32_t *cnt, *cand, *ti;
    double *part;
};
static RecheckBufs recheck_bufs(Ctx &x, int64_t m, int64_t nc_max, int p, int k) {
    RecheckBufs B;
    B.xsub = x.get<float>("knnr_xq", (size_t)m * p);
    B.xsubc = x.get<float>("knnr_xqc", (size_t)m * p);
    B.qn = x.get<float>("knnr_qn", (size_t)m);
    B.thr = x.get<float>("knnr_thr", (size_t)m);
    B.cnt = x.get<uint32_t>("knnr_scnt", (size_t)m);
    B.cand = x.get<uint32_t>("knnr_cand", (size_t)m * kRecheckCap);
    B.ti = x.get<uint32_t>("knnr_si", (size_t)m * k);
    B.td = x.get<float>("knnr_sd", (size_t)m * k);
    B.xcc = x.get<float>("knnr_xcc", (size_t)nc_max * p);
    B.cn = x.get<float>("knnr_cn", (size_t)nc_max);
    B.mean = x.get<float>("knnr_mean", (size_t)p);
    B.part = x.get<double>("knnr_part", (size_t)knn_centre_parts(nc_max) * p);
    return B;
}
// The same buffers for the queries from place a on.
static RecheckBufs recheck_slice(RecheckBufs B, int64_t a, int p, int k) {
    B.xsub += a * p;
    B.xsubc += a * p;
    B.qn += a;
    B.thr += a;
    B.cnt += a;
    B.cand += a * kRecheckCap;
    B.ti += a * k;
    B.td += a * k;
    return B;
}

// The queries a screen's certificate did not cover (m rows of the query set, listed
// on the device; p <= 256).  Their current k-th distance bounds the true one from
// above, so a threshold scan in FP32 over every candidate (rows centred on the
// candidates' mean, the GEMM screen's key; the threshold widened by the roundings of
// both distances and the key's bound) keeps every candidate that can be among the k
// nearest; an FP64 rank of those (at most 256 per query) gives the exact k, written
// into the queries' output rows.  Launches only: cnt (per query, the candidates under
// its threshold) tells afterwards which queries overflowed their lists.  `mean`: the
// candidates' column means if known (the screen's centring), else computed.
static void knn_recheck_launch(Ctx &x, const RecheckBufs &B, const float *dev_xc, int64_t nc, const float *dev_xq,
                               int p, int k, const uint32_t *rows, int64_t m, uint32_t *dev_idx, float *dev_dist,
                               const float *mean_known) {
    const bool graph = dev_xq == nullptr;
    constexpr int cap = kRecheckCap;
    float *mean = mean_known ? const_cast<float *>(mean_known) : B.mean;  // read only when known
    launch_gather_rows(graph ? dev_xc : dev_xq, p, rows, m, B.xsub, x.st);
    launch_knn_centre(dev_xc, nc, p, mean, mean_known == nullptr, B.part, B.xcc, x.st);
    launch_knn_centre(B.xsub, m, p, mean, false, B.part, B.xsubc, x.st);
    launch_knn_sqnorm(B.xcc, nc, p, B.cn, x.st);
    launch_knn_sqnorm(B.xsubc, m, p, B.qn, x.st);
    launch_fill_u32(B.cnt, m, 0u, x.st);
    launch_fill_u32(B.cand, m * cap, 0xFFFFFFFFu, x.st);
    const int scan = knn_scan_kind();
    if ((scan == 0 || scan == 3) && p <= 96) {
        launch_knn_scan_thresholds(dev_dist, rows, B.xsubc, m, k, p, B.thr, x.st);
        launch_knn_thr_mfma(B.xcc, B.cn, nc, p, B.xsubc, B.qn, m, B.thr, rows, graph ? 1 : 0, B.cnt, B.cand, cap,
                            x.st);
    } else if (scan != 2) {
        // the dot products of blocks of the rechecked rows with every candidate by
        // rocBLAS (<= 2 GiB per block), each block filtered in one pass
        launch_knn_scan_thresholds(dev_dist, rows, B.xsubc, m, k, p, B.thr, x.st);
        const int64_t ld = (nc + 3) / 4 * 4;
        const int64_t qb = std::max<int64_t>(1, std::min<int64_t>(m, ((int64_t)1 << 29) / ld));
        float *dots = x.get<float>("knnl_dots", (size_t)qb * ld);
        const float one = 1.0f, zero = 0.0f;
        for (int64_t q0 = 0; q0 < m; q0 += qb) {
            const int64_t nr = std::min(qb, m - q0);
            BLASCHK(rocblas_sgemm(x.blas, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)nc,
                                  (rocblas_int)nr, p, &one, B.xcc, p, B.xsubc + q0 * p, p, &zero, dots,
                                  (rocblas_int)ld));
            launch_knn_thr_filter(dots, ld, nc, B.cn, B.qn, B.thr, rows, graph ? 1 : 0, q0, nr, B.cnt, B.cand, cap,
                                  x.st);
        }
    } else {
        launch_knn_thr_scan(B.xcc, B.cn, nc, p, B.xsubc, B.qn, m, dev_dist, rows, k, graph ? 1 : 0, B.thr, B.cnt,
                            B.cand, cap, x.st);
    }
    launch_knn_rerank_sort(B.xsub, dev_xc, p, B.cand, cap, k, 0, m, B.ti, B.td, x.st);
    launch_knn_place(B.ti, B.td, m, k, k, rows, 0, dev_idx, dev_dist, x.st);
    x.check_launch();
}

// The places (0 .. m-1) of a finished recheck whose lists overflowed (more than 256
// candidates under the threshold: many points at almost one distance).
static std::vector<int64_t> recheck_overflows(Ctx &x, const RecheckBufs &B, int64_t m) {
    std::vector<uint32_t> hc((size_t)m);
    x.download(hc.data(), B.cnt, (size_t)m);
    x.sync();
    std::vector<int64_t> over;
    for (int64_t r = 0; r < m; ++r)
        if (hc[(size_t)r] > (uint32_t)kRecheckCap) over.push_back(r);
    return over;
}

static int64_t recheck_finish(Ctx &x, const RecheckBufs &B, const float *dev_xc, int64_t nc, const float *dev_xq, int p,
                              int k, const uint32_t *rows, int64_t m, uint32_t *dev_idx, float *dev_dist);

// The recheck of an FP16 screen's unproven queries on the screen's own operands: the
// threshold test in FP16 on MFMA with the FP16 bound folded into each threshold (no
// centring pass, a quarter of the FP32 test's MFMA K-steps), then the same FP64 rank
// and placement, and the GEMM recheck for overflowed lists.
static int64_t knn_recheck16(Ctx &x, const void *xa, const void *xb, int kpad, const KnnOrder &o, const float *dev_xc,
                             int64_t nc, const float *dev_xq, int p, int k, const uint32_t *rows, const uint32_t *pos,
                             int64_t m, uint32_t *dev_idx, float *dev_dist) {
    const bool graph = dev_xq == nullptr;
    const RecheckBufs B = recheck_bufs(x, m, 1, p, k);
    launch_gather_rows(graph ? dev_xc : dev_xq, p, rows, m, B.xsub, x.st);
    launch_fill_u32(B.cnt, m, 0u, x.st);
    launch_fill_u32(B.cand, m * kRecheckCap, 0xFFFFFFFFu, x.st);
    launch_knn_thr16(xa, xb, kpad, nc, o, B.xsub, dev_dist, rows, pos, m, k, p, graph ? 1 : 0, B.thr, B.cnt, B.cand,
                     kRecheckCap, x.st);
    launch_knn_rerank_sort(B.xsub, dev_xc, p, B.cand, kRecheckCap, k, 0, m, B.ti, B.td, x.st);
    launch_knn_place(B.ti, B.td, m, k, k, rows, 0, dev_idx, dev_dist, x.st);
    x.check_launch();
    return recheck_finish(x, B, dev_xc, nc, dev_xq, p, k, rows, m, dev_idx, dev_dist);
}

// The whole recheck of one query set: the threshold scan, then the GEMM recheck for
// the queries whose lists overflowed (and for rows wider than 256 columns).  Returns
// how many queries stay unproven.
static int64_t knn_recheck(Ctx &x, const float *dev_xc, int64_t nc, const float *dev_xq, int p, int k,
                           const uint32_t *rows, int64_t m, uint32_t *dev_idx, float *dev_dist,
                           const float *mean_known = nullptr) {
    if (p > 256) return knn_recheck_gemm(x, dev_xc, nc, dev_xq, p, k, rows, m, dev_idx, dev_dist);
    const RecheckBufs B = recheck_bufs(x, m, nc, p, k);
    knn_recheck_launch(x, B, dev_xc, nc, dev_xq, p, k, rows, m, dev_idx, dev_dist, mean_known);
    return recheck_finish(x, B, dev_xc, nc, dev_xq, p, k, rows, m, dev_idx, dev_dist);
}

// After a threshold-scan recheck: the GEMM recheck for the queries whose lists
// overflowed; returns how many queries stay unproven.
static int64_t recheck_finish(Ctx &x, const RecheckBufs &B, const float *dev_xc, int64_t nc, const float *dev_xq, int p,
                              int k, const uint32_t *rows, int64_t m, uint32_t *dev_idx, float *dev_dist) {
    const std::vector<int64_t> over = recheck_overflows(x, B, m);
    if (over.empty()) return 0;
    std::vector<uint32_t> hr((size_t)m), ov;
    x.download(hr.data(), rows, (size_t)m);
    x.sync();
    for (int64_t r : over) ov.push_back(hr[(size_t)r]);
    uint32_t *dl = x.get<uint32_t>("knnr_over", ov.size());
    HIPCHK(hipMemcpyAsync(dl, ov.data(), ov.size() * sizeof(uint32_t), hipMemcpyHostToDevice, x.st));
    x.sync();
    return knn_recheck_gemm(x, dev_xc, nc, dev_xq, p, k, dl, (int64_t)ov.size(), dev_idx, dev_dist);
}

static void knn_device(Ctx &x, const float *dev_xc, int64_t nc, const float *dev_xq, int64_t nq, int p, int k,
                       int prec, uint32_t *dev_idx, float *dev_dist, hipEvent_t *ev3) {
    const int kp = prec == CREST_PREC_FP64 ? CREST_PREC_FP32 : prec;
    const int kpad = knn_kpad(kp, p);
    x.knn_checked = x.knn_rechecked = x.knn_uncertified = 0;
    // k above the register lists, or rows wider than the register screens hold (134
    // columns at FP32, 137 split-BF16, 252 FP16): the any-width FP32 GEMM screen of
    // k + max(32, k/8) candidates per query, then the same exact rerank (no certificate)
    if (k > 64 || kpad < 0) {
        uint32_t *cnt = x.get<uint32_t>("knn_cert_cnt", 1);
        launch_fill_u32(cnt, 1, 0u, x.st);
        knn_device_large(x, dev_xc, nc, dev_xq, nq, p, k, dev_idx, dev_dist, ev3, cnt);
        x.knn_checked = dev_xq ? nq : nc;
        x.knn_uncertified = read_count(x, cnt);
        return;
    }
    const bool graph = dev_xq == nullptr;
    if (graph) nq = nc;
    const int L = k <= 16 ? 16 : k <= 32 ? 32 : 64;
    const int64_t nc_pad = knn_rows_padded(nc), nq_pad = knn_rows_padded(nq);
    const size_t eb = kp == CREST_PREC_FP32 ? 4 : 2;
    void *xa = x.get<uint8_t>("knn_xa", (size_t)nc_pad * kpad * eb);
    void *xb = x.get<uint8_t>("knn_xb", (size_t)nq_pad * kpad * eb);
    void *work = x.get<uint8_t>("knn_work", knn_work_bytes(nc, graph ? 0 : nq, p));
    float *cd = x.get<float>("knn_cd", (size_t)nq_pad * 2 * L);
    uint32_t *ci = x.get<uint32_t>("knn_ci", (size_t)nq_pad * 2 * L);
    if (ev3) HIPCHK(hipEventRecord(ev3[0], x.st));
    const KnnOrder o = launch_knn_prep(kp, dev_xc, nc, dev_xq, nq, p, kpad, work, xa, xb, x.st);
    if (ev3) HIPCHK(hipEventRecord(ev3[1], x.st));
    if (!launch_knn_screen(kp, xa, xb, nc, nq, kpad, L, graph ? 1 : 0, o, cd, ci, x.st))
        value_error("no kNN kernel for this width");
    if (ev3) HIPCHK(hipEventRecord(ev3[2], x.st));
    uint32_t *cnt = x.get<uint32_t>("knn_cert_cnt", 1);
    uint32_t *unc = x.get<uint32_t>("knn_cert_rows", (size_t)nq);
    uint32_t *upos = x.get<uint32_t>("knn_cert_pos", (size_t)nq);
    launch_fill_u32(cnt, 1, 0u, x.st);
    KnnCert ct = knn_cert(kp, p, kpad, cd, cnt, unc, nq);
    ct.pos = upos;
    launch_knn_rerank(graph ? dev_xc : dev_xq, nq, dev_xc, nc, p, o, graph ? 1 : 0, ci, 2 * L, k, dev_idx, dev_dist,
                      x.st, nullptr, nullptr, &ct);
    x.check_launch();
    const uint32_t m = read_count(x, cnt);
    x.knn_checked = nq;
    x.knn_rechecked = m;
    if (m && kp == CREST_PREC_FP16 && knn_scan_kind() == 0)
        x.knn_uncertified = knn_recheck16(x, xa, xb, kpad, o, dev_xc, nc, graph ? nullptr : dev_xq, p, k, unc, upos, m,
                                          dev_idx, dev_dist);
    else if (m)
        x.knn_uncertified =
            knn_recheck(x, dev_xc, nc, graph ? nullptr : dev_xq, p, k, unc, m, dev_idx, dev_dist, o.mean);
    if (ev3) HIPCHK(hipEventRecord(ev3[3], x.st));
    x.check_launch();
}

// mean/std of the HVGs from per-gene stats (pp.py:123-126, tl.py:31-34).
// blockIdx.y = subset y of a batch: stats + y n_genes 5, mean/sd + y d, n from cell_off.
__global__ void k_mean_std(const double *stats, const int32_t *gene_of_hvg, int d, double n, double *mean,
                           double *sd, int64_t n_genes, const int64_t *cell_off) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= d) return;
    if (cell_off) {
        const int y = blockIdx.y;
        stats += (int64_t)y * n_genes * 5;
        mean += (int64_t)y * d;
        sd += (int64_t)y * d;
        n = (double)(cell_off[y + 1] - cell_off[y]);
    }
    const int g = gene_of_hvg[h];
    const double m = stats[(int64_t)g * 5] / n;
    double var = 0.0;
    if (n > 1.0) {
        const double nm = n * m;
        const double nmm = nm * m;
        var = (stats[(int64_t)g * 5 + 1] - nmm) / (n - 1.0);
    }
    var = var > 0.0 ? var : 0.0;
    double s = sqrt(var);
    if (s == 0.0) s = 1.0;
    mean[h] = m;
    sd[h] = s;
}

__global__ void k_copy16(uint4 *dst, const uint4 *src, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_iota_i64(int64_t *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = i;
}
__global__ void k_iota_i32(int32_t *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = (int32_t)i;
}

}  // namespace crest

using namespace crest;

// ================================================================ C ABI
// ---------------------------------------------------------------- fused pipeline
// HVG index -> store gene from hvg_of_gene (store gene -> HVG index or -1): every
// index 0..d-1 used by exactly one gene.
static std::vector<int32_t> hvg_genes(const int32_t *hvg_of_gene, size_t n_genes, int d) {
    std::vector<int32_t> gene_of_hvg(d, -1);
    for (size_t g = 0; g < n_genes; ++g) {
        const int32_t h = hvg_of_gene[g];
        if (h >= d) value_error("hvg_of_gene values must be < d");
        if (h < 0) continue;
        if (gene_of_hvg[h] >= 0) value_error("hvg_of_gene maps two genes to HVG index " + std::to_string(h));
        gene_of_hvg[h] = (int32_t)g;
    }
    for (int h = 0; h < d; ++h)
        if (gene_of_hvg[h] < 0) value_error("every HVG index 0..d-1 must be used");
    return gene_of_hvg;
}

static void check_pipeline_args(const crest_pipeline_args *a, double *variance, double *loadings,
                                crest_pipeline_result *res) {
    if (!a || !res || !variance || !loadings) value_error("NULL argument");
    if (a->n_cells < 2 || a->d == 0) fail(CREST_E_RUNTIME, "need at least 2 cells and 1 gene");
    if (!a->hvg_of_gene || !a->dev_x_pca || !a->dev_knn_idx || !a->dev_knn_dist) value_error("NULL argument");
    if (a->prec_gram != CREST_PREC_FP64 && a->prec_gram != CREST_PREC_FP32 && a->prec_gram != CREST_PREC_BF16X3 &&
        a->prec_gram != CREST_PREC_FP16)
        value_error("unknown Gram precision");
    check_gram_prec(a->prec_gram, a->scale, a->max_value);
    hvg_genes(a->hvg_of_gene, a->n_genes, (int)a->d);
    if (a->prec_knn != CREST_PREC_FP64 && a->prec_knn != CREST_PREC_FP32 && a->prec_knn != CREST_PREC_BF16X3 &&
        a->prec_knn != CREST_PREC_FP16)
        value_error("unknown kNN screen precision");
}

// tl.pca + pp.neighbors over a chunk source: each(tmpl, fn) calls fn(chunk) on the
// library stream for every chunk of the CSR in cell order, where chunk is tmpl with
// its span/entry fields filled in (cells are rows through the identity cell map).
template <class Each>
static void run_pipeline(Ctx &x, const crest_pipeline_args *a, Each &&each, double *variance, double *loadings,
                         crest_pipeline_result *res) {
    const int64_t n = (int64_t)a->n_cells;
    const int d = (int)a->d;
    const int kp = (int)std::min<int64_t>({(int64_t)a->n_comps, (int64_t)d, n - 1});
    if (kp > 256) value_error("n_comps > 256 is not supported");
    const int kn = (int)std::max<int64_t>(std::min<int64_t>((int64_t)a->k, n - 1), 1);
    const std::vector<int32_t> gene_of_hvg = hvg_genes(a->hvg_of_gene, a->n_genes, d);
    const int d_pad = (int)round_up(d, kGramBlock);
    const int64_t tile_rows = a->chunk_rows ? (int64_t)round_up((int64_t)a->chunk_rows, 16) : 65536;
    // constant maps: identity cell map and gene map (every cell and gene kept)
    int64_t *cmap = x.get<int64_t>("pl_cmap", n);
    int32_t *gmap = x.get<int32_t>("pl_gmap", a->n_genes);
    hipLaunchKernelGGL(k_iota_i64, dim3(1024), dim3(256), 0, x.st, cmap, n);
    hipLaunchKernelGGL(k_iota_i32, dim3(256), dim3(256), 0, x.st, gmap, (int64_t)a->n_genes);
    const int32_t *sub = x.upload("pl_sub", a->hvg_of_gene, a->n_genes);
    const int32_t *goh = x.upload("pl_goh", gene_of_hvg.data(), d);
    double *stats = x.get<double>("pl_stats", a->n_genes * 5);
    double *mean = x.get<double>("pl_mean", d);
    double *sd = x.get<double>("pl_sd", d);
    double *dg = x.get<double>("pl_gram", (size_t)d_pad * d_pad);
    double *dc = x.get<double>("pl_colsum", d_pad);
    double *dv = x.get<double>("pl_var", kp);
    double *dl = x.get<double>("pl_load", (size_t)d * kp);
    double *dt = x.get<double>("pl_total", 1);
    float *shift = x.get<float>("pl_shift", kp);
    x.sync();

    DevChunk tmpl{};
    tmpl.cell_map = cmap;
    tmpl.n_cell_map = n;
    tmpl.gene_var = gmap;
    tmpl.n_gene_map = (int64_t)a->n_genes;
    tmpl.target_sum = a->target_sum;
    tmpl.log1p = a->log1p ? 1 : 0;
    DevScale sc{};
    sc.sub = sub;
    sc.scale = a->scale ? 1 : 0;
    sc.max_value = a->max_value;
    sc.d = d;
    sc.mean = mean;
    sc.sd = sd;

    hipEvent_t *ev = x.ev;
    HIPCHK(hipEventRecord(ev[0], x.st));
    // PASS 0: gene statistics -> mean/std (tl.py:24-34)
    if (a->scale) {
        launch_fill_f64(stats, (int64_t)a->n_genes * 5, 0.0, x.st);
        each(tmpl, [&](const DevChunk &c) { launch_gene_stats(c, stats, (int)a->n_genes, x.st); });
        hipLaunchKernelGGL(k_mean_std, dim3((d + 255) / 256), dim3(256), 0, x.st, stats, goh, d, (double)n, mean, sd,
                           (int64_t)0, (const int64_t *)nullptr);
    }
    HIPCHK(hipEventRecord(ev[1], x.st));
    // PASS 1: Gram + colsum (tl.py:68-71)
    launch_fill_f64(dg, (int64_t)d_pad * d_pad, 0.0, x.st);
    launch_fill_f64(dc, d_pad, 0.0, x.st);
    each(tmpl, [&](const DevChunk &c) { gram_rows(x, c, sc, a->prec_gram, d_pad, tile_rows, dg, dc); });
    HIPCHK(hipEventRecord(ev[2], x.st));
    // eigensolve (tl.py:72-73)
    pca_device(x, dg, d_pad, dc, d, (double)n, kp, dv, dl, dt);
    HIPCHK(hipEventRecord(ev[3], x.st));
    // PASS 2: shift prefill + projection (tl.py:76-80)
    launch_shift_fill(dc, dl, d, kp, (double)n, n, a->dev_x_pca, shift, x.st);
    each(tmpl, [&](const DevChunk &c) { launch_project(c, sc, dl, kp, a->dev_x_pca, 0, x.st); });
    x.check_launch();
    HIPCHK(hipEventRecord(ev[4], x.st));
    // kNN (pp.py:309)
    knn_device(x, a->dev_x_pca, n, nullptr, 0, kp, kn, a->prec_knn, a->dev_knn_idx, a->dev_knn_dist, ev + 5);
    x.sync();
    float ms[8] = {};
    for (int i = 0; i < 8; ++i) HIPCHK(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
    res->k_pcs = (size_t)kp;
    res->k_nn = (size_t)kn;
    res->ms_stats = ms[0];
    res->ms_gram = ms[1];
    res->ms_eig = ms[2];
    res->ms_project = ms[3];
    res->ms_knn_prep = ms[5];
    res->ms_knn_screen = ms[6];
    res->ms_knn_rerank = ms[7];
    float tot = 0.f;
    HIPCHK(hipEventElapsedTime(&tot, ev[0], ev[8]));
    res->ms_total = tot;
    x.download(variance, dv, kp);
    std::vector<double> hl((size_t)d * kp);
    x.download(hl.data(), dl, hl.size());
    x.download(&res->total_variance, dt, 1);
    x.sync();
    for (int i = 0; i < d; ++i)
        for (int j = 0; j < kp; ++j) loadings[(size_t)i * a->n_comps + j] = hl[(size_t)i * kp + j];
}

// memcpy split over up to 8 threads (host staging of streamed chunks).
static void par_copy(void *dst, const void *src, size_t bytes) {
    const size_t part_min = (size_t)8 << 20;
    const int nt = (int)std::min<size_t>(8, std::max<size_t>(1, bytes / part_min));
    if (nt <= 1) {
        if (bytes) std::memcpy(dst, src, bytes);
        return;
    }
    const size_t per = (bytes + nt - 1) / nt;
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t) {
        const size_t o = (size_t)t * per;
        if (o >= bytes) break;
        const size_t len = std::min(per, bytes - o);
        th.emplace_back([=] { std::memcpy((char *)dst + o, (const char *)src + o, len); });
    }
    for (auto &t : th) t.join();
}

// ---------------------------------------------------------------- batch of subsets
// Host-side plan of a batch: where each subset's rows sit in the densified tiles, the
// row ranges each workgroup takes, the Gram work items and the reduce items.
struct BatchTile {
    int64_t rows_pad = 0;
    std::vector<SegDesc> segs;
    std::vector<RowRange> dens;
    std::vector<GramItem> items;
    std::vector<RedItem> red;
    int64_t n_slots = 0;
};

constexpr int64_t kDensRows = 64;     // rows densified per workgroup
constexpr int64_t kStatsRows = 2048;  // rows per gene-statistics workgroup
constexpr int64_t kProjRows = 128;    // rows per projection workgroup

// Subsets laid out one after another in tiles of at most tile_rows rows (each segment
// padded to 16 rows); a subset that does not fit the rest of a tile continues in the
// next one.
static std::vector<BatchTile> plan_tiles(const std::vector<int64_t> &off, int64_t tile_rows) {
    std::vector<BatchTile> tiles(1);
    const int n_sub = (int)off.size() - 1;
    for (int y = 0; y < n_sub; ++y) {
        int64_t c = off[y];
        while (c < off[y + 1]) {
            if (tile_rows - tiles.back().rows_pad < 16) tiles.emplace_back();
            BatchTile &t = tiles.back();
            const int64_t take = std::min(off[y + 1] - c, tile_rows - t.rows_pad);
            const int64_t rows16 = round_up(take, 16);
            t.segs.push_back(SegDesc{c, take, t.rows_pad / kRowGroup, y, 0});
            for (int64_t r = 0; r < rows16; r += kDensRows)
                t.dens.push_back(RowRange{r, std::min(r + kDensRows, rows16), (int32_t)t.segs.size() - 1, 0});
            t.rows_pad += rows16;
            c += take;
        }
    }
    if (tiles.back().segs.empty()) tiles.pop_back();
    return tiles;
}

// Gram items of one tile: a segment's rows are cut into ceil(rows / gram_rows) ranges
// of equal length (whole 16-row steps).  A segment of one range adds its blocks
// straight into its subset's Gram (one item per block and launch, so no two items
// touch the same block); the ranges of a longer segment write partial slots that a
// reduce item sums in range order.  Ranges are dealt round-robin, longest first, to
// 8 queues that are then interleaved: with a persistent grid of a multiple of 8
// workgroups (dispatched round-robin over the 8 XCDs) every block of one row range
// runs on one XCD and reads those rows through that XCD's L2.
static void plan_items(BatchTile &t, int nblk, int64_t gram_rows) {
    struct Range {
        int32_t subset, n_split, j;
        int64_t g0, g1, slot0;
    };
    std::vector<Range> ranges;
    t.items.clear();
    t.red.clear();
    t.n_slots = 0;
    for (const SegDesc &sd : t.segs) {
        const int64_t rows16 = round_up(sd.n, 16);
        const int64_t groups = rows16 / kRowGroup;
        const int64_t want = std::max<int64_t>(1, ceil_div(rows16, gram_rows));
        const int64_t per = round_up(ceil_div(groups, want), 2);
        const int n_split = (int)ceil_div(groups, per);
        int64_t slot0 = -1;
        if (n_split > 1) {
            slot0 = t.n_slots;
            t.n_slots += (int64_t)nblk * n_split;
            for (int b = 0; b < nblk; ++b) t.red.push_back(RedItem{b, sd.subset, slot0 + (int64_t)b * n_split, n_split, 0});
        }
        for (int j = 0; j < n_split; ++j)
            ranges.push_back(Range{sd.subset, n_split, j, sd.grp0 + j * per, sd.grp0 + std::min(groups, (j + 1) * per),
                                   slot0});
    }
    std::stable_sort(ranges.begin(), ranges.end(),
                     [](const Range &a, const Range &b) { return a.g1 - a.g0 > b.g1 - b.g0; });
    std::vector<GramItem> q[8];
    for (size_t r = 0; r < ranges.size(); ++r) {
        const Range &R = ranges[r];
        for (int b = 0; b < nblk; ++b)
            q[r % 8].push_back(GramItem{b, R.subset, R.g0, R.g1, R.slot0 < 0 ? -1 : R.slot0 + (int64_t)b * R.n_split + R.j});
    }
    size_t qmax = 0;
    for (auto &v : q) qmax = std::max(qmax, v.size());
    t.items.assign(8 * qmax, GramItem{-1, 0, 0, 0, 0});
    for (size_t j = 0; j < qmax; ++j)
        for (int z = 0; z < 8; ++z)
            if (j < q[z].size()) t.items[8 * j + z] = q[z][j];
}

// Host tables packed into one device buffer (256-byte aligned pieces).
struct TablePack {
    std::vector<uint8_t> bytes;
    template <class T> size_t add(const std::vector<T> &v) {
        const size_t o = round_up((int64_t)bytes.size(), 256);
        bytes.resize(o + std::max<size_t>(v.size() * sizeof(T), 1));
        if (!v.empty()) std::memcpy(bytes.data() + o, v.data(), v.size() * sizeof(T));
        return o;
    }
};

// The autotuner's persistent cache: "key value" lines in the file named by
// CREST_TUNE_CACHE (read once per process, appended on every new choice).
static std::mutex g_tune_mu;
static std::map<std::string, int64_t> g_tune;
static bool g_tune_loaded = false;
static bool tune_lookup(const std::string &key, int64_t &v) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!g_tune_loaded) {
        g_tune_loaded = true;
        if (const char *f = std::getenv("CREST_TUNE_CACHE")) {
            if (FILE *fp = std::fopen(f, "r")) {
                char k[512];
                long long val;
                while (std::fscanf(fp, "%511s %lld", k, &val) == 2) g_tune[k] = val;
                std::fclose(fp);
            }
        }
    }
    auto it = g_tune.find(key);
    if (it == g_tune.end()) return false;
    v = it->second;
    return true;
}
static void tune_store(const std::string &key, int64_t v) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = v;
    if (const char *f = std::getenv("CREST_TUNE_CACHE"))
        if (FILE *fp = std::fopen(f, "a")) {
            std::fprintf(fp, "%s %lld\n", key.c_str(), (long long)v);
            std::fclose(fp);
        }
}

static void check_batch_args(const crest_batch_args *b, double *variance, double *loadings, double *total,
                             crest_batch_result *res) {
    if (!b || !variance || !loadings || !total || !res || !b->subset_off) value_error("NULL argument");
    const crest_pipeline_args *a = &b->pipe;
    if (!a->dev_indptr || !a->dev_genes || !a->dev_values) value_error("NULL argument");
    if (!a->hvg_of_gene || !a->dev_x_pca || !a->dev_knn_idx || !a->dev_knn_dist) value_error("NULL argument");
    if (b->n_subsets < 1 || b->n_subsets > (1u << 20)) value_error("n_subsets must be in 1 .. 2^20");
    if (a->d == 0) fail(CREST_E_RUNTIME, "need at least 1 gene");
    if (a->n_cells >= (size_t)UINT32_MAX) value_error("too many cells for one batch");
    if (a->prec_gram != CREST_PREC_FP64 && a->prec_gram != CREST_PREC_FP32 && a->prec_gram != CREST_PREC_BF16X3 &&
        a->prec_gram != CREST_PREC_FP16)
        value_error("unknown Gram precision");
    check_gram_prec(a->prec_gram, a->scale, a->max_value);
    if (a->prec_knn != CREST_PREC_FP64 && a->prec_knn != CREST_PREC_FP32 && a->prec_knn != CREST_PREC_BF16X3 &&
        a->prec_knn != CREST_PREC_FP16)
        value_error("unknown kNN screen precision");
    if (a->k < 1 || a->k > 64) value_error("the batched kNN needs 1 <= k <= 64");
    hvg_genes(a->hvg_of_gene, a->n_genes, (int)a->d);
    const int64_t kp = (int64_t)std::min<size_t>(a->n_comps, a->d);
    if (kp < 1 || kp > 256) value_error("n_comps must be in 1 .. 256");
    const int64_t need = std::max<int64_t>(kp, (int64_t)a->k) + 1;
    const int64_t *o = b->subset_off;
    if (o[0] != 0 || o[b->n_subsets] != (int64_t)a->n_cells) value_error("subset_off must run from 0 to n_cells");
    for (size_t i = 0; i < b->n_subsets; ++i)
        if (o[i + 1] - o[i] < need)
            value_error("subset " + std::to_string(i) + " has " + std::to_string(o[i + 1] - o[i]) +
                        " cells; a batch needs at least max(n_comps, k) + 1 = " + std::to_string(need) +
                        " per subset");
}

// Everything one call needs on the device, planned on the host.
struct BatchPlan {
    std::string pre;  // buffer-name prefix: graph-mode calls own their buffers ("bg_"), eager ones "bt_"
    int n_sub = 0, d = 0, d_pad = 0, kp = 0, kn = 0, fmt = 0, kprec = 0, L = 0, kpad = 0, nblk = 0, grid = 0;
    int64_t n = 0, max_rows = 0, n_pad = 0, max_pad = 0, tile_rows = 0, gram_rows = 0, eig_batch = 0, nnz = 0;
    std::vector<int64_t> off, row0, nrow;
    std::vector<BatchTile> tiles;
    // device tables
    const int64_t *d_off = nullptr, *d_row0 = nullptr, *d_nrow = nullptr;
    const SegDesc *d_sub_segs = nullptr;
    const RowRange *d_stats_ranges = nullptr, *d_proj_ranges = nullptr;
    int64_t n_stats_ranges = 0, n_proj_ranges = 0;
    const int32_t *d_blk_seg = nullptr, *d_order = nullptr;
    std::vector<const SegDesc *> d_tile_segs;
    std::vector<const RowRange *> d_tile_dens;
    std::vector<const GramItem *> d_tile_items;
    std::vector<const RedItem *> d_tile_red;
    int64_t max_slots = 0, max_tile_rows = 0;
    // the FP64 Gram by sparse pair products: 64-row groups (never straddling two
    // subsets), items over group ranges, reduce items for subsets of several ranges
    bool sparse = false;
    std::vector<SegDesc> sp_groups;
    std::vector<GramItem> sp_items;
    std::vector<RedItem> sp_red;
    int64_t sp_slots = 0;
    const SegDesc *d_sp_groups = nullptr;
    const GramItem *d_sp_items = nullptr;
    const RedItem *d_sp_red = nullptr;
};

// Sparse Gram work of a batch: each subset's rows in groups of 64, its groups cut
// into ceil(groups / range_groups) ranges of equal length; one item per (range,
// block), the items of the longest ranges first.  A subset of one range adds its
// blocks straight into its Gram; the ranges of a longer one write partial slots
// that a reduce item sums in range order.
static void plan_sparse(BatchPlan &P, int64_t range_groups) {
    struct Range {
        int32_t subset, n_split, j;
        int64_t g0, g1, slot0;
    };
    P.sp_groups.clear();
    P.sp_items.clear();
    P.sp_red.clear();
    P.sp_slots = 0;
    std::vector<Range> ranges;
    for (int y = 0; y < P.n_sub; ++y) {
        const int64_t ny = P.off[y + 1] - P.off[y], gy0 = (int64_t)P.sp_groups.size();
        for (int64_t r = 0; r < ny; r += kRowGroupSparse)
            P.sp_groups.push_back(SegDesc{P.off[y] + r, std::min<int64_t>(kRowGroupSparse, ny - r), 0, y, 0});
        const int64_t G = (int64_t)P.sp_groups.size() - gy0;
        if (G == 0) continue;
        const int n_split = (int)ceil_div(G, range_groups);
        const int64_t per = ceil_div(G, n_split);
        int64_t slot0 = -1;
        if (n_split > 1) {
            slot0 = P.sp_slots;
            P.sp_slots += (int64_t)P.nblk * n_split;
            for (int b = 0; b < P.nblk; ++b)
                P.sp_red.push_back(RedItem{b, y, slot0 + (int64_t)b * n_split, n_split, 0});
        }
        for (int j = 0; j < n_split; ++j)
            ranges.push_back(Range{y, n_split, j, gy0 + j * per, gy0 + std::min(G, (j + 1) * per), slot0});
    }
    std::stable_sort(ranges.begin(), ranges.end(),
                     [](const Range &a, const Range &b) { return a.g1 - a.g0 > b.g1 - b.g0; });
    for (const Range &R : ranges)
        for (int b = 0; b < P.nblk; ++b)
            P.sp_items.push_back(
                GramItem{b, R.subset, R.g0, R.g1, R.slot0 < 0 ? -1 : R.slot0 + (int64_t)b * R.n_split + R.j});
}

// Builds every host table and uploads them into one buffer (synchronously).
static void batch_tables(Ctx &x, BatchPlan &P) {
    TablePack tp;
    const size_t o_off = tp.add(P.off), o_row0 = tp.add(P.row0), o_nrow = tp.add(P.nrow);
    std::vector<SegDesc> sub_segs;
    std::vector<RowRange> stats_r, proj_r;
    for (int y = 0; y < P.n_sub; ++y) {
        const int64_t ny = P.off[y + 1] - P.off[y];
        sub_segs.push_back(SegDesc{P.off[y], ny, 0, y, 0});
        for (int64_t r = 0; r < ny; r += kStatsRows) stats_r.push_back(RowRange{r, std::min(ny, r + kStatsRows), y, 0});
        for (int64_t r = 0; r < ny; r += kProjRows) proj_r.push_back(RowRange{r, std::min(ny, r + kProjRows), y, 0});
    }
    P.n_stats_ranges = (int64_t)stats_r.size();
    P.n_proj_ranges = (int64_t)proj_r.size();
    const size_t o_segs = tp.add(sub_segs), o_sr = tp.add(stats_r), o_pr = tp.add(proj_r);
    // kNN: subset of each 256-row block of operand rows; 512-row query blocks ordered by
    // subset size (largest first, so the longest-running workgroups start first)
    std::vector<int32_t> blk_seg(P.n_pad / kSegBlock), order;
    std::vector<int> by_size(P.n_sub);
    for (int y = 0; y < P.n_sub; ++y) {
        by_size[y] = y;
        for (int64_t r = P.row0[y]; r < P.row0[y + 1]; r += kSegBlock) blk_seg[r / kSegBlock] = y;
    }
    std::stable_sort(by_size.begin(), by_size.end(), [&](int a, int b) { return P.nrow[a] > P.nrow[b]; });
    for (int y : by_size)
        for (int64_t r = P.row0[y]; r < P.row0[y + 1]; r += 512) order.push_back((int32_t)(r / 512));
    const size_t o_bs = tp.add(blk_seg), o_ord = tp.add(order);
    const size_t o_spg = tp.add(P.sp_groups), o_spi = tp.add(P.sp_items), o_spr = tp.add(P.sp_red);
    std::vector<size_t> o_ts, o_td, o_ti, o_tr;
    for (const BatchTile &t : P.tiles) {
        o_ts.push_back(tp.add(t.segs));
        o_td.push_back(tp.add(t.dens));
        o_ti.push_back(tp.add(t.items));
        o_tr.push_back(tp.add(t.red));
    }
    uint8_t *dev = x.get<uint8_t>(P.pre + "tables", tp.bytes.size());
    HIPCHK(hipMemcpyAsync(dev, tp.bytes.data(), tp.bytes.size(), hipMemcpyHostToDevice, x.st));
    x.sync();
    P.d_off = (const int64_t *)(dev + o_off);
    P.d_row0 = (const int64_t *)(dev + o_row0);
    P.d_nrow = (const int64_t *)(dev + o_nrow);
    P.d_sub_segs = (const SegDesc *)(dev + o_segs);
    P.d_stats_ranges = (const RowRange *)(dev + o_sr);
    P.d_proj_ranges = (const RowRange *)(dev + o_pr);
    P.d_blk_seg = (const int32_t *)(dev + o_bs);
    P.d_order = (const int32_t *)(dev + o_ord);
    P.d_sp_groups = (const SegDesc *)(dev + o_spg);
    P.d_sp_items = (const GramItem *)(dev + o_spi);
    P.d_sp_red = (const RedItem *)(dev + o_spr);
    P.d_tile_segs.clear();
    P.d_tile_dens.clear();
    P.d_tile_items.clear();
    P.d_tile_red.clear();
    for (size_t i = 0; i < P.tiles.size(); ++i) {
        P.d_tile_segs.push_back((const SegDesc *)(dev + o_ts[i]));
        P.d_tile_dens.push_back((const RowRange *)(dev + o_td[i]));
        P.d_tile_items.push_back((const GramItem *)(dev + o_ti[i]));
        P.d_tile_red.push_back((const RedItem *)(dev + o_tr[i]));
    }
}

// Device buffers of a batch (grow-only, named for the batch path alone).
struct BatchBufs {
    int64_t *cmap;
    int32_t *gmap;
    const int32_t *sub, *goh;
    double *stats, *mean, *sd, *gram, *colsum, *cov, *w, *e, *var, *load, *tot;
    double *tau, *topw, *y, *x;  // top-p eigensolver: reflector scalars, eigenvalues, T's and A's vectors
    double *fac;                 // and its LU factors and row swaps
    uint8_t *swp;
    int *tinfo;
    rocblas_int *info;
    float *shift;
    void *ta, *tb;
    double *partial;
    double *sp_vals;  // sparse Gram: packed scaled entries, their columns, group records
    uint8_t *sp_cols;
    uint32_t *sp_meta;
    void *xa, *xb, *kwork;
    float *cd;
    uint32_t *ci;
};

static BatchBufs batch_buffers(Ctx &x, const crest_batch_args *b, const BatchPlan &P,
                               const std::vector<int32_t> &gene_of_hvg) {
    const crest_pipeline_args *a = &b->pipe;
    BatchBufs B{};
    const size_t S = (size_t)P.n_sub, d = (size_t)P.d, dp = (size_t)P.d_pad;
    B.cmap = x.get<int64_t>(P.pre + "cmap", P.n);
    B.gmap = x.get<int32_t>(P.pre + "gmap", a->n_genes);
    hipLaunchKernelGGL(k_iota_i64, dim3(1024), dim3(256), 0, x.st, B.cmap, P.n);
    hipLaunchKernelGGL(k_iota_i32, dim3(256), dim3(256), 0, x.st, B.gmap, (int64_t)a->n_genes);
    B.sub = x.upload(P.pre + "sub", a->hvg_of_gene, a->n_genes);
    B.goh = x.upload(P.pre + "goh", gene_of_hvg.data(), d);
    B.stats = x.get<double>(P.pre + "stats", S * a->n_genes * 5);
    B.mean = x.get<double>(P.pre + "mean", S * d);
    B.sd = x.get<double>(P.pre + "sd", S * d);
    B.gram = x.get<double>(P.pre + "gram", S * dp * dp);
    B.colsum = x.get<double>(P.pre + "colsum", S * dp);
    const size_t eb = (size_t)P.eig_batch;
    B.cov = x.get<double>(P.pre + "cov", eb * d * d);
    B.w = x.get<double>(P.pre + "eig_w", eb * d);
    B.e = x.get<double>(P.pre + "eig_e", eb * d);
    B.info = x.get<rocblas_int>(P.pre + "eig_info", S);
    B.tau = x.get<double>(P.pre + "eig_tau", eb * d);
    B.topw = x.get<double>(P.pre + "eig_topw", eb * P.kp);
    B.y = x.get<double>(P.pre + "eig_y", eb * d * P.kp);
    B.x = x.get<double>(P.pre + "eig_x", eb * d * P.kp);
    B.tinfo = x.get<int>(P.pre + "eig_tinfo", eb);
    B.fac = x.get<double>(P.pre + "eig_fac", eb * 3 * d * P.kp);
    B.swp = x.get<uint8_t>(P.pre + "eig_swp", eb * d * P.kp);
    B.var = x.get<double>(P.pre + "var", S * P.kp);
    B.load = x.get<double>(P.pre + "load", S * d * P.kp);
    B.tot = x.get<double>(P.pre + "total", S);
    B.shift = x.get<float>(P.pre + "shift", S * P.kp);
    if (P.sparse) {
        B.ta = B.tb = nullptr;
        B.sp_vals = x.get<double>(P.pre + "sp_vals", (size_t)std::max<int64_t>(P.nnz, 1));
        B.sp_cols = x.get<uint8_t>(P.pre + "sp_cols", (size_t)std::max<int64_t>(P.nnz, 1));
        B.sp_meta = x.get<uint32_t>(P.pre + "sp_meta", sparse_meta_words((int64_t)P.sp_groups.size() * kRowGroupSparse,
                                                                          P.d_pad));
    } else {
        const size_t tile_elems = (size_t)P.max_tile_rows * dp;
        B.ta = x.get<uint8_t>(P.pre + "tile_a", tile_elems * tile_elem_bytes(P.fmt));
        B.tb = (P.fmt == TILE_BF16X2 || P.fmt == TILE_F16X2) ? x.get<uint8_t>(P.pre + "tile_b", tile_elems * 2)
                                                             : nullptr;
        B.sp_vals = nullptr;
        B.sp_cols = nullptr;
        B.sp_meta = nullptr;
    }
    B.partial = x.get<double>(P.pre + "partial", (size_t)std::max<int64_t>(std::max(P.max_slots, P.sp_slots), 1) *
                                                     kGramBlock * kGramBlock);
    const size_t kb = P.kprec == CREST_PREC_FP32 ? 4 : 2;
    B.xa = x.get<uint8_t>(P.pre + "knn_xa", (size_t)P.n_pad * P.kpad * kb);
    B.xb = x.get<uint8_t>(P.pre + "knn_xb", (size_t)P.n_pad * P.kpad * kb);
    B.kwork = x.get<uint8_t>(P.pre + "knn_work", knn_batch_work_bytes(P.n, P.n_pad, P.n_sub, P.kp));
    B.cd = x.get<float>(P.pre + "knn_cd", (size_t)P.n_pad * 2 * P.L);
    B.ci = x.get<uint32_t>(P.pre + "knn_ci", (size_t)P.n_pad * 2 * P.L);
    return B;
}

static DevChunk batch_chunk(const crest_batch_args *b, const BatchPlan &P, const BatchBufs &B) {
    const crest_pipeline_args *a = &b->pipe;
    DevChunk c{};
    c.span_off = a->dev_indptr;
    c.span_cell = nullptr;
    c.first_cell = 0;
    c.n_spans = P.n;
    c.genes = a->dev_genes;
    c.values = a->dev_values;
    c.cell_map = B.cmap;
    c.n_cell_map = P.n;
    c.gene_var = B.gmap;
    c.n_gene_map = (int64_t)a->n_genes;
    c.target_sum = a->target_sum;
    c.log1p = a->log1p ? 1 : 0;
    c.base = 0;
    c.nnz = P.nnz;
    return c;
}

static DevScale batch_scale(const crest_batch_args *b, const BatchPlan &P, const BatchBufs &B) {
    DevScale sc{};
    sc.sub = B.sub;
    sc.scale = b->pipe.scale ? 1 : 0;
    sc.max_value = b->pipe.max_value;
    sc.d = P.d;
    sc.mean = B.mean;
    sc.sd = B.sd;
    return sc;
}

// Stage launches (no host synchronisation inside, so they can be captured).
static void batch_stats(Ctx &x, const crest_batch_args *b, const BatchPlan &P, const BatchBufs &B) {
    if (!b->pipe.scale) return;
    const int64_t ng = (int64_t)b->pipe.n_genes;
    launch_fill_f64(B.stats, (int64_t)P.n_sub * ng * 5, 0.0, x.st);
    launch_gene_stats_ranges(batch_chunk(b, P, B), P.d_sub_segs, P.d_stats_ranges, P.n_stats_ranges, B.stats, (int)ng,
                             x.st);
    hipLaunchKernelGGL(k_mean_std, dim3((P.d + 255) / 256, P.n_sub), dim3(256), 0, x.st, B.stats, B.goh, P.d, 0.0,
                       B.mean, B.sd, ng, P.d_off);
}
static void batch_gram(Ctx &x, const crest_batch_args *b, const BatchPlan &P, const BatchBufs &B) {
    launch_fill_f64(B.gram, (int64_t)P.n_sub * P.d_pad * P.d_pad, 0.0, x.st);
    launch_fill_f64(B.colsum, (int64_t)P.n_sub * P.d_pad, 0.0, x.st);
    const DevChunk c = batch_chunk(b, P, B);
    const DevScale sc = batch_scale(b, P, B);
    if (P.sparse) {
        launch_sparse_pack_groups(c, sc, P.d_pad, P.d_sp_groups, (int64_t)P.sp_groups.size(), B.sp_vals, B.sp_cols,
                                  B.sp_meta, x.st);
        launch_gram_sparse_items(B.sp_vals, B.sp_cols, B.sp_meta, c.span_off, c.base, P.d_sp_groups, P.d_sp_items,
                                 (int64_t)P.sp_items.size(), P.d_pad, B.partial, B.gram, B.colsum, x.st);
        launch_gram_reduce_items(B.partial, P.d_sp_red, (int64_t)P.sp_red.size(), P.d_pad, B.gram, x.st);
        return;
    }
    for (size_t i = 0; i < P.tiles.size(); ++i) {
        const BatchTile &t = P.tiles[i];
        launch_densify_ranges(c, sc, P.fmt, B.ta, B.tb, P.d_pad, P.d_tile_segs[i], P.d_tile_dens[i],
                              (int64_t)t.dens.size(), B.colsum, B.gram, x.st);
        launch_gram_items(P.fmt, B.ta, B.tb, P.d_pad, P.d_tile_items[i], (int64_t)t.items.size(), P.grid, B.partial,
                          B.gram, x.st);
        launch_gram_reduce_items(B.partial, P.d_tile_red[i], (int64_t)t.red.size(), P.d_pad, B.gram, x.st);
    }
}
static void batch_project(Ctx &x, const crest_batch_args *b, const BatchPlan &P, const BatchBufs &B) {
    launch_shift_fill_batch(B.colsum, P.d_pad, B.load, P.d, P.kp, P.d_off, P.n_sub, P.max_rows, b->pipe.dev_x_pca,
                            B.shift, x.st);
    launch_project_ranges(batch_chunk(b, P, B), batch_scale(b, P, B), B.load, P.kp, P.d_sub_segs, P.d_proj_ranges,
                          P.n_proj_ranges, b->pipe.dev_x_pca, x.st);
}
static KnnOrder batch_knn_prep(Ctx &x, const crest_batch_args *b, const BatchPlan &P, const BatchBufs &B) {
    return launch_knn_prep_batch(P.kprec, b->pipe.dev_x_pca, P.n, P.d_off, P.n_sub, P.max_rows, P.d_row0, P.n_pad,
                                 P.max_pad, P.kp, P.kpad, B.kwork, B.xa, B.xb, x.st);
}
static void batch_knn_screen(Ctx &x, const BatchPlan &P, const BatchBufs &B) {
    const KnnSeg sg{P.d_row0, P.d_nrow, P.d_blk_seg, P.d_order};
    if (!launch_knn_screen_batch(P.kprec, B.xa, B.xb, P.n_pad, P.kpad, P.L, sg, B.cd, B.ci, x.st))
        value_error("no kNN kernel for this width");
}
static void batch_knn_rerank(Ctx &x, const crest_batch_args *b, const BatchPlan &P, const BatchBufs &B,
                             const KnnOrder &o) {
    // the certificate lists uncertified queries by data row (a fill kernel clears the
    // count, so the launch can be captured)
    uint32_t *cnt = x.get<uint32_t>(P.pre + "cert_cnt", 1);
    uint32_t *unc = x.get<uint32_t>(P.pre + "cert_rows", (size_t)P.n);
    launch_fill_u32(cnt, 1, 0u, x.st);
    const KnnCert ct = knn_cert(P.kprec, P.kp, P.kpad, B.cd, cnt, unc, P.n);
    launch_knn_rerank(b->pipe.dev_x_pca, P.n_pad, b->pipe.dev_x_pca, P.n_pad, P.kp, o, 1, B.ci, 2 * P.L, P.kn,
                      b->pipe.dev_knn_idx, b->pipe.dev_knn_dist, x.st, P.d_blk_seg, P.d_off, &ct);
}

// After the batch's kNN: the queries its certificate did not cover, rechecked subset by
// subset in FP32 among their own subset's cells (indices within the subset): one upload
// of every subset's local rows, the subsets' threshold scans launched back to back,
// one read of the list counts, and the GEMM recheck only for overflowed lists.
static void batch_knn_recheck(Ctx &x, const crest_batch_args *b, const BatchPlan &P, const BatchBufs &BB) {
    x.knn_checked = P.n;
    x.knn_rechecked = x.knn_uncertified = 0;
    const uint32_t m = read_count(x, x.get<uint32_t>(P.pre + "cert_cnt", 1));
    if (!m) return;
    std::vector<uint32_t> rows(m);
    x.download(rows.data(), x.get<uint32_t>(P.pre + "cert_rows", (size_t)P.n), m);
    x.sync();
    std::sort(rows.begin(), rows.end());
    // subsets in order: [a, e) of the sorted rows are subset y's, as local rows
    struct Part {
        int y;
        size_t a, e;
    };
    std::vector<Part> parts;
    std::vector<uint32_t> loc(m);
    int64_t nc_max = 0;
    for (size_t a = 0; a < rows.size();) {
        const int y = (int)(std::upper_bound(P.off.begin(), P.off.end(), (int64_t)rows[a]) - P.off.begin()) - 1;
        size_t e = a;
        while (e < rows.size() && (int64_t)rows[e] < P.off[y + 1]) {
            loc[e] = (uint32_t)(rows[e] - P.off[y]);
            ++e;
        }
        parts.push_back({y, a, e});
        nc_max = std::max(nc_max, P.off[y + 1] - P.off[y]);
        a = e;
    }
    uint32_t *dl = x.get<uint32_t>("knnr_rows", (size_t)m);
    HIPCHK(hipMemcpyAsync(dl, loc.data(), (size_t)m * sizeof(uint32_t), hipMemcpyHostToDevice, x.st));
    x.sync();
    float *xp = b->pipe.dev_x_pca;
    uint32_t *oi = b->pipe.dev_knn_idx;
    float *od = b->pipe.dev_knn_dist;
    const int p = P.kp, k = P.kn;
    if (p > 256) {  // rows too wide for the threshold scan: the GEMM recheck per subset
        for (const Part &t : parts)
            x.knn_uncertified += knn_recheck_gemm(x, xp + P.off[t.y] * p, P.off[t.y + 1] - P.off[t.y], nullptr, p, k,
                                                  dl + t.a, (int64_t)(t.e - t.a), oi + P.off[t.y] * k,
                                                  od + P.off[t.y] * k);
        x.knn_rechecked = m;
        return;
    }
    const RecheckBufs B = recheck_bufs(x, m, nc_max, p, k);
    if (p <= 96 && (knn_scan_kind() == 0 || knn_scan_kind() == 3)) {
        // all subsets at once: every row centred on its subset's mean (the screen's), the
        // rechecked rows gathered, one MFMA threshold test over a table of (subset, up to
        // 128 queries, the subset's candidates), one FP64 rank, indices placed within
        // the subsets
        std::vector<uint32_t> grow(m), ibase(m);
        std::vector<KnnScanBlock> blk;
        int64_t max_slices = 1;
        for (const Part &t : parts) {
            const int64_t c0 = P.off[t.y], c1 = P.off[t.y + 1];
            const int64_t slice = round_up(std::max<int64_t>(4096, ceil_div(c1 - c0, 16)), 64);
            max_slices = std::max(max_slices, ceil_div(c1 - c0, slice));
            for (size_t a = t.a; a < t.e; ++a) {
                grow[a] = rows[a];
                ibase[a] = (uint32_t)c0;
            }
            for (size_t a = t.a; a < t.e; a += 128)
                blk.push_back({(int64_t)a, c0, c1, slice, (int32_t)std::min<size_t>(128, t.e - a), 0});
        }
        uint32_t *dg = x.upload<uint32_t>("knnr_grows", grow.data(), m);
        uint32_t *db = x.upload<uint32_t>("knnr_ibase", ibase.data(), m);
        KnnScanBlock *dk = x.upload<KnnScanBlock>("knnr_blocks", blk.data(), blk.size());
        x.sync();  // the host vectors above are pageable and end with this block
        float *xall = x.get<float>("knnr_xall", (size_t)P.n * p);
        float *nall = x.get<float>("knnr_nall", (size_t)P.n);
        const float *means = knn_batch_means(BB.kwork, P.n, P.n_pad, P.n_sub, p);
        launch_knn_centre_seg(xp, P.n, P.d_off, P.n_sub, p, means, xall, x.st);
        launch_knn_sqnorm(xall, P.n, p, nall, x.st);
        launch_gather_rows(xp, p, dg, m, B.xsub, x.st);
        launch_gather_rows(xall, p, dg, m, B.xsubc, x.st);
        launch_gather_rows(nall, 1, dg, m, B.qn, x.st);
        launch_fill_u32(B.cnt, m, 0u, x.st);
        launch_fill_u32(B.cand, (int64_t)m * kRecheckCap, 0xFFFFFFFFu, x.st);
        launch_knn_scan_thresholds(od, dg, B.xsubc, m, k, p, B.thr, x.st);
        launch_knn_thr_mfma_blocks(xall, nall, p, B.xsubc, B.qn, B.thr, dg, dk, (int)blk.size(), max_slices, B.cnt,
                                   B.cand, kRecheckCap, x.st);
        launch_knn_rerank_sort(B.xsub, xp, p, B.cand, kRecheckCap, k, 0, m, B.ti, B.td, x.st);
        launch_knn_place(B.ti, B.td, m, k, k, dg, 0, oi, od, x.st, db);
        x.check_launch();
    } else {
        for (const Part &t : parts)
            knn_recheck_launch(x, recheck_slice(B, (int64_t)t.a, p, k), xp + P.off[t.y] * p,
                               P.off[t.y + 1] - P.off[t.y], nullptr, p, k, dl + t.a, (int64_t)(t.e - t.a),
                               oi + P.off[t.y] * k, od + P.off[t.y] * k, nullptr);
    }
    const std::vector<int64_t> over = recheck_overflows(x, B, m);
    for (const Part &t : parts) {
        std::vector<uint32_t> ov;
        for (int64_t r : over)
            if ((size_t)r >= t.a && (size_t)r < t.e) ov.push_back(loc[(size_t)r]);
        if (ov.empty()) continue;
        uint32_t *dv = x.get<uint32_t>("knnr_over", ov.size());
        HIPCHK(hipMemcpyAsync(dv, ov.data(), ov.size() * sizeof(uint32_t), hipMemcpyHostToDevice, x.st));
        x.sync();
        x.knn_uncertified += knn_recheck_gemm(x, xp + P.off[t.y] * p, P.off[t.y + 1] - P.off[t.y], nullptr, p, k, dv,
                                              (int64_t)ov.size(), oi + P.off[t.y] * k, od + P.off[t.y] * k);
    }
    x.knn_rechecked = m;
}
// Covariance, batched FP64 eigensolve and selection, eig_batch subsets per call
// (rocSOLVER: eager, never captured).
//
// Subsets with fewer cells than genes have a covariance of rank below n; with the
// top-p solver, a subset whose padded cell count m (256, 512, 1,024 or 1,536) is below
// d takes the n x n dual instead: Zc = its centred scaled rows in FP64 (m x d, rows
// beyond n zero), K = Zc Zc^T (rocBLAS dgemm), the top p eigenpairs (lambda, u) of K,
// then C's eigenpairs (lambda / (n - 1), Zc^T u normalised) — the same nonzero
// eigenpairs as C = Zc^T Zc / (n - 1) at O(m^3) instead of O(d^3).  The trace comes
// from the Gram diagonal with the covariance's own arithmetic.  CREST_BATCH_DUAL=0
// keeps every subset on the d x d covariance.
// Tridiagonalization of ng matrices of order n (lower, LAPACK dsytrd layout): the
// one-workgroup-per-matrix kernel for groups of at least 48 — one matrix of order 2,000
// takes ~300 ms on its workgroup, a group side by side until HBM saturates (162 take
// ~480 ms), against ~6.2 ms per matrix for rocSOLVER's batched dsytrd — and rocSOLVER
// for smaller groups.  CREST_SYTRD=wg|rocsolver forces either.
static void batch_sytrd(Ctx &x, const std::string &pre, double *A, int n, int ng, double *Dg, double *Eg,
                        double *tau, hipStream_t st, rocblas_handle blas) {
    const char *e = std::getenv("CREST_SYTRD");
    const std::string mode = e ? e : "auto";
    const bool wg = sytrd_wg_supported(n) && (mode == "wg" || (mode != "rocsolver" && ng >= 48));
    const int64_t nn = (int64_t)n * n;
    if (!wg) {
        BLASCHK(rocsolver_dsytrd_strided_batched(blas, rocblas_fill_lower, n, A, n, nn, Dg, n, Eg, n, tau, n, ng));
        return;
    }
    double *W = x.get<double>(pre + "sytrd_w", (size_t)ng * sytrd_wg_scratch(n));
    sytrd_set_attrs();
    launch_sytrd_lower(A, nn, n, Dg, n, Eg, n, tau, n, W, n, ng, st);
    x.check_launch();
}

static void batch_eig_impl(Ctx &x, const crest_batch_args *b, const BatchPlan &P, const BatchBufs &B);
// Any error leaves no work of this call queued on the side stream (its uploads read
// host vectors that die with the call).
static void batch_eig(Ctx &x, const crest_batch_args *b, const BatchPlan &P, const BatchBufs &B) {
    try {
        batch_eig_impl(x, b, P, B);
    } catch (...) {
        (void)hipStreamSynchronize(x.side);
        throw;
    }
}

// The batch's covariances by the filtered subspace iteration side by side (default
// wherever chefsi_topp would solve one of them; CREST_BATCH_EIG=tri keeps the
// tridiagonalization of k_sytrd_lower / rocSOLVER and the top-p kernels).
static bool batch_chefsi_enabled(int d, int kp) {
    const char *e = std::getenv("CREST_BATCH_EIG");
    if (e && std::string(e) == "tri") return false;
    return chefsi_block(d, kp) > 0;
}

static void batch_eig_impl(Ctx &x, const crest_batch_args *b, const BatchPlan &P, const BatchBufs &B) {
    const bool topp = use_topp(P.d, P.kp, (int)std::min<int64_t>(P.eig_batch, P.n_sub));
    if (topp) tri_set_attrs();
    const char *dual_env = std::getenv("CREST_BATCH_DUAL");
    // the dense centred rows are staged in LDS, four rows of d doubles per workgroup
    const bool dual_on = topp && P.d <= 4096 && !(dual_env && std::string(dual_env) == "0");
    static const int kBuckets[] = {256, 512, 1024, 1536};
    std::vector<int32_t> primal;
    std::vector<int32_t> dual[4];
    for (int y = 0; y < P.n_sub; ++y) {
        const int64_t ny = P.off[y + 1] - P.off[y];
        int bk = -1;
        if (dual_on)
            for (int i = 0; i < 4; ++i)
                if (ny <= kBuckets[i] && kBuckets[i] < P.d && P.kp <= kBuckets[i] && tri_topp_supported(kBuckets[i])) {
                    bk = i;
                    break;
                }
        if (bk < 0)
            primal.push_back(y);
        else
            dual[bk].push_back(y);
    }
    int32_t *ids_dev = x.get<int32_t>(P.pre + "eig_ids", (size_t)P.n_sub);
    std::vector<int32_t> all_ids;  // primal ids, then each bucket's, in one upload
    all_ids.insert(all_ids.end(), primal.begin(), primal.end());
    for (auto &v : dual) all_ids.insert(all_ids.end(), v.begin(), v.end());
    HIPCHK(hipMemcpyAsync(ids_dev, all_ids.data(), all_ids.size() * sizeof(int32_t), hipMemcpyHostToDevice, x.st));
    const int64_t dd = (int64_t)P.d * P.d;
    if (topp) launch_fill_u32(reinterpret_cast<uint32_t *>(B.info), P.n_sub, 0u, x.st);  // a fallback sets its own
    // a subset whose inverse iteration did not converge: full eigensolve of its covariance
    auto fallback = [&](int y) {
        launch_cov_batch(B.gram + (int64_t)y * P.d_pad * P.d_pad, B.colsum + (int64_t)y * P.d_pad, P.d, P.d_pad,
                         P.d_off + y, 1, B.cov, B.tot + y, x.st);
        BLASCHK(rocsolver_dsyevd(x.blas, rocblas_evect_original, rocblas_fill_upper, P.d, B.cov, P.d, B.w, B.e,
                                 B.info + y));
        launch_eig_select_batch(B.cov, B.w, P.d, P.kp, 1, B.var + (int64_t)y * P.kp, B.load + (int64_t)y * P.d * P.kp,
                                x.st);
        x.check_launch();
    };
    // ---- the n x n duals, bucket by bucket, on the side stream: they run while the
    // covariances below are tridiagonalized (one workgroup per matrix leaves CUs idle).
    // Every bucket and group has buffers of its own; they write only their own
    // subsets' variance, loadings and totals.
    struct DualGroup {
        const std::vector<int32_t> *L;
        int g0, ng;
        int *tinfo;
    };
    std::vector<DualGroup> dgs;
    std::vector<std::vector<SegDesc>> seg_host;  // kept until the join (async uploads)
    bool any_dual = false;
    for (auto &v : dual) any_dual = any_dual || !v.empty();
    if (any_dual) {
        HIPCHK(hipEventRecord(x.fork, x.st));
        HIPCHK(hipStreamWaitEvent(x.side, x.fork, 0));
        const DevChunk c = batch_chunk(b, P, B);
        const DevScale sc = batch_scale(b, P, B);
        int64_t base = (int64_t)primal.size();
        for (int bi = 0; bi < 4; ++bi) {
            const int m = kBuckets[bi];
            const std::vector<int32_t> &L = dual[bi];
            for (int g0 = 0; g0 < (int)L.size(); g0 += (int)P.eig_batch) {
                const int ng = (int)std::min<int64_t>(P.eig_batch, (int64_t)L.size() - g0);
                const int32_t *ids = ids_dev + base + g0;
                const std::string pf = P.pre + "dual" + std::to_string(bi) + "_" + std::to_string(g0) + "_";
                seg_host.emplace_back(ng);
                std::vector<SegDesc> &segs = seg_host.back();
                for (int i = 0; i < ng; ++i) {
                    const int y = L[g0 + i];
                    segs[i] = SegDesc{P.off[y], P.off[y + 1] - P.off[y], 0, y, 0};
                }
                SegDesc *segs_dev = x.get<SegDesc>(pf + "segs", (size_t)ng);
                HIPCHK(hipMemcpyAsync(segs_dev, segs.data(), segs.size() * sizeof(SegDesc), hipMemcpyHostToDevice,
                                      x.side));
                const int64_t mm = (int64_t)m * m, md = (int64_t)m * P.d, mp = (int64_t)m * P.kp;
                double *zc = x.get<double>(pf + "zc", (size_t)(ng * md));
                double *K = x.get<double>(pf + "k", (size_t)(ng * mm));
                double *V = x.get<double>(pf + "v", (size_t)ng * P.d * P.kp);
                double *ev = x.get<double>(pf + "ev", (size_t)ng * P.kp);
                double *Dg = x.get<double>(pf + "d", (size_t)ng * m);
                double *Eg = x.get<double>(pf + "e", (size_t)ng * m);
                double *tau = x.get<double>(pf + "tau", (size_t)ng * m);
                double *tw = x.get<double>(pf + "tw", (size_t)ng * P.kp);
                double *ty = x.get<double>(pf + "ty", (size_t)(ng * mp));
                double *tx = x.get<double>(pf + "tx", (size_t)(ng * mp));
                double *fac = x.get<double>(pf + "fac", (size_t)(3 * ng * mp));
                uint8_t *swp = x.get<uint8_t>(pf + "swp", (size_t)(ng * mp));
                int *tinfo = x.get<int>(pf + "tinfo", (size_t)ng);
                launch_dense_centred(c, sc, segs_dev, ng, m, B.colsum, P.d_pad, zc, x.side);
                const double one = 1.0, zero = 0.0;
                BLASCHK(rocblas_dgemm_strided_batched(x.blas_side, rocblas_operation_transpose, rocblas_operation_none,
                                                      m, m, P.d, &one, zc, P.d, md, zc, P.d, md, &zero, K, m, mm, ng));
                batch_sytrd(x, pf, K, m, ng, Dg, Eg, tau, x.side, x.blas_side);
                launch_fill_u32(reinterpret_cast<uint32_t *>(tinfo), ng, 0u, x.side);
                launch_tri_topp(K, mm, m, tau, m, Dg, m, Eg, m, m, P.kp, ng, tw, ty, tx, fac, swp, tinfo, x.side);
                BLASCHK(rocblas_dgemm_strided_batched(x.blas_side, rocblas_operation_none, rocblas_operation_none, P.d,
                                                      P.kp, m, &one, zc, P.d, md, tx, m, mp, &zero, V, P.d,
                                                      (int64_t)P.d * P.kp, ng));
                launch_dual_finish(V, tw, P.d, P.kp, P.d_off, ids, ng, ev, x.side);
                launch_eig_select_batch(V, ev, P.d, P.kp, ng, B.var, B.load, x.side, P.kp, ids);
                launch_trace_gram(B.gram, B.colsum, P.d, P.d_pad, P.d_off, ids, ng, B.tot, x.side);
                x.check_launch();
                dgs.push_back(DualGroup{&L, g0, ng, tinfo});
            }
            base += (int64_t)L.size();
        }
        HIPCHK(hipEventRecord(x.join, x.side));
    }
    // ---- the d x d covariances
    const bool chb = batch_chefsi_enabled(P.d, P.kp);
    if (chb) launch_fill_u32(reinterpret_cast<uint32_t *>(B.info), P.n_sub, 0u, x.st);
    for (int g0 = 0; g0 < (int)primal.size(); g0 += (int)P.eig_batch) {
        const int ng = (int)std::min<int64_t>(P.eig_batch, (int64_t)primal.size() - g0);
        const int32_t *ids = ids_dev + g0;
        launch_cov_batch(B.gram, B.colsum, P.d, P.d_pad, P.d_off, ng, B.cov, B.tot, x.st, ids);
        x.check_launch();
        if (chb) {  // the filtered subspace iteration, all ng side by side
            const std::vector<char> ok = chefsi_topp_batch(x, P.pre, B.cov, ng, P.d, P.kp, ids, B.var, B.load);
            for (int i = 0; i < ng; ++i)
                if (!ok[i]) fallback(primal[g0 + i]);
            continue;
        }
        if (topp) {
            batch_sytrd(x, P.pre, B.cov, P.d, ng, B.w, B.e, B.tau, x.st, x.blas);
            launch_fill_u32(reinterpret_cast<uint32_t *>(B.tinfo), ng, 0u, x.st);
            launch_tri_topp(B.cov, dd, P.d, B.tau, P.d, B.w, P.d, B.e, P.d, P.d, P.kp, ng, B.topw, B.y, B.x, B.fac,
                            B.swp, B.tinfo, x.st);
            launch_eig_select_batch(B.x, B.topw, P.d, P.kp, ng, B.var, B.load, x.st, P.kp, ids);
            x.check_launch();
            std::vector<int> hi(ng);
            x.download(hi.data(), B.tinfo, ng);
            x.sync();
            for (int i = 0; i < ng; ++i)
                if (hi[i]) fallback(primal[g0 + i]);
            continue;
        }
        // without the top-p solver no subset takes the dual, so primal[g0 + i] = g0 + i
        if (ng == 1)
            BLASCHK(rocsolver_dsyevd(x.blas, rocblas_evect_original, rocblas_fill_upper, P.d, B.cov, P.d, B.w, B.e,
                                     B.info + g0));
        else
            BLASCHK(rocsolver_dsyevd_strided_batched(x.blas, rocblas_evect_original, rocblas_fill_upper, P.d, B.cov,
                                                     P.d, dd, B.w, P.d, B.e, P.d, B.info + g0, ng));
        launch_eig_select_batch(B.cov, B.w, P.d, P.kp, ng, B.var, B.load, x.st, 0, ids);
        x.check_launch();
    }
    // ---- join: the duals' convergence flags, then their fallbacks on the main stream
    if (!dgs.empty()) {
        HIPCHK(hipStreamWaitEvent(x.st, x.join, 0));
        for (const DualGroup &g : dgs) {
            std::vector<int> hi(g.ng);
            x.download(hi.data(), g.tinfo, g.ng);
            x.sync();
            for (int i = 0; i < g.ng; ++i)
                if (hi[i]) fallback((*g.L)[g.g0 + i]);
        }
        x.sync();
    }
}

// gram_rows from the autotuner: the first tile is densified once, then the Gram
// launch is timed for each candidate range length (after one untimed run) and the
// fastest is kept, cached per shape in-process and in CREST_TUNE_CACHE.
static std::string tune_key(const BatchPlan &P) {
    int64_t lg = 0;
    while (((int64_t)1 << lg) < P.max_rows) ++lg;
    return "gram_rows:fmt" + std::to_string(P.fmt) + ":dpad" + std::to_string(P.d_pad) + ":tile" +
           std::to_string(P.tile_rows) + ":maxrows2^" + std::to_string(lg) + ":cu" + std::to_string(P.grid);
}
static int64_t tune_gram_rows(Ctx &x, const crest_batch_args *b, BatchPlan &P, const BatchBufs &B0, bool &ran) {
    ran = false;
    const std::string key = tune_key(P);
    int64_t v = 0;
    static const int64_t cand[] = {2048, 4096, 8192, 16384, 1 << 20};
    std::vector<int64_t> cands;
    for (int64_t c : cand)
        if (c < round_up(P.max_rows, 16) || cands.empty()) cands.push_back(c);
    if (cands.size() == 1) cands[0] = cand[sizeof(cand) / sizeof(cand[0]) - 1];
    if (cands.size() > 1) {
        ran = true;
        BatchTile t = P.tiles[0];
        // densify the first tile (its column sums and Gram diagonal are cleared again
        // before the real pass)
        batch_stats(x, b, P, B0);
        launch_fill_f64(B0.gram, (int64_t)P.n_sub * P.d_pad * P.d_pad, 0.0, x.st);
        launch_fill_f64(B0.colsum, (int64_t)P.n_sub * P.d_pad, 0.0, x.st);
        launch_densify_ranges(batch_chunk(b, P, B0), batch_scale(b, P, B0), P.fmt, B0.ta, B0.tb, P.d_pad,
                              P.d_tile_segs[0], P.d_tile_dens[0], (int64_t)t.dens.size(), B0.colsum, B0.gram, x.st);
        float best = 0.f;
        for (int64_t c : cands) {
            plan_items(t, P.nblk, c);
            TablePack tp;
            const size_t oi = tp.add(t.items), orr = tp.add(t.red);
            uint8_t *dev = x.get<uint8_t>("bt_tune_tables", tp.bytes.size());
            HIPCHK(hipMemcpyAsync(dev, tp.bytes.data(), tp.bytes.size(), hipMemcpyHostToDevice, x.st));
            double *part = x.get<double>("bt_tune_partial", (size_t)std::max<int64_t>(t.n_slots, 1) * kGramBlock *
                                                               kGramBlock);
            float ms = 0.f;
            for (int rep = 0; rep < 2; ++rep) {
                HIPCHK(hipEventRecord(x.ev[10], x.st));
                launch_gram_items(P.fmt, B0.ta, B0.tb, P.d_pad, (const GramItem *)(dev + oi), (int64_t)t.items.size(),
                                  P.grid, part, B0.gram, x.st);
                launch_gram_reduce_items(part, (const RedItem *)(dev + orr), (int64_t)t.red.size(), P.d_pad, B0.gram,
                                         x.st);
                HIPCHK(hipEventRecord(x.ev[11], x.st));
                x.check_launch();
                HIPCHK(hipEventSynchronize(x.ev[11]));
                HIPCHK(hipEventElapsedTime(&ms, x.ev[10], x.ev[11]));
            }
            if (v == 0 || ms < best) {
                best = ms;
                v = c;
            }
        }
    } else {
        v = cands[0];
    }
    tune_store(key, v);
    return v;
}

static std::string batch_key(const crest_batch_args *b, const BatchPlan &P) {
    std::string k;
    auto put = [&](const void *p, size_t n) { k.append((const char *)p, n); };
    put(&b->pipe, sizeof(crest_pipeline_args));
    put(b->pipe.hvg_of_gene, b->pipe.n_genes * sizeof(int32_t));
    put(P.off.data(), P.off.size() * sizeof(int64_t));
    const int64_t extra[5] = {P.tile_rows, P.gram_rows, P.eig_batch, (int64_t)P.grid, (int64_t)P.sparse};
    put(extra, sizeof(extra));
    const double cs = knn_cert_scale();  // the certificate's bound is a kernel argument of the graph
    put(&cs, sizeof(cs));
    return k;
}

static void run_batch(Ctx &x, const crest_batch_args *b, double *variance, double *loadings, double *total,
                      crest_batch_result *res) {
    const crest_pipeline_args *a = &b->pipe;
    BatchPlan P;
    P.n_sub = (int)b->n_subsets;
    P.n = (int64_t)a->n_cells;
    HIPCHK(hipMemcpy(&P.nnz, a->dev_indptr + P.n, 8, hipMemcpyDeviceToHost));
    P.d = (int)a->d;
    P.d_pad = (int)round_up(P.d, kGramBlock);
    P.kp = (int)std::min<size_t>(a->n_comps, a->d);
    P.kn = (int)a->k;
    P.fmt = tile_fmt(a->prec_gram);
    P.kprec = a->prec_knn == CREST_PREC_FP64 ? CREST_PREC_FP32 : a->prec_knn;
    P.L = P.kn <= 16 ? 16 : P.kn <= 32 ? 32 : 64;
    P.kpad = knn_kpad(P.kprec, P.kp);
    if (P.kpad < 0) value_error("too many components for the GPU kNN screen");
    P.nblk = (P.d_pad / kGramBlock) * (P.d_pad / kGramBlock + 1) / 2;
    P.off.assign(b->subset_off, b->subset_off + P.n_sub + 1);
    P.row0.assign(P.n_sub + 1, 0);
    P.nrow.assign(P.n_sub, 0);
    for (int y = 0; y < P.n_sub; ++y) {
        const int64_t ny = P.off[y + 1] - P.off[y];
        P.nrow[y] = ny;
        P.row0[y + 1] = P.row0[y] + knn_rows_padded(ny);
        P.max_rows = std::max(P.max_rows, ny);
        P.max_pad = std::max(P.max_pad, knn_rows_padded(ny));
    }
    P.n_pad = P.row0[P.n_sub];
    P.tile_rows = b->tile_rows ? round_up((int64_t)b->tile_rows, 16) : 524288;
    P.eig_batch = b->eig_batch ? (int64_t)b->eig_batch : 256;
    int n_cu = 0;
    HIPCHK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, x.dev));
    P.grid = (int)round_up((int64_t)n_cu * gram_items_occupancy(P.fmt), 8);
    P.tiles = plan_tiles(P.off, P.tile_rows);
    for (const BatchTile &t : P.tiles) P.max_tile_rows = std::max(P.max_tile_rows, t.rows_pad);

    const std::vector<int32_t> gene_of_hvg = hvg_genes(a->hvg_of_gene, a->n_genes, P.d);

    // Gram row ranges: given, cached by the autotuner for this shape, or tuned now (the
    // tuner needs the first tile's tables and buffers sized for the smallest candidate)
    bool tuned_now = false;
    P.pre = b->use_graph ? "bg_" : "bt_";
    P.gram_rows = b->gram_rows ? round_up((int64_t)b->gram_rows, 16) : 0;
    auto plan_all = [&](int64_t rows) {
        P.max_slots = 0;
        for (BatchTile &t : P.tiles) {
            plan_items(t, P.nblk, rows);
            P.max_slots = std::max(P.max_slots, t.n_slots);
        }
    };
    P.sparse = gram_sparse_enabled(a->prec_gram, P.d_pad);
    if (P.sparse) {  // gram_rows: rows per sparse Gram range (no tuning: one range per subset up to 16,384 rows)
        const int64_t rg = P.gram_rows ? std::max<int64_t>(1, P.gram_rows / kRowGroupSparse) : 256;
        P.gram_rows = rg * kRowGroupSparse;
        plan_sparse(P, rg);
        sparse_set_attrs();
    } else if (!P.gram_rows && !tune_lookup(tune_key(P), P.gram_rows)) {
        // a new shape: time the candidates (a graph captured for other arguments goes,
        // since the buffers it reads may be resized here)
        if (b->use_graph) x.drop_graphs();
        plan_all(2048);
        densify_set_attrs();
        batch_tables(x, P);
        BatchBufs B0 = batch_buffers(x, b, P, gene_of_hvg);
        P.gram_rows = tune_gram_rows(x, b, P, B0, tuned_now);
    }
    if (!P.sparse) plan_all(P.gram_rows);
    densify_set_attrs();
    const std::string key = batch_key(b, P);
    // graphs are replayed only by graph-mode calls with identical arguments; eager calls
    // use their own buffers and leave them alone
    const bool replay = b->use_graph && x.graph_key == key && x.graph[0] && x.graph[1];
    if (b->use_graph && !replay) x.drop_graphs();
    batch_tables(x, P);
    const BatchBufs B = batch_buffers(x, b, P, gene_of_hvg);
    (void)x.get<uint32_t>(P.pre + "cert_cnt", 1);  // the certificate's buffers exist before any capture
    (void)x.get<uint32_t>(P.pre + "cert_rows", (size_t)P.n);
    x.sync();

    hipEvent_t *ev = x.ev;
    float ms[8] = {};
    if (b->use_graph) {
        if (!replay) {
            KnnOrder o{};
            for (int gi = 0; gi < 2; ++gi) {
                hipGraph_t g = nullptr;
                HIPCHK(hipStreamBeginCapture(x.st, hipStreamCaptureModeRelaxed));
                try {
                    if (gi == 0) {
                        batch_stats(x, b, P, B);
                        batch_gram(x, b, P, B);
                    } else {
                        batch_project(x, b, P, B);
                        o = batch_knn_prep(x, b, P, B);
                        batch_knn_screen(x, P, B);
                        batch_knn_rerank(x, b, P, B, o);
                    }
                } catch (...) {
                    hipStreamEndCapture(x.st, &g);
                    if (g) hipGraphDestroy(g);
                    throw;
                }
                HIPCHK(hipStreamEndCapture(x.st, &g));
                const hipError_t e = hipGraphInstantiate(&x.graph[gi], g, nullptr, nullptr, 0);
                hipGraphDestroy(g);
                if (e != hipSuccess) {
                    x.drop_graphs();
                    fail(CREST_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
                }
            }
            x.graph_key = key;
        }
        HIPCHK(hipEventRecord(ev[0], x.st));
        HIPCHK(hipGraphLaunch(x.graph[0], x.st));
        HIPCHK(hipEventRecord(ev[1], x.st));
        batch_eig(x, b, P, B);
        HIPCHK(hipEventRecord(ev[2], x.st));
        HIPCHK(hipGraphLaunch(x.graph[1], x.st));
        HIPCHK(hipEventRecord(ev[3], x.st));
        x.sync();
        for (int i = 0; i < 3; ++i) HIPCHK(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
        res->ms_stats = 0.0;
        res->ms_gram = ms[0];
        res->ms_eig = ms[1];
        res->ms_project = ms[2];
        res->ms_knn_prep = res->ms_knn_screen = res->ms_knn_rerank = 0.0;
        float tot = 0.f;
        HIPCHK(hipEventElapsedTime(&tot, ev[0], ev[3]));
        res->ms_total = tot;
    } else {
        HIPCHK(hipEventRecord(ev[0], x.st));
        batch_stats(x, b, P, B);
        HIPCHK(hipEventRecord(ev[1], x.st));
        batch_gram(x, b, P, B);
        HIPCHK(hipEventRecord(ev[2], x.st));
        batch_eig(x, b, P, B);
        HIPCHK(hipEventRecord(ev[3], x.st));
        batch_project(x, b, P, B);
        HIPCHK(hipEventRecord(ev[4], x.st));
        const KnnOrder o = batch_knn_prep(x, b, P, B);
        HIPCHK(hipEventRecord(ev[5], x.st));
        batch_knn_screen(x, P, B);
        HIPCHK(hipEventRecord(ev[6], x.st));
        batch_knn_rerank(x, b, P, B, o);
        HIPCHK(hipEventRecord(ev[7], x.st));
        x.check_launch();
        x.sync();
        for (int i = 0; i < 7; ++i) HIPCHK(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
        res->ms_stats = ms[0];
        res->ms_gram = ms[1];
        res->ms_eig = ms[2];
        res->ms_project = ms[3];
        res->ms_knn_prep = ms[4];
        res->ms_knn_screen = ms[5];
        res->ms_knn_rerank = ms[6];
        float tot = 0.f;
        HIPCHK(hipEventElapsedTime(&tot, ev[0], ev[7]));
        res->ms_total = tot;
    }
    {
        const auto t0 = std::chrono::steady_clock::now();
        batch_knn_recheck(x, b, P, B);
        res->ms_total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() *
                         (x.knn_rechecked ? 1.0 : 0.0);
    }
    res->k_pcs = (size_t)P.kp;
    res->k_nn = (size_t)P.kn;
    res->tile_rows = (size_t)P.tile_rows;
    res->gram_rows = (size_t)P.gram_rows;
    res->graph_replayed = replay ? 1 : 0;
    res->gram_tuned = tuned_now ? 1 : 0;
    // results
    std::vector<rocblas_int> info(P.n_sub);
    x.download(info.data(), B.info, P.n_sub);
    std::vector<double> hv((size_t)P.n_sub * P.kp), hl((size_t)P.n_sub * P.d * P.kp);
    x.download(hv.data(), B.var, hv.size());
    x.download(hl.data(), B.load, hl.size());
    x.download(total, B.tot, P.n_sub);
    x.sync();
    for (int y = 0; y < P.n_sub; ++y)
        if (info[y] != 0)
            fail(CREST_E_RUNTIME, "eigensolver did not converge for subset " + std::to_string(y) + " (info " +
                                      std::to_string(info[y]) + ")");
    const size_t nc = b->pipe.n_comps;
    for (int y = 0; y < P.n_sub; ++y)
        for (int j = 0; j < P.kp; ++j) {
            variance[(size_t)y * nc + j] = hv[(size_t)y * P.kp + j];
            for (int i = 0; i < P.d; ++i)
                loadings[((size_t)y * P.d + i) * nc + j] = hl[((size_t)y * P.d + i) * P.kp + j];
        }
}

extern "C" {

int crest_set_device(int32_t device) {
    return guarded([&] {
        if (device < 0 || device >= 64) value_error("device index out of range");
        g_device = device;
    });
}

int crest_set_precision(int32_t prec) {
    return guarded([&] {
        if (prec != CREST_PREC_FP64 && prec != CREST_PREC_FP32 && prec != CREST_PREC_BF16X3 && prec != CREST_PREC_FP16)
            value_error("unknown precision");
        g_prec = prec;
    });
}

int32_t crest_get_precision(void) { return g_prec; }

int crest_set_knn_precision(int32_t prec) {
    return guarded([&] {
        if (prec != -1 && prec != CREST_PREC_FP64 && prec != CREST_PREC_FP32 && prec != CREST_PREC_BF16X3 &&
            prec != CREST_PREC_FP16)
            value_error("unknown kNN screen precision");
        g_knn_prec = prec;
    });
}

int32_t crest_get_knn_precision(void) {
    const int p = g_knn_prec;
    return p < 0 ? (int32_t)g_prec : p;
}

const char *crest_last_error(void) { return g_err.c_str(); }

int32_t crest_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int crest_knn_screen_bound(int32_t prec, size_t dim, double *coef, int32_t *kpad) {
    return guarded([&] {
        if (!coef || !kpad) value_error("output pointer is NULL");
        if (prec < -1 || prec > CREST_PREC_FP16) value_error("unknown precision");
        if (dim > (size_t)INT32_MAX) value_error("dim too large");
        const int kp = prec == CREST_PREC_FP64 ? CREST_PREC_FP32 : prec;
        *kpad = prec < 0 ? (int32_t)dim : knn_kpad(kp, (int)dim);
        const KnnCert c = prec < 0    ? knn_cert_gemm((int)dim, nullptr)
                          : *kpad < 0 ? KnnCert{}
                                      : knn_cert(kp, (int)dim, *kpad, nullptr, nullptr, nullptr, 0);
        coef[0] = c.c1;
        coef[1] = c.c2;
        coef[2] = c.c3;
        coef[3] = c.c4;
    });
}

int crest_knn_certificate(int64_t *checked, int64_t *rechecked, int64_t *uncertified) {
    return guarded([&] {
        if (!checked || !rechecked || !uncertified) value_error("output pointer is NULL");
        Ctx &x = ctx();
        std::lock_guard<std::mutex> lk(x.mu);
        *checked = x.knn_checked;
        *rechecked = x.knn_rechecked;
        *uncertified = x.knn_uncertified;
    });
}

int crest_release(void) {
    return guarded([&] {
        Ctx &x = ctx();
        std::lock_guard<std::mutex> lk(x.mu);
        x.sync();
        x.release();
    });
}

int crest_dev_alloc(size_t bytes, void **ptr) {
    return guarded([&] {
        if (!ptr) value_error("ptr is NULL");
        *ptr = nullptr;
        ctx();
        HIPCHK(hipMalloc(ptr, bytes ? bytes : 1));
    });
}

int crest_dev_free(void *ptr) {
    return guarded([&] {
        ctx();
        if (ptr) HIPCHK(hipFree(ptr));
    });
}

int crest_memcpy_htod(void *dst, const void *src, size_t bytes) {
    return guarded([&] {
        Ctx &x = ctx();
        if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, x.st));
        x.sync();
    });
}

int crest_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    return guarded([&] {
        Ctx &x = ctx();
        if (bytes) HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, x.st));
        x.sync();
    });
}

int crest_stream_synchronize(void *stream) {
    return guarded([&] {
        const int dev = g_device;
        if (dev < 0 || dev >= 64) value_error("device index out of range");
        HIPCHK(hipSetDevice(dev));
        HIPCHK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    });
}

int crest_dev_copy(void *dst, const void *src, size_t bytes) {
    return guarded([&] {
        if (bytes % 16) value_error("bytes must be a multiple of 16");
        Ctx &x = ctx();
        hipLaunchKernelGGL(k_copy16, dim3(4096), dim3(256), 0, x.st, (uint4 *)dst, (const uint4 *)src,
                           (int64_t)(bytes / 16));
        x.check_launch();
        x.sync();
    });
}

int crest_synchronize(void) {
    return guarded([&] {
        Ctx &x = ctx();
        x.sync();
    });
}

int crest_gene_stats(const crest_chunk *chunk, double *out, size_t out_len) {
    return guarded([&] {
        HostChunk h = check_chunk(chunk);
        if ((int64_t)out_len != h.n_vars * 5 || (out_len && !out)) value_error("out must be (n_vars, 5)");
        if (h.n_vars == 0 || h.n_spans == 0) return;
        Ctx &x = ctx();
        std::lock_guard<std::mutex> lk(x.mu);
        DevChunk c = upload_chunk(x, h);
        double *dev = x.get<double>("stats", out_len);
        launch_fill_f64(dev, (int64_t)out_len, 0.0, x.st);
        launch_gene_stats(c, dev, (int)h.n_vars, x.st);
        x.check_launch();
        std::vector<double> tmp(out_len);
        x.download(tmp.data(), dev, out_len);
        x.sync();
        for (size_t i = 0; i < out_len; ++i) out[i] += tmp[i];
    });
}

int crest_qc(const crest_chunk *chunk, const uint8_t *flag, size_t n_flag, double *cell_total, size_t n_cell_total,
             uint32_t *cell_ngenes, size_t n_cell_ngenes, double *cell_flag, size_t n_cell_flag, uint32_t *gene_ncells,
             size_t n_gene_ncells, double *gene_total, size_t n_gene_total) {
    return guarded([&] {
        HostChunk h = check_chunk(chunk);
        const int64_t nr = h.n_rows, nv = h.n_vars;
        // py.rs:242-244
        if ((int64_t)n_cell_total < nr || (int64_t)n_cell_ngenes < nr || (int64_t)n_cell_flag < nr ||
            (int64_t)n_gene_ncells != nv || (int64_t)n_gene_total != nv || (int64_t)n_flag != nv)
            value_error("output sizes do not match the maps");
        if ((nr && (!cell_total || !cell_ngenes || !cell_flag)) || (nv && (!gene_ncells || !gene_total || !flag)))
            value_error("NULL argument");
        if (h.n_spans == 0 || nr == 0 || nv == 0) return;
        Ctx &x = ctx();
        std::lock_guard<std::mutex> lk(x.mu);
        DevChunk c = upload_chunk(x, h);
        const uint8_t *df = x.upload("qc_flag", flag, (size_t)nv);
        double *rows = x.get<double>("qc_rows", (size_t)nr * 3);
        double *gt = x.get<double>("qc_gene_total", (size_t)nv);
        uint32_t *gn = x.get<uint32_t>("qc_gene_ncells", (size_t)nv);
        launch_fill_f64(rows, nr * 3, 0.0, x.st);
        launch_fill_f64(gt, nv, 0.0, x.st);
        launch_fill_u32(gn, nv, 0u, x.st);
        launch_qc(c, df, rows, rows + nr, rows + 2 * nr, gn, gt, x.st);
        x.check_launch();
        std::vector<double> hr((size_t)nr * 3), hgt((size_t)nv);
        std::vector<uint32_t> hgn((size_t)nv);
        x.download(hr.data(), rows, hr.size());
        x.download(hgt.data(), gt, hgt.size());
        x.download(hgn.data(), gn, hgn.size());
        x.sync();
        for (int64_t i = 0; i < nr; ++i) {
            cell_total[i] += hr[i];
            cell_ngenes[i] += (uint32_t)hr[nr + i];
            cell_flag[i] += hr[2 * nr + i];
        }
        for (int64_t g = 0; g < nv; ++g) {
            gene_ncells[g] += hgn[g];
            gene_total[g] += hgt[g];
        }
    });
}

int crest_gram_accumulate(const crest_chunk *chunk, const int32_t *sub_map, size_t n_sub_map, double *gram,
                          size_t gram_len, double *colsum, size_t d, int32_t scale, const double *mean, size_t n_mean,
                          const double *std_, size_t n_std, double max_value) {
    return guarded([&] {
        HostChunk h = check_chunk(chunk);
        if (gram_len != d * d || (d && (!gram || !colsum))) value_error("gram must be (d, d)");
        check_sub(sub_map, n_sub_map, h.n_vars, d);
        check_scale(scale, mean, n_mean, std_, n_std, d);
        if (d == 0 || h.n_spans == 0) return;
        Ctx &x = ctx();
        std::lock_guard<std::mutex> lk(x.mu);
        const int prec = g_prec;
        check_gram_prec(prec, scale, max_value);
        const int d_pad = (int)round_up((int64_t)d, kGramBlock);
        DevChunk c = upload_chunk(x, h);
        DevScale sc{};
        sc.sub = x.upload("sub_map", sub_map, n_sub_map);
        sc.scale = scale ? 1 : 0;
        sc.max_value = max_value;
        sc.d = (int)d;
        if (scale) {
            sc.mean = x.upload("mean", mean, d);
            sc.sd = x.upload("std", std_, d);
        }
        double *dg = x.get<double>("gram", (size_t)d_pad * d_pad);
        double *dc = x.get<double>("colsum", d_pad);
        launch_fill_f64(dg, (int64_t)d_pad * d_pad, 0.0, x.st);
        launch_fill_f64(dc, d_pad, 0.0, x.st);
        gram_rows(x, c, sc, prec, d_pad, 65536, dg, dc);
        std::vector<double> hg((size_t)d_pad * d_pad), hc(d);
        x.download(hg.data(), dg, hg.size());
        x.download(hc.data(), dc, d);
        x.sync();
        for (size_t a = 0; a < d; ++a) {
            colsum[a] += hc[a];
            for (size_t b = a; b < d; ++b) gram[a * d + b] += hg[a * d_pad + b];
        }
    });
}

// ---------------------------------------------------------------- gram_accumulate without allocation
// The scratch of crest_gram_accumulate_noalloc: the d_pad x d_pad Gram and colsum it
// accumulates into, one split-K partial per Gram block and the densified tile(s).
// Every GPU holds its own copy of this array from module load to process end; host
// code reaches it only through the per-device address from hipGetSymbolAddress.
namespace crest {
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
constexpr int kNoAllocMaxD = 2048;
}  // namespace crest
static __device__ __align__(256) unsigned char scratch[crest::kNoAllocScratchBytes];

namespace crest {
static std::mutex g_scratch_mu;
static void *g_scratch_dev[64] = {};

// Address of this device's scratch (looked up once per device, with it current).
static unsigned char *noalloc_scratch(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) fail(CREST_E_DEVICE, "no device address for the scratch array");
        g_scratch_dev[dev] = ptr;
    }
    return static_cast<unsigned char *>(g_scratch_dev[dev]);
}

// dev_gram[a * d + b] += g[a * d_pad + b] (a <= b < d), dev_colsum[a] += c[a] (a < d).
__global__ void k_noalloc_flush(const double *g, const double *c, int d, int d_pad, double *dev_gram,
                                double *dev_colsum) {
    const int64_t total = (int64_t)d * d;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int a = (int)(e / d), b = (int)(e % d);
        if (b >= a) dev_gram[e] += g[(int64_t)a * d_pad + b];
        if (b == 0) dev_colsum[a] += c[a];
    }
}
}  // namespace crest

int crest_gram_accumulate_noalloc(const crest_chunk *dev_chunk, const int32_t *dev_sub_map, size_t n_sub_map,
                                  double *dev_gram, double *dev_colsum, size_t d, int32_t scale,
                                  const double *dev_mean, const double *dev_std, double max_value, void *stream) {
    return guarded([&] {
        const crest_chunk *c = dev_chunk;
        if (!c) value_error("chunk is NULL");
        if (c->cells || !c->indptr || c->n_indptr < 1) value_error("the no-allocation form takes indptr chunks only");
        if ((c->nnz && (!c->genes || !c->values)) || (c->n_cell_map && !c->cell_map) ||
            (c->n_gene_map && !c->gene_map) || (n_sub_map && !dev_sub_map))
            value_error("chunk arrays must not be NULL");
        if (d && (!dev_gram || !dev_colsum)) value_error("gram must be (d, d)");
        if (scale && (!dev_mean || !dev_std)) value_error("scale=True needs mean and std");
        if (d > (size_t)kNoAllocMaxD)
            value_error("d exceeds the no-allocation scratch (at most " + std::to_string(kNoAllocMaxD) + " HVGs)");
        const int prec = g_prec;
        check_gram_prec(prec, scale, max_value);
        if (d == 0 || c->n_indptr < 2) return;
        const int dev = g_device;
        if (dev < 0 || dev >= 64) value_error("device index out of range");
        HIPCHK(hipSetDevice(dev));
        unsigned char *base = noalloc_scratch(dev);
        hipStream_t st = static_cast<hipStream_t>(stream);
        const int fmt = tile_fmt(prec);
        const int d_pad = (int)round_up((int64_t)d, kGramBlock);
        const int nb = d_pad / kGramBlock, nblk = nb * (nb + 1) / 2;
        // scratch layout: Gram, colsum, partials, tile(s), each 256-byte aligned
        size_t off = 0;
        auto carve = [&](size_t bytes) {
            unsigned char *p = base + off;
            off += (size_t)round_up((int64_t)bytes, 256);
            return p;
        };
        double *g = reinterpret_cast<double *>(carve((size_t)d_pad * d_pad * sizeof(double)));
        double *cs = reinterpret_cast<double *>(carve((size_t)d_pad * sizeof(double)));
        double *partial = reinterpret_cast<double *>(carve((size_t)nblk * kGramBlock * kGramBlock * sizeof(double)));
        const bool x2 = fmt == TILE_BF16X2 || fmt == TILE_F16X2;
        const size_t row_bytes = (size_t)d_pad * (x2 ? 4 : tile_elem_bytes(fmt));
        if (off + 512 > kNoAllocScratchBytes) value_error("d exceeds the no-allocation scratch");
        const int64_t tile_rows = (int64_t)((kNoAllocScratchBytes - off - 512) / row_bytes) / 16 * 16;
        if (tile_rows < 16) value_error("d exceeds the no-allocation scratch");
        void *ta = carve((size_t)tile_rows * d_pad * (x2 ? 2 : tile_elem_bytes(fmt)));
        void *tb = x2 ? carve((size_t)tile_rows * d_pad * 2) : nullptr;
        DevChunk ch{};
        ch.span_off = c->indptr;
        ch.span_cell = nullptr;
        ch.first_cell = c->first_cell;
        ch.n_spans = (int64_t)c->n_indptr - 1;
        ch.genes = c->genes;
        ch.values = c->values;
        ch.cell_map = c->cell_map;
        ch.n_cell_map = (int64_t)c->n_cell_map;
        ch.gene_var = c->gene_map;
        ch.n_gene_map = (int64_t)c->n_gene_map;
        ch.target_sum = c->target_sum;
        ch.log1p = c->log1p ? 1 : 0;
        ch.base = 0;
        ch.nnz = (int64_t)c->nnz;
        DevScale sc{};
        sc.sub = dev_sub_map;
        sc.scale = scale ? 1 : 0;
        sc.max_value = max_value;
        sc.d = (int)d;
        sc.mean = dev_mean;
        sc.sd = dev_std;
        // the Gram and colsum of this call start at zero (fill kernels, no memset)
        launch_fill_f64(g, (int64_t)d_pad * d_pad, 0.0, st);
        launch_fill_f64(cs, d_pad, 0.0, st);
        for (int64_t r0 = 0; r0 < ch.n_spans; r0 += tile_rows) {
            const int64_t r1 = std::min<int64_t>(ch.n_spans, r0 + tile_rows);
            const int64_t rows_pad = round_up(r1 - r0, 16);
            launch_densify(sub_chunk(ch, r0, r1), sc, fmt, ta, tb, rows_pad, d_pad, cs, g, st);
            launch_gram(fmt, ta, tb, rows_pad, d_pad, 1, partial, st);
            launch_gram_reduce(partial, 1, d_pad, g, st);
        }
        const int blocks = (int)std::min<int64_t>(ceil_div((int64_t)d * d, 256), 4096);
        hipLaunchKernelGGL(k_noalloc_flush, dim3(blocks), dim3(256), 0, st, g, cs, (int)d, d_pad, dev_gram,
                           dev_colsum);
        HIPCHK(hipGetLastError());
    });
}

int crest_pca_from_gram(const double *gram, size_t gram_len, const double *colsum, size_t d, size_t n, size_t k,
                        double *variance, double *loadings, double *total, size_t *k_out) {
    return guarded([&] {
        if (gram_len != d * d) value_error("gram must be (d, d)");
        if (n < 2 || d == 0) fail(CREST_E_RUNTIME, "need at least 2 cells and 1 gene");
        const size_t kk = std::min(k, d);
        if (k_out) *k_out = kk;
        Ctx &x = ctx();
        std::lock_guard<std::mutex> lk(x.mu);
        const double *dg = x.upload("pca_gram", gram, d * d);
        const double *dc = x.upload("pca_colsum", colsum, d);
        double *dv = x.get<double>("pca_var", std::max<size_t>(kk, 1));
        double *dl = x.get<double>("pca_load", std::max<size_t>(d * kk, 1));
        double *dt = x.get<double>("pca_total", 1);
        pca_device(x, dg, (int)d, dc, (int)d, (double)n, (int)kk, dv, dl, dt);
        x.download(variance, dv, kk);
        x.download(loadings, dl, d * kk);
        x.download(total, dt, 1);
        x.sync();
    });
}

int crest_project(const crest_chunk *chunk, const int32_t *sub_map, size_t n_sub_map, const double *loadings, size_t d,
                  size_t k, float *out, size_t out_len, int32_t scale, const double *mean, size_t n_mean,
                  const double *std_, size_t n_std, double max_value) {
    return guarded([&] {
        HostChunk h = check_chunk(chunk);
        check_sub(sub_map, n_sub_map, h.n_vars, d);
        if ((int64_t)out_len < h.n_rows * (int64_t)k) value_error("out must be (n_rows, k)");
        check_scale(scale, mean, n_mean, std_, n_std, d);
        if (k > 256) value_error("k > 256 is not supported by the GPU projection");
        if (k == 0 || h.n_spans == 0) return;
        // rows this chunk writes
        int64_t rmin = INT64_MAX, rmax = -1;
        for (int64_t i = 0; i < h.n_spans; ++i) {
            const int64_t r = host_row(h, i);
            if (r >= 0) {
                rmin = std::min(rmin, r);
                rmax = std::max(rmax, r);
            }
        }
        if (rmax < 0) return;
        Ctx &x = ctx();
        std::lock_guard<std::mutex> lk(x.mu);
        DevChunk c = upload_chunk(x, h);
        DevScale sc{};
        sc.sub = x.upload("sub_map", sub_map, n_sub_map);
        sc.scale = scale ? 1 : 0;
        sc.max_value = max_value;
        sc.d = (int)d;
        if (scale) {
            sc.mean = x.upload("mean", mean, d);
            sc.sd = x.upload("std", std_, d);
        }
        const double *V = x.upload("loadings", loadings, d * k);
        const size_t nrow = (size_t)(rmax - rmin + 1);
        float *dout = x.upload("proj_out", out + rmin * (int64_t)k, nrow * k);
        launch_project(c, sc, V, (int)k, dout, rmin, x.st);
        x.check_launch();
        x.download(out + rmin * (int64_t)k, dout, nrow * k);
        x.sync();
    });
}

int crest_knn_graph(const float *data, size_t n, size_t dim, size_t k, int32_t exact, uint64_t seed,
                    uint32_t *indices, float *distances, size_t *k_out) {
    (void)exact;
    (void)seed;
    return guarded([&] {
        // knn.rs:338-345
        size_t kk = 0;
        if (n > 1) kk = std::max<size_t>(std::min(k, n - 1), 1);
        if (k_out) *k_out = kk;
        if (n == 0 || kk == 0) return;
        if (dim == 0) value_error("data must have at least one column");
        if (n >= 0xFFFFFFFFull) value_error("too many points");
        Ctx &x = ctx();
        std::lock_guard<std::mutex> lk(x.mu);
        const float *dx = x.upload("knn_x", data, n * dim);
        uint32_t *di = x.get<uint32_t>("knn_idx", n * kk);
        float *dd = x.get<float>("knn_dist", n * kk);
        knn_device(x, dx, (int64_t)n, nullptr, 0, (int)dim, (int)kk, crest_get_knn_precision(), di, dd, nullptr);
        x.download(indices, di, n * kk);
        x.download(distances, dd, n * kk);
        x.sync();
    });
}

int crest_knn_query(const float *reference, size_t n_ref, const float *query, size_t n_query, size_t dim, size_t k,
                    int32_t exact, uint64_t seed, uint32_t *indices, float *distances, size_t *k_out) {
    (void)exact;
    (void)seed;
    return guarded([&] {
        // knn.rs:485-490
        const size_t kk = std::min(k, n_ref);
        if (k_out) *k_out = 0;
        if (n_query == 0 || kk == 0) return;
        if (dim == 0) value_error("data must have at least one column");
        if (n_ref >= 0xFFFFFFFFull || n_query >= 0xFFFFFFFFull) value_error("too many points");
        if (!reference || !query || !indices || !distances) value_error("NULL argument");
        Ctx &x = ctx();
        std::lock_guard<std::mutex> lk(x.mu);
        const float *dr = x.upload("knnq_ref", reference, n_ref * dim);
        const float *dq = x.upload("knnq_query", query, n_query * dim);
        uint32_t *di = x.get<uint32_t>("knn_idx", n_query * kk);
        float *dd = x.get<float>("knn_dist", n_query * kk);
        knn_device(x, dr, (int64_t)n_ref, dq, (int64_t)n_query, (int)dim, (int)kk, crest_get_knn_precision(), di, dd,
                   nullptr);
        x.download(indices, di, n_query * kk);
        x.download(distances, dd, n_query * kk);
        x.sync();
        if (k_out) *k_out = kk;
    });
}

int crest_connectivities(const uint32_t *indices, const float *distances, size_t n, size_t k, size_t n_neighbors,
                         uint32_t *rows, uint32_t *cols, float *weights, size_t cap, size_t *n_out) {
    return guarded([&] {
        if (n_out) *n_out = 0;
        if (n == 0 || k == 0) return;  // graph.rs:63-65
        if (!indices || !distances) value_error("indices and distances must be (n, k) arrays of equal shape");
        if (n * k >= 0xFFFFFFFFull) value_error("graph too large");
        Ctx &x = ctx();
        std::lock_guard<std::mutex> lk(x.mu);
        const int64_t m = (int64_t)(n * k);
        const uint32_t *di = x.upload("conn_idx", indices, m);
        const float *dd = x.upload("conn_dist", distances, m);
        uint64_t *k1 = x.get<uint64_t>("conn_k1", m), *k2 = x.get<uint64_t>("conn_k2", m);
        float *w1 = x.get<float>("conn_w1", m), *w2 = x.get<float>("conn_w2", m), *wu = x.get<float>("conn_wu", m);
        uint32_t *fl = x.get<uint32_t>("conn_flag", m), *ps = x.get<uint32_t>("conn_pos", m);
        double *part = x.get<double>("conn_part", 257);
        const size_t tb = conn_temp_bytes(m);
        void *tmp = x.get<uint8_t>("conn_tmp", tb);
        uint32_t *orow = x.get<uint32_t>("conn_rows", m), *ocol = x.get<uint32_t>("conn_cols", m);
        float *ow = x.get<float>("conn_w", m);
        uint32_t *cnt = x.get<uint32_t>("conn_cnt", 1);
        launch_connectivities(di, dd, (int64_t)n, (int)k, (int)n_neighbors, k1, k2, w1, w2, wu, fl, ps, part, tmp, tb,
                              orow, ocol, ow, cnt, x.st);
        x.check_launch();
        uint32_t hc = 0;
        x.download(&hc, cnt, 1);
        x.sync();
        if (hc > cap) value_error("output capacity too small");
        x.download(rows, orow, hc);
        x.download(cols, ocol, hc);
        x.download(weights, ow, hc);
        x.sync();
        if (n_out) *n_out = hc;
    });
}

int crest_materialize(const crest_chunk *chunk, uint32_t *rows, uint32_t *vars, float *values, size_t cap,
                      size_t *n_out) {
    return guarded([&] {
        HostChunk h = check_chunk(chunk);
        if (n_out) *n_out = 0;
        if (h.n_spans == 0) return;
        Ctx &x = ctx();
        std::lock_guard<std::mutex> lk(x.mu);
        DevChunk c = upload_chunk(x, h);
        int64_t *kept = x.get<int64_t>("mat_kept", h.n_spans);
        launch_materialize_count(c, kept, x.st);
        x.check_launch();
        std::vector<int64_t> hk(h.n_spans), off(h.n_spans);
        x.download(hk.data(), kept, hk.size());
        x.sync();
        int64_t tot = 0;
        for (int64_t i = 0; i < h.n_spans; ++i) {
            off[i] = tot;
            tot += hk[i];
        }
        if ((size_t)tot > cap) value_error("output capacity too small");
        if (n_out) *n_out = (size_t)tot;
        if (tot == 0) return;
        const int64_t *doff = x.upload("mat_off", off.data(), off.size());
        uint32_t *dr = x.get<uint32_t>("mat_rows", tot);
        uint32_t *dv = x.get<uint32_t>("mat_vars", tot);
        float *dx = x.get<float>("mat_vals", tot);
        launch_materialize(c, doff, dr, dv, dx, x.st);
        x.check_launch();
        x.download(rows, dr, tot);
        x.download(vars, dv, tot);
        x.download(values, dx, tot);
        x.sync();
    });
}

int crest_pipeline_run(const crest_pipeline_args *a, double *variance, double *loadings, crest_pipeline_result *res) {
    return guarded([&] {
        check_pipeline_args(a, variance, loadings, res);
        if (!a->dev_indptr || !a->dev_genes || !a->dev_values) value_error("NULL argument");
        Ctx &x = ctx();
        std::lock_guard<std::mutex> lk(x.mu);
        const int64_t n = (int64_t)a->n_cells;
        int64_t nnz = 0;
        HIPCHK(hipMemcpy(&nnz, a->dev_indptr + n, 8, hipMemcpyDeviceToHost));
        auto resident = [&](const DevChunk &tmpl, auto &&fn) {
            DevChunk c = tmpl;
            c.span_off = a->dev_indptr;
            c.span_cell = nullptr;
            c.first_cell = 0;
            c.n_spans = n;
            c.genes = a->dev_genes;
            c.values = a->dev_values;
            c.base = 0;
            c.nnz = nnz;
            fn(c);
        };
        run_pipeline(x, a, resident, variance, loadings, res);
    });
}

int crest_pipeline_run_streamed(const crest_pipeline_args *a, const int64_t *host_indptr,
                                const uint32_t *host_genes, const float *host_values, int32_t host_pinned,
                                size_t stream_rows, double *variance, double *loadings, crest_pipeline_result *res) {
    return guarded([&] {
        check_pipeline_args(a, variance, loadings, res);
        if (!host_indptr || !host_genes || !host_values) value_error("NULL argument");
        const int64_t n = (int64_t)a->n_cells;
        if (host_indptr[0] != 0) value_error("indptr must start at 0");
        const int64_t rows = stream_rows ? (int64_t)stream_rows : 32768;
        // the CSR's own consistency: every offset non-decreasing (a chunk's entries then
        // fit the device slots sized from its end offsets), then the largest chunk
        for (int64_t i = 1; i <= n; ++i)
            if (host_indptr[i] < host_indptr[i - 1]) value_error("indptr must be non-decreasing");
        int64_t max_ne = 0;
        for (int64_t lo = 0; lo < n; lo += rows) {
            const int64_t hi = std::min(n, lo + rows);
            max_ne = std::max(max_ne, host_indptr[hi] - host_indptr[lo]);
        }
        Ctx &x = ctx();
        std::lock_guard<std::mutex> lk(x.mu);
        // both slots sized up front, so nothing is reallocated while a copy is in flight
        int64_t *d_ip[2];
        uint32_t *d_g[2];
        float *d_v[2];
        uint8_t *h_stage[2] = {nullptr, nullptr};
        const size_t ip_bytes = (size_t)(rows + 1) * 8, g_bytes = (size_t)max_ne * 4;
        for (int b = 0; b < 2; ++b) {
            const std::string sfx = b ? "1" : "0";
            d_ip[b] = x.get<int64_t>("st_ip" + sfx, rows + 1);
            d_g[b] = x.get<uint32_t>("st_g" + sfx, max_ne);
            d_v[b] = x.get<float>("st_v" + sfx, max_ne);
            if (!host_pinned) h_stage[b] = (uint8_t *)x.host("stage" + sfx, ip_bytes + 2 * g_bytes);
        }
        x.sync();
        HIPCHK(hipStreamSynchronize(x.copy));
        auto streamed = [&](const DevChunk &tmpl, auto &&fn) {
            int64_t i = 0;
            for (int64_t lo = 0; lo < n; lo += rows, ++i) {
                const int64_t hi = std::min(n, lo + rows), nr = hi - lo;
                const int b = (int)(i & 1);
                const int64_t e0 = host_indptr[lo], ne = host_indptr[hi] - e0;
                const void *sip = host_indptr + lo, *sg = host_genes + e0, *sv = host_values + e0;
                if (!host_pinned) {  // wait until the slot's previous upload has left the staging buffer
                    HIPCHK(hipEventSynchronize(x.copied[b]));
                    uint8_t *st = h_stage[b];
                    par_copy(st, sip, (size_t)(nr + 1) * 8);
                    par_copy(st + ip_bytes, sg, (size_t)ne * 4);
                    par_copy(st + ip_bytes + g_bytes, sv, (size_t)ne * 4);
                    sip = st;
                    sg = st + ip_bytes;
                    sv = st + ip_bytes + g_bytes;
                }
                // the device slot is free once the kernels of its previous chunk are done
                HIPCHK(hipStreamWaitEvent(x.copy, x.consumed[b], 0));
                HIPCHK(hipMemcpyAsync(d_ip[b], sip, (size_t)(nr + 1) * 8, hipMemcpyHostToDevice, x.copy));
                if (ne) {
                    HIPCHK(hipMemcpyAsync(d_g[b], sg, (size_t)ne * 4, hipMemcpyHostToDevice, x.copy));
                    HIPCHK(hipMemcpyAsync(d_v[b], sv, (size_t)ne * 4, hipMemcpyHostToDevice, x.copy));
                }
                HIPCHK(hipEventRecord(x.copied[b], x.copy));
                HIPCHK(hipStreamWaitEvent(x.st, x.copied[b], 0));
                DevChunk c = tmpl;
                c.span_off = d_ip[b];
                c.span_cell = nullptr;
                c.first_cell = (uint32_t)lo;
                c.n_spans = nr;
                c.genes = d_g[b];
                c.values = d_v[b];
                c.base = e0;
                c.nnz = ne;
                fn(c);
                HIPCHK(hipEventRecord(x.consumed[b], x.st));
            }
        };
        run_pipeline(x, a, streamed, variance, loadings, res);
        HIPCHK(hipStreamSynchronize(x.copy));
    });
}

int crest_host_alloc(size_t bytes, void **ptr) {
    return guarded([&] {
        if (!ptr) value_error("ptr is NULL");
        *ptr = nullptr;
        Ctx &x = ctx();
        (void)x;
        HIPCHK(hipHostMalloc(ptr, std::max<size_t>(bytes, 1), hipHostMallocDefault));
    });
}

int crest_host_free(void *ptr) {
    return guarded([&] {
        if (ptr) HIPCHK(hipHostFree(ptr));
    });
}

int crest_pipeline_run_batched(const crest_batch_args *b, double *variance, double *loadings, double *total,
                               crest_batch_result *res) {
    return guarded([&] {
        check_batch_args(b, variance, loadings, total, res);
        Ctx &x = ctx();
        std::lock_guard<std::mutex> lk(x.mu);
        run_batch(x, b, variance, loadings, total, res);
    });
}

}  // extern "C"

// ================================================================ sessions
// The device state of one tl.pca / highly_variable_genes run over a chunk source:
// gene statistics, scale parameters, Gram, colsum, eigenpairs and scores stay in
// device memory between pushes; each push stages its chunk through one of two pinned
// slots and the copy stream, so the upload of chunk i+1 overlaps the kernels of chunk i.
struct crest_session {
    int dev = 0;
    int64_t n_vars = 0, n_rows = 0;
    std::vector<void *> owned;  // device allocations freed with the session
    // chunk maps, cached by the host array (pointer, length) they came from
    const int64_t *cmap_src = nullptr;
    size_t cmap_len = 0;
    int64_t *d_cmap = nullptr;
    size_t d_cmap_cap = 0;
    const int32_t *gmap_src = nullptr;
    size_t gmap_len = 0;
    int32_t *d_gmap = nullptr;
    size_t d_gmap_cap = 0;
    double *d_stats = nullptr;
    int64_t stats_pushed = 0;
    // the Gram pass
    bool gram_ready = false;
    int d = 0, d_pad = 0, prec = 0, scale = 0;
    double max_value = 0.0;
    int32_t *d_sub = nullptr;
    double *d_mean = nullptr, *d_sd = nullptr, *d_gram = nullptr, *d_colsum = nullptr;
    // the eigensolve and projection
    bool solved = false, projecting = false;
    int k = 0;
    double *d_var = nullptr, *d_load = nullptr, *d_total = nullptr;
    float *d_shift = nullptr, *d_x = nullptr;
};

namespace crest {

static void *session_alloc(crest_session *S, size_t bytes) {
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, std::max<size_t>(bytes, 256)));
    S->owned.push_back(p);
    return p;
}
static void session_free(crest_session *S, void *p) {
    if (!p) return;
    auto it = std::find(S->owned.begin(), S->owned.end(), p);
    if (it != S->owned.end()) S->owned.erase(it);
    HIPCHK(hipFree(p));
}
static Ctx &session_ctx(crest_session *S) {
    if (!S) value_error("session is NULL");
    const int keep = g_device;
    g_device = S->dev;
    Ctx &x = ctx();
    g_device = keep;
    return x;
}

// The chunk's maps in device memory (uploaded when the host array differs from the
// one cached), with the session's extents checked.
static void session_maps(Ctx &x, crest_session *S, const crest_chunk *c) {
    if (c->cell_map != S->cmap_src || c->n_cell_map != S->cmap_len) {
        if (map_extent(c->cell_map, c->n_cell_map) > S->n_rows)
            value_error("cell_map maps cells beyond the session's n_rows");
        x.sync();  // kernels in flight may read the previous map
        if (S->d_cmap_cap < c->n_cell_map) {
            session_free(S, S->d_cmap);
            S->d_cmap = (int64_t *)session_alloc(S, c->n_cell_map * sizeof(int64_t));
            S->d_cmap_cap = c->n_cell_map;
        }
        if (c->n_cell_map) HIPCHK(hipMemcpy(S->d_cmap, c->cell_map, c->n_cell_map * 8, hipMemcpyHostToDevice));
        S->cmap_src = c->cell_map;
        S->cmap_len = c->n_cell_map;
    }
    if (c->gene_map != S->gmap_src || c->n_gene_map != S->gmap_len) {
        if (map_extent(c->gene_map, c->n_gene_map) != S->n_vars)
            value_error("gene_map must map onto the session's n_vars genes");
        x.sync();
        if (S->d_gmap_cap < c->n_gene_map) {
            session_free(S, S->d_gmap);
            S->d_gmap = (int32_t *)session_alloc(S, c->n_gene_map * sizeof(int32_t));
            S->d_gmap_cap = c->n_gene_map;
        }
        if (c->n_gene_map) HIPCHK(hipMemcpy(S->d_gmap, c->gene_map, c->n_gene_map * 4, hipMemcpyHostToDevice));
        S->gmap_src = c->gene_map;
        S->gmap_len = c->n_gene_map;
    }
}

// Stage one chunk into the next slot and hand back its device view; the caller
// launches its kernels on x.st and then calls session_consumed.
static DevChunk session_stage(Ctx &x, crest_session *S, const crest_chunk *c, int &slot_index) {
    const HostChunk h = check_chunk_geometry(c);
    session_maps(x, S, c);
    const int si = x.stage_next;
    x.stage_next ^= 1;
    Ctx::StageSlot &sl = x.stage[si];
    if (sl.used) HIPCHK(hipEventSynchronize(sl.consumed));  // its last kernels and copy are done
    const size_t nnz = c->nnz;
    const bool cells = c->cells != nullptr;
    const size_t n_off = cells ? h.span_off.size() : c->n_indptr;
    auto up = [](size_t b) { return (size_t)round_up((int64_t)b, 256); };
    const size_t o_g = 0, o_v = o_g + up(nnz * 4), o_s = o_v + up(nnz * 4), o_c = o_s + up(n_off * 8);
    const size_t total = o_c + (cells ? up(h.span_cell.size() * 4) : 0);
    if (sl.pin_bytes < total) {
        if (sl.pin) HIPCHK(hipHostFree(sl.pin));
        sl.pin = nullptr;
        sl.pin_bytes = 0;
        HIPCHK(hipHostMalloc((void **)&sl.pin, total, hipHostMallocDefault));
        sl.pin_bytes = total;
    }
    if (sl.dev_bytes < total) {
        if (sl.dev) HIPCHK(hipFree(sl.dev));
        sl.dev = nullptr;
        sl.dev_bytes = 0;
        HIPCHK(hipMalloc((void **)&sl.dev, total));
        sl.dev_bytes = total;
    }
    if (nnz) {
        par_copy(sl.pin + o_g, c->genes, nnz * 4);
        par_copy(sl.pin + o_v, c->values, nnz * 4);
    }
    if (cells) {
        std::memcpy(sl.pin + o_s, h.span_off.data(), n_off * 8);
        std::memcpy(sl.pin + o_c, h.span_cell.data(), h.span_cell.size() * 4);
    } else {
        std::memcpy(sl.pin + o_s, c->indptr, n_off * 8);
    }
    HIPCHK(hipMemcpyAsync(sl.dev, sl.pin, total, hipMemcpyHostToDevice, x.copy));
    HIPCHK(hipEventRecord(sl.copied, x.copy));
    HIPCHK(hipStreamWaitEvent(x.st, sl.copied, 0));
    DevChunk d{};
    d.genes = reinterpret_cast<const uint32_t *>(sl.dev + o_g);
    d.values = reinterpret_cast<const float *>(sl.dev + o_v);
    d.span_off = reinterpret_cast<const int64_t *>(sl.dev + o_s);
    d.span_cell = cells ? reinterpret_cast<const uint32_t *>(sl.dev + o_c) : nullptr;
    d.first_cell = cells ? 0 : c->first_cell;
    d.n_spans = h.n_spans;
    d.base = h.base;
    d.nnz = (int64_t)nnz;
    d.cell_map = S->d_cmap;
    d.n_cell_map = (int64_t)c->n_cell_map;
    d.gene_var = S->d_gmap;
    d.n_gene_map = (int64_t)c->n_gene_map;
    d.target_sum = c->target_sum;
    d.log1p = c->log1p ? 1 : 0;
    slot_index = si;
    return d;
}
static void session_consumed(Ctx &x, int si) {
    x.check_launch();
    HIPCHK(hipEventRecord(x.stage[si].consumed, x.st));
    x.stage[si].used = true;
}
static DevScale session_scale(const crest_session *S) {
    DevScale sc{};
    sc.sub = S->d_sub;
    sc.scale = S->scale;
    sc.max_value = S->max_value;
    sc.d = S->d;
    sc.mean = S->d_mean;
    sc.sd = S->d_sd;
    return sc;
}

}  // namespace crest

int crest_session_create(size_t n_vars, size_t n_rows, crest_session **out) {
    return guarded([&] {
        if (!out) value_error("out is NULL");
        *out = nullptr;
        if (n_vars == 0 || n_vars > (size_t)INT32_MAX) value_error("n_vars must be in 1 .. 2^31 - 1");
        if (n_rows > (size_t)UINT32_MAX) value_error("too many rows for a session");
        Ctx &x = ctx();
        std::lock_guard<std::mutex> lk(x.mu);
        std::unique_ptr<crest_session> S(new crest_session());
        S->dev = x.dev;
        S->n_vars = (int64_t)n_vars;
        S->n_rows = (int64_t)n_rows;
        try {
            S->d_stats = (double *)session_alloc(S.get(), n_vars * 5 * sizeof(double));
            launch_fill_f64(S->d_stats, (int64_t)n_vars * 5, 0.0, x.st);
            x.check_launch();
        } catch (...) {
            for (void *p : S->owned) hipFree(p);
            throw;
        }
        *out = S.release();
    });
}

int crest_session_destroy(crest_session *S) {
    return guarded([&] {
        if (!S) return;
        Ctx &x = session_ctx(S);
        std::lock_guard<std::mutex> lk(x.mu);
        x.sync();
        HIPCHK(hipStreamSynchronize(x.copy));
        for (void *p : S->owned) hipFree(p);
        delete S;
    });
}

int crest_session_push_stats(crest_session *S, const crest_chunk *chunk) {
    return guarded([&] {
        Ctx &x = session_ctx(S);
        std::lock_guard<std::mutex> lk(x.mu);
        int si = 0;
        const DevChunk c = session_stage(x, S, chunk, si);
        launch_gene_stats(c, S->d_stats, (int)S->n_vars, x.st);
        session_consumed(x, si);
        ++S->stats_pushed;
    });
}

int crest_session_stats(crest_session *S, double *out, size_t out_len) {
    return guarded([&] {
        Ctx &x = session_ctx(S);
        if (!out || out_len != (size_t)S->n_vars * 5) value_error("out must be (n_vars, 5)");
        std::lock_guard<std::mutex> lk(x.mu);
        x.download(out, S->d_stats, out_len);
        x.sync();
    });
}

int crest_session_hvg_seurat(crest_session *S, int64_t n_top, int32_t n_bins, double min_mean, double max_mean,
                             double min_disp, double max_disp, double *means, double *dispersions,
                             double *dispersions_norm, uint8_t *highly_variable, double *mean_log,
                             double *std_log) {
    return guarded([&] {
        Ctx &x = session_ctx(S);
        if (!means || !dispersions || !dispersions_norm || !highly_variable || !mean_log || !std_log)
            value_error("NULL argument");
        if (!hvg_bins_supported(n_bins)) value_error("n_bins must be in 1 .. 256");
        std::lock_guard<std::mutex> lk(x.mu);
        const int64_t G = S->n_vars;
        int32_t *codes = x.get<int32_t>("hvg_codes", G);
        double *work = x.get<double>("hvg_work", (size_t)G * 6);
        uint8_t *hv = x.get<uint8_t>("hvg_flag", G);
        double *o_means = work + G, *o_disp = work + 2 * G, *o_dn = work + 3 * G, *o_ml = work + 4 * G,
               *o_sl = work + 5 * G;
        launch_hvg_seurat(S->d_stats, G, (double)S->n_rows, n_bins, n_top, min_mean, max_mean, min_disp, max_disp,
                          codes, work, o_means, o_disp, o_dn, hv, o_ml, o_sl, x.st);
        x.check_launch();
        x.download(means, o_means, G);
        x.download(dispersions, o_disp, G);
        x.download(dispersions_norm, o_dn, G);
        x.download(highly_variable, hv, G);
        x.download(mean_log, o_ml, G);
        x.download(std_log, o_sl, G);
        x.sync();
    });
}

int crest_session_begin_gram(crest_session *S, const int32_t *sub_map, size_t n_sub_map, size_t d, int32_t scale,
                             const double *mean, const double *std_, double max_value) {
    return guarded([&] {
        Ctx &x = session_ctx(S);
        if (!sub_map || (int64_t)n_sub_map != S->n_vars) value_error("sub_map must have length n_vars and values < d");
        if (d == 0) fail(CREST_E_RUNTIME, "need at least 2 cells and 1 gene");
        for (size_t i = 0; i < n_sub_map; ++i)
            if (sub_map[i] >= (int64_t)d) value_error("sub_map must have length n_vars and values < d");
        const bool derive = scale && !mean && !std_;
        if (scale && !derive && (!mean || !std_)) value_error("scale=True needs mean and std");
        if (derive && S->stats_pushed == 0) value_error("no statistics were pushed to derive mean and std from");
        const int prec = g_prec;
        check_gram_prec(prec, scale, max_value);
        std::vector<int32_t> goh;
        if (derive) goh = hvg_genes(sub_map, n_sub_map, (int)d);
        std::lock_guard<std::mutex> lk(x.mu);
        x.sync();  // kernels of an earlier Gram pass may still read these buffers
        for (void *p : {(void *)S->d_sub, (void *)S->d_mean, (void *)S->d_sd, (void *)S->d_gram, (void *)S->d_colsum})
            session_free(S, p);
        S->gram_ready = S->solved = S->projecting = false;
        S->cmap_src = nullptr;  // each pass reads its maps afresh
        S->gmap_src = nullptr;
        S->d = (int)d;
        S->d_pad = (int)round_up((int64_t)d, kGramBlock);
        S->prec = prec;
        S->scale = scale ? 1 : 0;
        S->max_value = max_value;
        S->d_sub = (int32_t *)session_alloc(S, n_sub_map * 4);
        S->d_mean = (double *)session_alloc(S, d * 8);
        S->d_sd = (double *)session_alloc(S, d * 8);
        S->d_gram = (double *)session_alloc(S, (size_t)S->d_pad * S->d_pad * 8);
        S->d_colsum = (double *)session_alloc(S, (size_t)S->d_pad * 8);
        HIPCHK(hipMemcpy(S->d_sub, sub_map, n_sub_map * 4, hipMemcpyHostToDevice));
        if (scale && !derive) {
            HIPCHK(hipMemcpy(S->d_mean, mean, d * 8, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(S->d_sd, std_, d * 8, hipMemcpyHostToDevice));
        } else if (derive) {
            int32_t *d_goh = x.get<int32_t>("sess_goh", d);
            HIPCHK(hipMemcpyAsync(d_goh, goh.data(), d * 4, hipMemcpyHostToDevice, x.st));
            hipLaunchKernelGGL(k_mean_std, dim3((unsigned)ceil_div((int64_t)d, 256)), dim3(256), 0, x.st, S->d_stats,
                               d_goh, (int)d, (double)S->n_rows, S->d_mean, S->d_sd, (int64_t)0,
                               (const int64_t *)nullptr);
        }
        launch_fill_f64(S->d_gram, (int64_t)S->d_pad * S->d_pad, 0.0, x.st);
        launch_fill_f64(S->d_colsum, S->d_pad, 0.0, x.st);
        x.check_launch();
        S->gram_ready = true;
    });
}

int crest_session_push_gram(crest_session *S, const crest_chunk *chunk) {
    return guarded([&] {
        Ctx &x = session_ctx(S);
        if (!S->gram_ready) value_error("crest_session_begin_gram first");
        std::lock_guard<std::mutex> lk(x.mu);
        int si = 0;
        const DevChunk c = session_stage(x, S, chunk, si);
        gram_rows(x, c, session_scale(S), S->prec, S->d_pad, 65536, S->d_gram, S->d_colsum);
        session_consumed(x, si);
    });
}

int crest_session_solve(crest_session *S, size_t k, double *variance, double *loadings, double *total,
                        double *colsum, double *mean, double *std_) {
    return guarded([&] {
        Ctx &x = session_ctx(S);
        if (!S->gram_ready) value_error("crest_session_begin_gram first");
        if (S->n_rows < 2) fail(CREST_E_RUNTIME, "need at least 2 cells and 1 gene");
        if (k < 1 || k > (size_t)S->d || k > 256) value_error("k must be in 1 .. min(d, 256)");
        if (!variance || !loadings || !total || !colsum) value_error("NULL argument");
        std::lock_guard<std::mutex> lk(x.mu);
        const int d = S->d, kk = (int)k;
        if (S->k != kk || !S->d_var) {
            for (void *p : {(void *)S->d_var, (void *)S->d_load, (void *)S->d_total, (void *)S->d_shift,
                            (void *)S->d_x})
                session_free(S, p);
            S->d_var = (double *)session_alloc(S, k * 8);
            S->d_load = (double *)session_alloc(S, (size_t)d * k * 8);
            S->d_total = (double *)session_alloc(S, 8);
            S->d_shift = (float *)session_alloc(S, k * 4);
            S->d_x = (float *)session_alloc(S, (size_t)S->n_rows * k * 4);
        }
        S->k = kk;
        S->solved = false;
        pca_device(x, S->d_gram, S->d_pad, S->d_colsum, d, (double)S->n_rows, kk, S->d_var, S->d_load, S->d_total);
        x.download(variance, S->d_var, k);
        x.download(loadings, S->d_load, (size_t)d * k);
        x.download(total, S->d_total, 1);
        x.download(colsum, S->d_colsum, d);
        if (mean) x.download(mean, S->d_mean, d);
        if (std_) x.download(std_, S->d_sd, d);
        x.sync();
        S->solved = true;
    });
}

int crest_session_begin_project(crest_session *S, const float *prefill, size_t k) {
    return guarded([&] {
        Ctx &x = session_ctx(S);
        if (!S->solved) value_error("crest_session_solve first");
        if (!prefill || k != (size_t)S->k) value_error("prefill must have k entries");
        std::lock_guard<std::mutex> lk(x.mu);
        HIPCHK(hipMemcpy(S->d_shift, prefill, k * 4, hipMemcpyHostToDevice));
        launch_fill_rows(S->d_shift, S->k, S->n_rows, S->d_x, x.st);
        x.check_launch();
        S->projecting = true;
        S->cmap_src = nullptr;  // each pass reads its maps afresh
        S->gmap_src = nullptr;
    });
}

int crest_session_push_project(crest_session *S, const crest_chunk *chunk) {
    return guarded([&] {
        Ctx &x = session_ctx(S);
        if (!S->projecting) value_error("crest_session_begin_project first");
        std::lock_guard<std::mutex> lk(x.mu);
        int si = 0;
        const DevChunk c = session_stage(x, S, chunk, si);
        launch_project(c, session_scale(S), S->d_load, S->k, S->d_x, 0, x.st);
        session_consumed(x, si);
    });
}

int crest_session_x_pca(crest_session *S, float *out, size_t out_len) {
    return guarded([&] {
        Ctx &x = session_ctx(S);
        if (!S->projecting) value_error("crest_session_begin_project first");
        if (!out || out_len != (size_t)S->n_rows * S->k) value_error("out must be (n_rows, k)");
        std::lock_guard<std::mutex> lk(x.mu);
        x.download(out, S->d_x, out_len);
        x.sync();
    });
}
