/*
 * ORACLE — test infrastructure, never the product.  A plain-C restatement of the
 * reference's CPU algorithms for the counts -> scale -> PCA -> kNN path, used by
 * tests/ as the parity checker and by bench.py's cpu_baseline leg as the timed
 * CPU baseline.  Nothing in crest_amd/ may load it.
 *
 * Each function cites the reference lines it restates.  Summation orders follow
 * the reference's sequential order within one rayon part (one part here); the
 * OpenMP paths parallelise exactly where the reference does (thread-owned Gram row
 * blocks, query blocks of the kNN) without changing any per-element order.
 *
 * log1p of the lazy transform (kernels.rs:81 `y.ln_1p()`, i.e. the C library's
 * log1pf): mode 1 (default) calls log1pf, the reference's own rounding; mode 0 is
 * the correctly rounded float of the double log1p, which on glibc 2.35 differs from
 * mode 1 by one ulp on 9,149,209 of the floats >= 0.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct oc_chunk {
    const uint32_t *genes;
    const float *values;
    size_t nnz;
    const int64_t *cell_map;
    size_t n_cell_map;
    const int32_t *gene_map;
    size_t n_gene_map;
    double target_sum;
    int32_t log1p;
    const uint32_t *cells;
    const int64_t *indptr;
    size_t n_indptr;
    uint32_t first_cell;
} oc_chunk;

static int g_log1p_mode = 1;
void oc_set_log1p_mode(int mode) { g_log1p_mode = mode; }

static float ln1p_f32(float y) { return g_log1p_mode ? log1pf(y) : (float)log1p((double)y); }

/* ------------------------------------------------------------------ spans */
/* kernels.rs:36-59 cell_spans: (cell id, lo, hi) per cell present in the chunk. */
typedef struct span { uint32_t cell; size_t lo, hi; } span;

static size_t oc_spans(const oc_chunk *v, span **out) {
    span *s = NULL;
    size_t n = 0;
    if (v->indptr) {
        const int64_t base = v->n_indptr ? v->indptr[0] : 0;
        n = v->n_indptr ? v->n_indptr - 1 : 0;
        s = (span *)malloc((n ? n : 1) * sizeof(span));
        for (size_t i = 0; i < n; ++i) {
            s[i].cell = v->first_cell + (uint32_t)i;
            s[i].lo = (size_t)(v->indptr[i] - base);
            s[i].hi = (size_t)(v->indptr[i + 1] - base);
        }
    } else {
        s = (span *)malloc((v->nnz ? v->nnz : 1) * sizeof(span));
        size_t p = 0;
        while (p < v->nnz) {
            uint32_t c = v->cells[p];
            size_t q = p + 1;
            while (q < v->nnz && v->cells[q] == c) ++q;
            s[n].cell = c;
            s[n].lo = p;
            s[n].hi = q;
            ++n;
            p = q;
        }
    }
    *out = s;
    return n;
}

/* kernels.rs:61-85 ChunkView::cell: kept entries (var, transformed value) of one
 * cell into buf; returns the row or -1. */
static int64_t oc_cell(const oc_chunk *v, const span *sp, uint32_t *bv, float *bx, size_t *bn) {
    *bn = 0;
    if ((size_t)sp->cell >= v->n_cell_map) return -1;
    const int64_t row = v->cell_map[sp->cell];
    if (row < 0) return -1;
    double total = 0.0;
    for (size_t e = sp->lo; e < sp->hi; ++e) {
        const uint32_t g = v->genes[e];
        const int32_t m = (size_t)g < v->n_gene_map ? v->gene_map[g] : -1;
        if (m >= 0) {
            bv[*bn] = (uint32_t)m;
            bx[*bn] = v->values[e];
            ++*bn;
            total += (double)v->values[e];
        }
    }
    if (v->target_sum > 0.0 || v->log1p) {
        const double scale = (v->target_sum > 0.0 && total > 0.0) ? v->target_sum / total : 1.0;
        for (size_t i = 0; i < *bn; ++i) {
            const float y = v->target_sum > 0.0 ? (float)((double)bx[i] * scale) : bx[i];
            bx[i] = v->log1p ? ln1p_f32(y) : y;
        }
    }
    return row;
}

static size_t max_span(const span *s, size_t n) {
    size_t m = 1;
    for (size_t i = 0; i < n; ++i)
        if (s[i].hi - s[i].lo > m) m = s[i].hi - s[i].lo;
    return m;
}

/* kernels.rs:122-139 materialize (one part: chunk order). */
size_t oc_materialize(const oc_chunk *v, uint32_t *rows, uint32_t *vars, float *vals) {
    span *s;
    const size_t ns = oc_spans(v, &s);
    const size_t mx = max_span(s, ns);
    uint32_t *bv = malloc(mx * sizeof(uint32_t));
    float *bx = malloc(mx * sizeof(float));
    size_t out = 0;
    for (size_t i = 0; i < ns; ++i) {
        size_t bn;
        const int64_t row = oc_cell(v, &s[i], bv, bx, &bn);
        if (row < 0) continue;
        for (size_t j = 0; j < bn; ++j) {
            rows[out] = (uint32_t)row;
            vars[out] = bv[j];
            vals[out] = bx[j];
            ++out;
        }
    }
    free(bv);
    free(bx);
    free(s);
    return out;
}

/* kernels.rs:141-160 gene_stats: out[g*5 + .] += [x, x^2, expm1 x, expm1(x)^2, 1]. */
void oc_gene_stats(const oc_chunk *v, double *out) {
    span *s;
    const size_t ns = oc_spans(v, &s);
    const size_t mx = max_span(s, ns);
    uint32_t *bv = malloc(mx * sizeof(uint32_t));
    float *bx = malloc(mx * sizeof(float));
    for (size_t i = 0; i < ns; ++i) {
        size_t bn;
        if (oc_cell(v, &s[i], bv, bx, &bn) < 0) continue;
        for (size_t j = 0; j < bn; ++j) {
            const double x = (double)bx[j], ex = expm1(x);
            double *o = out + (size_t)bv[j] * 5;
            o[0] += x;
            o[1] += x * x;
            o[2] += ex;
            o[3] += ex * ex;
            o[4] += 1.0;
        }
    }
    free(bv);
    free(bx);
    free(s);
}

/* kernels.rs:162-195 qc: per row (cell_map) the total, the count and the flagged
 * total of the kept non-zero values; per gene the count and the total.  Sequential,
 * in chunk order (the reference's order inside a cell). */
void oc_qc(const oc_chunk *v, const uint8_t *flag, double *cell_total, uint32_t *cell_ngenes, double *cell_flag,
           uint32_t *gene_ncells, double *gene_total) {
    span *s;
    const size_t ns = oc_spans(v, &s);
    const size_t mx = max_span(s, ns);
    uint32_t *bv = malloc(mx * sizeof(uint32_t));
    float *bx = malloc(mx * sizeof(float));
    for (size_t i = 0; i < ns; ++i) {
        size_t bn;
        const int64_t row = oc_cell(v, &s[i], bv, bx, &bn);
        if (row < 0) continue;
        double t = 0.0, f = 0.0;
        uint32_t n = 0;
        for (size_t j = 0; j < bn; ++j) {
            const float x = bx[j];
            if (x == 0.0f) continue;
            t += (double)x;
            n += 1;
            if (flag[bv[j]]) f += (double)x;
            gene_ncells[bv[j]] += 1;
            gene_total[bv[j]] += (double)x;
        }
        cell_total[row] += t;
        cell_ngenes[row] += n;
        cell_flag[row] += f;
    }
    free(bv);
    free(bx);
    free(s);
}

/* kernels.rs:209-218 Transform::apply */
static double clip_mv(double z, double mv) {
    if (mv > 0.0) {
        if (z < -mv) return -mv;
        if (z > mv) return mv;
    }
    return z;
}
static double apply_t(int scale, const double *mean, const double *sd, double mv, size_t g, float x) {
    if (!scale) return (double)x;
    const double mu = mean[g], s = sd[g];
    return clip_mv(((double)x - mu) / s, mv) - clip_mv(-mu / s, mv);
}

typedef struct row_ent { uint32_t g; double x; } row_ent;
static int cmp_ent(const void *a, const void *b) {
    const row_ent *p = a, *q = b;
    return p->g < q->g ? -1 : (p->g > q->g ? 1 : 0);
}

/* kernels.rs:221-309 gram_accumulate.  Rows of (hvg, s) sorted by hvg; colsum in
 * entry order; the upper triangle G[a][b] += s_a s_b row by row, with gene rows a
 * owned by threads (as the reference's row blocks), so each entry sees the
 * sequential order whatever the thread count. */
void oc_gram_accumulate(const oc_chunk *v, const int32_t *sub, double *gram, double *colsum, size_t d, int scale,
                        const double *mean, const double *sd, double mv, int n_threads) {
    span *s;
    const size_t ns = oc_spans(v, &s);
    const size_t mx = max_span(s, ns);
    uint32_t *bv = malloc(mx * sizeof(uint32_t));
    float *bx = malloc(mx * sizeof(float));
    size_t cap = v->nnz ? v->nnz : 1, ne = 0;
    row_ent *ent = malloc(cap * sizeof(row_ent));
    size_t *starts = malloc((ns + 1) * sizeof(size_t));
    size_t nr = 0;
    starts[0] = 0;
    for (size_t i = 0; i < ns; ++i) {
        size_t bn;
        if (oc_cell(v, &s[i], bv, bx, &bn) < 0) continue;
        const size_t st = ne;
        for (size_t j = 0; j < bn; ++j) {
            const int32_t m = sub[bv[j]];
            if (m >= 0) {
                ent[ne].g = (uint32_t)m;
                ent[ne].x = apply_t(scale, mean, sd, mv, (size_t)m, bx[j]);
                ++ne;
            }
        }
        if (ne > st) {
            int sorted = 1;
            for (size_t j = st + 1; j < ne; ++j)
                if (ent[j - 1].g > ent[j].g) sorted = 0;
            if (!sorted) { /* stable: equal genes keep their order (none in canonical CSR) */
                for (size_t a = st + 1; a < ne; ++a) {
                    row_ent t = ent[a];
                    size_t b = a;
                    while (b > st && ent[b - 1].g > t.g) {
                        ent[b] = ent[b - 1];
                        --b;
                    }
                    ent[b] = t;
                }
            }
            starts[++nr] = ne;
        }
    }
    for (size_t e = 0; e < ne; ++e) colsum[ent[e].g] += ent[e].x;
    if (n_threads < 1) n_threads = 1;
#pragma omp parallel num_threads(n_threads)
    {
        const int t = omp_get_thread_num(), nt = omp_get_num_threads();
        for (size_t r = 0; r < nr; ++r) {
            const size_t lo = starts[r], hi = starts[r + 1];
            for (size_t k = lo; k < hi; ++k) {
                const size_t a = ent[k].g;
                if ((int)(a % (size_t)nt) != t) continue;
                const double va = ent[k].x;
                double *o = gram + a * d;
                for (size_t m = k; m < hi; ++m) o[ent[m].g] += va * ent[m].x;
            }
        }
    }
    free(ent);
    free(starts);
    free(bv);
    free(bx);
    free(s);
}

/* The colsum and the Gram diagonal of gram_accumulate alone (kernels.rs:258-260 and
 * the a == b terms of :291-308): s1[m] += s, s2[m] += s * s in entry order, for a
 * checker that needs only the covariance trace. */
void oc_scaled_moments(const oc_chunk *v, const int32_t *sub, double *s1, double *s2, int scale, const double *mean,
                       const double *sd, double mv) {
    span *s;
    const size_t ns = oc_spans(v, &s);
    const size_t mx = max_span(s, ns);
    uint32_t *bv = malloc((mx ? mx : 1) * sizeof(uint32_t));
    float *bx = malloc((mx ? mx : 1) * sizeof(float));
    for (size_t i = 0; i < ns; ++i) {
        size_t bn;
        if (oc_cell(v, &s[i], bv, bx, &bn) < 0) continue;
        for (size_t j = 0; j < bn; ++j) {
            const int32_t m = sub[bv[j]];
            if (m < 0) continue;
            const double x = apply_t(scale, mean, sd, mv, (size_t)m, bx[j]);
            s1[m] += x;
            s2[m] += x * x;
        }
    }
    free(bv);
    free(bx);
    free(s);
}

/* kernels.rs:319-323: cov = (G[min,max] - c_a c_b / n) / (n - 1); total = trace. */
void oc_cov_from_gram(const double *gram, const double *colsum, size_t d, size_t n, double *cov, double *total) {
    const double nf = (double)n;
    double t = 0.0;
    for (size_t i = 0; i < d; ++i)
        for (size_t j = 0; j < d; ++j) {
            const size_t a = i <= j ? i : j, b = i <= j ? j : i;
            cov[i * d + j] = (gram[a * d + b] - colsum[a] * colsum[b] / nf) / (nf - 1.0);
        }
    for (size_t i = 0; i < d; ++i) t += cov[i * d + i];
    *total = t;
}

/* kernels.rs:350-371 project: out[row] += (f32) Σ_e s_e V[h_e, :] (f64, entry order). */
void oc_project(const oc_chunk *v, const int32_t *sub, const double *V, size_t k, float *out, int scale,
                const double *mean, const double *sd, double mv) {
    span *s;
    const size_t ns = oc_spans(v, &s);
    const size_t mx = max_span(s, ns);
    uint32_t *bv = malloc(mx * sizeof(uint32_t));
    float *bx = malloc(mx * sizeof(float));
    double *acc = malloc((k ? k : 1) * sizeof(double));
    for (size_t i = 0; i < ns; ++i) {
        size_t bn;
        const int64_t row = oc_cell(v, &s[i], bv, bx, &bn);
        if (row < 0) continue;
        for (size_t j = 0; j < k; ++j) acc[j] = 0.0;
        for (size_t j = 0; j < bn; ++j) {
            const int32_t m = sub[bv[j]];
            if (m < 0) continue;
            const double x = apply_t(scale, mean, sd, mv, (size_t)m, bx[j]);
            const double *vr = V + (size_t)m * k;
            for (size_t c = 0; c < k; ++c) acc[c] += x * vr[c];
        }
        float *o = out + (size_t)row * k;
        for (size_t c = 0; c < k; ++c) o[c] += (float)acc[c];
    }
    free(acc);
    free(bv);
    free(bx);
    free(s);
}

/* ------------------------------------------------------------------ kNN */
/* knn.rs:48-66 sq_dist: 8 f32 partial sums, summed in order, then the tail. */
static float sq_dist(const float *a, const float *b, size_t p) {
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const size_t m = p - p % 8;
    for (size_t i = 0; i < m; i += 8)
        for (int l = 0; l < 8; ++l) {
            const float t = a[i + l] - b[i + l];
            acc[l] += t * t;
        }
    float s = acc[0];
    for (int l = 1; l < 8; ++l) s += acc[l];
    for (size_t i = m; i < p; ++i) {
        const float t = a[i] - b[i];
        s += t * t;
    }
    return s;
}

/* knn.rs:68-87 push_topk */
static void push_topk(float *dist, uint32_t *idx, size_t k, float d, uint32_t j) {
    if (d >= dist[k - 1]) return;
    for (size_t i = 0; i < k; ++i)
        if (idx[i] == j) return;
    size_t p = k - 1;
    while (p > 0 && dist[p - 1] > d) {
        dist[p] = dist[p - 1];
        idx[p] = idx[p - 1];
        --p;
    }
    dist[p] = d;
    idx[p] = j;
}

typedef struct dpair { float d; uint32_t j; } dpair;
static int cmp_pair(const void *a, const void *b) {
    const dpair *p = a, *q = b;
    if (p->d < q->d) return -1;
    if (p->d > q->d) return 1;
    return p->j < q->j ? -1 : (p->j > q->j ? 1 : 0);
}

/* f32 dot product in eight partial sums (the screen's GEMM; faer's own order is
 * not restated). */
static float dot8(const float *a, const float *b, size_t p) {
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const size_t m = p - p % 8;
    for (size_t i = 0; i < m; i += 8)
        for (int l = 0; l < 8; ++l) acc[l] += a[i + l] * b[i + l];
    float s = acc[0];
    for (int l = 1; l < 8; ++l) s += acc[l];
    for (size_t i = m; i < p; ++i) s += a[i] * b[i];
    return s;
}

/* knn.rs:93-151,337-369: exact brute force for the queries q0..q1 (all n as
 * candidates) in 256-query blocks and 1024-candidate tiles, screen d = max(qn + cn
 * - 2 q.c, 0), push_topk, then exact f32 rerank and (distance, index) sort.
 * idx/dist are [(q1-q0), k] with k = the clamped k of knn.rs:342. */
void oc_knn(const float *x, size_t n, size_t p, size_t k, size_t q0, size_t q1, uint32_t *idx, float *dist,
            int n_threads) {
    if (n == 0 || k == 0) return;
    float *norms = malloc(n * sizeof(float));
    for (size_t i = 0; i < n; ++i) {
        float s = 0.0f;
        for (size_t f = 0; f < p; ++f) s += x[i * p + f] * x[i * p + f];
        norms[i] = s;
    }
    const size_t nq = q1 - q0, qb = 256, tile = 1024;
    const size_t nblk = (nq + qb - 1) / qb;
    if (n_threads < 1) n_threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads)
    for (size_t b = 0; b < nblk; ++b) {
        const size_t lo = q0 + b * qb, hi = lo + qb < q1 ? lo + qb : q1;
        float *td = dist + (lo - q0) * k;
        uint32_t *ti = idx + (lo - q0) * k;
        for (size_t i = 0; i < (hi - lo) * k; ++i) {
            td[i] = INFINITY;
            ti[i] = UINT32_MAX;
        }
        for (size_t c0 = 0; c0 < n; c0 += tile) {
            const size_t c1 = c0 + tile < n ? c0 + tile : n;
            for (size_t q = lo; q < hi; ++q) {
                const float qn = norms[q];
                float *d = td + (q - lo) * k;
                uint32_t *j = ti + (q - lo) * k;
                for (size_t c = c0; c < c1; ++c) {
                    if (c == q) continue;
                    float v = qn + norms[c] - 2.0f * dot8(x + q * p, x + c * p, p);
                    if (v < 0.0f) v = 0.0f;
                    if (v < d[k - 1]) push_topk(d, j, k, v, (uint32_t)c);
                }
            }
        }
        /* exact distances, re-sorted (knn.rs:356-367) */
        dpair *pr = malloc(k * sizeof(dpair));
        for (size_t q = lo; q < hi; ++q) {
            float *d = td + (q - lo) * k;
            uint32_t *j = ti + (q - lo) * k;
            for (size_t s = 0; s < k; ++s) {
                pr[s].j = j[s];
                pr[s].d = j[s] == UINT32_MAX ? INFINITY : sqrtf(sq_dist(x + q * p, x + (size_t)j[s] * p, p));
            }
            qsort(pr, k, sizeof(dpair), cmp_pair);
            for (size_t s = 0; s < k; ++s) {
                d[s] = pr[s].d;
                j[s] = pr[s].j;
            }
        }
        free(pr);
    }
    free(norms);
}

/* The reference's exact distance (knn.rs:48-66) of arbitrary pairs, for checks. */
void oc_pair_dist(const float *x, size_t p, const uint32_t *a, const uint32_t *b, size_t m, float *out) {
    for (size_t i = 0; i < m; ++i) out[i] = sqrtf(sq_dist(x + (size_t)a[i] * p, x + (size_t)b[i] * p, p));
}

/* ------------------------------------------------------------------ connectivities */
/* umap/graph.rs:29-58 smooth_knn_dist (local_connectivity = 1). */
static void smooth_knn_dist(const float *dists, size_t k, size_t n_neighbors, double mean_all, double *rho_out,
                            double *sigma_out) {
    const double target = log2((double)n_neighbors);
    double rho = 0.0;
    for (size_t i = 0; i < k; ++i)
        if (dists[i] > 0.0f) {
            rho = (double)dists[i];
            break;
        }
    double lo = 0.0, hi = INFINITY, mid = 1.0;
    for (int it = 0; it < 64; ++it) {
        double psum = 0.0;
        for (size_t i = 0; i < k; ++i) {
            const double v = (double)dists[i] - rho;
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
    for (size_t i = 0; i < k; ++i) s += (double)dists[i];
    const double mean_i = s / (double)(k + 1);
    const double floor_ = rho > 0.0 ? 1e-3 * mean_i : 1e-3 * mean_all;
    *rho_out = rho;
    *sigma_out = mid > floor_ ? mid : floor_;
}

typedef struct kentry { uint64_t key; float w; } kentry;
static int cmp_key(const void *a, const void *b) {
    const kentry *p = a, *q = b;
    return p->key < q->key ? -1 : (p->key > q->key ? 1 : 0);
}

/* umap/graph.rs:61-107 fuzzy_simplicial_set + py.rs:92-97 (edges with source < target,
 * in key order).  Returns the number of edges written. */
size_t oc_connectivities(const uint32_t *idx, const float *dist, size_t n, size_t k, size_t n_neighbors,
                         uint32_t *rows, uint32_t *cols, float *w) {
    if (n == 0 || k == 0) return 0;
    double sa = 0.0;
    for (size_t i = 0; i < n * k; ++i) sa += (double)dist[i];
    const double mean_all = sa / (double)(n * (k + 1));
    kentry *e = malloc(n * k * sizeof(kentry));
    size_t m = 0;
    for (size_t i = 0; i < n; ++i) {
        double rho, sigma;
        smooth_knn_dist(dist + i * k, k, n_neighbors, mean_all, &rho, &sigma);
        for (size_t t = 0; t < k; ++t) {
            const uint32_t j = idx[i * k + t];
            if (j == UINT32_MAX || (size_t)j == i) continue;
            const double v = (double)dist[i * k + t] - rho;
            const double ww = (v <= 0.0 || sigma == 0.0) ? 1.0 : exp(-v / sigma);
            const uint64_t a = (uint32_t)i < j ? (uint64_t)i : (uint64_t)j;
            const uint64_t b = (uint32_t)i < j ? (uint64_t)j : (uint64_t)i;
            e[m].key = (a << 32) | b;
            e[m].w = (float)ww;
            ++m;
        }
    }
    qsort(e, m, sizeof(kentry), cmp_key);
    size_t out = 0, p = 0;
    while (p < m) {
        const uint64_t key = e[p].key;
        float ww = e[p].w;
        size_t q = p + 1;
        while (q < m && e[q].key == key) {
            const float b = e[q].w;
            ww = ww + b - ww * b;
            ++q;
        }
        if (ww > 0.0f) {
            rows[out] = (uint32_t)(key >> 32);
            cols[out] = (uint32_t)(key & 0xffffffffu);
            w[out] = ww;
            ++out;
        }
        p = q;
    }
    free(e);
    return out;
}

/* ------------------------------------------------------------------ FP64 exact kNN check */
/* Checker helper (no reference counterpart): per row r of a block of FP64 dot
 * products G[r][c] (query r with candidate col0 + c), the squared distance
 * d = (qn[r] + cn[c]) - 2 G[r][c]; keep the m smallest as a list sorted by
 * (distance, candidate index), skipping the row's own index self[r] (none when self
 * is NULL).  best_d / best_i [rows, m] carry the lists across blocks (start with
 * +inf / -1).  Ties keep the smaller index, so the lists equal a full sort by
 * (distance, index). */
void oc_topk_update(const double *G, const double *qn, const double *cn, size_t rows, size_t cols, int64_t col0,
                    const int64_t *self, size_t m, double *best_d, int64_t *best_i, int threads) {
#pragma omp parallel for schedule(static) num_threads(threads)
    for (size_t r = 0; r < rows; ++r) {
        const double *row = G + r * cols;
        double *bd = best_d + r * m;
        int64_t *bi = best_i + r * m;
        const int64_t me = self ? self[r] : -1;
        const double q = qn[r];
        double worst = bd[m - 1];
        for (size_t c = 0; c < cols; ++c) {
            const double d = (q + cn[c]) - 2.0 * row[c];
            if (!(d <= worst)) continue;
            const int64_t j = col0 + (int64_t)c;
            if (j == me) continue;
            if (d == worst && bi[m - 1] >= 0 && j > bi[m - 1]) continue;
            size_t pos = m - 1;
            while (pos > 0 && (bd[pos - 1] > d || (bd[pos - 1] == d && bi[pos - 1] > j))) {
                bd[pos] = bd[pos - 1];
                bi[pos] = bi[pos - 1];
                --pos;
            }
            bd[pos] = d;
            bi[pos] = j;
            worst = bd[m - 1];
        }
    }
}

/* Exact FP64 squared distances sum_j (q_j - x_j)^2 of m (query, candidate) pairs of
 * rows of x [n, p] (f32 rows widened to f64, summed in column order). */
void oc_pair_dist_f64(const float *x, size_t p, const int64_t *a, const int64_t *b, size_t m, double *out,
                      int threads) {
#pragma omp parallel for schedule(static) num_threads(threads)
    for (size_t i = 0; i < m; ++i) {
        const float *u = x + (size_t)a[i] * p, *v = x + (size_t)b[i] * p;
        double s = 0.0;
        for (size_t j = 0; j < p; ++j) {
            const double t = (double)u[j] - (double)v[j];
            s += t * t;
        }
        out[i] = s;
    }
}

/* Checker helper: the FP64 screen of oc_topk_update for query rows q[0..m) of x
 * [n, p] against every row of x, without materialising dot products: blocks of 32
 * queries stream all candidates once (queries transposed so the 32 dot products of
 * one candidate are independent vector lanes).  d = (qn + cn) - 2 q.x in FP64 with
 * fused multiply-adds (a screen: the kept candidates are recomputed exactly by the
 * caller).  best_d / best_i [m, mk] as oc_topk_update; self excluded. */
#define OC_QB 32
void oc_knn_screen_f64(const float *x, size_t n, size_t p, const int64_t *q, size_t m, size_t mk, double *best_d,
                       int64_t *best_i, int threads) {
    double *cn = (double *)malloc(n * sizeof(double));
#pragma omp parallel for schedule(static) num_threads(threads)
    for (size_t c = 0; c < n; ++c) {
        double s = 0.0;
        for (size_t j = 0; j < p; ++j) s = fma((double)x[c * p + j], (double)x[c * p + j], s);
        cn[c] = s;
    }
    const size_t nb = (m + OC_QB - 1) / OC_QB;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (size_t b = 0; b < nb; ++b) {
        const size_t q0 = b * OC_QB, nq = (m - q0) < OC_QB ? (m - q0) : OC_QB;
        double *qt = (double *)calloc(p * OC_QB, sizeof(double));
        double qn[OC_QB];
        int64_t self[OC_QB];
        for (size_t i = 0; i < OC_QB; ++i) {
            qn[i] = INFINITY;
            self[i] = -1;
        }
        for (size_t i = 0; i < nq; ++i) {
            const float *r = x + (size_t)q[q0 + i] * p;
            qn[i] = cn[q[q0 + i]];
            self[i] = q[q0 + i];
            for (size_t j = 0; j < p; ++j) qt[j * OC_QB + i] = (double)r[j];
        }
        double worst[OC_QB];
        for (size_t i = 0; i < nq; ++i) worst[i] = best_d[(q0 + i) * mk + mk - 1];
        for (size_t c = 0; c < n; ++c) {
            const float *xc = x + c * p;
            double acc[OC_QB];
            for (size_t i = 0; i < OC_QB; ++i) acc[i] = 0.0;
            for (size_t j = 0; j < p; ++j) {
                const double v = (double)xc[j];
                const double *t = qt + j * OC_QB;
                for (size_t i = 0; i < OC_QB; ++i) acc[i] = fma(t[i], v, acc[i]);
            }
            for (size_t i = 0; i < nq; ++i) {
                const double d = (qn[i] + cn[c]) - 2.0 * acc[i];
                if (!(d <= worst[i]) || (int64_t)c == self[i]) continue;
                double *bd = best_d + (q0 + i) * mk;
                int64_t *bi = best_i + (q0 + i) * mk;
                if (d == worst[i] && bi[mk - 1] >= 0 && (int64_t)c > bi[mk - 1]) continue;
                size_t pos = mk - 1;
                while (pos > 0 && (bd[pos - 1] > d || (bd[pos - 1] == d && bi[pos - 1] > (int64_t)c))) {
                    bd[pos] = bd[pos - 1];
                    bi[pos] = bi[pos - 1];
                    --pos;
                }
                bd[pos] = d;
                bi[pos] = (int64_t)c;
                worst[i] = bd[mk - 1];
            }
        }
        free(qt);
    }
    free(cn);
}
