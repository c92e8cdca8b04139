/*
 * Seeded negative-binomial count generator for the synthetic benchmark inputs.
 * Entry (cell c, gene g) draws one uniform from a counter-based hash of (seed, c, g)
 * and inverts the NB(mean mu_cg, size r) CDF with it, so the matrix is a pure
 * function of the seed, whatever the thread count.  Two passes: row counts, then
 * fill (the fill regenerates the same values).
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>

static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline double uniform(uint64_t seed, uint64_t cell, uint64_t gene, uint64_t n_genes) {
    const uint64_t z = mix64(seed * 0x9E3779B97F4A7C15ull + mix64(cell * n_genes + gene + 1));
    return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

/* NB(mean mu, size r): P(0) = (r/(r+mu))^r, P(k)/P(k-1) = (k-1+r)/k * mu/(r+mu). */
static inline uint32_t nb_inv(double u, double mu, double r) {
    if (mu <= 0.0) return 0;
    const double q = mu / (r + mu);
    double p = pow(r / (r + mu), r);
    double cdf = p;
    if (u < cdf) return 0;
    uint32_t k = 0;
    while (u >= cdf && k < 100000000u) {
        ++k;
        p *= ((double)k - 1.0 + r) / (double)k * q;
        cdf += p;
        if (p < 1e-300) break; /* past the mode: the rest of the mass is below rounding */
    }
    return k;
}

/* prob[cluster * n_genes + g] = m_{cl,g} / Σ_g m_{cl,g};  mu = lib[c] * prob. */
void synth_count(int64_t n_cells, int64_t n_genes, const int32_t *cluster, const double *lib, const double *prob,
                 double size, uint64_t seed, int64_t *row_nnz, int threads) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
    for (int64_t c = 0; c < n_cells; ++c) {
        const double *pr = prob + (int64_t)cluster[c] * n_genes;
        int64_t cnt = 0;
        for (int64_t g = 0; g < n_genes; ++g) {
            const double mu = lib[c] * pr[g];
            const double u = uniform(seed, (uint64_t)c, (uint64_t)g, (uint64_t)n_genes);
            const double p0 = pow(size / (size + mu), size);
            if (u >= p0) ++cnt;
        }
        row_nnz[c] = cnt;
    }
}

void synth_fill(int64_t n_cells, int64_t n_genes, const int32_t *cluster, const double *lib, const double *prob,
                double size, uint64_t seed, const int64_t *indptr, uint32_t *indices, float *data, int threads) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
    for (int64_t c = 0; c < n_cells; ++c) {
        const double *pr = prob + (int64_t)cluster[c] * n_genes;
        int64_t o = indptr[c];
        for (int64_t g = 0; g < n_genes; ++g) {
            const double mu = lib[c] * pr[g];
            const double u = uniform(seed, (uint64_t)c, (uint64_t)g, (uint64_t)n_genes);
            const uint32_t k = nb_inv(u, mu, size);
            if (k) {
                indices[o] = (uint32_t)g;
                data[o] = (float)k;
                ++o;
            }
        }
    }
}
