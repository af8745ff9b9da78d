/*
 * Synthetic single-cell counts for the benchmarks and tests (host C, OpenMP).
 *
 * The model: gene means mu_g ~ LogNormal(mu_log, sigma_log) times `mean_scale`
 * (chosen so the layer reaches the asked density), cell size factors
 * s_c ~ LogNormal(0, size_sigma), cell types t_c uniform over `n_types`, each type
 * with N(0, 1) natural-log fold changes on a random `marker_share` of the genes, and
 * counts ~ NegativeBinomial(mean m = mu_g s_c exp(lfc_{t_c,g}), dispersion phi):
 * variance m + phi m^2, P(0) = (1 + m phi)^(-1/phi), drawn by inverting the CDF.
 *
 * Every draw is a pure function of (seed, stream, index) through SplitMix64, so the
 * matrix does not depend on the thread count. Output is CSR (u64 indptr, u32 gene
 * indices increasing within a row, f32 counts).
 *
 * Also: kyto's input layer, normalize_total(target_sum) then log1p, with kyto's
 * steps (crates/kyto-exec/src/elementwise.rs:262-269 divisor, kernels.rs DivideRows
 * f32 division, kernels/log1p.rs: ln(1 + x) evaluated in f64 and rounded once).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline uint64_t sm_mix(uint64_t v) {
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ULL;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBULL;
    return v ^ (v >> 31);
}

/* A uniform in (0, 1) from (seed, stream, index). */
static inline double uni(uint64_t seed, uint64_t stream, uint64_t index) {
    uint64_t h = sm_mix(seed ^ sm_mix(stream * 0x9E3779B97F4A7C15ULL + sm_mix(index + 0x632BE59BD9B4E019ULL)));
    return ((double)(h >> 11) + 0.5) * 0x1p-53;
}

static inline double normal(uint64_t seed, uint64_t stream, uint64_t index) {
    double u1 = uni(seed, stream, 2 * index), u2 = uni(seed, stream, 2 * index + 1);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

enum { S_GENE_MEAN = 1, S_SIZE = 2, S_TYPE = 3, S_MARKER = 4, S_LFC = 5, S_COUNT = 6 };

typedef struct {
    uint64_t n_cells;
    uint32_t n_genes;
    uint32_t n_types;
    uint64_t seed;
    double mu_log, sigma_log, size_sigma, phi, marker_share, mean_scale;
} ks_spec;

/* Per-gene means (scaled) and per-type multiplicative folds (n_types x n_genes). */
static void gene_tables(const ks_spec *s, double *mu, double *fold) {
    for (uint32_t g = 0; g < s->n_genes; g++)
        mu[g] = s->mean_scale * exp(s->mu_log + s->sigma_log * normal(s->seed, S_GENE_MEAN, g));
    for (uint32_t t = 0; t < s->n_types; t++)
        for (uint32_t g = 0; g < s->n_genes; g++) {
            uint64_t i = (uint64_t)t * s->n_genes + g;
            double f = 1.0;
            if (uni(s->seed, S_MARKER, i) < s->marker_share) f = exp(normal(s->seed, S_LFC, i));
            fold[i] = f;
        }
}

static inline double cell_size(const ks_spec *s, uint64_t c) {
    return exp(s->size_sigma * normal(s->seed, S_SIZE, c));
}
static inline uint32_t cell_type(const ks_spec *s, uint64_t c) {
    return (uint32_t)(uni(s->seed, S_TYPE, c) * s->n_types) % s->n_types;
}

/* One count: NB(mean m, dispersion phi) by CDF inversion of the uniform u. */
static inline uint32_t nb_count(double m, double phi, double u) {
    if (m <= 0.0) return 0;
    /* P(0) >= 1 - m (Bernoulli's inequality): most draws stop here. */
    if (u < 1.0 - m) return 0;
    double r = 1.0 / phi;
    double p0 = exp(-r * log1p(m * phi));
    if (u < p0) return 0;
    double q = m * phi / (1.0 + m * phi); /* success share */
    double pk = p0, cdf = p0;
    uint32_t k = 0;
    while (cdf <= u && k < 100000) {
        pk *= (k + r) / (double)(k + 1) * q;
        k++;
        cdf += pk;
        if (pk < 1e-300) break;
    }
    return k;
}

/* Count of stored values per row (pass 1) or the values themselves (pass 2). */
static void generate(const ks_spec *s, const double *mu, const double *fold, uint64_t *row_nnz,
                     const uint64_t *indptr, uint32_t *indices, float *values) {
    #pragma omp parallel for schedule(dynamic, 256)
    for (uint64_t c = 0; c < s->n_cells; c++) {
        double size = cell_size(s, c);
        const double *f = fold + (uint64_t)cell_type(s, c) * s->n_genes;
        uint64_t at = indptr ? indptr[c] : 0, cnt = 0;
        for (uint32_t g = 0; g < s->n_genes; g++) {
            double m = mu[g] * size * f[g];
            double u = uni(s->seed, S_COUNT, c * (uint64_t)s->n_genes + g);
            uint32_t k = nb_count(m, s->phi, u);
            if (k) {
                if (indptr) {
                    indices[at] = g;
                    values[at] = (float)k;
                    at++;
                }
                cnt++;
            }
        }
        if (row_nnz) row_nnz[c] = cnt;
    }
}

/* Expected share of stored values over the first `sample_cells` cells. */
double ks_expected_density(const ks_spec *s, uint64_t sample_cells) {
    double *mu = (double *)malloc(sizeof(double) * s->n_genes);
    double *fold = (double *)malloc(sizeof(double) * s->n_genes * s->n_types);
    gene_tables(s, mu, fold);
    uint64_t nc = sample_cells < s->n_cells ? sample_cells : s->n_cells;
    double total = 0.0;
    #pragma omp parallel for reduction(+ : total) schedule(static)
    for (uint64_t c = 0; c < nc; c++) {
        double size = cell_size(s, c);
        const double *f = fold + (uint64_t)cell_type(s, c) * s->n_genes;
        double row = 0.0;
        for (uint32_t g = 0; g < s->n_genes; g++) {
            double m = mu[g] * size * f[g];
            row += 1.0 - exp(-log1p(m * s->phi) / s->phi);
        }
        total += row;
    }
    free(mu);
    free(fold);
    return total / ((double)nc * s->n_genes);
}

/* Pass 1: nnz per row into indptr[1..] (prefix-summed); returns nnz. */
uint64_t ks_count(const ks_spec *s, uint64_t *indptr) {
    double *mu = (double *)malloc(sizeof(double) * s->n_genes);
    double *fold = (double *)malloc(sizeof(double) * s->n_genes * s->n_types);
    gene_tables(s, mu, fold);
    generate(s, mu, fold, indptr + 1, NULL, NULL, NULL);
    indptr[0] = 0;
    for (uint64_t c = 0; c < s->n_cells; c++) indptr[c + 1] += indptr[c];
    free(mu);
    free(fold);
    return indptr[s->n_cells];
}

/* Pass 2: fill indices and values given indptr from ks_count. */
void ks_fill(const ks_spec *s, const uint64_t *indptr, uint32_t *indices, float *values) {
    double *mu = (double *)malloc(sizeof(double) * s->n_genes);
    double *fold = (double *)malloc(sizeof(double) * s->n_genes * s->n_types);
    gene_tables(s, mu, fold);
    generate(s, mu, fold, NULL, indptr, indices, values);
    free(mu);
    free(fold);
}

/* One block of rows [first, first + count) as a CSR of its own: indptr (count + 1,
 * from 0), then, when its nnz fits `capacity`, the indices and values. Returns the
 * block's nnz (a block larger than capacity is counted but not filled). The rows are
 * those of the whole matrix: draws depend on (seed, cell, gene) only. */
uint64_t ks_block(const ks_spec *s, uint64_t first, uint64_t count, uint64_t *indptr, uint32_t *indices,
                  float *values, uint64_t capacity) {
    double *mu = (double *)malloc(sizeof(double) * s->n_genes);
    double *fold = (double *)malloc(sizeof(double) * s->n_genes * s->n_types);
    gene_tables(s, mu, fold);
    uint64_t *row_nnz = (uint64_t *)malloc(sizeof(uint64_t) * (count + 1));
    #pragma omp parallel for schedule(dynamic, 64)
    for (uint64_t i = 0; i < count; i++) {
        const uint64_t c = first + i;
        double size = cell_size(s, c);
        const double *f = fold + (uint64_t)cell_type(s, c) * s->n_genes;
        uint64_t cnt = 0;
        for (uint32_t g = 0; g < s->n_genes; g++) {
            double u = uni(s->seed, S_COUNT, c * (uint64_t)s->n_genes + g);
            if (nb_count(mu[g] * size * f[g], s->phi, u)) cnt++;
        }
        row_nnz[i] = cnt;
    }
    indptr[0] = 0;
    for (uint64_t i = 0; i < count; i++) indptr[i + 1] = indptr[i] + row_nnz[i];
    const uint64_t nnz = indptr[count];
    if (nnz <= capacity) {
        #pragma omp parallel for schedule(dynamic, 64)
        for (uint64_t i = 0; i < count; i++) {
            const uint64_t c = first + i;
            double size = cell_size(s, c);
            const double *f = fold + (uint64_t)cell_type(s, c) * s->n_genes;
            uint64_t at = indptr[i];
            for (uint32_t g = 0; g < s->n_genes; g++) {
                double u = uni(s->seed, S_COUNT, c * (uint64_t)s->n_genes + g);
                uint32_t k = nb_count(mu[g] * size * f[g], s->phi, u);
                if (k) {
                    indices[at] = g;
                    values[at] = (float)k;
                    at++;
                }
            }
        }
    }
    free(row_nnz);
    free(mu);
    free(fold);
    return nnz;
}

/* Cell types (for tests that look at clusters). */
void ks_cell_types(const ks_spec *s, uint32_t *types) {
    for (uint64_t c = 0; c < s->n_cells; c++) types[c] = cell_type(s, c);
}

/* normalize_total(target_sum) then log1p, in place on `values`. */
void ks_normalize_log1p(const uint64_t *indptr, uint64_t n_rows, double target_sum, float *values) {
    #pragma omp parallel for schedule(static)
    for (uint64_t r = 0; r < n_rows; r++) {
        double lanes[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        uint64_t s = indptr[r], e = indptr[r + 1];
        for (uint64_t i = s; i < e; i++) lanes[(i - s) % 8] += (double)values[i];
        double h0 = lanes[0] + lanes[4], h1 = lanes[1] + lanes[5], h2 = lanes[2] + lanes[6],
               h3 = lanes[3] + lanes[7];
        double total = (h0 + h2) + (h1 + h3);
        float divisor = (float)total / (float)target_sum;
        if (divisor == 0.0f) divisor = 1.0f;
        for (uint64_t i = s; i < e; i++) {
            float v = values[i] / divisor;
            values[i] = (float)log1p((double)v);
        }
    }
}
