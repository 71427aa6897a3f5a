/*
 * Deterministic synthetic corpora (benchmark and test data; no reference
 * counterpart -- the reference benchmarks read MS MARCO files instead,
 * benchmarks/index_search_benchmark.cpp:197-280).
 *
 * Two passes: nsparse_synth_row_nnz gives every row's length (build indptr from
 * it), nsparse_synth_fill writes the rows. Row r of a corpus depends only on
 * (spec, r), so any slice [first_row, first_row + n_rows) can be made alone.
 */
#ifndef NSPARSE_SYNTH_H
#define NSPARSE_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { NSPARSE_SYNTH_NNZ_LOGNORMAL = 0, NSPARSE_SYNTH_NNZ_LOGUNIFORM = 1 };

typedef struct nsparse_synth_spec {
    uint64_t seed;
    int dim;          /* vocabulary size; term ids are Zipf ranks minus one */
    double zipf_s;    /* Zipf exponent */
    int nnz_mode;     /* LOGNORMAL: round(exp(ln(a) + b*N(0,1))); LOGUNIFORM: round(exp(U(ln a, ln b))) */
    double nnz_a, nnz_b;
    int nnz_min, nnz_max; /* clip range for row lengths (also capped at dim) */
    double w_mu, w_sigma; /* weights: float(exp(w_mu + w_sigma*N(0,1))) */
} nsparse_synth_spec;

int nsparse_synth_row_nnz(const nsparse_synth_spec* spec, int64_t first_row, int64_t n_rows,
                          int32_t* row_nnz, int n_threads);
int nsparse_synth_fill(const nsparse_synth_spec* spec, int64_t first_row, int64_t n_rows,
                       const int64_t* indptr, uint16_t* indices, float* values, int n_threads);

#ifdef __cplusplus
}
#endif

#endif /* NSPARSE_SYNTH_H */
