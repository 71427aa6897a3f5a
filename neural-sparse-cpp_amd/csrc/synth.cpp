// Deterministic synthetic sparse corpora for tests and benchmarks.
//
// Every row draws from its own splitmix64 stream keyed by (seed, row), so a
// corpus can be generated in parallel and in pieces and still come out
// byte-identical on any x86-64 host running this library. Term ids follow a
// Zipf(s) law over ranks 1..dim (rank r -> term id r-1), drawn without repeats
// inside a row and emitted in ascending term order (the reference's tests
// build rows from std::map, so rows are sorted by term there too). Weights are
// FP32 lognormal(mu, sigma). Row lengths are either lognormal (rounded, then
// clipped) or log-uniform over [a, b].
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <thread>
#include <vector>

#include "capi_util.h"
#include "nsparse_synth.h"

namespace nsparse_amd {
namespace {

inline uint64_t splitmix64(uint64_t& state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct RowRng {
    uint64_t s;
    RowRng(uint64_t seed, uint64_t stream, uint64_t row) {
        uint64_t k = seed * 0xD1B54A32D192ED03ULL ^ (stream + 0x632BE59BD9B4E019ULL);
        uint64_t a = splitmix64(k);
        uint64_t r = row ^ a;
        s = splitmix64(r) ^ a;
    }
    // Uniform double in [0, 1) with 53 random bits.
    double uniform() { return static_cast<double>(splitmix64(s) >> 11) * 0x1.0p-53; }
    // Uniform double in (0, 1].
    double uniform_pos() { return static_cast<double>((splitmix64(s) >> 11) + 1) * 0x1.0p-53; }
    double normal() {
        const double u1 = uniform_pos();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct ZipfTable {
    std::vector<double> cdf;     // cdf[i] = P(term <= i)
    std::vector<int32_t> guide;  // guide[j] = first i with cdf[i] > j / G
    ZipfTable(int dim, double s) : cdf(dim), guide(dim + 1) {
        double h = 0.0;
        for (int i = 0; i < dim; ++i) h += std::pow(static_cast<double>(i + 1), -s);
        double acc = 0.0;
        for (int i = 0; i < dim; ++i) {
            acc += std::pow(static_cast<double>(i + 1), -s) / h;
            cdf[i] = acc;
        }
        cdf[dim - 1] = 1.0;
        const int g = dim;
        int i = 0;
        for (int j = 0; j <= g; ++j) {
            const double u = static_cast<double>(j) / g;
            while (i < dim - 1 && cdf[i] <= u) ++i;
            guide[j] = i;
        }
    }
    int sample(double u) const {
        const int g = static_cast<int>(guide.size()) - 1;
        int i = guide[static_cast<int>(u * g)];
        while (cdf[i] <= u) ++i;
        return i;
    }
};

int row_length(RowRng& rng, const nsparse_synth_spec& spec) {
    double x;
    if (spec.nnz_mode == NSPARSE_SYNTH_NNZ_LOGNORMAL) {
        x = std::exp(std::log(spec.nnz_a) + spec.nnz_b * rng.normal());
    } else {
        const double la = std::log(spec.nnz_a), lb = std::log(spec.nnz_b);
        x = std::exp(la + (lb - la) * rng.uniform());
    }
    long n = std::lround(x);
    n = std::max<long>(n, spec.nnz_min);
    n = std::min<long>(n, spec.nnz_max);
    n = std::min<long>(n, spec.dim);
    return static_cast<int>(n);
}

void check_spec(const nsparse_synth_spec* spec) {
    if (spec == nullptr) throw std::invalid_argument("spec must not be null");
    if (spec->dim <= 0 || spec->dim > 65536) throw std::invalid_argument("dim must be in [1, 65536]");
    if (spec->nnz_min < 0 || spec->nnz_max < spec->nnz_min)
        throw std::invalid_argument("invalid nnz clip range");
    if (spec->nnz_mode != NSPARSE_SYNTH_NNZ_LOGNORMAL && spec->nnz_mode != NSPARSE_SYNTH_NNZ_LOGUNIFORM)
        throw std::invalid_argument("unknown nnz mode");
    if (!(spec->nnz_a > 0.0)) throw std::invalid_argument("nnz_a must be positive");
}

template <class F>
void parallel_rows(int64_t n_rows, int n_threads, F&& f) {
    if (n_threads <= 0) n_threads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    n_threads = static_cast<int>(std::min<int64_t>(n_threads, std::max<int64_t>(1, n_rows / 256)));
    if (n_threads <= 1) {
        f(0, n_rows);
        return;
    }
    std::vector<std::thread> pool;
    std::vector<std::exception_ptr> errors(n_threads);
    const int64_t chunk = (n_rows + n_threads - 1) / n_threads;
    for (int t = 0; t < n_threads; ++t) {
        const int64_t lo = t * chunk, hi = std::min<int64_t>(n_rows, lo + chunk);
        if (lo >= hi) break;
        pool.emplace_back([&f, &errors, t, lo, hi] {
            try {
                f(lo, hi);
            } catch (...) {
                errors[t] = std::current_exception();
            }
        });
    }
    for (auto& th : pool) th.join();
    for (auto& e : errors)
        if (e) std::rethrow_exception(e);
}

}  // namespace
}  // namespace nsparse_amd

using namespace nsparse_amd;

extern "C" int nsparse_synth_row_nnz(const nsparse_synth_spec* spec, int64_t first_row, int64_t n_rows,
                                     int32_t* row_nnz, int n_threads) {
    return guarded([&] {
        check_spec(spec);
        if (n_rows < 0 || (n_rows > 0 && row_nnz == nullptr)) throw std::invalid_argument("bad row range");
        parallel_rows(n_rows, n_threads, [&](int64_t lo, int64_t hi) {
            for (int64_t r = lo; r < hi; ++r) {
                RowRng rng(spec->seed, 0, static_cast<uint64_t>(first_row + r));
                row_nnz[r] = row_length(rng, *spec);
            }
        });
    });
}

extern "C" int nsparse_synth_fill(const nsparse_synth_spec* spec, int64_t first_row, int64_t n_rows,
                                  const int64_t* indptr, uint16_t* indices, float* values, int n_threads) {
    return guarded([&] {
        check_spec(spec);
        if (n_rows < 0) throw std::invalid_argument("bad row range");
        if (n_rows == 0) return;
        if (indptr == nullptr || indices == nullptr || values == nullptr)
            throw std::invalid_argument("output pointers must not be null");
        const ZipfTable zipf(spec->dim, spec->zipf_s);
        parallel_rows(n_rows, n_threads, [&](int64_t lo, int64_t hi) {
            std::vector<uint32_t> stamp(spec->dim, 0);
            uint32_t epoch = 0;
            std::vector<int> terms;
            for (int64_t r = lo; r < hi; ++r) {
                RowRng len_rng(spec->seed, 0, static_cast<uint64_t>(first_row + r));
                const int n = row_length(len_rng, *spec);
                if (indptr[r + 1] - indptr[r] != n)
                    throw std::invalid_argument("indptr does not match the generated row lengths");
                RowRng rng(spec->seed, 1, static_cast<uint64_t>(first_row + r));
                ++epoch;
                terms.clear();
                while (static_cast<int>(terms.size()) < n) {
                    const int t = zipf.sample(rng.uniform());
                    if (stamp[t] == epoch) continue;
                    stamp[t] = epoch;
                    terms.push_back(t);
                }
                std::sort(terms.begin(), terms.end());
                uint16_t* out_i = indices + indptr[r];
                float* out_v = values + indptr[r];
                for (int j = 0; j < n; ++j) {
                    out_i[j] = static_cast<uint16_t>(terms[j]);
                    out_v[j] = static_cast<float>(std::exp(spec->w_mu + spec->w_sigma * rng.normal()));
                }
            }
        });
    });
}
