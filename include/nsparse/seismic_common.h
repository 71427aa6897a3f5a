// nsparse/seismic_common.h:36-60: the SEISMIC build parameters and their
// defaults (lambda and beta -1: derived from the corpus at build time).
#ifndef NSPARSE_DROPIN_SEISMIC_COMMON_H
#define NSPARSE_DROPIN_SEISMIC_COMMON_H

#include <cstdio>
#include <string>

namespace nsparse {

constexpr int kRandomSeed = -1;

// inverted_list_batch_size and batch_file_output_path steer the reference's
// out-of-core CPU build, which writes lists to files in batches; the device
// build holds every list in HBM at once and does not read them.
struct SeismicClusterParameters {
    int lambda;
    int beta;
    float alpha;
    int inverted_list_batch_size = 1;
    char* batch_file_output_path = nullptr;
    // kRandomSeed draws a fresh seed per build; any other value makes the
    // build's centroid draws, and so its layout, repeatable.
    int seed = kRandomSeed;
};

namespace detail {

constexpr int kDefaultLambda = -1;
constexpr int kDefaultBeta = -1;
constexpr float kDefaultAlpha = 0.4F;
constexpr SeismicClusterParameters kDefaultSeismicClusterParams = {
    kDefaultLambda, kDefaultBeta, kDefaultAlpha, 1, nullptr, kRandomSeed};

// The index_factory keys of a parameter set ("lambda=..|beta=..|alpha=..|seed=..");
// alpha in 9 significant digits, which a float survives unchanged.
inline std::string cluster_keys(const SeismicClusterParameters& p) {
    char buf[160];
    std::snprintf(buf, sizeof(buf), "lambda=%d|beta=%d|alpha=%.9g|seed=%d", p.lambda, p.beta,
                  static_cast<double>(p.alpha), p.seed);
    return buf;
}

}  // namespace detail

}  // namespace nsparse

#endif  // NSPARSE_DROPIN_SEISMIC_COMMON_H
