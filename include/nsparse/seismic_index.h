// nsparse/seismic_index.h:24-60: SeismicSearchParameters and SeismicIndex
// (FP32 values), built and searched by the device library.
#ifndef NSPARSE_DROPIN_SEISMIC_INDEX_H
#define NSPARSE_DROPIN_SEISMIC_INDEX_H

#include <array>

#include "nsparse/index.h"
#include "nsparse/seismic_common.h"
#include "nsparse/types.h"

namespace nsparse {

struct SeismicSearchParameters : public SearchParameters {
    int cut = 10;
    float heap_factor = 1.0F;
    SeismicSearchParameters(int cut, float heap_factor) : cut(cut), heap_factor(heap_factor) {}
    SeismicSearchParameters() = default;
    void to_c(nsparse_search_params* p) const override {
        p->kind = NSPARSE_PARAMS_SEISMIC;
        p->cut = cut;
        p->heap_factor = heap_factor;
    }
};

class SeismicIndex : public Index {
public:
    static constexpr std::array<char, 4> name = {'S', 'E', 'I', 'S'};

    explicit SeismicIndex(int dim) : SeismicIndex(dim, detail::kDefaultSeismicClusterParams) {}
    SeismicIndex(int dim, SeismicClusterParameters parameter)
        : Index(dim, "seismic," + detail::cluster_keys(parameter)) {}
    SeismicIndex(detail::AdoptHandle tag, nsparse_index* handle) : Index(tag, handle) {}
    ~SeismicIndex() override = default;
    std::array<char, 4> id() const override { return name; }
};

}  // namespace nsparse

#endif  // NSPARSE_DROPIN_SEISMIC_INDEX_H
