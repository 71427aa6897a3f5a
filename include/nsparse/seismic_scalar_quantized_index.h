// nsparse/seismic_scalar_quantized_index.h:27-50: SeismicSQSearchParameters
// and SeismicScalarQuantizedIndex (uint8 codes), built and searched by the
// device library.
#ifndef NSPARSE_DROPIN_SEISMIC_SQ_INDEX_H
#define NSPARSE_DROPIN_SEISMIC_SQ_INDEX_H

#include <array>
#include <cstdio>
#include <string>

#include "nsparse/index.h"
#include "nsparse/seismic_common.h"
#include "nsparse/seismic_index.h"
#include "nsparse/types.h"
#include "nsparse/utils/scalar_quantizer.h"

namespace nsparse {

struct SeismicSQSearchParameters : public SeismicSearchParameters {
    float vmin;
    float vmax;
    SeismicSQSearchParameters(float vmin, float vmax, int cut, float heap_factor)
        : SeismicSearchParameters(cut, heap_factor), vmin(vmin), vmax(vmax) {}
    void to_c(nsparse_search_params* p) const override {
        SeismicSearchParameters::to_c(p);
        p->kind = NSPARSE_PARAMS_SEISMIC_SQ;
        p->vmin = vmin;
        p->vmax = vmax;
    }
};

class SeismicScalarQuantizedIndex : public Index {
public:
    static constexpr std::array<char, 4> name = {'S', 'E', 'S', 'Q'};

    // An 8-bit quantizer over [0, 1] and the default build parameters, as the
    // reference's one-argument constructor.
    explicit SeismicScalarQuantizedIndex(int dim)
        : SeismicScalarQuantizedIndex(QuantizerType::QT_8bit, 0.0F, 1.0F, detail::kDefaultSeismicClusterParams,
                                      dim) {}
    SeismicScalarQuantizedIndex(QuantizerType quantizer_type, float vmin, float vmax,
                                SeismicClusterParameters parameter, int dim)
        : Index(dim, description(quantizer_type, vmin, vmax, parameter)) {}
    SeismicScalarQuantizedIndex(detail::AdoptHandle tag, nsparse_index* handle) : Index(tag, handle) {}
    ~SeismicScalarQuantizedIndex() override = default;
    std::array<char, 4> id() const override { return name; }

private:
    static std::string description(QuantizerType qt, float vmin, float vmax, const SeismicClusterParameters& p) {
        char buf[96];
        std::snprintf(buf, sizeof(buf), "seismic_sq,quantizer=%s|vmin=%.9g|vmax=%.9g|",
                      qt == QuantizerType::QT_8bit ? "8bit" : "16bit", static_cast<double>(vmin),
                      static_cast<double>(vmax));
        return buf + detail::cluster_keys(p);
    }
};

}  // namespace nsparse

#endif  // NSPARSE_DROPIN_SEISMIC_SQ_INDEX_H
