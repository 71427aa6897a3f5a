// nsparse/index_factory.h:16: index_factory(dimension, description), the
// reference's description grammar unchanged. The returned index is owned by
// the caller and is of the class its description names (SeismicIndex,
// SeismicScalarQuantizedIndex, BrutalIndex, InvertedIndex, IDMapIndex).
#ifndef NSPARSE_DROPIN_INDEX_FACTORY_H
#define NSPARSE_DROPIN_INDEX_FACTORY_H

#include <cstring>
#include <stdexcept>
#include <string>

#include "nsparse/brutal_index.h"
#include "nsparse/id_map_index.h"
#include "nsparse/index.h"
#include "nsparse/inverted_index.h"
#include "nsparse/seismic_index.h"
#include "nsparse/seismic_scalar_quantized_index.h"

namespace nsparse {

namespace detail {

// The class of a handle the library made, by its fourcc; takes the handle
// over (and frees it if no class matches).
inline Index* adopt_index(nsparse_index* h) {
    char fourcc[4];
    const int status = nsparse_index_id(h, fourcc);
    if (status != NSPARSE_OK) {
        nsparse_free(h);
        check(status);
    }
    const std::string id(fourcc, 4);
    if (id == "SEIS") return new SeismicIndex(AdoptHandle{}, h);
    if (id == "SESQ") return new SeismicScalarQuantizedIndex(AdoptHandle{}, h);
    if (id == "BRUT") return new BrutalIndex(AdoptHandle{}, h);
    if (id == "INVT") return new InvertedIndex(AdoptHandle{}, h);
    if (id == "IDMP") return new IDMapIndex(AdoptHandle{}, h);
    nsparse_free(h);
    throw std::runtime_error("unknown index type " + id);
}

}  // namespace detail

inline Index* index_factory(int dimension, const char* description) {
    nsparse_index* h = nullptr;
    detail::check(nsparse_index_factory(dimension, description, &h));
    return detail::adopt_index(h);
}

}  // namespace nsparse

#endif  // NSPARSE_DROPIN_INDEX_FACTORY_H
