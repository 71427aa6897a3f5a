// nsparse/inverted_index.h:28-60: InvertedIndex, exact MaxScore search over
// plain posting lists, built on the device.
#ifndef NSPARSE_DROPIN_INVERTED_INDEX_H
#define NSPARSE_DROPIN_INVERTED_INDEX_H

#include <array>

#include "nsparse/index.h"

namespace nsparse {

class InvertedIndex : public Index {
public:
    static constexpr std::array<char, 4> name = {'I', 'N', 'V', 'T'};
    explicit InvertedIndex(int dim) : Index(dim, "inverted") {}
    InvertedIndex(detail::AdoptHandle tag, nsparse_index* handle) : Index(tag, handle) {}
    ~InvertedIndex() override = default;
    std::array<char, 4> id() const override { return name; }
};

}  // namespace nsparse

#endif  // NSPARSE_DROPIN_INVERTED_INDEX_H
