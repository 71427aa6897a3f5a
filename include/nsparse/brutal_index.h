// nsparse/brutal_index.h:24-40: BrutalIndex, exact scoring of every stored
// row (no build step).
#ifndef NSPARSE_DROPIN_BRUTAL_INDEX_H
#define NSPARSE_DROPIN_BRUTAL_INDEX_H

#include <array>

#include "nsparse/index.h"

namespace nsparse {

class BrutalIndex : public Index {
public:
    static constexpr std::array<char, 4> name = {'B', 'R', 'U', 'T'};
    explicit BrutalIndex(int dim = 0) : Index(dim, "brutal") {}
    BrutalIndex(detail::AdoptHandle tag, nsparse_index* handle) : Index(tag, handle) {}
    ~BrutalIndex() override = default;
    std::array<char, 4> id() const override { return name; }
};

}  // namespace nsparse

#endif  // NSPARSE_DROPIN_BRUTAL_INDEX_H
