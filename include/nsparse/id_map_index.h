// nsparse/id_map_index.h:78-109: IDMapIndex, external ids over a delegate
// index. Like the reference's, it takes ownership of the delegate and deletes
// it when destroyed; the delegate object stays alive until then but holds no
// index of its own (its calls throw std::invalid_argument).
#ifndef NSPARSE_DROPIN_ID_MAP_INDEX_H
#define NSPARSE_DROPIN_ID_MAP_INDEX_H

#include <array>
#include <memory>

#include "nsparse/index.h"

namespace nsparse {

class IDMapIndex : public Index {
public:
    static constexpr std::array<char, 4> name = {'I', 'D', 'M', 'P'};
    // No delegate: every call throws std::invalid_argument.
    IDMapIndex() = default;
    explicit IDMapIndex(Index* delegate) : delegate_(delegate) {
        if (delegate == nullptr) throw std::invalid_argument("IDMapIndex needs a delegate index");
        dimension_ = delegate->get_dimension();
        detail::check(nsparse_idmap_wrap(delegate->handle(), &handle_));
    }
    IDMapIndex(detail::AdoptHandle tag, nsparse_index* handle) : Index(tag, handle) {}
    ~IDMapIndex() override = default;
    std::array<char, 4> id() const override { return name; }

private:
    std::unique_ptr<Index> delegate_;
};

}  // namespace nsparse

#endif  // NSPARSE_DROPIN_ID_MAP_INDEX_H
