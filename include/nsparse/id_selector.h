// nsparse/id_selector.h:22-86: the selector family. A search passes a
// selector to the device as its ids (Set, Array), as a negation of another
// (Not), or, for any other IDSelector subclass, as a membership bitmap the
// adapter evaluates over the index's doc ids.
#ifndef NSPARSE_DROPIN_ID_SELECTOR_H
#define NSPARSE_DROPIN_ID_SELECTOR_H

#include <algorithm>
#include <cstddef>
#include <unordered_set>
#include <vector>

#include "nsparse/types.h"

namespace nsparse {

class IDSelector {
public:
    virtual ~IDSelector() = default;
    virtual bool is_member(idx_t id) const = 0;
    bool operator()(idx_t id) const { return is_member(id); }
};

class IDSelectorEnumerable : public IDSelector {
public:
    ~IDSelectorEnumerable() override = default;
    virtual std::vector<idx_t> ids() const = 0;
    virtual std::vector<idx_t> ordered_ids() const {
        auto vec = ids();
        std::sort(vec.begin(), vec.end());
        return vec;
    }
    virtual size_t size() const = 0;
};

class SetIDSelector : public IDSelectorEnumerable {
public:
    explicit SetIDSelector(size_t n, const idx_t* indices) : ids_(indices, indices + n) {}
    bool is_member(idx_t id) const override { return ids_.count(id) != 0; }
    std::vector<idx_t> ids() const override { return std::vector<idx_t>(ids_.begin(), ids_.end()); }
    size_t size() const override { return ids_.size(); }

private:
    std::unordered_set<idx_t> ids_;
};

class ArrayIDSelector : public IDSelectorEnumerable {
public:
    explicit ArrayIDSelector(size_t n, const idx_t* indices) : ids_(indices, indices + n) {}
    bool is_member(idx_t id) const override { return std::find(ids_.begin(), ids_.end(), id) != ids_.end(); }
    std::vector<idx_t> ids() const override { return ids_; }
    size_t size() const override { return ids_.size(); }

private:
    std::vector<idx_t> ids_;
};

// Non-owning, as the reference's.
class NotIDSelector : public IDSelector {
public:
    explicit NotIDSelector(IDSelector* selector) : delegate_(selector) {}
    bool is_member(idx_t id) const override { return !delegate_->is_member(id); }
    const IDSelector* delegate() const { return delegate_; }

private:
    IDSelector* delegate_;
};

}  // namespace nsparse

#endif  // NSPARSE_DROPIN_ID_SELECTOR_H
