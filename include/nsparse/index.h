// nsparse/index.h:21-70: SearchParameters and the Index base class, as an
// adapter over the C ABI. An Index built by this header owns one
// nsparse_index handle; its calls forward to the device library and re-throw
// the library's status codes as the exception types the reference throws
// (std::invalid_argument, std::runtime_error, std::out_of_range,
// std::bad_alloc).
#ifndef NSPARSE_DROPIN_INDEX_H
#define NSPARSE_DROPIN_INDEX_H

#include <array>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <exception>
#include <new>
#include <stdexcept>
#include <string>
#include <typeinfo>
#include <vector>

#include "nsparse/id_selector.h"
#include "nsparse/types.h"
#include "nsparse_c.h"

namespace nsparse {

namespace detail {

// Re-throws a C ABI status as the reference's exception type.
inline void check(int status) {
    if (status == NSPARSE_OK) return;
    const std::string msg = nsparse_last_error();
    switch (status) {
        case NSPARSE_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
        case NSPARSE_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
        case NSPARSE_ERR_BAD_ALLOC: throw std::bad_alloc();
        default: throw std::runtime_error(msg);
    }
}

// Tag of the constructors that take over a handle the library made.
struct AdoptHandle {};

// An IDSelector tree as nsparse_selector nodes. Set and Array selectors go
// over as their ids, Not as a node over its delegate, an enumerable selector
// of another type as its ordered ids when size() <= k (the reference's
// exact-match gate, seismic_common.h:117-128) and any other selector as a
// callback into its is_member. The nodes live as long as this object.
class SelectorMarshal {
public:
    SelectorMarshal(const IDSelector* selector, int k) : k_(k) {
        if (selector != nullptr) root_ = convert(selector);
    }
    SelectorMarshal(const SelectorMarshal&) = delete;
    SelectorMarshal& operator=(const SelectorMarshal&) = delete;

    const nsparse_selector* get() const { return root_; }
    // An exception an is_member callback threw during the call, re-thrown.
    void rethrow_callback_error() const {
        if (error_) std::rethrow_exception(error_);
    }

private:
    struct Callback {
        const IDSelector* selector;
        SelectorMarshal* owner;
    };

    static int call_is_member(const void* ctx, nsparse_idx_t id) {
        const auto* cb = static_cast<const Callback*>(ctx);
        try {
            return cb->selector->is_member(id) ? 1 : 0;
        } catch (...) {
            if (!cb->owner->error_) cb->owner->error_ = std::current_exception();
            return -1;
        }
    }

    nsparse_selector* ids_node(int kind, std::vector<idx_t> ids) {
        ids_.push_back(std::move(ids));
        nsparse_selector& s = nodes_.emplace_back();
        s.kind = kind;
        s.n = static_cast<int64_t>(ids_.back().size());
        s.ids = ids_.back().data();
        return &s;
    }

    nsparse_selector* convert(const IDSelector* sel) {
        const std::type_info& t = typeid(*sel);
        if (t == typeid(SetIDSelector))
            return ids_node(NSPARSE_SELECTOR_SET, static_cast<const SetIDSelector*>(sel)->ids());
        if (t == typeid(ArrayIDSelector))
            return ids_node(NSPARSE_SELECTOR_ARRAY, static_cast<const ArrayIDSelector*>(sel)->ids());
        if (t == typeid(NotIDSelector)) {
            const IDSelector* inner = static_cast<const NotIDSelector*>(sel)->delegate();
            if (inner == nullptr) throw std::invalid_argument("NotIDSelector needs an inner selector");
            nsparse_selector* in = convert_inner(inner);
            nsparse_selector& s = nodes_.emplace_back();
            s.kind = NSPARSE_SELECTOR_NOT;
            s.inner = in;
            return &s;
        }
        if (const auto* en = dynamic_cast<const IDSelectorEnumerable*>(sel))
            if (en->size() <= static_cast<size_t>(k_)) return ids_node(NSPARSE_SELECTOR_ARRAY, en->ordered_ids());
        return callback_node(sel);
    }

    // Below a Not only membership matters: no exact-match gate applies.
    nsparse_selector* convert_inner(const IDSelector* sel) {
        const std::type_info& t = typeid(*sel);
        if (t == typeid(SetIDSelector) || t == typeid(ArrayIDSelector) || t == typeid(NotIDSelector))
            return convert(sel);
        return callback_node(sel);
    }

    nsparse_selector* callback_node(const IDSelector* sel) {
        callbacks_.push_back(Callback{sel, this});
        nsparse_selector& s = nodes_.emplace_back();
        s.kind = NSPARSE_SELECTOR_CALLBACK;
        s.is_member = &SelectorMarshal::call_is_member;
        s.ctx = &callbacks_.back();
        return &s;
    }

    int k_;
    nsparse_selector* root_ = nullptr;
    std::deque<nsparse_selector> nodes_;
    std::deque<std::vector<idx_t>> ids_;
    std::deque<Callback> callbacks_;
    std::exception_ptr error_;
};

}  // namespace detail

struct SearchParameters {
    virtual ~SearchParameters() = default;
    const IDSelector* get_id_selector() const { return id_selector; }
    void set_id_selector(IDSelector* selector) { id_selector = selector; }
    // The C ABI's parameter block for these parameters (selector aside).
    virtual void to_c(nsparse_search_params* p) const {
        p->kind = NSPARSE_PARAMS_BASE;
        p->cut = 10;
        p->heap_factor = 1.0F;
    }

private:
    IDSelector* id_selector = nullptr;
};

class Index {
public:
    explicit Index(int dim = 0) : dimension_(dim) {}
    virtual ~Index() {
        if (handle_ != nullptr) nsparse_free(handle_);
    }
    Index(const Index&) = delete;
    Index& operator=(const Index&) = delete;

    virtual std::array<char, 4> id() const {
        std::array<char, 4> out{};
        detail::check(nsparse_index_id(handle_, out.data()));
        return out;
    }
    virtual void add(idx_t n, const idx_t* indptr, const term_t* indices, const float* values) {
        detail::check(nsparse_add(handle_, n, indptr, indices, values));
    }
    virtual void build() { detail::check(nsparse_build(handle_)); }
    // Pre-allocated outputs: n * k, padded with (-1.0f, -1) (index.cpp:32-49).
    virtual void search(idx_t n, const idx_t* indptr, const term_t* indices, const float* values, int k,
                        float* distances, idx_t* labels, SearchParameters* search_parameters = nullptr) {
        nsparse_search_params p{};
        SearchParameters base;
        (search_parameters != nullptr ? search_parameters : &base)->to_c(&p);
        const detail::SelectorMarshal sel(search_parameters != nullptr ? search_parameters->get_id_selector() : nullptr,
                                          k);
        p.selector = sel.get();
        const int status = nsparse_search(handle_, n, indptr, indices, values, k, distances, labels, &p);
        sel.rethrow_callback_error();
        detail::check(status);
    }

    int get_dimension() const { return dimension_; }
    virtual size_t num_vectors() const {
        int64_t n = 0;
        detail::check(nsparse_num_vectors(handle_, &n));
        return static_cast<size_t>(n);
    }
    virtual void add_with_ids(idx_t n, const idx_t* indptr, const term_t* indices, const float* values,
                              const idx_t* ids) {
        detail::check(nsparse_add_with_ids(handle_, n, indptr, indices, values, ids));
    }
    // Loads rows from a CSR interchange file (three int64 sizes, int64 row
    // offsets, int32 term ids, float32 weights).
    virtual void read_csr(const char* file_path, Residency residency = Residency::kInMemory) {
        detail::check(nsparse_read_csr(handle_, file_path, static_cast<int>(residency)));
    }

    // The C ABI handle this index forwards to (for nsparse_search_device and
    // the other batched entry points).
    nsparse_index* handle() const { return handle_; }

protected:
    Index(detail::AdoptHandle, nsparse_index* handle) : dimension_(0), handle_(handle) {
        if (handle_ != nullptr) detail::check(nsparse_get_dimension(handle_, &dimension_));
    }
    // Makes an index from an index_factory description.
    Index(int dim, const std::string& description) : dimension_(dim) {
        detail::check(nsparse_index_factory(dim, description.c_str(), &handle_));
    }
    int dimension_;
    nsparse_index* handle_ = nullptr;
};

}  // namespace nsparse

#endif  // NSPARSE_DROPIN_INDEX_H
