// Host implementation of the index family, the factory and the reference's
// serialized layout. All compute runs on the GPU (search.hip, build.hip); the
// host validates arguments, keeps the reference's error behaviour and moves
// bytes between the caller and HBM.
#include "index.h"

#include <algorithm>
#include <functional>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_set>

#include "io.h"
#include "kernels.h"

namespace nsparse_amd {

namespace {

void throw_if_not_positive(int64_t v) {
    if (v <= 0) throw std::invalid_argument("value must be positive");
}
template <class... P>
void throw_if_any_null(P*... ptrs) {
    if (((ptrs == nullptr) || ...)) throw std::invalid_argument("unexpected nullptr");
}

uint32_t fourcc_of(const std::array<char, 4>& id) {
    return static_cast<uint32_t>(static_cast<uint8_t>(id[0])) |
           static_cast<uint32_t>(static_cast<uint8_t>(id[1])) << 8 |
           static_cast<uint32_t>(static_cast<uint8_t>(id[2])) << 16 |
           static_cast<uint32_t>(static_cast<uint8_t>(id[3])) << 24;
}

// Validates a caller's CSR batch: indptr non-decreasing from indptr[0], term
// ids below dim. The reference trusts the caller here; a bad term id there is
// an out-of-bounds write, here it is an invalid_argument.
void check_csr(idx_t n, const idx_t* indptr, const term_t* indices, int dim, const char* what) {
    for (idx_t i = 0; i < n; ++i)
        if (indptr[i + 1] < indptr[i]) throw std::invalid_argument(std::string(what) + " indptr is not monotonic");
    if (indptr[0] < 0) throw std::invalid_argument(std::string(what) + " indptr must not be negative");
    for (int64_t j = indptr[0]; j < indptr[n]; ++j)
        if (indices[j] >= dim) throw std::invalid_argument(std::string(what) + " term id out of range");
}

// Evaluates an nsparse_selector the way the reference's IDSelector family does
// (id_selector.h:22-86).
class SelectorEval {
public:
    explicit SelectorEval(const nsparse_selector* s) : s_(s) {
        if (s == nullptr) return;
        if (s->kind == NSPARSE_SELECTOR_ARRAY || s->kind == NSPARSE_SELECTOR_SET) {
            if (s->n < 0 || (s->n > 0 && s->ids == nullptr)) throw std::invalid_argument("selector ids are null");
            set_.insert(s->ids, s->ids + s->n);
        } else if (s->kind == NSPARSE_SELECTOR_NOT) {
            if (s->inner == nullptr) throw std::invalid_argument("NotIDSelector needs an inner selector");
            inner_ = std::make_unique<SelectorEval>(s->inner);
        } else if (s->kind == NSPARSE_SELECTOR_BITMAP) {
            if (s->n < 0 || (s->n > 0 && s->bitmap == nullptr)) throw std::invalid_argument("selector bitmap is null");
        } else if (s->kind == NSPARSE_SELECTOR_CALLBACK) {
            if (s->is_member == nullptr) throw std::invalid_argument("selector callback is null");
        } else {
            throw std::invalid_argument("unknown selector kind");
        }
    }
    bool active() const { return s_ != nullptr; }
    bool is_member(int64_t id) const {
        switch (s_->kind) {
            case NSPARSE_SELECTOR_ARRAY:
            case NSPARSE_SELECTOR_SET: return set_.count(static_cast<idx_t>(id)) != 0;
            case NSPARSE_SELECTOR_NOT: return !inner_->is_member(id);
            case NSPARSE_SELECTOR_CALLBACK: {
                const int r = s_->is_member(s_->ctx, static_cast<idx_t>(id));
                if (r < 0) throw std::runtime_error("IDSelector::is_member failed");
                return r != 0;
            }
            default: return id >= 0 && id < s_->n && ((s_->bitmap[id >> 5] >> (id & 31)) & 1U);
        }
    }
    bool enumerable() const {
        return s_ != nullptr && (s_->kind == NSPARSE_SELECTOR_ARRAY || s_->kind == NSPARSE_SELECTOR_SET);
    }
    // IDSelectorEnumerable::size(): ArrayIDSelector counts duplicates, SetIDSelector does not.
    size_t size() const { return s_->kind == NSPARSE_SELECTOR_ARRAY ? static_cast<size_t>(s_->n) : set_.size(); }
    // ordered_ids(): ids() sorted; ArrayIDSelector keeps duplicates.
    std::vector<idx_t> ordered_ids() const {
        std::vector<idx_t> v;
        if (s_->kind == NSPARSE_SELECTOR_ARRAY)
            v.assign(s_->ids, s_->ids + s_->n);
        else
            v.assign(set_.begin(), set_.end());
        std::sort(v.begin(), v.end());
        return v;
    }

private:
    const nsparse_selector* s_;
    std::unordered_set<idx_t> set_;
    std::unique_ptr<SelectorEval> inner_;
};

SearchRequest request_from(const nsparse_search_params* p) {
    SearchRequest r;
    if (p == nullptr) return r;
    if (p->kind == NSPARSE_PARAMS_SEISMIC || p->kind == NSPARSE_PARAMS_SEISMIC_SQ) {
        r.cut = p->cut;
        r.heap_factor = p->heap_factor;
    } else if (p->kind != NSPARSE_PARAMS_BASE) {
        throw std::invalid_argument("unknown search parameters kind");
    }
    if (p->kind == NSPARSE_PARAMS_SEISMIC_SQ) {
        r.has_sq_range = true;
        r.vmin = p->vmin;
        r.vmax = p->vmax;
    }
    r.selector = p->selector;
    return r;
}

void check_quantizer_range(float vmin, float vmax) {
    if (!std::isfinite(vmin) || !std::isfinite(vmax)) throw std::invalid_argument("vmin and vmax must be finite");
    if (vmax <= vmin) throw std::invalid_argument("vmax must be greater than vmin");
}

}  // namespace

void require_device() {
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        throw device_error("no HIP device available: this index runs its search and build on the GPU");
}

void* ScratchPool::acquire(size_t bytes, size_t& slot) {
    std::lock_guard<std::mutex> lock(mu_);
    for (size_t i = 0; i < slots_.size(); ++i) {
        Slot& s = slots_[i];
        if (s.busy || s.buf.bytes() < bytes) continue;
        const hipError_t e = hipEventQuery(s.done);
        if (e == hipErrorNotReady) continue;
        hip_check(e, "hipEventQuery");
        s.busy = true;
        slot = i;
        return s.buf.get();
    }
    Slot s;
    s.buf.allocate(bytes);
    hip_check(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "hipEventCreate");
    s.busy = true;
    slots_.push_back(std::move(s));
    slot = slots_.size() - 1;
    return slots_.back().buf.get();
}

void ScratchPool::release(size_t slot, hipStream_t stream) {
    std::lock_guard<std::mutex> lock(mu_);
    Slot& s = slots_.at(slot);
    const hipError_t e = hipEventRecord(s.done, stream);
    s.busy = false;
    hip_check(e, "hipEventRecord");
}

void Index::add_with_ids(idx_t, const idx_t*, const term_t*, const float*, const idx_t*) {
    throw std::runtime_error("add_with_ids not implemented in Index");
}
void Index::build() { throw std::runtime_error("not implemented"); }

// ============================================================== SeismicIndex

SeismicIndex::SeismicIndex(int dim, bool sq, float vmin, float vmax, Params p, bool w16)
    : Index(dim), sq_(sq), w16_(sq && w16), vmin_(vmin), vmax_(vmax), params_(p) {
    if (sq_) check_quantizer_range(vmin, vmax);
    fwd_.esz = sq_ ? (w16_ ? 2 : 1) : 4;
    fwd_.dim = static_cast<size_t>(dim);
}

std::array<char, 4> SeismicIndex::id() const {
    return sq_ ? std::array<char, 4>{'S', 'E', 'S', 'Q'} : std::array<char, 4>{'S', 'E', 'I', 'S'};
}

// SeismicIndex::add (seismic_index.cpp:127-142) / SeismicScalarQuantizedIndex::add
// (seismic_scalar_quantized_index.cpp:159-177) + SparseVectors::add_vectors
// (sparse_vectors.cpp:92-126). SQ codes are produced on the GPU.
void SeismicIndex::add(idx_t n, const idx_t* indptr, const term_t* indices, const float* values) {
    throw_if_not_positive(n);
    throw_if_any_null(indptr, indices, values);
    if (dim_ <= 0) throw std::invalid_argument("value must be positive");
    check_csr(n, indptr, indices, dim_, "add:");
    const int64_t nnz = indptr[n];
    if (static_cast<int64_t>(fwd_.indices.size()) + nnz > std::numeric_limits<idx_t>::max())
        throw std::invalid_argument("too many non-zeros for 32-bit offsets");
    if (fwd_.indptr.empty()) fwd_.indptr.push_back(0);
    const idx_t off = fwd_.indptr.back();
    fwd_.indices.insert(fwd_.indices.end(), indices, indices + nnz);
    const size_t old_vals = fwd_.values.size();
    fwd_.values.resize(old_vals + static_cast<size_t>(nnz) * fwd_.esz);
    if (sq_) {
        require_device();
        encode_sq_on_device(values, static_cast<size_t>(nnz), vmin_, vmax_, fwd_.values.data() + old_vals,
                            w16_ ? 16 : 8);
    } else {
        std::memcpy(fwd_.values.data() + old_vals, values, static_cast<size_t>(nnz) * 4);
    }
    for (idx_t i = 1; i <= n; ++i) fwd_.indptr.push_back(indptr[i] + off);
    has_vectors_ = fwd_.n() > 0;
    dev_fwd_valid_ = false;
}

void SeismicIndex::set_lists(std::vector<HostList>&& lists) {
    lists_ = std::move(lists);
    host_lists_valid_ = true;
    dev_lists_valid_ = false;
    n_lists_ = static_cast<int32_t>(lists_.size());
    max_clusters_ = 0;
    for (auto& l : lists_) max_clusters_ = std::max<int>(max_clusters_, static_cast<int>(l.n_clusters));
}

const std::vector<HostList>& SeismicIndex::lists() {
    if (!host_lists_valid_) download_lists();
    return lists_;
}

// Index::build for SEISMIC (seismic_common.h:146-184), entirely on the GPU.
void SeismicIndex::build() {
    if (!has_vectors_) throw std::invalid_argument("build needs vectors: call add first");
    require_device();
    ensure_device();
    DeviceLists out;
    build_on_device(view_, fwd_.n(), static_cast<int64_t>(fwd_.indices.size()), params_.lambda, params_.beta,
                    params_.alpha, params_.seed, out);
    d_list_cluster_base_ = std::move(out.list_cluster_base);
    d_cluster_doc_ptr_ = std::move(out.cluster_doc_ptr);
    d_docs_ = std::move(out.docs);
    d_list_term_base_ = std::move(out.list_term_base);
    d_sum_term_ids_ = std::move(out.sum_term_ids);
    d_sum_term_ptr_ = std::move(out.sum_term_ptr);
    d_sum_cluster_ = std::move(out.sum_cluster);
    d_sum_value_ = std::move(out.sum_value);
    d_list_entry_base_ = std::move(out.list_entry_base);
    n_lists_ = out.n_lists;
    max_clusters_ = out.max_clusters;
    build_stats_ = out.stats;
    lists_.clear();
    host_lists_valid_ = false;
    dev_lists_valid_ = true;
    view_.n_lists = n_lists_;
    view_.max_clusters = max_clusters_;
    view_.list_cluster_base = d_list_cluster_base_.get<int32_t>();
    view_.cluster_doc_ptr = d_cluster_doc_ptr_.get<int64_t>();
    view_.docs = d_docs_.get<int32_t>();
    view_.list_term_base = d_list_term_base_.get<int32_t>();
    view_.sum_term_ids = d_sum_term_ids_.get<uint16_t>();
    view_.sum_term_ptr = d_sum_term_ptr_.get<int64_t>();
    view_.sum_cluster = d_sum_cluster_.get<uint16_t>();
    view_.sum_value = d_sum_value_.get<void>();
    refresh_block_rows();
    refresh_summary_bitmaps();
    max_list_docs_ = max_list_docs(view_);
}

// Block-contiguous row copies for the uint8 search (device_index.h). They cost
// about 2.4 B per non-zero of every posting on a 30522-term vocabulary with
// rows sorted by term (16-byte words of 7-8 slots); NSPARSE_BLOCK_ROWS=0 turns them
// off, and they are skipped when HBM cannot hold them. The search reads the
// forward index directly when they are absent.
void SeismicIndex::refresh_block_rows() {
    view_.row_ref = nullptr;
    view_.blk_rows = nullptr;
    view_.blk_words = 0;
    const char* e = std::getenv("NSPARSE_BLOCK_ROWS");
    const bool on = u8() && !(e != nullptr && std::string(e) == "0");
    const int64_t postings = static_cast<int64_t>(d_docs_.bytes() / 4);
    int64_t words = 0;
    if (!on || !build_block_rows(view_, postings, d_row_ref_, d_blk_rows_, words)) {
        d_row_ref_.release();
        d_blk_rows_.release();
        return;
    }
    view_.row_ref = d_row_ref_.get<int64_t>();
    view_.blk_rows = d_blk_rows_.get<uint4>();
    view_.blk_words = words;
}

// Per-list summary-term bitmaps (device_index.h): the search finds a query
// term's summary column with one bitmap word and one rank prefix instead of
// scanning the list's summary terms. About 6 B per 32 vocabulary terms per
// list (175 MB for 30522 lists of a 30522-term vocabulary);
// NSPARSE_SUMMARY_BITMAPS=0 turns them off, and they are skipped when HBM
// cannot hold them. Both paths find the same columns.
void SeismicIndex::refresh_summary_bitmaps() {
    view_.sum_bits = nullptr;
    view_.sum_rank = nullptr;
    view_.sum_nw = 0;
    const char* e = std::getenv("NSPARSE_SUMMARY_BITMAPS");
    if ((e != nullptr && std::string(e) == "0") || !build_summary_bitmaps(view_, d_sum_bits_, d_sum_rank_)) {
        d_sum_bits_.release();
        d_sum_rank_.release();
        return;
    }
    view_.sum_bits = d_sum_bits_.get<uint32_t>();
    view_.sum_rank = d_sum_rank_.get<uint16_t>();
    view_.sum_nw = (dim_ + 31) / 32;
}

void SeismicIndex::ensure_device() {
    std::lock_guard<std::mutex> lock(dev_mu_);
    if (!dev_fwd_valid_) {
        // vector loads read whole 4-element words, so pad every array by 8 elements
        d_indptr_.upload(fwd_.indptr.data(), fwd_.indptr.size() * 4, 64);
        d_indices_.upload(fwd_.indices.data(), fwd_.indices.size() * 2, 64);
        d_values_.upload(fwd_.values.data(), fwd_.values.size(), 64);
        // the forward-index view before the lists: upload_lists derives the
        // block rows from it
        view_.n_docs = static_cast<int32_t>(fwd_.n());
        view_.dim = dim_;
        view_.esz = static_cast<int32_t>(fwd_.esz);
        view_.indptr = d_indptr_.get<int32_t>();
        view_.indices = d_indices_.get<uint16_t>();
        view_.values = d_values_.get<void>();
        dev_fwd_valid_ = true;
    }
    if (!d_err_.bytes()) {
        d_err_.allocate(4);
        hip_check(hipMemset(d_err_.get(), 0, 4), "hipMemset");
    }
    if (!dev_lists_valid_ && host_lists_valid_) upload_lists();
}

const DevIndexView& SeismicIndex::view() {
    ensure_device();
    return view_;
}

// Flattens the per-list structures into the HBM layout of device_index.h,
// validating everything a kernel will dereference.
void SeismicIndex::upload_lists() {
    const int64_t n_docs = fwd_.n();
    const int32_t L = static_cast<int32_t>(lists_.size());
    std::vector<int32_t> lcb(L + 1, 0), ltb(L + 1, 0);
    std::vector<int64_t> leb(L + 1, 0);
    int64_t C = 0, D = 0, T = 0, E = 0;
    for (int32_t l = 0; l < L; ++l) {
        const HostList& h = lists_[l];
        if (h.n_clusters > 65536) throw std::runtime_error("posting list has more clusters than 16-bit ids allow");
        if (h.n_clusters > 0) {
            if (h.offsets.size() != h.n_clusters + 1)
                throw std::runtime_error("posting list offsets disagree with its cluster count");
            if (h.offsets.front() != 0 || static_cast<size_t>(h.offsets.back()) != h.docs.size())
                throw std::runtime_error("posting list offsets do not span its docs");
            for (size_t c = 0; c < h.n_clusters; ++c)
                if (h.offsets[c + 1] < h.offsets[c]) throw std::runtime_error("posting list offsets not monotonic");
        }
        for (idx_t d : h.docs)
            if (d < 0 || d >= n_docs) throw std::runtime_error("posting list references a doc out of range");
        if (!h.term_ids.empty()) {
            if (h.term_ptr.size() != h.term_ids.size() + 1 || h.term_ptr.front() != 0 ||
                static_cast<size_t>(h.term_ptr.back()) != h.csc_cluster.size())
                throw std::runtime_error("summary term offsets are inconsistent");
            for (size_t t = 0; t < h.term_ids.size(); ++t) {
                if (h.term_ptr[t + 1] < h.term_ptr[t]) throw std::runtime_error("summary term offsets not monotonic");
                if (h.term_ids[t] >= dim_) throw std::runtime_error("summary term id out of range");
            }
        } else if (!h.csc_cluster.empty()) {
            throw std::runtime_error("summary entries without terms");
        }
        for (uint16_t c : h.csc_cluster)
            if (c >= h.n_clusters) throw std::runtime_error("summary entry names a cluster out of range");
        if (h.csc_value.size() != h.csc_cluster.size() * fwd_.esz)
            throw std::runtime_error("summary values disagree with the element size");
        if (!h.csc_cluster.empty() && h.esz != fwd_.esz)
            throw std::runtime_error("summary element size disagrees with the index");
        lcb[l + 1] = lcb[l] + static_cast<int32_t>(h.n_clusters);
        ltb[l + 1] = ltb[l] + static_cast<int32_t>(h.term_ids.size());
        leb[l + 1] = leb[l] + static_cast<int64_t>(h.csc_cluster.size());
        C += h.n_clusters;
        D += h.docs.size();
        T += h.term_ids.size();
        E += h.csc_cluster.size();
    }
    if (C >= std::numeric_limits<int32_t>::max() || T >= std::numeric_limits<int32_t>::max())
        throw std::runtime_error("index too large for 32-bit cluster/term tables");
    std::vector<int64_t> cdp(C + 1, 0), stp(T + 1, 0);
    std::vector<int32_t> docs(D);
    std::vector<uint16_t> tids(T), scl(E);
    std::vector<uint8_t> sval(E * fwd_.esz);
    int64_t dc = 0, cc = 0, tc = 0, ec = 0;
    for (int32_t l = 0; l < L; ++l) {
        const HostList& h = lists_[l];
        for (size_t c = 0; c < h.n_clusters; ++c) cdp[cc + c] = dc + h.offsets[c];
        cc += h.n_clusters;
        std::copy(h.docs.begin(), h.docs.end(), docs.begin() + dc);
        dc += h.docs.size();
        cdp[cc] = dc;
        for (size_t t = 0; t < h.term_ids.size(); ++t) {
            tids[tc + t] = h.term_ids[t];
            stp[tc + t] = ec + h.term_ptr[t];
        }
        tc += h.term_ids.size();
        std::copy(h.csc_cluster.begin(), h.csc_cluster.end(), scl.begin() + ec);
        std::copy(h.csc_value.begin(), h.csc_value.end(), sval.begin() + ec * fwd_.esz);
        ec += h.csc_cluster.size();
        stp[tc] = ec;
    }
    d_list_cluster_base_.upload(lcb);
    d_cluster_doc_ptr_.upload(cdp);
    d_docs_.upload(docs);
    d_list_term_base_.upload(ltb);
    d_sum_term_ids_.upload(tids);
    d_sum_term_ptr_.upload(stp);
    d_sum_cluster_.upload(scl);
    d_sum_value_.upload(sval, 64);
    d_list_entry_base_.upload(leb);
    n_lists_ = L;
    view_.n_lists = L;
    view_.max_clusters = max_clusters_;
    view_.list_cluster_base = d_list_cluster_base_.get<int32_t>();
    view_.cluster_doc_ptr = d_cluster_doc_ptr_.get<int64_t>();
    view_.docs = d_docs_.get<int32_t>();
    view_.list_term_base = d_list_term_base_.get<int32_t>();
    view_.sum_term_ids = d_sum_term_ids_.get<uint16_t>();
    view_.sum_term_ptr = d_sum_term_ptr_.get<int64_t>();
    view_.sum_cluster = d_sum_cluster_.get<uint16_t>();
    view_.sum_value = d_sum_value_.get<void>();
    dev_lists_valid_ = true;
    refresh_block_rows();
    refresh_summary_bitmaps();
    max_list_docs_ = max_list_docs(view_);
}

// Rebuilds the per-list host structures from the HBM layout (after a GPU build).
void SeismicIndex::download_lists() {
    const int32_t L = n_lists_;
    auto lcb = d_list_cluster_base_.to_host<int32_t>(L + 1);
    auto ltb = d_list_term_base_.to_host<int32_t>(L + 1);
    auto leb = d_list_entry_base_.to_host<int64_t>(L + 1);
    const int64_t C = lcb[L], T = ltb[L], E = leb[L];
    auto cdp = d_cluster_doc_ptr_.to_host<int64_t>(C + 1);
    auto docs = d_docs_.to_host<int32_t>(cdp[C]);
    auto tids = d_sum_term_ids_.to_host<uint16_t>(T);
    auto stp = d_sum_term_ptr_.to_host<int64_t>(T + 1);
    auto scl = d_sum_cluster_.to_host<uint16_t>(E);
    auto sval = d_sum_value_.to_host<uint8_t>(E * fwd_.esz);
    std::vector<HostList> out(L);
    for (int32_t l = 0; l < L; ++l) {
        HostList& h = out[l];
        h.esz = fwd_.esz;
        const int32_t c0 = lcb[l], c1 = lcb[l + 1];
        h.n_clusters = static_cast<size_t>(c1 - c0);
        if (c1 > c0) {
            const int64_t d0 = cdp[c0];
            h.docs.assign(docs.begin() + d0, docs.begin() + cdp[c1]);
            for (int32_t c = c0; c <= c1; ++c) h.offsets.push_back(static_cast<idx_t>(cdp[c] - d0));
        }
        const int32_t t0 = ltb[l], t1 = ltb[l + 1];
        const int64_t e0 = leb[l], e1 = leb[l + 1];
        if (t1 > t0) {
            h.term_ids.assign(tids.begin() + t0, tids.begin() + t1);
            for (int32_t t = t0; t < t1; ++t) h.term_ptr.push_back(static_cast<idx_t>(stp[t] - e0));
            h.term_ptr.push_back(static_cast<idx_t>(e1 - e0));
        }
        h.csc_cluster.assign(scl.begin() + e0, scl.begin() + e1);
        h.csc_value.assign(sval.begin() + e0 * fwd_.esz, sval.begin() + e1 * fwd_.esz);
    }
    lists_ = std::move(out);
    host_lists_valid_ = true;
}

// Index::search + SeismicIndex::search / SeismicScalarQuantizedIndex::search.
void SeismicIndex::search(idx_t n, const idx_t* indptr, const term_t* indices, const float* values, int k,
                          float* distances, idx_t* labels, const nsparse_search_params* params) {
    throw_if_not_positive(n);
    throw_if_not_positive(k);
    throw_if_any_null(indptr, indices, values, labels, distances);
    const SearchRequest req = request_from(params);
    search_host(n, indptr, indices, values, k, distances, labels, req, nullptr);
}

void SeismicIndex::search_host(idx_t n, const idx_t* indptr, const term_t* indices, const float* values, int k,
                               float* distances, idx_t* labels, const SearchRequest& req, int64_t* stats_out) {
    // SeismicIndex::search returns empty per-query vectors without vectors
    // (seismic_index.cpp:156-159), and Index::search then copies nothing.
    if (!has_vectors_) return;
    if (req.has_sq_range && sq_) check_quantizer_range(req.vmin, req.vmax);
    check_csr(n, indptr, indices, dim_, "query");
    SelectorEval sel(req.selector);
    // Exact-match gate (seismic_common.h:117-128).
    if (sel.enumerable() && sel.size() <= static_cast<size_t>(k)) {
        exact_search(n, indptr, indices, values, k, distances, labels, req, sel.ordered_ids());
        return;
    }
    if (req.cut <= 0) throw std::invalid_argument("value must be positive");
    require_device();
    ensure_device();
    // Queries longer than the per-query workgroup's tables take the spread
    // search, whose tables are sized for the batch's longest query.
    idx_t longest = 0;
    for (idx_t i = 0; i < n; ++i) longest = std::max(longest, indptr[i + 1] - indptr[i]);
    const int qcap = static_cast<int>(std::min<idx_t>(longest, INT32_MAX));
    const bool long_queries = longest > kDeviceQueryMax;
    if (long_queries && (stats_out != nullptr || !u8() || view_.row_ref == nullptr ||
                         !search_spread_fits(view_, max_clusters_, k, req.cut, max_list_docs_, qcap)))
        throw std::invalid_argument("query has more non-zeros than the device path supports (" +
                                    std::to_string(kDeviceQueryMax) +
                                    "; longer ones need a uint8 index with block rows and a cut of at most 64)");
    // One upload of the queries (rebased so the device copy starts at 0) and
    // the selector bitmap, one download of the results and the error flag.
    const int64_t q0 = indptr[0], qnnz = indptr[n] - q0;
    const size_t nk = static_cast<size_t>(n) * k;
    auto st = stages_.take();
    st->begin();
    const int s_qip = st->add(HostStage::kIn, static_cast<size_t>(n + 1) * 4);
    const int s_qix = st->add(HostStage::kIn, static_cast<size_t>(qnnz) * 2);
    const int s_qv = st->add(HostStage::kIn, static_cast<size_t>(qnnz) * 4);
    const int64_t nd = fwd_.n();
    const int s_sel = sel.active() ? st->add(HostStage::kIn, static_cast<size_t>((nd + 31) / 32) * 4) : -1;
    const int s_out_d = st->add(HostStage::kOut, nk * 4);
    const int s_out_l = st->add(HostStage::kOut, nk * 4);
    const int s_err = st->add(HostStage::kOut, 4);
    const int s_stats =
        stats_out != nullptr ? st->add(HostStage::kOut, static_cast<size_t>(n) * kStatFields * 8) : -1;
    const bool spread = long_queries || use_spread(n, k, req.cut, stats_out != nullptr, st->stream());
    const int s_spread =
        spread ? st->add(HostStage::kScratch,
                         search_spread_scratch_bytes(view_, max_clusters_, k, req.cut, max_list_docs_, n, qcap))
               : -1;
    st->commit();
    int32_t* qip = st->host<int32_t>(s_qip);
    for (idx_t i = 0; i <= n; ++i) qip[i] = static_cast<int32_t>(indptr[i] - q0);
    st->put(s_qix, indices + q0, static_cast<size_t>(qnnz) * 2);
    st->put(s_qv, values + q0, static_cast<size_t>(qnnz) * 4);
    if (s_sel >= 0) {
        uint32_t* bm = st->host<uint32_t>(s_sel);
        std::fill(bm, bm + (nd + 31) / 32, 0U);
        for (int64_t d = 0; d < nd; ++d)
            if (sel.is_member(d)) bm[d >> 5] |= 1U << (d & 31);
    }
    st->upload();
    SearchArgs a{};
    a.ix = view_;
    a.q_indptr = st->dev<int32_t>(s_qip);
    a.q_indices = st->dev<uint16_t>(s_qix);
    a.q_values = st->dev<float>(s_qv);
    a.ks = nullptr;
    a.k_max = k;
    a.cut = req.cut;
    a.heap_factor = req.heap_factor;
    const float qmin = req.has_sq_range ? req.vmin : vmin_;
    const float qmax = req.has_sq_range ? req.vmax : vmax_;
    a.q_vmin = qmin;
    a.q_scale = code_max() / (qmax - qmin);
    {
        // ScalarQuantizer::decode_dot_product (utils/scalar_quantizer.h:85-93)
        const float ingest_range = vmax_ - vmin_;
        const float query_range = qmax - qmin;
        const float max_q = code_max();
        a.decode_scale = (ingest_range * query_range) / (max_q * max_q);
    }
    a.sel_of_query = nullptr;
    a.sel_bitmaps = nullptr;
    a.sel_all = s_sel >= 0 ? st->dev<uint32_t>(s_sel) : nullptr;
    a.out_d = st->dev<float>(s_out_d);
    a.out_l = st->dev<int32_t>(s_out_l);
    a.stats = s_stats >= 0 ? st->dev<int64_t>(s_stats) : nullptr;
    a.error_flag = st->dev<int32_t>(s_err);
    a.n_queries = n;
    a.query_base = 0;
    if (spread)
        launch_search_spread(a, max_clusters_, k, max_list_docs_, st->dev<void>(s_spread), st->stream(), qcap);
    else
        launch_search_checked(a, max_clusters_, k, u8(), st->stream());
    st->download_and_wait();
    const int32_t err = *st->host<int32_t>(s_err);
    if (err != 0) throw std::invalid_argument(device_error_message(err));
    std::memcpy(distances, st->host(s_out_d), nk * 4);
    std::memcpy(labels, st->host(s_out_l), nk * 4);
    if (stats_out != nullptr) std::memcpy(stats_out, st->host(s_stats), static_cast<size_t>(n) * kStatFields * 8);
}

void SeismicIndex::search_device(int64_t n, const int32_t* d_indptr, const uint16_t* d_indices,
                                 const float* d_values, const int32_t* d_ks, int k_max, const SearchRequest& req,
                                 const int32_t* d_sel_of_query, const uint32_t* const* d_bitmaps, int32_t n_bitmaps,
                                 float* d_out_d, int32_t* d_out_l, int64_t* d_stats, hipStream_t stream,
                                 int32_t* d_err, int64_t spread_cap) {
    if (!has_vectors_ && n > 0 && k_max > 0) return;
    SearchArgs a = search_args(n, d_indptr, d_indices, d_values, d_ks, k_max, req, d_sel_of_query, d_bitmaps,
                               n_bitmaps, d_out_d, d_out_l, d_stats);
    if (d_err != nullptr) a.error_flag = d_err;
    if (use_spread(n, k_max, req.cut, d_stats != nullptr, stream, spread_cap)) {
        size_t slot = 0;
        void* scratch = spread_scratch_.acquire(
            search_spread_scratch_bytes(view_, max_clusters_, k_max, req.cut, max_list_docs_, n), slot);
        try {
            launch_search_spread(a, max_clusters_, k_max, max_list_docs_, scratch, stream);
        } catch (...) {
            spread_scratch_.release(slot, stream);
            throw;
        }
        spread_scratch_.release(slot, stream);
        return;
    }
    launch_search_checked(a, max_clusters_, k_max, u8(), stream);
}

// A launch of a few queries leaves most CUs idle with one workgroup per query
// (about 1.5 ms for any batch up to a hundred at 8.8M docs); the spread search
// puts each query on the whole GPU, 0.13 ms for one query and 0.8 ms for 32.
// It reads every doc of the cut lists, so on a GPU shared with other launches
// it costs more than it saves. Batches of at most cap queries take it (default
// 32, for a caller alone on the GPU); NSPARSE_SPREAD_MAX_QUERIES overrides the
// cap and NSPARSE_SPREAD=0 turns it off. Launches under stream capture, or with
// traversal counters, keep the per-query workgroup.
bool SeismicIndex::use_spread(int64_t n, int k_max, int cut, bool stats, hipStream_t stream, int64_t cap) const {
    const char* off = std::getenv("NSPARSE_SPREAD");
    if (off != nullptr && std::string(off) == "0") return false;
    const char* e = std::getenv("NSPARSE_SPREAD_MAX_QUERIES");
    const int64_t max_n = e != nullptr ? std::atoll(e) : cap >= 0 ? cap : 32;
    if (stats || !u8() || n > max_n || view_.row_ref == nullptr) return false;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
    return search_spread_fits(view_, max_clusters_, k_max, cut, max_list_docs_);
}

void SeismicIndex::search_device_noalloc(int64_t n, const int32_t* d_indptr, const uint16_t* d_indices,
                                         const float* d_values, const int32_t* d_ks, int k_max,
                                         const SearchRequest& req, const int32_t* d_sel_of_query,
                                         const uint32_t* const* d_bitmaps, int32_t n_bitmaps, float* d_out_d,
                                         int32_t* d_out_l, int32_t* d_err, hipStream_t stream) {
    if (d_err == nullptr) throw std::invalid_argument("the allocation-free search needs a device error flag");
    if (!has_vectors_ && n > 0 && k_max > 0) return;
    {
        // ensure_device (called by search_args) must find nothing to upload
        std::lock_guard<std::mutex> lock(dev_mu_);
        if (!dev_fwd_valid_ || (!dev_lists_valid_ && host_lists_valid_) || !d_err_.bytes())
            throw std::invalid_argument(
                "the allocation-free search needs the index resident on the device: build it, or search it once");
    }
    SearchArgs a = search_args(n, d_indptr, d_indices, d_values, d_ks, k_max, req, d_sel_of_query, d_bitmaps,
                               n_bitmaps, d_out_d, d_out_l, nullptr);
    a.error_flag = d_err;
    launch_search_noalloc(a, max_clusters_, k_max, u8(), stream);
}

SearchArgs SeismicIndex::search_args(int64_t n, const int32_t* d_indptr, const uint16_t* d_indices,
                                     const float* d_values, const int32_t* d_ks, int k_max, const SearchRequest& req,
                                     const int32_t* d_sel_of_query, const uint32_t* const* d_bitmaps,
                                     int32_t n_bitmaps, float* d_out_d, int32_t* d_out_l, int64_t* d_stats) {
    if (n <= 0) throw std::invalid_argument("value must be positive");
    if (k_max <= 0) throw std::invalid_argument("value must be positive");
    throw_if_any_null(d_indptr, d_indices, d_values, d_out_d, d_out_l);
    if (!has_vectors_) throw std::invalid_argument("the index holds no vectors");
    if (req.cut <= 0) throw std::invalid_argument("value must be positive");
    if (req.has_sq_range && sq_) check_quantizer_range(req.vmin, req.vmax);
    require_device();
    ensure_device();
    SearchArgs a{};
    a.ix = view_;
    a.q_indptr = d_indptr;
    a.q_indices = d_indices;
    a.q_values = d_values;
    a.ks = d_ks;
    a.k_max = k_max;
    a.cut = req.cut;
    a.heap_factor = req.heap_factor;
    const float qmin = req.has_sq_range ? req.vmin : vmin_;
    const float qmax = req.has_sq_range ? req.vmax : vmax_;
    a.q_vmin = qmin;
    a.q_scale = code_max() / (qmax - qmin);
    const float ingest_range = vmax_ - vmin_;
    const float query_range = qmax - qmin;
    const float max_q = code_max();
    a.decode_scale = (ingest_range * query_range) / (max_q * max_q);
    a.sel_of_query = d_sel_of_query;
    if (d_sel_of_query != nullptr && (d_bitmaps == nullptr || n_bitmaps <= 0))
        throw std::invalid_argument("per-query selectors need a bitmap table");
    a.sel_bitmaps = d_bitmaps;
    a.n_bitmaps = d_sel_of_query != nullptr ? n_bitmaps : 0;
    a.sel_all = nullptr;
    a.out_d = d_out_d;
    a.out_l = d_out_l;
    a.stats = d_stats;
    a.error_flag = d_err_.get<int32_t>();
    a.n_queries = n;
    a.query_base = 0;
    return a;
}

// ---- SearchPlan

namespace {

float time_launch(const std::function<void()>& launch, hipStream_t stream, int reps) {
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    float best = std::numeric_limits<float>::infinity();
    for (int r = 0; r < reps; ++r) {
        hip_check(hipEventRecord(e0, stream), "hipEventRecord");
        launch();
        hip_check(hipEventRecord(e1, stream), "hipEventRecord");
        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
        float ms = 0.0F;
        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
        best = std::min(best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return best;
}

}  // namespace

SearchPlan::SearchPlan(SeismicIndex& index, const SearchArgs& args, int k_max, const SearchConfig* config,
                       hipStream_t stream)
    : args_(args), k_max_(k_max), max_clusters_(index.max_clusters()), sq_(index.u8()) {
    scratch_.allocate(search_scratch_bytes(args.n_queries));
    const int dim = args.ix.dim;
    if (config != nullptr) {
        config_ = resolve_search_config(dim, max_clusters_, k_max_, sq_, *config);
    } else {
        // every candidate schedule on this very batch: one warm-up, best of two
        std::vector<SearchConfig> cands;
        for (int lookup : {1, 0})
            for (int block : {1024, 512, 256})
                for (int sched : {0, 1, 2}) {
                    SearchConfig c;
                    c.lookup = lookup;
                    c.block = block;
                    c.persistent = sched > 0 ? 1 : 0;
                    c.longest_first = sched == 2 ? 1 : 0;
                    c = resolve_search_config(dim, max_clusters_, k_max_, sq_, c);
                    bool seen = false;
                    for (const auto& o : cands)
                        seen = seen || (o.lookup == c.lookup && o.block == c.block && o.persistent == c.persistent &&
                                        o.longest_first == c.longest_first);
                    if (!seen) cands.push_back(c);
                }
        float best = std::numeric_limits<float>::infinity();
        for (const auto& c : cands) {
            auto go = [&] { launch_search(args_, max_clusters_, k_max_, sq_, c, scratch_.get(), stream); };
            go();
            const float ms = time_launch(go, stream, 2);
            tuning_.push_back({c, ms});
            if (ms < best) {
                best = ms;
                config_ = c;
            }
        }
    }
    // capture the launch sequence (queue reset, optional ordering, search)
    hipStream_t cs = nullptr;
    hip_check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
    try {
        hip_check(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
        launch_search(args_, max_clusters_, k_max_, sq_, config_, scratch_.get(), cs);
        hip_check(hipStreamEndCapture(cs, &graph_), "hipStreamEndCapture");
        hip_check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(cs, &g);
        if (g != nullptr) (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(cs);
        throw;
    }
    (void)hipStreamDestroy(cs);
}

SearchPlan::~SearchPlan() {
    if (exec_ != nullptr) (void)hipGraphExecDestroy(exec_);
    if (graph_ != nullptr) (void)hipGraphDestroy(graph_);
}

void SearchPlan::launch(hipStream_t stream) { hip_check(hipGraphLaunch(exec_, stream), "hipGraphLaunch"); }

int32_t SeismicIndex::last_device_error() {
    if (!d_err_.bytes()) return 0;
    int32_t e = 0;
    d_err_.download(&e, 4);
    return e;
}

// ExactMatcher::search (exact_matcher.h:23-58) on the GPU.
void SeismicIndex::exact_search(idx_t n, const idx_t* indptr, const term_t* indices, const float* values, int k,
                                float* distances, idx_t* labels, const SearchRequest& req,
                                const std::vector<idx_t>& ids) {
    for (idx_t d : ids)
        if (d < 0 || d >= fwd_.n()) throw std::invalid_argument("selector id out of range");
    require_device();
    ensure_device();
    for (idx_t i = 0; i < n; ++i)
        if (indptr[i + 1] - indptr[i] > kDeviceQueryMax)
            throw std::invalid_argument("query has more non-zeros than the device path supports");
    const int64_t q0 = indptr[0], qnnz = indptr[n] - q0;
    const size_t nk = static_cast<size_t>(n) * k;
    auto st = stages_.take();
    st->begin();
    const int s_qip = st->add(HostStage::kIn, static_cast<size_t>(n + 1) * 4);
    const int s_qix = st->add(HostStage::kIn, static_cast<size_t>(qnnz) * 2);
    const int s_qv = st->add(HostStage::kIn, static_cast<size_t>(qnnz) * 4);
    const int s_ids = st->add(HostStage::kIn, ids.size() * 4);
    const int s_out_d = st->add(HostStage::kOut, nk * 4);
    const int s_out_l = st->add(HostStage::kOut, nk * 4);
    st->commit();
    int32_t* qip = st->host<int32_t>(s_qip);
    for (idx_t i = 0; i <= n; ++i) qip[i] = static_cast<int32_t>(indptr[i] - q0);
    st->put(s_qix, indices + q0, static_cast<size_t>(qnnz) * 2);
    st->put(s_qv, values + q0, static_cast<size_t>(qnnz) * 4);
    st->put(s_ids, ids.data(), ids.size() * 4);
    st->upload();
    const float qmin = req.has_sq_range ? req.vmin : vmin_;
    const float qmax = req.has_sq_range ? req.vmax : vmax_;
    const float ingest_range = vmax_ - vmin_;
    const float query_range = qmax - qmin;
    const float max_q = code_max();
    ExactArgs a{};
    a.ix = view_;
    a.q_indptr = st->dev<int32_t>(s_qip);
    a.q_indices = st->dev<uint16_t>(s_qix);
    a.q_values = st->dev<float>(s_qv);
    a.ids = st->dev<int32_t>(s_ids);
    a.n_ids = static_cast<int32_t>(ids.size());
    a.k = k;
    a.q_vmin = qmin;
    a.q_scale = code_max() / (qmax - qmin);
    a.decode_scale = (ingest_range * query_range) / (max_q * max_q);
    a.out_d = st->dev<float>(s_out_d);
    a.out_l = st->dev<int32_t>(s_out_l);
    a.n_queries = n;
    launch_exact_checked(a, u8(), st->stream());
    st->download_and_wait();
    std::memcpy(distances, st->host(s_out_d), nk * 4);
    std::memcpy(labels, st->host(s_out_l), nk * 4);
}

// ---------------------------------------------------------------- layout IO
// SeismicScalarQuantizedIndex::write_header (seismic_scalar_quantized_index.cpp:421-428),
// SparseVectors::serialize (sparse_vectors.cpp:206-231),
// SeismicInvertedListsWriter::serialize (io/seismic_invlists_writer.h:36-43),
// InvertedListClusters::serialize (cluster/inverted_list_clusters.cpp:333-363).
void SeismicIndex::write_payload(ByteWriter& w) const {
    if (sq_) {
        w.scalar<uint8_t>(w16_ ? 1 : 0);  // QuantizerType: QT_8bit 0, QT_16bit 1
        w.scalar<float>(vmin_);
        w.scalar<float>(vmax_);
    }
    const size_t count = has_vectors_ ? static_cast<size_t>(fwd_.n()) : 0;
    w.scalar<size_t>(count);
    if (count > 0) {
        w.scalar<size_t>(static_cast<size_t>(dim_));
        w.scalar<size_t>(fwd_.esz);
        w.padded(fwd_.indptr.data(), count + 1);
        const size_t nnz = static_cast<size_t>(fwd_.indptr[count]);
        w.padded(fwd_.indices.data(), nnz);
        w.padded(fwd_.values.data(), nnz * fwd_.esz, fwd_.esz);
    }
    const auto& L = const_cast<SeismicIndex*>(this)->lists();
    w.scalar<size_t>(L.size());
    for (const auto& h : L) {
        w.scalar<size_t>(h.docs.size());
        w.padded(h.docs.data(), h.docs.size());
        w.scalar<size_t>(h.offsets.size());
        w.padded(h.offsets.data(), h.offsets.size());
        w.scalar<size_t>(h.n_clusters);
        w.scalar<size_t>(h.esz);
        w.scalar<size_t>(h.term_ids.size());
        if (!h.term_ids.empty()) {
            w.padded(h.term_ids.data(), h.term_ids.size());
            w.padded(h.term_ptr.data(), h.term_ids.size() + 1);
        }
        w.scalar<size_t>(h.csc_cluster.size());
        w.padded(h.csc_cluster.data(), h.csc_cluster.size());
        w.padded(h.csc_value.data(), h.csc_value.size(), h.esz);
    }
}

// The matching readers: SparseVectors::deserialize (sparse_vectors.cpp:233-253),
// InvertedListClusters::deserialize (inverted_list_clusters.cpp:365-389).
void SeismicIndex::read_payload(ByteReader& r) {
    if (sq_) {
        const uint8_t qt = r.scalar<uint8_t>();
        const float lo = r.scalar<float>();
        const float hi = r.scalar<float>();
        if (qt != 0 && qt != 1) throw std::runtime_error("index file declares an unknown quantizer type");
        check_quantizer_range(lo, hi);
        w16_ = qt == 1;
        vmin_ = lo;
        vmax_ = hi;
    }
    const size_t count = r.scalar<size_t>();
    fwd_ = HostCsr{};
    fwd_.esz = sq_ ? (w16_ ? 2 : 1) : 4;
    fwd_.dim = static_cast<size_t>(dim_);
    if (count > 0) {
        const size_t fdim = r.scalar<size_t>();
        const size_t esz = r.scalar<size_t>();
        (void)fdim;
        if (esz != fwd_.esz)
            throw std::runtime_error(sq_ ? "index file's element size disagrees with its quantizer type"
                                         : "index file's element size is not the seismic index's");
        if (count >= static_cast<size_t>(std::numeric_limits<idx_t>::max()))
            throw std::runtime_error("index file declares too many vectors");
        fwd_.indptr = r.padded<idx_t>(count + 1);
        const int64_t nnz = fwd_.indptr[count];
        if (fwd_.indptr[0] != 0 || nnz < 0) throw std::runtime_error("index file has a corrupt indptr");
        for (size_t i = 0; i < count; ++i)
            if (fwd_.indptr[i + 1] < fwd_.indptr[i]) throw std::runtime_error("index file has a corrupt indptr");
        fwd_.indices = r.padded<term_t>(static_cast<size_t>(nnz));
        for (term_t t : fwd_.indices)
            if (t >= dim_) throw std::runtime_error("index file has a term id out of range");
        fwd_.values = r.padded<uint8_t>(static_cast<size_t>(nnz) * esz, esz);
    }
    has_vectors_ = count > 0;
    dev_fwd_valid_ = false;
    const size_t n_lists = r.scalar<size_t>();
    if (n_lists > r.remaining()) throw std::runtime_error("index file is truncated");
    std::vector<HostList> lists(n_lists);
    for (auto& h : lists) {
        const size_t n_docs = r.scalar<size_t>();
        h.docs = r.padded<idx_t>(n_docs);
        const size_t n_off = r.scalar<size_t>();
        h.offsets = r.padded<idx_t>(n_off);
        h.n_clusters = r.scalar<size_t>();
        h.esz = r.scalar<size_t>();
        const size_t n_terms = r.scalar<size_t>();
        if (n_terms > 0) {
            h.term_ids = r.padded<term_t>(n_terms);
            h.term_ptr = r.padded<idx_t>(n_terms + 1);
        }
        const size_t nnz = r.scalar<size_t>();
        h.csc_cluster = r.padded<uint16_t>(nnz);
        if (h.esz != 1 && h.esz != 2 && h.esz != 4 && h.esz != 8)
            throw std::runtime_error("index file declares an invalid element width");
        if (nnz > r.remaining() / h.esz) throw std::runtime_error("index file is truncated");
        h.csc_value = r.padded<uint8_t>(nnz * h.esz, h.esz);
    }
    set_lists(std::move(lists));
    // An index read from disk carries the library defaults (seismic_common.h:58-59).
    params_ = Params{-1, -1, 0.4F, -1};
}

// ========================================================= staged host queries
namespace {
// Segments of a brute-force or inverted search over host arrays.
struct StagedQueries {
    int qip, qix, qv, extra, out_d, out_l, scratch;
};

// Lays a batch of host queries (rebased so the device copy starts at 0), an
// optional extra input (an allow bitmap), the two outputs and a kernel's
// scratch out in a stage, and fills the input mirror.
StagedQueries stage_queries(HostStage& st, idx_t n, const idx_t* indptr, const term_t* indices, const float* values,
                            int k, size_t scratch_bytes, const void* extra = nullptr, size_t extra_bytes = 0) {
    const int64_t q0 = indptr[0], qnnz = indptr[n] - q0;
    const size_t nk = static_cast<size_t>(n) * k;
    StagedQueries q{};
    st.begin();
    q.qip = st.add(HostStage::kIn, static_cast<size_t>(n + 1) * 4);
    q.qix = st.add(HostStage::kIn, static_cast<size_t>(qnnz) * 2);
    q.qv = st.add(HostStage::kIn, static_cast<size_t>(qnnz) * 4);
    q.extra = extra != nullptr ? st.add(HostStage::kIn, extra_bytes) : -1;
    q.out_d = st.add(HostStage::kOut, nk * 4);
    q.out_l = st.add(HostStage::kOut, nk * 4);
    q.scratch = st.add(HostStage::kScratch, scratch_bytes);
    st.commit();
    int32_t* qip = st.host<int32_t>(q.qip);
    for (idx_t i = 0; i <= n; ++i) qip[i] = static_cast<int32_t>(indptr[i] - q0);
    st.put(q.qix, indices + q0, static_cast<size_t>(qnnz) * 2);
    st.put(q.qv, values + q0, static_cast<size_t>(qnnz) * 4);
    if (extra != nullptr) st.put(q.extra, extra, extra_bytes);
    return q;
}
}  // namespace

// ============================================================== BrutalIndex

BrutalIndex::BrutalIndex(int dim) : Index(dim) {
    fwd_.esz = 4;
    fwd_.dim = dim > 0 ? static_cast<size_t>(dim) : 0;
}

// BrutalIndex::add (brutal_index.cpp:27-42): FP32 rows appended as given.
void BrutalIndex::add(idx_t n, const idx_t* indptr, const term_t* indices, const float* values) {
    throw_if_not_positive(n);
    throw_if_any_null(indptr, indices, values);
    if (dim_ <= 0) throw std::invalid_argument("value must be positive");
    check_csr(n, indptr, indices, dim_, "add:");
    const int64_t nnz = indptr[n];
    if (static_cast<int64_t>(fwd_.indices.size()) + nnz > std::numeric_limits<idx_t>::max())
        throw std::invalid_argument("too many non-zeros for 32-bit offsets");
    if (fwd_.indptr.empty()) fwd_.indptr.push_back(0);
    const idx_t off = fwd_.indptr.back();
    fwd_.indices.insert(fwd_.indices.end(), indices, indices + nnz);
    const size_t old_vals = fwd_.values.size();
    fwd_.values.resize(old_vals + static_cast<size_t>(nnz) * 4);
    std::memcpy(fwd_.values.data() + old_vals, values, static_cast<size_t>(nnz) * 4);
    for (idx_t i = 1; i <= n; ++i) fwd_.indptr.push_back(indptr[i] + off);
    dev_valid_ = false;
}

// BrutalIndex::search / single_query (brutal_index.cpp:44-96): every document
// is a candidate (zero scores included), the k best by (score desc, doc id
// asc) come back, padded with (-1.0f, -1) when the corpus is smaller than k.
// The selector is not read. Without vectors nothing is written, as the
// reference's empty per-query results leave the outputs untouched.
void BrutalIndex::search(idx_t n, const idx_t* indptr, const term_t* indices, const float* values, int k,
                         float* distances, idx_t* labels, const nsparse_search_params*) {
    throw_if_not_positive(n);
    throw_if_not_positive(k);
    throw_if_any_null(indptr, indices, values, labels, distances);
    if (fwd_.n() == 0) return;
    check_csr(n, indptr, indices, dim_, "query");
    if (k > kBruteKMax)
        throw std::invalid_argument("brutal search supports k <= " + std::to_string(kBruteKMax) + " on the device");
    for (idx_t i = 0; i < n; ++i)
        if (indptr[i + 1] - indptr[i] > kDeviceQueryMax)
            throw std::invalid_argument("query has more non-zeros than the device path supports (" +
                                        std::to_string(kDeviceQueryMax) + ")");
    require_device();
    {
        std::lock_guard<std::mutex> lock(dev_mu_);
        if (!dev_valid_) {
            d_indptr_.upload(fwd_.indptr.data(), fwd_.indptr.size() * 4, 64);
            d_indices_.upload(fwd_.indices.data(), fwd_.indices.size() * 2, 64);
            d_values_.upload(fwd_.values.data(), fwd_.values.size(), 64);
            dev_valid_ = true;
        }
    }
    auto st = stages_.take();
    const StagedQueries q = stage_queries(*st, n, indptr, indices, values, k, brute_search_scratch_bytes(fwd_.n(), n));
    st->upload();
    launch_brute_search(d_indptr_.get<int32_t>(), d_indices_.get<uint16_t>(), d_values_.get<float>(), fwd_.n(), dim_,
                        st->dev<int32_t>(q.qip), st->dev<uint16_t>(q.qix), st->dev<float>(q.qv), n, k,
                        st->dev<float>(q.out_d), st->dev<int32_t>(q.out_l), st->dev<float>(q.scratch), st->stream());
    st->download_and_wait();
    std::memcpy(distances, st->host(q.out_d), static_cast<size_t>(n) * k * 4);
    std::memcpy(labels, st->host(q.out_l), static_cast<size_t>(n) * k * 4);
}

// write_index / read_index (io/index_io.cpp:120-141): BrutalIndex is not an
// IndexIO.
void BrutalIndex::write_payload(ByteWriter&) const {
    throw std::runtime_error("Index does not support serialization");
}

void BrutalIndex::read_payload(ByteReader&) { throw std::runtime_error("Index does not support serialization"); }

// ============================================================== InvertedIndex

InvertedIndex::InvertedIndex(int dim) : Index(dim) {
    pending_.esz = 4;
    pending_.dim = dim > 0 ? static_cast<size_t>(dim) : 0;
}

// InvertedIndex::add (inverted_index.cpp:218-234): FP32 rows appended as given;
// term ids are checked by build(), as the reference's are.
void InvertedIndex::add(idx_t n, const idx_t* indptr, const term_t* indices, const float* values) {
    throw_if_not_positive(n);
    throw_if_any_null(indptr, indices, values);
    for (idx_t i = 0; i < n; ++i)
        if (indptr[i + 1] < indptr[i]) throw std::invalid_argument("add: indptr is not monotonic");
    if (indptr[0] < 0) throw std::invalid_argument("add: indptr must not be negative");
    const int64_t nnz = indptr[n];
    if (static_cast<int64_t>(pending_.indices.size()) + nnz > std::numeric_limits<idx_t>::max())
        throw std::invalid_argument("too many non-zeros for 32-bit offsets");
    if (pending_.indptr.empty()) pending_.indptr.push_back(0);
    const idx_t off = pending_.indptr.back();
    pending_.indices.insert(pending_.indices.end(), indices, indices + nnz);
    const size_t old_vals = pending_.values.size();
    pending_.values.resize(old_vals + static_cast<size_t>(nnz) * 4);
    std::memcpy(pending_.values.data() + old_vals, values, static_cast<size_t>(nnz) * 4);
    for (idx_t i = 1; i <= n; ++i) pending_.indptr.push_back(indptr[i] + off);
    n_vectors_ += n;
}

// InvertedIndex::build (inverted_index.cpp:236-252) on the device. Like the
// reference's, it consumes the vectors added since the last build; with none,
// where the reference dereferences its released vectors, it raises.
void InvertedIndex::build() {
    if (pending_.indptr.empty()) throw std::runtime_error("no vectors added since the last build");
    if (dim_ <= 0) throw std::invalid_argument("value must be positive");
    for (term_t t : pending_.indices)
        if (t >= dim_) throw std::invalid_argument("term_id out of range");  // ArrayInvertedLists::add_entries
    require_device();
    const int64_t n = pending_.n(), nnz = static_cast<int64_t>(pending_.indices.size());
    {
        DeviceBuffer d_ip, d_ix, d_v;
        d_ip.upload(pending_.indptr.data(), pending_.indptr.size() * 4);
        d_ix.upload(pending_.indices.data(), pending_.indices.size() * 2);
        d_v.upload(pending_.values.data(), pending_.values.size());
        build_postings_on_device(d_ip.get<int32_t>(), d_ix.get<uint16_t>(), d_v.get<float>(), n, nnz, dim_, dev_);
    }
    pending_ = HostCsr{};
    pending_.esz = 4;
    pending_.dim = static_cast<size_t>(dim_);
    list_docs_ = n;
    n_lists_ = dim_;
    built_ = true;
    dev_valid_ = true;
    host_valid_ = false;
}

void InvertedIndex::ensure_device() {
    std::lock_guard<std::mutex> lock(mu_);
    if (dev_valid_) return;
    require_device();
    dev_.list_ptr.upload(h_list_ptr_);
    dev_.docs.upload(h_docs_, 64);
    dev_.vals.upload(h_vals_, 64);
    dev_.n_postings = static_cast<int64_t>(h_docs_.size());
    dev_.duplicate_docs = false;
    for (size_t t = 0; t + 1 < h_list_ptr_.size() && !dev_.duplicate_docs; ++t)
        for (int64_t i = h_list_ptr_[t] + 1; i < h_list_ptr_[t + 1]; ++i)
            if (h_docs_[i] == h_docs_[i - 1]) {
                dev_.duplicate_docs = true;
                break;
            }
    build_window_starts(dev_, list_docs_, static_cast<int>(h_list_ptr_.size()) - 1);
    dev_valid_ = true;
}

void InvertedIndex::ensure_host() const {
    std::lock_guard<std::mutex> lock(mu_);
    if (host_valid_) return;
    const size_t L = static_cast<size_t>(n_lists_);
    h_list_ptr_ = dev_.list_ptr.to_host<int64_t>(L + 1);
    h_docs_ = dev_.docs.to_host<idx_t>(static_cast<size_t>(dev_.n_postings));
    h_vals_ = dev_.vals.to_host<float>(static_cast<size_t>(dev_.n_postings));
    h_max_ = dev_.max_score.to_host<float>(L);
    host_valid_ = true;
}

// InvertedIndex::search (inverted_index.cpp:254-293): the selector is read off
// the base SearchParameters, whatever its kind, and filters every doc.
void InvertedIndex::search(idx_t n, const idx_t* indptr, const term_t* indices, const float* values, int k,
                           float* distances, idx_t* labels, const nsparse_search_params* params) {
    throw_if_not_positive(n);
    throw_if_not_positive(k);
    throw_if_any_null(indptr, indices, values, labels, distances);
    const SearchRequest req = request_from(params);
    SelectorEval sel(req.selector);
    if (!built_) return;
    std::vector<uint32_t> allow;
    if (sel.active()) {
        allow.assign(static_cast<size_t>((list_docs_ + 31) / 32), 0);
        for (int64_t d = 0; d < list_docs_; ++d)
            if (sel.is_member(d)) allow[d >> 5] |= 1U << (d & 31);
    }
    search_allowed(n, indptr, indices, values, k, distances, labels, sel.active() ? &allow : nullptr);
}

// An index that was never built leaves the outputs as they are: the
// reference's per-query results are then empty (inverted_index.cpp:258-261).
void InvertedIndex::search_allowed(idx_t n, const idx_t* indptr, const term_t* indices, const float* values, int k,
                                   float* distances, idx_t* labels, const std::vector<uint32_t>* allow) {
    throw_if_not_positive(n);
    throw_if_not_positive(k);
    throw_if_any_null(indptr, indices, values, labels, distances);
    if (!built_) return;
    check_csr(n, indptr, indices, dim_, "query");
    if (k > kBruteKMax)
        throw std::invalid_argument("inverted search supports k <= " + std::to_string(kBruteKMax) + " on the device");
    for (idx_t i = 0; i < n; ++i)
        if (indptr[i + 1] - indptr[i] > kDeviceQueryMax)
            throw std::invalid_argument("query has more non-zeros than the device path supports (" +
                                        std::to_string(kDeviceQueryMax) + ")");
    if (list_docs_ == 0) {
        std::fill(distances, distances + static_cast<int64_t>(n) * k, -1.0F);
        std::fill(labels, labels + static_cast<int64_t>(n) * k, -1);
        return;
    }
    ensure_device();
    auto st = stages_.take();
    const StagedQueries q = stage_queries(*st, n, indptr, indices, values, k, brute_search_scratch_bytes(list_docs_, n),
                                          allow != nullptr ? allow->data() : nullptr,
                                          allow != nullptr ? allow->size() * 4 : 0);
    st->upload();
    launch_inverted_search(dev_, list_docs_, st->dev<int32_t>(q.qip), st->dev<uint16_t>(q.qix), st->dev<float>(q.qv),
                           n, k, allow != nullptr ? st->dev<uint32_t>(q.extra) : nullptr, st->dev<float>(q.out_d),
                           st->dev<int32_t>(q.out_l), st->dev<float>(q.scratch), st->stream());
    st->download_and_wait();
    std::memcpy(distances, st->host(q.out_d), static_cast<size_t>(n) * k * 4);
    std::memcpy(labels, st->host(q.out_l), static_cast<size_t>(n) * k * 4);
}

// InvertedIndex::write_index (inverted_index.cpp:367-402): document count,
// list count, element width, then each list's doc ids and value bytes, then
// the per-term maxima; plain writes, no alignment padding.
void InvertedIndex::write_payload(ByteWriter& w) const {
    w.scalar<size_t>(static_cast<size_t>(n_vectors_));
    const size_t n_terms = built_ ? static_cast<size_t>(n_lists_) : 0;
    w.scalar<size_t>(n_terms);
    w.scalar<size_t>(4);
    if (built_) ensure_host();
    for (size_t t = 0; t < n_terms; ++t) {
        const int64_t a = h_list_ptr_[t], b = h_list_ptr_[t + 1];
        w.scalar<size_t>(static_cast<size_t>(b - a));
        if (b > a) {
            w.write(h_docs_.data() + a, static_cast<size_t>(b - a) * 4);
            w.write(h_vals_.data() + a, static_cast<size_t>(b - a) * 4);
        }
    }
    const size_t ns = built_ ? h_max_.size() : 0;
    w.scalar<size_t>(ns);
    if (ns > 0) w.write(h_max_.data(), ns * 4);
}

// InvertedIndex::read_index (inverted_index.cpp:404-447). Each list must hold
// non-negative doc ids in ascending order, as every writer produces them: the
// device search relies on it, so a file breaking it is rejected as corrupt.
void InvertedIndex::read_payload(ByteReader& r) {
    n_vectors_ = static_cast<int64_t>(r.scalar<size_t>());
    const size_t n_terms = r.scalar<size_t>();
    const size_t esz = r.scalar<size_t>();
    if (esz != 4) throw std::runtime_error("inverted index file declares an unsupported element width");
    const size_t L = std::max<size_t>(n_terms, dim_ > 0 ? static_cast<size_t>(dim_) : 0);
    h_list_ptr_.assign(L + 1, 0);
    h_docs_.clear();
    h_vals_.clear();
    int64_t max_doc = -1;
    for (size_t t = 0; t < n_terms; ++t) {
        const size_t len = r.scalar<size_t>();
        if (len > r.remaining() / 8) throw std::runtime_error("index file is truncated");
        if (h_docs_.size() + len > static_cast<size_t>(std::numeric_limits<idx_t>::max()))
            throw std::runtime_error("inverted index holds more postings than 32-bit offsets address");
        const size_t at = h_docs_.size();
        h_docs_.resize(at + len);
        h_vals_.resize(at + len);
        r.read(h_docs_.data() + at, len * 4);
        r.read(h_vals_.data() + at, len * 4);
        for (size_t i = at; i < at + len; ++i) {
            if (h_docs_[i] < 0 || (i > at && h_docs_[i] < h_docs_[i - 1]))
                throw std::runtime_error("corrupt index file: a posting list is not in ascending doc order");
        }
        if (len > 0) max_doc = std::max<int64_t>(max_doc, h_docs_[at + len - 1]);
        h_list_ptr_[t + 1] = static_cast<int64_t>(h_docs_.size());
    }
    for (size_t t = n_terms; t < L; ++t) h_list_ptr_[t + 1] = h_list_ptr_[t];
    const size_t ns = r.scalar<size_t>();
    if (ns > r.remaining() / 4) throw std::runtime_error("index file is truncated");
    std::vector<float> maxima(ns);
    r.read(maxima.data(), ns * 4);
    h_max_ = std::move(maxima);
    n_lists_ = static_cast<int64_t>(n_terms);
    built_ = n_terms > 0;
    list_docs_ = max_doc + 1;
    host_valid_ = true;
    dev_valid_ = false;
}

// ============================================================== IDMapIndex

IDMapIndex::IDMapIndex(std::unique_ptr<Index> delegate) : Index(delegate->dimension()), delegate_(std::move(delegate)) {}

void IDMapIndex::add(idx_t n, const idx_t* indptr, const term_t* indices, const float* values) {
    delegate_->add(n, indptr, indices, values);
}

// IDMapIndex::add_with_ids (id_map_index.cpp:63-75).
void IDMapIndex::add_with_ids(idx_t n, const idx_t* indptr, const term_t* indices, const float* values,
                              const idx_t* ids) {
    const int64_t old = delegate_->num_vectors();
    delegate_->add(n, indptr, indices, values);
    if (ids == nullptr) throw std::invalid_argument("unexpected nullptr");
    i2e_.resize(old + n);
    for (idx_t i = 0; i < n; ++i) {
        i2e_[old + i] = ids[i];
        e2i_[ids[i]] = static_cast<idx_t>(old + i);
    }
}

void IDMapIndex::build() { delegate_->build(); }

// IDMapIndex::search (id_map_index.cpp:29-60): the selector is evaluated on
// external ids, the labels are mapped back to external ids.
void IDMapIndex::search(idx_t n, const idx_t* indptr, const term_t* indices, const float* values, int k,
                        float* distances, idx_t* labels, const nsparse_search_params* params) {
    if (dynamic_cast<BrutalIndex*>(delegate_.get()) != nullptr) {
        // the delegate ignores selectors; only the labels are mapped
        delegate_->search(n, indptr, indices, values, k, distances, labels, params);
        for (int64_t i = 0; i < static_cast<int64_t>(n) * k; ++i)
            if (labels[i] >= 0) labels[i] = i2e_.at(labels[i]);
        return;
    }
    if (auto* inv = dynamic_cast<InvertedIndex*>(delegate_.get())) {
        // the selector tests external ids: IDSelectorEnumerableWithIDMap /
        // is_member through the map (id_map_index.h:28-76)
        throw_if_not_positive(n);
        throw_if_not_positive(k);
        throw_if_any_null(indptr, indices, values, labels, distances);
        SelectorEval sel(request_from(params).selector);
        if (!inv->built()) return;
        std::vector<uint32_t> allow;
        if (sel.active()) {
            const int64_t nd = inv->list_docs();
            allow.assign(static_cast<size_t>((nd + 31) / 32), 0);
            for (int64_t d = 0; d < nd; ++d) {
                const int64_t ext = d < static_cast<int64_t>(i2e_.size()) ? i2e_[d] : d;
                if (sel.is_member(ext)) allow[d >> 5] |= 1U << (d & 31);
            }
        }
        inv->search_allowed(n, indptr, indices, values, k, distances, labels, sel.active() ? &allow : nullptr);
        for (int64_t i = 0; i < static_cast<int64_t>(n) * k; ++i)
            if (labels[i] >= 0) labels[i] = i2e_.at(labels[i]);
        return;
    }
    auto* seismic = dynamic_cast<SeismicIndex*>(delegate_.get());
    if (seismic == nullptr) throw std::runtime_error("IDMapIndex delegate is not a SEISMIC index");
    throw_if_not_positive(n);
    throw_if_not_positive(k);
    throw_if_any_null(indptr, indices, values, labels, distances);
    SearchRequest req = request_from(params);
    SelectorEval sel(req.selector);
    nsparse_selector mapped{};
    std::vector<idx_t> mapped_ids;
    std::vector<uint32_t> bitmap;
    if (sel.active()) {
        const int64_t nd = delegate_->num_vectors();
        if (sel.enumerable()) {
            // IDSelectorEnumerableWithIDMap: ids() keeps the external ids the map
            // knows, size() stays the delegate's (id_map_index.h:28-76).
            if (sel.size() <= static_cast<size_t>(k)) {
                std::vector<idx_t> ext = sel.ordered_ids();
                for (idx_t e : ext) {
                    auto it = e2i_.find(e);
                    if (it != e2i_.end()) mapped_ids.push_back(it->second);
                }
                // ordered_ids() of the wrapper sorts the mapped (internal) ids.
                std::sort(mapped_ids.begin(), mapped_ids.end());
                if (!seismic->vectors()) return;
                // The gate uses the delegate's size(), as the reference's wrapper does.
                seismic->exact_search(n, indptr, indices, values, k, distances, labels, req, mapped_ids);
                for (int64_t i = 0; i < static_cast<int64_t>(n) * k; ++i)
                    if (labels[i] >= 0) labels[i] = i2e_.at(labels[i]);
                return;
            }
        }
        bitmap.assign((nd + 31) / 32, 0);
        for (int64_t d = 0; d < nd; ++d) {
            const int64_t ext = d < static_cast<int64_t>(i2e_.size()) ? i2e_[d] : d;
            if (sel.is_member(ext)) bitmap[d >> 5] |= 1U << (d & 31);
        }
        mapped.kind = NSPARSE_SELECTOR_BITMAP;
        mapped.n = nd;
        mapped.bitmap = bitmap.data();
        req.selector = &mapped;
    }
    seismic->search_host(n, indptr, indices, values, k, distances, labels, req, nullptr);
    for (int64_t i = 0; i < static_cast<int64_t>(n) * k; ++i)
        if (labels[i] >= 0) labels[i] = i2e_.at(labels[i]);
}

// IDMapIndex::write_index / read_index (id_map_index.cpp:76-109).
void IDMapIndex::write_payload(ByteWriter& w) const {
    w.scalar<size_t>(i2e_.size());
    w.write(i2e_.data(), i2e_.size() * sizeof(idx_t));
    write_index(*delegate_, w);
}

void IDMapIndex::read_payload(ByteReader& r) {
    const size_t m = r.scalar<size_t>();
    if (m > r.remaining() / sizeof(idx_t)) throw std::runtime_error("index file is truncated");
    i2e_.resize(m);
    r.read(i2e_.data(), m * sizeof(idx_t));
    delegate_ = read_index(r);
    dim_ = delegate_->dimension();
    e2i_.clear();
    for (size_t i = 0; i < m; ++i) e2i_[i2e_[i]] = static_cast<idx_t>(i);
}

// ============================================================== factory + IO

namespace {
std::vector<std::string> split(const std::string& s, char d) {
    std::vector<std::string> out;
    size_t start = 0, end = s.find(d);
    while (end != std::string::npos) {
        out.push_back(s.substr(start, end - start));
        start = end + 1;
        end = s.find(d, start);
    }
    out.push_back(s.substr(start));
    return out;
}
std::string trim(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t\n\r");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\n\r");
    return s.substr(a, b - a + 1);
}
}  // namespace

// index_factory (index_factory.cpp:63-153): same grammar, same defaults, unknown
// keys ignored.
std::unique_ptr<Index> index_factory(int dim, const char* description) {
    if (description == nullptr || std::strlen(description) == 0)
        throw std::invalid_argument("Description cannot be null or empty");
    const std::string desc(description);
    const auto segments = split(desc, ',');
    const std::string type = trim(segments[0]);
    std::vector<std::pair<std::string, std::string>> params;
    if (segments.size() > 1)
        for (const auto& tok : split(segments[1], '|')) {
            const auto kv = split(tok, '=');
            if (kv.size() == 2) params.emplace_back(trim(kv[0]), trim(kv[1]));
        }
    auto get = [&](const std::string& key, const std::string& def) {
        for (const auto& [k, v] : params)
            if (k == key) return v;
        return def;
    };
    if (type == "brutal") return std::make_unique<BrutalIndex>(dim);
    if (type == "inverted") return std::make_unique<InvertedIndex>(dim);
    if (type == "seismic" || type == "seismic_sq") {
        SeismicIndex::Params p;
        bool sq = type == "seismic_sq";
        bool w16 = false;
        float vmin = 0.0F, vmax = 1.0F;
        if (sq) {
            const std::string qs = get("quantizer", "8bit");
            vmin = std::stof(get("vmin", "0.0"));
            vmax = std::stof(get("vmax", "1.0"));
            p.lambda = std::stoi(get("lambda", "10"));
            p.beta = std::stoi(get("beta", "5"));
            p.alpha = std::stof(get("alpha", "0.5"));
            p.seed = std::stoi(get("seed", "-1"));
            check_quantizer_range(vmin, vmax);
            w16 = qs == "16bit";  // any other value is 8-bit, as in the reference
        } else {
            p.lambda = std::stoi(get("lambda", "10"));
            p.beta = std::stoi(get("beta", "5"));
            p.alpha = std::stof(get("alpha", "0.5"));
            p.seed = std::stoi(get("seed", "-1"));
        }
        return std::make_unique<SeismicIndex>(dim, sq, vmin, vmax, p, w16);
    }
    if (type == "idmap") {
        if (segments.size() < 2) throw std::invalid_argument("idmap requires a delegate index type");
        std::string d;
        for (size_t i = 1; i < segments.size(); ++i) {
            if (i > 1) d += ",";
            d += segments[i];
        }
        return std::make_unique<IDMapIndex>(index_factory(dim, d.c_str()));
    }
    throw std::invalid_argument("Unknown index type: " + type);
}

// detail::write_index (io/index_io.cpp:88-95, 118-131): fourcc, dim, payload.
void write_index(const Index& index, ByteWriter& w) {
    w.scalar<uint32_t>(fourcc_of(index.id()));
    w.scalar<int>(index.dimension());
    index.write_payload(w);
}

// detail::read_index (io/index_io.cpp:97-116, 133-163).
std::unique_ptr<Index> read_index(ByteReader& r) {
    const uint32_t id = r.scalar<uint32_t>();
    const int dim = r.scalar<int>();
    std::unique_ptr<Index> ix;
    if (id == fourcc_of({'S', 'E', 'I', 'S'}))
        ix = std::make_unique<SeismicIndex>(dim, false, 0.0F, 1.0F, SeismicIndex::Params{-1, -1, 0.4F, -1});
    else if (id == fourcc_of({'S', 'E', 'S', 'Q'}))
        ix = std::make_unique<SeismicIndex>(dim, true, 0.0F, 1.0F, SeismicIndex::Params{-1, -1, 0.4F, -1});
    else if (id == fourcc_of({'I', 'D', 'M', 'P'}))
        ix = std::make_unique<IDMapIndex>();
    else if (id == fourcc_of({'B', 'R', 'U', 'T'}))
        ix = std::make_unique<BrutalIndex>(dim);  // whose read_payload throws, as the reference's read does
    else if (id == fourcc_of({'I', 'N', 'V', 'T'}))
        ix = std::make_unique<InvertedIndex>(dim);
    else
        throw std::runtime_error("Unknown index type");
    ix->read_payload(r);
    return ix;
}

}  // namespace nsparse_amd
