// Host side of the index: the reference's nsparse::Index family
// (index.h:30-68, seismic_index.h, seismic_scalar_quantized_index.h,
// id_map_index.h) re-built around an HBM-resident copy of the index.
#pragma once

#include <array>
#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "device_index.h"
#include "device_mem.h"
#include "host_stage.h"
#include "kernels.h"
#include "nsparse_c.h"

namespace nsparse_amd {

using idx_t = int32_t;
using term_t = uint16_t;

// Forward index on the host: the reference's SparseVectors CSR
// (sparse_vectors.h:105-110) with values as raw bytes of width esz.
struct HostCsr {
    size_t esz = 1;
    size_t dim = 0;
    std::vector<idx_t> indptr;
    std::vector<term_t> indices;
    std::vector<uint8_t> values;
    int64_t n() const { return indptr.empty() ? 0 : static_cast<int64_t>(indptr.size()) - 1; }
};

// One posting list's clusters + transposed summaries: InvertedListClusters
// (cluster/inverted_list_clusters.h:75-83).
struct HostList {
    std::vector<idx_t> docs, offsets;
    size_t n_clusters = 0;
    size_t esz = 4;
    std::vector<term_t> term_ids;
    std::vector<idx_t> term_ptr;
    std::vector<uint16_t> csc_cluster;
    std::vector<uint8_t> csc_value;
};

class ByteWriter;
class ByteReader;

struct SearchRequest {
    int cut = 10;
    float heap_factor = 1.0F;
    bool has_sq_range = false;
    float vmin = 0.0F, vmax = 1.0F;
    const nsparse_selector* selector = nullptr;
};

class Index {
public:
    explicit Index(int dim) : dim_(dim) {}
    virtual ~Index() = default;
    virtual std::array<char, 4> id() const = 0;
    int dimension() const { return dim_; }
    virtual int64_t num_vectors() const = 0;
    virtual void add(idx_t n, const idx_t* indptr, const term_t* indices, const float* values) = 0;
    virtual void add_with_ids(idx_t, const idx_t*, const term_t*, const float*, const idx_t*);
    virtual void build();
    // Index::search (index.cpp:32-49): validates, then fills n*k outputs.
    virtual void search(idx_t n, const idx_t* indptr, const term_t* indices, const float* values, int k,
                        float* distances, idx_t* labels, const nsparse_search_params* params) = 0;
    virtual void write_payload(ByteWriter& w) const = 0;
    virtual void read_payload(ByteReader& r) = 0;
    virtual const HostCsr* vectors() const { return nullptr; }

protected:
    int dim_;
};

// SeismicIndex (FP32) and SeismicScalarQuantizedIndex (8-bit or 16-bit codes)
// in one class: they differ only in element width, query encoding and score
// decoding.
class SeismicIndex : public Index {
public:
    struct Params {
        int lambda = 10, beta = 5;
        float alpha = 0.5F;
        int seed = -1;
    };
    // sq: scalar-quantized (SESQ); w16: its codes are 16-bit (QT_16bit)
    SeismicIndex(int dim, bool sq, float vmin, float vmax, Params p, bool w16 = false);
    std::array<char, 4> id() const override;
    int64_t num_vectors() const override { return fwd_.n(); }
    void add(idx_t n, const idx_t* indptr, const term_t* indices, const float* values) override;
    void build() override;
    void search(idx_t n, const idx_t* indptr, const term_t* indices, const float* values, int k, float* distances,
                idx_t* labels, const nsparse_search_params* params) override;
    void write_payload(ByteWriter& w) const override;
    void read_payload(ByteReader& r) override;
    const HostCsr* vectors() const override { return has_vectors_ ? &fwd_ : nullptr; }

    // Batched search over queries already in HBM (nsparse_search_device).
    // d_err: the launch's device error flag; nullptr uses the index's (which
    // last_device_error reads). Launches that may overlap pass flags of their own.
    // spread_cap >= 0 sets the largest launch that takes the spread search in
    // place of the default (a caller that knows the GPU's load picks it).
    void search_device(int64_t n, const int32_t* d_indptr, const uint16_t* d_indices, const float* d_values,
                       const int32_t* d_ks, int k_max, const SearchRequest& req, const int32_t* d_sel_of_query,
                       const uint32_t* const* d_bitmaps, int32_t n_bitmaps, float* d_out_d, int32_t* d_out_l,
                       int64_t* d_stats, hipStream_t stream, int32_t* d_err = nullptr, int64_t spread_cap = -1);
    // search_device without any device allocation: the index must already be
    // resident on the device (after build(), or after a first search of an
    // index read from a file), the error code goes to d_err, and the
    // schedule's scratch is the library's static device array.
    void search_device_noalloc(int64_t n, const int32_t* d_indptr, const uint16_t* d_indices, const float* d_values,
                               const int32_t* d_ks, int k_max, const SearchRequest& req,
                               const int32_t* d_sel_of_query, const uint32_t* const* d_bitmaps, int32_t n_bitmaps,
                               float* d_out_d, int32_t* d_out_l, int32_t* d_err, hipStream_t stream);
    // The kernel arguments of a batched search (validated), as search_device uses them.
    SearchArgs search_args(int64_t n, const int32_t* d_indptr, const uint16_t* d_indices, const float* d_values,
                           const int32_t* d_ks, int k_max, const SearchRequest& req, const int32_t* d_sel_of_query,
                           const uint32_t* const* d_bitmaps, int32_t n_bitmaps, float* d_out_d, int32_t* d_out_l,
                           int64_t* d_stats);
    // The search proper over host queries (IDMap passes its mapped selector).
    void search_host(idx_t n, const idx_t* indptr, const term_t* indices, const float* values, int k,
                     float* distances, idx_t* labels, const SearchRequest& req, int64_t* stats_out);
    // ExactMatcher path over an explicit sorted id list (IDMap passes mapped ids).
    void exact_search(idx_t n, const idx_t* indptr, const term_t* indices, const float* values, int k,
                      float* distances, idx_t* labels, const SearchRequest& req, const std::vector<idx_t>& ids);
    // Error code a kernel left from the last search_device call (0 = none).
    int32_t last_device_error();
    bool sq() const { return sq_; }
    // uint8 codes: the search kernels' code-table and block-row paths
    bool u8() const { return sq_ && !w16_; }
    const DevIndexView& view();
    int max_clusters() const { return max_clusters_; }
    // Counters of the last build() (zero for an index read from a file).
    const BuildStats& build_stats() const { return build_stats_; }
    void set_lists(std::vector<HostList>&& lists);
    const std::vector<HostList>& lists();

private:
    void ensure_device();
    void upload_lists();
    void download_lists();
    void refresh_block_rows();
    void refresh_summary_bitmaps();
    // Whether a launch of n queries runs as the spread search (search_spread.hip).
    bool use_spread(int64_t n, int k_max, int cut, bool stats, hipStream_t stream, int64_t cap = -1) const;
    // kMax8bit / kMax16bit of the quantizer (utils/scalar_quantizer.h): the
    // largest code, as float
    float code_max() const { return w16_ ? 65535.0F : 255.0F; }

    bool sq_;
    bool w16_ = false;
    float vmin_ = 0.0F, vmax_ = 1.0F;
    Params params_;
    HostCsr fwd_;
    bool has_vectors_ = false;
    // Lists: host copy and/or device copy; whichever is missing is made on demand.
    std::vector<HostList> lists_;
    bool host_lists_valid_ = true;
    bool dev_lists_valid_ = false;
    bool dev_fwd_valid_ = false;
    int32_t n_lists_ = 0;
    int max_clusters_ = 0;
    int64_t max_list_docs_ = 0;  // most docs of any device posting list
    BuildStats build_stats_{};
    // device copies
    DeviceBuffer d_indptr_, d_indices_, d_values_;
    DeviceBuffer d_list_cluster_base_, d_cluster_doc_ptr_, d_docs_, d_list_term_base_, d_sum_term_ids_,
        d_sum_term_ptr_, d_sum_cluster_, d_sum_value_, d_list_entry_base_;
    DeviceBuffer d_row_ref_, d_blk_rows_;                  // block-contiguous row copies (uint8)
    DeviceBuffer d_sum_bits_, d_sum_rank_;                  // per-list summary-term bitmaps
    DeviceBuffer d_err_;
    DevIndexView view_{};
    // Guards the lazy upload of host state, so concurrent searches on one
    // built index (which the reference allows) find it resident exactly once.
    std::mutex dev_mu_;
    // Staging for searches over host arrays, one stage per concurrent caller.
    HostStagePool stages_;
    // Scratch of spread searches launched by search_device.
    ScratchPool spread_scratch_;
};

// BrutalIndex (brutal_index.h:22-44, brutal_index.cpp): FP32 vectors kept as
// they are added, every query scored against every document on the device.
// Like the reference it has no build step (Index::build throws), ignores a
// selector, and does not serialize.
class BrutalIndex : public Index {
public:
    explicit BrutalIndex(int dim);
    std::array<char, 4> id() const override { return {'B', 'R', 'U', 'T'}; }
    int64_t num_vectors() const override { return fwd_.n(); }
    void add(idx_t n, const idx_t* indptr, const term_t* indices, const float* values) override;
    void search(idx_t n, const idx_t* indptr, const term_t* indices, const float* values, int k, float* distances,
                idx_t* labels, const nsparse_search_params* params) override;
    void write_payload(ByteWriter& w) const override;
    void read_payload(ByteReader& r) override;
    const HostCsr* vectors() const override { return fwd_.n() > 0 ? &fwd_ : nullptr; }

private:
    HostCsr fwd_;
    bool dev_valid_ = false;
    DeviceBuffer d_indptr_, d_indices_, d_values_;
    std::mutex dev_mu_;  // the lazy upload, as in SeismicIndex
    HostStagePool stages_;
};

// InvertedIndex (inverted_index.h:25-61, inverted_index.cpp): FP32 posting
// lists built on the device from the vectors added since the last build (which
// build then releases), searched exactly: the k best positive scores, ranked
// score desc then doc id asc -- what the reference's MaxScore traversal
// returns for non-negative weights. The lists' doc ids number that build's
// documents from 0, as in the reference; num_vectors counts every add.
class InvertedIndex : public Index {
public:
    explicit InvertedIndex(int dim);
    std::array<char, 4> id() const override { return {'I', 'N', 'V', 'T'}; }
    int64_t num_vectors() const override { return n_vectors_; }
    void add(idx_t n, const idx_t* indptr, const term_t* indices, const float* values) override;
    void build() override;
    void search(idx_t n, const idx_t* indptr, const term_t* indices, const float* values, int k, float* distances,
                idx_t* labels, const nsparse_search_params* params) override;
    void write_payload(ByteWriter& w) const override;
    void read_payload(ByteReader& r) override;
    // The search with the selector already resolved to a bitmap over the
    // lists' doc ids (nullptr: every doc); IDMapIndex passes its mapped one.
    void search_allowed(idx_t n, const idx_t* indptr, const term_t* indices, const float* values, int k,
                        float* distances, idx_t* labels, const std::vector<uint32_t>* allow);
    bool built() const { return built_; }
    // Doc ids the lists can hold: [0, list_docs()).
    int64_t list_docs() const { return list_docs_; }

private:
    void ensure_device();
    void ensure_host() const;

    HostCsr pending_;  // vectors_ of the reference
    int64_t n_vectors_ = 0;
    bool built_ = false;
    int64_t list_docs_ = 0;
    int64_t n_lists_ = 0;  // lists written: dim after a build, the file's count after a read
    // Lists: on the device after a build, on the host after a read; each side
    // is made from the other on demand.
    mutable bool host_valid_ = false;
    bool dev_valid_ = false;
    mutable std::vector<int64_t> h_list_ptr_;
    mutable std::vector<idx_t> h_docs_;
    mutable std::vector<float> h_vals_, h_max_;  // h_max_: max_term_scores_
    DevicePostings dev_;  // max_score is filled by a build only
    mutable std::mutex mu_;  // the lazy copies between host and device
    HostStagePool stages_;
};

// A batched search over fixed device buffers, replayed from a hipGraph: the
// schedule (lookup table, workgroup width, persistent queue, longest-first
// order) is given or picked by timing every candidate on the batch itself;
// each launch replays the captured kernels, so the caller may rewrite the query
// buffers (same sizes) between launches. Results equal search_device's.
class SearchPlan {
public:
    SearchPlan(SeismicIndex& index, const SearchArgs& args, int k_max, const SearchConfig* config, hipStream_t stream);
    ~SearchPlan();
    SearchPlan(const SearchPlan&) = delete;
    SearchPlan& operator=(const SearchPlan&) = delete;
    void launch(hipStream_t stream);
    const SearchConfig& config() const { return config_; }
    const std::vector<std::pair<SearchConfig, float>>& tuning() const { return tuning_; }

private:
    SearchArgs args_;
    int k_max_;
    int max_clusters_;
    bool sq_;
    SearchConfig config_{};
    std::vector<std::pair<SearchConfig, float>> tuning_;  // candidate -> best ms
    DeviceBuffer scratch_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

// IDMapIndex (id_map_index.h:78-109, id_map_index.cpp).
class IDMapIndex : public Index {
public:
    explicit IDMapIndex(std::unique_ptr<Index> delegate);
    IDMapIndex() : Index(0) {}
    std::array<char, 4> id() const override { return {'I', 'D', 'M', 'P'}; }
    int64_t num_vectors() const override { return delegate_ ? delegate_->num_vectors() : 0; }
    void add(idx_t n, const idx_t* indptr, const term_t* indices, const float* values) override;
    void add_with_ids(idx_t n, const idx_t* indptr, const term_t* indices, const float* values,
                      const idx_t* ids) override;
    void build() override;
    void search(idx_t n, const idx_t* indptr, const term_t* indices, const float* values, int k, float* distances,
                idx_t* labels, const nsparse_search_params* params) override;
    void write_payload(ByteWriter& w) const override;
    void read_payload(ByteReader& r) override;
    const HostCsr* vectors() const override { return delegate_ ? delegate_->vectors() : nullptr; }
    Index* delegate() { return delegate_.get(); }

private:
    std::unique_ptr<Index> delegate_;
    std::vector<idx_t> i2e_;
    std::unordered_map<idx_t, idx_t> e2i_;
};

std::unique_ptr<Index> index_factory(int dim, const char* description);
void write_index(const Index& index, ByteWriter& w);
std::unique_ptr<Index> read_index(ByteReader& r);

}  // namespace nsparse_amd
