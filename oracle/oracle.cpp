// ORACLE -- test infrastructure, not product code.
//
// A CPU restatement of neural-sparse-cpp's SEISMIC / SEISMIC-SQ build, search
// and index writer, used only as the checker: by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg. Nothing in the
// shipped library links or calls it.
//
// It follows the reference line by line (file:line cited per function) and uses
// the same libstdc++ routines the reference uses (std::ranges::sort,
// std::priority_queue, std::mt19937, std::uniform_int_distribution,
// std::lround), so with tie_mode == 0 it reproduces the reference's tie order
// as well. tie_mode == 1 is the deterministic variant the device path
// implements: exact float ties are broken by id (doc id in lambda pruning and
// in the result top-k, term id in summary pruning). Both modes take identical
// traversal decisions; they differ only where two values compare equal.
//
// Pinning (what each part is checked against) is listed in the header comment
// of tests/test_oracle.py. Parts pinned only by invariants there are marked
// "parity unpinned" in that file.
#include <algorithm>
#include <array>
#include <atomic>
#include <bit>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <numeric>
#include <queue>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#include <immintrin.h>
#include <omp.h>

namespace oracle {

using idx_t = int32_t;
using term_t = uint16_t;
constexpr idx_t kInvalid = -1;

// ---------------------------------------------------------------- quantizer
// utils/scalar_quantizer.h:32-44, 52-93, 97-126 (8-bit and 16-bit codes).
struct Quantizer {
    float vmin = 0.0F, vmax = 1.0F;
    int bits = 8;  // QuantizerType: QT_8bit (8) or QT_16bit (16)
    Quantizer() = default;
    Quantizer(float lo, float hi, int b = 8) : vmin(lo), vmax(hi), bits(b) {
        if (!std::isfinite(lo) || !std::isfinite(hi)) throw std::invalid_argument("vmin and vmax must be finite");
        if (hi <= lo) throw std::invalid_argument("vmax must be greater than vmin");
    }
    // kMax8bit / kMax16bit: numeric_limits<uint8_t / uint16_t>::max() as float
    float kmax() const { return bits == 16 ? 65535.0F : 255.0F; }
    size_t bytes() const { return bits == 16 ? 2 : 1; }
    // encode_8bit / encode_16bit (:97-101, :109-113)
    uint32_t encode(float val) const {
        const float kMax = kmax();
        float scaled = (val - vmin) * (kMax / (vmax - vmin));
        scaled = std::max(0.0F, std::min(kMax, scaled));
        return static_cast<uint32_t>(std::lround(scaled));
    }
    // codes of n values, bytes() each, little-endian
    void encode_n(const float* v, size_t n, uint8_t* out) const {
        for (size_t i = 0; i < n; ++i) {
            const uint32_t c = encode(v[i]);
            if (bits == 16) {
                const uint16_t c16 = static_cast<uint16_t>(c);
                std::memcpy(out + 2 * i, &c16, 2);
            } else {
                out[i] = static_cast<uint8_t>(c);
            }
        }
    }
    // decode_dot_product (:85-93): the ingest quantizer's width sets max_q
    float decode_dot_product(float s, const Quantizer& q) const {
        const float ingest_range = vmax - vmin;
        const float query_range = q.vmax - q.vmin;
        const float max_q = kmax();
        const float scale = (ingest_range * query_range) / (max_q * max_q);
        return s * scale;
    }
};

// ---------------------------------------------------------------- storage
// sparse_vectors.cpp:92-126 (add_vectors: copy + rebase offsets).
struct Fwd {
    size_t esz = 0, dim = 0;
    std::vector<idx_t> indptr;
    std::vector<term_t> indices;
    std::vector<uint8_t> values;
    size_t n() const { return indptr.empty() ? 0 : indptr.size() - 1; }
    void add(const idx_t* ip, size_t ip_size, const term_t* idx, size_t nnz, const uint8_t* w) {
        if (ip_size < 2) return;
        if (indptr.empty()) indptr.push_back(0);
        indices.insert(indices.end(), idx, idx + nnz);
        values.insert(values.end(), w, w + nnz * esz);
        const idx_t off = indptr.back();
        for (size_t i = 1; i < ip_size; ++i) indptr.push_back(ip[i] + off);
    }
};

// cluster/inverted_list_clusters.h:75-83.
struct ListClusters {
    std::vector<idx_t> docs, offsets;
    size_t n_clusters = 0, esz = 4;
    std::vector<term_t> term_ids;
    std::vector<idx_t> term_ptr;
    std::vector<uint16_t> csc_cluster;
    std::vector<uint8_t> csc_value;
};

struct Stats {  // traversal counters, per query
    int64_t lists, terms_found, csc_entries, clusters_tested, clusters_opened, docs_read, docs_scored,
        nnz_scored;
};

// ---------------------------------------------------------------- top-k
// utils/ranker.h:34-138 (TopKHolder with a std::priority_queue min-heap).
template <class T>
struct HeapTopK {
    struct Item {
        float score;
        T id;
    };
    struct Cmp {
        bool operator()(const Item& a, const Item& b) const { return std::greater<float>()(a.score, b.score); }
    };
    int k;
    std::priority_queue<Item, std::vector<Item>, Cmp> pq;
    float threshold = std::numeric_limits<float>::infinity();
    explicit HeapTopK(int kk) : k(kk) {
        if (k <= 0) throw std::invalid_argument("value must be positive");
        std::vector<Item> v;
        v.reserve(k);
        pq = std::priority_queue<Item, std::vector<Item>, Cmp>(Cmp(), std::move(v));
    }
    void add(float score, T id) {
        if (static_cast<int>(pq.size()) >= k && score <= threshold) return;
        if (static_cast<int>(pq.size()) < k) {
            pq.push({score, id});
            if (static_cast<int>(pq.size()) == k) threshold = pq.top().score;
        } else {
            pq.pop();
            pq.push({score, id});
            threshold = pq.top().score;
        }
    }
    bool full() const { return static_cast<int>(pq.size()) == k; }
    float peek() const { return pq.top().score; }
    // ranker.h:96-108
    std::vector<T> ids_descending() {
        std::vector<T> ret(pq.size());
        int i = static_cast<int>(pq.size()) - 1;
        while (!pq.empty() && i >= 0) {
            ret[i--] = pq.top().id;
            pq.pop();
        }
        return ret;
    }
    // ranker.h:110-124
    void items_descending(std::vector<float>& s, std::vector<T>& ids) {
        const size_t n = pq.size();
        s.assign(n, 0.0F);
        ids.assign(n, T{});
        for (int i = static_cast<int>(n) - 1; i >= 0; --i) {
            s[i] = pq.top().score;
            ids[i] = pq.top().id;
            pq.pop();
        }
    }
};

// Same contract with exact ties broken by id: keeps the k best of all distinct
// (score, id) pairs added under (score desc, id asc). Its kth score is always
// the heap's (the k largest scores form one multiset however ties are kept).
template <class T>
struct IdTopK {
    int k;
    std::vector<std::pair<float, T>> v;  // sorted best first
    explicit IdTopK(int kk) : k(kk) {
        if (k <= 0) throw std::invalid_argument("value must be positive");
    }
    static bool better(const std::pair<float, T>& a, const std::pair<float, T>& b) {
        return a.first > b.first || (a.first == b.first && a.second < b.second);
    }
    void add(float score, T id) {
        std::pair<float, T> it{score, id};
        if (static_cast<int>(v.size()) == k && !better(it, v.back())) return;
        auto pos = std::lower_bound(v.begin(), v.end(), it, better);
        v.insert(pos, it);
        if (static_cast<int>(v.size()) > k) v.pop_back();
    }
    bool full() const { return static_cast<int>(v.size()) == k; }
    float peek() const { return v.back().first; }
    void items_descending(std::vector<float>& s, std::vector<T>& ids) {
        s.clear();
        ids.clear();
        for (auto& p : v) {
            s.push_back(p.first);
            ids.push_back(p.second);
        }
    }
};

// utils/vector_process.h:18-30 (query cut). Ties follow the libstdc++ heap in
// both modes: the device path reproduces that heap exactly.
template <class W>
std::vector<term_t> top_k_tokens(const term_t* indices, const W* weights, int size, int k) {
    if (k >= size) return std::vector<term_t>(indices, indices + size);
    HeapTopK<term_t> h(k);
    for (int i = 0; i < size; ++i) h.add(static_cast<float>(weights[i]), indices[i]);
    return h.ids_descending();
}

// seismic_common.h:105-115 (first list: ranges::sort by summary score desc).
std::vector<size_t> reorder_clusters(const std::vector<float>& s, bool first_list) {
    std::vector<size_t> order(s.size());
    std::iota(order.begin(), order.end(), 0);
    if (first_list) std::ranges::sort(order, [&](size_t a, size_t b) { return s[a] > s[b]; });
    return order;
}

// ---------------------------------------------------------------- build
// seismic_common.h:130-144
int calculate_lambda(int lambda, size_t n) {
    if (lambda == -1) return std::max(static_cast<int>(0.0005F * static_cast<float>(n)), 160);
    return lambda;
}
int calculate_beta(int beta, int lambda) {
    if (beta == -1) return static_cast<int>(static_cast<float>(lambda) * 0.1F);
    return beta;
}

float value_as_float(const uint8_t* p, size_t esz) {
    if (esz == 4) {
        float f;
        std::memcpy(&f, p, 4);
        return f;
    }
    if (esz == 2) {
        uint16_t c;
        std::memcpy(&c, p, 2);
        return static_cast<float>(c);
    }
    return static_cast<float>(*p);
}

// invlists/inverted_lists.cpp:110-132
std::vector<idx_t> prune(const std::vector<idx_t>& docs, const std::vector<uint8_t>& codes, size_t esz,
                         size_t lambda, int tie_mode) {
    const size_t n = docs.size();
    if (lambda <= 0 || n == 0 || lambda >= n) return docs;
    std::vector<std::pair<float, idx_t>> p;
    p.reserve(n);
    for (size_t i = 0; i < n; ++i) p.emplace_back(value_as_float(codes.data() + i * esz, esz), docs[i]);
    auto cmp = [](const auto& a, const auto& b) { return a.first > b.first; };
    if (tie_mode == 0)
        std::ranges::sort(p, cmp);
    else
        std::stable_sort(p.begin(), p.end(), cmp);  // ties keep list order = doc id order
    std::vector<idx_t> kept;
    kept.reserve(lambda);
    for (size_t i = 0; i < lambda; ++i) kept.push_back(p[i].second);
    return kept;
}

// cluster/random_kmeans.cpp:28-63
std::vector<std::vector<idx_t>> select_centroids(std::vector<idx_t> docs, size_t n_clusters, uint32_t seed) {
    const size_t n = docs.size();
    if (n_clusters > n || n_clusters <= 0) return {};
    std::mt19937 gen(seed);
    std::vector<std::vector<idx_t>> clusters(n_clusters);
    for (size_t i = 0; i < n_clusters; i++) {
        std::uniform_int_distribution<> distrib(0, n - i - 1);
        size_t r = distrib(gen);
        clusters[i].push_back(docs[r]);
        std::swap(docs[r], docs[n - 1 - i]);
    }
    return clusters;
}
size_t boundary_check_n_clusters(size_t n_docs, size_t n_clusters) {
    if (n_clusters <= 0) n_clusters = static_cast<size_t>(std::sqrt(n_docs));
    n_clusters = n_clusters > n_docs ? n_docs : n_clusters;
    return std::max(static_cast<size_t>(1), n_clusters);
}

// cluster/kmeans_utils.cpp:51-158 (centroid CSC + argmax assignment).
template <class T, class Acc>
void map_docs_to_clusters(const Fwd& f, const std::vector<idx_t>& docs, std::vector<std::vector<idx_t>>& clusters) {
    if (clusters.empty() || docs.empty()) return;
    const idx_t* ip = f.indptr.data();
    const term_t* ix = f.indices.data();
    const T* vals = reinterpret_cast<const T*>(f.values.data());
    const size_t nc = clusters.size();
    size_t max_term = 0, nnz = 0;
    for (auto& c : clusters)
        for (idx_t j = ip[c.at(0)]; j < ip[c.at(0) + 1]; ++j) {
            max_term = std::max<size_t>(max_term, ix[j]);
            ++nnz;
        }
    const size_t n_cols = nnz == 0 ? 0 : max_term + 1;
    std::vector<idx_t> tp(n_cols + 1, 0);
    for (auto& c : clusters)
        for (idx_t j = ip[c.at(0)]; j < ip[c.at(0) + 1]; ++j) tp[ix[j] + 1]++;
    for (size_t t = 0; t < n_cols; ++t) tp[t + 1] += tp[t];
    std::vector<uint32_t> ccl(nnz);
    std::vector<T> cw(nnz);
    std::vector<idx_t> cur(tp.begin(), tp.end() - 1);
    for (size_t c = 0; c < nc; ++c) {
        const idx_t cen = clusters[c].at(0);
        for (idx_t j = ip[cen]; j < ip[cen + 1]; ++j) {
            const idx_t pos = cur[ix[j]]++;
            ccl[pos] = static_cast<uint32_t>(c);
            cw[pos] = vals[j];
        }
    }
    std::vector<idx_t> cids;
    for (auto& c : clusters) cids.push_back(c.at(0));
    std::ranges::sort(cids);
    std::vector<Acc> sims(nc, Acc(0));
    for (const idx_t d : docs) {
        if (std::ranges::binary_search(cids, d)) continue;
        std::ranges::fill(sims, Acc(0));
        for (idx_t j = ip[d]; j < ip[d + 1]; ++j) {
            const size_t t = ix[j];
            if (t >= n_cols) continue;
            const Acc dv = static_cast<Acc>(vals[j]);
            for (idx_t p = tp[t]; p < tp[t + 1]; ++p) sims[ccl[p]] += dv * static_cast<Acc>(cw[p]);
        }
        size_t best_c = 0;
        Acc best = sims[0];
        for (size_t c = 1; c < nc; ++c)
            if (sims[c] > best) {
                best = sims[c];
                best_c = c;
            }
        clusters[best_c].push_back(d);
    }
}

// cluster/inverted_list_clusters.cpp:37-65 (alpha-mass cut), 72-124 (max-pool).
template <class T>
void summarize(const Fwd& f, ListClusters& lc, float alpha, int tie_mode, std::vector<std::vector<term_t>>& s_terms,
               std::vector<std::vector<T>>& s_vals) {
    const size_t nc = lc.offsets.size() <= 1 ? 0 : lc.offsets.size() - 1;
    s_terms.assign(nc, {});
    s_vals.assign(nc, {});
    std::vector<T> acc(f.dim, T(0));
    std::vector<uint32_t> epoch(f.dim, 0);
    std::vector<term_t> touched;
    uint32_t cur = 0;
    const T* vals = reinterpret_cast<const T*>(f.values.data());
    for (size_t i = 0; i < nc; ++i) {
        ++cur;
        touched.clear();
        float sum = 0.0F;
        for (idx_t q = lc.offsets[i]; q < lc.offsets[i + 1]; ++q) {
            const idx_t d = lc.docs[q];
            for (idx_t j = f.indptr[d]; j < f.indptr[d + 1]; ++j) {
                const term_t t = f.indices[j];
                const T v = vals[j];
                if (epoch[t] != cur) {
                    epoch[t] = cur;
                    const T value = std::max(T(0), v);
                    acc[t] = value;
                    touched.push_back(t);
                    sum += value;
                } else if (v > acc[t]) {
                    sum += v - acc[t];
                    acc[t] = v;
                }
            }
        }
        std::vector<std::pair<term_t, T>> pairs;
        for (term_t t : touched) pairs.emplace_back(t, acc[t]);
        if (tie_mode == 0)
            std::ranges::sort(pairs, [](const auto& a, const auto& b) { return a.second > b.second; });
        else
            std::ranges::sort(pairs, [](const auto& a, const auto& b) {
                return a.second > b.second || (a.second == b.second && a.first < b.first);
            });
        float addup = 0.0F;
        for (size_t j = 0; j < pairs.size(); ++j) {
            addup += pairs[j].second;
            if (addup / sum >= alpha) {
                pairs.erase(pairs.begin() + j + 1, pairs.end());
                break;
            }
        }
        std::ranges::sort(pairs, [](const auto& a, const auto& b) { return a.first < b.first; });
        for (auto& [t, v] : pairs) {
            s_terms[i].push_back(t);
            s_vals[i].push_back(v);
        }
    }
}

// cluster/inverted_list_clusters.cpp:230-293 (term-major transpose).
template <class T>
void build_transpose(ListClusters& lc, const std::vector<std::vector<term_t>>& st, const std::vector<std::vector<T>>& sv) {
    lc.n_clusters = st.size();
    lc.term_ids.clear();
    lc.term_ptr.clear();
    lc.csc_cluster.clear();
    lc.csc_value.clear();
    if (lc.n_clusters == 0) return;
    std::vector<term_t> ids;
    for (auto& v : st) ids.insert(ids.end(), v.begin(), v.end());
    const size_t nnz = ids.size();
    std::ranges::sort(ids);
    ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
    auto col = [&](term_t t) { return static_cast<size_t>(std::lower_bound(ids.begin(), ids.end(), t) - ids.begin()); };
    std::vector<idx_t> tp(ids.size() + 1, 0);
    for (auto& v : st)
        for (term_t t : v) tp[col(t) + 1]++;
    for (size_t t = 0; t < ids.size(); ++t) tp[t + 1] += tp[t];
    std::vector<uint16_t> cc(nnz);
    std::vector<uint8_t> cv(nnz * sizeof(T));
    std::vector<idx_t> cursor(tp.begin(), tp.end() - 1);
    for (size_t c = 0; c < st.size(); ++c)
        for (size_t j = 0; j < st[c].size(); ++j) {
            const idx_t pos = cursor[col(st[c][j])]++;
            cc[pos] = static_cast<uint16_t>(c);
            std::memcpy(cv.data() + pos * sizeof(T), &sv[c][j], sizeof(T));
        }
    lc.term_ids = std::move(ids);
    lc.term_ptr = std::move(tp);
    lc.csc_cluster = std::move(cc);
    lc.csc_value = std::move(cv);
}

template <class T, class Acc>
ListClusters build_one_list(const Fwd& f, const std::vector<idx_t>& kept, int beta, uint32_t seed, float alpha,
                            int tie_mode) {
    ListClusters lc;
    lc.esz = f.esz;
    // RandomKMeans::train, random_kmeans.cpp:65-78
    std::vector<std::vector<idx_t>> clusters;
    if (!kept.empty()) {
        const size_t nc = boundary_check_n_clusters(kept.size(), static_cast<size_t>(beta));
        clusters = select_centroids(kept, nc, seed);
        map_docs_to_clusters<T, Acc>(f, kept, clusters);
    }
    // InvertedListClusters(docs), inverted_list_clusters.cpp:193-208
    if (!clusters.empty()) {
        lc.offsets.push_back(0);
        for (auto& c : clusters) {
            lc.docs.insert(lc.docs.end(), c.begin(), c.end());
            lc.offsets.push_back(static_cast<idx_t>(lc.docs.size()));
        }
    }
    std::vector<std::vector<term_t>> st;
    std::vector<std::vector<T>> sv;
    summarize<T>(f, lc, alpha, tie_mode, st, sv);
    build_transpose<T>(lc, st, sv);
    return lc;
}

// seismic_common.h:146-184
std::vector<ListClusters> build_lists(const Fwd& f, int lambda_p, int beta_p, float alpha, int seed, int tie_mode,
                                      int n_threads) {
    const size_t dim = f.dim;
    std::vector<std::vector<idx_t>> ldocs(dim);
    std::vector<std::vector<uint8_t>> lcodes(dim);
    for (size_t i = 0; i < f.n(); ++i)
        for (idx_t j = f.indptr[i]; j < f.indptr[i + 1]; ++j) {
            const term_t t = f.indices[j];
            if (t >= dim) throw std::invalid_argument("term_id out of range");
            ldocs[t].push_back(static_cast<idx_t>(i));
            lcodes[t].insert(lcodes[t].end(), f.values.begin() + static_cast<size_t>(j) * f.esz,
                             f.values.begin() + static_cast<size_t>(j + 1) * f.esz);
        }
    const int lambda = calculate_lambda(lambda_p, f.n());
    const int beta = calculate_beta(beta_p, lambda);
    const uint32_t base_seed = seed == -1 ? std::random_device{}() : static_cast<uint32_t>(seed);
    std::vector<ListClusters> out(dim);
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t idx = 0; idx < static_cast<int64_t>(dim); ++idx) {
        auto kept = prune(ldocs[idx], lcodes[idx], f.esz, static_cast<size_t>(lambda), tie_mode);
        const uint32_t s = base_seed + static_cast<uint32_t>(idx);
        if (f.esz == 4)
            out[idx] = build_one_list<float, float>(f, kept, beta, s, alpha, tie_mode);
        else if (f.esz == 2)
            out[idx] = build_one_list<uint16_t, int64_t>(f, kept, beta, s, alpha, tie_mode);
        else
            out[idx] = build_one_list<uint8_t, int64_t>(f, kept, beta, s, alpha, tie_mode);
    }
    return out;
}

// ---------------------------------------------------------------- search
struct Index {
    bool sq = false;
    Quantizer q;
    int lambda = 10, beta = 5, seed = -1;
    float alpha = 0.5F;
    Fwd fwd;
    std::vector<ListClusters> lists;
    int tie_mode = 0;
    bool has_vectors = false;
};

// InvertedListClusters::score_summaries_transposed, inverted_list_clusters.cpp:295-331
template <class QT>
void score_summaries(const ListClusters& lc, const term_t* qi, const QT* qv, size_t qn, std::vector<float>& out,
                     Stats& st) {
    out.assign(lc.n_clusters, 0.0F);
    if (lc.n_clusters == 0 || lc.term_ids.empty()) return;
    const QT* cv = reinterpret_cast<const QT*>(lc.csc_value.data());
    for (size_t i = 0; i < qn; ++i) {
        auto it = std::lower_bound(lc.term_ids.begin(), lc.term_ids.end(), qi[i]);
        if (it == lc.term_ids.end() || *it != qi[i]) continue;
        const size_t col = static_cast<size_t>(it - lc.term_ids.begin());
        const float v = static_cast<float>(qv[i]);
        st.terms_found++;
        st.csc_entries += lc.term_ptr[col + 1] - lc.term_ptr[col];
        for (idx_t j = lc.term_ptr[col]; j < lc.term_ptr[col + 1]; ++j)
            out[lc.csc_cluster[j]] += v * static_cast<float>(cv[j]);
    }
}

// id_selector.h:22-86; kind 4 is the IDMapIndex wrapper of id_map_index.h:27-75
// (IDSelectorWithIDMap / IDSelectorEnumerableWithIDMap): membership of an
// internal id is the inner selector's on its external id; an enumerable inner
// selector stays enumerable, with the inner size() and, as ids(), the internal
// ids of the external ids the map knows.
struct Selector {
    int kind = -1;  // -1 none, 0 array, 1 set, 2 not, 3 bitmap, 4 idmap
    std::vector<idx_t> ids;
    std::unordered_set<idx_t> set;
    const uint32_t* bitmap = nullptr;
    int64_t nbits = 0;
    const Selector* inner = nullptr;
    const std::vector<idx_t>* i2e = nullptr;                  // kind 4
    const std::unordered_map<idx_t, idx_t>* e2i = nullptr;    // kind 4
    bool is_member(idx_t id) const {
        switch (kind) {
            case 0: return std::find(ids.begin(), ids.end(), id) != ids.end();
            case 1: return set.count(id) != 0;
            case 2: return !inner->is_member(id);
            case 3: return id >= 0 && id < nbits && ((bitmap[id >> 5] >> (id & 31)) & 1U);
            case 4:
                // the reference indexes its map unchecked; a doc without an
                // external id is outside what it defines
                if (id < 0 || static_cast<size_t>(id) >= i2e->size())
                    throw std::out_of_range("doc without an external id under an IDMap selector");
                return inner->is_member((*i2e)[id]);
            default: return true;
        }
    }
    bool enumerable() const { return kind == 0 || kind == 1 || (kind == 4 && inner->enumerable()); }
    size_t size() const { return kind == 0 ? ids.size() : kind == 1 ? set.size() : inner->size(); }
    // IDSelectorEnumerable::ordered_ids (id_selector.h:33-37): ids(), sorted
    std::vector<idx_t> ordered_ids() const {
        std::vector<idx_t> out;
        if (kind == 0) {
            out = ids;
        } else if (kind == 1) {
            out.assign(set.begin(), set.end());
        } else if (kind == 4) {
            for (idx_t e : inner->ordered_ids()) {
                auto it = e2i->find(e);
                if (it != e2i->end()) out.push_back(it->second);
            }
        }
        std::ranges::sort(out);
        return out;
    }
};

template <class TopK>
void finish(TopK& h, int k, bool sq, const Quantizer& iq, const Quantizer& qq, float* d, idx_t* l) {
    std::vector<float> s;
    std::vector<idx_t> ids;
    h.items_descending(s, ids);
    if (sq)
        for (auto& x : s) x = iq.decode_dot_product(x, qq);
    for (int i = 0; i < k; ++i) {
        d[i] = i < static_cast<int>(s.size()) ? s[i] : -1.0F;
        l[i] = i < static_cast<int>(ids.size()) ? ids[i] : kInvalid;
    }
}

// The visited-doc set of one search thread: open addressing with linear
// probing over doc ids (>= 0; -1 marks a free slot), emptied between queries
// through the list of slots it filled. Plays the part of the reference's
// per-thread absl::flat_hash_set (seismic_scalar_quantized_index.cpp:279-281);
// only membership matters to the traversal.
class DocSet {
public:
    explicit DocSet(size_t expected) { rehash(std::max<size_t>(1024, std::bit_ceil(2 * expected))); }
    // true if d was not yet in the set
    bool insert(idx_t d) {
        if (2 * (used_.size() + 1) > slots_.size()) rehash(2 * slots_.size());
        uint32_t h = (static_cast<uint32_t>(d) * 0x9E3779B1u) >> shift_;
        while (true) {
            const idx_t v = slots_[h];
            if (v == d) return false;
            if (v < 0) {
                slots_[h] = d;
                used_.push_back(h);
                return true;
            }
            h = (h + 1) & mask_;
        }
    }
    void clear() {
        for (const uint32_t h : used_) slots_[h] = -1;
        used_.clear();
    }

private:
    void rehash(size_t cap) {
        std::vector<idx_t> keys;
        keys.reserve(used_.size());
        for (const uint32_t h : used_) keys.push_back(slots_[h]);
        slots_.assign(cap, -1);
        used_.clear();
        mask_ = static_cast<uint32_t>(cap - 1);
        shift_ = 32 - std::countr_zero(cap);
        for (const idx_t d : keys) insert(d);
    }
    std::vector<idx_t> slots_;
    std::vector<uint32_t> used_;
    uint32_t mask_ = 0;
    int shift_ = 0;
};

// uint8 dot product of a doc row with the dense query codes (integer sums:
// any order gives the generic loop's result, utils/distance.h:39-47). The
// AVX-512 form gathers 16 query codes at a time (4-byte loads at byte offsets,
// so `dense` carries 3 bytes of padding) and is taken when the CPU has it,
// as the reference's build picks its SIMD dot products (distance_simd.h:15-25).
__attribute__((target("avx512f,avx512bw"))) int dot_u8_avx512(const term_t* idx, const uint8_t* val, size_t n,
                                                              const uint8_t* dense) {
    __m512i acc = _mm512_setzero_si512();
    const __m512i low = _mm512_set1_epi32(0xFF);
    size_t j = 0;
    for (; j + 16 <= n; j += 16) {
        const __m512i ix = _mm512_cvtepu16_epi32(_mm256_loadu_si256(reinterpret_cast<const __m256i*>(idx + j)));
        const __m512i q = _mm512_and_si512(_mm512_i32gather_epi32(ix, dense, 1), low);
        const __m512i c = _mm512_cvtepu8_epi32(_mm_loadu_si128(reinterpret_cast<const __m128i*>(val + j)));
        acc = _mm512_add_epi32(acc, _mm512_mullo_epi32(q, c));
    }
    int r = _mm512_reduce_add_epi32(acc);
    for (; j < n; ++j) r += val[j] * dense[idx[j]];
    return r;
}

inline int dot_u8_generic(const term_t* idx, const uint8_t* val, size_t n, const uint8_t* dense) {
    int r = 0;
    for (size_t j = 0; j < n; ++j) r += val[j] * dense[idx[j]];
    return r;
}

// 1: take the AVX-512 dot when the CPU has it; 0: always the generic loop
std::atomic<int> g_simd{1};

inline bool cpu_has_avx512() {
    static const bool has = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw");
    return has && g_simd.load(std::memory_order_relaxed) != 0;
}

// The leading cache lines of a row's term ids and values, fetched one doc
// ahead of scoring (seismic_scalar_quantized_index.cpp:107-122: distance 1,
// 4 lines).
inline void prefetch_row_head(const Fwd& f, idx_t d) {
    constexpr size_t kLines = 4;
    const idx_t s0 = f.indptr[d];
    const size_t len = static_cast<size_t>(f.indptr[d + 1] - s0);
    const char* ip = reinterpret_cast<const char*>(f.indices.data() + s0);
    const char* vp = reinterpret_cast<const char*>(f.values.data() + static_cast<size_t>(s0) * f.esz);
    const size_t il = std::min(kLines, (len * sizeof(term_t) + 63) / 64);
    const size_t vl = std::min(kLines, (len * f.esz + 63) / 64);
    for (size_t l = 0; l < il; ++l) __builtin_prefetch(ip + 64 * l);
    for (size_t l = 0; l < vl; ++l) __builtin_prefetch(vp + 64 * l);
}

// seismic_scalar_quantized_index.cpp:67-135 / seismic_index.cpp:53-118 and
// single_query :307-357 / :231-274.
template <class TopK, class VT>
void single_query(const Index& ix, const std::vector<VT>& dense, const term_t* qi, const VT* qv, size_t qn,
                  const std::vector<term_t>& cuts, float hf, const Selector* sel, TopK& heap, Stats& st,
                  DocSet& visited) {
    visited.clear();
    const bool avx512 = cpu_has_avx512();
    std::vector<float> scores;
    bool first_list = true;
    const Fwd& f = ix.fwd;
    for (const term_t term : cuts) {
        if (term >= ix.lists.size()) continue;
        const ListClusters& lc = ix.lists[term];
        st.lists++;
        if (lc.n_clusters != 0) {
            score_summaries<VT>(lc, qi, qv, qn, scores, st);
            auto order = reorder_clusters(scores, first_list);
            for (size_t c : order) {
                st.clusters_tested++;
                if (heap.full() && (scores[c] * hf < heap.peek())) {
                    if (first_list) break;
                    continue;
                }
                st.clusters_opened++;
                const idx_t p1 = lc.offsets[c + 1];
                for (idx_t p = lc.offsets[c]; p < p1; ++p) {
                    const idx_t d = lc.docs[p];
                    if (p + 1 < p1) prefetch_row_head(f, lc.docs[p + 1]);
                    st.docs_read++;
                    if (!visited.insert(d)) continue;
                    if (sel != nullptr && !sel->is_member(d)) continue;
                    const idx_t s0 = f.indptr[d], s1 = f.indptr[d + 1];
                    st.docs_scored++;
                    st.nnz_scored += s1 - s0;
                    float score;
                    if constexpr (std::is_same_v<VT, float>) {
                        const float* w = reinterpret_cast<const float*>(f.values.data());
                        float r = 0.0F;
                        for (idx_t j = s0; j < s1; ++j) r += w[j] * dense[f.indices[j]];
                        score = r;
                    } else if constexpr (std::is_same_v<VT, uint16_t>) {
                        // dot_product_uint16_dense (utils/distance.h:49-57): int64 sum
                        const uint16_t* w = reinterpret_cast<const uint16_t*>(f.values.data());
                        int64_t r = 0;
                        for (idx_t j = s0; j < s1; ++j) r += static_cast<int64_t>(w[j]) * dense[f.indices[j]];
                        score = static_cast<float>(r);
                    } else {
                        const size_t n = static_cast<size_t>(s1 - s0);
                        const int r = avx512 ? dot_u8_avx512(f.indices.data() + s0, f.values.data() + s0, n,
                                                             dense.data())
                                             : dot_u8_generic(f.indices.data() + s0, f.values.data() + s0, n,
                                                              dense.data());
                        score = static_cast<float>(r);
                    }
                    heap.add(score, d);
                }
            }
        }
        first_list = false;
    }
}

// exact_matcher.h:23-41
template <class TopK>
void exact_query(const Index& ix, const Selector& sel, const std::vector<uint8_t>& qbytes_dense, int k, TopK& h) {
    const std::vector<idx_t> ids = sel.ordered_ids();
    const Fwd& f = ix.fwd;
    for (idx_t d : ids) {
        if (d < 0 || static_cast<size_t>(d) >= f.n()) throw std::invalid_argument("selector id out of range");
        float score;
        if (f.esz == 4) {
            const float* w = reinterpret_cast<const float*>(f.values.data());
            const float* dn = reinterpret_cast<const float*>(qbytes_dense.data());
            float r = 0.0F;
            for (idx_t j = f.indptr[d]; j < f.indptr[d + 1]; ++j) r += w[j] * dn[f.indices[j]];
            score = r;
        } else if (f.esz == 2) {
            const uint16_t* w = reinterpret_cast<const uint16_t*>(f.values.data());
            const uint16_t* dn = reinterpret_cast<const uint16_t*>(qbytes_dense.data());
            int64_t r = 0;
            for (idx_t j = f.indptr[d]; j < f.indptr[d + 1]; ++j) r += static_cast<int64_t>(w[j]) * dn[f.indices[j]];
            score = static_cast<float>(r);
        } else {
            int r = 0;
            for (idx_t j = f.indptr[d]; j < f.indptr[d + 1]; ++j) r += f.values[j] * qbytes_dense[f.indices[j]];
            score = static_cast<float>(r);
        }
        h.add(score, d);
    }
}

template <class TopK>
void search_impl(const Index& ix, int n, const idx_t* ip, const term_t* qi, const float* qv, int k, int cut, float hf,
                 const Quantizer& qq, const Selector* sel, float* D, idx_t* L, Stats* stats, int n_threads) {
    const size_t nnz = ip[n];
    const bool sq = ix.sq;
    const bool w16 = sq && ix.q.bits == 16;
    std::vector<uint8_t> codes;
    std::vector<uint16_t> codes16;
    if (w16) {
        codes16.resize(nnz);
        for (size_t i = 0; i < nnz; ++i) codes16[i] = static_cast<uint16_t>(qq.encode(qv[i]));
    } else if (sq) {
        codes.resize(nnz);
        for (size_t i = 0; i < nnz; ++i) codes[i] = static_cast<uint8_t>(qq.encode(qv[i]));
    }
    for (size_t i = 0; i < nnz; ++i)
        if (qi[i] >= ix.fwd.dim) throw std::invalid_argument("query term id out of range");
    // exact-match gate: seismic_common.h:117-128
    if (sel != nullptr && sel->enumerable() && sel->size() <= static_cast<size_t>(k)) {
        const size_t esz = ix.fwd.esz;
        for (int qidx = 0; qidx < n; ++qidx) {
            std::vector<uint8_t> dense(ix.fwd.dim * esz, 0);
            for (idx_t j = ip[qidx]; j < ip[qidx + 1]; ++j) {
                if (w16)
                    std::memcpy(dense.data() + static_cast<size_t>(qi[j]) * 2, &codes16[j], 2);
                else if (sq)
                    dense[qi[j]] = codes[j];
                else
                    std::memcpy(dense.data() + static_cast<size_t>(qi[j]) * 4, qv + j, 4);
            }
            TopK h(k);
            exact_query(ix, *sel, dense, k, h);
            std::vector<float> s;
            std::vector<idx_t> ids;
            h.items_descending(s, ids);
            for (int i = 0; i < k; ++i) {
                float v = -1.0F;
                idx_t l = kInvalid;
                if (i < static_cast<int>(s.size())) {
                    v = sq ? ix.q.decode_dot_product(s[i], qq) : s[i];
                    l = ids[i];
                }
                D[static_cast<size_t>(qidx) * k + i] = v;
                L[static_cast<size_t>(qidx) * k + i] = l;
            }
        }
        return;
    }
    if (cut <= 0) throw std::invalid_argument("cut must be positive");
    if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel
    {
        std::vector<float> densef;
        std::vector<uint8_t> denseq;
        std::vector<uint16_t> dense16;
        DocSet visited(static_cast<size_t>(std::max(k, 1)) * 4096);
        if (w16)
            dense16.assign(ix.fwd.dim, 0);
        else if (sq)
            denseq.assign(ix.fwd.dim + 3, 0);  // + the 4-byte gathers' padding
        else
            densef.assign(ix.fwd.dim, 0.0F);
#pragma omp for schedule(dynamic, 64)
        for (int qidx = 0; qidx < n; ++qidx) {
            const idx_t s0 = ip[qidx];
            const size_t len = ip[qidx + 1] - s0;
            Stats st{};
            TopK heap(k);
            if (w16) {
                const uint16_t* qc = codes16.data() + s0;
                auto cuts = top_k_tokens<uint16_t>(qi + s0, qc, static_cast<int>(len), cut);
                for (size_t i = 0; i < len; ++i) dense16[qi[s0 + i]] = qc[i];
                single_query<TopK, uint16_t>(ix, dense16, qi + s0, qc, len, cuts, hf, sel, heap, st, visited);
                for (size_t i = 0; i < len; ++i) dense16[qi[s0 + i]] = 0;
            } else if (sq) {
                const uint8_t* qc = codes.data() + s0;
                auto cuts = top_k_tokens<uint8_t>(qi + s0, qc, static_cast<int>(len), cut);
                for (size_t i = 0; i < len; ++i) denseq[qi[s0 + i]] = qc[i];
                single_query<TopK, uint8_t>(ix, denseq, qi + s0, qc, len, cuts, hf, sel, heap, st, visited);
                for (size_t i = 0; i < len; ++i) denseq[qi[s0 + i]] = 0;
            } else {
                const float* qf = qv + s0;
                auto cuts = top_k_tokens<float>(qi + s0, qf, static_cast<int>(len), cut);
                for (size_t i = 0; i < len; ++i) densef[qi[s0 + i]] = qf[i];
                single_query<TopK, float>(ix, densef, qi + s0, qf, len, cuts, hf, sel, heap, st, visited);
                for (size_t i = 0; i < len; ++i) densef[qi[s0 + i]] = 0.0F;
            }
            finish(heap, k, sq, ix.q, qq, D + static_cast<size_t>(qidx) * k, L + static_cast<size_t>(qidx) * k);
            if (stats != nullptr) stats[qidx] = st;
        }
    }
}

// ---------------------------------------------------------------- writer
// io/index_io.cpp:88-95, seismic_scalar_quantized_index.cpp:421-428,
// sparse_vectors.cpp:206-231, io/seismic_invlists_writer.h:36-43,
// cluster/inverted_list_clusters.cpp:333-363, io/align.h:43-88.
struct Writer {
    std::vector<uint8_t> buf;
    void raw(const void* p, size_t n) {
        const uint8_t* b = static_cast<const uint8_t*>(p);
        buf.insert(buf.end(), b, b + n);
    }
    template <class T>
    void scalar(T v) {
        raw(&v, sizeof(T));
    }
    void pad(size_t a) {
        const size_t n = (a - buf.size() % a) % a;
        buf.insert(buf.end(), n, 0);
    }
    template <class T>
    void padded(const T* p, size_t count, size_t align = alignof(T)) {
        pad(align);
        if (count > 0) raw(p, sizeof(T) * count);
    }
};

void write_index(const Index& ix, Writer& w) {
    const char* id = ix.sq ? "SESQ" : "SEIS";
    uint32_t fcc = static_cast<uint32_t>(id[0]) | static_cast<uint32_t>(id[1]) << 8 |
                   static_cast<uint32_t>(id[2]) << 16 | static_cast<uint32_t>(id[3]) << 24;
    w.scalar<uint32_t>(fcc);
    w.scalar<int>(static_cast<int>(ix.fwd.dim));
    if (ix.sq) {
        w.scalar<uint8_t>(ix.q.bits == 16 ? 1 : 0);  // QuantizerType: QT_8bit 0, QT_16bit 1
        w.scalar<float>(ix.q.vmin);
        w.scalar<float>(ix.q.vmax);
    }
    const size_t count = ix.has_vectors ? ix.fwd.n() : 0;
    w.scalar<size_t>(count);
    if (count > 0) {
        w.scalar<size_t>(ix.fwd.dim);
        w.scalar<size_t>(ix.fwd.esz);
        w.padded(ix.fwd.indptr.data(), count + 1);
        const size_t nnz = ix.fwd.indptr[count];
        w.padded(ix.fwd.indices.data(), nnz);
        w.padded(ix.fwd.values.data(), nnz * ix.fwd.esz, ix.fwd.esz);
    }
    w.scalar<size_t>(ix.lists.size());
    for (const auto& lc : ix.lists) {
        w.scalar<size_t>(lc.docs.size());
        w.padded(lc.docs.data(), lc.docs.size());
        w.scalar<size_t>(lc.offsets.size());
        w.padded(lc.offsets.data(), lc.offsets.size());
        w.scalar<size_t>(lc.n_clusters);
        w.scalar<size_t>(lc.esz);
        w.scalar<size_t>(lc.term_ids.size());
        if (!lc.term_ids.empty()) {
            w.padded(lc.term_ids.data(), lc.term_ids.size());
            w.padded(lc.term_ptr.data(), lc.term_ids.size() + 1);
        }
        w.scalar<size_t>(lc.csc_cluster.size());
        w.padded(lc.csc_cluster.data(), lc.csc_cluster.size());
        w.padded(lc.csc_value.data(), lc.csc_value.size(), lc.esz);
    }
}

thread_local std::string g_err;

template <class F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return 1;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 2;
    }
}

Selector make_selector(int kind, int64_t n, const idx_t* ids, const uint32_t* bitmap, int negate, Selector& inner) {
    Selector s;
    if (kind < 0) return s;
    Selector base;
    base.kind = kind;
    if (kind == 0) base.ids.assign(ids, ids + n);
    if (kind == 1) base.set.insert(ids, ids + n);
    if (kind == 3) {
        base.bitmap = bitmap;
        base.nbits = n;
    }
    if (!negate) return base;
    inner = std::move(base);
    s.kind = 2;
    s.inner = &inner;
    return s;
}

// ---------------------------------------------------------------- InvertedIndex
// inverted_index.cpp: FP32 posting lists in doc order and a MaxScore traversal
// over windows of 4096 doc ids. Each step follows the reference, so scores are
// summed in its order (essential terms into the window in ascending bound
// order, then the non-essential terms from the highest bound down) and the
// pruning tests see the same float values.
struct InvIndex {
    int dim = 0;
    int tie_mode = 1;
    int64_t n_vectors = 0;  // num_vectors_ (inverted_index.cpp:233)
    Fwd pending;            // vectors_, released by build (:239)
    bool built = false;     // inverted_lists_ != nullptr
    std::vector<std::vector<idx_t>> docs;
    std::vector<std::vector<float>> vals;
    std::vector<float> max_score;  // max_term_scores_
};

// InvertedIndex::add (inverted_index.cpp:218-234).
void inv_add(InvIndex& ix, int n, const idx_t* indptr, const term_t* indices, const float* values) {
    if (n <= 0) throw std::invalid_argument("value must be positive");
    if (indptr == nullptr || indices == nullptr || values == nullptr) throw std::invalid_argument("unexpected nullptr");
    if (ix.pending.indptr.empty()) {
        ix.pending.esz = 4;
        ix.pending.dim = static_cast<size_t>(ix.dim);
    }
    ix.pending.add(indptr, static_cast<size_t>(n) + 1, indices, static_cast<size_t>(indptr[n]),
                   reinterpret_cast<const uint8_t*>(values));
    ix.n_vectors += n;
}

// ArrayInvertedLists::build_inverted_lists (invlists/inverted_lists.cpp:159-180)
// and InvertedIndex::build (inverted_index.cpp:236-252): postings appended doc
// by doc in row order, then the per-term maximum starting from 0
// (InvertedList::max_value, inverted_lists.cpp:88-96).
void inv_build(InvIndex& ix) {
    // The reference dereferences its released vectors here; the oracle refuses.
    if (ix.pending.indptr.empty()) throw std::runtime_error("no vectors added since the last build");
    const Fwd& f = ix.pending;
    ix.docs.assign(ix.dim, {});
    ix.vals.assign(ix.dim, {});
    for (size_t d = 0; d < f.n(); ++d)
        for (idx_t j = f.indptr[d]; j < f.indptr[d + 1]; ++j) {
            const term_t t = f.indices[j];
            if (t >= ix.dim) throw std::invalid_argument("term_id out of range");
            float v;
            std::memcpy(&v, f.values.data() + static_cast<size_t>(j) * 4, 4);
            ix.docs[t].push_back(static_cast<idx_t>(d));
            ix.vals[t].push_back(v);
        }
    ix.max_score.assign(ix.dim, 0.0F);
    for (int t = 0; t < ix.dim; ++t)
        for (float v : ix.vals[t])
            if (v > ix.max_score[t]) ix.max_score[t] = v;
    ix.pending = Fwd{};
    ix.built = true;
}

constexpr idx_t kNoDoc = std::numeric_limits<idx_t>::max();
constexpr int kInvWindow = 1 << 12;  // kScoreWindowSize (inverted_index.cpp:31)

// DirectTermScorer (inverted_index.cpp:38-86): a cursor over one posting list.
struct TermCursor {
    const idx_t* docs = nullptr;
    const float* vals = nullptr;
    int size = 0;
    int pos = 0;
    float weight = 0.0F;
    float bound = 0.0F;  // weight * max posting value
    idx_t doc() const { return pos < 0 || pos >= size ? kNoDoc : docs[pos]; }
    // advance_to: first posting >= target at or after the cursor.
    void seek(idx_t target) {
        if (pos >= size) return;
        if (pos >= 0 && docs[pos] >= target) return;
        const int from = pos < 0 ? 0 : pos + 1;
        pos = static_cast<int>(std::lower_bound(docs + from, docs + size, target) - docs);
    }
};

// InvertedIndex::single_query (inverted_index.cpp:295-365) with build_scorers
// (:88-122), score_essential_terms (:124-150) and evaluate_window_candidates
// (:152-212).
template <class TopK>
void inv_query(const InvIndex& ix, const term_t* qi, const float* qv, int size, int k, const Selector* sel, float* D,
               idx_t* L) {
    for (int i = 0; i < k; ++i) {
        D[i] = -1.0F;
        L[i] = kInvalid;
    }
    if (size == 0) return;
    TopK heap(k);
    std::vector<TermCursor> cur(size);
    for (int i = 0; i < size; ++i) {
        const term_t t = qi[i];
        cur[i].docs = ix.docs[t].data();
        cur[i].vals = ix.vals[t].data();
        cur[i].size = static_cast<int>(ix.docs[t].size());
        cur[i].weight = qv[i];
        cur[i].bound = qv[i] * ix.max_score[t];
    }
    std::sort(cur.begin(), cur.end(), [](const TermCursor& a, const TermCursor& b) { return a.bound < b.bound; });
    std::vector<float> prefix(size + 1, 0.0F);
    for (int i = 0; i < size; ++i) prefix[i + 1] = prefix[i] + cur[i].bound;
    std::vector<float> acc(kInvWindow, 0.0F);
    std::array<uint64_t, kInvWindow / 64> touched{};
    int first = 0;  // first essential cursor
    for (;;) {
        if (heap.full()) {
            const float thr = heap.peek();
            first = 0;
            for (int i = 0; i < size; ++i) {
                if (prefix[i + 1] < thr)
                    first = i + 1;
                else
                    break;
            }
        }
        idx_t low = kNoDoc;
        for (int i = first; i < size; ++i) low = std::min(low, cur[i].doc());
        if (low == kNoDoc) break;
        const idx_t base = low / kInvWindow * kInvWindow, end = base + kInvWindow;
        touched.fill(0);
        for (int i = first; i < size; ++i) {
            TermCursor& c = cur[i];
            while (c.pos < c.size) {
                const idx_t d = c.docs[c.pos];
                if (d >= end) break;
                if (d >= base) {
                    const int slot = d - base;
                    touched[slot >> 6] |= uint64_t{1} << (slot & 63);
                    acc[slot] += c.weight * c.vals[c.pos];
                }
                ++c.pos;
            }
        }
        float thr = heap.full() ? heap.peek() : 0.0F;
        const float rest = prefix[first];
        for (int w = 0; w < kInvWindow / 64; ++w) {
            for (uint64_t bits = touched[w]; bits != 0; bits &= bits - 1) {
                const int slot = w * 64 + std::countr_zero(bits);
                const float part = acc[slot];
                acc[slot] = 0.0F;
                const idx_t d = base + slot;
                if (part + rest <= thr) continue;
                if (sel != nullptr && !sel->is_member(d)) continue;
                float total = part;
                for (int j = first - 1; j >= 0; --j) {
                    TermCursor& c = cur[j];
                    c.seek(d);
                    if (c.doc() == d) total += c.weight * c.vals[c.pos];
                    if (total + prefix[j] <= thr) break;
                }
                heap.add(total, d);
                if (heap.full()) thr = heap.peek();
            }
        }
    }
    std::vector<float> s;
    std::vector<idx_t> ids;
    heap.items_descending(s, ids);
    for (size_t i = 0; i < s.size() && i < static_cast<size_t>(k); ++i) {
        D[i] = s[i];
        L[i] = ids[i];
    }
}

// InvertedIndex::write_index (inverted_index.cpp:367-402) after the header
// (io/index_io.cpp:88-95): plain field writes, no alignment padding.
void inv_write(const InvIndex& ix, Writer& w) {
    w.scalar<uint32_t>(static_cast<uint32_t>('I') | static_cast<uint32_t>('N') << 8 |
                       static_cast<uint32_t>('V') << 16 | static_cast<uint32_t>('T') << 24);
    w.scalar<int>(ix.dim);
    w.scalar<size_t>(static_cast<size_t>(ix.n_vectors));
    const size_t n_terms = ix.built ? ix.docs.size() : 0;
    w.scalar<size_t>(n_terms);
    w.scalar<size_t>(4);
    for (size_t t = 0; t < n_terms; ++t) {
        w.scalar<size_t>(ix.docs[t].size());
        if (!ix.docs[t].empty()) {
            w.raw(ix.docs[t].data(), ix.docs[t].size() * 4);
            w.raw(ix.vals[t].data(), ix.vals[t].size() * 4);
        }
    }
    const size_t ns = ix.built ? ix.max_score.size() : 0;
    w.scalar<size_t>(ns);
    if (ns > 0) w.raw(ix.max_score.data(), ns * 4);
}

}  // namespace oracle

using namespace oracle;

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

// Selects the uint8 dot product form of later searches (1 = AVX-512 when
// available, 0 = generic); returns whether the CPU has AVX-512.
int oracle_set_simd(int on) {
    g_simd.store(on != 0 ? 1 : 0);
    return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") ? 1 : 0;
}

// sq: 0 = seismic (FP32), 1 = seismic_sq 8-bit, 2 = seismic_sq 16-bit.
int oracle_create(int dim, int sq, float vmin, float vmax, int lambda, int beta, float alpha, int seed, int tie_mode,
                  void** out) {
    return guard([&] {
        auto* ix = new Index;
        ix->sq = sq != 0;
        if (ix->sq) ix->q = Quantizer(vmin, vmax, sq == 2 ? 16 : 8);
        ix->lambda = lambda;
        ix->beta = beta;
        ix->alpha = alpha;
        ix->seed = seed;
        ix->tie_mode = tie_mode;
        ix->fwd.esz = ix->sq ? ix->q.bytes() : 4;
        ix->fwd.dim = static_cast<size_t>(dim);
        *out = ix;
    });
}

void oracle_free(void* h) { delete static_cast<Index*>(h); }

int oracle_add(void* h, int n, const idx_t* indptr, const term_t* indices, const float* values) {
    return guard([&] {
        auto* ix = static_cast<Index*>(h);
        if (n <= 0) throw std::invalid_argument("value must be positive");
        const size_t nnz = indptr[n];
        if (ix->sq) {
            std::vector<uint8_t> codes(nnz * ix->q.bytes());
            ix->q.encode_n(values, nnz, codes.data());
            ix->fwd.add(indptr, n + 1, indices, nnz, codes.data());
        } else {
            ix->fwd.add(indptr, n + 1, indices, nnz, reinterpret_cast<const uint8_t*>(values));
        }
        ix->has_vectors = ix->fwd.n() > 0;
    });
}

int oracle_build(void* h, int n_threads) {
    return guard([&] {
        auto* ix = static_cast<Index*>(h);
        ix->lists = build_lists(ix->fwd, ix->lambda, ix->beta, ix->alpha, ix->seed, ix->tie_mode, n_threads);
    });
}

// sel_kind: -1 none, 0 array, 1 set, 3 bitmap; sel_negate wraps it in NotIDSelector.
// qsq: 1 => SeismicSQSearchParameters(qvmin, qvmax, ...), else the index range.
int oracle_search(void* h, int n, const idx_t* indptr, const term_t* indices, const float* values, int k, int cut,
                  float hf, int qsq, float qvmin, float qvmax, int sel_kind, int64_t sel_n, const idx_t* sel_ids,
                  const uint32_t* sel_bitmap, int sel_negate, float* D, idx_t* L, int64_t* stats, int n_threads) {
    return guard([&] {
        auto* ix = static_cast<Index*>(h);
        if (n <= 0 || k <= 0) throw std::invalid_argument("value must be positive");
        if (!ix->has_vectors) return;
        Quantizer qq = ix->q;
        if (ix->sq && qsq) qq = Quantizer(qvmin, qvmax, ix->q.bits);
        Selector inner;
        Selector sel = make_selector(sel_kind, sel_n, sel_ids, sel_bitmap, sel_negate, inner);
        const Selector* sp = sel.kind < 0 ? nullptr : &sel;
        auto* st = reinterpret_cast<Stats*>(stats);
        if (ix->tie_mode == 0)
            search_impl<HeapTopK<idx_t>>(*ix, n, indptr, indices, values, k, cut, hf, qq, sp, D, L, st, n_threads);
        else
            search_impl<IdTopK<idx_t>>(*ix, n, indptr, indices, values, k, cut, hf, qq, sp, D, L, st, n_threads);
    });
}

// IDMapIndex::search (id_map_index.cpp:29-60) over this index as the
// delegate: i2e[internal id] = external id (add_with_ids order,
// id_map_index.cpp:63-75; a repeated external id maps to its last doc). The
// selector, given on external ids, is wrapped as the reference wraps it, and
// every label >= 0 is mapped to its external id.
int oracle_search_idmap(void* h, const idx_t* i2e, int64_t n_map, int n, const idx_t* indptr, const term_t* indices,
                        const float* values, int k, int cut, float hf, int qsq, float qvmin, float qvmax,
                        int sel_kind, int64_t sel_n, const idx_t* sel_ids, const uint32_t* sel_bitmap,
                        int sel_negate, float* D, idx_t* L, int n_threads) {
    return guard([&] {
        auto* ix = static_cast<Index*>(h);
        if (n <= 0 || k <= 0) throw std::invalid_argument("value must be positive");
        if (!ix->has_vectors) return;
        Quantizer qq = ix->q;
        if (ix->sq && qsq) qq = Quantizer(qvmin, qvmax, ix->q.bits);
        const std::vector<idx_t> map(i2e, i2e + n_map);
        std::unordered_map<idx_t, idx_t> rev;
        for (int64_t i = 0; i < n_map; ++i) rev[map[i]] = static_cast<idx_t>(i);
        Selector inner_not;
        Selector user = make_selector(sel_kind, sel_n, sel_ids, sel_bitmap, sel_negate, inner_not);
        Selector wrap;
        const Selector* sp = nullptr;
        if (user.kind >= 0) {
            wrap.kind = 4;
            wrap.inner = &user;
            wrap.i2e = &map;
            wrap.e2i = &rev;
            sp = &wrap;
        }
        if (ix->tie_mode == 0)
            search_impl<HeapTopK<idx_t>>(*ix, n, indptr, indices, values, k, cut, hf, qq, sp, D, L, nullptr,
                                         n_threads);
        else
            search_impl<IdTopK<idx_t>>(*ix, n, indptr, indices, values, k, cut, hf, qq, sp, D, L, nullptr,
                                       n_threads);
        for (int64_t i = 0; i < static_cast<int64_t>(n) * k; ++i) {
            if (L[i] < 0) continue;
            if (static_cast<size_t>(L[i]) >= map.size()) throw std::out_of_range("result doc without an external id");
            L[i] = map[static_cast<size_t>(L[i])];
        }
    });
}

// Summary scores of the oracle's lists, as the search computes them
// (score_summaries, restating inverted_list_clusters.cpp:295-331): query q, all
// its entries in CSR order (SQ: codes of the index's quantizer), against list
// lists[q]. out holds cap floats per query; out_n[q] receives the cluster count.
int oracle_score_summaries(void* h, int nq, const idx_t* q_indptr, const term_t* q_indices, const float* q_values,
                           const int32_t* lists, int cap, float* out, int32_t* out_n) {
    return guard([&] {
        auto* ix = static_cast<Index*>(h);
        std::vector<float> scores;
        for (int qi = 0; qi < nq; ++qi) {
            const idx_t s0 = q_indptr[qi];
            const size_t len = static_cast<size_t>(q_indptr[qi + 1] - s0);
            const ListClusters& lc = ix->lists.at(static_cast<size_t>(lists[qi]));
            Stats st{};
            if (ix->sq && ix->q.bits == 16) {
                std::vector<uint16_t> codes(len);
                for (size_t i = 0; i < len; ++i) codes[i] = static_cast<uint16_t>(ix->q.encode(q_values[s0 + i]));
                score_summaries<uint16_t>(lc, q_indices + s0, codes.data(), len, scores, st);
            } else if (ix->sq) {
                std::vector<uint8_t> codes(len);
                for (size_t i = 0; i < len; ++i) codes[i] = static_cast<uint8_t>(ix->q.encode(q_values[s0 + i]));
                score_summaries<uint8_t>(lc, q_indices + s0, codes.data(), len, scores, st);
            } else {
                score_summaries<float>(lc, q_indices + s0, q_values + s0, len, scores, st);
            }
            out_n[qi] = static_cast<int32_t>(scores.size());
            for (size_t c = 0; c < scores.size() && c < static_cast<size_t>(cap); ++c)
                out[static_cast<size_t>(qi) * cap + c] = scores[c];
        }
    });
}

// Replaces the index's posting lists with hand-made ones (known-answer tests):
// list l has nc[l] clusters; cluster c (numbered across all lists) holds
// cdocs[c] doc ids from `docs` and a summary of cterms[c] (term, value) pairs
// from sum_terms / sum_vals (terms ascending; SQ values are codes 0..255). The
// term-major transpose is built as build_transpose does.
int oracle_set_lists(void* h, int n_lists, const int32_t* nc, const int32_t* cdocs, const idx_t* docs,
                     const int32_t* cterms, const term_t* sum_terms, const float* sum_vals) {
    return guard([&] {
        auto* ix = static_cast<Index*>(h);
        std::vector<ListClusters> out(static_cast<size_t>(n_lists));
        size_t c = 0, d = 0, t = 0;
        for (int l = 0; l < n_lists; ++l) {
            ListClusters& lc = out[l];
            lc.esz = ix->fwd.esz;
            std::vector<std::vector<term_t>> st;
            std::vector<std::vector<float>> svf;
            std::vector<std::vector<uint8_t>> svq;
            std::vector<std::vector<uint16_t>> svw;
            if (nc[l] > 0) lc.offsets.push_back(0);
            for (int j = 0; j < nc[l]; ++j, ++c) {
                lc.docs.insert(lc.docs.end(), docs + d, docs + d + cdocs[c]);
                d += static_cast<size_t>(cdocs[c]);
                lc.offsets.push_back(static_cast<idx_t>(lc.docs.size()));
                st.emplace_back(sum_terms + t, sum_terms + t + cterms[c]);
                svf.emplace_back(sum_vals + t, sum_vals + t + cterms[c]);
                svq.emplace_back();
                svw.emplace_back();
                for (int e = 0; e < cterms[c]; ++e) {
                    svq.back().push_back(static_cast<uint8_t>(sum_vals[t + e]));
                    svw.back().push_back(static_cast<uint16_t>(sum_vals[t + e]));
                }
                t += static_cast<size_t>(cterms[c]);
            }
            if (ix->sq && ix->q.bits == 16)
                build_transpose<uint16_t>(lc, st, svw);
            else if (ix->sq)
                build_transpose<uint8_t>(lc, st, svq);
            else
                build_transpose<float>(lc, st, svf);
        }
        ix->lists = std::move(out);
    });
}

// Reads a write_index byte stream (SEIS / SESQ) into an oracle index, the
// inverse of write_index above (sparse_vectors.cpp:233-253,
// inverted_list_clusters.cpp:365-389). The stream is trusted: it comes from the
// library under test or from this file's writer.
int oracle_load(const uint8_t* buf, size_t size, int tie_mode, void** out) {
    return guard([&] {
        size_t pos = 0;
        auto rd = [&](void* dst, size_t n) {
            if (n > size - pos) throw std::runtime_error("truncated index stream");
            std::memcpy(dst, buf + pos, n);
            pos += n;
        };
        auto scalar = [&](auto& v) { rd(&v, sizeof(v)); };
        auto skip = [&](size_t a) { pos += (a - pos % a) % a; };
        auto arr = [&](auto& vec, size_t n, size_t a) {
            skip(a);
            vec.resize(n);
            rd(vec.data(), n * sizeof(vec[0]));
        };
        uint32_t fcc = 0;
        int dim = 0;
        scalar(fcc);
        scalar(dim);
        auto* ix = new Index;
        std::unique_ptr<Index> hold(ix);
        ix->tie_mode = tie_mode;
        ix->sq = fcc == (uint32_t('S') | uint32_t('E') << 8 | uint32_t('S') << 16 | uint32_t('Q') << 24);
        if (ix->sq) {
            uint8_t qt = 0;
            float lo = 0, hi = 1;
            scalar(qt);
            scalar(lo);
            scalar(hi);
            if (qt > 1) throw std::runtime_error("index file declares an unknown quantizer type");
            ix->q = Quantizer(lo, hi, qt == 1 ? 16 : 8);
        }
        ix->fwd.esz = ix->sq ? ix->q.bytes() : 4;
        ix->fwd.dim = static_cast<size_t>(dim);
        size_t count = 0;
        scalar(count);
        if (count > 0) {
            size_t fdim = 0, esz = 0;
            scalar(fdim);
            scalar(esz);
            arr(ix->fwd.indptr, count + 1, 4);
            const size_t nnz = static_cast<size_t>(ix->fwd.indptr[count]);
            arr(ix->fwd.indices, nnz, 2);
            arr(ix->fwd.values, nnz * esz, esz);
        }
        ix->has_vectors = count > 0;
        size_t n_lists = 0;
        scalar(n_lists);
        ix->lists.resize(n_lists);
        for (auto& l : ix->lists) {
            size_t n = 0;
            scalar(n);
            arr(l.docs, n, 4);
            scalar(n);
            arr(l.offsets, n, 4);
            scalar(l.n_clusters);
            scalar(l.esz);
            size_t nt = 0;
            scalar(nt);
            if (nt > 0) {
                arr(l.term_ids, nt, 2);
                arr(l.term_ptr, nt + 1, 4);
            }
            scalar(n);
            arr(l.csc_cluster, n, 2);
            arr(l.csc_value, n * l.esz, l.esz);
        }
        *out = hold.release();
    });
}

// Copies the forward index and posting lists of `src` into `dst` (same kind),
// so one build can be searched under both tie modes.
int oracle_copy_index(void* dst, const void* src) {
    return guard([&] {
        auto* d = static_cast<Index*>(dst);
        const auto* s = static_cast<const Index*>(src);
        d->fwd = s->fwd;
        d->lists = s->lists;
        d->has_vectors = s->has_vectors;
        d->q = s->q;
    });
}

int oracle_serialize(void* h, uint8_t** out, size_t* size) {
    return guard([&] {
        Writer w;
        write_index(*static_cast<Index*>(h), w);
        *size = w.buf.size();
        *out = static_cast<uint8_t*>(std::malloc(w.buf.size()));
        std::memcpy(*out, w.buf.data(), w.buf.size());
    });
}

void oracle_buffer_free(uint8_t* p) { std::free(p); }

// ---- component hooks for the pinning tests ----

int oracle_top_k_tokens_u8(const term_t* idx, const uint8_t* w, int n, int k, term_t* out, int* out_n) {
    return guard([&] {
        auto r = top_k_tokens<uint8_t>(idx, w, n, k);
        std::copy(r.begin(), r.end(), out);
        *out_n = static_cast<int>(r.size());
    });
}

int oracle_top_k_tokens_f32(const term_t* idx, const float* w, int n, int k, term_t* out, int* out_n) {
    return guard([&] {
        auto r = top_k_tokens<float>(idx, w, n, k);
        std::copy(r.begin(), r.end(), out);
        *out_n = static_cast<int>(r.size());
    });
}

int oracle_cluster_order(const float* scores, int n, int64_t* out) {
    return guard([&] {
        auto o = reorder_clusters(std::vector<float>(scores, scores + n), true);
        for (int i = 0; i < n; ++i) out[i] = static_cast<int64_t>(o[i]);
    });
}

int oracle_sq_encode(float vmin, float vmax, const float* v, size_t n, uint8_t* out) {
    return guard([&] { Quantizer(vmin, vmax).encode_n(v, n, out); });
}

// 16-bit codes (QT_16bit), n uint16 values into out.
int oracle_sq_encode16(float vmin, float vmax, const float* v, size_t n, uint16_t* out) {
    return guard([&] { Quantizer(vmin, vmax, 16).encode_n(v, n, reinterpret_cast<uint8_t*>(out)); });
}

float oracle_sq_decode(float ivmin, float ivmax, float qvmin, float qvmax, float s) {
    return Quantizer(ivmin, ivmax).decode_dot_product(s, Quantizer(qvmin, qvmax));
}

float oracle_sq_decode16(float ivmin, float ivmax, float qvmin, float qvmax, float s) {
    return Quantizer(ivmin, ivmax, 16).decode_dot_product(s, Quantizer(qvmin, qvmax, 16));
}

int oracle_select_centroids(int n_docs, int n_clusters, uint32_t seed, int32_t* out_positions) {
    return guard([&] {
        std::vector<idx_t> docs(n_docs);
        std::iota(docs.begin(), docs.end(), 0);
        auto c = select_centroids(docs, static_cast<size_t>(n_clusters), seed);
        for (size_t i = 0; i < c.size(); ++i) out_positions[i] = c[i][0];
    });
}

uint32_t oracle_mt19937_nth(uint32_t seed, int n) {
    std::mt19937 g(seed);
    uint32_t v = 0;
    for (int i = 0; i < n; ++i) v = static_cast<uint32_t>(g());
    return v;
}

int oracle_lambda_beta(int lambda, int beta, int64_t n_vectors, int* out_lambda, int* out_beta) {
    return guard([&] {
        *out_lambda = calculate_lambda(lambda, static_cast<size_t>(n_vectors));
        *out_beta = calculate_beta(beta, *out_lambda);
    });
}

// ---- InvertedIndex ----

int oracle_inv_create(int dim, int tie_mode, void** out) {
    return guard([&] {
        auto* ix = new InvIndex;
        ix->dim = dim;
        ix->tie_mode = tie_mode;
        *out = ix;
    });
}

void oracle_inv_free(void* h) { delete static_cast<InvIndex*>(h); }

int oracle_inv_add(void* h, int n, const idx_t* indptr, const term_t* indices, const float* values) {
    return guard([&] { inv_add(*static_cast<InvIndex*>(h), n, indptr, indices, values); });
}

int oracle_inv_build(void* h) {
    return guard([&] { inv_build(*static_cast<InvIndex*>(h)); });
}

int64_t oracle_inv_num_vectors(void* h) { return static_cast<InvIndex*>(h)->n_vectors; }

// InvertedIndex::search (inverted_index.cpp:254-293): queries in parallel;
// an index that was never built writes nothing. Selector arguments as for
// oracle_search.
int oracle_inv_search(void* h, int n, const idx_t* indptr, const term_t* indices, const float* values, int k,
                      int sel_kind, int64_t sel_n, const idx_t* sel_ids, const uint32_t* sel_bitmap, int sel_negate,
                      float* D, idx_t* L, int n_threads) {
    return guard([&] {
        const auto* ix = static_cast<const InvIndex*>(h);
        if (n <= 0 || k <= 0) throw std::invalid_argument("value must be positive");
        if (!ix->built) return;
        Selector inner;
        Selector sel = make_selector(sel_kind, sel_n, sel_ids, sel_bitmap, sel_negate, inner);
        const Selector* sp = sel.kind < 0 ? nullptr : &sel;
        for (int64_t j = indptr[0]; j < indptr[n]; ++j)
            if (indices[j] >= ix->dim) throw std::invalid_argument("query term id out of range");
        if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 16)
        for (int q = 0; q < n; ++q) {
            const idx_t s0 = indptr[q];
            const int len = indptr[q + 1] - s0;
            if (ix->tie_mode == 0)
                inv_query<HeapTopK<idx_t>>(*ix, indices + s0, values + s0, len, k, sp, D + int64_t{q} * k,
                                           L + int64_t{q} * k);
            else
                inv_query<IdTopK<idx_t>>(*ix, indices + s0, values + s0, len, k, sp, D + int64_t{q} * k,
                                         L + int64_t{q} * k);
        }
    });
}

int oracle_inv_serialize(void* h, uint8_t** out, size_t* size) {
    return guard([&] {
        Writer w;
        inv_write(*static_cast<InvIndex*>(h), w);
        *size = w.buf.size();
        *out = static_cast<uint8_t*>(std::malloc(w.buf.size()));
        std::memcpy(*out, w.buf.data(), w.buf.size());
    });
}

}  // extern "C"
