// ORACLE (test infrastructure): a thin C entry layer over the reference's own
// build-side sources (cluster/, invlists/, sparse_vectors.cpp, io/), compiled
// straight from /root/reference into oracle/_ref/ by oracle/Makefile. Used only
// by tests/ to check the restatement in oracle/oracle.cpp (build bytes, summary
// scores, SQ codes, dot products). It calls the reference's functions; the
// only code here is argument plumbing plus the
// per-list loop of build_inverted_lists_clusters (seismic_common.h:146-184) and
// the header writes of write_index (io/index_io.cpp:88-95,
// seismic_scalar_quantized_index.cpp:421-428), which live in files that pull in
// Abseil (absent here) and so cannot be compiled themselves.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "nsparse/cluster/inverted_list_clusters.h"
#include "nsparse/cluster/random_kmeans.h"
#include "nsparse/invlists/inverted_lists.h"
#include "nsparse/io/buffered_io.h"
#include "nsparse/sparse_vectors.h"
#include "nsparse/utils/distance.h"
#include "nsparse/utils/scalar_quantizer.h"

using namespace nsparse;

namespace {
thread_local std::string g_err;
template <class F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// The reference's forward index and per-term InvertedListClusters, built with
// final lambda and beta (already resolved): the per-list loop of
// build_inverted_lists_clusters (seismic_common.h:146-184).
struct RefBuilt {
    SparseVectors vectors;
    ScalarQuantizer q;
    std::vector<InvertedListClusters> lists;
};

std::unique_ptr<RefBuilt> ref_build(int dim, int sq, float vmin, float vmax, int n, const int32_t* indptr,
                                    const uint16_t* indices, const float* values, int lambda, int beta, float alpha,
                                    uint32_t base_seed) {
    const size_t esz = sq == 2 ? 2 : sq ? 1 : 4;
    const size_t nnz = indptr[n];
    auto b = std::unique_ptr<RefBuilt>(new RefBuilt{
        SparseVectors({.element_size = esz, .dimension = static_cast<size_t>(dim)}),
        ScalarQuantizer(sq == 2 ? QuantizerType::QT_16bit : QuantizerType::QT_8bit, sq ? vmin : 0.0F,
                        sq ? vmax : 1.0F),
        {}});
    std::vector<uint8_t> codes;
    if (sq) {
        codes.resize(nnz * esz);
        b->q.encode(values, codes.data(), nnz);
    } else {
        codes.resize(nnz * 4);
        std::memcpy(codes.data(), values, nnz * 4);
    }
    b->vectors.add_vectors(indptr, n + 1, indices, nnz, codes.data(), codes.size());
    auto inv = ArrayInvertedLists::build_inverted_lists(dim, esz, &b->vectors);
    b->lists.resize(inv->size());
    for (size_t idx = 0; idx < inv->size(); ++idx) {
        auto& il = (*inv)[idx];
        const auto doc_ids = il.prune_and_keep_doc_ids(lambda);
        InvertedListClusters c(
            detail::RandomKMeans::train(&b->vectors, doc_ids, beta, base_seed + static_cast<uint32_t>(idx)));
        c.summarize(&b->vectors, alpha);
        b->lists[idx] = std::move(c);
        il.clear();
    }
    return b;
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// Builds a SEIS (sq == 0) or SESQ (sq == 1: 8-bit, sq == 2: 16-bit) index with final lambda and
// beta (already resolved) and writes its full write_index byte stream.
int ref_build_serialize(int dim, int sq, float vmin, float vmax, int n, const int32_t* indptr,
                        const uint16_t* indices, const float* values, int lambda, int beta, float alpha,
                        uint32_t base_seed, uint8_t** out, size_t* out_size) {
    return guard([&] {
        auto b = ref_build(dim, sq, vmin, vmax, n, indptr, indices, values, lambda, beta, alpha, base_seed);
        const SparseVectors& vectors = b->vectors;
        const ScalarQuantizer& q = b->q;
        const std::vector<InvertedListClusters>& lists = b->lists;
        BufferedIOWriter w;
        const char* id = sq ? "SESQ" : "SEIS";
        uint32_t fcc = static_cast<uint32_t>(id[0]) | static_cast<uint32_t>(id[1]) << 8 |
                       static_cast<uint32_t>(id[2]) << 16 | static_cast<uint32_t>(id[3]) << 24;
        w.write(&fcc, sizeof(uint32_t), 1);
        int d = dim;
        w.write(&d, sizeof(int), 1);
        if (sq) {
            auto t = q.get_quantizer_type();
            w.write(&t, sizeof(QuantizerType), 1);
            float lo = q.get_min(), hi = q.get_max();
            w.write(&lo, sizeof(float), 1);
            w.write(&hi, sizeof(float), 1);
        }
        vectors.serialize(&w);
        size_t n_lists = lists.size();
        w.write(&n_lists, sizeof(size_t), 1);
        for (const auto& l : lists) l.serialize(&w);
        *out_size = w.size();
        *out = static_cast<uint8_t*>(std::malloc(w.size()));
        std::memcpy(*out, w.data().data(), w.size());
    });
}

void ref_buffer_free(uint8_t* p) { std::free(p); }

// InvertedListClusters::score_summaries_transposed (inverted_list_clusters.cpp:295-331)
// of the reference's own build: query q (all its entries, in CSR order) against
// the summaries of list lists[q]. As the reference's search does, SQ queries go
// in as codes of the index's quantizer (seismic_scalar_quantized_index.cpp:
// 182-191, 87-88), FP32 queries as their float bytes. out holds cap floats per
// query; out_n[q] receives the list's cluster count (scores past cap are dropped).
int ref_score_summaries(int dim, int sq, float vmin, float vmax, int n, const int32_t* indptr,
                        const uint16_t* indices, const float* values, int lambda, int beta, float alpha,
                        uint32_t base_seed, int nq, const int32_t* q_indptr, const uint16_t* q_indices,
                        const float* q_values, const int32_t* lists, int cap, float* out, int32_t* out_n) {
    return guard([&] {
        auto b = ref_build(dim, sq, vmin, vmax, n, indptr, indices, values, lambda, beta, alpha, base_seed);
        std::vector<float> scores;
        for (int qi = 0; qi < nq; ++qi) {
            const int32_t s0 = q_indptr[qi];
            const size_t len = static_cast<size_t>(q_indptr[qi + 1] - s0);
            std::vector<uint8_t> qbytes;
            if (sq) {
                qbytes.resize(len * (sq == 2 ? 2 : 1));
                b->q.encode(q_values + s0, qbytes.data(), len);
            } else {
                qbytes.resize(len * 4);
                std::memcpy(qbytes.data(), q_values + s0, len * 4);
            }
            const auto& lc = b->lists.at(static_cast<size_t>(lists[qi]));
            lc.score_summaries_transposed(q_indices + s0, qbytes.data(), len, scores);
            out_n[qi] = static_cast<int32_t>(scores.size());
            for (size_t c = 0; c < scores.size() && c < static_cast<size_t>(cap); ++c)
                out[static_cast<size_t>(qi) * cap + c] = scores[c];
        }
    });
}

// The posting lists and per-term maxima of InvertedIndex::build
// (inverted_index.cpp:236-252), made by the reference's own
// ArrayInvertedLists::build_inverted_lists and InvertedList::max_value
// (invlists/inverted_lists.cpp:88-96, 159-180), written in the field order of
// InvertedIndex::write_index (inverted_index.cpp:367-402) after the header.
// inverted_index.cpp itself pulls in Abseil, so only that loop is restated here.
int ref_inverted_serialize(int dim, int n, const int32_t* indptr, const uint16_t* indices, const float* values,
                           uint8_t** out, size_t* out_size) {
    return guard([&] {
        const size_t nnz = indptr[n];
        SparseVectors vectors({.element_size = 4, .dimension = static_cast<size_t>(dim)});
        vectors.add_vectors(indptr, static_cast<size_t>(n) + 1, indices, nnz,
                            reinterpret_cast<const uint8_t*>(values), nnz * 4);
        auto lists = ArrayInvertedLists::build_inverted_lists(static_cast<size_t>(dim), 4, &vectors);
        std::vector<uint8_t> w;
        auto put = [&](const void* p, size_t bytes) {
            const auto* b = static_cast<const uint8_t*>(p);
            w.insert(w.end(), b, b + bytes);
        };
        const uint32_t fcc = 'I' | 'N' << 8 | 'V' << 16 | 'T' << 24;
        put(&fcc, 4);
        put(&dim, 4);
        const size_t nv = static_cast<size_t>(n), n_terms = lists->size(), esz = 4;
        put(&nv, 8);
        put(&n_terms, 8);
        put(&esz, 8);
        std::vector<float> maxima(n_terms);
        for (size_t t = 0; t < n_terms; ++t) {
            const auto& l = (*lists)[t];
            const size_t len = l.get_doc_ids().size();
            put(&len, 8);
            if (len > 0) {
                put(l.get_doc_ids().data(), len * 4);
                put(l.get_codes().data(), l.get_codes().size());
            }
            maxima[t] = l.max_value();
        }
        put(&n_terms, 8);
        put(maxima.data(), n_terms * 4);
        *out_size = w.size();
        *out = static_cast<uint8_t*>(std::malloc(w.size()));
        std::memcpy(*out, w.data(), w.size());
    });
}

// InvertedList::prune_and_keep_doc_ids on a hand-made list (inverted_lists.cpp:110-132).
int ref_prune(int n, const int32_t* docs, const uint8_t* codes, int esz, int64_t lambda, int32_t* out, int* out_n) {
    return guard([&] {
        InvertedList il(static_cast<size_t>(esz));
        il.add_entries(n, docs, codes);
        auto kept = il.prune_and_keep_doc_ids(static_cast<size_t>(lambda));
        std::copy(kept.begin(), kept.end(), out);
        *out_n = static_cast<int>(kept.size());
    });
}

// ScalarQuantizer encode / decode_dot_product (utils/scalar_quantizer.h).
int ref_sq_encode(float vmin, float vmax, const float* v, size_t n, uint8_t* out) {
    return guard([&] { ScalarQuantizer(QuantizerType::QT_8bit, vmin, vmax).encode(v, out, n); });
}
float ref_sq_decode(float ivmin, float ivmax, float qvmin, float qvmax, float s) {
    ScalarQuantizer a(QuantizerType::QT_8bit, ivmin, ivmax), b(QuantizerType::QT_8bit, qvmin, qvmax);
    return a.decode_dot_product(s, b);
}
int ref_sq_encode16(float vmin, float vmax, const float* v, size_t n, uint16_t* out) {
    return guard([&] {
        ScalarQuantizer(QuantizerType::QT_16bit, vmin, vmax).encode(v, reinterpret_cast<uint8_t*>(out), n);
    });
}
float ref_sq_decode16(float ivmin, float ivmax, float qvmin, float qvmax, float s) {
    ScalarQuantizer a(QuantizerType::QT_16bit, ivmin, ivmax), b(QuantizerType::QT_16bit, qvmin, qvmax);
    return a.decode_dot_product(s, b);
}

// detail::dot_product_{uint8,uint16,float}_dense (utils/distance.h:29-57), generic build.
float ref_dot_u8(const uint16_t* idx, const uint8_t* w, size_t len, const uint8_t* dense) {
    return detail::dot_product_uint8_dense(idx, w, len, dense);
}
float ref_dot_f32(const uint16_t* idx, const float* w, size_t len, const float* dense) {
    return detail::dot_product_float_dense(idx, w, len, dense);
}
float ref_dot_u16(const uint16_t* idx, const uint16_t* w, size_t len, const uint16_t* dense) {
    return detail::dot_product_uint16_dense(idx, w, len, dense);
}

}  // extern "C"
