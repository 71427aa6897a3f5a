// A C++ caller written against the reference's public headers
// (nsparse/index.h, index_factory.h, seismic_index.h,
// seismic_scalar_quantized_index.h, id_map_index.h, id_selector.h,
// io/index_io.h), compiled unchanged against include/nsparse/ and linked to
// the device library. The calls follow benchmarks/index_search_benchmark.cpp:
// 122-147,203-212 of the reference. Test infrastructure: the pytest driver
// writes the inputs, runs a mode, and checks what it writes.
//
//   dropin_check api
//   dropin_check build  <dim> <description> <docs.bin> <out.index>
//   dropin_check search <index> <queries.bin> <ids.bin> <vmax> <k> <out_prefix>
//   dropin_check idmap  <dim> <description> <docs.bin> <ext_ids.bin> <queries.bin> <sel_ids.bin> <k> <out_prefix>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <typeinfo>
#include <vector>

#include "nsparse/brutal_index.h"
#include "nsparse/id_map_index.h"
#include "nsparse/id_selector.h"
#include "nsparse/index.h"
#include "nsparse/index_factory.h"
#include "nsparse/inverted_index.h"
#include "nsparse/io/index_io.h"
#include "nsparse/seismic_common.h"
#include "nsparse/seismic_index.h"
#include "nsparse/seismic_scalar_quantized_index.h"
#include "nsparse/types.h"

namespace {

struct Csr {
    int32_t n = 0;
    std::vector<nsparse::idx_t> indptr;
    std::vector<nsparse::term_t> indices;
    std::vector<float> values;
};

// int32 rows, int64 nnz, int32 indptr[rows + 1], uint16 indices[nnz], float values[nnz]
Csr read_csr_bin(const char* path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    Csr m;
    int64_t nnz = 0;
    f.read(reinterpret_cast<char*>(&m.n), 4);
    f.read(reinterpret_cast<char*>(&nnz), 8);
    m.indptr.resize(m.n + 1);
    m.indices.resize(nnz);
    m.values.resize(nnz);
    f.read(reinterpret_cast<char*>(m.indptr.data()), (m.n + 1) * 4);
    f.read(reinterpret_cast<char*>(m.indices.data()), nnz * 2);
    f.read(reinterpret_cast<char*>(m.values.data()), nnz * 4);
    if (!f) throw std::runtime_error(std::string("short file ") + path);
    return m;
}

// int32 count, int32 ids[count]
std::vector<nsparse::idx_t> read_ids(const char* path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error(std::string("cannot open ") + path);
    int32_t n = 0;
    f.read(reinterpret_cast<char*>(&n), 4);
    std::vector<nsparse::idx_t> ids(n);
    f.read(reinterpret_cast<char*>(ids.data()), static_cast<std::streamsize>(n) * 4);
    if (!f) throw std::runtime_error(std::string("short file ") + path);
    return ids;
}

// float distances[n * k], then int32 labels[n * k]
void write_results(const std::string& path, const std::vector<float>& d, const std::vector<nsparse::idx_t>& l) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(d.data()), static_cast<std::streamsize>(d.size() * 4));
    f.write(reinterpret_cast<const char*>(l.data()), static_cast<std::streamsize>(l.size() * 4));
    if (!f) throw std::runtime_error("cannot write " + path);
}

void expect(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "FAILED: %s\n", what);
        std::exit(1);
    }
}

template <class E, class F>
void expect_throw(F&& f, const char* what) {
    try {
        f();
    } catch (const E&) {
        return;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "FAILED: %s threw %s (%s)\n", what, typeid(e).name(), e.what());
        std::exit(1);
    }
    std::fprintf(stderr, "FAILED: %s did not throw\n", what);
    std::exit(1);
}

// Every id not divisible by three: a selector the library cannot enumerate.
class NotMultipleOfThree : public nsparse::IDSelector {
public:
    bool is_member(nsparse::idx_t id) const override { return id % 3 != 0; }
};

// An enumerable selector of a type of its own.
class FewIds : public nsparse::IDSelectorEnumerable {
public:
    explicit FewIds(std::vector<nsparse::idx_t> ids) : ids_(std::move(ids)) {}
    bool is_member(nsparse::idx_t id) const override {
        for (auto i : ids_)
            if (i == id) return true;
        return false;
    }
    std::vector<nsparse::idx_t> ids() const override { return ids_; }
    size_t size() const override { return ids_.size(); }

private:
    std::vector<nsparse::idx_t> ids_;
};

class Throwing : public nsparse::IDSelector {
public:
    bool is_member(nsparse::idx_t) const override { throw std::domain_error("selector failed"); }
};

int mode_api() {
    const int dim = 30522;
    // index_factory hands back the class its description names.
    std::unique_ptr<nsparse::Index> s(nsparse::index_factory(dim, "seismic,lambda=100|beta=10|alpha=0.4"));
    expect(dynamic_cast<nsparse::SeismicIndex*>(s.get()) != nullptr, "seismic is a SeismicIndex");
    expect(s->id() == nsparse::SeismicIndex::name, "seismic fourcc");
    expect(s->get_dimension() == dim, "dimension");
    std::unique_ptr<nsparse::Index> q(
        nsparse::index_factory(dim, "seismic_sq,quantizer=8bit|vmin=0|vmax=3|lambda=100|beta=10|alpha=0.4"));
    expect(dynamic_cast<nsparse::SeismicScalarQuantizedIndex*>(q.get()) != nullptr, "seismic_sq class");
    expect(q->id() == nsparse::SeismicScalarQuantizedIndex::name, "seismic_sq fourcc");
    std::unique_ptr<nsparse::Index> b(nsparse::index_factory(dim, "brutal"));
    expect(dynamic_cast<nsparse::BrutalIndex*>(b.get()) != nullptr, "brutal class");
    std::unique_ptr<nsparse::Index> v(nsparse::index_factory(dim, "inverted"));
    expect(dynamic_cast<nsparse::InvertedIndex*>(v.get()) != nullptr, "inverted class");
    std::unique_ptr<nsparse::Index> m(nsparse::index_factory(dim, "idmap,seismic,lambda=100|beta=10|alpha=0.4"));
    expect(dynamic_cast<nsparse::IDMapIndex*>(m.get()) != nullptr, "idmap class");
    expect(m->id() == nsparse::IDMapIndex::name, "idmap fourcc");

    // Constructors, as callers of the classes write them.
    nsparse::SeismicClusterParameters cp{200, 20, 0.4F};
    cp.seed = 0;
    nsparse::SeismicIndex si(dim, cp);
    nsparse::SeismicIndex sd(dim);
    nsparse::SeismicScalarQuantizedIndex sq(nsparse::QuantizerType::QT_8bit, 0.0F, 2.5F, cp, dim);
    nsparse::SeismicScalarQuantizedIndex sqd(dim);
    nsparse::BrutalIndex br(dim);
    nsparse::InvertedIndex inv(dim);
    expect(si.id() == nsparse::SeismicIndex::name && sq.id() == nsparse::SeismicScalarQuantizedIndex::name,
           "class fourccs");
    auto* owned = new nsparse::SeismicIndex(dim, cp);
    nsparse::IDMapIndex idm(owned);  // takes ownership, as the reference's does
    expect(idm.get_dimension() == dim, "idmap dimension");

    // add, num_vectors (three rows; FP32 adds stay on the host)
    const nsparse::idx_t ip[] = {0, 2, 3, 5};
    const nsparse::term_t ix[] = {1, 7, 3, 2, 9};
    const float vv[] = {0.5F, 1.0F, 2.0F, 0.25F, 0.75F};
    si.add(3, ip, ix, vv);
    expect(si.num_vectors() == 3, "num_vectors after add");
    // (a SEISMIC-SQ add quantizes on the device: the GPU modes cover it)
    const nsparse::idx_t ids[] = {40, 10, 30};
    idm.add_with_ids(3, ip, ix, vv, ids);
    expect(idm.num_vectors() == 3, "idmap num_vectors after add_with_ids");

    // The reference's exception types.
    expect_throw<std::invalid_argument>([&] { delete nsparse::index_factory(dim, "bogus"); }, "unknown type");
    expect_throw<std::invalid_argument>([&] { delete nsparse::index_factory(dim, ""); }, "empty description");
    nsparse::SeismicScalarQuantizedIndex q16(nsparse::QuantizerType::QT_16bit, 0.0F, 1.0F, cp, dim);
    expect(q16.id() == nsparse::SeismicScalarQuantizedIndex::name, "16-bit quantizer");
    expect_throw<std::runtime_error>([&] { si.add_with_ids(3, ip, ix, vv, ids); }, "add_with_ids on SeismicIndex");
    expect_throw<std::invalid_argument>([&] { si.add(0, ip, ix, vv); }, "add of zero rows");
    expect_throw<std::invalid_argument>([&] { si.add(3, nullptr, ix, vv); }, "add with a null array");
    expect_throw<std::invalid_argument>([&] { owned->add(3, ip, ix, vv); }, "a delegate taken by an IDMap");
    expect_throw<std::runtime_error>([&] { delete nsparse::read_index("/nonexistent/file.index"); }, "missing file");

    // SearchParameters family
    nsparse::SeismicSearchParameters sp(3, 1.0F);
    nsparse::SeismicSQSearchParameters sqp(0.0F, 2.0F, 5, 0.9F);
    nsparse_search_params c{};
    sp.to_c(&c);
    expect(c.kind == NSPARSE_PARAMS_SEISMIC && c.cut == 3 && c.heap_factor == 1.0F, "SeismicSearchParameters");
    sqp.to_c(&c);
    expect(c.kind == NSPARSE_PARAMS_SEISMIC_SQ && c.cut == 5 && c.vmax == 2.0F, "SeismicSQSearchParameters");
    nsparse::SetIDSelector set(3, ids);
    sp.set_id_selector(&set);
    expect(sp.get_id_selector() == &set, "selector getter");
    std::printf("api ok\n");
    return 0;
}

int mode_build(int dim, const char* desc, const char* docs_path, const char* out) {
    const Csr docs = read_csr_bin(docs_path);
    nsparse::Index* index = nsparse::index_factory(dim, desc);
    index->add(docs.n, docs.indptr.data(), docs.indices.data(), docs.values.data());
    index->build();
    nsparse::write_index(index, const_cast<char*>(out));
    delete index;
    std::printf("build ok\n");
    return 0;
}

int mode_search(const char* index_path, const char* q_path, const char* ids_path, float vmax, int k,
                const std::string& prefix) {
    const Csr qs = read_csr_bin(q_path);
    const std::vector<nsparse::idx_t> allowed = read_ids(ids_path);
    std::unique_ptr<nsparse::Index> index(
        nsparse::read_index(const_cast<char*>(index_path), nsparse::IndexIoFlag::kUseMmap));
    std::unique_ptr<nsparse::Index> copy(nsparse::read_index(index_path));
    const size_t nk = static_cast<size_t>(qs.n) * k;
    std::vector<float> d(nk);
    std::vector<nsparse::idx_t> l(nk);
    auto run = [&](nsparse::Index* ix, nsparse::SearchParameters* p, const char* name) {
        std::fill(d.begin(), d.end(), 0.0F);
        std::fill(l.begin(), l.end(), 0);
        ix->search(qs.n, qs.indptr.data(), qs.indices.data(), qs.values.data(), k, d.data(), l.data(), p);
        write_results(prefix + name + ".bin", d, l);
    };
    nsparse::SeismicSearchParameters params(/*cut=*/10, /*heap_factor=*/0.8F);
    run(index.get(), &params, "plain");
    run(copy.get(), &params, "plain_copy");
    run(index.get(), nullptr, "defaults");
    nsparse::SeismicSQSearchParameters ranged(0.0F, 2.0F * vmax, 10, 0.8F);
    run(index.get(), &ranged, "sq_range");
    nsparse::SetIDSelector set(allowed.size(), allowed.data());
    params.set_id_selector(&set);
    run(index.get(), &params, "set");
    nsparse::NotIDSelector not_set(&set);
    params.set_id_selector(&not_set);
    run(index.get(), &params, "not_set");
    NotMultipleOfThree custom;
    params.set_id_selector(&custom);
    run(index.get(), &params, "custom");
    nsparse::NotIDSelector not_custom(&custom);
    params.set_id_selector(&not_custom);
    run(index.get(), &params, "not_custom");
    FewIds few({allowed[5], allowed[1], allowed[5], allowed[9]});
    params.set_id_selector(&few);
    run(index.get(), &params, "few");
    Throwing bad;
    params.set_id_selector(&bad);
    expect_throw<std::domain_error>([&] { run(index.get(), &params, "unused"); }, "throwing selector");
    std::printf("search ok\n");
    return 0;
}

int mode_idmap(int dim, const char* desc, const char* docs_path, const char* ext_path, const char* q_path,
               const char* sel_path, int k, const std::string& prefix) {
    const Csr docs = read_csr_bin(docs_path);
    const std::vector<nsparse::idx_t> ext = read_ids(ext_path);
    const Csr qs = read_csr_bin(q_path);
    const std::vector<nsparse::idx_t> sel = read_ids(sel_path);
    std::unique_ptr<nsparse::Index> index(new nsparse::IDMapIndex(nsparse::index_factory(dim, desc)));
    index->add_with_ids(docs.n, docs.indptr.data(), docs.indices.data(), docs.values.data(), ext.data());
    index->build();
    const size_t nk = static_cast<size_t>(qs.n) * k;
    std::vector<float> d(nk);
    std::vector<nsparse::idx_t> l(nk);
    nsparse::SeismicSearchParameters params(10, 0.9F);
    index->search(qs.n, qs.indptr.data(), qs.indices.data(), qs.values.data(), k, d.data(), l.data(), &params);
    write_results(prefix + "plain.bin", d, l);
    nsparse::SetIDSelector set(sel.size(), sel.data());
    params.set_id_selector(&set);
    index->search(qs.n, qs.indptr.data(), qs.indices.data(), qs.values.data(), k, d.data(), l.data(), &params);
    write_results(prefix + "set.bin", d, l);
    NotMultipleOfThree custom;  // evaluated on external ids
    params.set_id_selector(&custom);
    index->search(qs.n, qs.indptr.data(), qs.indices.data(), qs.values.data(), k, d.data(), l.data(), &params);
    write_results(prefix + "custom.bin", d, l);
    std::printf("idmap ok\n");
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const std::string mode = argc > 1 ? argv[1] : "";
        if (mode == "api") return mode_api();
        if (mode == "build" && argc == 6) return mode_build(std::atoi(argv[2]), argv[3], argv[4], argv[5]);
        if (mode == "search" && argc == 8)
            return mode_search(argv[2], argv[3], argv[4], std::strtof(argv[5], nullptr), std::atoi(argv[6]), argv[7]);
        if (mode == "idmap" && argc == 10)
            return mode_idmap(std::atoi(argv[2]), argv[3], argv[4], argv[5], argv[6], argv[7], std::atoi(argv[8]),
                              argv[9]);
        std::fprintf(stderr, "usage: see the header of dropin_check.cpp\n");
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "FAILED: %s: %s\n", typeid(e).name(), e.what());
        return 1;
    }
}
