// Step-for-step re-implementations of the two libstdc++ 11 orderings whose tie
// behaviour decides which clusters and which query terms SEISMIC visits:
//
//  * std::sort (introsort: median-of-3 Hoare partition, depth limit 2*lg(n),
//    heapsort fallback, final insertion sort with threshold 16) as called by
//    std::ranges::sort in reorder_clusters (seismic_common.h:105-115);
//  * std::priority_queue push/pop (push_heap / pop_heap / adjust_heap) as used
//    by TopKHolder inside top_k_tokens (utils/ranker.h:58-108,
//    utils/vector_process.h:18-30).
//
// The algorithms' steps must match libstdc++'s exactly for equal keys to land
// where the reference puts them, so the control flow below mirrors
// bits/stl_algo.h and bits/stl_heap.h of GCC 11; the code itself is written for
// a single GPU lane working in LDS (iterative, no recursion) and compiles for
// the host too, where the tests compare it against the real std::ranges::sort
// and std::priority_queue.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define NS_HD __host__ __device__ __forceinline__
#else
#define NS_HD inline
#endif

namespace nsparse_amd {

// A (key, payload) element. comp(a, b) := a.key > b.key, i.e. the reference's
// `summary_scores[a] > summary_scores[b]` (descending) or TopKHolder's
// std::greater<float> on scores.
struct KeyItem {
    float key;
    uint32_t id;
};

NS_HD bool greater_key(const KeyItem& a, const KeyItem& b) { return a.key > b.key; }

NS_HD int floor_log2(int n) {
    int r = 0;
    while (n > 1) {
        n >>= 1;
        ++r;
    }
    return r;
}

// The heap routines take any random-access handle F over KeyItem (a pointer,
// or on the GPU a heap held in a wavefront's registers).

// bits/stl_heap.h __push_heap
template <class F>
NS_HD void heap_push_hole(F f, int hole, int top, KeyItem value) {
    int parent = (hole - 1) / 2;
    while (hole > top && greater_key(f[parent], value)) {
        f[hole] = f[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    f[hole] = value;
}

// bits/stl_heap.h __adjust_heap
template <class F>
NS_HD void heap_adjust(F f, int hole, int len, KeyItem value) {
    const int top = hole;
    int second = hole;
    while (second < (len - 1) / 2) {
        second = 2 * (second + 1);
        if (greater_key(f[second], f[second - 1])) second--;
        f[hole] = f[second];
        hole = second;
    }
    if ((len & 1) == 0 && second == (len - 2) / 2) {
        second = 2 * (second + 1);
        f[hole] = f[second - 1];
        hole = second - 1;
    }
    heap_push_hole(f, hole, top, value);
}

// std::priority_queue<Item, vector, greater-on-score>::push
template <class F>
NS_HD void pq_push(F heap, int& size, KeyItem v) {
    heap[size] = v;
    ++size;
    heap_push_hole(heap, size - 1, 0, v);
}

// std::priority_queue::pop (std::pop_heap, then pop_back)
template <class F>
NS_HD void pq_pop(F heap, int& size) {
    if (size > 1) {
        const int last = size - 1;
        const KeyItem value = heap[last];
        heap[last] = heap[0];
        heap_adjust(heap, 0, last, value);
    }
    --size;
}

// TopKHolder<term_t>::add followed by top_k_descending, for the query cut:
// keeps `k` items of the `n` (weights[i], terms[i]) pairs and writes their
// terms best-first into `out`; returns the count. `heap` needs k slots.
// When k >= n the reference returns all terms in input order.
template <class GetW, class GetT, class F>
NS_HD int topk_tokens(int n, int k, GetW weight, GetT term, F heap, uint16_t* out) {
    if (k >= n) {
        for (int i = 0; i < n; ++i) out[i] = static_cast<uint16_t>(term(i));
        return n;
    }
    int size = 0;
    float threshold = __builtin_inff();
    for (int i = 0; i < n; ++i) {
        const float s = weight(i);
        if (size >= k && s <= threshold) continue;
        if (size < k) {
            pq_push(heap, size, KeyItem{s, static_cast<uint32_t>(term(i))});
            if (size == k) threshold = KeyItem(heap[0]).key;
        } else {
            pq_pop(heap, size);
            pq_push(heap, size, KeyItem{s, static_cast<uint32_t>(term(i))});
            threshold = KeyItem(heap[0]).key;
        }
    }
    const int total = size;
    for (int idx = total - 1; idx >= 0; --idx) {
        out[idx] = static_cast<uint16_t>(KeyItem(heap[0]).id);
        pq_pop(heap, size);
    }
    return total;
}

// ---- std::sort(first, last, comp) with comp = greater_key ----

// bits/stl_algo.h __move_median_to_first
NS_HD void median_to_first(KeyItem* A, int result, int a, int b, int c) {
    int pick;
    if (greater_key(A[a], A[b])) {
        if (greater_key(A[b], A[c]))
            pick = b;
        else if (greater_key(A[a], A[c]))
            pick = c;
        else
            pick = a;
    } else if (greater_key(A[a], A[c]))
        pick = a;
    else if (greater_key(A[b], A[c]))
        pick = c;
    else
        pick = b;
    const KeyItem t = A[result];
    A[result] = A[pick];
    A[pick] = t;
}

// bits/stl_algo.h __unguarded_partition_pivot (+ __unguarded_partition)
NS_HD int partition_pivot(KeyItem* A, int first, int last) {
    const int mid = first + (last - first) / 2;
    median_to_first(A, first, first + 1, mid, last - 1);
    const KeyItem pivot = A[first];
    int lo = first + 1, hi = last;
    while (true) {
        while (greater_key(A[lo], pivot)) ++lo;
        --hi;
        while (greater_key(pivot, A[hi])) --hi;
        if (!(lo < hi)) return lo;
        const KeyItem t = A[lo];
        A[lo] = A[hi];
        A[hi] = t;
        ++lo;
    }
}

// bits/stl_algo.h __partial_sort(first, last, last): __make_heap + __sort_heap
NS_HD void heap_sort_range(KeyItem* A, int first, int last) {
    KeyItem* f = A + first;
    const int len = last - first;
    if (len >= 2) {
        int parent = (len - 2) / 2;
        while (true) {
            heap_adjust(f, parent, len, f[parent]);
            if (parent == 0) break;
            parent--;
        }
    }
    int end = len;
    while (end > 1) {
        --end;
        const KeyItem value = f[end];
        f[end] = f[0];
        heap_adjust(f, 0, end, value);
    }
}

// bits/stl_algo.h __unguarded_linear_insert
NS_HD void unguarded_linear_insert(KeyItem* A, int last) {
    const KeyItem val = A[last];
    int next = last - 1;
    while (greater_key(val, A[next])) {
        A[last] = A[next];
        last = next;
        --next;
    }
    A[last] = val;
}

// bits/stl_algo.h __insertion_sort
NS_HD void insertion_sort(KeyItem* A, int first, int last) {
    if (first == last) return;
    for (int i = first + 1; i != last; ++i) {
        if (greater_key(A[i], A[first])) {
            const KeyItem val = A[i];
            for (int j = i; j > first; --j) A[j] = A[j - 1];
            A[first] = val;
        } else {
            unguarded_linear_insert(A, i);
        }
    }
}

constexpr int kIntrosortThreshold = 16;  // _S_threshold
constexpr int kIntrosortStack = 40;  // > 2*lg(65536) + 1 pending right parts

// bits/stl_algo.h __sort: __introsort_loop (its recursion on the right part
// turned into an explicit stack -- the parts are disjoint, so the order in
// which they are finished does not change the result) + __final_insertion_sort.
NS_HD void introsort_desc(KeyItem* A, int n) {
    if (n <= 0) return;
    int st_lo[kIntrosortStack], st_hi[kIntrosortStack], st_d[kIntrosortStack];
    int sp = 0;
    st_lo[sp] = 0;
    st_hi[sp] = n;
    st_d[sp] = floor_log2(n) * 2;
    ++sp;
    while (sp > 0) {
        --sp;
        int lo = st_lo[sp], hi = st_hi[sp], depth = st_d[sp];
        while (hi - lo > kIntrosortThreshold) {
            if (depth == 0) {
                heap_sort_range(A, lo, hi);
                break;
            }
            --depth;
            const int cut = partition_pivot(A, lo, hi);
            st_lo[sp] = cut;
            st_hi[sp] = hi;
            st_d[sp] = depth;
            ++sp;
            hi = cut;
        }
    }
    if (n > kIntrosortThreshold) {
        insertion_sort(A, 0, kIntrosortThreshold);
        for (int i = kIntrosortThreshold; i != n; ++i) unguarded_linear_insert(A, i);
    } else {
        insertion_sort(A, 0, n);
    }
}

}  // namespace nsparse_amd
