// Continuous-batching search server over one device-resident SEISMIC index.
//
// Requests (one query each, with its own k and optional selector bitmap) are
// queued by any thread. A dispatcher thread packs the oldest pending requests
// into batches and runs each batch on one of n_streams default-priority
// streams: host staging (pinned) -> HBM, one batched search launch, results
// back to pinned memory. A batch is admitted only while the bytes of every
// batch in flight (queries + outputs) stay within the scratch budget, and only
// when a stream is free; requests that arrive while others run are admitted as
// soon as a stream and budget free up, so their work overlaps on the device.
// Every query is searched exactly as nsparse_search_device searches it alone.
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "device_mem.h"
#include "index.h"

namespace nsparse_amd {

struct ServerConfig {
    int32_t max_batch = 512;
    int32_t n_streams = 4;
    int64_t scratch_budget = int64_t{256} << 20;
    int32_t k_max = 100;
    SearchRequest req;
};

struct ServerStats {
    int64_t completed = 0;
    int64_t batches = 0;
    int64_t max_inflight_bytes = 0;
    int64_t max_batch_seen = 0;
    int64_t deferred_by_budget = 0;  // admission rounds that stopped at the budget
    int64_t max_concurrent_batches = 0;
};

class Server {
public:
    Server(SeismicIndex& index, const ServerConfig& cfg);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    // Copies a selector bitmap (bit d set = doc d allowed) into HBM; returns its id.
    int32_t add_bitmap(const uint32_t* words, int64_t n_words);
    // Queues one query; arrival_ns < 0 stamps it now (ns since the server started).
    int64_t submit(int32_t nnz, const uint16_t* indices, const float* values, int32_t k, int32_t bitmap_id,
                   int64_t arrival_ns);
    // Blocks until the request is done; copies its k results and returns the
    // latency (completion - arrival) in microseconds.
    double wait(int64_t ticket, float* distances, int32_t* labels);
    int64_t now_ns() const;
    int32_t k_max() const { return cfg_.k_max; }
    ServerStats stats();
    // Bytes of HBM scratch one request of nnz non-zeros holds while in flight.
    int64_t request_bytes(int32_t nnz) const;

private:
    struct Request {
        std::vector<uint16_t> indices;
        std::vector<float> values;
        int32_t k = 0;
        int32_t bitmap_id = -1;
        int64_t arrival_ns = 0;
        int64_t done_ns = -1;
        std::vector<float> dist;
        std::vector<int32_t> lab;
        bool collected = false;
    };
    struct Slot {
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        bool busy = false;
        int64_t bytes = 0;
        std::vector<int64_t> tickets;
        // One pinned block and one device block for a batch's inputs, one of
        // each for its outputs, sized for max_batch full-length queries. A
        // batch packs its arrays from the start of each block, so it moves in
        // one copy each way.
        char *h_in = nullptr, *h_out = nullptr;
        DeviceBuffer d_in, d_out;
        // byte offsets of the current batch's arrays: inputs (query offsets,
        // k, selector, term ids, values) and outputs (error flag at 0,
        // distances, labels)
        size_t off_ks = 0, off_sel = 0, off_ix = 0, off_v = 0, in_bytes = 0;
        size_t off_lab = 0, out_bytes = 0;
    };

    void run();
    bool reap();
    bool admit();
    void launch(Slot& s, const std::vector<int64_t>& batch, int64_t spread_cap);

    SeismicIndex& index_;
    ServerConfig cfg_;
    std::chrono::steady_clock::time_point epoch_;
    std::vector<Slot> slots_;
    std::vector<DeviceBuffer> bitmaps_;
    DeviceBuffer d_bitmap_table_;
    std::vector<const uint32_t*> bitmap_ptrs_;

    std::mutex mu_;
    std::condition_variable cv_submit_, cv_done_;
    std::deque<Request> requests_;  // ticket - first_ticket_ indexes it
    int64_t first_ticket_ = 0;
    std::deque<int64_t> pending_;
    int64_t inflight_bytes_ = 0;
    // queries in flight up to which batches of up to 64 take the spread search
    // (NSPARSE_SERVE_SPREAD_LOAD)
    int64_t spread_load_ = 128;
    ServerStats stats_;
    std::string error_;  // set if the dispatcher failed; every wait then reports it
    std::atomic<bool> stop_{false};
    std::thread worker_;
};

}  // namespace nsparse_amd
