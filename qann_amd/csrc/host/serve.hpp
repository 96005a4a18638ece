// A serving loop over one index: requests of any size, k, nprobe and filter
// queue up; worker threads admit them into searches as device memory allows,
// coalescing queued requests into shared launches (each query keeps its own
// k, nprobe and filter, qann_hip_search_requests), and run several launches
// at once, each on its own HIP stream and scratch. The reference has no such loop; its queries are synchronous calls
// (bindings/python_bindings.cpp:599-642), which one launch here reproduces
// bit for bit whatever else shares it.
#pragma once

#include "indexes.hpp"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <future>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace qh {

// One request: raw queries (the server prepares them as batch_query does),
// its k and nprobe (0: the index's own), and an optional slot filter, shared
// so that many requests can carry the same one.
struct ServeRequest {
    std::vector<float> queries; // nq * dim
    uint64_t nq = 0;
    uint32_t k = 10;
    uint32_t nprobe = 0;
    std::shared_ptr<const SlotFilter> filter;
};

struct ServeResult {
    std::vector<int64_t> ids; // nq * k, padded with -1
    std::vector<float> dists; // nq * k, padded with +inf
    double latency_ms = 0;    // from submit to the results being ready
    uint64_t launch_queries = 0; // queries of the launch the request ran in
};

struct ServeStats {
    uint64_t requests = 0, queries = 0, launches = 0;
    double p50_ms = 0, p90_ms = 0, p99_ms = 0, mean_ms = 0, max_ms = 0;
    double seconds = 0; // from the first submit to the last completion
};

class Server {
public:
    // workers: searches run at once; max_batch: queries one launch may take;
    // memory_fraction: the share of the GPU's free memory (at start) that the
    // launches running at once may hold as scratch.
    Server(std::shared_ptr<DenseIndex> index, int workers, uint64_t max_batch, double memory_fraction);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    std::future<ServeResult> submit(ServeRequest req);
    ServeStats stats(bool reset);
    uint64_t memory_budget() const { return budget_; }
    const DenseIndex& index() const { return *index_; }
    void close();

private:
    using Clock = std::chrono::steady_clock;
    struct Pending {
        ServeRequest req;
        std::promise<ServeResult> promise;
        Clock::time_point t0;
    };
    void run();
    void execute(std::vector<Pending>& group, uint64_t nq);
    uint64_t estimate(uint64_t nq, uint32_t k, uint32_t nprobe, bool filtered);

    std::shared_ptr<DenseIndex> index_;
    uint64_t max_batch_, budget_ = 0, inflight_ = 0;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<Pending> queue_;
    bool stop_ = false;
    std::vector<std::thread> workers_;
    // stats
    std::vector<double> latencies_;
    uint64_t queries_ = 0, launches_ = 0;
    Clock::time_point first_{}, last_{};
    bool any_ = false;
};

// An open-loop load for the serving configuration: requests arrive as a
// Poisson process of `rate` per second for `seconds`; each takes nq rows of
// the pool (log-uniform in [1, max_nq]), a k from ks and an nprobe from
// nprobes, and with probability filtered_fraction one of `filters`. Returns
// the server's statistics over exactly these requests.
struct LoadSpec {
    double rate = 100, seconds = 5;
    uint64_t seed = 1;
    uint32_t max_nq = 256;
    std::vector<uint32_t> ks{10}, nprobes{0};
    double filtered_fraction = 0;
    std::vector<std::shared_ptr<const SlotFilter>> filters;
};
ServeStats run_load(Server& server, const float* pool, uint64_t pool_rows, const LoadSpec& spec);

} // namespace qh
