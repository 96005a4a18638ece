// The serving loop of serve.hpp.
#include "serve.hpp"

#include "../../../include/qann_hip.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <shared_mutex>
#include <stdexcept>

namespace qh {

namespace {

// The index and the base it wraps, locked shared as queries lock them
// (lock_chain, bindings/python_bindings.cpp:97-113).
std::vector<std::shared_lock<std::shared_mutex>> read_chain(const DenseIndex& idx) {
    std::vector<std::shared_lock<std::shared_mutex>> locks;
    locks.emplace_back(idx.mu);
    if (auto* r = dynamic_cast<const RefineIndex*>(&idx)) locks.emplace_back(r->base()->mu);
    return locks;
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const size_t rank = static_cast<size_t>(std::ceil(q * static_cast<double>(v.size())));
    return v[std::min(v.size() - 1, rank == 0 ? 0 : rank - 1)];
}

} // namespace

Server::Server(std::shared_ptr<DenseIndex> index, int workers, uint64_t max_batch, double memory_fraction)
    : index_(std::move(index)), max_batch_(std::max<uint64_t>(max_batch, 1)) {
    if (!index_) throw std::invalid_argument("null index");
    if (workers < 1) throw std::invalid_argument("workers must be >= 1");
    if (!(memory_fraction > 0.0 && memory_fraction <= 1.0)) throw std::invalid_argument("memory_fraction must be in (0, 1]");
    uint64_t free_bytes = 0, total = 0;
    check_status(qann_hip_device_memory(&free_bytes, &total));
    budget_ = static_cast<uint64_t>(static_cast<double>(free_bytes) * memory_fraction);
    // an index that cannot be searched (an untrained IVF, an unsupported
    // scoring mode) raises here rather than in a worker
    (void)estimate(1, 1, 0, false);
    for (int w = 0; w < workers; ++w) workers_.emplace_back([this] { run(); });
}

Server::~Server() { close(); }

void Server::close() {
    {
        std::lock_guard<std::mutex> g(mu_);
        if (stop_ && workers_.empty()) return;
        stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
    workers_.clear();
}

std::future<ServeResult> Server::submit(ServeRequest req) {
    if (req.queries.size() != req.nq * index_->dim())
        throw std::invalid_argument("queries must hold nq rows of the index's dimension");
    Pending p{std::move(req), {}, Clock::now()};
    std::future<ServeResult> f = p.promise.get_future();
    {
        std::lock_guard<std::mutex> g(mu_);
        if (stop_) throw std::runtime_error("the server is closed");
        if (!any_) {
            first_ = p.t0;
            any_ = true;
        }
        queue_.push_back(std::move(p));
    }
    cv_.notify_one();
    return f;
}

uint64_t Server::estimate(uint64_t nq, uint32_t k, uint32_t nprobe, bool filtered) {
    auto locks = read_chain(*index_);
    return index_->scratch_bytes(nq, k, nprobe, filtered);
}

void Server::run() {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
        if (queue_.empty()) {
            if (stop_) return;
            continue;
        }
        const ServeRequest& head = queue_.front().req;
        uint64_t bytes = 0;
        try {
            bytes = estimate(head.nq, head.k, head.nprobe, head.filter != nullptr);
        } catch (...) {
            // the request cannot be planned: it fails alone, the loop goes on
            Pending bad = std::move(queue_.front());
            queue_.pop_front();
            bad.promise.set_exception(std::current_exception());
            continue;
        }
        // admit by free memory: a launch waits while others hold the budget
        if (inflight_ > 0 && inflight_ + bytes > budget_) {
            cv_.wait(lk);
            continue;
        }
        std::vector<Pending> group;
        group.push_back(std::move(queue_.front()));
        queue_.pop_front();
        uint64_t nq = group[0].req.nq;
        uint32_t k_max = group[0].req.k, np_max = group[0].req.nprobe;
        bool any_filter = group[0].req.filter != nullptr;
        // queued requests that share the launch, oldest first: any k, nprobe
        // and filter (each query keeps its own), while the launch's queries
        // stay within max_batch and its scratch within the budget
        for (auto it = queue_.begin(); it != queue_.end() && nq < max_batch_;) {
            const ServeRequest& r = it->req;
            if (nq + r.nq > max_batch_) {
                ++it;
                continue;
            }
            const uint32_t k2 = std::max(k_max, r.k), np2 = std::max(np_max, r.nprobe);
            const bool f2 = any_filter || r.filter != nullptr;
            uint64_t b = 0;
            try {
                b = estimate(nq + r.nq, k2, np2, f2);
            } catch (...) {
                break; // the launch keeps what it has; the rest is planned again later
            }
            if (inflight_ > 0 && inflight_ + b > budget_) break;
            nq += r.nq;
            k_max = k2;
            np_max = np2;
            any_filter = f2;
            bytes = b;
            group.push_back(std::move(*it));
            it = queue_.erase(it);
        }
        inflight_ += bytes;
        ++launches_;
        lk.unlock();
        execute(group, nq);
        lk.lock();
        inflight_ -= bytes;
        cv_.notify_all();
    }
}

void Server::execute(std::vector<Pending>& group, uint64_t nq) {
    const uint64_t d = index_->dim();
    std::vector<float> raw(nq * d);
    std::vector<DenseIndex::Request> reqs(group.size());
    std::vector<uint64_t> at(group.size() + 1, 0); // each request's first result entry
    uint64_t row = 0;
    for (size_t g = 0; g < group.size(); ++g) {
        const ServeRequest& r = group[g].req;
        std::copy(r.queries.begin(), r.queries.end(), raw.begin() + row * d);
        row += r.nq;
        reqs[g].nq = r.nq;
        reqs[g].k = r.k;
        reqs[g].nprobe = r.nprobe;
        reqs[g].filter = r.filter.get();
        at[g + 1] = at[g] + r.nq * r.k;
    }
    std::vector<int64_t> ids(at.back());
    std::vector<float> dists(at.back());
    try {
        auto locks = read_chain(*index_);
        const std::vector<float> q = index_->prepare(raw.data(), nq);
        index_->search_requests(q.data(), reqs.data(), reqs.size(), ids.data(), dists.data());
    } catch (...) {
        for (Pending& p : group) p.promise.set_exception(std::current_exception());
        return;
    }
    const Clock::time_point t1 = Clock::now();
    std::vector<double> lat;
    for (size_t g = 0; g < group.size(); ++g) {
        Pending& p = group[g];
        ServeResult r;
        r.ids.assign(ids.begin() + at[g], ids.begin() + at[g + 1]);
        r.dists.assign(dists.begin() + at[g], dists.begin() + at[g + 1]);
        r.latency_ms = std::chrono::duration<double, std::milli>(t1 - p.t0).count();
        r.launch_queries = nq;
        lat.push_back(r.latency_ms);
        p.promise.set_value(std::move(r));
    }
    std::lock_guard<std::mutex> g(mu_);
    latencies_.insert(latencies_.end(), lat.begin(), lat.end());
    queries_ += nq;
    last_ = std::max(last_, t1);
}

ServeStats Server::stats(bool reset) {
    std::lock_guard<std::mutex> g(mu_);
    ServeStats s;
    s.requests = latencies_.size();
    s.queries = queries_;
    s.launches = launches_;
    s.p50_ms = percentile(latencies_, 0.50);
    s.p90_ms = percentile(latencies_, 0.90);
    s.p99_ms = percentile(latencies_, 0.99);
    double sum = 0.0, mx = 0.0;
    for (double v : latencies_) {
        sum += v;
        mx = std::max(mx, v);
    }
    s.mean_ms = latencies_.empty() ? 0.0 : sum / static_cast<double>(latencies_.size());
    s.max_ms = mx;
    s.seconds = any_ ? std::chrono::duration<double>(last_ - first_).count() : 0.0;
    if (reset) {
        latencies_.clear();
        queries_ = launches_ = 0;
        any_ = false;
    }
    return s;
}

ServeStats run_load(Server& server, const float* pool, uint64_t pool_rows, const LoadSpec& spec) {
    if (spec.rate <= 0 || spec.seconds <= 0) throw std::invalid_argument("rate and seconds must be positive");
    if (spec.ks.empty() || spec.nprobes.empty()) throw std::invalid_argument("ks and nprobes must not be empty");
    if (spec.filtered_fraction > 0 && spec.filters.empty()) throw std::invalid_argument("filtered requests need filters");
    if (pool_rows == 0) throw std::invalid_argument("the query pool is empty");
    const uint64_t d = server.index().dim();
    const uint32_t max_nq = static_cast<uint32_t>(std::min<uint64_t>(std::max<uint32_t>(spec.max_nq, 1), pool_rows));
    std::mt19937_64 rng(spec.seed);
    std::exponential_distribution<double> gap(spec.rate);
    std::uniform_real_distribution<double> unit(0.0, 1.0);
    std::vector<std::future<ServeResult>> futures;
    server.stats(true);
    const auto start = std::chrono::steady_clock::now();
    double t = 0.0;
    for (;;) {
        t += gap(rng);
        if (t > spec.seconds) break;
        ServeRequest r;
        // nq log-uniform over [1, max_nq]
        r.nq = std::clamp<uint64_t>(static_cast<uint64_t>(std::lround(std::exp(unit(rng) * std::log(double(max_nq))))), 1,
                                    max_nq);
        r.k = spec.ks[rng() % spec.ks.size()];
        r.nprobe = spec.nprobes[rng() % spec.nprobes.size()];
        if (unit(rng) < spec.filtered_fraction) r.filter = spec.filters[rng() % spec.filters.size()];
        const uint64_t first = rng() % (pool_rows - r.nq + 1);
        r.queries.assign(pool + first * d, pool + (first + r.nq) * d);
        std::this_thread::sleep_until(start + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                                  std::chrono::duration<double>(t)));
        futures.push_back(server.submit(std::move(r)));
    }
    for (auto& f : futures) f.get();
    return server.stats(false);
}

} // namespace qh
