// Serving: continuous batching of position lookups over tables resident in
// HBM.
//
// Requests (table id, board) enter a queue.  A worker thread owns `slots`
// batch slots, each with its own HIP stream and pinned buffers.  Whenever a
// slot is free and requests are waiting, it takes up to max_batch waiting
// requests of one table, launches H2D + lookup kernel + D2H on its stream
// and moves on; finished slots are retired by polling their events.  New
// arrivals therefore never wait for the batches already on the GPU, and the
// batches of several slots overlap on the device.  Tables are admitted
// against a byte budget.  A Poisson load generator measures QPS and latency
// percentiles from submit to result visible to the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <random>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "egtb.h"
#include "egtb_runtime.h"
#include "egtb_serving.h"

namespace egtb {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket;
    uint64_t board;
    int32_t table;
};

struct Completion {
    uint64_t ticket;
    float out4[4];
    int32_t status;
    Clock::time_point done;
};

class Server {
  public:
    Server(const egtb_server_options &o);
    ~Server();

    int32_t attach(TableSet *t);
    void detach(int32_t id);
    void submit(int32_t table, const uint64_t *boards, size_t n, uint64_t *tickets);
    size_t poll(size_t max, uint64_t *tickets, float *out4, int32_t *status, int wait_us,
                std::vector<Clock::time_point> *done_times = nullptr);
    void stats(uint64_t *batches, uint64_t *queries, uint64_t *max_batch) const {
        *batches = batches_.load();
        *queries = queries_.load();
        *max_batch = max_batch_seen_.load();
    }
    uint64_t next_ticket() {
        std::lock_guard<std::mutex> lk(mu_);
        return next_ticket_;
    }

  private:
    struct Slot {
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        uint64_t *h_boards = nullptr;
        float *h_out = nullptr;
        int32_t *h_status = nullptr;
        uint64_t *d_boards = nullptr;
        float *d_out = nullptr;
        int32_t *d_status = nullptr;
        std::vector<uint64_t> tickets;
        int32_t table = -1;
        bool busy = false;
    };
    void run();
    bool launch_slot(Slot &s);
    void retire_slot(Slot &s);
    // answer tickets with EGTB_LOOKUP_ERROR (NaN rates): a batch that could
    // not run, or requests left queued when the worker stopped
    void complete_with_error(const uint64_t *tickets, size_t n);

    int dev_;
    size_t max_batch_;
    int64_t budget_;
    std::vector<Slot> slots_;
    std::mutex mu_;  // queue, tables
    std::condition_variable cv_;
    std::deque<Request> queue_;
    std::map<int32_t, TableSet *> tables_;
    std::map<int32_t, int> inflight_;  // batches in flight per table
    std::set<int32_t> detaching_;      // tables that refuse new requests
    std::string fail_msg_;             // why the worker stopped (under mu_)
    bool failed_ = false;              // the worker stopped on an error (under mu_)
    int32_t next_table_ = 0;
    int64_t table_bytes_ = 0;
    uint64_t next_ticket_ = 1;
    std::mutex cmu_;  // completions
    std::condition_variable ccv_;
    std::deque<Completion> done_;
    std::atomic<bool> stop_{false};
    std::atomic<uint64_t> batches_{0}, queries_{0}, max_batch_seen_{0};
    std::thread worker_;
};

Server::Server(const egtb_server_options &o)
    : dev_(o.device), max_batch_(o.max_batch ? o.max_batch : 4096), budget_(o.memory_budget) {
    DeviceCtx &C = device_ctx(dev_);
    (void)C;
    configure_serving_kernels();
    const int n = o.slots > 0 ? o.slots : 4;
    slots_.resize(n);
    for (Slot &s : slots_) {
        HIPCHK(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
        HIPCHK(hipHostMalloc((void **)&s.h_boards, max_batch_ * 8, hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void **)&s.h_out, max_batch_ * 16, hipHostMallocDefault));
        HIPCHK(hipHostMalloc((void **)&s.h_status, max_batch_ * 4, hipHostMallocDefault));
        HIPCHK(hipMalloc((void **)&s.d_boards, max_batch_ * 8));
        HIPCHK(hipMalloc((void **)&s.d_out, max_batch_ * 16));
        HIPCHK(hipMalloc((void **)&s.d_status, max_batch_ * 4));
        s.tickets.reserve(max_batch_);
    }
    worker_ = std::thread([this] { run(); });
}

Server::~Server() {
    stop_.store(true);
    cv_.notify_all();
    if (worker_.joinable()) worker_.join();
    (void)hipSetDevice(dev_);
    for (Slot &s : slots_) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        (void)hipFree(s.d_boards);
        (void)hipFree(s.d_out);
        (void)hipFree(s.d_status);
        (void)hipHostFree(s.h_boards);
        (void)hipHostFree(s.h_out);
        (void)hipHostFree(s.h_status);
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.stream) (void)hipStreamDestroy(s.stream);
    }
}

int32_t Server::attach(TableSet *t) {
    if (!t) throw Error(EGTB_ERR_ARG, "table is NULL");
    if (t->dev != dev_) throw Error(EGTB_ERR_ARG, "table lives on another device");
    std::lock_guard<std::mutex> lk(mu_);
    if (budget_ > 0 && table_bytes_ + (int64_t)t->bytes > budget_)
        throw Error(EGTB_ERR_MEMORY, "table of " + std::to_string(t->bytes) + " bytes exceeds the serving budget (" +
                                         std::to_string(budget_ - table_bytes_) + " bytes free)");
    const int32_t id = next_table_++;
    tables_[id] = t;
    table_bytes_ += (int64_t)t->bytes;
    return id;
}

void Server::detach(int32_t id) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = tables_.find(id);
    if (it == tables_.end() || detaching_.count(id)) throw Error(EGTB_ERR_ARG, "unknown table id");
    for (const Request &r : queue_)
        if (r.table == id) throw Error(EGTB_ERR_ARG, "table has queued requests");
    // from here on submit refuses the table; wait for its batches in flight
    // (a failed batch gives its count back, and the worker retires every
    // busy slot when it stops)
    detaching_.insert(id);
    while (inflight_[id] > 0) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(50));
        lk.lock();
    }
    table_bytes_ -= (int64_t)tables_.at(id)->bytes;
    tables_.erase(id);
    inflight_.erase(id);
    detaching_.erase(id);
}

void Server::submit(int32_t table, const uint64_t *boards, size_t n, uint64_t *tickets) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_.load())
            throw Error(failed_ ? EGTB_ERR_HIP : EGTB_ERR_ARG,
                        "the serving worker has stopped" + (failed_ ? ": " + fail_msg_ : std::string()));
        if (!tables_.count(table) || detaching_.count(table)) throw Error(EGTB_ERR_ARG, "unknown table id");
        for (size_t i = 0; i < n; ++i) {
            const uint64_t t = next_ticket_++;
            queue_.push_back(Request{t, boards[i], table});
            if (tickets) tickets[i] = t;
        }
    }
    cv_.notify_one();
}

void Server::complete_with_error(const uint64_t *tickets, size_t n) {
    if (!n) return;
    const auto now = Clock::now();
    {
        std::lock_guard<std::mutex> lk(cmu_);
        for (size_t i = 0; i < n; ++i) {
            Completion c;
            c.ticket = tickets[i];
            for (float &v : c.out4) v = NAN;
            c.status = EGTB_LOOKUP_ERROR;
            c.done = now;
            done_.push_back(c);
        }
    }
    ccv_.notify_all();
}

bool Server::launch_slot(Slot &s) {
    TableSet *ts = nullptr;
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (queue_.empty()) return false;
        const int32_t table = queue_.front().table;
        s.tickets.clear();
        size_t m = 0;
        // take up to max_batch requests of the oldest request's table, in order
        for (auto it = queue_.begin(); it != queue_.end() && m < max_batch_;) {
            if (it->table == table) {
                s.h_boards[m++] = it->board;
                s.tickets.push_back(it->ticket);
                it = queue_.erase(it);
            } else {
                ++it;
            }
        }
        s.table = table;
        auto t = tables_.find(table);
        if (t != tables_.end()) {
            ts = t->second;
            inflight_[table]++;
        }
    }
    if (!ts) {  // the table went away while its requests waited
        complete_with_error(s.tickets.data(), s.tickets.size());
        return true;
    }
    const size_t m = s.tickets.size();
    try {
        HIPCHK(hipMemcpyAsync(s.d_boards, s.h_boards, m * 8, hipMemcpyHostToDevice, s.stream));
        launch_lookup<float>(s.stream, *ts, s.d_boards, m, s.d_out, s.d_status, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(s.h_out, s.d_out, m * 16, hipMemcpyDeviceToHost, s.stream));
        HIPCHK(hipMemcpyAsync(s.h_status, s.d_status, m * 4, hipMemcpyDeviceToHost, s.stream));
        HIPCHK(hipEventRecord(s.done, s.stream));
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(mu_);
            inflight_[s.table]--;
        }
        complete_with_error(s.tickets.data(), s.tickets.size());
        throw;
    }
    s.busy = true;
    batches_++;
    queries_ += m;
    uint64_t prev = max_batch_seen_.load();
    while (m > prev && !max_batch_seen_.compare_exchange_weak(prev, m)) {
    }
    return true;
}

void Server::retire_slot(Slot &s) {
    const auto now = Clock::now();
    const size_t m = s.tickets.size();
    {
        std::lock_guard<std::mutex> lk(cmu_);
        for (size_t i = 0; i < m; ++i) {
            Completion c;
            c.ticket = s.tickets[i];
            std::memcpy(c.out4, s.h_out + 4 * i, 16);
            c.status = s.h_status[i];
            c.done = now;
            done_.push_back(c);
        }
    }
    ccv_.notify_all();
    {
        std::lock_guard<std::mutex> lk(mu_);
        inflight_[s.table]--;
    }
    s.busy = false;
}

void Server::run() {
    (void)hipSetDevice(dev_);
    while (!stop_.load()) {
        bool did = false;
        for (Slot &s : slots_) {
            if (s.busy && hipEventQuery(s.done) == hipSuccess) {
                retire_slot(s);
                did = true;
            }
        }
        for (Slot &s : slots_) {
            if (!s.busy) {
                try {
                    if (launch_slot(s)) did = true;
                } catch (const std::exception &e) {
                    std::fprintf(stderr, "egtb server: batch launch failed: %s\n", e.what());
                    std::lock_guard<std::mutex> lk(mu_);
                    failed_ = true;
                    fail_msg_ = e.what();
                    stop_.store(true);
                }
            }
        }
        if (!did) {
            bool any_busy = false;
            for (Slot &s : slots_) any_busy |= s.busy;
            if (any_busy) {
                std::this_thread::yield();
            } else {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait_for(lk, std::chrono::microseconds(200), [this] { return !queue_.empty() || stop_.load(); });
            }
        }
    }
    for (Slot &s : slots_)
        if (s.busy) {
            (void)hipEventSynchronize(s.done);
            retire_slot(s);
        }
    // requests still queued get an error answer; submit refuses new ones
    // from now on (it checks stop_ under mu_, which was set before this)
    std::vector<uint64_t> left;
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (const Request &r : queue_) left.push_back(r.ticket);
        queue_.clear();
    }
    complete_with_error(left.data(), left.size());
}

size_t Server::poll(size_t max, uint64_t *tickets, float *out4, int32_t *status, int wait_us,
                    std::vector<Clock::time_point> *done_times) {
    std::unique_lock<std::mutex> lk(cmu_);
    if (done_.empty() && wait_us > 0)
        ccv_.wait_for(lk, std::chrono::microseconds(wait_us), [this] { return !done_.empty(); });
    if (done_.empty() && stop_.load()) {
        lk.unlock();
        std::lock_guard<std::mutex> g(mu_);
        if (failed_) throw Error(EGTB_ERR_HIP, "the serving worker has stopped: " + fail_msg_);
        return 0;
    }
    size_t k = 0;
    while (k < max && !done_.empty()) {
        const Completion &c = done_.front();
        if (tickets) tickets[k] = c.ticket;
        if (out4) std::memcpy(out4 + 4 * k, c.out4, 16);
        if (status) status[k] = c.status;
        if (done_times) done_times->push_back(c.done);
        done_.pop_front();
        ++k;
    }
    return k;
}

// Poisson arrivals from n_games sources; each query's board is drawn from
// the pool by a seeded uniform index.  Returns the latency distribution.
static void loadgen(Server &srv, const egtb_loadgen_options &o, egtb_loadgen_report &rep) {
    if (!o.pool || !o.pool_size) throw Error(EGTB_ERR_ARG, "empty query pool");
    if (!(o.qps > 0) || !(o.duration_s > 0)) throw Error(EGTB_ERR_ARG, "qps and duration must be positive");
    const uint64_t total = (uint64_t)std::llround(o.qps * o.duration_s);
    std::mt19937_64 rng(o.seed);
    std::exponential_distribution<double> gap(o.qps);
    std::uniform_int_distribution<uint64_t> pick(0, o.pool_size - 1);
    std::vector<Clock::time_point> t_submit(total);
    std::vector<double> lat;
    lat.reserve(total);
    // tickets are sequential; the generator is the only submitter meanwhile
    const uint64_t first_ticket = srv.next_ticket();
    uint64_t fallbacks = 0;
    std::atomic<bool> gen_done{false}, stalled{false};
    std::thread collector([&] {
        std::vector<uint64_t> tk(4096);
        std::vector<int32_t> st(4096);
        std::vector<Clock::time_point> dt;
        uint64_t got = 0;
        auto last_progress = Clock::now();
        while (got < total) {
            dt.clear();
            const size_t k = srv.poll(tk.size(), tk.data(), nullptr, st.data(), 1000, &dt);
            if (k) last_progress = Clock::now();
            else if (gen_done.load() && Clock::now() - last_progress > std::chrono::seconds(10)) {
                stalled.store(true);
                break;
            }
            for (size_t i = 0; i < k; ++i) {
                const uint64_t idx = tk[i] - first_ticket;
                if (idx < total) {
                    lat.push_back(std::chrono::duration<double, std::milli>(dt[i] - t_submit[idx]).count());
                    fallbacks += st[i] == EGTB_LOOKUP_FALLBACK;
                }
            }
            got += k;
            if (k == 0 && gen_done.load() && got >= total) break;
        }
    });
    const auto t0 = Clock::now();
    double t_next = 0.0;
    uint64_t tkt = 0;
    for (uint64_t i = 0; i < total; ++i) {
        t_next += gap(rng);
        const auto due = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(t_next));
        while (Clock::now() < due) {
        }
        const uint64_t board = o.pool[pick(rng)];
        // stamp before the request becomes visible to the worker
        t_submit[i] = Clock::now();
        srv.submit(o.table_id, &board, 1, &tkt);
    }
    const auto t_gen = Clock::now();
    gen_done.store(true);
    collector.join();
    const auto t_end = Clock::now();
    if (stalled.load()) throw Error(EGTB_ERR_INTERNAL, "load generator: completions stopped arriving");
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) {
        if (lat.empty()) return 0.0;
        size_t k = (size_t)std::ceil(q * (double)lat.size()) ;
        k = k == 0 ? 0 : k - 1;
        return lat[std::min(k, lat.size() - 1)];
    };
    std::memset(&rep, 0, sizeof rep);
    rep.qps_offered = o.qps;
    rep.completed = lat.size();
    rep.qps_achieved = (double)lat.size() / std::chrono::duration<double>(t_end - t0).count();
    rep.p50_ms = pct(0.50);
    rep.p90_ms = pct(0.90);
    rep.p99_ms = pct(0.99);
    rep.p999_ms = pct(0.999);
    rep.max_ms = lat.empty() ? 0.0 : lat.back();
    double sum = 0;
    for (double v : lat) sum += v;
    rep.mean_ms = lat.empty() ? 0.0 : sum / (double)lat.size();
    rep.fallbacks = fallbacks;
    rep.submit_seconds = std::chrono::duration<double>(t_gen - t0).count();
    uint64_t b, q, mb;
    srv.stats(&b, &q, &mb);
    rep.batches = b;
    rep.max_batch = mb;
}

}  // namespace egtb

// ===========================================================================
// C ABI
// ===========================================================================

using namespace egtb;

struct egtb_server {
    std::unique_ptr<Server> s;
};

template <typename F>
static int guarded_v(F &&f) {
    try {
        g_last_error.clear();
        f();
        return EGTB_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return EGTB_ERR_INTERNAL;
    }
}

extern "C" {

int egtb_server_create(const egtb_server_options *o, egtb_server **out) {
    if (!o || !out) return EGTB_ERR_ARG;
    *out = nullptr;
    return guarded_v([&] {
        auto sv = std::make_unique<egtb_server>();
        sv->s = std::make_unique<Server>(*o);
        *out = sv.release();
    });
}

int egtb_server_attach(egtb_server *sv, egtb_table *t, int32_t *table_id) {
    return guarded_v([&] {
        if (!sv || !table_id) throw Error(EGTB_ERR_ARG, "NULL argument");
        *table_id = sv->s->attach(reinterpret_cast<TableSet *>(t));
    });
}

int egtb_server_detach(egtb_server *sv, int32_t table_id) {
    return guarded_v([&] {
        if (!sv) throw Error(EGTB_ERR_ARG, "server is NULL");
        sv->s->detach(table_id);
    });
}

int egtb_server_submit(egtb_server *sv, int32_t table_id, const uint64_t *boards, size_t n, uint64_t *tickets) {
    return guarded_v([&] {
        if (!sv || (!boards && n)) throw Error(EGTB_ERR_ARG, "NULL argument");
        sv->s->submit(table_id, boards, n, tickets);
    });
}

int egtb_server_poll(egtb_server *sv, size_t max, uint64_t *tickets, float *out4, int32_t *status, size_t *n_out,
                     int32_t wait_us) {
    return guarded_v([&] {
        if (!sv || !n_out) throw Error(EGTB_ERR_ARG, "NULL argument");
        *n_out = sv->s->poll(max, tickets, out4, status, wait_us);
    });
}

int egtb_server_stats(const egtb_server *sv, uint64_t *batches, uint64_t *queries, uint64_t *max_batch) {
    if (!sv || !batches || !queries || !max_batch) return EGTB_ERR_ARG;
    sv->s->stats(batches, queries, max_batch);
    return EGTB_OK;
}

int egtb_server_loadgen(egtb_server *sv, const egtb_loadgen_options *o, egtb_loadgen_report *rep) {
    return guarded_v([&] {
        if (!sv || !o || !rep) throw Error(EGTB_ERR_ARG, "NULL argument");
        loadgen(*sv->s, *o, *rep);
    });
}

void egtb_server_destroy(egtb_server *sv) { delete sv; }

}  // extern "C"
