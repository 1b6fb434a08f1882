// server.hip -- continuous-batching LP server behind the C ABI (cpdlp_server_*).
//
// No counterpart in the reference, which solves one LP per `solve` call
// (src/algorithms/common.jl:186-209); this is the served form of that same solve.
// Requests are copied at submit (or borrowed: cpdlp_server_submit_borrowed).  Small
// ones (m + n <= batch_limit) go to a batch engine (cpdlp_batch_*: one workgroup per LP,
// iterates and restart state resident in its HBM arena between rounds), medium ones
// (m + n <= batch2_limit) to a second batch engine with rounds of its own, so that their
// longer rounds never hold up the small ones; both admit as arena room frees up.  Larger
// ones go to plain engines (cpdlp_create / cpdlp_solve, one HIP stream each) run by worker
// threads, in lanes by nonzero count with their own shortest-job-first queues, under an
// HBM budget.  The caller drives the batch routes with cpdlp_server_pump, or the library
// does on one thread per route after cpdlp_server_start; plain workers run on their own.
// Finished requests are reported by cpdlp_server_poll and handed over (and freed) by
// cpdlp_server_get.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>

#include "common.hpp"
#include "layout.hpp"

namespace cpdlp {
namespace {

double snow() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// HBM of one plain-engine solve (matrices in two layouts, FP64 and FP32 values,
// vectors and iterates), for the plain route's budget
int64_t plain_footprint(int64_t m, int64_t n, int64_t nnz) { return 40 * nnz + 250 * (m + n); }

struct Request {
    int64_t id = 0, m = 0, n = 0, nnz = 0;
    std::vector<int32_t> colptr, rowval;
    std::vector<double> nzval, c, lv, uv, lc, uc, x0, y0;
    bool warm = false;
    bool borrowed = false;         // the caller's arrays, valid until the request is handed over
    cpdlp_lp blp{};
    const double *bx0 = nullptr, *by0 = nullptr;
    int lane = -1;                 // -1: batch route
    double t_submit = 0.0, t_done = 0.0;
    // result
    int32_t code = CPDLP_OK;
    std::string message;
    std::vector<double> x, y;
    cpdlp_stats stats{};

    cpdlp_lp view() const {
        if (borrowed) return blp;
        cpdlp_lp lp{};
        lp.m = m;
        lp.n = n;
        lp.index_base = 0;
        lp.index_bits = 32;
        lp.colptr = colptr.data();
        lp.rowval = rowval.data();
        lp.nzval = nzval.data();
        lp.c = c.data();
        lp.lv = lv.data();
        lp.uv = uv.data();
        lp.lc = lc.data();
        lp.uc = uc.data();
        return lp;
    }
    const double* start_x() const { return borrowed ? bx0 : warm ? x0.data() : nullptr; }
    const double* start_y() const { return borrowed ? by0 : warm ? y0.data() : nullptr; }
    void drop_input() {
        std::vector<int32_t>().swap(colptr);
        std::vector<int32_t>().swap(rowval);
        for (auto* v : {&nzval, &c, &lv, &uv, &lc, &uc, &x0, &y0}) std::vector<double>().swap(*v);
    }
};

struct Lane {
    double max_nnz = 0.0;
    int workers = 1;
    std::deque<Request*> waiting;  // shortest job first at admission
    std::deque<Request*> ready;    // admitted, not yet taken by a worker
    int running = 0;               // admitted and not finished
    std::vector<std::thread> threads;
};

// one batch engine and its queues
struct BatchRoute {
    cpdlp_batch* b = nullptr;
    int32_t limit = 0;             // m + n upper bound of its requests
    int32_t checks = 20;           // checks per round for its smallest instances
    int32_t code = -1;             // the route cpdlp_server_get reports: -1 or -2
    std::deque<Request*> waiting;
    std::unordered_map<int64_t, Request*> in_batch;   // batch id -> request
    int64_t admitting = 0;         // taken off waiting by a pump, not yet placed
};

}  // namespace

struct Server {
    cpdlp_params P;
    cpdlp_server_config cfg;
    BatchRoute routes[2];          // small (-1) and medium (-2) requests
    std::mutex mu;
    std::condition_variable cv;
    bool stop = false;
    int64_t next_id = 1;
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    std::vector<std::unique_ptr<Lane>> lanes;   // Lane holds threads: not movable
    int64_t plain_used = 0;
    std::deque<int64_t> done;
    // the optional driver threads (cpdlp_server_start), one per batch route
    std::thread drivers[2];
    bool driving = false;
    int32_t driver_code = CPDLP_OK;
    std::string driver_error;

    Server(const cpdlp_params& p, const cpdlp_server_config& c) : P(p), cfg(c) {
        if (cfg.batch_limit < 0 || cfg.batch_limit > 65536)
            throw Error(CPDLP_ERR_ARGUMENT, "batch_limit must be in [0, 65536]");
        if (cfg.checks_per_round <= 0) throw Error(CPDLP_ERR_ARGUMENT, "checks_per_round must be positive");
        if (cfg.n_lanes < 1 || cfg.n_lanes > CPDLP_SERVER_MAX_LANES)
            throw Error(CPDLP_ERR_ARGUMENT, "n_lanes must be in [1, CPDLP_SERVER_MAX_LANES]");
        if (cfg.plain_bytes <= 0) throw Error(CPDLP_ERR_ARGUMENT, "plain_bytes must be positive");
        for (int k = 0; k < cfg.n_lanes; ++k)
            if (cfg.lane_workers[k] < 1 || cfg.lane_workers[k] > 16)
                throw Error(CPDLP_ERR_ARGUMENT, "lane_workers must be in [1, 16]");
        if (cfg.batch2_limit < 0 || cfg.batch2_limit > 65536)
            throw Error(CPDLP_ERR_ARGUMENT, "batch2_limit must be in [0, 65536]");
        if (cfg.batch2_checks < 0) throw Error(CPDLP_ERR_ARGUMENT, "batch2_checks must not be negative");
        const int32_t limits[2] = {cfg.batch_limit, cfg.batch2_limit > cfg.batch_limit ? cfg.batch2_limit : 0};
        const int64_t bytes[2] = {cfg.batch_bytes, cfg.batch2_bytes};
        for (int k = 0; k < 2; ++k) {
            BatchRoute& br = routes[k];
            br.limit = limits[k];
            br.checks = k == 1 && cfg.batch2_checks > 0 ? cfg.batch2_checks : cfg.checks_per_round;
            br.code = -1 - k;
            if (br.limit <= 0) continue;
            const int rc = cpdlp_batch_create(&P, bytes[k], &br.b);
            if (rc != CPDLP_OK) {   // no destructor runs for a constructor that throws
                const std::string msg = std::string("batch route: ") + cpdlp_last_error();
                for (BatchRoute& made : routes)
                    if (made.b) cpdlp_batch_destroy(made.b);
                throw Error(rc, msg);
            }
        }
        for (int k = 0; k < cfg.n_lanes; ++k) {
            lanes.push_back(std::make_unique<Lane>());
            lanes[k]->max_nnz = cfg.lane_max_nnz[k];
            lanes[k]->workers = cfg.lane_workers[k];
        }
        for (int k = 0; k < cfg.n_lanes; ++k)
            for (int w = 0; w < lanes[k]->workers; ++w) lanes[k]->threads.emplace_back([this, k] { work(k); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> g(mu);
            stop = true;
        }
        cv.notify_all();
        for (auto& d : drivers)
            if (d.joinable()) d.join();
        for (auto& ln : lanes)
            for (auto& t : ln->threads) t.join();
        for (auto& br : routes)
            if (br.b) cpdlp_batch_destroy(br.b);
    }

    // one plain solve on a worker thread; fixed CSR format, so that a request's result
    // does not depend on the autotuner's timing
    void work(int k) {
        Lane& ln = *lanes[k];
        for (;;) {
            Request* r;
            {
                std::unique_lock<std::mutex> g(mu);
                cv.wait(g, [&] { return stop || !ln.ready.empty(); });
                if (ln.ready.empty()) return;   // stop with nothing left to run
                r = ln.ready.front();
                ln.ready.pop_front();
            }
            cpdlp_params p = P;
            p.format = CPDLP_FMT_CSR;
            const cpdlp_lp lp = r->view();
            cpdlp_handle* h = nullptr;
            int rc = CPDLP_OK;
            try {
                rc = cpdlp_create(&lp, &p, &h);
                if (rc == CPDLP_OK)
                    rc = cpdlp_solve(h, r->start_x(), r->start_y());
                if (rc == CPDLP_OK) {
                    r->x.assign((size_t)r->n, 0.0);
                    r->y.assign((size_t)r->m, 0.0);
                    rc = cpdlp_get_solution(h, r->x.data(), r->y.data());
                }
                if (rc == CPDLP_OK) rc = cpdlp_get_stats(h, &r->stats);
                if (rc != CPDLP_OK) r->message = cpdlp_last_error();
            } catch (const std::bad_alloc&) {
                rc = CPDLP_ERR_MEMORY;
                r->message = "host allocation failed";
            }
            if (h) cpdlp_destroy(h);
            r->code = rc;
            r->drop_input();
            std::lock_guard<std::mutex> g(mu);
            r->t_done = snow();
            ln.running -= 1;
            plain_used -= plain_footprint(r->m, r->n, r->nnz);
            done.push_back(r->id);
        }
    }

    int route(const Request& r) const {
        for (const BatchRoute& br : routes)
            if (br.b && r.m + r.n <= br.limit) return br.code;
        for (int k = 0; k < (int)lanes.size(); ++k)
            if ((double)r.nnz <= lanes[k]->max_nnz) return k;
        return (int)lanes.size() - 1;
    }

    int64_t submit(const cpdlp_lp& lp, const double* x0, const double* y0, bool borrow) {
        auto r = std::make_unique<Request>();
        if (borrow) {
            // nothing is copied: the engines validate the arrays when they take the request
            if (lp.m < 0 || lp.n < 0) throw Error(CPDLP_ERR_DIMENSION, "negative matrix size");
            if (lp.index_base != 0 && lp.index_base != 1) throw Error(CPDLP_ERR_ARGUMENT, "index_base must be 0 or 1");
            if (lp.index_bits != 32 && lp.index_bits != 64) throw Error(CPDLP_ERR_ARGUMENT, "index_bits must be 32 or 64");
            if (!lp.colptr) throw Error(CPDLP_ERR_ARGUMENT, "colptr is NULL");
            r->borrowed = true;
            r->blp = lp;
            r->bx0 = x0;
            r->by0 = y0;
            r->m = lp.m;
            r->n = lp.n;
            r->nnz = (lp.index_bits == 32 ? (int64_t)((const int32_t*)lp.colptr)[lp.n]
                                          : ((const int64_t*)lp.colptr)[lp.n]) - lp.index_base;
            r->lane = route(*r);
            r->t_submit = snow();
            std::lock_guard<std::mutex> g(mu);
            r->id = next_id++;
            Request* p = r.get();
            reqs[p->id] = std::move(r);
            enqueue_locked(p);
            return p->id;
        }
        read_csc(lp.m, lp.n, lp.index_base, lp.index_bits, lp.colptr, lp.rowval, r->colptr, r->rowval);
        r->m = lp.m;
        r->n = lp.n;
        r->nnz = r->colptr.empty() ? 0 : r->colptr.back();
        if (r->nnz > 0 && !lp.nzval) throw Error(CPDLP_ERR_ARGUMENT, "nzval is NULL");
        for (const void* v : {(const void*)lp.c, (const void*)lp.lv, (const void*)lp.uv})
            if (r->n > 0 && !v) throw Error(CPDLP_ERR_ARGUMENT, "NULL variable vector");
        for (const void* v : {(const void*)lp.lc, (const void*)lp.uc})
            if (r->m > 0 && !v) throw Error(CPDLP_ERR_ARGUMENT, "NULL constraint vector");
        auto copy = [](std::vector<double>& dst, const double* src, int64_t len) {
            dst.assign((size_t)len, 0.0);
            if (src && len) std::memcpy(dst.data(), src, (size_t)len * 8);
        };
        copy(r->nzval, lp.nzval, r->nnz);
        copy(r->c, lp.c, r->n);
        copy(r->lv, lp.lv, r->n);
        copy(r->uv, lp.uv, r->n);
        copy(r->lc, lp.lc, r->m);
        copy(r->uc, lp.uc, r->m);
        if (x0 || y0) {   // a missing half of the start is zero, as in cpdlp_solve
            r->warm = true;
            copy(r->x0, x0, r->n);
            copy(r->y0, y0, r->m);
        }
        r->lane = route(*r);
        r->t_submit = snow();
        std::lock_guard<std::mutex> g(mu);
        r->id = next_id++;
        Request* p = r.get();
        reqs[p->id] = std::move(r);
        enqueue_locked(p);
        return p->id;
    }

    void enqueue_locked(Request* p) {
        if (p->lane < 0) routes[-1 - p->lane].waiting.push_back(p);
        else lanes[p->lane]->waiting.push_back(p);
    }

    // plain admission (caller holds mu): shortest job first within a lane; a job waits
    // while it would pass the HBM budget and another plain job is running
    void admit_plain() {
        bool woke = false;
        for (auto& lp : lanes) {
            Lane& ln = *lp;
            std::stable_sort(ln.waiting.begin(), ln.waiting.end(),
                             [](const Request* a, const Request* b) { return a->nnz < b->nnz; });
            while (!ln.waiting.empty() && ln.running < ln.workers) {
                Request* r = ln.waiting.front();
                const int64_t fp = plain_footprint(r->m, r->n, r->nnz);
                int busy = 0;
                for (auto& l2 : lanes) busy += l2->running;
                if (busy > 0 && plain_used + fp > cfg.plain_bytes) break;
                ln.waiting.pop_front();
                ln.running += 1;
                plain_used += fp;
                ln.ready.push_back(r);
                woke = true;
            }
        }
        if (woke) cv.notify_all();
    }

    int64_t pending_locked() const {
        // admitting may count a request twice for a moment (never zero times)
        int64_t k = 0;
        for (const BatchRoute& br : routes) k += (int64_t)br.waiting.size() + (int64_t)br.in_batch.size() + br.admitting;
        for (auto& ln : lanes) k += (int64_t)ln->waiting.size() + ln->running;
        return k;
    }

    // one round of a batch route: admit, advance by its checks per round, harvest
    void pump_route(BatchRoute& br) {
        if (!br.b) return;
        std::deque<Request*> to_batch;
        {
            std::lock_guard<std::mutex> g(mu);
            to_batch.swap(br.waiting);
            br.admitting = (int64_t)to_batch.size();
        }
        // batch admission as arena room allows; the rest waits, in order
        std::deque<Request*> left;
        while (!to_batch.empty()) {
            Request* r = to_batch.front();
            if (!left.empty()) {   // keep arrival order behind a request that did not fit
                left.push_back(r);
                to_batch.pop_front();
                continue;
            }
            const cpdlp_lp lp = r->view();
            int64_t bid = 0;
            const int rc = cpdlp_batch_add_start(br.b, &lp, r->start_x(), r->start_y(), &bid);
            to_batch.pop_front();
            if (rc == CPDLP_ERR_MEMORY) {
                left.push_back(r);
            } else if (rc != CPDLP_OK) {   // invalid for the batch engine: reported, not fatal
                r->code = rc;
                r->message = cpdlp_last_error();
                r->drop_input();
                std::lock_guard<std::mutex> g(mu);
                r->t_done = snow();
                done.push_back(r->id);
            } else {
                r->drop_input();
                std::lock_guard<std::mutex> g(mu);
                br.in_batch[bid] = r;
            }
        }
        bool any;
        {
            std::lock_guard<std::mutex> g(mu);
            for (auto it = left.rbegin(); it != left.rend(); ++it) br.waiting.push_front(*it);
            br.admitting = 0;
            any = !br.in_batch.empty();
        }
        if (!any) return;
        int64_t running = 0;
        int rc = cpdlp_batch_run(br.b, br.checks, &running);
        if (rc != CPDLP_OK) throw Error(rc, std::string("batch route: ") + cpdlp_last_error());
        std::vector<int64_t> ids(64);
        for (;;) {
            int64_t cnt = 0;
            rc = cpdlp_batch_poll(br.b, ids.data(), (int64_t)ids.size(), &cnt);
            if (rc != CPDLP_OK) throw Error(rc, std::string("batch route: ") + cpdlp_last_error());
            if (cnt == 0) break;
            const double now = snow();
            for (int64_t k = 0; k < cnt; ++k) {
                Request* r;
                {
                    std::lock_guard<std::mutex> g(mu);
                    r = br.in_batch.at(ids[k]);
                }
                r->x.assign((size_t)r->n, 0.0);
                r->y.assign((size_t)r->m, 0.0);
                rc = cpdlp_batch_get(br.b, ids[k], r->x.data(), r->y.data(), &r->stats);
                if (rc == CPDLP_OK) rc = cpdlp_batch_release(br.b, ids[k]);
                if (rc != CPDLP_OK) throw Error(rc, std::string("batch route: ") + cpdlp_last_error());
                std::lock_guard<std::mutex> g(mu);
                br.in_batch.erase(ids[k]);
                r->t_done = now;
                done.push_back(r->id);
            }
        }
    }

    // one scheduling round of every route (the caller's cpdlp_server_pump)
    int64_t pump() {
        {
            std::lock_guard<std::mutex> g(mu);
            admit_plain();
        }
        for (BatchRoute& br : routes) pump_route(br);
        std::lock_guard<std::mutex> g(mu);
        admit_plain();   // plain jobs that finished during the round free budget
        return pending_locked();
    }

    // pumps batch route k (route 0 also admits plain jobs) until the server is destroyed;
    // idles 200 us when the route has nothing to run.  A failure stops every driver and
    // is reported by every later call.
    void drive(int k) {
        BatchRoute& br = routes[k];
        for (;;) {
            {
                std::lock_guard<std::mutex> g(mu);
                if (stop || driver_code != CPDLP_OK) return;
                if (k == 0) admit_plain();
            }
            try {
                pump_route(br);
            } catch (const Error& e) {
                std::lock_guard<std::mutex> g(mu);
                driver_code = e.code;
                driver_error = e.what();
                return;
            } catch (const std::exception& e) {
                std::lock_guard<std::mutex> g(mu);
                driver_code = CPDLP_ERR_ARGUMENT;
                driver_error = e.what();
                return;
            }
            bool idle;
            {
                std::lock_guard<std::mutex> g(mu);
                idle = br.in_batch.empty() && br.waiting.empty();
            }
            if (idle) std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
    }

    void start() {
        std::lock_guard<std::mutex> g(mu);
        if (driving) return;
        driving = true;
        drivers[0] = std::thread([this] { drive(0); });   // also admits the plain lanes
        if (routes[1].b) drivers[1] = std::thread([this] { drive(1); });
    }

    void check_driver() {
        std::lock_guard<std::mutex> g(mu);
        if (driver_code != CPDLP_OK) throw Error(driver_code, "server driver stopped: " + driver_error);
    }

    // cpdlp_server_pump: one round, or with the driver running only the pending count
    int64_t pump_call() {
        check_driver();
        {
            std::lock_guard<std::mutex> g(mu);
            if (driving) return pending_locked();
        }
        return pump();
    }

    int64_t poll(int64_t* ids, int64_t cap) {
        check_driver();
        std::lock_guard<std::mutex> g(mu);
        int64_t k = 0;
        while (!done.empty() && k < cap) {
            if (ids) ids[k] = done.front();
            done.pop_front();
            ++k;
        }
        return k;
    }

    // hands a finished request's result over and frees it; returns the request's code
    int get(int64_t id, double* x, double* y, cpdlp_stats* st, double* seconds, int32_t* route_out) {
        std::unique_ptr<Request> r;
        {
            std::lock_guard<std::mutex> g(mu);
            auto it = reqs.find(id);
            if (it == reqs.end()) throw Error(CPDLP_ERR_ARGUMENT, "unknown request id");
            if (it->second->t_done == 0.0) throw Error(CPDLP_ERR_STATE, "request still running");
            r = std::move(it->second);
            reqs.erase(it);
        }
        if (seconds) *seconds = r->t_done - r->t_submit;
        if (route_out) *route_out = r->lane;
        if (r->code != CPDLP_OK) throw Error(r->code, r->message);
        if (x && r->n) std::memcpy(x, r->x.data(), (size_t)r->n * 8);
        if (y && r->m) std::memcpy(y, r->y.data(), (size_t)r->m * 8);
        if (st) *st = r->stats;
        return CPDLP_OK;
    }
};

}  // namespace cpdlp

struct cpdlp_server {
    std::unique_ptr<cpdlp::Server> s;
};

using cpdlp::Error;
using cpdlp::guarded;

#define SERVER_OR_THROW(s) \
    if (!(s) || !(s)->s) throw Error(CPDLP_ERR_ARGUMENT, "NULL server")

extern "C" {

int cpdlp_server_default_config(cpdlp_server_config* c) {
    return guarded([&] {
        if (!c) throw Error(CPDLP_ERR_ARGUMENT, "NULL config");
        std::memset(c, 0, sizeof(*c));
        c->batch_bytes = (int64_t)16 << 30;
        c->plain_bytes = (int64_t)64 << 30;
        c->batch_limit = 8192;
        c->checks_per_round = 20;
        c->n_lanes = 2;
        c->lane_max_nnz[0] = 1e6;
        c->lane_workers[0] = 3;   // 2 / 3 / 4 measured at 300 QPS: 3 served the most small and medium
        c->lane_max_nnz[1] = INFINITY;
        c->lane_workers[1] = 1;
        c->batch2_bytes = (int64_t)2 << 30;
        c->batch2_limit = 0;   // off: measured, it did not pay on the 70/25/5 mix
        c->batch2_checks = 0;
    });
}

int cpdlp_server_create(const cpdlp_params* params, const cpdlp_server_config* cfg, cpdlp_server** out) {
    return guarded([&] {
        if (!out) throw Error(CPDLP_ERR_ARGUMENT, "NULL argument");
        cpdlp_params p;
        if (params) p = *params;
        else cpdlp_default_params(&p, CPDLP_ALGO_PDLP);
        cpdlp_server_config c;
        if (cfg) c = *cfg;
        else cpdlp_server_default_config(&c);
        auto* s = new cpdlp_server;
        try {
            s->s = std::make_unique<cpdlp::Server>(p, c);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int cpdlp_server_submit(cpdlp_server* s, const cpdlp_lp* lp, const double* x0, const double* y0,
                        int64_t* id) {
    return guarded([&] {
        SERVER_OR_THROW(s);
        if (!lp || !id) throw Error(CPDLP_ERR_ARGUMENT, "NULL argument");
        *id = s->s->submit(*lp, x0, y0, false);
    });
}

int cpdlp_server_submit_borrowed(cpdlp_server* s, const cpdlp_lp* lp, const double* x0,
                                 const double* y0, int64_t* id) {
    return guarded([&] {
        SERVER_OR_THROW(s);
        if (!lp || !id) throw Error(CPDLP_ERR_ARGUMENT, "NULL argument");
        *id = s->s->submit(*lp, x0, y0, true);
    });
}

int cpdlp_server_pump(cpdlp_server* s, int64_t* pending) {
    return guarded([&] {
        SERVER_OR_THROW(s);
        const int64_t k = s->s->pump_call();
        if (pending) *pending = k;
    });
}

int cpdlp_server_poll(cpdlp_server* s, int64_t* ids, int64_t cap, int64_t* count) {
    return guarded([&] {
        SERVER_OR_THROW(s);
        const int64_t k = s->s->poll(ids, cap);
        if (count) *count = k;
    });
}

int cpdlp_server_get(cpdlp_server* s, int64_t id, double* x, double* y, cpdlp_stats* st,
                     double* seconds, int32_t* route) {
    return guarded([&] {
        SERVER_OR_THROW(s);
        s->s->get(id, x, y, st, seconds, route);
    });
}

int cpdlp_server_start(cpdlp_server* s) {
    return guarded([&] {
        SERVER_OR_THROW(s);
        s->s->start();
    });
}

void cpdlp_server_destroy(cpdlp_server* s) { delete s; }

}  // extern "C"
