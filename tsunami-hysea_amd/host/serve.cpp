// serve.cpp -- request loop for batched scenarios (include/thysea_serve.h) over the
// C-ABI of libthysea.so.  No counterpart in the reference (one scenario per process).
//
// The main thread drives the batch: it queues one captured step sequence, and while
// the GPU runs it retires finished scenarios and admits arrived ones (a scenario
// admitted during a launch joins the next one).  Worker threads do the work that
// touches one context only: creating and uploading it and preparing its run
// (thy_batch_prepare: metrics, initial dt, faults due at t = 0) ahead of admission, and
// after retirement reading out its max height, arrival times and gauges and
// destroying it.
#include "../../include/thysea_serve.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <numeric>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

class Pool {
  public:
    explicit Pool(int n) {
        for (int k = 0; k < n; k++) threads_.emplace_back([this] { loop(); });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : threads_) t.join();
    }
    template <typename F>
    auto submit(F f) -> std::future<decltype(f())> {
        auto task = std::make_shared<std::packaged_task<decltype(f())()>>(std::move(f));
        auto fut = task->get_future();
        {
            std::lock_guard<std::mutex> g(m_);
            q_.emplace_back([task] { (*task)(); });
        }
        cv_.notify_one();
        return fut;
    }

  private:
    void loop() {
        for (;;) {
            std::function<void()> job;
            {
                std::unique_lock<std::mutex> l(m_);
                cv_.wait(l, [this] { return stop_ || !q_.empty(); });
                if (q_.empty()) return;
                job = std::move(q_.front());
                q_.pop_front();
            }
            job();
        }
    }
    std::vector<std::thread> threads_;
    std::deque<std::function<void()>> q_;
    std::mutex m_;
    std::condition_variable cv_;
    bool stop_ = false;
};

template <typename T>
bool ready(const std::future<T> &f) {
    return f.wait_for(std::chrono::seconds(0)) == std::future_status::ready;
}

struct Opts {
    int capacity, K, workers, burst;
    size_t budget;
};

Opts opts_of(const thy_serve_options *o) {
    Opts r;
    r.capacity = (o && o->capacity > 0) ? o->capacity : 512;
    r.K = (o && o->steps_per_sync > 0) ? o->steps_per_sync : 48;
    r.workers = (o && o->workers > 0) ? o->workers : 8;
    r.burst = (o && o->burst > 0) ? o->burst : 8;
    r.budget = o ? o->hbm_budget : 0;
    return r;
}

// the request's run configuration without callbacks or snapshots
thy_run_config run_of(const thy_request &q, int K) {
    thy_run_config c = q.run;
    c.save_interval = -1.0;
    c.on_snapshot = nullptr;
    c.on_okada = nullptr;
    c.on_step = nullptr;
    c.user = nullptr;
    if (c.steps_per_sync <= 0) c.steps_per_sync = K;
    return c;
}

// context created, uploaded and (batched) prepared; nullptr with the status on failure
struct Prepared {
    thy_ctx *ctx = nullptr;
    int status = THY_OK;
};

Prepared create_upload(const thy_request &q) {
    Prepared p;
    if (!q.state1 && (!q.h || !q.H || !q.qx || !q.qy)) {
        p.status = THY_ERR_INVALID;
        return p;
    }
    p.status = thy_create(&q.grid, &q.params, &p.ctx);
    if (p.status == THY_OK) {
        if (q.state1) {
            p.status = thy_upload_state(p.ctx, q.state1, q.state2);
        } else {
            // pair the four arrays as the reference's double2 state
            const size_t n = (size_t)q.grid.nx * q.grid.ny;
            std::vector<double> s1, s2;
            try {
                s1.resize(2 * n);
                s2.resize(2 * n);
            } catch (const std::bad_alloc &) {
                p.status = THY_ERR_HOST_MEMORY;
            }
            if (p.status == THY_OK) {
                for (size_t c = 0; c < n; c++) {
                    s1[2 * c] = q.h[c];
                    s1[2 * c + 1] = q.H[c];
                    s2[2 * c] = q.qx[c];
                    s2[2 * c + 1] = q.qy[c];
                }
                p.status = thy_upload_state(p.ctx, s1.data(), s2.data());
            }
        }
    }
    if (p.status != THY_OK && p.ctx) {
        thy_destroy(p.ctx);
        p.ctx = nullptr;
    }
    return p;
}

// max height, arrival times and gauges into the caller's buffers
int read_out(thy_ctx *ctx, const thy_request &q, thy_request_result &r, int n_samples) {
    int e = THY_OK;
    if (r.max_height && (e = thy_get_field(ctx, THY_MAX_HEIGHT, r.max_height)) != THY_OK) return e;
    if (r.arrival_times && (e = thy_get_field(ctx, THY_ARRIVAL_TIMES, r.arrival_times)) != THY_OK) return e;
    r.n_samples = n_samples;
    const int ng = q.run.n_gauges;
    if (n_samples > 0 && ng > 0 && (r.g_time || r.g_eta || r.g_u || r.g_v)) {
        // the library copies every recorded sample: stage them, then keep what fits
        std::vector<float> t(n_samples), e3[3];
        for (auto &v : e3) v.resize((size_t)n_samples * ng);
        if ((e = thy_get_gauges(ctx, t.data(), e3[0].data(), e3[1].data(), e3[2].data(), nullptr, nullptr,
                                nullptr)) != THY_OK)
            return e;
        const int keep = std::min(n_samples, std::max(0, r.gauge_capacity));
        if (r.g_time) std::copy(t.begin(), t.begin() + keep, r.g_time);
        float *dst[3] = {r.g_eta, r.g_u, r.g_v};
        for (int k = 0; k < 3; k++)
            if (dst[k]) std::copy(e3[k].begin(), e3[k].begin() + (size_t)keep * ng, dst[k]);
    }
    return THY_OK;
}

}  // namespace

extern "C" {

int thy_serve_batched(const thy_request *req, int n, const thy_serve_options *opt, thy_request_result *out,
                      double *gpu_seconds) {
    if (n < 0 || (n > 0 && (!req || !out))) return THY_ERR_INVALID;
    const Opts o = opts_of(opt);
    thy_batch *b = nullptr;
    int rc = thy_batch_create(o.capacity, o.budget, o.K, &b);
    if (rc != THY_OK) {
        thy_batch_destroy(b);
        return rc;
    }
    const Clock::time_point t0 = Clock::now();
    auto now = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return req[a].arrival_s < req[c].arrival_s; });
    for (int i = 0; i < n; i++) {
        out[i].status = THY_OK;
        out[i].admitted_s = out[i].finished_s = -1.0;
        out[i].steps = 0;
        out[i].n_samples = 0;
    }
    std::vector<int> slot_req(o.capacity, -1);
    std::vector<thy_ctx *> slot_ctx(o.capacity, nullptr);
    std::deque<int> pending;  // finished (or failed) slots not yet retired
    std::map<int, std::future<Prepared>> prepared;
    std::vector<std::pair<int, std::future<double>>> readouts;
    int nxt = 0, nsub = 0, done = 0, n_live = 0;
    bool in_flight = false;
    int err = THY_OK;
    {
        Pool pool(o.workers);
        auto submit_prepares = [&] {
            // up to two requests per worker in preparation, in arrival order
            while (nsub < n && req[order[nsub]].arrival_s <= now() && nsub - nxt < 2 * o.workers) {
                const int i = order[nsub++];
                const thy_run_config cfg = run_of(req[i], o.K);
                prepared[i] = pool.submit([&req, i, cfg]() {
                    Prepared p = create_upload(req[i]);
                    if (p.ctx && (p.status = thy_batch_prepare(p.ctx, &cfg)) != THY_OK) {
                        thy_destroy(p.ctx);
                        p.ctx = nullptr;
                    }
                    return p;
                });
            }
        };
        auto finish_now = [&](int i, int status) {
            out[i].status = status;
            out[i].finished_s = now();
            done++;
        };
        // 1 admitted (or resolved), 0 slots or budget full, -1 not prepared yet
        auto admit = [&](bool block) -> int {
            const int i = order[nxt];
            auto it = prepared.find(i);
            if (it == prepared.end() || (!block && !ready(it->second))) return -1;
            Prepared p = it->second.get();
            prepared.erase(it);
            if (!p.ctx) {
                nxt++;
                finish_now(i, p.status);
                return 1;
            }
            int slot = -1;
            const int r = thy_batch_add(b, p.ctx, nullptr, &slot);
            if (r == THY_ERR_BUDGET) {
                if (n_live == 0) {  // can never fit
                    thy_destroy(p.ctx);
                    nxt++;
                    finish_now(i, THY_ERR_BUDGET);
                    return 1;
                }
                // keep it prepared for a later attempt
                std::promise<Prepared> pr;
                pr.set_value(p);
                prepared[i] = pr.get_future();
                return 0;
            }
            if (r != THY_OK) {
                thy_destroy(p.ctx);
                nxt++;
                finish_now(i, r);
                return 1;
            }
            slot_req[slot] = i;
            slot_ctx[slot] = p.ctx;
            n_live++;
            out[i].admitted_s = now();
            nxt++;
            return 1;
        };
        auto retire = [&](int slot) {
            const int i = slot_req[slot];
            thy_ctx *ctx = slot_ctx[slot];
            int state = 0;
            thy_run_result res{};
            thy_batch_status(b, slot, &state, &res);
            thy_batch_remove(b, slot);
            slot_req[slot] = -1;
            slot_ctx[slot] = nullptr;
            n_live--;
            out[i].steps = res.steps;
            const bool failed = state == THY_SLOT_FAILED;
            readouts.emplace_back(i, pool.submit([&, i, ctx, failed, res]() {
                const int r = failed ? THY_ERR_NONFINITE : read_out(ctx, req[i], out[i], res.n_samples);
                out[i].status = r;
                thy_destroy(ctx);
                return now();
            }));
        };
        auto collect = [&](bool block) {
            std::vector<std::pair<int, std::future<double>>> rest;
            for (auto &rf : readouts) {
                if (block || ready(rf.second)) {
                    out[rf.first].finished_s = rf.second.get();
                    done++;
                } else {
                    rest.push_back(std::move(rf));
                }
            }
            readouts.swap(rest);
        };
        while (done < n && err == THY_OK) {
            submit_prepares();
            int k = 0;
            for (;;) {
                collect(false);
                if (in_flight) {
                    int d = 0;
                    if ((err = thy_batch_ready(b, &d)) != THY_OK) break;
                    if (d) break;
                } else if (k >= o.burst) {
                    break;
                }
                if (!pending.empty()) {
                    const int s = pending.front();
                    pending.pop_front();
                    retire(s);
                    continue;
                }
                if (nxt < n && req[order[nxt]].arrival_s <= now()) {
                    const int a = admit(!in_flight);
                    if (a == 1) {
                        k++;
                        submit_prepares();
                        continue;
                    }
                    if (a == 0) break;
                    submit_prepares();  // it arrived after the last submission, or is still being prepared
                    std::this_thread::sleep_for(std::chrono::microseconds(200));
                    continue;
                }
                break;
            }
            if (err != THY_OK) break;
            if (in_flight) {
                int running = 0;
                if ((err = thy_batch_wait(b, &running)) != THY_OK) break;
                in_flight = false;
                for (int s = 0; s < o.capacity; s++) {
                    if (slot_req[s] < 0) continue;
                    int st = 0;
                    thy_batch_status(b, s, &st, nullptr);
                    if ((st == THY_SLOT_FINISHED || st == THY_SLOT_FAILED) &&
                        std::find(pending.begin(), pending.end(), s) == pending.end())
                        pending.push_back(s);
                }
            }
            bool any_running = false;
            for (int s = 0; s < o.capacity && !any_running; s++) {
                if (slot_req[s] < 0) continue;
                int st = 0;
                thy_batch_status(b, s, &st, nullptr);
                any_running = st == THY_SLOT_RUNNING;
            }
            if (any_running) {
                if ((err = thy_batch_launch(b)) != THY_OK) break;
                in_flight = true;
            } else if (pending.empty()) {
                if (!readouts.empty()) {
                    collect(true);
                } else if (nxt < n) {
                    const double wait = std::min(0.01, std::max(0.0, req[order[nxt]].arrival_s - now()));
                    std::this_thread::sleep_for(std::chrono::duration<double>(wait));
                }
            }
        }
        // on an error: finish what is in flight and release every context
        if (in_flight) thy_batch_wait(b, nullptr);
        collect(true);
        for (auto &pf : prepared) {
            Prepared p = pf.second.get();
            if (p.ctx) thy_destroy(p.ctx);
        }
        for (int s = 0; s < o.capacity; s++)
            if (slot_ctx[s]) {
                thy_batch_remove(b, s);
                thy_destroy(slot_ctx[s]);
            }
    }
    if (gpu_seconds) thy_batch_info(b, nullptr, nullptr, gpu_seconds);
    thy_batch_destroy(b);
    return err;
}

int thy_serve_sequential(const thy_request *req, int n, const thy_serve_options *opt, thy_request_result *out) {
    if (n < 0 || (n > 0 && (!req || !out))) return THY_ERR_INVALID;
    const Opts o = opts_of(opt);
    const Clock::time_point t0 = Clock::now();
    auto now = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return req[a].arrival_s < req[c].arrival_s; });
    for (int i : order) {
        const double wait = req[i].arrival_s - now();
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
        thy_request_result &r = out[i];
        r.admitted_s = now();
        r.steps = 0;
        r.n_samples = 0;
        Prepared p = create_upload(req[i]);
        r.status = p.status;
        if (p.ctx) {
            const thy_run_config cfg = run_of(req[i], o.K);
            thy_run_result res{};
            r.status = thy_run(p.ctx, &cfg, &res);
            r.steps = res.steps;
            if (r.status == THY_OK) r.status = read_out(p.ctx, req[i], r, res.n_samples);
            thy_destroy(p.ctx);
        }
        r.finished_s = now();
    }
    return THY_OK;
}

}  // extern "C"
