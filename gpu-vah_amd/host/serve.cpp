// serve.cpp -- the serving form's request loop (include/vah_serve.h).
//
// Continuous batching over the batch engine: first-come first-served
// admission against the engine's slot and HBM-budget limits, one batch of
// steps per poll for every resident event, retirement as the engine stops
// each event, latency from arrival to retirement.  The loop talks to the
// engine through vah_serve_backend so that its scheduling can be exercised
// against a stand-in as well.
#include <math.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <numeric>
#include <thread>
#include <vector>

#include "../../include/vah_serve.h"

namespace {

double steady_now() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// numpy's default percentile: linear interpolation between closest ranks
double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q / 100.0 * (double)(v.size() - 1);
    const size_t lo = (size_t)floor(pos), hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (pos - (double)lo) * (v[hi] - v[lo]);
}

// the batch engine as a backend
struct BatchBackend {
    vah_batch *b;
};

int batch_admit(void *u, const vah_request *r, int *slot) {
    vah_batch *b = static_cast<BatchBackend *>(u)->b;
    if (!r->e) return vah_batch_admit_glauber(b, &r->ev, &r->glauber, r->pl_ratio, slot);
    if (r->pl_ratio > 0) {
        const size_t n = (size_t)r->ev.nx * r->ev.ny * r->ev.nz;
        std::vector<double> pl(n);
        for (size_t m = 0; m < n; ++m) pl[m] = r->pl_ratio * r->e[m];
        return vah_batch_admit(b, &r->ev, r->e, pl.data(), nullptr, nullptr, nullptr, nullptr, nullptr, slot);
    }
    return vah_batch_admit(b, &r->ev, r->e, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, slot);
}
int batch_step(void *u, int k) { return vah_batch_step(static_cast<BatchBackend *>(u)->b, k); }
int batch_poll(void *u, vah_batch_status *out, int cap, int *n) {
    return vah_batch_poll(static_cast<BatchBackend *>(u)->b, out, cap, n);
}
int batch_retire(void *u, int slot) { return vah_batch_retire(static_cast<BatchBackend *>(u)->b, slot); }

}  // namespace

extern "C" {

int vah_serve_run(const vah_serve_backend *be, const vah_request *reqs, int n, int steps_per_poll,
                  vah_retire_fn on_retire, void *user, vah_request_result *results, vah_serve_stats *stats) {
    if (!be || !be->admit || !be->step || !be->poll || !be->retire || n < 0 || (n > 0 && !reqs) ||
        steps_per_poll < 1)
        return VAH_ERR_ARG;
    auto now = [be] { return be->now ? be->now(be->user) : steady_now(); };
    auto sleep_for = [be](double s) {
        if (be->sleep)
            be->sleep(be->user, s);
        else if (s > 0)
            std::this_thread::sleep_for(std::chrono::duration<double>(s));
    };
    std::vector<vah_request_result> res(n);
    for (int i = 0; i < n; ++i) {
        res[i] = vah_request_result{reqs[i].ev.tag, -1, -1, 0, 0.0, -1.0, -1.0, -1.0};
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [reqs](int a, int b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
    std::deque<int> pending(order.begin(), order.end()), waiting;
    std::map<int, int> resident;  // slot -> request
    std::vector<vah_batch_status> st(std::max(n, 1));
    int peak = 0, polls = 0;
    const double start = now();
    while (!pending.empty() || !waiting.empty() || !resident.empty()) {
        const double t = now() - start;
        while (!pending.empty() && reqs[pending.front()].arrival_s <= t) {
            waiting.push_back(pending.front());
            pending.pop_front();
        }
        // first come, first served: the head of the queue waits for room
        while (!waiting.empty()) {
            const int i = waiting.front();
            int slot = -1;
            const int rc = be->admit(be->user, &reqs[i], &slot);
            if (rc == VAH_ERR_BUDGET || rc == VAH_ERR_FULL) break;
            if (rc != VAH_OK) return rc;
            waiting.pop_front();
            res[i].admitted_s = now() - start;
            res[i].slot = slot;
            resident[slot] = i;
        }
        peak = std::max(peak, (int)resident.size());
        if (resident.empty()) {
            if (!waiting.empty()) {
                // nothing resident and the head still does not fit: it never will
                const int i = waiting.front();
                waiting.pop_front();
                res[i].finished_s = now() - start;
                res[i].state = VAH_REQUEST_REJECTED;
                continue;
            }
            if (!pending.empty()) sleep_for(std::max(0.0, reqs[pending.front()].arrival_s - (now() - start)));
            continue;
        }
        int rc = be->step(be->user, steps_per_poll);
        if (rc != VAH_OK) return rc;
        int got = 0;
        if ((rc = be->poll(be->user, st.data(), (int)st.size(), &got)) != VAH_OK) return rc;
        ++polls;
        const double tp = now() - start;
        for (int k = 0; k < got; ++k) {
            const vah_batch_status &s = st[k];
            auto it = resident.find(s.slot);
            if (s.state == VAH_EVENT_LIVE || it == resident.end()) continue;
            const int i = it->second;
            resident.erase(it);
            res[i].finished_s = tp;
            res[i].nstep = s.nstep;
            res[i].tau = s.tau;
            res[i].state = s.state;
            res[i].latency_s = tp - reqs[i].arrival_s;
            if (on_retire) on_retire(user, &res[i]);
            if ((rc = be->retire(be->user, s.slot)) != VAH_OK) return rc;
        }
        if (be->on_poll) be->on_poll(be->user, tp, (int)resident.size(), (int)waiting.size());
    }
    const double total = now() - start;
    std::vector<double> lat;
    int rejected = 0;
    for (int i = 0; i < n; ++i) {
        if (res[i].state == VAH_REQUEST_REJECTED)
            ++rejected;
        else
            lat.push_back(res[i].latency_s);
    }
    if (results) std::copy(res.begin(), res.end(), results);
    if (stats) {
        stats->served = n - rejected;
        stats->rejected = rejected;
        stats->peak_resident = peak;
        stats->polls = polls;
        stats->seconds = total;
        stats->events_per_s = total > 0 ? (n - rejected) / total : 0.0;
        stats->p50_latency_s = percentile(lat, 50);
        stats->p99_latency_s = percentile(lat, 99);
    }
    return VAH_OK;
}

int vah_serve(vah_batch *b, const vah_request *reqs, int n, int steps_per_poll, vah_retire_fn on_retire, void *user,
              vah_request_result *results, vah_serve_stats *stats) {
    if (!b) return VAH_ERR_ARG;
    BatchBackend bb{b};
    vah_serve_backend be{&bb, batch_admit, batch_step, batch_poll, batch_retire, nullptr, nullptr, nullptr};
    return vah_serve_run(&be, reqs, n, steps_per_poll, on_retire, user, results, stats);
}

}  // extern "C"
