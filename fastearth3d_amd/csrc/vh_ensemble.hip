// vh_ensemble.hip -- an uneven ensemble of independent models advanced
// together on one GPU (host side only; every member keeps its own stream and
// device state and runs the single-model kernels).
//
// The reference has no in-process batching: an ensemble there is a set of
// separate solid_earth instances (src/vilma_coupling.f90:127-176) stepped one
// after another by the host. Here a pool of host threads, each bound to the
// device the ensemble was created on, takes members from a shared queue and
// steps them on their own streams, so that the kernels of several members
// occupy the GPU at once: a member of lmax 64 fills a few percent of the
// 256 CUs, and the host work of one member (launches, its guard read-back)
// overlaps the device work of the others. Members are taken longest first
// (by the measured duration of their previous work, or by nlm * ne with a
// weight for lateral viscosity before anything was measured).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "vilma_hip.h"

struct vh_ensemble {
    std::vector<vh_model*> m;
    std::vector<size_t> ng;       // grid points of each member
    std::vector<double> cost;     // per-call cost estimate (seconds, or the static weight)
    std::vector<double> meas;     // seconds of each member's task in the current job (-1: not run)
    int dev = 0;
    std::string err;
    int failed = -1;
    // worker pool
    std::vector<std::thread> th;
    std::mutex mu;
    std::condition_variable cv_go, cv_done;
    long gen = 0;                 // job generation
    int running = 0;              // workers still busy with the current job
    bool quit = false;
    std::function<bool(int)> job; // runs member i; false on failure
    std::vector<int> order;
    std::atomic<int> next{0};
    std::atomic<bool> abort{false};
};

namespace {

void worker(vh_ensemble* e) {
    hipSetDevice(e->dev);
    long seen = 0;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(e->mu);
            e->cv_go.wait(lk, [&] { return e->quit || e->gen != seen; });
            if (e->quit) return;
            seen = e->gen;
        }
        for (;;) {
            if (e->abort.load()) break;
            int k = e->next.fetch_add(1);
            if (k >= (int)e->order.size()) break;
            int i = e->order[k];
            auto t0 = std::chrono::steady_clock::now();
            bool ok = e->job(i);
            double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (!ok) {
                std::lock_guard<std::mutex> lk(e->mu);
                if (e->failed < 0) {
                    e->failed = i;
                    const char* msg = vh_last_error(e->m[i]);
                    e->err = "ensemble member " + std::to_string(i) + ": " + (msg ? msg : "failed");
                }
                e->abort.store(true);
                break;
            }
            e->meas[i] = s;
        }
        std::lock_guard<std::mutex> lk(e->mu);
        if (--e->running == 0) e->cv_done.notify_all();
    }
}

// run job(i) for every member on the pool, longest first; per_call divides
// the measured time of a task into the per-call cost kept for ordering
int dispatch(vh_ensemble* e, std::function<bool(int)> job, int per_call) {
    const int n = (int)e->m.size();
    e->order.resize(n);
    std::iota(e->order.begin(), e->order.end(), 0);
    std::stable_sort(e->order.begin(), e->order.end(), [&](int a, int b) { return e->cost[a] > e->cost[b]; });
    e->meas.assign(n, -1.0);
    {
        std::lock_guard<std::mutex> lk(e->mu);
        e->job = std::move(job);
        e->next.store(0);
        e->abort.store(false);
        e->failed = -1;
        e->err.clear();
        e->running = (int)e->th.size();
        e->gen++;
    }
    e->cv_go.notify_all();
    {
        std::unique_lock<std::mutex> lk(e->mu);
        e->cv_done.wait(lk, [&] { return e->running == 0; });
        e->job = nullptr;
    }
    for (int i = 0; i < n; i++)
        if (e->meas[i] >= 0.0) e->cost[i] = e->meas[i] / per_call;
    return e->failed < 0 ? 0 : 1;
}

}  // namespace

extern "C" {

int vh_ensemble_create(vh_model* const* members, int n_members, int n_workers, vh_ensemble** out) {
    if (!out) return 2;
    *out = nullptr;
    if (!members || n_members <= 0) return 2;
    vh_ensemble* e = new vh_ensemble();
    hipGetDevice(&e->dev);
    e->m.assign(members, members + n_members);
    for (int i = 0; i < n_members; i++) {
        int nphi = 0, nlat = 0, nlm = 0, ne = 0, ne3d = 0;
        if (!e->m[i] || vh_get_dims(e->m[i], VH_D_NPHI, &nphi) || vh_get_dims(e->m[i], VH_D_NLAT, &nlat) ||
            vh_get_dims(e->m[i], VH_D_NLM, &nlm) || vh_get_dims(e->m[i], VH_D_NE, &ne) ||
            vh_get_dims(e->m[i], VH_D_NE3D, &ne3d)) {
            delete e;
            return 2;
        }
        e->ng.push_back((size_t)nphi * nlat);
        // static weight until a measurement exists: spectral size x elements,
        // lateral viscosity (tensor transforms over the 3-D elements) x 6
        e->cost.push_back(1.0e3 * (double)nlm * ne * (ne3d > 0 ? 6.0 : 1.0));
    }
    int nw = n_workers > 0 ? n_workers : 16;
    nw = std::max(1, std::min(nw, n_members));
    for (int w = 0; w < nw; w++) e->th.emplace_back(worker, e);
    *out = e;
    return 0;
}

int vh_ensemble_update(vh_ensemble* e, const double* const* h_ice, const double* dt_yr) {
    if (!e || !h_ice || !dt_yr) return 2;
    return dispatch(e, [=](int i) { return vh_update(e->m[i], h_ice[i], dt_yr[i]) == 0; }, 1);
}

int vh_ensemble_update_device(vh_ensemble* e, const double* const* h_ice_device, const double* dt_yr) {
    if (!e || !h_ice_device || !dt_yr) return 2;
    return dispatch(e, [=](int i) { return vh_update_device(e->m[i], h_ice_device[i], dt_yr[i]) == 0; }, 1);
}

int vh_ensemble_run_device(vh_ensemble* e, int ncalls, const double* const* ice_hist_device, const double* dt_yr) {
    if (!e || !ice_hist_device || !dt_yr || ncalls < 0) return 2;
    if (ncalls == 0) return 0;
    return dispatch(e, [=](int i) {
        for (int k = 0; k < ncalls; k++) {
            if (e->abort.load()) return true;   // another member failed: stop early, not an error of this one
            if (vh_update_device(e->m[i], ice_hist_device[i] + (size_t)k * e->ng[i], dt_yr[i])) return false;
        }
        return true;
    }, ncalls);
}

int vh_ensemble_tune(vh_ensemble* e, int ncalls, const double* const* ice_hist_device, const double* dt_yr,
                     int* chosen_sle_blocks) {
    if (!e || !ice_hist_device || !dt_yr || ncalls <= 0) return 2;
    // candidates of the sea-level solve's block count: one block per work
    // item (the single-model optimum), 64 and 32 blocks (room for the other
    // members' kernels beside each solve)
    const int cand[3] = {0, 64, 32};
    const int n = (int)e->m.size();
    double best_t = 0.0;
    int best = 0;
    for (int c = 0; c < 3; c++) {
        for (int i = 0; i < n; i++)
            if (vh_set_tuning(e->m[i], VH_T_SLE_BLOCKS, cand[c])) return 2;
        std::vector<const double*> at(n);
        for (int i = 0; i < n; i++) at[i] = ice_hist_device[i] + (size_t)c * ncalls * e->ng[i];
        auto t0 = std::chrono::steady_clock::now();
        if (vh_ensemble_run_device(e, ncalls, at.data(), dt_yr)) return 1;
        const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (c == 0 || t < best_t) { best_t = t; best = cand[c]; }
    }
    for (int i = 0; i < n; i++) vh_set_tuning(e->m[i], VH_T_SLE_BLOCKS, best);
    if (chosen_sle_blocks) *chosen_sle_blocks = best;
    return 0;
}

int vh_ensemble_failed_member(const vh_ensemble* e) { return e ? e->failed : -1; }

const char* vh_ensemble_last_error(const vh_ensemble* e) { return e ? e->err.c_str() : "null ensemble"; }

int vh_ensemble_destroy(vh_ensemble* e) {
    if (!e) return 0;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        e->quit = true;
    }
    e->cv_go.notify_all();
    for (auto& t : e->th) t.join();
    delete e;
    return 0;
}

}  // extern "C"
