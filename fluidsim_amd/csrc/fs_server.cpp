// fs_server.cpp -- the serving loop: many resident simulations stepped concurrently on one GPU.
//
// Built only on the grid ABI (include/fluidsim_hip.h).  `lanes` worker threads each own a single-stream
// fs_context (one HIP stream -- one hardware queue -- and pinned read-back words); a lane steps its resident
// instances in rotation, one step each per round, so instances on different lanes overlap on the device
// while each instance's kernels stay ordered on its lane's stream.  With a batch size set, a round steps
// the lane's small instances together (one fs_batch_step) and each larger one with its own fs_step.  Admission is first come,
// first served against an HBM budget counted with fs_grid_estimate_bytes; an instance's grid is created
// and uploaded by its lane when it is first stepped, and destroyed (its device memory freed) when it
// finishes, after its final state has been copied to host memory for fs_server_result.
//
// Every instance owns all of its device state and every kernel of a step reads only that state, so a
// request's results are the same whether it runs alone or among others.
#include "fs_internal.hpp"
#include "../../include/fluidsim_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// Nearest-rank percentile of a copy of `v` (q in [0, 1]); 0 for an empty list.
double percentile(std::vector<float> v, double q)
{
    if (v.empty())
        return 0.0;
    const size_t k = std::min(v.size() - 1, (size_t)std::ceil(q * (double)v.size()) - (q > 0.0 ? 1 : 0));
    std::nth_element(v.begin(), v.begin() + (long)k, v.end());
    return v[k];
}

struct ServerError : std::runtime_error {
    fs_status code;
    ServerError(fs_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void need(bool cond, fs_status code, const std::string& message)
{
    if (!cond)
        throw ServerError(code, message);
}

template <class F> fs_status guard(F&& f)
{
    try {
        f();
        fsim::set_last_error("");
        return FS_OK;
    } catch (const ServerError& e) {
        fsim::set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        fsim::set_last_error("out of host memory");
        return FS_E_RUNTIME;
    } catch (const std::exception& e) {
        fsim::set_last_error(e.what());
        return FS_E_RUNTIME;
    }
}

struct Request {
    long long id = 0;
    fs_request p{};   // scalar parameters; the pointers below replace the caller's
    std::vector<double> rho_l, mu_l, rho_g, mu_g, fraction, u, v, lmass, gmass;
    std::vector<unsigned char> solid;
    size_t estimate = 0;
    int state = FS_REQUEST_QUEUED;
    int lane = -1;
    fs_grid* grid = nullptr;
    int steps_done = 0, pressure_iterations = 0, viscosity_iterations = 0, physical_steps = 0;
    double max_divergence = 0.0;
    Clock::time_point submitted, admitted, finished;
    std::vector<float> step_ms;
    size_t device_bytes = 0;
    double setup_ms = 0.0;      // materialise: create, configure, upload
    double teardown_ms = 0.0;   // retire: download the final state, free the instance
    std::string error;
    // eviction: marked by admission, carried out by the lane at the next step boundary
    bool evict = false;
    bool live = false;          // its grid exists (set and cleared under the server's lock)
    int steps_since_admit = 0;
    int evictions = 0;
    std::string checkpoint_metadata;            // empty unless the instance was evicted
    std::vector<unsigned char> checkpoint_payload;
    // final state, kept on the host once the instance has left the device
    std::vector<double> out_fraction, out_u, out_v, out_p, out_lmass, out_gmass;
};

struct Lane {
    int index = 0;
    fs_context* ctx = nullptr;
    std::vector<Request*> active;
    std::condition_variable wake;
    std::thread thread;
};

}  // namespace

struct fs_server {
    int device = 0;
    size_t budget = 0;
    long long batch_cells = 0;   // instances of at most this many cells step batched
    int evict_after = 0;         // resident steps before an instance may be evicted (0: never)
    size_t evicting_bytes = 0;   // estimates of the instances marked for eviction
    long long evictions = 0;
    std::mutex m;
    std::condition_variable finished_cv;
    std::deque<Request*> queue;
    std::map<long long, std::unique_ptr<Request>> requests;
    std::vector<std::unique_ptr<Lane>> lanes;
    size_t resident_bytes = 0, peak_bytes = 0;
    long long next_id = 1, submitted = 0, admitted = 0, completed = 0, failed = 0;
    bool stopping = false;
    // measurement window
    Clock::time_point window_start = Clock::now();
    long long window_steps = 0;
    std::vector<float> window_step_ms;
    double window_setup_ms = 0.0, window_teardown_ms = 0.0;

    // Admits waiting requests in arrival order while they fit; when the head does not, may mark resident
    // instances for eviction.  Caller holds m.
    void admit_locked()
    {
        while (!queue.empty() && !stopping) {
            Request* r = queue.front();
            if (resident_bytes + r->estimate > budget) {
                mark_evictions_locked(r->estimate);
                break;
            }
            queue.pop_front();
            // small (batched) and larger instances are balanced across the lanes separately
            const bool small = batched(r);
            auto load = [&](const Lane& l) {
                size_t n = 0;
                for (const Request* a : l.active)
                    n += batched(a) == small;
                return n;
            };
            Lane* best = lanes.front().get();
            for (auto& l : lanes)
                if (load(*l) < load(*best))
                    best = l.get();
            resident_bytes += r->estimate;
            peak_bytes = std::max(peak_bytes, resident_bytes);
            r->state = FS_REQUEST_RUNNING;
            r->lane = best->index;
            r->steps_since_admit = 0;
            if (r->evictions == 0) {   // a readmitted instance keeps its first admission time
                r->admitted = Clock::now();
                ++admitted;
            }
            best->active.push_back(r);
            best->wake.notify_one();
        }
    }

    // Marks the longest-resident instances that have run evict_after steps since their admission, enough
    // of them to make room for `need` bytes -- or none, if all of them together would not.  Caller holds m.
    void mark_evictions_locked(size_t need)
    {
        if (evict_after <= 0)
            return;
        size_t room = budget - resident_bytes + evicting_bytes;   // free once the marked ones have left
        if (room >= need)
            return;
        std::vector<Request*> candidates;
        size_t total = 0;
        for (auto& l : lanes)
            for (Request* a : l->active)
                if (!a->evict && a->live && a->steps_since_admit >= evict_after && a->steps_done < a->p.steps) {
                    candidates.push_back(a);
                    total += a->estimate;
                }
        if (room + total < need)
            return;
        std::sort(candidates.begin(), candidates.end(),
            [](const Request* a, const Request* b) { return a->admitted < b->admitted; });
        for (Request* a : candidates) {
            if (room >= need)
                break;
            a->evict = true;
            evicting_bytes += a->estimate;
            room += a->estimate;
        }
    }

    // Checkpoints a marked instance to host memory, frees its grid and queues it again.  On a failure to
    // save, the request ends as failed.
    void evict(Lane& lane, Request* r)
    {
        const auto t0 = Clock::now();
        size_t meta_bytes = 0, payload_bytes = 0;
        std::string meta;
        std::vector<unsigned char> payload;
        fs_status st = fs_checkpoint_size(r->grid, &meta_bytes, &payload_bytes);
        if (st == FS_OK) {
            meta.assign(meta_bytes, '\0');
            payload.resize(payload_bytes);
            st = fs_checkpoint_save(r->grid, &meta[0], meta_bytes, payload.data(), payload_bytes);
        }
        if (st != FS_OK) {
            retire(lane, r, false, std::string("eviction failed: ") + fs_last_error());
            return;
        }
        meta.resize(std::strlen(meta.c_str()));
        fs_grid_destroy(r->grid);
        const double ms = ms_between(t0, Clock::now());
        std::lock_guard<std::mutex> lock(m);
        r->grid = nullptr;
        r->live = false;
        r->teardown_ms += ms;
        window_teardown_ms += ms;
        r->checkpoint_metadata = std::move(meta);
        r->checkpoint_payload = std::move(payload);
        r->evict = false;
        evicting_bytes -= r->estimate;
        r->state = FS_REQUEST_QUEUED;
        ++r->evictions;
        ++evictions;
        lane.active.erase(std::find(lane.active.begin(), lane.active.end(), r));
        resident_bytes -= r->estimate;
        queue.push_back(r);
        admit_locked();
    }

    bool batched(const Request* r) const
    {
        return batch_cells > 0 && (long long)r->p.width * r->p.height <= batch_cells;
    }

    // Ends a request on its lane: keeps the final state (if done), frees the grid, admits more.
    void retire(Lane& lane, Request* r, bool ok, std::string error)
    {
        size_t bytes = 0;
        const auto t0 = Clock::now();
        if (r->grid) {
            (void)fs_grid_device_bytes(r->grid, &bytes);
            if (ok) {
                const size_t N = (size_t)r->p.width * r->p.height;
                r->out_fraction.resize(N);
                r->out_u.resize((size_t)(r->p.width + 1) * r->p.height);
                r->out_v.resize((size_t)r->p.width * (r->p.height + 1));
                r->out_p.resize(N);
                r->out_lmass.resize(r->rho_l.size() * N);
                r->out_gmass.resize(r->rho_g.size() * N);
                if (fs_download(r->grid, r->out_fraction.data(), r->out_u.data(), r->out_v.data(), r->out_p.data(),
                        r->out_lmass.empty() ? nullptr : r->out_lmass.data(),
                        r->out_gmass.empty() ? nullptr : r->out_gmass.data()) != FS_OK) {
                    ok = false;
                }
            }
            if (!ok && error.empty())
                error = fs_last_error();
            fs_grid_destroy(r->grid);
        }
        const double teardown = ms_between(t0, Clock::now());
        std::lock_guard<std::mutex> lock(m);
        r->grid = nullptr;
        r->live = false;
        if (r->evict) {   // marked, but it ended first
            r->evict = false;
            evicting_bytes -= r->estimate;
        }
        r->checkpoint_metadata.clear();
        r->checkpoint_payload.clear();
        r->checkpoint_payload.shrink_to_fit();
        r->device_bytes = bytes;
        r->teardown_ms += teardown;
        window_teardown_ms += teardown;
        r->finished = Clock::now();
        r->state = ok ? FS_REQUEST_DONE : FS_REQUEST_FAILED;
        r->error = ok ? "" : error;
        // the inputs are no longer needed
        r->fraction.clear();
        r->u.clear();
        r->v.clear();
        r->lmass.clear();
        r->gmass.clear();
        if (ok)
            ++completed;
        else
            ++failed;
        lane.active.erase(std::find(lane.active.begin(), lane.active.end(), r));
        resident_bytes -= r->estimate;
        admit_locked();
        finished_cv.notify_all();
    }

    // Creates, configures and uploads an instance on its lane's context.  Returns false on failure.
    bool materialise(Lane& lane, Request* r, std::string& error)
    {
        auto ok = [&](fs_status s) {
            if (s != FS_OK)
                error = fs_last_error();
            return s == FS_OK;
        };
        const fs_request& p = r->p;
        fs_grid* g = nullptr;
        if (!ok(fs_grid_create(lane.ctx, p.width, p.height, p.cell_size, &g)))
            return false;
        r->grid = g;
        auto ptr = [](const std::vector<double>& v) { return v.empty() ? nullptr : v.data(); };
        return ok(fs_configure_materials(g, (int)r->rho_l.size(), ptr(r->rho_l), ptr(r->mu_l), (int)r->rho_g.size(),
                   ptr(r->rho_g), ptr(r->mu_g))) &&
               ok(fs_set_gravity(g, p.gravity_y)) && ok(fs_set_surface_tension(g, p.surface_tension)) &&
               ok(fs_set_transport(g, (fs_transport)p.transport)) &&
               ok(fs_set_preconditioner(g, (fs_preconditioner)p.preconditioner)) &&
               (r->checkpoint_metadata.empty()
                       ? ok(fs_upload(g, ptr(r->fraction), ptr(r->u), ptr(r->v), nullptr, ptr(r->lmass), ptr(r->gmass)))
                       : ok(fs_checkpoint_load(g, r->checkpoint_metadata.c_str(), r->checkpoint_metadata.size(),
                             r->checkpoint_payload.data(), r->checkpoint_payload.size()))) &&
               (r->solid.empty() || ok(fs_set_solid_mask(g, r->solid.data()))) &&
               ok(fs_set_multigrid_precision(g, (fs_mg_precision)p.multigrid_precision)) &&
               ok(fs_set_relative_tolerance(g, p.relative_tolerance)) &&
               (p.smoother_sweeps == 0 ||
                   ok(fs_set_pressure_smoother(g, (fs_pressure_smoother)p.pressure_smoother, p.smoother_sweeps,
                       p.coarse_sweeps ? p.coarse_sweeps : 4))) &&
               ok(fs_set_viscosity_smoother(g, p.viscosity_sweeps));
    }

    // Records one finished step of `r` (ms of lane time); ends the request when it failed, was not
    // accepted, or has run all its steps.
    void finish_step(Lane& lane, Request* r, fs_status st, const fs_projection_result& res, float ms)
    {
        if (st != FS_OK) {
            retire(lane, r, false, fs_last_error());
            return;
        }
        {
            // counters are read by reporting threads under the same lock
            std::lock_guard<std::mutex> lock(m);
            r->step_ms.push_back(ms);
            ++r->steps_done;
            ++r->steps_since_admit;
            r->pressure_iterations += res.iterations;
            r->viscosity_iterations += res.viscosity_iterations;
            r->physical_steps += res.physical_steps;
            r->max_divergence = res.max_divergence;
            ++window_steps;
            window_step_ms.push_back(ms);
            // an instance that has now run its quantum may make room for the waiting head
            if (evict_after > 0 && r->steps_since_admit >= evict_after && !queue.empty())
                admit_locked();
        }
        if (!(res.converged || res.relaxed || res.precision_limited)) {
            retire(lane, r, false,
                "projection not accepted at step " + std::to_string(r->steps_done) + " (max divergence " +
                    std::to_string(res.max_divergence) + ")");
            return;
        }
        if (r->steps_done >= r->p.steps)
            retire(lane, r, true, "");
    }

    void run_lane(Lane& lane)
    {
        (void)hipSetDevice(device);
        std::vector<Request*> round, group, solo;
        for (;;) {
            {
                std::unique_lock<std::mutex> lock(m);
                lane.wake.wait(lock, [&] { return stopping || !lane.active.empty(); });
                if (stopping)
                    break;
                round = lane.active;   // instances admitted meanwhile join the next round
            }
            group.clear();
            solo.clear();
            for (Request* r : round) {
                std::string error;
                if (!r->grid) {
                    const auto t0 = Clock::now();
                    const bool made = materialise(lane, r, error);
                    const double setup = ms_between(t0, Clock::now());
                    {
                        std::lock_guard<std::mutex> lock(m);
                        r->setup_ms += setup;
                        window_setup_ms += setup;
                    }
                    if (!made) {
                        retire(lane, r, false, error);
                        continue;
                    }
                    std::lock_guard<std::mutex> lock(m);
                    r->live = true;
                    r->checkpoint_metadata.clear();
                    r->checkpoint_payload.clear();
                    r->checkpoint_payload.shrink_to_fit();
                }
                if (r->steps_done >= r->p.steps) {
                    retire(lane, r, true, "");
                    continue;
                }
                bool leave;
                {
                    std::lock_guard<std::mutex> lock(m);
                    leave = r->evict;
                }
                if (leave) {
                    evict(lane, r);
                    continue;
                }
                (batched(r) ? group : solo).push_back(r);
            }
            if (!group.empty()) {
                const size_t n = group.size();
                std::vector<fs_grid*> grids(n);
                std::vector<double> dts(n), tolerances(n);
                std::vector<int> budgets(n);
                std::vector<fs_projection_result> results(n);
                std::vector<fs_status> statuses(n, FS_OK);
                for (size_t k = 0; k < n; ++k) {
                    grids[k] = group[k]->grid;
                    dts[k] = group[k]->p.dt;
                    budgets[k] = group[k]->p.max_iterations;
                    tolerances[k] = group[k]->p.divergence_tolerance;
                }
                const auto t0 = Clock::now();
                const fs_status st = fs_batch_step((int)n, grids.data(), dts.data(), budgets.data(), tolerances.data(),
                    results.data(), statuses.data());
                const float ms = (float)ms_between(t0, Clock::now());
                const std::string message = fs_last_error();   // the lowest failing instance's message
                if (st != FS_OK && st != FS_E_RUNTIME) {
                    // the batch itself was refused: every instance of it fails with that message
                    for (Request* r : group)
                        retire(lane, r, false, message);
                } else {
                    bool first = true;
                    for (size_t k = 0; k < n; ++k) {
                        if (statuses[k] != FS_OK) {
                            retire(lane, group[k], false, first ? message : "step failed inside a batched step");
                            first = false;
                            continue;
                        }
                        finish_step(lane, group[k], FS_OK, results[k], ms);
                    }
                }
            }
            for (Request* r : solo) {
                fs_projection_result res{};
                const auto t0 = Clock::now();
                const fs_status st = fs_step(r->grid, r->p.dt, r->p.max_iterations, r->p.divergence_tolerance,
                    nullptr, &res);
                finish_step(lane, r, st, res, (float)ms_between(t0, Clock::now()));
            }
        }
        // shutting down: unfinished instances end as failed
        std::vector<Request*> left;
        {
            std::lock_guard<std::mutex> lock(m);
            left = lane.active;
        }
        for (Request* r : left)
            retire(lane, r, false, "server shut down");
    }

    Request* find_locked(long long id)
    {
        auto it = requests.find(id);
        need(it != requests.end(), FS_E_INVALID_ARGUMENT, "unknown request id");
        return it->second.get();
    }

    void report_locked(const Request* r, fs_request_report* out) const
    {
        fs_request_report rep{};
        rep.state = r->state;
        rep.lane = r->lane;
        rep.steps_done = r->steps_done;
        rep.pressure_iterations = r->pressure_iterations;
        rep.viscosity_iterations = r->viscosity_iterations;
        rep.physical_steps = r->physical_steps;
        rep.max_divergence = r->max_divergence;
        const auto now = Clock::now();
        const bool admitted_yet = r->state != FS_REQUEST_QUEUED || r->evictions > 0;
        const bool ended = r->state == FS_REQUEST_DONE || r->state == FS_REQUEST_FAILED;
        rep.queue_ms = ms_between(r->submitted, admitted_yet ? r->admitted : now);
        rep.run_ms = admitted_yet ? ms_between(r->admitted, ended ? r->finished : now) : 0.0;
        rep.step_p50_ms = percentile(r->step_ms, 0.5);
        rep.step_p99_ms = percentile(r->step_ms, 0.99);
        rep.step_max_ms = r->step_ms.empty() ? 0.0 : *std::max_element(r->step_ms.begin(), r->step_ms.end());
        rep.estimated_bytes = r->estimate;
        rep.device_bytes = r->device_bytes;
        rep.setup_ms = r->setup_ms;
        rep.teardown_ms = r->teardown_ms;
        rep.evictions = r->evictions;
        std::snprintf(rep.error, sizeof rep.error, "%s", r->error.c_str());
        *out = rep;
    }
};

extern "C" {

fs_status fs_server_create(const fs_server_config* config, fs_server** out)
{
    return guard([&] {
        need(config != nullptr && out != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        need(config->lanes >= 1 && config->lanes <= 16, FS_E_INVALID_ARGUMENT, "lanes must be in 1..16");
        auto s = std::make_unique<fs_server>();
        need(config->batch_cells >= 0, FS_E_INVALID_ARGUMENT, "batch_cells must be non-negative");
        need(config->evict_after_steps >= 0, FS_E_INVALID_ARGUMENT, "evict_after_steps must be non-negative");
        s->device = config->device;
        s->budget = config->hbm_budget_bytes;
        s->batch_cells = config->batch_cells;
        s->evict_after = config->evict_after_steps;
        if (s->budget == 0) {
            need(hipSetDevice(config->device) == hipSuccess, FS_E_DEVICE, "no such HIP device");
            size_t free_bytes = 0, total = 0;
            need(hipMemGetInfo(&free_bytes, &total) == hipSuccess, FS_E_DEVICE, "hipMemGetInfo failed");
            s->budget = free_bytes / 5 * 4;
        }
        for (int i = 0; i < config->lanes; ++i) {
            auto lane = std::make_unique<Lane>();
            lane->index = i;
            if (fs_context_create_ex(config->device, FS_CONTEXT_SINGLE_STREAM | FS_CONTEXT_PRIVATE_QUEUE, &lane->ctx) !=
                FS_OK) {
                const std::string message = fs_last_error();
                for (auto& l : s->lanes)
                    fs_context_destroy(l->ctx);
                throw ServerError(FS_E_DEVICE, message);
            }
            s->lanes.push_back(std::move(lane));
        }
        fs_server* server = s.get();
        for (auto& l : s->lanes) {
            Lane* lane = l.get();
            lane->thread = std::thread([server, lane] { server->run_lane(*lane); });
        }
        *out = s.release();
    });
}

fs_status fs_server_destroy(fs_server* server)
{
    return guard([&] {
        if (!server)
            return;
        {
            std::lock_guard<std::mutex> lock(server->m);
            server->stopping = true;
            for (Request* r : server->queue) {
                r->state = FS_REQUEST_FAILED;
                r->error = "server shut down";
                r->finished = Clock::now();
            }
            server->queue.clear();
            for (auto& l : server->lanes)
                l->wake.notify_all();
        }
        for (auto& l : server->lanes)
            if (l->thread.joinable())
                l->thread.join();
        for (auto& l : server->lanes)
            fs_context_destroy(l->ctx);
        server->finished_cv.notify_all();
        delete server;
    });
}

fs_status fs_server_submit(fs_server* server, const fs_request* request, long long* id)
{
    return guard([&] {
        need(server != nullptr && request != nullptr && id != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        const fs_request& q = *request;
        need(q.width > 0 && q.height > 0 && (long long)q.width * q.height <= (1LL << 28), FS_E_INVALID_ARGUMENT,
            "grid dimensions must be positive (at most 2^28 cells)");
        need(std::isfinite(q.cell_size) && q.cell_size > 0.0, FS_E_INVALID_ARGUMENT, "cell size must be positive");
        need(q.steps >= 0, FS_E_INVALID_ARGUMENT, "steps must be non-negative");
        need(std::isfinite(q.dt) && q.dt > 0.0, FS_E_INVALID_ARGUMENT, "dt must be finite and positive");
        need(q.n_liquid >= 0 && q.n_gas >= 0 && q.n_liquid <= 64 && q.n_gas <= 64, FS_E_INVALID_ARGUMENT,
            "at most 64 components per phase");
        need((q.n_liquid == 0 || q.liquid_density) && (q.n_gas == 0 || q.gas_density), FS_E_INVALID_ARGUMENT,
            "component densities are required");
        need(q.transport == FS_TRANSPORT_REFERENCE || q.transport == FS_TRANSPORT_NONE ||
                 q.transport == FS_TRANSPORT_SEMI_LAGRANGIAN,
            FS_E_INVALID_ARGUMENT, "unknown transport mode");
        need(q.preconditioner == FS_PRECOND_DIAGONAL || q.preconditioner == FS_PRECOND_MULTIGRID,
            FS_E_INVALID_ARGUMENT, "unknown preconditioner");
        need(std::isfinite(q.surface_tension) && q.surface_tension >= 0.0, FS_E_INVALID_ARGUMENT,
            "Surface tension must be finite and nonnegative");
        need(q.multigrid_precision >= FS_MG_FP32 && q.multigrid_precision <= FS_MG_BF16_VECTORS,
            FS_E_INVALID_ARGUMENT, "unknown multigrid precision");
        if (q.smoother_sweeps != 0) {
            const bool cheb = q.pressure_smoother == FS_SMOOTHER_CHEBYSHEV;
            need(q.pressure_smoother == FS_SMOOTHER_RED_BLACK || cheb, FS_E_INVALID_ARGUMENT, "unknown pressure smoother");
            need(q.smoother_sweeps >= 1 && q.smoother_sweeps <= (cheb ? 8 : 4) && q.coarse_sweeps >= 0 &&
                     q.coarse_sweeps <= 4,
                FS_E_INVALID_ARGUMENT, "invalid smoother sweep counts");
            need(!cheb || q.multigrid_precision == FS_MG_FP32, FS_E_INVALID_ARGUMENT,
                "the Chebyshev smoother runs FP32 multigrid levels");
        }
        need(q.viscosity_sweeps == 0 || q.viscosity_sweeps == 2 || q.viscosity_sweeps == 4 || q.viscosity_sweeps == 6,
            FS_E_INVALID_ARGUMENT, "viscosity smoother sweeps must be 0, 2, 4 or 6");
        if (q.solid) {
            const std::string problem = fsim::solid_mask_problem(q.solid, q.width, q.height);
            need(problem.empty(), FS_E_INVALID_ARGUMENT, problem);
        }
        auto r = std::make_unique<Request>();
        r->p = q;
        const size_t N = (size_t)q.width * q.height;
        auto copy = [](std::vector<double>& dst, const double* src, size_t n) {
            if (src)
                dst.assign(src, src + n);
        };
        copy(r->rho_l, q.liquid_density, (size_t)q.n_liquid);
        copy(r->rho_g, q.gas_density, (size_t)q.n_gas);
        copy(r->mu_l, q.liquid_viscosity, (size_t)q.n_liquid);
        copy(r->mu_g, q.gas_viscosity, (size_t)q.n_gas);
        copy(r->fraction, q.liquid_fraction, N);
        copy(r->u, q.u, (size_t)(q.width + 1) * q.height);
        copy(r->v, q.v, (size_t)q.width * (q.height + 1));
        copy(r->lmass, q.liquid_mass, (size_t)q.n_liquid * N);
        copy(r->gmass, q.gas_mass, (size_t)q.n_gas * N);
        r->p.liquid_density = r->p.liquid_viscosity = r->p.gas_density = r->p.gas_viscosity = nullptr;
        r->p.liquid_fraction = r->p.u = r->p.v = r->p.liquid_mass = r->p.gas_mass = nullptr;
        if (q.solid)
            r->solid.assign(q.solid, q.solid + N);
        r->p.solid = nullptr;
        bool viscous = false;
        for (double mu : r->mu_l)
            viscous |= mu > 0.0;
        for (double mu : r->mu_g)
            viscous |= mu > 0.0;
        const int features = (q.transport == FS_TRANSPORT_REFERENCE ? FS_FEATURE_TRANSPORT : 0) |
                             (viscous ? FS_FEATURE_VISCOSITY : 0) |
                             (q.surface_tension > 0.0 ? FS_FEATURE_SURFACE_TENSION : 0) |
                             (q.solid ? FS_FEATURE_SOLID_MASK : 0) |
                             (q.multigrid_precision != FS_MG_FP32 ? FS_FEATURE_MG16 : 0);
        need(fs_grid_estimate_bytes(q.width, q.height, q.n_liquid, q.n_gas, features, &r->estimate) == FS_OK,
            FS_E_INVALID_ARGUMENT, "cannot size the request");
        need(r->estimate <= server->budget, FS_E_INVALID_ARGUMENT, "request is larger than the HBM budget");
        std::lock_guard<std::mutex> lock(server->m);
        need(!server->stopping, FS_E_LOGIC, "server is shutting down");
        r->id = server->next_id++;
        r->submitted = Clock::now();
        Request* raw = r.get();
        server->requests[raw->id] = std::move(r);
        server->queue.push_back(raw);
        ++server->submitted;
        server->admit_locked();
        *id = raw->id;
    });
}

fs_status fs_server_wait(fs_server* server, long long id, double timeout_s, fs_request_report* report)
{
    return guard([&] {
        need(server != nullptr, FS_E_INVALID_ARGUMENT, "null server");
        std::unique_lock<std::mutex> lock(server->m);
        Request* r = server->find_locked(id);
        auto ended = [&] { return r->state == FS_REQUEST_DONE || r->state == FS_REQUEST_FAILED; };
        if (timeout_s < 0.0)
            server->finished_cv.wait(lock, ended);
        else
            server->finished_cv.wait_for(lock, std::chrono::duration<double>(timeout_s), ended);
        if (report)
            server->report_locked(r, report);
    });
}

fs_status fs_server_query(fs_server* server, long long id, fs_request_report* report)
{
    return fs_server_wait(server, id, 0.0, report);
}

fs_status fs_server_result(fs_server* server, long long id, double* liquid_fraction, double* u, double* v,
    double* pressure, double* liquid_mass, double* gas_mass)
{
    return guard([&] {
        need(server != nullptr, FS_E_INVALID_ARGUMENT, "null server");
        std::lock_guard<std::mutex> lock(server->m);
        const Request* r = server->find_locked(id);
        need(r->state == FS_REQUEST_DONE, FS_E_LOGIC, "the request has not finished successfully");
        auto put = [](double* dst, const std::vector<double>& src) {
            if (dst && !src.empty())
                std::memcpy(dst, src.data(), src.size() * sizeof(double));
        };
        put(liquid_fraction, r->out_fraction);
        put(u, r->out_u);
        put(v, r->out_v);
        put(pressure, r->out_p);
        put(liquid_mass, r->out_lmass);
        put(gas_mass, r->out_gmass);
    });
}

fs_status fs_server_release(fs_server* server, long long id)
{
    return guard([&] {
        need(server != nullptr, FS_E_INVALID_ARGUMENT, "null server");
        std::lock_guard<std::mutex> lock(server->m);
        const Request* r = server->find_locked(id);
        need(r->state == FS_REQUEST_DONE || r->state == FS_REQUEST_FAILED, FS_E_LOGIC,
            "the request is still queued or running");
        server->requests.erase(id);
    });
}

fs_status fs_server_drain(fs_server* server)
{
    return guard([&] {
        need(server != nullptr, FS_E_INVALID_ARGUMENT, "null server");
        std::unique_lock<std::mutex> lock(server->m);
        server->finished_cv.wait(lock, [&] {
            if (!server->queue.empty())
                return false;
            for (auto& l : server->lanes)
                if (!l->active.empty())
                    return false;
            return true;
        });
    });
}

fs_status fs_server_get_stats(fs_server* server, fs_server_stats* stats)
{
    return guard([&] {
        need(server != nullptr && stats != nullptr, FS_E_INVALID_ARGUMENT, "null pointer");
        std::lock_guard<std::mutex> lock(server->m);
        fs_server_stats s{};
        s.submitted = server->submitted;
        s.admitted = server->admitted;
        s.completed = server->completed;
        s.failed = server->failed;
        s.steps = server->window_steps;
        s.queued = (int)server->queue.size();
        for (auto& l : server->lanes)
            s.resident += (int)l->active.size();
        s.budget_bytes = server->budget;
        s.resident_bytes = server->resident_bytes;
        s.peak_resident_bytes = server->peak_bytes;
        s.elapsed_s = std::chrono::duration<double>(Clock::now() - server->window_start).count();
        s.steps_per_s = s.elapsed_s > 0.0 ? (double)server->window_steps / s.elapsed_s : 0.0;
        s.step_p50_ms = percentile(server->window_step_ms, 0.5);
        s.step_p99_ms = percentile(server->window_step_ms, 0.99);
        s.setup_ms_total = server->window_setup_ms;
        s.teardown_ms_total = server->window_teardown_ms;
        s.evictions = server->evictions;
        *stats = s;
    });
}

fs_status fs_server_reset_stats(fs_server* server)
{
    return guard([&] {
        need(server != nullptr, FS_E_INVALID_ARGUMENT, "null server");
        std::lock_guard<std::mutex> lock(server->m);
        server->window_start = Clock::now();
        server->window_steps = 0;
        server->window_step_ms.clear();
        server->window_setup_ms = server->window_teardown_ms = 0.0;
        server->peak_bytes = server->resident_bytes;
    });
}

}  // extern "C"
