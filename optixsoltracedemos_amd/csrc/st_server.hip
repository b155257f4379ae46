// st_server.hip - continuous-batching trace service (st_server_*).
//
// No counterpart in the reference, which runs one blocking launch per scene
// (lib/soltrace_system.cpp:145-170). Jobs reference device-resident scenes;
// a worker thread owns one stream and runs rounds:
//   1. admit queued jobs (FIFO) while a slot is free and the summed HBM
//      footprint of resident jobs stays within the budget; a job larger than
//      the whole budget is rejected;
//   2. set up every newly admitted job: copy the scene's elements, re-aim its
//      mirrors for the job's sun (FP64 bisector at the scene's last receiver),
//      pack frames, refit the BVH, build the sun plane and its candidate grid
//      and, for jobs of kServerListRays rays or more, the first-reflection
//      candidate lists (each mirror's re-aim candidates, kept per scene, tested
//      against the job's cones) - all on the device, in buffers pooled per scene;
//   3. launch ONE trace kernel over a slice of the active jobs, shortest
//      remaining job first, at most chunk_rays per job and round_rays in all
//      (jobs are chunk lists of a single launch, so work of different jobs
//      overlaps and a short job does not wait behind long ones);
//   4. retire jobs whose rays are all traced (statistics copied to the host,
//      per-job scene copy back to the pool, outputs kept until released).
// Rounds are pipelined: round k+1 is planned and queued on the stream while
// round k runs, and round k is retired (its event waited for) only after that,
// so the device does not idle between rounds. Set-up is queued the same way: a
// batch of admitted jobs has its set-up work and the read-back of its grid
// totals queued, and joins the rounds once that read-back has arrived. Each round uses one of two
// pinned staging slots (job records, chunk prefix, finished jobs' statistics
// and sun frames); a slot is reused two rounds later, after its round retired.
// Ray draws are keyed by (seed, job_id, ray index) and the candidate lists
// give the BVH's hits, so a job's result does not depend on the slices, on the
// lists or on which other jobs shared its launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "st_api_internal.h"

using namespace st;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// device copy of a scene for one job (re-aimed), its sun plane, candidate
// grid and reflection lists; pooled per scene
struct WorkBufs {
    DevBuf elems, aim_in, ref_aabb, bounds, prims, nodes, sun, frames, grid_start, grid_items, list_range, list_items;
    DevBuf flux_copies;   // copies of a small flux map, gathered into the job's flux when it completes
    uint32_t copies = 1;
};
// a job's outputs, kept until st_server_release; pooled
struct OutBufs {
    DevBuf flux, hits, stats;
};

constexpr uint64_t kServerListRays = 1ull << 21;  // jobs this long get first-reflection lists

struct ServerJob {
    int64_t handle = 0;
    st_job_desc desc{};
    st_scene scene = nullptr;
    int32_t status = ST_OK;
    bool done = false, cancel = false, admitted = false, ready = false;
    double submit_s = 0, start_s = 0, finish_s = 0, power_per_ray = 0;
    uint64_t dispatched = 0;
    uint64_t stats[ST_STAT_COUNT] = {0};
    double aim_host[6] = {0, 0, 0, 0, 0, 0};  // sun unit vector + aim target (source of an async upload)
    size_t footprint = 0, scene_bytes = 0;
    bool lists = false;
    std::unique_ptr<WorkBufs> work;
    std::unique_ptr<OutBufs> out;
    JobDev dev{};
};

}  // namespace

struct st_server_s {
    st::ContextRef ctx_ref;
    st_context ctx = nullptr;
    st_server_config cfg{};
    std::vector<st_scene> scenes;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<ServerJob*> pending;
    std::vector<ServerJob*> active;
    std::unordered_map<int64_t, std::unique_ptr<ServerJob>> jobs;
    int64_t next_handle = 1;
    bool stop = false;
    std::thread worker;
    hipStream_t stream = nullptr, fetch_stream = nullptr;
    DevBuf d_jobs, d_prefix, d_counter;
    DevBuf d_round_stats[2], d_round_frames[2];  // per round slot: results of the jobs it finishes
    // pinned staging of the two round slots
    JobDev* h_jobs[2] = {nullptr, nullptr};
    uint64_t* h_prefix[2] = {nullptr, nullptr};
    uint64_t* h_stats[2] = {nullptr, nullptr};   // ST_STAT_COUNT per job the round finishes
    SunFrame* h_frames[2] = {nullptr, nullptr};  // one per job the round finishes
    hipEvent_t round_done[2] = {nullptr, nullptr};
    // grid set-up scratch shared by the jobs set up in a round (worker stream)
    DevBuf d_grid_counts, d_grid_cursor, d_grid_stats;
    uint32_t* h_grid_stats = nullptr;  // pinned, 2 per active slot
    // buffer pools: scene copies per scene slot (worker thread only), outputs (under mu)
    std::unordered_map<int32_t, std::vector<std::unique_ptr<WorkBufs>>> work_pool;
    // per scene slot: first-reflection candidates over the re-aims of its jobs
    // (every job re-aims the mirrors at the scene's receiver; worker thread only)
    std::unordered_map<int32_t, std::unique_ptr<st::ReaimCache>> reaim;
    std::vector<std::unique_ptr<OutBufs>> out_pool;
    st_server_stats stats{};
    std::string worker_error;
};

namespace {

// device bytes of a job: the re-aimed scene copy with its sun plane, grid
// (at most kGridMaxAverage entries per cell), copies of a small flux map and,
// for long jobs, frames and reflection lists; plus the outputs
size_t job_scene_bytes(st_scene sc, const st_job_desc& d) {
    const size_t n = (size_t)sc->n, nn = (size_t)sc->n_nodes;
    size_t b = n * (sizeof(st_element_f64) + 6 * sizeof(double) + 6 * sizeof(float) + 4 * sizeof(float4)) +
               nn * 4 * sizeof(float4) + sizeof(SunFrame) + 6 * sizeof(double) +
               (kGridCells + 1 + (size_t)kGridMaxAverage * kGridCells) * sizeof(uint32_t);
    if (d.num_rays >= kServerListRays)
        b += n * (9 * sizeof(double) + sizeof(uint2) + kMaxBounceList * sizeof(int32_t));
    const size_t nbins = (size_t)std::max(d.flux_nx, 0) * (size_t)std::max(d.flux_ny, 0);
    const uint32_t copies = flux_copies(nbins, d.num_rays);
    if (copies > 1) b += (size_t)copies * nbins * sizeof(int64_t);
    return b;
}

size_t job_footprint(st_scene sc, const st_job_desc& d) {
    const size_t n = (size_t)sc->n;
    const size_t outputs = (size_t)std::max(d.flux_nx, 0) * (size_t)std::max(d.flux_ny, 0) * sizeof(int64_t) +
                           n * sizeof(uint64_t) + ST_STAT_COUNT * sizeof(uint64_t);
    return job_scene_bytes(sc, d) + outputs;
}

// First half of a job's device set-up on the server stream: scene copy, aim,
// frames, refit, sun plane, zeroed outputs, and the count pass of its
// candidate grid (totals to d_grid_stats[2 * slot]).
int setup_job(st_server srv, ServerJob* j, int slot) {
    st_scene sc = j->scene;
    const size_t n = (size_t)sc->n, nn = (size_t)sc->n_nodes;
    hipStream_t st = srv->stream;
    // the job's copy of the scene follows the scene's own queued set-up work
    if (int rc0 = order_before_trace(sc->ctx, st); rc0 != ST_OK) return rc0;
    const st_job_desc& d = j->desc;
    auto& pool = srv->work_pool[d.scene_slot];
    if (!pool.empty()) {
        j->work = std::move(pool.back());
        pool.pop_back();
    } else {
        j->work = std::make_unique<WorkBufs>();
    }
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (!srv->out_pool.empty()) {
            j->out = std::move(srv->out_pool.back());
            srv->out_pool.pop_back();
        }
    }
    if (!j->out) j->out = std::make_unique<OutBufs>();
    WorkBufs& w = *j->work;
    OutBufs& o = *j->out;
    j->lists = d.num_rays >= kServerListRays;
    int rc;
    if ((rc = w.elems.ensure(sizeof(st_element_f64) * n, "job elements")) != ST_OK) return rc;
    if ((rc = w.aim_in.ensure(sizeof(double) * 6, "job sun")) != ST_OK) return rc;
    if ((rc = w.ref_aabb.ensure(sizeof(double) * 6 * n, "job AABBs")) != ST_OK) return rc;
    if ((rc = w.bounds.ensure(sizeof(float) * 6 * n, "job bounds")) != ST_OK) return rc;
    if ((rc = w.prims.ensure(sizeof(float4) * 4 * n, "job primitives")) != ST_OK) return rc;
    if ((rc = w.nodes.ensure(sizeof(float4) * 4 * nn, "job nodes")) != ST_OK) return rc;
    if ((rc = w.sun.ensure(sizeof(SunFrame), "job sun frame")) != ST_OK) return rc;
    if ((rc = w.grid_start.ensure(sizeof(uint32_t) * (kGridCells + 1), "job grid")) != ST_OK) return rc;
    if (j->lists) {
        if ((rc = w.frames.ensure(sizeof(double) * 9 * n, "job frames")) != ST_OK) return rc;
        if ((rc = w.list_range.ensure(sizeof(uint2) * n, "job reflection ranges")) != ST_OK) return rc;
        if ((rc = w.list_items.ensure(sizeof(int32_t) * kMaxBounceList * n, "job reflection lists")) != ST_OK)
            return rc;
    }
    const size_t nbins = (size_t)std::max(d.flux_nx, 0) * (size_t)std::max(d.flux_ny, 0);
    if (nbins && (rc = o.flux.ensure(sizeof(int64_t) * nbins, "job flux")) != ST_OK) return rc;
    if ((rc = o.hits.ensure(sizeof(uint64_t) * n, "job hits")) != ST_OK) return rc;
    if ((rc = o.stats.ensure(sizeof(uint64_t) * ST_STAT_COUNT, "job stats")) != ST_OK) return rc;
    const double l = std::sqrt(d.sun.vector[0] * d.sun.vector[0] + d.sun.vector[1] * d.sun.vector[1] +
                               d.sun.vector[2] * d.sun.vector[2]);
    double* in = j->aim_host;
    for (int k = 0; k < 3; ++k) in[k] = d.sun.vector[k] / l;
    for (int32_t i = sc->n - 1; i >= 0; --i)
        if (sc->host_elems[i].role == ST_ROLE_RECEIVER) {
            for (int k = 0; k < 3; ++k) in[3 + k] = sc->host_elems[i].origin[k];
            break;
        }
    hipError_t e;
    if ((e = hipMemcpyAsync(w.aim_in.ptr, in, sizeof(j->aim_host), hipMemcpyHostToDevice, st)) != hipSuccess)
        return hip_fail(e, "job sun upload");
    // the job's copy of the elements, every mirror re-aimed at the job's sun
    if (d.flags & ST_JOB_KEEP_AIM)
        e = hipMemcpyAsync(w.elems.ptr, sc->d_elems.ptr, sizeof(st_element_f64) * n, hipMemcpyDeviceToDevice, st);
    else
        e = launch_aim_copy(st, (const st_element_f64*)sc->d_elems.ptr, (st_element_f64*)w.elems.ptr, sc->n,
                            (const double*)w.aim_in.ptr, (const double*)w.aim_in.ptr + 3);
    if (e != hipSuccess) return hip_fail(e, "job elements");
    if ((e = launch_frames(st, (const st_element_f64*)w.elems.ptr, sc->n, (const int32_t*)sc->d_elem_slot.ptr, nullptr,
                           nullptr, j->lists ? (double*)w.frames.ptr : nullptr, (double*)w.ref_aabb.ptr,
                           (float*)w.bounds.ptr, (float4*)w.prims.ptr, nullptr, nullptr)) != hipSuccess)
        return hip_fail(e, "job frames");
    if ((e = launch_refit(st, (const int32_t*)sc->d_child.ptr, (const int32_t*)sc->d_level_nodes.ptr,
                          sc->level_offsets.data(), (int32_t)sc->level_offsets.size() - 1, sc->n_nodes, sc->n, 1,
                          (const float*)w.bounds.ptr, (float4*)w.nodes.ptr)) != hipSuccess)
        return hip_fail(e, "job refit");
    st_sun_desc sun = d.sun;
    if (!(sun.dni > 0)) sun.dni = 1000.0;
    if ((e = launch_sun_plane(st, (const double*)w.ref_aabb.ptr, sc->n, 1, (const double*)w.aim_in.ptr, &sun,
                              (SunFrame*)w.sun.ptr)) != hipSuccess)
        return hip_fail(e, "job sun plane");
    if (nbins && (e = launch_fill_u64(st, o.flux.ptr, nbins, 0)) != hipSuccess) return hip_fail(e, "job zero");
    if ((e = launch_fill_u64(st, o.hits.ptr, n, 0)) != hipSuccess) return hip_fail(e, "job zero");
    if ((e = launch_fill_u64(st, o.stats.ptr, ST_STAT_COUNT, 0)) != hipSuccess) return hip_fail(e, "job zero");
    if (sc->n < kGridMinElements) return ST_OK;  // no grid (finish_setup ignores the slot's totals)
    if ((e = launch_fill_u64(st, srv->d_grid_counts.ptr, kGridCells / 2, 0)) != hipSuccess)
        return hip_fail(e, "job grid clear");
    if ((e = launch_grid_count(st, (const float*)w.bounds.ptr, (const float4*)w.prims.ptr,
                               (const uint32_t*)sc->d_meta.ptr, sc->n, 1, (const SunFrame*)w.sun.ptr,
                               std::tan(sun_cone_angle(sun)), (uint32_t*)srv->d_grid_counts.ptr,
                               (uint32_t*)w.grid_start.ptr, (uint32_t*)srv->d_grid_stats.ptr + 2 * slot, nullptr,
                               nullptr, 0)) != hipSuccess)
        return hip_fail(e, "job grid count");
    return ST_OK;
}

// Second half, after the grid totals are on the host: the grid's fill pass
// (when its lists pay), the reflection lists of long jobs, the job record.
int finish_setup(st_server srv, ServerJob* j, uint32_t grid_total, uint32_t grid_longest) {
    st_scene sc = j->scene;
    hipStream_t st = srv->stream;
    const st_job_desc& d = j->desc;
    WorkBufs& w = *j->work;
    OutBufs& o = *j->out;
    st_sun_desc sun = d.sun;
    if (!(sun.dni > 0)) sun.dni = 1000.0;
    hipError_t e;
    int rc;
    const bool grid = sc->n >= kGridMinElements && grid_lists_pay(grid_total, grid_longest);
    if (grid) {
        if ((rc = w.grid_items.ensure(sizeof(int32_t) * grid_total, "job grid items")) != ST_OK) return rc;
        if ((e = launch_fill_u64(st, srv->d_grid_cursor.ptr, kGridCells / 2, 0)) != hipSuccess)
            return hip_fail(e, "job grid cursor clear");
        if ((e = launch_grid_fill(st, (const float*)w.bounds.ptr, (const float4*)w.prims.ptr,
                                  (const uint32_t*)sc->d_meta.ptr, sc->n, 1, (const SunFrame*)w.sun.ptr,
                                  std::tan(sun_cone_angle(sun)), (const uint32_t*)w.grid_start.ptr,
                                  (uint32_t*)srv->d_grid_cursor.ptr, (int32_t*)w.grid_items.ptr, grid_total)) !=
            hipSuccess)
            return hip_fail(e, "job grid fill");
    }
    // a job that re-aims the mirrors at the receiver: its lists test the
    // scene's re-aim candidates instead of walking the BVH
    st::ReaimCache* ra = nullptr;
    if (j->lists && !(d.flags & ST_JOB_KEEP_AIM)) {
        auto& slot_cache = srv->reaim[d.scene_slot];
        if (!slot_cache) slot_cache = std::make_unique<st::ReaimCache>();
        ra = slot_cache.get();
        if ((rc = ensure_reaim(*ra, sc, j->aim_host + 3, sun_cone_angle(sun), st)) != ST_OK) return rc;
    }
    if (j->lists &&
        (e = launch_cone_lists(st, (const st_element_f64*)w.elems.ptr, sc->n, sc->n, sc->n, sc->n_nodes,
                               (const int32_t*)sc->d_elem_slot.ptr, (const uint32_t*)sc->d_meta.ptr,
                               (const double*)w.frames.ptr, (const float*)w.bounds.ptr, (const float4*)w.nodes.ptr,
                               (const SunFrame*)w.sun.ptr, sun_cone_angle(sun), nullptr, nullptr, nullptr,
                               (uint2*)w.list_range.ptr, (int32_t*)w.list_items.ptr,
                               ra ? (const uint2*)ra->range.ptr : nullptr,
                               ra ? (const int32_t*)ra->items.ptr : nullptr)) != hipSuccess)
        return hip_fail(e, "job reflection lists");
    st_trace_desc td;
    memset(&td, 0, sizeof(td));
    td.num_rays = d.num_rays;
    td.seed = d.seed;
    td.job_id = d.job_id;
    td.max_depth = d.max_depth;
    td.flags = d.flags & ST_TRACE_FLUX_DIRECT_SUN;
    td.flux_element = -1;
    td.flux_nx = d.flux_nx;
    td.flux_ny = d.flux_ny;
    st_outputs out;
    memset(&out, 0, sizeof(out));
    const size_t nbins = (size_t)std::max(d.flux_nx, 0) * (size_t)std::max(d.flux_ny, 0);
    out.flux = nbins ? (int64_t*)o.flux.ptr : nullptr;
    out.element_hits = (uint64_t*)o.hits.ptr;
    out.stats = (uint64_t*)o.stats.ptr;
    if ((rc = fill_job(sc, &td, &out, w.prims.ptr, w.nodes.ptr, w.sun.ptr, &j->dev)) != ST_OK) return rc;
    w.copies = 1;
    if (j->dev.flux) {
        if ((rc = route_flux_copies(w.flux_copies, &j->dev, 1, nbins, d.num_rays, st, &w.copies)) != ST_OK)
            return rc;
        // a pooled buffer may hold a cancelled job's partial sums
        if (w.copies > 1 && (e = launch_fill_u64(st, w.flux_copies.ptr, (uint64_t)w.copies * nbins, 0)) != hipSuccess)
            return hip_fail(e, "job flux copies clear");
    }
    if (grid) {
        j->dev.grid_start = (const uint32_t*)w.grid_start.ptr;
        j->dev.grid_items = (const int32_t*)w.grid_items.ptr;
    }
    if (j->lists) {
        j->dev.bounce_range = (const uint2*)w.list_range.ptr;
        j->dev.bounce_items = (const int32_t*)w.list_items.ptr;
    }
    return ST_OK;
}

// scene copy back to its scene's pool (worker thread)
void release_scene_copy(st_server srv, ServerJob* j) {
    if (j->work) srv->work_pool[j->desc.scene_slot].push_back(std::move(j->work));
}

// outputs back to the pool (caller holds srv->mu)
void release_outputs(st_server srv, ServerJob* j) {
    if (j->out) srv->out_pool.push_back(std::move(j->out));
}

void finish_job(st_server srv, ServerJob* j, int32_t status) {
    // caller holds srv->mu
    j->status = status;
    j->done = true;
    j->finish_s = now_s();
    srv->stats.resident_bytes -= j->scene_bytes;
    if (status != ST_OK) srv->stats.resident_bytes -= (j->footprint - j->scene_bytes);
    if (status == ST_OK) srv->stats.completed++;
    else if (status == ST_ERR_CANCELLED) srv->stats.cancelled++;
}

void worker_main(st_server srv) {
    (void)hipSetDevice(srv->ctx->device);
    hipStream_t st = srv->stream;
    // the round in flight: queued on the stream, not yet retired
    struct Inflight {
        bool valid = false;
        int slot = 0;
        std::vector<ServerJob*> jobs, finished;
    } fl;
    int next_slot = 0;
    // admitted jobs waiting for set-up, and the batch whose grid totals are on
    // their way back to h_grid_stats (one batch at a time)
    std::vector<ServerJob*> unset, setting;
    hipEvent_t setup_ev = nullptr;
    bool setup_pending = false;
    if (hipEventCreateWithFlags(&setup_ev, hipEventDisableTiming) != hipSuccess) setup_ev = nullptr;
    // caller holds srv->mu: every active job fails (the stream has faulted)
    auto fail_all = [&](hipError_t e) {
        srv->worker_error = hipGetErrorString(e);
        for (ServerJob* j : srv->active) {
            release_scene_copy(srv, j);
            release_outputs(srv, j);
            finish_job(srv, j, ST_ERR_DEVICE);
        }
        srv->active.clear();
        fl.valid = false;
        unset.clear();
        setting.clear();
        setup_pending = false;
        srv->cv_done.notify_all();
    };
    // waits for the round in flight and retires the jobs it finished
    auto retire = [&]() -> bool {
        if (!fl.valid) return true;
        const hipError_t e = hipEventSynchronize(srv->round_done[fl.slot]);
        std::lock_guard<std::mutex> lk(srv->mu);
        fl.valid = false;
        if (e != hipSuccess) {
            fail_all(e);
            return false;
        }
        for (size_t i = 0; i < fl.finished.size(); ++i) {
            ServerJob* j = fl.finished[i];
            memcpy(j->stats, srv->h_stats[fl.slot] + i * ST_STAT_COUNT, sizeof(j->stats));
            const SunFrame& F = srv->h_frames[fl.slot][i];
            j->power_per_ray = F.dni * F.area / (double)j->desc.num_rays;
            release_scene_copy(srv, j);
            finish_job(srv, j, ST_OK);
            srv->active.erase(std::find(srv->active.begin(), srv->active.end(), j));
        }
        if (!fl.finished.empty()) srv->cv_done.notify_all();
        return true;
    };
    auto in_flight = [&](const ServerJob* j) {
        return (fl.valid && std::find(fl.jobs.begin(), fl.jobs.end(), j) != fl.jobs.end()) ||
               std::find(setting.begin(), setting.end(), j) != setting.end();
    };
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            // sleep unless there is active work or the head of the queue can be decided
            // (admitted, rejected or cancelled); a head waiting for HBM wakes on release
            srv->cv_work.wait(lk, [&] {
                if (srv->stop || !srv->active.empty()) return true;
                if (srv->pending.empty()) return false;
                const ServerJob* h = srv->pending.front();
                return h->cancel || h->footprint > srv->cfg.hbm_budget_bytes ||
                       srv->stats.resident_bytes + h->footprint <= srv->cfg.hbm_budget_bytes;
            });
            if (srv->stop) break;
            // drop cancelled active jobs that the round in flight does not use
            // (the others go once it has retired)
            for (size_t i = 0; i < srv->active.size();) {
                ServerJob* j = srv->active[i];
                if (j->cancel && !in_flight(j)) {
                    release_scene_copy(srv, j);
                    release_outputs(srv, j);
                    finish_job(srv, j, ST_ERR_CANCELLED);
                    srv->active.erase(srv->active.begin() + i);
                    unset.erase(std::remove(unset.begin(), unset.end(), j), unset.end());
                    srv->cv_done.notify_all();
                } else {
                    ++i;
                }
            }
            // admission, FIFO
            while (!srv->pending.empty() && (int)srv->active.size() < srv->cfg.max_active_jobs) {
                ServerJob* j = srv->pending.front();
                if (j->cancel) {
                    srv->pending.pop_front();
                    j->status = ST_ERR_CANCELLED;
                    j->done = true;
                    j->finish_s = now_s();
                    srv->stats.cancelled++;
                    srv->cv_done.notify_all();
                    continue;
                }
                if (j->footprint > srv->cfg.hbm_budget_bytes) {
                    srv->pending.pop_front();
                    j->status = ST_ERR_REJECTED;
                    j->done = true;
                    j->finish_s = now_s();
                    srv->stats.rejected++;
                    srv->cv_done.notify_all();
                    continue;
                }
                if (srv->stats.resident_bytes + j->footprint > srv->cfg.hbm_budget_bytes) break;  // wait for memory
                srv->pending.pop_front();
                srv->stats.resident_bytes += j->footprint;
                srv->stats.peak_resident_bytes = std::max(srv->stats.peak_resident_bytes, srv->stats.resident_bytes);
                srv->stats.admitted++;
                j->admitted = true;
                j->start_s = now_s();
                srv->active.push_back(j);
                unset.push_back(j);
            }
        }
        const double t0 = now_s();
        auto drop = [&](ServerJob* j, int rc) {
            std::lock_guard<std::mutex> lk(srv->mu);
            release_scene_copy(srv, j);
            release_outputs(srv, j);
            finish_job(srv, j, rc);
            srv->active.erase(std::find(srv->active.begin(), srv->active.end(), j));
            srv->cv_done.notify_all();
        };
        // second half of a batch's set-up, once its grid totals are on the
        // host: grids, lists and job records (queued), then the jobs join rounds
        auto finish_batch = [&]() {
            const hipError_t e = hipEventSynchronize(setup_ev);
            for (size_t i = 0; i < setting.size(); ++i) {
                const int rc = e == hipSuccess ? finish_setup(srv, setting[i], srv->h_grid_stats[2 * i],
                                                              srv->h_grid_stats[2 * i + 1])
                                               : hip_fail(e, "job set-up");
                if (rc != ST_OK) drop(setting[i], rc);
                else setting[i]->ready = true;
            }
            setting.clear();
            setup_pending = false;
        };
        if (setup_pending && hipEventQuery(setup_ev) != hipErrorNotReady) finish_batch();
        // first half of the next batch: set-up work and the read-back of its
        // grid totals, queued behind the rounds already on the stream
        if (!setup_pending && !unset.empty()) {
            for (ServerJob* j : unset) {
                const int rc = setup_job(srv, j, (int)setting.size());
                if (rc != ST_OK) drop(j, rc);
                else setting.push_back(j);
            }
            unset.clear();
            if (!setting.empty()) {
                hipError_t e = setup_ev ? hipSuccess : hipErrorInvalidResourceHandle;
                if (e == hipSuccess)
                    e = hipMemcpyAsync(srv->h_grid_stats, srv->d_grid_stats.ptr, sizeof(uint32_t) * 2 * setting.size(),
                                       hipMemcpyDeviceToHost, st);
                if (e == hipSuccess) e = hipEventRecord(setup_ev, st);
                if (e == hipSuccess) {
                    setup_pending = true;
                } else {
                    for (ServerJob* j : setting) drop(j, hip_fail(e, "job set-up"));
                    setting.clear();
                }
            }
        }
        // one launch over slices of the active jobs: shortest remaining first,
        // at most chunk_rays per job and round_rays in all
        std::vector<ServerJob*> cand;
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            for (ServerJob* j : srv->active)
                if (j->ready && !j->cancel && j->dispatched < j->desc.num_rays) cand.push_back(j);
        }
        if (cand.empty()) {
            // nothing to queue: wait for a set-up batch, else for the round in flight
            if (setup_pending) finish_batch();
            else retire();
            continue;
        }
        std::stable_sort(cand.begin(), cand.end(), [](const ServerJob* x, const ServerJob* y) {
            return x->desc.num_rays - x->dispatched < y->desc.num_rays - y->dispatched;
        });
        std::vector<ServerJob*> round;
        std::vector<uint64_t> slice;
        uint64_t round_rays = 0;
        const uint64_t budget = srv->cfg.round_rays;
        for (ServerJob* j : cand) {
            if (round_rays >= budget) break;
            const uint64_t sl = std::min<uint64_t>({srv->cfg.chunk_rays, j->desc.num_rays - j->dispatched,
                                                    budget - round_rays});
            round.push_back(j);
            slice.push_back(sl);
            round_rays += sl;
        }
        int32_t depth = 0, min_n = INT32_MAX;
        for (ServerJob* j : round) {
            depth = std::max(depth, j->scene->topo.depth);
            min_n = std::min(min_n, j->scene->n);
        }
        bool lists = false;  // long jobs have first-reflection lists and no pair lists
        for (ServerJob* j : round) lists = lists || j->lists;
        LaunchShape sh = pick_shape(round_rays, srv->cfg.block_size, srv->cfg.grid_blocks, ST_KERNEL_AUTO, depth,
                                    nullptr, min_n, 0);
        use_grouped_bvh(sh, lists);
        // slot `slot` was last used two rounds ago, and that round has retired
        const int slot = next_slot;
        next_slot ^= 1;
        JobDev* hj = srv->h_jobs[slot];
        uint64_t* hp = srv->h_prefix[slot];
        hp[0] = 0;
        for (size_t i = 0; i < round.size(); ++i) {
            JobDev J = round[i]->dev;
            J.ray_begin = round[i]->dispatched;
            J.ray_count = slice[i];
            J.record_base = J.ray_begin;
            set_trace_order(&J, round[i]->desc.sun.position_sampling);
            hj[i] = J;
            hp[i + 1] = hp[i] + (slice[i] + sh.chunk - 1) / sh.chunk;
        }
        // the device copies are ordered after the round in flight on the stream
        hipError_t e = hipMemcpyAsync(srv->d_jobs.ptr, hj, sizeof(JobDev) * round.size(), hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(srv->d_prefix.ptr, hp, sizeof(uint64_t) * (round.size() + 1), hipMemcpyHostToDevice, st);
        JobDev dummy;
        memset(&dummy, 0, sizeof(dummy));
        if (e == hipSuccess)
            e = launch_trace(st, (const JobDev*)srv->d_jobs.ptr, (const uint64_t*)srv->d_prefix.ptr,
                             (int32_t)round.size(), hp[round.size()], (uint32_t*)srv->d_counter.ptr, sh, dummy);
        // statistics and sun frames of the jobs this round completes
        std::vector<ServerJob*> finished;
        for (size_t i = 0; i < round.size(); ++i) {
            round[i]->dispatched += slice[i];
            if (round[i]->dispatched >= round[i]->desc.num_rays) finished.push_back(round[i]);
        }
        std::vector<const uint64_t*> fin_stats;
        std::vector<const SunFrame*> fin_suns;
        for (size_t i = 0; i < finished.size() && e == hipSuccess; ++i) {
            ServerJob* j = finished[i];
            if (j->work->copies > 1)
                e = launch_flux_gather(st, j->dev.flux, j->work->copies, j->dev.flux_bins, 1, (int64_t*)j->out->flux.ptr);
            fin_stats.push_back((const uint64_t*)j->out->stats.ptr);
            fin_suns.push_back((const SunFrame*)j->work->sun.ptr);
        }
        // their statistics and sun frames: gathered on the device, two copies
        if (e == hipSuccess && !finished.empty()) {
            const int32_t nf = (int32_t)finished.size();
            e = launch_round_results(st, fin_stats.data(), fin_suns.data(), nf, (uint64_t*)srv->d_round_stats[slot].ptr,
                                     (SunFrame*)srv->d_round_frames[slot].ptr);
            if (e == hipSuccess)
                e = hipMemcpyAsync(srv->h_stats[slot], srv->d_round_stats[slot].ptr, sizeof(uint64_t) * ST_STAT_COUNT * nf,
                                   hipMemcpyDeviceToHost, st);
            if (e == hipSuccess)
                e = hipMemcpyAsync(srv->h_frames[slot], srv->d_round_frames[slot].ptr, sizeof(SunFrame) * nf,
                                   hipMemcpyDeviceToHost, st);
        }
        if (e == hipSuccess) e = hipEventRecord(srv->round_done[slot], st);
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(st);
            std::lock_guard<std::mutex> lk(srv->mu);
            fail_all(e);
            continue;
        }
        // the new round is queued: retire the previous one while it runs
        const bool ok = retire();
        fl.valid = ok;
        fl.slot = slot;
        fl.jobs = std::move(round);
        fl.finished = std::move(finished);
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stats.launches++;
        srv->stats.rays_traced += round_rays;
        srv->stats.busy_s += now_s() - t0;
    }
    retire();
    if (setup_pending) (void)hipEventSynchronize(setup_ev);
    if (setup_ev) (void)hipEventDestroy(setup_ev);
}

void fill_result(const ServerJob* j, st_job_result* r) {
    memset(r, 0, sizeof(*r));
    r->status = j->status;
    r->done = j->done ? 1 : 0;
    r->submit_s = j->submit_s;
    r->start_s = j->start_s;
    r->finish_s = j->finish_s;
    r->power_per_ray = j->power_per_ray;
    memcpy(r->stats, j->stats, sizeof(r->stats));
}

}  // namespace

#define HIP_TRY(expr, what)                                \
    do {                                                   \
        hipError_t e__ = (expr);                           \
        if (e__ != hipSuccess) return hip_fail(e__, what); \
    } while (0)

extern "C" int st_server_create(st_context ctx, const st_server_config* cfg, st_server* out) {
    if (!ctx || !cfg || !out) return fail(ST_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->max_active_jobs <= 0 || cfg->max_active_jobs > 4096)
        return fail(ST_ERR_INVALID, "max_active_jobs must be in [1, 4096]");
    if (cfg->hbm_budget_bytes == 0) return fail(ST_ERR_INVALID, "zero HBM budget");
    if (cfg->block_size != 0 && cfg->block_size != 64 && cfg->block_size != 128 && cfg->block_size != 256)
        return fail(ST_ERR_INVALID, "block_size must be 0, 64, 128 or 256");
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    auto srv = std::make_unique<st_server_s>();
    srv->ctx = ctx;
    srv->ctx_ref.hold(ctx);
    srv->cfg = *cfg;
    if (srv->cfg.chunk_rays == 0) srv->cfg.chunk_rays = 8u << 20;
    if (srv->cfg.round_rays == 0) srv->cfg.round_rays = 4 * srv->cfg.chunk_rays;
    HIP_TRY(hipStreamCreateWithFlags(&srv->stream, hipStreamNonBlocking), "server stream");
    HIP_TRY(hipStreamCreateWithFlags(&srv->fetch_stream, hipStreamNonBlocking), "server fetch stream");
    const size_t J = (size_t)srv->cfg.max_active_jobs;
    int rc;
    if ((rc = srv->d_jobs.ensure(sizeof(JobDev) * J, "server jobs")) != ST_OK) return rc;
    if ((rc = srv->d_prefix.ensure(sizeof(uint64_t) * (J + 1), "server prefix")) != ST_OK) return rc;
    if ((rc = srv->d_counter.ensure(sizeof(uint32_t) * 2, "server counter")) != ST_OK) return rc;
    HIP_TRY(hipMemset(srv->d_counter.ptr, 0, sizeof(uint32_t) * 2), "counter clear");
    for (int k = 0; k < 2; ++k) {
        HIP_TRY(hipHostMalloc((void**)&srv->h_jobs[k], sizeof(JobDev) * J, hipHostMallocDefault), "pinned jobs");
        HIP_TRY(hipHostMalloc((void**)&srv->h_prefix[k], sizeof(uint64_t) * (J + 1), hipHostMallocDefault),
                "pinned prefix");
        HIP_TRY(hipHostMalloc((void**)&srv->h_stats[k], sizeof(uint64_t) * ST_STAT_COUNT * J, hipHostMallocDefault),
                "pinned job statistics");
        HIP_TRY(hipHostMalloc((void**)&srv->h_frames[k], sizeof(SunFrame) * J, hipHostMallocDefault),
                "pinned sun frames");
        HIP_TRY(hipEventCreateWithFlags(&srv->round_done[k], hipEventDisableTiming), "round event");
        if ((rc = srv->d_round_stats[k].ensure(sizeof(uint64_t) * ST_STAT_COUNT * J, "round statistics")) != ST_OK)
            return rc;
        if ((rc = srv->d_round_frames[k].ensure(sizeof(SunFrame) * J, "round sun frames")) != ST_OK) return rc;
    }
    HIP_TRY(hipHostMalloc((void**)&srv->h_grid_stats, sizeof(uint32_t) * 2 * J, hipHostMallocDefault), "pinned grid stats");
    if ((rc = srv->d_grid_counts.ensure(sizeof(uint32_t) * kGridCells, "server grid counts")) != ST_OK) return rc;
    if ((rc = srv->d_grid_cursor.ensure(sizeof(uint32_t) * kGridCells, "server grid cursor")) != ST_OK) return rc;
    if ((rc = srv->d_grid_stats.ensure(sizeof(uint32_t) * 2 * J, "server grid stats")) != ST_OK) return rc;
    HIP_TRY(hipDeviceSynchronize(), "server init sync");
    st_server s = srv.release();
    s->worker = std::thread(worker_main, s);
    *out = s;
    return ST_OK;
}

extern "C" int st_server_destroy(st_server srv) {
    if (!srv) return fail(ST_ERR_INVALID, "null server");
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    if (srv->worker.joinable()) srv->worker.join();
    (void)hipSetDevice(srv->ctx->device);
    (void)hipStreamSynchronize(srv->stream);
    srv->jobs.clear();
    for (int k = 0; k < 2; ++k) {
        if (srv->h_jobs[k]) (void)hipHostFree(srv->h_jobs[k]);
        if (srv->h_prefix[k]) (void)hipHostFree(srv->h_prefix[k]);
        if (srv->h_stats[k]) (void)hipHostFree(srv->h_stats[k]);
        if (srv->h_frames[k]) (void)hipHostFree(srv->h_frames[k]);
        if (srv->round_done[k]) (void)hipEventDestroy(srv->round_done[k]);
    }
    if (srv->h_grid_stats) (void)hipHostFree(srv->h_grid_stats);
    (void)hipStreamDestroy(srv->stream);
    (void)hipStreamDestroy(srv->fetch_stream);
    delete srv;
    return ST_OK;
}

extern "C" int st_server_add_scene(st_server srv, st_scene scene, int32_t* slot) {
    if (!srv || !scene || !slot) return fail(ST_ERR_INVALID, "null argument");
    if (scene->ctx != srv->ctx) return fail(ST_ERR_INVALID, "scene belongs to another context");
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->scenes.push_back(scene);
    *slot = (int32_t)srv->scenes.size() - 1;
    return ST_OK;
}

extern "C" int st_server_submit(st_server srv, const st_job_desc* job, int64_t* handle) {
    if (!srv || !job || !handle) return fail(ST_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    if (job->scene_slot < 0 || job->scene_slot >= (int32_t)srv->scenes.size())
        return fail(ST_ERR_INVALID, "unknown scene slot %d", job->scene_slot);
    if (job->num_rays == 0) return fail(ST_ERR_INVALID, "num_rays must be positive");
    if (job->max_depth < 1 || job->max_depth > kMaxDepthLimit) return fail(ST_ERR_INVALID, "bad max_depth %d", job->max_depth);
    if (job->flux_nx < 0 || job->flux_ny < 0 || (int64_t)job->flux_nx * job->flux_ny > (1 << 26))
        return fail(ST_ERR_INVALID, "bad flux grid");
    if (job->flags & ST_TRACE_RECORDS) return fail(ST_ERR_INVALID, "records are not served");
    const double* v = job->sun.vector;
    if (!(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] > 0)) return fail(ST_ERR_INVALID, "zero sun vector");
    if (job->sun.shape != ST_SUN_PILLBOX && job->sun.shape != ST_SUN_GAUSSIAN) return fail(ST_ERR_INVALID, "bad sun shape");
    if (!(job->sun.param >= 0) || job->sun.param > 0.2) return fail(ST_ERR_INVALID, "bad sun angle");
    if (job->sun.position_sampling == ST_POS_HALTON && job->num_rays > kHaltonMaxRays)
        return fail(ST_ERR_INVALID, "Halton positions take a 32-bit ray index (num_rays > 2^32)");
    st_scene sc = srv->scenes[job->scene_slot];
    bool has_receiver = false;
    for (const auto& e : sc->host_elems) has_receiver |= e.role == ST_ROLE_RECEIVER;
    if (!has_receiver) return fail(ST_ERR_INVALID, "scene has no receiver to aim at");
    auto j = std::make_unique<ServerJob>();
    j->handle = srv->next_handle++;
    j->desc = *job;
    j->scene = sc;
    j->submit_s = now_s();
    j->footprint = job_footprint(sc, *job);
    j->scene_bytes = job_scene_bytes(sc, *job);
    ServerJob* raw = j.get();
    srv->jobs[j->handle] = std::move(j);
    srv->pending.push_back(raw);
    srv->stats.submitted++;
    *handle = raw->handle;
    srv->cv_work.notify_all();
    return ST_OK;
}

extern "C" int st_server_poll(st_server srv, int64_t h, st_job_result* r) {
    if (!srv || !r) return fail(ST_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->jobs.find(h);
    if (it == srv->jobs.end()) return fail(ST_ERR_INVALID, "unknown job %lld", (long long)h);
    fill_result(it->second.get(), r);
    return ST_OK;
}

extern "C" int st_server_wait(st_server srv, int64_t h, st_job_result* r) {
    if (!srv || !r) return fail(ST_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->jobs.find(h);
    if (it == srv->jobs.end()) return fail(ST_ERR_INVALID, "unknown job %lld", (long long)h);
    ServerJob* j = it->second.get();
    srv->cv_done.wait(lk, [&] { return j->done; });
    fill_result(j, r);
    return ST_OK;
}

extern "C" int st_server_fetch(st_server srv, int64_t h, int64_t* flux_host, uint64_t* hits_host) {
    if (!srv) return fail(ST_ERR_INVALID, "null server");
    ServerJob* j = nullptr;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->jobs.find(h);
        if (it == srv->jobs.end()) return fail(ST_ERR_INVALID, "unknown job %lld", (long long)h);
        j = it->second.get();
        if (!j->done) return fail(ST_ERR_STATE, "job %lld has not finished", (long long)h);
        if (j->status != ST_OK) return fail(j->status, "job %lld did not complete", (long long)h);
    }
    HIP_TRY(hipSetDevice(srv->ctx->device), "hipSetDevice");
    const size_t nbins = (size_t)std::max(j->desc.flux_nx, 0) * (size_t)std::max(j->desc.flux_ny, 0);
    if (flux_host && nbins)
        HIP_TRY(hipMemcpyAsync(flux_host, j->out->flux.ptr, nbins * sizeof(int64_t), hipMemcpyDeviceToHost, srv->fetch_stream),
                "fetch flux");
    if (hits_host)
        HIP_TRY(hipMemcpyAsync(hits_host, j->out->hits.ptr, (size_t)j->scene->n * sizeof(uint64_t), hipMemcpyDeviceToHost,
                               srv->fetch_stream),
                "fetch hits");
    HIP_TRY(hipStreamSynchronize(srv->fetch_stream), "fetch sync");
    return ST_OK;
}

extern "C" int st_server_cancel(st_server srv, int64_t h) {
    if (!srv) return fail(ST_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->jobs.find(h);
    if (it == srv->jobs.end()) return fail(ST_ERR_INVALID, "unknown job %lld", (long long)h);
    if (!it->second->done) it->second->cancel = true;
    srv->cv_work.notify_all();
    return ST_OK;
}

extern "C" int st_server_release(st_server srv, int64_t h) {
    if (!srv) return fail(ST_ERR_INVALID, "null server");
    std::unique_ptr<ServerJob> j;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->jobs.find(h);
        if (it == srv->jobs.end()) return fail(ST_ERR_INVALID, "unknown job %lld", (long long)h);
        if (!it->second->done) return fail(ST_ERR_STATE, "job %lld is still running; cancel it first", (long long)h);
        if (it->second->status == ST_OK) srv->stats.resident_bytes -= (it->second->footprint - it->second->scene_bytes);
        j = std::move(it->second);
        srv->jobs.erase(it);
        release_outputs(srv, j.get());  // the output buffers serve later jobs
        srv->cv_work.notify_all();      // memory freed: admission may proceed
    }
    return ST_OK;
}

extern "C" int st_server_get_stats(st_server srv, st_server_stats* out) {
    if (!srv || !out) return fail(ST_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    *out = srv->stats;
    return ST_OK;
}
