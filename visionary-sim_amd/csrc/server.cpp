// server.cpp -- continuous batching of simulation scenes on one world.
//
// Requests (a scene, a frame count, optional per-frame grab targets and an
// arrival time) wait in arrival order. Every iteration the server admits the
// requests that have arrived while the world's pools -- the HBM budget -- can
// hold them and enqueues ONE batched frame of every admitted scene. Frames
// are pipelined: frame k+1 is enqueued while frame k runs, and only then does
// the host wait, on frame k's event, to book it, so the device always has the
// next frame queued. A scene whose last frame has been enqueued is removed at
// once (its final positions copied out first, stream-ordered, when results
// are kept): the world's removal does not wait for the device, and the next
// admission into its ranges waits on the device for the frames before it.
// Newcomers are admitted on the upload stream while frames run. A scene's
// frame latency runs from the moment the frame became eligible (the scene
// admitted, or its previous frame done) to the completion of the batch that
// ran it. The reference steps one scene per browser tab and has no
// counterpart.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vsim.h"
#include "world_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

// Owning copy of a vs_scene_desc and everything it points at.
struct SceneCopy {
    std::vector<vs_body_desc> bodies;
    std::vector<std::vector<uint32_t>> cells, surface, statics;
    std::vector<std::vector<uint8_t>> occupancy;
    std::vector<std::vector<float>> rest, inv, mean, quat, scale;
    std::vector<uint32_t> driven;
    std::vector<float> offsets;
    vs_scene_desc desc{};
    uint64_t nodes = 0;

    template <typename T>
    static std::vector<T> take(const T *p, size_t n)
    {
        return p && n ? std::vector<T>(p, p + n) : std::vector<T>();
    }

    explicit SceneCopy(const vs_scene_desc &d)
    {
        desc = d;
        bodies.assign(d.bodies, d.bodies + d.n_bodies);
        for (vs_body_desc &b : bodies) {
            const size_t nn = (size_t)(b.cell_dims[0] + 1) * (b.cell_dims[1] + 1) * (b.cell_dims[2] + 1);
            nodes += nn;
            cells.push_back(take(b.cells, b.n_cells));
            surface.push_back(take(b.surface, b.n_surface));
            rest.push_back(take(b.rest, 3 * nn));
            inv.push_back(take(b.inv_mass, nn));
            mean.push_back(take(b.splat_mean, 3 * (size_t)b.n_splats));
            quat.push_back(take(b.splat_quat, 4 * (size_t)b.n_splats));
            scale.push_back(take(b.splat_scale, 3 * (size_t)b.n_splats));
            occupancy.push_back(take(b.occupancy, b.occupancy ? (size_t)b.cell_dims[0] * b.cell_dims[1] * b.cell_dims[2] : 0));
            statics.push_back(take(b.static_nodes, b.n_static));
        }
        for (size_t i = 0; i < bodies.size(); i++) {
            bodies[i].cells = cells[i].data();
            bodies[i].surface = surface[i].data();
            bodies[i].rest = rest[i].data();
            bodies[i].inv_mass = inv[i].data();
            bodies[i].splat_mean = mean[i].empty() ? nullptr : mean[i].data();
            bodies[i].splat_quat = quat[i].empty() ? nullptr : quat[i].data();
            bodies[i].splat_scale = scale[i].empty() ? nullptr : scale[i].data();
            bodies[i].occupancy = occupancy[i].empty() ? nullptr : occupancy[i].data();
            bodies[i].static_nodes = statics[i].empty() ? nullptr : statics[i].data();
        }
        driven = take(d.driven, d.n_driven);
        offsets = take(d.driven_offset, 3 * (size_t)d.n_driven);
        desc.bodies = bodies.data();
        desc.driven = driven.empty() ? nullptr : driven.data();
        desc.driven_offset = offsets.empty() ? nullptr : offsets.data();
    }
};

struct Request {
    std::unique_ptr<SceneCopy> scene;
    double arrival = 0;
    uint32_t frames = 0, done = 0, enqueued = 0;
    std::vector<float> targets;  // (frames + 1) x 3, or empty
    int64_t scene_id = -1;
    double admitted = -1, finished = -1, eligible = 0;
    std::vector<float> result;   // final positions, 3 planes
    float *staging = nullptr;    // pinned: the final positions' copy in flight
};

// A frame in flight: the requests it steps, the event after it.
struct Flight {
    std::vector<size_t> batch;
    uint32_t slot = 0;
};

}  // namespace

struct vs_server {
    vs_world *world = nullptr;
    vs_server_config cfg{};
    std::vector<Request> requests;
    std::vector<float> latencies_ms;
    std::string error;
};

extern "C" {

int vs_server_create(vs_world *world, const vs_server_config *cfg, vs_server **out)
{
    if (!world || !cfg || !out || cfg->substeps == 0) return VS_E_INVALID;
    vs_server *s = new vs_server();
    s->world = world;
    s->cfg = *cfg;
    *out = s;
    return VS_OK;
}

void vs_server_destroy(vs_server *s) { delete s; }

const char *vs_server_error(const vs_server *s) { return s ? s->error.c_str() : "no server"; }

int vs_server_submit(vs_server *s, const vs_scene_desc *desc, double arrival_s, uint32_t frames, const float *targets,
                     uint32_t *request_id)
{
    if (!s || !desc || !frames || !(arrival_s >= 0)) return VS_E_INVALID;
    if (!s->requests.empty() && arrival_s < s->requests.back().arrival) {
        s->error = "requests must be submitted in arrival order";
        return VS_E_INVALID;
    }
    Request r;
    r.scene.reset(new SceneCopy(*desc));
    r.arrival = arrival_s;
    r.frames = frames;
    if (targets) r.targets.assign(targets, targets + 3 * ((size_t)frames + 1));
    s->requests.push_back(std::move(r));
    if (request_id) *request_id = (uint32_t)(s->requests.size() - 1);
    return VS_OK;
}

int vs_server_run(vs_server *s, int32_t realtime, vs_server_report *report)
{
    if (!s || !report) return VS_E_INVALID;
    vs_world *w = s->world;
    const auto t0 = Clock::now();
    auto now = [&]() { return std::chrono::duration<double>(Clock::now() - t0).count(); };
    std::deque<size_t> waiting;
    std::vector<size_t> active;
    size_t next = 0;
    uint64_t node_substeps = 0, frames = 0;
    uint32_t iterations = 0, max_active = 0;
    s->latencies_ms.clear();

    auto admit = [&](double t) -> int {
        while (next < s->requests.size() && (!realtime || s->requests[next].arrival <= t)) waiting.push_back(next++);
        while (!waiting.empty()) {
            Request &r = s->requests[waiting.front()];
            uint32_t id = 0;
            const int rc = vs_world_add_scene(w, &r.scene->desc, &id);
            if (rc == VS_E_NOMEM) break;  // the budget is full: wait for retirements
            if (rc) {
                s->error = std::string("admission failed: ") + vs_world_error(w);
                return rc;
            }
            r.scene_id = id;
            r.admitted = now();
            r.eligible = r.admitted;
            active.push_back(waiting.front());
            waiting.pop_front();
        }
        return VS_OK;
    };

    // Frames complete at the GPU time of their events, measured from an
    // epoch event recorded on the idle stream: the host may book a frame
    // well after it ran (an admission in between), and its latency must not
    // count that.
    if (int rc = vs_world_synchronize(w)) return rc;
    struct Epoch {
        hipEvent_t e = nullptr;
        ~Epoch()
        {
            if (e) (void)hipEventDestroy(e);
        }
    } epoch;
    if (hipEventCreate(&epoch.e) != hipSuccess || hipEventRecord(epoch.e, vsim::world_stream(w)) != hipSuccess) {
        s->error = "server epoch event";
        return VS_E_DEVICE;
    }
    const double epoch_t = now();
    std::deque<Flight> flights;  // oldest first; at most two
    // Waits for the oldest frame in flight and books it.
    auto book = [&]() -> int {
        const Flight f = std::move(flights.front());
        flights.pop_front();
        const hipEvent_t ev = vsim::world_event(w, f.slot);
        float ms = 0.f;
        if (!ev || hipEventSynchronize(ev) != hipSuccess || hipEventElapsedTime(&ms, epoch.e, ev) != hipSuccess) {
            s->error = "waiting for a frame";
            return VS_E_DEVICE;
        }
        const double done_t = epoch_t + 1e-3 * (double)ms;
        iterations++;
        max_active = std::max<uint32_t>(max_active, (uint32_t)f.batch.size());
        for (size_t k : f.batch) {
            Request &r = s->requests[k];
            s->latencies_ms.push_back((float)((done_t - r.eligible) * 1e3));
            r.eligible = done_t;
            r.done++;
            frames++;
            node_substeps += r.scene->nodes * s->cfg.substeps;
            if (r.done < r.frames) continue;
            r.finished = done_t;
            if (r.staging) r.result.assign(r.staging, r.staging + 3 * r.scene->nodes);
        }
        return VS_OK;
    };
    struct Staging {
        std::vector<float *> p;
        ~Staging()
        {
            for (float *q : p) (void)hipHostFree(q);
        }
    } staging;

    if (int rc = admit(now())) return rc;
    while (next < s->requests.size() || !waiting.empty() || !active.empty() || !flights.empty()) {
        if (active.empty() && flights.empty()) {
            if (waiting.empty() && realtime && next < s->requests.size()) {
                const double wait = s->requests[next].arrival - now();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
            }
            if (int rc = admit(now())) return rc;
            if (active.empty()) {
                if (!waiting.empty()) {
                    s->error = "a waiting scene does not fit the empty world";
                    return VS_E_NOMEM;
                }
                continue;
            }
        }
        if (!active.empty()) {
            // Grab targets of the frame about to be enqueued.
            for (size_t k : active) {
                Request &r = s->requests[k];
                if (r.targets.empty()) continue;
                const float *cur = &r.targets[3 * (size_t)(r.enqueued + 1)], *prev = &r.targets[3 * (size_t)r.enqueued];
                const double fdt = (double)s->cfg.substeps * (double)r.scene->desc.dt;
                float vel[3];
                for (int a = 0; a < 3; a++) vel[a] = (float)(((double)cur[a] - (double)prev[a]) / fdt);
                if (int rc = vs_world_set_target(w, (uint32_t)r.scene_id, cur, vel)) return rc;
            }
            if (int rc = vs_world_frame(w, s->cfg.substeps, s->cfg.splats, s->cfg.graph)) {
                s->error = vs_world_error(w);
                return rc;
            }
            // Scenes whose last frame this is leave the world now; their
            // ranges are reused only after this frame (and the copy of their
            // final positions) has run.
            std::vector<size_t> still;
            for (size_t k : active) {
                Request &r = s->requests[k];
                if (++r.enqueued < r.frames) {
                    still.push_back(k);
                    continue;
                }
                if (s->cfg.keep_results) {
                    float *buf = nullptr;
                    if (hipHostMalloc(reinterpret_cast<void **>(&buf), sizeof(float) * 3 * r.scene->nodes) != hipSuccess) {
                        s->error = "pinned staging of a result";
                        return VS_E_NOMEM;
                    }
                    staging.p.push_back(buf);
                    r.staging = buf;
                    if (int rc = vsim::world_read_positions_async(w, (uint32_t)r.scene_id, buf)) return rc;
                }
                if (int rc = vs_world_remove_scene(w, (uint32_t)r.scene_id)) return rc;
                r.scene_id = -1;
            }
            Flight f;
            f.batch = active;
            if (int rc = vs_world_record(w, &f.slot)) return rc;
            flights.push_back(std::move(f));
            active = still;
            // Newcomers are admitted while the device runs; their first frame
            // is the next batch.
            if (int rc = admit(now())) return rc;
        }
        // Keep one frame queued behind the one running; with nothing to
        // enqueue, drain.
        while (flights.size() > 1 || (!flights.empty() && active.empty())) {
            if (int rc = book()) return rc;
            if (int rc = admit(now())) return rc;
            if (!active.empty()) break;
        }
    }
    const double wall = now();
    std::vector<float> sorted = s->latencies_ms;
    std::sort(sorted.begin(), sorted.end());
    auto pct = [&](double q) -> double {
        if (sorted.empty()) return 0;
        const size_t i = std::min(sorted.size() - 1, (size_t)std::ceil(q * (double)sorted.size()) - (q > 0 ? 1 : 0));
        return sorted[i];
    };
    double qsum = 0, qmax = 0, msum = 0;
    for (const Request &r : s->requests) {
        const double q = r.admitted - (realtime ? r.arrival : 0.0);
        qsum += q;
        qmax = std::max(qmax, q);
    }
    for (float v : sorted) msum += v;
    *report = vs_server_report{};
    report->wall_s = wall;
    report->node_substeps = node_substeps;
    report->frames = frames;
    report->requests = (uint32_t)s->requests.size();
    report->iterations = iterations;
    report->max_active = max_active;
    report->p50_ms = pct(0.50);
    report->p99_ms = pct(0.99);
    report->mean_ms = sorted.empty() ? 0 : msum / (double)sorted.size();
    report->max_ms = sorted.empty() ? 0 : sorted.back();
    report->mean_queue_s = s->requests.empty() ? 0 : qsum / (double)s->requests.size();
    report->max_queue_s = qmax;
    uint32_t completed = 0;
    for (const Request &r : s->requests) completed += r.finished >= 0;
    report->completed = completed;
    return VS_OK;
}

int vs_server_latencies(const vs_server *s, float *ms, uint64_t cap, uint64_t *count)
{
    if (!s || !count) return VS_E_INVALID;
    *count = s->latencies_ms.size();
    if (ms) std::memcpy(ms, s->latencies_ms.data(), sizeof(float) * std::min<uint64_t>(cap, *count));
    return VS_OK;
}

int vs_server_result(const vs_server *s, uint32_t request_id, float *pos)
{
    if (!s || request_id >= s->requests.size() || !pos) return VS_E_INVALID;
    const Request &r = s->requests[request_id];
    if (r.result.empty()) return VS_E_INVALID;
    std::memcpy(pos, r.result.data(), sizeof(float) * r.result.size());
    return VS_OK;
}

int vs_server_times(const vs_server *s, uint32_t request_id, double *arrival, double *admitted, double *finished)
{
    if (!s || request_id >= s->requests.size()) return VS_E_INVALID;
    const Request &r = s->requests[request_id];
    if (arrival) *arrival = r.arrival;
    if (admitted) *admitted = r.admitted;
    if (finished) *finished = r.finished;
    return VS_OK;
}

}  // extern "C"
