// world.cpp -- layer 2 of vsim.h: many scenes resident on one device, stepped
// as one batch per pass.
//
// Device memory is a set of pools sized once from the HBM budget (node planes,
// cells, bodies, scene uniforms, contact records and buckets, driven records,
// splats). A scene owns one range in each pool. Tile tables list the work of
// every admitted scene per pass; a pass is one launch over its table whatever
// the number of scenes, and a frame is captured once into a hipGraph and
// replayed. The plain path launches the same kernels over one scene's slice of
// each table, the way the reference dispatches one scene (stepper.js:101-108).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "kernels.h"
#include "world_internal.h"

namespace {

using vsim::GridView;
using vsim::NodeView;
using vsim::PairLog;
using vsim::Shape;
using vsim::SplatView;
using vsim::SurfaceView;
using vsim::TileView;

constexpr uint32_t BUCKET_SLOTS = 32;  // contacts.wgsl:75, bodyContact.js:35
// Extension slots a bucket: an overfull bucket's first 32 + EXT_SLOTS
// arrivals are found without the overflow list (384 B a bucket).
constexpr uint32_t EXT_SLOTS = 96;
constexpr int P_NODE = 0, P_CELL = 1, P_SURF = 9, P_DRIVE = 10, P_SPLAT = 11, P_ROW = 12, P_REST = 16, P_REST0 = 17,
              P_BODY = 18, P_LAST = 19, NPASS = 20;
constexpr uint32_t BODY_SWEEP_NODES = 11 * 11 * 11;  // vsim::k_elastic_bodies' LDS

// The four colour-pair (row class) tables.
constexpr bool is_row_pass(int p) { return p >= P_ROW && p < P_ROW + 4; }
constexpr int NODE_PLANES = 13;        // p xyz, prev xyz, vel xyz, w, rest xyz
constexpr int SURF_PLANES = 9;
constexpr int DRIVE_PLANES = 4;
constexpr int SPLAT_OUT_PLANES = 16;    // mean xyz, cov 6, rotation xyzw, axis lengths xyz
constexpr uint32_t EVENT_POOL = 8192;

// First-fit range allocator with coalescing.
struct Pool {
    uint64_t cap = 0, used = 0;
    std::map<uint64_t, uint64_t> free_;
    void init(uint64_t c)
    {
        cap = c;
        used = 0;
        free_.clear();
        if (c) free_[0] = c;
    }
    bool alloc(uint64_t n, uint64_t &out)
    {
        out = 0;
        if (n == 0) return true;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            out = it->first;
            const uint64_t rest = it->second - n, next = it->first + n;
            free_.erase(it);
            if (rest) free_[next] = rest;
            used += n;
            return true;
        }
        return false;
    }
    void release(uint64_t b, uint64_t n)
    {
        if (!n) return;
        used -= n;
        auto next = free_.lower_bound(b);
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == b) {
                b = prev->first;
                n += prev->second;
                free_.erase(prev);
            }
        }
        if (next != free_.end() && b + n == next->first) {
            n += next->second;
            free_.erase(next);
        }
        free_[b] = n;
    }
};

struct BodyRec {
    uint64_t node_begin = 0, node_count = 0;
    uint64_t surf_begin = 0, surf_count = 0;  // its contact records (0 without contact)
    uint64_t color_begin[8] = {}, color_count[8] = {};
    uint64_t splat_begin = 0, splat_count = 0;
    float damping = 0;
    uint32_t cell_dims[3] = {};
    bool dense = false;  // every cell of the lattice box occupied
};

struct SceneRec {
    bool live = false;
    vs_scene_params params{};
    uint64_t body_begin = 0, body_count = 0;
    uint64_t node_begin = 0, node_count = 0;
    uint64_t cell_begin = 0, cell_count = 0;
    uint64_t surf_begin = 0, surf_count = 0;
    uint64_t bucket_begin = 0, bucket_count = 0;
    uint64_t drive_begin = 0, drive_count = 0;
    uint64_t splat_begin = 0, splat_count = 0;
    bool contact = false, stretch = false;
    uint32_t branch = 0;  // chain of a two-chain frame (vs_world_config.branches)
    std::vector<BodyRec> bodies;
    std::vector<uint32_t> splat_perm;  // stored splat -> caller's splat index
};

template <typename T>
struct HostPinned {
    T *p = nullptr;
    size_t n = 0;
    ~HostPinned()
    {
        if (p) (void)hipHostFree(p);
    }
};

}  // namespace

struct vs_world {
    vs_world_config cfg{};
    std::string error;
    hipStream_t stream = nullptr, upload = nullptr;
    // Two-chain frames: the second chain's stream, the fork and join events,
    // and, while a chain is enqueued, its stream and index (-1: one chain).
    hipStream_t stream2 = nullptr, cur_stream = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    int cur_branch = -1;
    hipEvent_t upload_done = nullptr, tables_done = nullptr, params_done = nullptr;
    // Recorded on `stream` by each removal: the work enqueued before it may
    // still read the removed scene's ranges, so admissions, which may hand
    // those ranges out again, wait for it on the upload stream.
    hipEvent_t released = nullptr;
    std::vector<hipEvent_t> events;
    hipEvent_t tune[2] = {};
    uint32_t next_event = 0;

    Pool nodes, cells, bodies, scenes, surface, buckets, driven, splats;
    float *d_nodes = nullptr;
    uint32_t *d_cells = nullptr;
    vs_body *d_bodies = nullptr;
    vs_scene_params *d_scenes = nullptr;
    uint32_t *d_surf = nullptr;
    float4 *d_surf_pay = nullptr;  // per record: node position and radius, then prev position and inverse mass, at bin time
    uint32_t *d_grid = nullptr;
    uint2 *d_ovf = nullptr;            // records that found their bucket full: (record, grid bucket)
    uint32_t *d_ovf_count = nullptr;   // 4 words: records listed, buckets listed, list full (settle empties them), spare;
                                       // 4 more for the parity-1 grid
    uint2 *d_ovb = nullptr;            // overfull buckets this substep: (grid bucket, scene)
    uint32_t *d_ext = nullptr;         // EXT_SLOTS more slots a bucket (arrivals 32 .. 32 + EXT_SLOTS - 1)
    uint32_t *d_summary = nullptr;     // per bucket: slots held and sole body, for collide (GridView::summary)
    // The parity-1 grid's bucket counters, then its summaries (2 x max_buckets):
    // with the bin folded into the body sweep, substep s bins into grid s & 1
    // while the same launch settles substep s - 1 in the other; slots,
    // extension slots and list entries are shared.
    uint32_t *d_parity1 = nullptr;
    uint32_t *d_drive = nullptr;
    float *d_splat_in = nullptr, *d_splat_out = nullptr;
    int32_t *d_pairlog = nullptr;
    uint32_t pair_cap = 0;

    // Tile tables: per pass a 16-byte header then tiles; per-scene counts for
    // the plain path.
    uint32_t *d_tiles[NPASS] = {};
    uint64_t tile_cap[NPASS] = {};
    uint32_t *d_scene_counts = nullptr;  // [NPASS][max_scenes]
    uint32_t *d_branch_counts = nullptr; // [NPASS][2]
    HostPinned<uint32_t> h_branch_counts;
    uint32_t branch_first[NPASS][2] = {}, branch_ntiles[NPASS][2] = {};
    HostPinned<uint32_t> h_tiles[NPASS];
    // The colour-pair tables once more, in the order the batched launches
    // take them (rows of the most lanes first, node tiles last); the tables
    // above keep each scene's tiles together for the per-scene launches.
    uint32_t *d_rows_by_work[4] = {};
    HostPinned<uint32_t> h_rows_by_work[4];
    HostPinned<uint32_t> h_scene_counts;
    std::vector<uint32_t> tile_count;                 // [NPASS]
    std::vector<std::vector<uint32_t>> scene_first;   // [NPASS][max_scenes]
    std::vector<std::vector<uint32_t>> scene_ntiles;  // [NPASS][max_scenes]
    HostPinned<vs_scene_params> h_scenes;
    std::vector<vs_body> h_bodies;

    std::vector<SceneRec> recs;
    bool tables_dirty = true, params_dirty = true, any_stretch = false, any_contact = false;
    // Every contact scene's bodies are body sweeps that run the node passes:
    // each substep's settle rides on the next substep's sweep.
    bool settle_fold = false;
    uint32_t grid_parity = 0;  // the grid the contact launches being enqueued use
    std::vector<std::vector<char>> pending;  // host copies kept alive until uploaded

    // Captured frame.
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint32_t> graph_key;
    uint32_t captures = 0;

    uint32_t splat_mode = VS_SPLAT_COV;  // outputs of the splat pass being enqueued
    uint32_t block_node = 256, block_cells = 256, block_surf = 256, block_drive = 64, block_splat = 256;
    uint32_t block_rest = 256;   // the node pass of P_REST: 64 when its runs are short (folded scenes)
    uint32_t block_rest0 = 256;  // the frame's first predict over P_REST0, likewise
    uint32_t block_body = 256;   // P_BODY
    uint32_t block_last = 256;   // the frame's last finalize over P_LAST, as block_rest
    uint32_t tile_nodes = 256, tile_cells = 256, tile_surf = 256, tile_drive = 64, tile_splat = 256;
};

namespace {

int fail(vs_world *w, int code, const std::string &m)
{
    if (w) w->error = m;
    return code;
}

#define HIPW(w, call)                                                                  \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess) return fail((w), VS_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

float *plane(vs_world *w, int k) { return w->d_nodes + (size_t)k * w->cfg.max_nodes; }

// Splat outputs of a frame: 0 none, else VS_SPLAT_COV and/or VS_SPLAT_AXES
// (any other non-zero value means the covariances).
uint32_t splat_mode(int32_t splats)
{
    if (!splats) return 0;
    const uint32_t m = (uint32_t)splats & (VS_SPLAT_COV | VS_SPLAT_AXES);
    return m ? m : VS_SPLAT_COV;
}

NodeView node_view(vs_world *w)
{
    return NodeView{ plane(w, 0), plane(w, 1), plane(w, 2), plane(w, 3), plane(w, 4),
                     plane(w, 5), plane(w, 6), plane(w, 7), plane(w, 8), plane(w, 9) };
}

SurfaceView surf_view(vs_world *w)
{
    const size_t S = w->cfg.max_surface;
    uint32_t *s = w->d_surf;
    return SurfaceView{ s,
                        s + S,
                        reinterpret_cast<float *>(s + 2 * S),
                        s + 3 * S,
                        s + 4 * S,
                        reinterpret_cast<float *>(s + 5 * S),
                        reinterpret_cast<float *>(s + 6 * S),
                        reinterpret_cast<float *>(s + 7 * S),
                        reinterpret_cast<float *>(s + 8 * S),
                        w->d_surf_pay,
                        w->d_surf_pay + w->cfg.max_surface };
}

// A bucket is overfull past at least one slot, so no more than half the
// records can make buckets overfull.
uint32_t ovb_cap(const vs_world *w) { return (uint32_t)(w->cfg.max_surface / 2u + 1u); }

GridView grid_view(vs_world *w, uint32_t parity)
{
    const size_t B = w->cfg.max_buckets;
    return GridView{ parity ? w->d_parity1 : w->d_grid,
                     w->d_grid + B,
                     w->d_ovf,
                     w->d_ovf_count + (parity ? 4 : 0),
                     (uint32_t)w->cfg.max_surface,
                     w->d_ovb,
                     ovb_cap(w),
                     w->d_ext,
                     EXT_SLOTS,
                     parity ? w->d_parity1 + B : w->d_summary };
}

GridView grid_view(vs_world *w) { return grid_view(w, w->grid_parity); }

vsim::DriveView drive_view(vs_world *w)
{
    const size_t D = w->cfg.max_driven;
    uint32_t *d = w->d_drive;
    return vsim::DriveView{ d, reinterpret_cast<float *>(d + D), reinterpret_cast<float *>(d + 2 * D),
                            reinterpret_cast<float *>(d + 3 * D) };
}

SplatView splat_view(vs_world *w)
{
    const size_t G = w->cfg.max_splats;
    float *in = w->d_splat_in, *out = w->d_splat_out;
    return SplatView{ in + 4 * G,  in + 5 * G,     in + 6 * G,     reinterpret_cast<const float4 *>(in),
                      in + 7 * G,  in + 8 * G,     in + 9 * G,     out,
                      out + G,     out + 2 * G,    out + 3 * G,    out + 9 * G,
                      out + 13 * G, (uint32_t)G,   w->splat_mode };
}

uint32_t pass_block(const vs_world *w, int p)
{
    if (p == P_NODE) return w->block_node;
    if (p == P_REST) return w->block_rest;
    if (p == P_LAST) return w->block_last;
    if (p == P_REST0) return w->block_rest0;
    if ((p >= P_CELL && p < P_CELL + 8) || is_row_pass(p)) return w->block_cells;
    if (p == P_BODY) return w->block_body;
    if (p == P_SURF) return w->block_surf;
    if (p == P_DRIVE) return w->block_drive;
    return w->block_splat;
}

uint32_t pass_tile(const vs_world *w, int p)
{
    if (p == P_NODE || p == P_REST || p == P_REST0 || p == P_LAST) return w->tile_nodes;
    if (p >= P_CELL && p < P_CELL + 8) return w->tile_cells;
    if (p == P_SURF) return w->tile_surf;
    if (p == P_DRIVE) return w->tile_drive;
    return w->tile_splat;
}

// A small dense body of a scene whose node passes do not ride on its colour
// passes (contact, tensile sweeps or drive in between) runs its eight colours
// in one launch, in LDS (k_elastic_bodies).
bool sweep_body(const vs_world *w, const SceneRec &r, const BodyRec &b)
{
    return !w->cfg.separate_colours && !w->cfg.colour_launches && !(r.params.flags & VS_SCENE_FOLD) && b.dense &&
           b.node_count <= BODY_SWEEP_NODES;
}

// A dense body whose rows fit one workgroup runs its colours in pairs, by
// rows (k_elastic_rows); any other body runs one colour per pass.
bool row_body(const vs_world *w, const BodyRec &b)
{
    return !w->cfg.separate_colours && b.dense && b.cell_dims[0] / 2u + 1u <= w->block_cells;
}

// The tile table of an elastic launch: colours 0..7, then the four row
// classes (colours 2c and 2c+1) as 8..11, then whole small bodies as 12.
int elastic_table(int color) { return color < 8 ? P_CELL + color : color < 12 ? P_ROW + (color - 8) : P_BODY; }

// A row body of a VS_SCENE_FOLD scene has the nodes its last colour pair
// writes finalized and predicted inside that pass (unless the node passes
// are kept separate).
bool fold_body(const vs_world *w, const SceneRec &r, const BodyRec &b)
{
    return !w->cfg.separate_node_passes && (r.params.flags & VS_SCENE_FOLD) && row_body(w, b) && !sweep_body(w, r, b);
}

// A row body of a VS_SCENE_FOLD_PREDICT scene has the nodes its first colour
// pair writes predicted inside that pass at a frame's first substep.
// A sweep body of a scene with nothing between its predict and its first
// colour (no drive) has its node passes in its body sweep: the previous
// substep's finalize and this one's predict, or a frame's first predict.
bool sweep_fold_body(const vs_world *w, const SceneRec &r, const BodyRec &b)
{
    return sweep_body(w, r, b) && !w->cfg.separate_node_passes && (r.params.flags & VS_SCENE_FOLD_PREDICT);
}

// A contact scene whose bodies are all body sweeps that run the node passes
// has each substep's settle applied by the next substep's sweep, after its
// load and before the finalize (nothing runs between them).
bool scene_settle_fold(const vs_world *w, const SceneRec &r)
{
    if (!r.live || !r.contact || r.bodies.empty()) return false;
    for (const BodyRec &b : r.bodies)
        if (!sweep_fold_body(w, r, b)) return false;
    return true;
}

bool fold_predict_body(const vs_world *w, const SceneRec &r, const BodyRec &b)
{
    return !w->cfg.separate_node_passes && (r.params.flags & VS_SCENE_FOLD_PREDICT) && row_body(w, b) &&
           !sweep_body(w, r, b);
}

// The nodes of a row body that class `cls` leaves to a node pass: class 3
// (the last pair) writes node rows y = 1 .. 2 floor(cdy/2), class 0 (the
// first) y = 0 .. 2 ceil(cdy/2) - 1, likewise along z (vsim::Leftover).
uint32_t leftover_count(const BodyRec &b, uint32_t cls)
{
    return vsim::Leftover(b.cell_dims, cls).total;
}

// Rebuilds every tile table from the admitted scenes, in scene-id order.
int rebuild_tables(vs_world *w)
{
    HIPW(w, hipEventSynchronize(w->tables_done));
    const uint32_t S = w->cfg.max_scenes;
    w->any_stretch = false;
    w->any_contact = false;
    // Two-chain frames: scenes split into two groups of about equal work
    // (cells and nodes), largest first to the lighter group. Contact scenes
    // all go to chain 0: they share the grid's overflow lists and counters.
    {
        uint64_t load[2] = { 0, 0 };
        std::vector<uint32_t> order;
        for (uint32_t s = 0; s < S; s++) {
            SceneRec &r = w->recs[s];
            r.branch = 0;
            if (!r.live) continue;
            if (w->cfg.branches >= 2 && r.contact) load[0] += r.cell_count + r.node_count;
            else order.push_back(s);
        }
        if (w->cfg.branches >= 2) {
            std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
                return w->recs[a].cell_count + w->recs[a].node_count > w->recs[b].cell_count + w->recs[b].node_count;
            });
            for (uint32_t s : order) {
                const uint32_t b = load[1] < load[0] ? 1u : 0u;
                w->recs[s].branch = b;
                load[b] += w->recs[s].cell_count + w->recs[s].node_count;
            }
        }
    }
    for (int p = 0; p < NPASS; p++) {
        vs_tile *tiles = reinterpret_cast<vs_tile *>(w->h_tiles[p].p + 4);
        uint64_t n = 0;
        const uint32_t T = pass_tile(w, p);
        // Chain 0's scenes first, then chain 1's, each in scene-id order.
        for (uint32_t pass_br = 0; pass_br < 2; pass_br++) {
            w->branch_first[p][pass_br] = (uint32_t)n;
            for (uint32_t s = 0; s < S; s++) {
                SceneRec &r = w->recs[s];
                if (r.branch != pass_br) continue;
                w->scene_first[p][s] = (uint32_t)n;
                const uint64_t start = n;
                if (r.live) {
                    auto emit = [&](uint64_t begin, uint64_t count, uint32_t body) {
                        for (uint64_t o = 0; o < count; o += T) {
                            if (n >= w->tile_cap[p]) return false;
                            tiles[n++] = vs_tile{ (uint32_t)(begin + o), (uint32_t)std::min<uint64_t>(T, count - o), body,
                                                  s };
                        }
                        return true;
                    };
                    bool ok = true;
                    if (p == P_NODE) {
                        for (size_t b = 0; b < r.bodies.size(); b++)
                            ok &= emit(r.bodies[b].node_begin, r.bodies[b].node_count, (uint32_t)(r.body_begin + b));
                    } else if (p == P_REST || p == P_REST0 || p == P_LAST) {
                        // P_LAST: the nodes of bodies whose last pair does not
                        // finalize them; P_REST: those, but for bodies whose body
                        // sweep finalizes and predicts them at the next substep's
                        // start; P_REST0: the nodes of bodies whose first pair or
                        // body sweep does not predict them (a folded body's
                        // leftover nodes go to node tiles of that pair's launch).
                        for (size_t b = 0; b < r.bodies.size() && ok; b++) {
                            const BodyRec &br = r.bodies[b];
                            const bool skip = p == P_LAST ? fold_body(w, r, br)
                                              : p == P_REST ? fold_body(w, r, br) || sweep_fold_body(w, r, br)
                                                            : fold_predict_body(w, r, br) || sweep_fold_body(w, r, br);
                            if (skip) continue;
                            ok &= emit(br.node_begin, br.node_count, (uint32_t)(r.body_begin + b));
                        }
                    } else if (p < P_CELL + 8) {
                        const int c = p - P_CELL;
                        for (size_t b = 0; b < r.bodies.size(); b++)
                            if (!row_body(w, r.bodies[b]) && !sweep_body(w, r, r.bodies[b]))
                                ok &= emit(r.bodies[b].color_begin[c], r.bodies[b].color_count[c],
                                           (uint32_t)(r.body_begin + b));
                    } else if (p == P_BODY) {
                        for (size_t b = 0; b < r.bodies.size() && ok; b++)
                            if (sweep_body(w, r, r.bodies[b])) {
                                // One tile a body, carrying its contact records.
                                if (n >= w->tile_cap[p]) {
                                    ok = false;
                                    break;
                                }
                                tiles[n++] = vs_tile{ (uint32_t)r.bodies[b].surf_begin, (uint32_t)r.bodies[b].surf_count,
                                                      (uint32_t)(r.body_begin + b), s };
                            }
                    } else if (is_row_pass(p)) {
                        // Whole rows of the class, as many as fit one workgroup.
                        const uint32_t cls = (uint32_t)(p - P_ROW);
                        for (size_t b = 0; b < r.bodies.size() && ok; b++) {
                            const BodyRec &br = r.bodies[b];
                            if (!row_body(w, br) || sweep_body(w, r, br)) continue;
                            const uint64_t ny = (br.cell_dims[1] + 1u - (cls & 1u)) / 2u;
                            const uint64_t nz = (br.cell_dims[2] + 1u - (cls >> 1)) / 2u;
                            const uint64_t per = w->block_cells / (br.cell_dims[0] / 2u + 1u);
                            const uint64_t rows = ny * nz;
                            const uint32_t id = (uint32_t)(r.body_begin + b);
                            for (uint64_t o = 0; o < rows; o += per) {
                                if (n >= w->tile_cap[p]) {
                                    ok = false;
                                    break;
                                }
                                tiles[n++] = vs_tile{ (uint32_t)o, (uint32_t)std::min<uint64_t>(per, rows - o), id, s };
                            }
                            // The nodes the first pair predicts (class 0) or the
                            // last pair finalizes (class 3) that its rows leave
                            // out, as node tiles of the same launch: full
                            // workgroups of entries of the body's leftover list.
                            const bool nodes = cls == 3 ? fold_body(w, r, br) : cls == 0 && fold_predict_body(w, r, br);
                            const uint32_t left = ok && nodes ? leftover_count(br, cls) : 0u;
                            for (uint32_t o = 0; o < left; o += w->block_cells) {
                                if (n >= w->tile_cap[p]) {
                                    ok = false;
                                    break;
                                }
                                tiles[n++] = vs_tile{ o, std::min(w->block_cells, left - o) | VS_NODE_TILE, id, s };
                            }
                        }
                    } else if (p == P_SURF) {
                        if (r.contact) ok &= emit(r.surf_begin, r.surf_count, (uint32_t)r.body_begin);
                    } else if (p == P_DRIVE) {
                        ok &= emit(r.drive_begin, r.drive_count, (uint32_t)r.body_begin);
                    } else {
                        for (size_t b = 0; b < r.bodies.size(); b++)
                            ok &= emit(r.bodies[b].splat_begin, r.bodies[b].splat_count, (uint32_t)(r.body_begin + b));
                    }
                    if (!ok) return fail(w, VS_E_NOMEM, "tile table overflow");
                    w->any_stretch |= r.stretch;
                    w->any_contact |= r.contact;
                }
                w->scene_ntiles[p][s] = (uint32_t)(n - start);
                w->h_scene_counts.p[(size_t)p * S + s] = (uint32_t)(n - start);
            }
            w->branch_ntiles[p][pass_br] = (uint32_t)n - w->branch_first[p][pass_br];
            w->h_branch_counts.p[2 * p + pass_br] = w->branch_ntiles[p][pass_br];
        }
        w->tile_count[p] = (uint32_t)n;
        if (p == P_BODY) {
            // Workgroups of 256 lanes: a body's colour has at most 125 cells,
            // but its loads of the nodes into LDS, node passes and stores
            // run four times as wide as with 64 (the 512-body workload's
            // sweep: 37.2 us a substep with 64 lanes, 32.7 with 128, 30.8
            // with 256).
            w->block_body = 256u;
        }
        if (p == P_REST || p == P_REST0 || p == P_LAST) {
            // Tables of small bodies (tiles of under 128 nodes on average)
            // waste less with 64-lane workgroups than with 256.
            uint64_t items = 0;
            for (uint64_t i = 0; i < n; i++) items += tiles[i].count;
            (p == P_REST ? w->block_rest : p == P_REST0 ? w->block_rest0 : w->block_last) =
                n && items < 128 * n ? 64u : 256u;
        }
        w->h_tiles[p].p[0] = (uint32_t)n;
        w->h_tiles[p].p[1] = w->h_tiles[p].p[2] = w->h_tiles[p].p[3] = 0;
        HIPW(w, hipMemcpyAsync(w->d_tiles[p], w->h_tiles[p].p, 16 + n * sizeof(vs_tile), hipMemcpyHostToDevice,
                               w->stream));
        if (is_row_pass(p)) {
            // Tiles of one launch touch disjoint nodes, so their order is
            // free: the batched launch dispatches the rows of the most lanes
            // first and the short tiles (a body's last rows, node tiles)
            // last, where they fill the launch's tail.
            uint32_t *h = w->h_rows_by_work[p - P_ROW].p;
            vs_tile *sorted = reinterpret_cast<vs_tile *>(h + 4);
            std::memcpy(h, w->h_tiles[p].p, 16 + n * sizeof(vs_tile));
            auto lanes = [&](const vs_tile &t) -> uint64_t {
                if (t.count & VS_NODE_TILE) return 0;
                return (uint64_t)t.count * (w->h_bodies[t.body].cell_dims[0] / 2u + 1u);
            };
            // Within each chain's run of the table.
            for (int br = 0; br < 2; br++) {
                vs_tile *first = sorted + w->branch_first[p][br];
                std::stable_sort(first, first + w->branch_ntiles[p][br],
                                 [&](const vs_tile &a, const vs_tile &b) { return lanes(a) > lanes(b); });
            }
            HIPW(w, hipMemcpyAsync(w->d_rows_by_work[p - P_ROW], h, 16 + n * sizeof(vs_tile), hipMemcpyHostToDevice,
                                   w->stream));
        }
    }
    {
        bool any = false, all = true;
        for (uint32_t s = 0; s < S; s++) {
            const SceneRec &r = w->recs[s];
            if (!r.live || !r.contact) continue;
            any = true;
            all = all && scene_settle_fold(w, r);
        }
        w->settle_fold = any && all;
    }
    HIPW(w, hipMemcpyAsync(w->d_scene_counts, w->h_scene_counts.p, sizeof(uint32_t) * NPASS * S,
                           hipMemcpyHostToDevice, w->stream));
    HIPW(w, hipMemcpyAsync(w->d_branch_counts, w->h_branch_counts.p, sizeof(uint32_t) * NPASS * 2,
                           hipMemcpyHostToDevice, w->stream));
    HIPW(w, hipEventRecord(w->tables_done, w->stream));
    w->tables_dirty = false;
    return VS_OK;
}

int upload_params(vs_world *w)
{
    HIPW(w, hipEventSynchronize(w->params_done));
    for (uint32_t s = 0; s < w->cfg.max_scenes; s++) {
        if (w->recs[s].live) w->h_scenes.p[s] = w->recs[s].params;
    }
    HIPW(w, hipMemcpyAsync(w->d_scenes, w->h_scenes.p, sizeof(vs_scene_params) * w->cfg.max_scenes,
                           hipMemcpyHostToDevice, w->stream));
    HIPW(w, hipEventRecord(w->params_done, w->stream));
    w->params_dirty = false;
    return VS_OK;
}

int prepare(vs_world *w)
{
    // Work admitted on the upload stream must land before the next frame.
    HIPW(w, hipStreamWaitEvent(w->stream, w->upload_done, 0));
    if (w->tables_dirty) {
        if (int rc = rebuild_tables(w)) return rc;
    }
    if (w->params_dirty) {
        if (int rc = upload_params(w)) return rc;
    }
    return VS_OK;
}

// A tile view over the whole table (scene < 0) or one scene's slice of it.
TileView tile_view(vs_world *w, int p, int scene)
{
    // XCD-chunked launches give each XCD one contiguous run of the table, so
    // they keep the scene order (a region stays on one XCD from pass to pass).
    const int br = scene < 0 ? w->cur_branch : -1;
    if (scene < 0 && is_row_pass(p) && !w->cfg.xcd_chunked) {
        uint32_t *d = w->d_rows_by_work[p - P_ROW];
        if (br >= 0)
            return TileView{ w->d_branch_counts + 2 * p + br,
                             reinterpret_cast<const vs_tile *>(d + 4) + w->branch_first[p][br] };
        return TileView{ d, reinterpret_cast<const vs_tile *>(d + 4) };
    }
    const vs_tile *tiles = reinterpret_cast<const vs_tile *>(w->d_tiles[p] + 4);
    if (br >= 0) return TileView{ w->d_branch_counts + 2 * p + br, tiles + w->branch_first[p][br] };
    if (scene < 0) return TileView{ w->d_tiles[p], tiles };
    return TileView{ w->d_scene_counts + (size_t)p * w->cfg.max_scenes + scene, tiles + w->scene_first[p][scene] };
}

uint32_t grid_for(vs_world *w, int p, int scene)
{
    const uint32_t n = scene >= 0 ? w->scene_ntiles[p][scene]
                       : w->cur_branch >= 0 ? w->branch_ntiles[p][w->cur_branch]
                                            : w->tile_count[p];
    if (n == 0) return 0;
    // Persistent: a fixed grid walks the table, so a captured frame stays
    // valid whatever the admitted scenes (only a pass appearing or vanishing
    // changes the shape).
    if (scene < 0 && w->cfg.persistent) return w->cfg.persistent;
    return n;
}

// Events around launches, for vs_world_profile.
struct Profile {
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, size_t>> marks;  // kind, index of the start event
    size_t used = 0;
    hipEvent_t next()
    {
        if (used == pool.size()) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[used++];
    }
    ~Profile()
    {
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    }
};

// Enqueues one launch of a pass over the whole table (scene < 0) or one
// scene's slice: `color` picks the elastic colour, `limit_only` the
// stretch-only sweeps. A pass with no tiles enqueues nothing.
int enqueue_pass(vs_world *w, int kind, int color, int scene, uint32_t limit_only, uint32_t node_mode = 0)
{
    using namespace vsim;
    auto shape = [&](int p) { return Shape{ grid_for(w, p, scene), pass_block(w, p), scene < 0 ? w->cfg.xcd_chunked : 0u }; };
    const NodeView n = node_view(w);
    hipStream_t st = w->cur_stream ? w->cur_stream : w->stream;
    Shape sh{};
    switch (kind) {
    case VS_P_PREDICT: {
        // color 1: only the nodes the first colour pair does not predict.
        const int p = color ? P_REST0 : P_NODE;
        sh = shape(p);
        if (sh.grid) HIPW(w, launch_predict(st, sh, n, w->d_scenes, tile_view(w, p, scene)));
        return VS_OK;
    }
    case VS_P_DRIVE:
        sh = shape(P_DRIVE);
        if (sh.grid) HIPW(w, launch_drive(st, sh, n, w->d_scenes, drive_view(w), tile_view(w, P_DRIVE, scene), -1));
        return VS_OK;
    case VS_P_ELASTIC: {
        sh = shape(elastic_table(color));
        if (!sh.grid) return VS_OK;
        // The tensile limit's code only where a body of the launch has one.
        const bool stretch = scene >= 0 ? w->recs[scene].stretch : w->any_stretch;
        if (color == 12)
            HIPW(w, launch_elastic_bodies(st, sh, n, w->d_bodies, w->d_scenes, tile_view(w, P_BODY, scene), limit_only,
                                          node_mode & 3u, stretch, surf_view(w),
                                          grid_view(w, node_mode & 8u ? w->grid_parity ^ 1u : w->grid_parity),
                                          grid_view(w, w->grid_parity), (node_mode >> 2) & 3u));
        else if (color >= 8)
            HIPW(w, launch_elastic_rows(st, sh, n, w->d_bodies, w->d_scenes, tile_view(w, elastic_table(color), scene),
                                        (uint32_t)(color - 8), limit_only, node_mode, stretch));
        else
            HIPW(w, launch_elastic(st, sh, n, w->d_bodies, w->d_scenes, w->d_cells, tile_view(w, P_CELL + color, scene),
                                   limit_only, stretch));
        return VS_OK;
    }
    case VS_P_FINALIZE:
        // Every node, or, after folded colour pairs, the ones they leave.
        sh = shape(P_LAST);
        if (sh.grid) HIPW(w, launch_finalize(st, sh, n, w->d_bodies, w->d_scenes, tile_view(w, P_LAST, scene)));
        return VS_OK;
    case VS_P_FINALIZE_PREDICT:
        sh = shape(P_REST);
        if (sh.grid)
            HIPW(w, launch_finalize_predict(st, sh, n, w->d_bodies, w->d_scenes, tile_view(w, P_REST, scene)));
        return VS_OK;
    case VS_P_SPLAT:
        // Contiguous tile runs per XCD (the grid rounded up to a multiple of
        // 8; surplus workgroups find no tile).
        sh = shape(P_SPLAT);
        sh.grid = (sh.grid + 7u) / 8u * 8u;
        sh.xcd = 1;
        if (sh.grid)
            HIPW(w, launch_splats(st, sh, n.px, n.py, n.pz, w->d_bodies, splat_view(w), tile_view(w, P_SPLAT, scene)));
        return VS_OK;
    default:
        break;
    }
    sh = shape(P_SURF);
    if (!sh.grid) return VS_OK;
    const SurfaceView sv = surf_view(w);
    const GridView g = grid_view(w);
    const TileView t = tile_view(w, P_SURF, scene);
    switch (kind) {
    case VS_P_BIN: HIPW(w, launch_contact_bin(st, sh, n, w->d_scenes, sv, g, t)); break;
    case VS_P_SORT: HIPW(w, launch_contact_sort(st, sh, w->d_scenes, sv, g, t)); break;
    case VS_P_COLLIDE: {
        const PairLog log{ w->d_pairlog,
                           w->d_pairlog ? reinterpret_cast<uint32_t *>(w->d_pairlog + w->cfg.max_surface) : nullptr,
                           w->pair_cap };
        HIPW(w, launch_contact_collide(st, sh, n, w->d_bodies, w->d_scenes, sv, g, t, log, w->tile_surf));
        break;
    }
    case VS_P_SETTLE: HIPW(w, launch_contact_settle(st, sh, n, w->d_scenes, sv, g, t)); break;
    default: return fail(w, VS_E_INVALID, "unknown pass kind");
    }
    return VS_OK;
}

// Enqueues one frame: `substeps` substeps in Stepper.sequence order
// (stepper.js:101-108, solve order simulator.js:83), then the splats.
int enqueue_frame(vs_world *w, uint32_t substeps, uint32_t splats, int scene, Profile *prof = nullptr)
{
    if (splats) w->splat_mode = splats;
    bool stretch = w->any_stretch, contact = w->any_contact;
    if (scene >= 0) {
        stretch = w->recs[scene].stretch;
        contact = w->recs[scene].contact;
    }
    const bool merge = !w->cfg.separate_node_passes;
    const bool settle_fold = scene >= 0 ? scene_settle_fold(w, w->recs[scene]) : w->settle_fold;
    // The launch that takes a substep's settle covers the contact scenes
    // (with two chains, they are all chain 0's).
    const bool settle_here = settle_fold && (scene >= 0 || w->cur_branch != 1);
    // Without the tensile sweeps the same launch also bins the substep's
    // records (after its colours): substep s bins into grid s & 1 while the
    // launch settles substep s - 1 in the other, and sort, collide and
    // settle of substep s use grid s & 1.
    const bool bin_fold = settle_fold && !stretch;
    hipStream_t st = w->cur_stream ? w->cur_stream : w->stream;
    // With a profile, events bracket every launch of a non-empty pass.
    auto pass = [&](int kind, int color, uint32_t limit_only, uint32_t node_mode = 0) -> int {
        const int p = kind == VS_P_ELASTIC ? elastic_table(color)
                      : kind == VS_P_DRIVE ? P_DRIVE
                      : kind == VS_P_SPLAT ? P_SPLAT
                      : kind == VS_P_PREDICT ? (color ? P_REST0 : P_NODE)
                      : kind == VS_P_FINALIZE ? P_LAST
                      : kind == VS_P_FINALIZE_PREDICT ? P_REST
                                                      : P_SURF;
        const bool timed = prof && grid_for(w, p, scene) != 0;
        if (timed) {
            prof->marks.emplace_back(kind, prof->used);
            if (kind == VS_P_ELASTIC && color >= 8 && color < 12)
                prof->marks.emplace_back(node_mode ? VS_P_PAIR_NODES : VS_P_PAIR, prof->used);
            HIPW(w, hipEventRecord(prof->next(), st));
        }
        if (int rc = enqueue_pass(w, kind, color, scene, limit_only, node_mode)) return rc;
        if (timed) HIPW(w, hipEventRecord(prof->next(), st));
        return VS_OK;
    };
    // Nothing runs between one substep's finalize and the next one's
    // predict, so inside a frame the two share one node pass. In a scene
    // with nothing between its last colour and its finalize either, the last
    // colour-pair pass runs both for the nodes it writes, and the node pass
    // takes the rest. At the frame's first substep, in a scene with no drive
    // between predict and the first colour, the first colour-pair pass
    // predicts the nodes it holds and the predict pass takes the rest.
    for (uint32_t s = 0; s < substeps; s++) {
        const bool last = s + 1 == substeps;
        w->grid_parity = bin_fold ? s & 1u : 0u;
        if (s == 0 || !merge)
            if (int rc = pass(VS_P_PREDICT, merge ? 1 : 0, 0)) return rc;
        if (int rc = pass(VS_P_DRIVE, 0, 0)) return rc;
        // Colours in order; row class c (colours 2c, 2c+1 of the row bodies)
        // goes with the two colour passes of the other bodies.
        for (int sweep = 0; sweep < (stretch ? 4 : 1); sweep++) {
            // Small bodies' eight colours in one launch (their own nodes
            // only); at a sweep's start the bodies that fold their node
            // passes there predict (a frame's first substep) or finalize
            // the previous substep and predict.
            // (With the settle folded, the previous substep's settle runs
            // there too: node_mode bit 4; with the bin folded, this
            // substep's bin after the colours: bit 8.)
            const uint32_t sweep_mode = merge && sweep == 0 ? (s == 0 ? 3u : 1u) : 0u;
            if (int rc = pass(VS_P_ELASTIC, 12, sweep > 0 ? 1u : 0u,
                              sweep_mode | (settle_here && s > 0 && sweep == 0 ? 4u : 0u) |
                                  (settle_here && bin_fold ? 8u : 0u)))
                return rc;
            for (int c = 0; c < 4; c++)
                for (int e : { 8 + c, 2 * c, 2 * c + 1 }) {
                    uint32_t fold = merge && sweep == 0 && e == 11 ? (last ? 2u : 1u) : 0u;
                    if (merge && sweep == 0 && s == 0 && e == 8) fold = 3u;
                    if (int rc = pass(VS_P_ELASTIC, e, sweep > 0 ? 1u : 0u, fold)) return rc;
                }
        }
        if (contact)
            for (int kind : { VS_P_BIN, VS_P_SORT, VS_P_COLLIDE, VS_P_SETTLE }) {
                if (kind == VS_P_SETTLE && settle_fold && !last) continue;  // the next sweep's
                if (kind == VS_P_BIN && bin_fold) continue;                 // the sweep's
                if (int rc = pass(kind, 0, 0)) return rc;
            }
        if (int rc = pass(merge && !last ? VS_P_FINALIZE_PREDICT : VS_P_FINALIZE, 0, 0)) return rc;
    }
    w->grid_parity = 0;
    if (splats)
        if (int rc = pass(VS_P_SPLAT, 0, 0)) return rc;
    return VS_OK;
}

void release_ranges(vs_world *w, const SceneRec &r)
{
    w->bodies.release(r.body_begin, r.body_count);
    w->nodes.release(r.node_begin, r.node_count);
    w->cells.release(r.cell_begin, r.cell_count);
    w->surface.release(r.surf_begin, r.surf_count);
    w->buckets.release(r.bucket_begin, r.bucket_count);
    w->driven.release(r.drive_begin, r.drive_count);
    w->splats.release(r.splat_begin, r.splat_count);
}

template <typename T>
const T *keep(vs_world *w, std::vector<T> &&v)
{
    std::vector<char> bytes(v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(bytes.data(), v.data(), bytes.size());
    w->pending.push_back(std::move(bytes));
    return reinterpret_cast<const T *>(w->pending.back().data());
}

// Per-body sizes of an admission, and the device temporaries of a lattice
// built from occupancy.
struct BodyPlan {
    uint32_t n_cells = 0, n_surface = 0, color_count[8] = {};
    float origin[3] = {}, h = 0.f;
    bool device = false;
    uint8_t *d_occ = nullptr;
    uint32_t *d_counts = nullptr, *d_offsets = nullptr;
    uint32_t tiles = 0, node_tiles = 0, tile_base[9] = {};
    std::vector<uint32_t> counts, offsets;
};

// buildLattice's inverse mass of a node with k occupied incident cells: the
// f64 corner mass density*h*h*h/8 added k times into an f32 slot, then
// 1/mass in f64, stored f32 (lattice.js:190-233).
void mass_table(double density, double h, float out[9])
{
    const double corner = density * h * h * h / 8;
    float mass = 0.f;
    out[0] = 0.f;
    for (int k = 1; k <= 8; k++) {
        mass = (float)((double)mass + corner);
        out[k] = mass > 0.f ? (float)(1.0 / (double)mass) : 0.f;
    }
}

vsim::LatticeView lattice_view(const vs_body_desc &bd, const BodyPlan &pl, uint64_t node_offset, uint64_t body)
{
    vsim::LatticeView L{};
    L.occ = pl.d_occ;
    for (int a = 0; a < 3; a++) {
        L.cell_dims[a] = bd.cell_dims[a];
        L.origin[a] = bd.origin64[a];
        L.velocity[a] = bd.velocity[a];
    }
    L.node_offset = (uint32_t)node_offset;
    L.body = (uint32_t)body;
    L.radius = bd.contact_radius;
    L.h = bd.h64;
    mass_table(bd.density, bd.h64, L.wtab);
    for (int c = 0; c < 9; c++) L.tile_base[c] = pl.tile_base[c];
    L.tile_count = pl.d_counts;
    L.tile_offset = pl.d_offsets;
    return L;
}

float damp_factor(float damping, float dt)
{
    // exp(-damping * dt) of finalize.wgsl:72: the f32 product of the two
    // uniforms, exponentiated in double and rounded once.
    const float a = -(damping * dt);
    return (float)std::exp((double)a);
}

}  // namespace

extern "C" {

static int world_init(vs_world *W);

int vs_world_create(const vs_world_config *cfg, vs_world **out)
{
    if (!cfg || !out) return VS_E_INVALID;
    *out = nullptr;
    vs_world *w = new vs_world();
    w->cfg = *cfg;
    const int rc = world_init(w);
    if (rc) {
        vs_world_destroy(w);
        return rc;
    }
    *out = w;
    return VS_OK;
}

static int world_init(vs_world *w)
{
    vs_world_config &c = w->cfg;
    if (c.max_scenes == 0 || c.max_bodies == 0 || c.max_nodes == 0) return VS_E_INVALID;
    if (c.max_bodies >= vsim::SUMMARY_MIXED) return VS_E_INVALID;  // body ids fit a bucket summary
    // Kernels address node planes by 32-bit byte offsets.
    if (c.max_nodes >= (1ull << 29)) return fail(w, VS_E_INVALID, "max_nodes must stay below 2^29");
    if (c.max_cells == 0) c.max_cells = 1;
    if (c.max_surface == 0) c.max_surface = 1;
    if (c.max_buckets == 0) c.max_buckets = 1;
    if (c.max_driven == 0) c.max_driven = 1;
    if (c.max_splats == 0) c.max_splats = 1;
    if (c.tile_nodes) w->tile_nodes = c.tile_nodes;
    if (c.tile_cells) w->tile_cells = c.tile_cells;
    if (c.block_cells) w->block_cells = c.block_cells;
    if (w->block_cells != 64 && w->block_cells != 128 && w->block_cells != 256 && w->block_cells != 512) return VS_E_INVALID;
    hipError_t e = hipSetDevice(c.device);
    if (e != hipSuccess) return VS_E_DEVICE;
    vs_world *W = w;
#define ALLOC(ptr, bytes) HIPW(W, hipMalloc(reinterpret_cast<void **>(&(ptr)), (bytes)))
    ALLOC(W->d_nodes, sizeof(float) * NODE_PLANES * c.max_nodes);
    ALLOC(W->d_cells, sizeof(uint32_t) * c.max_cells);
    ALLOC(W->d_bodies, sizeof(vs_body) * c.max_bodies);
    ALLOC(W->d_scenes, sizeof(vs_scene_params) * c.max_scenes);
    ALLOC(W->d_surf, sizeof(uint32_t) * SURF_PLANES * c.max_surface);
    ALLOC(W->d_surf_pay, 2 * sizeof(float4) * c.max_surface);
    ALLOC(W->d_grid, sizeof(uint32_t) * (1 + BUCKET_SLOTS) * c.max_buckets);
    ALLOC(W->d_ovf, sizeof(uint2) * c.max_surface);
    ALLOC(W->d_ovf_count, 8 * sizeof(uint32_t));
    ALLOC(W->d_ovb, sizeof(uint2) * ovb_cap(W));
    ALLOC(W->d_ext, sizeof(uint32_t) * EXT_SLOTS * c.max_buckets);
    ALLOC(W->d_summary, sizeof(uint32_t) * c.max_buckets);
    ALLOC(W->d_parity1, sizeof(uint32_t) * 2 * c.max_buckets);
    ALLOC(W->d_drive, sizeof(uint32_t) * DRIVE_PLANES * c.max_driven);
    ALLOC(W->d_splat_in, sizeof(float) * 10 * c.max_splats);
    ALLOC(W->d_splat_out, sizeof(float) * SPLAT_OUT_PLANES * c.max_splats);
    ALLOC(W->d_scene_counts, sizeof(uint32_t) * NPASS * c.max_scenes);
    ALLOC(W->d_branch_counts, sizeof(uint32_t) * NPASS * 2);
    HIPW(W, hipMemset(W->d_nodes, 0, sizeof(float) * NODE_PLANES * c.max_nodes));
    HIPW(W, hipMemset(W->d_grid, 0, sizeof(uint32_t) * c.max_buckets));
    HIPW(W, hipMemset(W->d_summary, 0, sizeof(uint32_t) * c.max_buckets));
    HIPW(W, hipMemset(W->d_parity1, 0, sizeof(uint32_t) * 2 * c.max_buckets));
    HIPW(W, hipMemset(W->d_ovf_count, 0, 8 * sizeof(uint32_t)));
    HIPW(W, hipMemset(W->d_scenes, 0, sizeof(vs_scene_params) * c.max_scenes));
    W->nodes.init(c.max_nodes);
    W->cells.init(c.max_cells);
    W->bodies.init(c.max_bodies);
    W->scenes.init(c.max_scenes);
    W->surface.init(c.max_surface);
    W->buckets.init(c.max_buckets);
    W->driven.init(c.max_driven);
    W->splats.init(c.max_splats);
    // A table needs a tile per `tile` items plus one partial tile per body.
    W->tile_cap[P_NODE] = c.max_nodes / W->tile_nodes + c.max_bodies + 1;
    // P_REST and P_REST0 list whole bodies (those whose colour pairs do not
    // take their node passes), like P_NODE.
    W->tile_cap[P_REST] = c.max_nodes / W->tile_nodes + c.max_bodies + 1;
    W->tile_cap[P_REST0] = W->tile_cap[P_REST];
    W->tile_cap[P_LAST] = W->tile_cap[P_REST];
    W->tile_cap[P_BODY] = c.max_bodies + 1;
    W->tile_cap[P_SURF] = c.max_surface / W->tile_surf + c.max_scenes + 1;
    W->tile_cap[P_DRIVE] = c.max_driven / W->tile_drive + c.max_scenes + 1;
    W->tile_cap[P_SPLAT] = c.max_splats / W->tile_splat + c.max_bodies + 1;
    // Colour populations are uneven for small or odd lattices: size every
    // colour table for all cells (tiles are 16 bytes, cheap to over-provision).
    for (int k = 0; k < 8; k++) W->tile_cap[P_CELL + k] = c.max_cells / 64 + c.max_bodies + 64;
    // A full row tile holds at least block/2 >= 32 cells of its class.
    // Plus, for classes 0 and 3, node tiles of the rows' leftover nodes (a
    // body's leftover list in pieces of block_cells >= 64 entries).
    for (int k = 0; k < 4; k++)
        W->tile_cap[P_ROW + k] = c.max_cells / 32 + c.max_nodes / 64 + 2 * c.max_bodies + 64;
    for (int p = 0; p < NPASS; p++) {
        const size_t bytes = 16 + W->tile_cap[p] * sizeof(vs_tile);
        ALLOC(W->d_tiles[p], bytes);
        HIPW(W, hipHostMalloc(reinterpret_cast<void **>(&W->h_tiles[p].p), bytes, hipHostMallocDefault));
        std::memset(W->h_tiles[p].p, 0, 16);
        HIPW(W, hipMemset(W->d_tiles[p], 0, 16));
        if (is_row_pass(p)) {
            ALLOC(W->d_rows_by_work[p - P_ROW], bytes);
            HIPW(W, hipHostMalloc(reinterpret_cast<void **>(&W->h_rows_by_work[p - P_ROW].p), bytes,
                                  hipHostMallocDefault));
            std::memset(W->h_rows_by_work[p - P_ROW].p, 0, 16);
            HIPW(W, hipMemset(W->d_rows_by_work[p - P_ROW], 0, 16));
        }
    }
    HIPW(W, hipHostMalloc(reinterpret_cast<void **>(&W->h_scene_counts.p), sizeof(uint32_t) * NPASS * c.max_scenes,
                          hipHostMallocDefault));
    HIPW(W, hipHostMalloc(reinterpret_cast<void **>(&W->h_branch_counts.p), sizeof(uint32_t) * NPASS * 2,
                          hipHostMallocDefault));
    HIPW(W, hipHostMalloc(reinterpret_cast<void **>(&W->h_scenes.p), sizeof(vs_scene_params) * c.max_scenes,
                          hipHostMallocDefault));
    std::memset(W->h_scenes.p, 0, sizeof(vs_scene_params) * c.max_scenes);
#undef ALLOC
    W->tile_count.assign(NPASS, 0);
    W->scene_first.assign(NPASS, std::vector<uint32_t>(c.max_scenes, 0));
    W->scene_ntiles.assign(NPASS, std::vector<uint32_t>(c.max_scenes, 0));
    W->recs.resize(c.max_scenes);
    W->h_bodies.resize(c.max_bodies);
    HIPW(W, hipStreamCreateWithFlags(&W->stream, hipStreamNonBlocking));
    HIPW(W, hipStreamCreateWithFlags(&W->upload, hipStreamNonBlocking));
    HIPW(W, hipStreamCreateWithFlags(&W->stream2, hipStreamNonBlocking));
    HIPW(W, hipEventCreateWithFlags(&W->fork, hipEventDisableTiming));
    HIPW(W, hipEventCreateWithFlags(&W->join, hipEventDisableTiming));
    HIPW(W, hipEventCreateWithFlags(&W->upload_done, hipEventDisableTiming));
    HIPW(W, hipEventCreateWithFlags(&W->tables_done, hipEventDisableTiming));
    HIPW(W, hipEventCreateWithFlags(&W->params_done, hipEventDisableTiming));
    HIPW(W, hipEventCreateWithFlags(&W->released, hipEventDisableTiming));
    HIPW(W, hipEventRecord(W->released, W->stream));
    HIPW(W, hipEventRecord(W->upload_done, W->upload));
    HIPW(W, hipEventRecord(W->tables_done, W->stream));
    HIPW(W, hipEventRecord(W->params_done, W->stream));
    W->events.resize(EVENT_POOL);
    for (auto &ev : W->events) HIPW(W, hipEventCreate(&ev));
    for (auto &ev : W->tune) HIPW(W, hipEventCreate(&ev));
    HIPW(W, hipDeviceSynchronize());
    return VS_OK;
}

void vs_world_destroy(vs_world *w)
{
    if (!w) return;
    (void)hipSetDevice(w->cfg.device);
    if (w->stream) (void)hipStreamSynchronize(w->stream);
    if (w->stream2) (void)hipStreamSynchronize(w->stream2);
    if (w->upload) (void)hipStreamSynchronize(w->upload);
    if (w->exec) (void)hipGraphExecDestroy(w->exec);
    if (w->graph) (void)hipGraphDestroy(w->graph);
    for (auto &ev : w->events)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : { w->upload_done, w->tables_done, w->params_done, w->released, w->tune[0], w->tune[1], w->fork,
                           w->join })
        if (ev) (void)hipEventDestroy(ev);
    void *ptrs[] = { w->d_nodes, w->d_cells, w->d_bodies, w->d_scenes, w->d_surf, w->d_surf_pay,
                     w->d_grid,  w->d_ovf, w->d_ovf_count, w->d_ovb, w->d_ext, w->d_summary, w->d_parity1, w->d_drive, w->d_splat_in, w->d_splat_out,
                     w->d_pairlog, w->d_scene_counts, w->d_branch_counts };
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (int p = 0; p < NPASS; p++)
        if (w->d_tiles[p]) (void)hipFree(w->d_tiles[p]);
    for (uint32_t *d : w->d_rows_by_work)
        if (d) (void)hipFree(d);
    if (w->stream) (void)hipStreamDestroy(w->stream);
    if (w->stream2) (void)hipStreamDestroy(w->stream2);
    if (w->upload) (void)hipStreamDestroy(w->upload);
    delete w;
}

const char *vs_world_error(const vs_world *w) { return w ? w->error.c_str() : "no world"; }

int vs_world_add_scene(vs_world *w, const vs_scene_desc *d, uint32_t *scene_id)
{
    if (!w || !d || !scene_id || (d->n_bodies && !d->bodies)) return VS_E_INVALID;
    if (d->n_bodies == 0) return fail(w, VS_E_INVALID, "a scene needs at least one body");
    if (!(d->dt > 0.0f)) return fail(w, VS_E_INVALID, "stepDuration must be a finite positive number");
    HIPW(w, hipSetDevice(w->cfg.device));
    hipStream_t up = w->upload;
    // Ranges of removed scenes are written only after the frames enqueued
    // before their removal.
    HIPW(w, hipStreamWaitEvent(up, w->released, 0));
    // Retire host copies whose uploads have landed.
    if (hipEventQuery(w->upload_done) == hipSuccess) w->pending.clear();
    // Device temporaries of lattice builds, released in stream order on the
    // upload stream whichever way this returns.
    struct Temps {
        hipStream_t s;
        std::vector<void *> p;
        ~Temps()
        {
            for (void *q : p) (void)hipFreeAsync(q, s);
        }
    } temps{ up, {} };
    auto temp = [&](void **out, size_t bytes) -> hipError_t {
        const hipError_t e = hipMallocAsync(out, std::max<size_t>(bytes, 4), up);
        if (e == hipSuccess) temps.p.push_back(*out);
        return e;
    };
    // Sizes. A body either brings its lattice (buildLattice on the host) or
    // an occupancy grid the device turns into one; for those, a count pass
    // runs first so that the pools can be sized.
    std::vector<BodyPlan> plan(d->n_bodies);
    SceneRec r;
    uint64_t nodes = 0, ncells = 0, nsurf = 0, nsplats = 0;
    bool counting = false;
    for (uint32_t b = 0; b < d->n_bodies; b++) {
        const vs_body_desc &bd = d->bodies[b];
        BodyPlan &pl = plan[b];
        const uint64_t cells_total = (uint64_t)bd.cell_dims[0] * bd.cell_dims[1] * bd.cell_dims[2];
        const uint64_t nn = (uint64_t)(bd.cell_dims[0] + 1) * (bd.cell_dims[1] + 1) * (bd.cell_dims[2] + 1);
        if (bd.occupancy) {
            if (!(bd.h64 > 0.0) || !(bd.density > 0.0)) return fail(w, VS_E_INVALID, "cell size and density must be positive");
            if (cells_total == 0 || nn >= (1ull << 31)) return fail(w, VS_E_INVALID, "lattice dimensions out of range");
            for (uint32_t i = 0; i < bd.n_static; i++) {
                if (!bd.static_nodes || bd.static_nodes[i] >= nn) return fail(w, VS_E_INVALID, "static node outside the lattice");
            }
            pl.device = true;
            for (int a = 0; a < 3; a++) pl.origin[a] = (float)bd.origin64[a];
            pl.h = (float)bd.h64;
            pl.tiles = vsim::lattice_tiles(bd.cell_dims, pl.tile_base, &pl.node_tiles);
            HIPW(w, temp(reinterpret_cast<void **>(&pl.d_occ), cells_total));
            HIPW(w, temp(reinterpret_cast<void **>(&pl.d_counts), 4ull * pl.tiles));
            HIPW(w, temp(reinterpret_cast<void **>(&pl.d_offsets), 4ull * pl.tiles));
            const uint8_t *ho = keep(w, std::vector<uint8_t>(bd.occupancy, bd.occupancy + cells_total));
            HIPW(w, hipMemcpyAsync(pl.d_occ, ho, cells_total, hipMemcpyHostToDevice, up));
            vsim::LatticeView L = lattice_view(bd, pl, 0, 0);
            HIPW(w, vsim::launch_lattice_lists(up, L, pl.tiles, 0));
            pl.counts.resize(pl.tiles);
            HIPW(w, hipMemcpyAsync(pl.counts.data(), pl.d_counts, 4ull * pl.tiles, hipMemcpyDeviceToHost, up));
            counting = true;
        } else {
            if (!bd.rest || !bd.inv_mass || (bd.n_cells && !bd.cells)) return fail(w, VS_E_INVALID, "body arrays missing");
            uint64_t colored = 0;
            for (int c = 0; c < 8; c++) colored += bd.color_count[c];
            if (colored != bd.n_cells) return fail(w, VS_E_INVALID, "colour counts do not sum to the cell count");
            if (!(bd.h > 0.0f)) return fail(w, VS_E_INVALID, "cell size must be positive");
            for (uint32_t i = 0; i < bd.n_cells; i++) {
                if (bd.cells[i] >= cells_total) return fail(w, VS_E_INVALID, "cell index outside the lattice");
            }
            for (uint32_t i = 0; i < bd.n_surface; i++) {
                if (bd.surface[i] >= nn) return fail(w, VS_E_INVALID, "surface node outside the lattice");
            }
            pl.n_cells = bd.n_cells;
            pl.n_surface = bd.n_surface;
            for (int c = 0; c < 8; c++) pl.color_count[c] = bd.color_count[c];
            for (int a = 0; a < 3; a++) pl.origin[a] = bd.origin[a];
            pl.h = bd.h;
        }
        nodes += nn;
        nsplats += bd.n_splats;
        r.stretch |= bd.max_stretch > 1.0f;
    }
    if (counting) {
        HIPW(w, hipStreamSynchronize(up));
        for (BodyPlan &pl : plan) {
            if (!pl.device) continue;
            // Exclusive offsets: one running sum over the colour tiles (so it
            // carries each colour's base), another over the node tiles.
            pl.offsets.resize(pl.tiles);
            uint32_t at = 0;
            for (uint32_t t = 0; t < pl.tile_base[8]; t++) {
                pl.offsets[t] = at;
                at += pl.counts[t];
            }
            pl.n_cells = at;
            for (int c = 0; c < 8; c++) {
                uint32_t n = 0;
                for (uint32_t t = pl.tile_base[c]; t < pl.tile_base[c + 1]; t++) n += pl.counts[t];
                pl.color_count[c] = n;
            }
            at = 0;
            for (uint32_t t = pl.tile_base[8]; t < pl.tiles; t++) {
                pl.offsets[t] = at;
                at += pl.counts[t];
            }
            pl.n_surface = at;
        }
    }
    for (const BodyPlan &pl : plan) {
        ncells += pl.n_cells;
        if (d->contact) nsurf += pl.n_surface;
    }
    for (uint32_t i = 0; i < d->n_driven; i++) {
        if (d->driven[i] >= nodes) return fail(w, VS_E_INVALID, "driven node outside the scene");
    }
    r.contact = d->contact == 2 || (d->contact && d->n_bodies >= 2);
    // Contact grid (broadphase.js:46-58): one bucket per contact diameter
    // over the tank, a ring of margin buckets, origin one bucket below min.
    double edge = d->bucket_size;
    uint32_t dims[3] = { 1, 1, 1 };
    double origin[3] = { 0, 0, 0 };
    uint64_t nbuckets = 0;
    if (r.contact && d->grid_dims[0] && d->grid_dims[1] && d->grid_dims[2]) {
        for (int a = 0; a < 3; a++) {
            dims[a] = d->grid_dims[a];
            origin[a] = d->grid_origin[a];
        }
        nbuckets = (uint64_t)dims[0] * dims[1] * dims[2];
    } else if (r.contact) {
        if (!(edge > 0)) {
            edge = 0;
            for (int a = 0; a < 3; a++) edge = std::max(edge, (double)d->tank_max[a] - (double)d->tank_min[a]);
        }
        for (int a = 0; a < 3; a++) {
            const double extent = (double)d->tank_max[a] - (double)d->tank_min[a];
            dims[a] = (uint32_t)std::max(1.0, std::ceil(extent / edge)) + 2u;
            origin[a] = (double)d->tank_min[a] - edge;
        }
        nbuckets = (uint64_t)dims[0] * dims[1] * dims[2];
    }
    // Ranges.
    uint32_t id = 0;
    for (; id < w->cfg.max_scenes && w->recs[id].live; id++) {
    }
    if (id == w->cfg.max_scenes) return fail(w, VS_E_NOMEM, "no free scene slot");
    const bool ok = w->bodies.alloc(d->n_bodies, r.body_begin) && (r.body_count = d->n_bodies, true) &&
                    w->nodes.alloc(nodes, r.node_begin) && (r.node_count = nodes, true) &&
                    w->cells.alloc(ncells, r.cell_begin) && (r.cell_count = ncells, true) &&
                    w->surface.alloc(nsurf, r.surf_begin) && (r.surf_count = nsurf, true) &&
                    w->buckets.alloc(nbuckets, r.bucket_begin) && (r.bucket_count = nbuckets, true) &&
                    w->driven.alloc(d->n_driven, r.drive_begin) && (r.drive_count = d->n_driven, true) &&
                    w->splats.alloc(nsplats, r.splat_begin) && (r.splat_count = nsplats, true);
    if (!ok) {
        release_ranges(w, r);
        return fail(w, VS_E_NOMEM, "scene does not fit the HBM budget");
    }

    vs_scene_params &P = r.params;
    P = vs_scene_params{};
    P.dt = d->dt;
    for (int a = 0; a < 3; a++) {
        P.gravity[a] = d->gravity[a];
        P.tank_min[a] = d->tank_min[a];
        P.tank_max[a] = d->tank_max[a];
        P.grid_origin[a] = (float)origin[a];
        P.grid_dims[a] = dims[a];
    }
    for (int a = 0; a < 4; a++) P.sphere[a] = d->sphere[a];
    // Contact's bucket bound: no pair reaches farther than twice the widest
    // record radius; the margin covers the f32 rounding of bucket coordinates
    // and of the reach test over a grid of this extent (a few hundred ulps).
    {
        double widest = 0, extent = 0;
        for (uint32_t b = 0; b < d->n_bodies; b++) widest = std::max(widest, (double)d->bodies[b].contact_radius);
        for (int a = 0; a < 3; a++) extent = std::max(extent, std::fabs(origin[a]) + edge * dims[a]);
        P.reach_max = (float)(2.0 * widest);
        P.skip_margin = (float)(1e-5 * (extent + 2.0 * widest));
    }
    P.gain = 1.0f;  // GAIN, bodyContact.js:48
    P.bucket_size = (float)edge;
    P.bucket_count = (uint32_t)nbuckets;
    P.grid_base = (uint32_t)r.bucket_begin;
    P.response_time = d->response_time > 0 ? d->response_time : 0.05f;  // RESPONSE_TIME, kinematicDrive.js:65
    P.drive_law = d->drive_law;
    P.surf_base = (uint32_t)r.surf_begin;
    P.surf_count = (uint32_t)nsurf;
    P.bucket_slots = BUCKET_SLOTS;
    P.flags = (!r.contact && !r.stretch ? VS_SCENE_FOLD : 0u) | (d->n_driven == 0 ? VS_SCENE_FOLD_PREDICT : 0u);

    const size_t N = w->cfg.max_nodes;
    uint64_t node_at = r.node_begin, cell_at = r.cell_begin, splat_at = r.splat_begin, surf_at = r.surf_begin;
    const size_t S = w->cfg.max_surface;
    // Hard-held nodes carry no mass (kinematicDrive.js:358-364): a host-built
    // body's inv_mass arrives with them zeroed, a device-built one gets them
    // zeroed here, with its static nodes.
    const bool hard = d->n_driven && d->drive_law == 0;
    // Host-built bodies that follow each other are staged together and go up
    // as one copy per plane: a scene of many small bodies is a handful of
    // copies, not a dozen per body.
    struct Staging {
        uint64_t node0 = 0, cell0 = 0, surf0 = 0;
        std::vector<float> pos[3], vel[3], inv, radius;
        std::vector<uint32_t> cells, snode, sbody;
    } st;
    auto flush = [&]() -> int {
        const uint64_t nn = st.inv.size();
        if (nn) {
            for (int a = 0; a < 3; a++) {
                const float *hx = keep(w, std::move(st.pos[a]));
                for (int k : { 0, 3, 10 })
                    HIPW(w, hipMemcpyAsync(w->d_nodes + (k + a) * N + st.node0, hx, nn * 4, hipMemcpyHostToDevice, up));
                const float *hv = keep(w, std::move(st.vel[a]));
                HIPW(w, hipMemcpyAsync(w->d_nodes + (6 + a) * N + st.node0, hv, nn * 4, hipMemcpyHostToDevice, up));
            }
            const float *hw = keep(w, std::move(st.inv));
            HIPW(w, hipMemcpyAsync(w->d_nodes + 9 * N + st.node0, hw, nn * 4, hipMemcpyHostToDevice, up));
        }
        if (!st.cells.empty()) {
            const size_t nc = st.cells.size();
            const uint32_t *hc = keep(w, std::move(st.cells));
            HIPW(w, hipMemcpyAsync(w->d_cells + st.cell0, hc, nc * 4, hipMemcpyHostToDevice, up));
        }
        if (!st.snode.empty()) {
            const size_t ns = st.snode.size();
            const uint32_t *hn = keep(w, std::move(st.snode));
            const uint32_t *hb = keep(w, std::move(st.sbody));
            const float *hr = keep(w, std::move(st.radius));
            HIPW(w, hipMemcpyAsync(w->d_surf + st.surf0, hn, ns * 4, hipMemcpyHostToDevice, up));
            HIPW(w, hipMemcpyAsync(w->d_surf + S + st.surf0, hb, ns * 4, hipMemcpyHostToDevice, up));
            HIPW(w, hipMemcpyAsync(w->d_surf + 2 * S + st.surf0, hr, ns * 4, hipMemcpyHostToDevice, up));
        }
        st = Staging{};
        return VS_OK;
    };
    r.bodies.resize(d->n_bodies);
    for (uint32_t b = 0; b < d->n_bodies; b++) {
        const vs_body_desc &bd = d->bodies[b];
        const BodyPlan &pl = plan[b];
        BodyRec &br = r.bodies[b];
        const uint32_t nd[3] = { bd.cell_dims[0] + 1, bd.cell_dims[1] + 1, bd.cell_dims[2] + 1 };
        const uint64_t nn = (uint64_t)nd[0] * nd[1] * nd[2];
        br.node_begin = node_at;
        br.node_count = nn;
        br.damping = bd.damping;
        for (int a = 0; a < 3; a++) br.cell_dims[a] = bd.cell_dims[a];
        br.dense = (uint64_t)pl.n_cells == (uint64_t)bd.cell_dims[0] * bd.cell_dims[1] * bd.cell_dims[2];
        vs_body &vb = w->h_bodies[r.body_begin + b];
        vb = vs_body{};
        for (int a = 0; a < 3; a++) {
            vb.cell_dims[a] = bd.cell_dims[a];
            vb.node_dims[a] = nd[a];
            vb.origin[a] = pl.origin[a];
        }
        vb.node_offset = (uint32_t)node_at;
        vb.node_count = (uint32_t)nn;
        vb.scene = id;
        vb.h = pl.h;
        vb.alpha_dev = bd.alpha_dev;
        vb.alpha_hyd = bd.alpha_hyd;
        vb.gamma = bd.gamma;
        vb.relaxation = bd.relaxation;
        vb.max_stretch = bd.max_stretch;
        vb.friction = bd.friction;
        vb.damp = damp_factor(bd.damping, d->dt);
        const bool surface = d->contact && pl.n_surface;
        if (pl.device) {
            if (int rc = flush()) return rc;
            vsim::LatticeView L = lattice_view(bd, pl, node_at, r.body_begin + b);
            L.cells = w->d_cells + cell_at;
            if (surface) {
                L.s_node = w->d_surf + surf_at;
                L.s_body = w->d_surf + S + surf_at;
                L.s_radius = reinterpret_cast<float *>(w->d_surf + 2 * S + surf_at);
            }
            const uint32_t *hoff = keep(w, std::vector<uint32_t>(pl.offsets));
            HIPW(w, hipMemcpyAsync(pl.d_offsets, hoff, 4ull * pl.tiles, hipMemcpyHostToDevice, up));
            HIPW(w, vsim::launch_lattice_lists(up, L, surface ? pl.tiles : pl.tile_base[8], 1));
            HIPW(w, vsim::launch_lattice_nodes(up, L, node_view(w), plane(w, 10), plane(w, 11), plane(w, 12)));
            std::vector<uint32_t> zero;
            for (uint32_t i = 0; i < bd.n_static; i++) zero.push_back((uint32_t)(node_at + bd.static_nodes[i]));
            const uint64_t first = node_at - r.node_begin;
            for (uint32_t i = 0; hard && i < d->n_driven; i++) {
                if (d->driven[i] >= first && d->driven[i] < first + nn) zero.push_back((uint32_t)(r.node_begin + d->driven[i]));
            }
            if (!zero.empty()) {
                uint32_t *dz = nullptr;
                HIPW(w, temp(reinterpret_cast<void **>(&dz), 4 * zero.size()));
                const size_t nz = zero.size();
                const uint32_t *hz = keep(w, std::move(zero));
                HIPW(w, hipMemcpyAsync(dz, hz, 4 * nz, hipMemcpyHostToDevice, up));
                HIPW(w, vsim::launch_zero_mass(up, plane(w, 9), dz, (uint32_t)nz));
            }
        } else {
            // Node planes: p = prev = rest, vel, w, rest; the cell list; the
            // surface records, body by body (bodyContact.js:72-85).
            if (st.inv.empty()) {
                st.node0 = node_at;
                st.cell0 = cell_at;
                st.surf0 = surf_at;
            }
            for (int a = 0; a < 3; a++) {
                std::vector<float> &x = st.pos[a];
                for (uint64_t i = 0; i < nn; i++) x.push_back(bd.rest[3 * i + a]);
                st.vel[a].insert(st.vel[a].end(), nn, bd.velocity[a]);
            }
            st.inv.insert(st.inv.end(), bd.inv_mass, bd.inv_mass + nn);
            st.cells.insert(st.cells.end(), bd.cells, bd.cells + bd.n_cells);
            if (surface) {
                for (uint32_t i = 0; i < pl.n_surface; i++) st.snode.push_back((uint32_t)(node_at + bd.surface[i]));
                st.sbody.insert(st.sbody.end(), pl.n_surface, (uint32_t)(r.body_begin + b));
                st.radius.insert(st.radius.end(), pl.n_surface, bd.contact_radius);
            }
        }
        // Cells, colour by colour.
        uint64_t src = 0;
        for (int c = 0; c < 8; c++) {
            br.color_begin[c] = cell_at + src;
            br.color_count[c] = pl.color_count[c];
            src += pl.color_count[c];
        }
        cell_at += pl.n_cells;
        br.surf_begin = surf_at;
        br.surf_count = d->contact ? pl.n_surface : 0;
        if (d->contact) surf_at += pl.n_surface;
        // Splats.
        br.splat_begin = splat_at;
        br.splat_count = bd.n_splats;
        if (bd.n_splats) {
            // Stored in cell order (x fastest) so that the lanes of a wave
            // gather the same corner nodes; `splat_perm` maps the stored
            // order back to the caller's for readback.
            const size_t G = w->cfg.max_splats;
            const uint32_t ns = bd.n_splats;
            std::vector<uint64_t> key(ns);
            const double inv = 1.0 / (double)pl.h;
            for (uint32_t i = 0; i < ns; i++) {
                uint64_t c[3];
                for (int a = 0; a < 3; a++) {
                    double g = std::floor(((double)bd.splat_mean[3 * i + a] - (double)pl.origin[a]) * inv);
                    g = std::min(std::max(g, 0.0), (double)bd.cell_dims[a] - 1.0);
                    c[a] = (uint64_t)g;
                }
                key[i] = c[0] + (uint64_t)bd.cell_dims[0] * (c[1] + (uint64_t)bd.cell_dims[1] * c[2]);
            }
            std::vector<uint32_t> order(ns);
            for (uint32_t i = 0; i < ns; i++) order[i] = i;
            std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return key[x] < key[y]; });
            std::vector<float> planes(6 * (size_t)ns), quat(4 * (size_t)ns);
            for (uint32_t j = 0; j < ns; j++) {
                const uint32_t i = order[j];
                for (int a = 0; a < 3; a++) {
                    planes[a * (size_t)ns + j] = bd.splat_mean[3 * (size_t)i + a];
                    planes[(3 + a) * (size_t)ns + j] = bd.splat_scale[3 * (size_t)i + a];
                }
                for (int a = 0; a < 4; a++) quat[4 * (size_t)j + a] = bd.splat_quat[4 * (size_t)i + a];
                r.splat_perm.push_back((uint32_t)(splat_at - r.splat_begin) + i);
            }
            const float *hp = keep(w, std::move(planes));
            const float *hq = keep(w, std::move(quat));
            HIPW(w, hipMemcpyAsync(w->d_splat_in + 4 * splat_at, hq, (size_t)ns * 16, hipMemcpyHostToDevice, up));
            for (int k = 0; k < 6; k++)
                HIPW(w, hipMemcpyAsync(w->d_splat_in + (4 + k) * G + splat_at, hp + (size_t)k * ns, (size_t)ns * 4,
                                       hipMemcpyHostToDevice, up));
        }
        splat_at += bd.n_splats;
        node_at += nn;
    }
    if (int rc = flush()) return rc;
    const vs_body *hb_rec = keep(w, std::vector<vs_body>(w->h_bodies.begin() + r.body_begin,
                                                         w->h_bodies.begin() + r.body_begin + d->n_bodies));
    HIPW(w, hipMemcpyAsync(w->d_bodies + r.body_begin, hb_rec, sizeof(vs_body) * d->n_bodies, hipMemcpyHostToDevice, up));
    if (nbuckets) HIPW(w, hipMemsetAsync(w->d_grid + r.bucket_begin, 0, nbuckets * 4, up));
    if (nbuckets) HIPW(w, hipMemsetAsync(w->d_summary + r.bucket_begin, 0, nbuckets * 4, up));
    if (nbuckets) HIPW(w, hipMemsetAsync(w->d_parity1 + r.bucket_begin, 0, nbuckets * 4, up));
    if (nbuckets) HIPW(w, hipMemsetAsync(w->d_parity1 + w->cfg.max_buckets + r.bucket_begin, 0, nbuckets * 4, up));
    if (d->n_driven) {
        const size_t D = w->cfg.max_driven;
        std::vector<uint32_t> dn(d->n_driven);
        std::vector<float> off(3 * (size_t)d->n_driven);
        for (uint32_t i = 0; i < d->n_driven; i++) {
            dn[i] = (uint32_t)(r.node_begin + d->driven[i]);
            for (int a = 0; a < 3; a++) off[a * d->n_driven + i] = d->driven_offset[3 * i + a];
        }
        const uint32_t *hd = keep(w, std::move(dn));
        const float *ho = keep(w, std::move(off));
        HIPW(w, hipMemcpyAsync(w->d_drive + r.drive_begin, hd, (size_t)d->n_driven * 4, hipMemcpyHostToDevice, up));
        for (int a = 0; a < 3; a++)
            HIPW(w, hipMemcpyAsync(w->d_drive + (1 + a) * D + r.drive_begin, ho + (size_t)a * d->n_driven,
                                   (size_t)d->n_driven * 4, hipMemcpyHostToDevice, up));
    }
    HIPW(w, hipEventRecord(w->upload_done, up));
    r.live = true;
    w->recs[id] = std::move(r);
    w->tables_dirty = true;
    w->params_dirty = true;
    *scene_id = id;
    return VS_OK;
}

int vs_world_remove_scene(vs_world *w, uint32_t id)
{
    if (!w || id >= w->cfg.max_scenes || !w->recs[id].live) return VS_E_INVALID;
    // Frames already enqueued may still read the scene's ranges; the next
    // admission may hand them out again, and its uploads wait for this event.
    // The host does not wait: removal keeps the device busy.
    HIPW(w, hipSetDevice(w->cfg.device));
    HIPW(w, hipEventRecord(w->released, w->stream));
    release_ranges(w, w->recs[id]);
    w->recs[id] = SceneRec{};
    w->tables_dirty = true;
    return VS_OK;
}

int vs_world_reset_scene(vs_world *w, uint32_t id)
{
    if (!w || id >= w->cfg.max_scenes || !w->recs[id].live) return VS_E_INVALID;
    const SceneRec &r = w->recs[id];
    const size_t N = w->cfg.max_nodes;
    HIPW(w, hipStreamWaitEvent(w->stream, w->upload_done, 0));
    for (int a = 0; a < 3; a++) {
        HIPW(w, hipMemcpyAsync(w->d_nodes + a * N + r.node_begin, w->d_nodes + (10 + a) * N + r.node_begin,
                               r.node_count * 4, hipMemcpyDeviceToDevice, w->stream));
        HIPW(w, hipMemcpyAsync(w->d_nodes + (3 + a) * N + r.node_begin, w->d_nodes + (10 + a) * N + r.node_begin,
                               r.node_count * 4, hipMemcpyDeviceToDevice, w->stream));
        HIPW(w, hipMemsetAsync(w->d_nodes + (6 + a) * N + r.node_begin, 0, r.node_count * 4, w->stream));
    }
    return VS_OK;
}

int vs_world_set_target(vs_world *w, uint32_t id, const float target[3], const float velocity[3])
{
    if (!w || id >= w->cfg.max_scenes || !w->recs[id].live || !target || !velocity) return VS_E_INVALID;
    for (int a = 0; a < 3; a++) {
        w->recs[id].params.target[a] = target[a];
        w->recs[id].params.target_vel[a] = velocity[a];
    }
    w->params_dirty = true;
    return VS_OK;
}

// Enqueues one batched frame: on the world's stream, or, with two chains,
// each chain's scenes on a stream of its own between a fork and a join (the
// scenes of one chain share nothing with the other's: their launches may run
// side by side and fill each other's tails).
static int enqueue_batched_frame(vs_world *w, uint32_t substeps, uint32_t mode)
{
    bool two = w->cfg.branches >= 2;
    if (two) {
        bool used[2] = { false, false };
        for (const SceneRec &r : w->recs)
            if (r.live) used[r.branch] = true;
        two = used[0] && used[1];
    }
    if (!two) return enqueue_frame(w, substeps, mode, -1);
    HIPW(w, hipEventRecord(w->fork, w->stream));
    HIPW(w, hipStreamWaitEvent(w->stream2, w->fork, 0));
    int rc = VS_OK;
    for (int b = 0; b < 2 && !rc; b++) {
        w->cur_branch = b;
        w->cur_stream = b ? w->stream2 : w->stream;
        rc = enqueue_frame(w, substeps, mode, -1);
    }
    w->cur_branch = -1;
    w->cur_stream = nullptr;
    // Joined even after a failed enqueue, so no work is left outside the
    // world's stream.
    HIPW(w, hipEventRecord(w->join, w->stream2));
    HIPW(w, hipStreamWaitEvent(w->stream, w->join, 0));
    return rc;
}

int vs_world_frame(vs_world *w, uint32_t substeps, int32_t splats, int32_t graph)
{
    if (!w) return VS_E_INVALID;
    HIPW(w, hipSetDevice(w->cfg.device));
    if (int rc = prepare(w)) return rc;
    const uint32_t mode = splat_mode(splats);
    if (!graph) return enqueue_batched_frame(w, substeps, mode);
    std::vector<uint32_t> key = { substeps, mode, (uint32_t)w->any_stretch, (uint32_t)w->any_contact,
                                  w->cfg.separate_node_passes, w->pair_cap,
                                  w->block_cells,
                                  w->cfg.xcd_chunked, w->cfg.separate_colours, w->block_rest, w->block_rest0,
                                  w->cfg.colour_launches, w->block_body, w->block_last, w->cfg.branches,
                                  (uint32_t)w->settle_fold };
    for (int p = 0; p < NPASS; p++) {
        key.push_back(grid_for(w, p, -1));
        for (int b = 0; b < 2; b++) key.push_back(w->branch_ntiles[p][b] != 0);
    }
    if (!w->exec || key != w->graph_key) {
        if (w->exec) HIPW(w, hipGraphExecDestroy(w->exec));
        if (w->graph) HIPW(w, hipGraphDestroy(w->graph));
        w->exec = nullptr;
        w->graph = nullptr;
        HIPW(w, hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_batched_frame(w, substeps, mode);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(w->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(w, VS_E_DEVICE, std::string("frame capture: ") + hipGetErrorString(e));
        w->graph = g;
        HIPW(w, hipGraphInstantiate(&w->exec, w->graph, nullptr, nullptr, 0));
        w->graph_key = key;
        w->captures++;
    }
    HIPW(w, hipGraphLaunch(w->exec, w->stream));
    return VS_OK;
}

int vs_world_frame_scene(vs_world *w, uint32_t id, uint32_t substeps, int32_t splats)
{
    if (!w || id >= w->cfg.max_scenes || !w->recs[id].live) return VS_E_INVALID;
    HIPW(w, hipSetDevice(w->cfg.device));
    if (int rc = prepare(w)) return rc;
    return enqueue_frame(w, substeps, splat_mode(splats), (int)id);
}

int vs_world_synchronize(vs_world *w)
{
    if (!w) return VS_E_INVALID;
    HIPW(w, hipStreamSynchronize(w->upload));
    HIPW(w, hipStreamSynchronize(w->stream));
    w->pending.clear();
    return VS_OK;
}

int vs_world_record(vs_world *w, uint32_t *slot)
{
    if (!w || !slot) return VS_E_INVALID;
    const uint32_t k = w->next_event++ % EVENT_POOL;
    HIPW(w, hipEventRecord(w->events[k], w->stream));
    *slot = k;
    return VS_OK;
}

int vs_world_elapsed(vs_world *w, uint32_t from, uint32_t to, float *ms)
{
    if (!w || !ms || from >= EVENT_POOL || to >= EVENT_POOL) return VS_E_INVALID;
    HIPW(w, hipEventSynchronize(w->events[to]));
    HIPW(w, hipEventElapsedTime(ms, w->events[from], w->events[to]));
    return VS_OK;
}

int vs_world_read_nodes(vs_world *w, uint32_t id, float *pos, float *vel)
{
    if (!w || id >= w->cfg.max_scenes || !w->recs[id].live) return VS_E_INVALID;
    if (int rc = vs_world_synchronize(w)) return rc;
    const SceneRec &r = w->recs[id];
    const size_t N = w->cfg.max_nodes;
    for (int a = 0; a < 3; a++) {
        if (pos) HIPW(w, hipMemcpy(pos + a * r.node_count, w->d_nodes + a * N + r.node_begin, r.node_count * 4, hipMemcpyDeviceToHost));
        if (vel) HIPW(w, hipMemcpy(vel + a * r.node_count, w->d_nodes + (6 + a) * N + r.node_begin, r.node_count * 4, hipMemcpyDeviceToHost));
    }
    return VS_OK;
}

int vs_world_read_splats(vs_world *w, uint32_t id, float *mean, float *cov)
{
    if (!w || id >= w->cfg.max_scenes || !w->recs[id].live) return VS_E_INVALID;
    if (int rc = vs_world_synchronize(w)) return rc;
    const SceneRec &r = w->recs[id];
    const size_t G = w->cfg.max_splats, n = r.splat_count;
    std::vector<float> tmp(n);
    for (int k = 0; k < 9; k++) {
        float *dst = k < 3 ? (mean ? mean + k * n : nullptr) : (cov ? cov + (k - 3) * n : nullptr);
        if (!dst || !n) continue;
        HIPW(w, hipMemcpy(tmp.data(), w->d_splat_out + k * G + r.splat_begin, n * 4, hipMemcpyDeviceToHost));
        for (size_t j = 0; j < n; j++) dst[r.splat_perm[j]] = tmp[j];
    }
    return VS_OK;
}

int vs_world_body_info(vs_world *w, uint32_t id, uint32_t body, vs_body_info *out)
{
    if (!w || !out || id >= w->cfg.max_scenes || !w->recs[id].live || body >= w->recs[id].bodies.size())
        return VS_E_INVALID;
    const BodyRec &br = w->recs[id].bodies[body];
    *out = vs_body_info{};
    out->node_count = (uint32_t)br.node_count;
    out->n_surface = (uint32_t)br.surf_count;
    for (int c = 0; c < 8; c++) {
        out->color_count[c] = (uint32_t)br.color_count[c];
        out->n_cells += (uint32_t)br.color_count[c];
    }
    return VS_OK;
}

int vs_world_read_lattice(vs_world *w, uint32_t id, uint32_t body, float *rest, float *inv_mass, uint32_t *cells,
                          uint32_t *surface)
{
    vs_body_info info;
    if (int rc = vs_world_body_info(w, id, body, &info)) return rc;
    if (int rc = vs_world_synchronize(w)) return rc;
    const BodyRec &br = w->recs[id].bodies[body];
    const size_t n = br.node_count;
    for (int a = 0; a < 3 && rest; a++)
        HIPW(w, hipMemcpy(rest + a * n, plane(w, 10 + a) + br.node_begin, n * 4, hipMemcpyDeviceToHost));
    if (inv_mass) HIPW(w, hipMemcpy(inv_mass, plane(w, 9) + br.node_begin, n * 4, hipMemcpyDeviceToHost));
    if (cells && info.n_cells)
        HIPW(w, hipMemcpy(cells, w->d_cells + br.color_begin[0], info.n_cells * 4ull, hipMemcpyDeviceToHost));
    if (surface && info.n_surface) {
        HIPW(w, hipMemcpy(surface, w->d_surf + br.surf_begin, info.n_surface * 4ull, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < info.n_surface; i++) surface[i] -= (uint32_t)br.node_begin;
    }
    return VS_OK;
}

int vs_world_read_splat_axes(vs_world *w, uint32_t id, float *mean, float *rot, float *axis)
{
    if (!w || id >= w->cfg.max_scenes || !w->recs[id].live) return VS_E_INVALID;
    if (int rc = vs_world_synchronize(w)) return rc;
    const SceneRec &r = w->recs[id];
    const size_t G = w->cfg.max_splats, n = r.splat_count;
    std::vector<float> tmp(n);
    const int planes[10] = { 0, 1, 2, 9, 10, 11, 12, 13, 14, 15 };
    for (int k = 0; k < 10; k++) {
        float *dst = k < 3 ? (mean ? mean + k * n : nullptr)
                     : k < 7 ? (rot ? rot + (k - 3) * n : nullptr)
                             : (axis ? axis + (k - 7) * n : nullptr);
        if (!dst || !n) continue;
        HIPW(w, hipMemcpy(tmp.data(), w->d_splat_out + planes[k] * G + r.splat_begin, n * 4, hipMemcpyDeviceToHost));
        for (size_t j = 0; j < n; j++) dst[r.splat_perm[j]] = tmp[j];
    }
    return VS_OK;
}

int vs_world_log_pairs(vs_world *w, uint32_t cap)
{
    if (!w) return VS_E_INVALID;
    HIPW(w, hipStreamSynchronize(w->stream));
    if (w->d_pairlog) HIPW(w, hipFree(w->d_pairlog));
    w->d_pairlog = nullptr;
    w->pair_cap = 0;
    if (cap) {
        HIPW(w, hipMalloc(reinterpret_cast<void **>(&w->d_pairlog), sizeof(uint32_t) * w->cfg.max_surface * (1 + (size_t)cap)));
        HIPW(w, hipMemset(w->d_pairlog, 0, sizeof(uint32_t) * w->cfg.max_surface * (1 + (size_t)cap)));
        w->pair_cap = cap;
    }
    return VS_OK;
}

int vs_world_read_contacts(vs_world *w, uint32_t id, uint32_t *keys, int32_t *counts, uint32_t *pairs)
{
    if (!w || id >= w->cfg.max_scenes || !w->recs[id].live) return VS_E_INVALID;
    if (int rc = vs_world_synchronize(w)) return rc;
    const SceneRec &r = w->recs[id];
    if (!r.surf_count) return VS_OK;
    const size_t S = w->cfg.max_surface;
    if (keys) HIPW(w, hipMemcpy(keys, w->d_surf + 3 * S + r.surf_begin, r.surf_count * 4, hipMemcpyDeviceToHost));
    if (w->d_pairlog) {
        std::vector<int32_t> own;
        if (pairs && !counts) {
            own.resize(r.surf_count);
            counts = own.data();
        }
        if (counts) HIPW(w, hipMemcpy(counts, w->d_pairlog + r.surf_begin, r.surf_count * 4, hipMemcpyDeviceToHost));
        if (pairs) {
            HIPW(w, hipMemcpy(pairs, reinterpret_cast<uint32_t *>(w->d_pairlog + S) + r.surf_begin * w->pair_cap,
                              r.surf_count * w->pair_cap * 4, hipMemcpyDeviceToHost));
            // The log holds world record ids: hand back the scene's own.
            for (uint64_t s = 0; s < r.surf_count; s++) {
                const uint32_t n = std::min<uint32_t>((uint32_t)std::max(counts[s], 0), w->pair_cap);
                for (uint32_t k = 0; k < n; k++) pairs[s * w->pair_cap + k] -= (uint32_t)r.surf_begin;
            }
        }
    }
    return VS_OK;
}

int vs_world_stats_get(const vs_world *w, vs_world_stats *out)
{
    if (!w || !out) return VS_E_INVALID;
    *out = vs_world_stats{};
    for (const SceneRec &r : w->recs) {
        if (!r.live) continue;
        out->scenes++;
        out->bodies += r.body_count;
        out->nodes += r.node_count;
        out->cells += r.cell_count;
        out->surface += r.surf_count;
        out->buckets += r.bucket_count;
        out->splats += r.splat_count;
        out->driven += r.drive_count;
    }
    const vs_world_config &c = w->cfg;
    out->bytes_reserved = 4ull * NODE_PLANES * c.max_nodes + 4ull * c.max_cells + sizeof(vs_body) * c.max_bodies +
                          sizeof(vs_scene_params) * c.max_scenes + (4ull * SURF_PLANES + 32) * c.max_surface +
                          4ull * (4 + BUCKET_SLOTS + EXT_SLOTS) * c.max_buckets + 4ull * DRIVE_PLANES * c.max_driven +
                          4ull * (10 + SPLAT_OUT_PLANES) * c.max_splats;
    out->graph_captures = w->captures;
    return VS_OK;
}

int vs_world_profile(vs_world *w, uint32_t substeps, int32_t splats, float *ms, uint32_t *launches)
{
    if (!w || !ms || !launches) return VS_E_INVALID;
    HIPW(w, hipSetDevice(w->cfg.device));
    if (int rc = prepare(w)) return rc;
    Profile prof;
    if (int rc = enqueue_frame(w, substeps, splat_mode(splats), -1, &prof)) return rc;
    HIPW(w, hipStreamSynchronize(w->stream));
    for (int k = 0; k < VS_P_COUNT; k++) {
        ms[k] = 0.f;
        launches[k] = 0;
    }
    for (const auto &m : prof.marks) {
        float t = 0.f;
        HIPW(w, hipEventElapsedTime(&t, prof.pool[m.second], prof.pool[m.second + 1]));
        ms[m.first] += t;
        launches[m.first]++;
    }
    return VS_OK;
}

int vs_world_time_pass(vs_world *w, int32_t kind, uint32_t rounds, float *ms, uint32_t *launches)
{
    if (!w || !ms || !launches || rounds == 0 || kind < 0 || kind >= VS_P_PAIR) return VS_E_INVALID;
    // Repeated contact launches would leave the grid's counters and the
    // overflow list counted up (only the node planes are restored), so the
    // contact kinds are timed inside whole frames (vs_world_profile) only.
    if (kind >= VS_P_BIN && kind <= VS_P_SETTLE) {
        return fail(w, VS_E_INVALID, "time_pass: contact passes are timed in whole frames (vs_world_profile)");
    }
    HIPW(w, hipSetDevice(w->cfg.device));
    if (int rc = prepare(w)) return rc;
    const int p = kind == VS_P_ELASTIC ? P_CELL
                  : kind == VS_P_DRIVE ? P_DRIVE
                  : kind == VS_P_SPLAT ? P_SPLAT
                  : kind == VS_P_PREDICT ? P_NODE
                  : kind == VS_P_FINALIZE ? P_LAST
                  : kind == VS_P_FINALIZE_PREDICT ? P_REST
                                                  : P_SURF;
    const int kinds = kind == VS_P_ELASTIC ? 13 : 1;  // elastic: colours 0..7, row classes 8..11, bodies 12
    auto table = [&](int c) { return kind == VS_P_ELASTIC ? elastic_table(c) : p; };
    uint32_t per_round = 0;
    for (int c = 0; c < kinds; c++) per_round += grid_for(w, table(c), -1) != 0;
    *ms = 0.f;
    *launches = 0;
    if (per_round == 0) return VS_OK;
    const size_t bytes = sizeof(float) * 10 * w->cfg.max_nodes;
    void *snap = nullptr;
    HIPW(w, hipMalloc(&snap, bytes));
    int rc = VS_OK;
    hipEvent_t a = w->tune[0], b = w->tune[1];
    auto run = [&]() -> int {
        HIPW(w, hipMemcpyAsync(snap, w->d_nodes, bytes, hipMemcpyDeviceToDevice, w->stream));
        HIPW(w, hipEventRecord(a, w->stream));
        int e = VS_OK;
        for (uint32_t r = 0; r < rounds && !e; r++)
            for (int c = 0; c < kinds && !e; c++) e = enqueue_pass(w, kind, c, -1, 0);
        HIPW(w, hipEventRecord(b, w->stream));
        // Restored whether or not every launch went in.
        HIPW(w, hipMemcpyAsync(w->d_nodes, snap, bytes, hipMemcpyDeviceToDevice, w->stream));
        HIPW(w, hipStreamSynchronize(w->stream));
        if (e) return e;
        HIPW(w, hipEventElapsedTime(ms, a, b));
        return VS_OK;
    };
    rc = run();
    (void)hipFree(snap);
    if (rc) return rc;
    *launches = rounds * per_round;
    return VS_OK;
}

int vs_world_autotune(vs_world *w, uint32_t substeps, char *report, uint32_t report_bytes)
{
    if (!w || substeps == 0) return VS_E_INVALID;
    HIPW(w, hipSetDevice(w->cfg.device));
    if (int rc = prepare(w)) return rc;
    // The search steps the live scenes; snapshot the node planes, start every
    // candidate from them (so each is timed on the same state, not on where
    // the previous candidates left the scenes), and put them back afterwards
    // so tuning leaves the simulation where it was.
    const size_t bytes = sizeof(float) * 10 * w->cfg.max_nodes;
    void *snap = nullptr;
    HIPW(w, hipMalloc(&snap, bytes));
    HIPW(w, hipMemcpyAsync(snap, w->d_nodes, bytes, hipMemcpyDeviceToDevice, w->stream));
    struct Cand {
        uint32_t block, tile, persistent, xcd, colours, branches;
        float ms;
    };
    // Launch shapes for one colour per pass; colour pairs by rows take a
    // whole number of rows per workgroup, so only the block size, the grid
    // and the tile order vary for them.
    std::vector<Cand> cands;
    // Colour pairs also as one chain or two side by side.
    for (uint32_t colours : { 0u, 1u })
        for (uint32_t block : { 64u, 128u, 256u, 512u })
            for (uint32_t mult : { 1u, 2u, 4u })
                for (uint32_t persistent : { 0u, 2048u })
                    for (uint32_t xcd : { 0u, 1u })
                        for (uint32_t branches : { 1u, 2u }) {
                            if ((xcd && !persistent) || (!colours && mult > 1) || (colours && branches > 1)) continue;
                            cands.push_back(Cand{ block, block * mult, persistent, xcd, colours, branches, 0.f });
                        }
    const Cand saved{ w->block_cells, w->tile_cells, w->cfg.persistent, w->cfg.xcd_chunked, w->cfg.separate_colours,
                      w->cfg.branches, 0.f };
    auto apply = [&](const Cand &c) {
        w->block_cells = c.block;
        w->tile_cells = c.tile;
        w->cfg.persistent = c.persistent;
        w->cfg.xcd_chunked = c.xcd;
        w->cfg.separate_colours = c.colours;
        w->cfg.branches = c.branches;
        w->tables_dirty = true;
    };
    // Frames as the batched path runs them (captured once per shape, then
    // replayed): a warm-up frame (the capture), then two trials of three
    // frames each from the snapshot; the faster trial counts (candidates
    // differ by about 1%, the run-to-run noise of one trial).
    auto time = [&](Cand &c) -> int {
        apply(c);
        if (int rc = prepare(w)) return rc;
        c.ms = 0.f;
        // Frames with the splat output the world last produced (how a
        // frame runs: in two chains each chain's splat pass overlaps the
        // other's last substep).
        const int32_t splats = (int32_t)w->splat_mode;
        if (int rc = vs_world_frame(w, substeps, splats, 1)) return rc;  // warm-up
        for (int trial = 0; trial < 2; trial++) {
            HIPW(w, hipMemcpyAsync(w->d_nodes, snap, bytes, hipMemcpyDeviceToDevice, w->stream));
            hipEvent_t a = w->tune[0], b = w->tune[1];
            HIPW(w, hipEventRecord(a, w->stream));
            for (int rep = 0; rep < 3; rep++)
                if (int rc = vs_world_frame(w, substeps, splats, 1)) return rc;
            HIPW(w, hipEventRecord(b, w->stream));
            HIPW(w, hipEventSynchronize(b));
            float ms = 0.f;
            HIPW(w, hipEventElapsedTime(&ms, a, b));
            c.ms = trial == 0 ? ms : std::min(c.ms, ms);
        }
        return VS_OK;
    };
    int rc = VS_OK;
    for (Cand &c : cands)
        if ((rc = time(c))) break;
    HIPW(w, hipMemcpyAsync(w->d_nodes, snap, bytes, hipMemcpyDeviceToDevice, w->stream));
    HIPW(w, hipStreamSynchronize(w->stream));
    HIPW(w, hipFree(snap));
    if (rc) {
        apply(saved);
        return rc;
    }
    const Cand *best = &cands[0];
    for (const Cand &c : cands)
        if (c.ms < best->ms) best = &c;
    apply(*best);
    if (report && report_bytes) {
        auto one = [](const Cand &c) {
            char buf[224];
            std::snprintf(buf, sizeof buf,
                          "{\"block\":%u,\"tile\":%u,\"persistent\":%u,\"xcd\":%u,\"separate_colours\":%u,"
                          "\"branches\":%u,\"ms\":%.4f}",
                          c.block, c.tile, c.persistent, c.xcd, c.colours, c.branches, c.ms);
            return std::string(buf);
        };
        std::string out = "{\"candidates\":[";
        for (size_t i = 0; i < cands.size(); i++) out += (i ? "," : "") + one(cands[i]);
        out += "],\"best\":" + one(*best) + "}";
        std::snprintf(report, report_bytes, "%s", out.c_str());
    }
    return VS_OK;
}

}  // extern "C"

namespace vsim {

int world_read_positions_async(vs_world *w, uint32_t id, float *pos)
{
    if (!w || id >= w->cfg.max_scenes || !w->recs[id].live || !pos) return VS_E_INVALID;
    HIPW(w, hipSetDevice(w->cfg.device));
    const SceneRec &r = w->recs[id];
    const size_t N = w->cfg.max_nodes;
    for (int a = 0; a < 3; a++)
        HIPW(w, hipMemcpyAsync(pos + a * r.node_count, w->d_nodes + a * N + r.node_begin, r.node_count * 4,
                               hipMemcpyDeviceToHost, w->stream));
    return VS_OK;
}

hipStream_t world_stream(vs_world *w) { return w ? w->stream : nullptr; }

hipEvent_t world_event(vs_world *w, uint32_t slot) { return w && slot < EVENT_POOL ? w->events[slot] : nullptr; }

}  // namespace vsim
