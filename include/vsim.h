/*
 * vsim.h -- C ABI of the MI355X solver for the visionary-sim substep.
 *
 * Two layers, both plain C (pointers, sizes, integer handles):
 *
 *  1. The compute backend: device buffers and named kernel launches. It is the
 *     drop-in for the kernel package's `ComputeBackend`
 *     (packages/visionary-sim/src/sim/compute/backend.js:114-163), which is the
 *     only place the reference touches the GPU. The JS host binds it through
 *     the Node-API addon in visionary-sim_amd/node/.
 *
 *  2. The world: many scenes resident on the device at once, stepped as one
 *     batch per pass (one persistent launch per colour across every scene,
 *     a hipGraph per frame), plus the plain per-scene launch path and a
 *     continuous-batching server. The reference has no counterpart beyond its
 *     co-location of bodies in one dispatch (sim/state/particles.js:18-21,
 *     sim/discretize/lattice/elasticCells.js:143-163).
 *
 * Every entry point returns 0 on success and a negative VS_E_* code on
 * failure; vs_last_error() names the failure in the same words the reference's
 * backend uses for its Error messages where one exists.
 *
 * Node state is structure-of-arrays: a "node buffer" holds planes of `plane`
 * floats each (x plane, then y, then z), not the reference's vec4f records
 * (sim/state/state.wgsl:16-19). Buffer layout is private to the solver in the
 * reference too (deformation/bind.js:38-47).
 */
#ifndef VSIM_H
#define VSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */
#define VS_OK 0
#define VS_E_INVALID -1   /* bad argument, unknown kernel or parameter     */
#define VS_E_DEVICE -2    /* HIP runtime error                              */
#define VS_E_NOMEM -3     /* allocation beyond the device or budget limit    */
#define VS_E_DESTROYED -4 /* use after destroy                              */

/* ------------------------------------------------------- shared records */

/* Per-body constants. Replaces CellParams (lattice/wgsl/cells.wgsl:18-31,
 * packed by elasticCells.js:222-249) and NodeParams (state/state.wgsl:24-33,
 * packed by particles.js:145-159), plus the lattice origin the splat sampler
 * reads (lattice/sample.wgsl:17-21). 80 bytes. `damp` is exp(-damping * dt),
 * the factor finalize.wgsl:72 applies, formed on the host once per dt. */
typedef struct vs_body {
    uint32_t cell_dims[3];
    uint32_t node_offset;
    uint32_t node_dims[3];
    uint32_t scene;
    float h, alpha_dev, alpha_hyd, gamma;
    float relaxation, max_stretch, friction, damp;
    float origin[3];
    uint32_t node_count;
} vs_body;

/* Per-scene uniforms: the Params blocks of predict.wgsl:1-5,
 * finalize.wgsl:1-6, contacts.wgsl:1-8 and drive.wgsl:1-7 in one record,
 * plus the sphere collider and two contact bounds. 160 bytes. */
typedef struct vs_scene_params {
    float dt;
    float gravity[3];
    float tank_min[3];
    float gain;
    float tank_max[3];
    float bucket_size;
    float grid_origin[3];
    uint32_t bucket_count;
    uint32_t grid_dims[3];
    uint32_t grid_base;    /* first bucket of this scene in the grid buffers */
    float target[3];       /* grab target (drive.wgsl:2)                     */
    float response_time;
    float target_vel[3];
    uint32_t drive_law;    /* 0 hard, 1 compliant                            */
    uint32_t surf_base;    /* first surface record of this scene             */
    uint32_t surf_count;
    uint32_t bucket_slots; /* 32, BUCKET_SLOTS of contacts.wgsl:75           */
    uint32_t flags;        /* VS_SCENE_*                                     */
    float sphere[4];       /* sphere collider: centre xyz, radius (<= 0: none).
                              No reference counterpart: finalize moves a node
                              inside it out along the radius, with the tank
                              planes' friction (finalize.wgsl:29-42)          */
    float reach_max;       /* contact: twice the largest record radius; 0 = no bound */
    float skip_margin;     /* contact: rounding allowance of the bucket bound, metres.
                              A neighbour bucket farther than reach_max +
                              skip_margin from a record along an axis holds no
                              mate in reach, and collide may pass it by        */
    uint32_t pad1[2];
} vs_scene_params;

/* A scene whose finalize follows its last colour directly (no contact and
 * no tensile-limit sweeps between them): the last colour-pair pass may
 * finalize, and predict the next substep, for the nodes it writes. */
#define VS_SCENE_FOLD 1u
/* A scene with no kinematic drive (nothing between its predict and its first
 * colour): at a frame's first substep the first colour-pair pass may predict
 * the nodes it writes. */
#define VS_SCENE_FOLD_PREDICT 2u

/* One unit of work for a launch: `count` consecutive items from `begin` in the
 * pass's item array (nodes, cells, surface records, driven records, splats),
 * all of one body (or, for contact and drive, one scene). A tile buffer is a
 * 16-byte header {count, 0, 0, 0} followed by the tiles. In a colour-pair
 * table a tile is `count` rows of cells from row `begin`, or, with
 * VS_NODE_TILE set in `count`, entries `begin` .. `begin + count - 1` of the
 * body's list of the nodes those rows leave out (in node-id order). */
typedef struct vs_tile {
    uint32_t begin, count, body, scene;
} vs_tile;
#define VS_NODE_TILE 0x80000000u

/* ------------------------------------------------- layer 1: the backend */

typedef struct vs_backend vs_backend;

/* Kernel ids, keyed by the reference's pass labels. */
#define VS_K_PREDICT 0          /* sim-predict          predict.wgsl:19-36        */
#define VS_K_FINALIZE 1         /* sim-finalize         finalize.wgsl:44-74       */
#define VS_K_ELASTIC 2          /* sim-elastic-cells    project.wgsl:35-184       */
#define VS_K_CONTACT_BIN 3      /* sim-contact-bin      bin.wgsl:9-23             */
#define VS_K_CONTACT_SORT 4     /* sim-contact-sort     sort.wgsl:14-53           */
#define VS_K_CONTACT_COLLIDE 5  /* sim-contact-collide  collide.wgsl:16-125       */
#define VS_K_CONTACT_SETTLE 6   /* sim-contact-settle   settle.wgsl:20-36         */
#define VS_K_DRIVE_HARD 7       /* sim-drive-hard       driveHard.wgsl:15-35      */
#define VS_K_DRIVE_COMPLIANT 8  /* sim-drive-compliant  driveCompliant.wgsl:24-41 */
#define VS_K_SPLAT 9            /* splat deformation    sample.wgsl:40-80 + covariance.wgsl:30-42 */
#define VS_K_PICK 10            /* sim-pick-ray         nearestNodeToRay, pick.js:41-68           */
#define VS_K_GRIP 11            /* sim-grip-radius      withinRadius, grip.js:99-138              */
#define VS_K_COUNT 12

/* The question a grab asks (simulator.js:383-425), the `params` buffer of
 * VS_K_PICK and VS_K_GRIP. f64, as the JS computes it. */
typedef struct vs_grab_params {
    double origin[3];      /* pick: ray origin                                  */
    double direction[3];   /* pick: unit ray direction                          */
    double radius;         /* pick: aim tolerance; grip: hand radius, metres    */
    uint32_t anchor;       /* grip: the node the ray picked                     */
    uint32_t begin, end;   /* grip: the anchor's body's node range (bodyOfNode) */
    uint32_t pad0;
} vs_grab_params;

/* Outputs of the splat pass (a bit mask). VS_SPLAT_AXES is the engine's splat
 * format: the covariance taken apart into a unit rotation (xyzw) and three
 * axis lengths by four sweeps of cyclic Jacobi, the step that follows the
 * deformation in the reference (apps/simulator/src/render/ellipsoid.wgsl:32-91). */
#define VS_SPLAT_COV 1
#define VS_SPLAT_AXES 2

/* Buffer slots of a launch. Node buffers hold 3 planes of `plane` floats. */
#define VS_B_POS 0        /* nodePos  (3 planes)                              */
#define VS_B_PREV 1       /* nodePrev (3 planes)                              */
#define VS_B_VEL 2        /* nodeVel  (3 planes)                              */
#define VS_B_INVMASS 3    /* nodeInvMass (pick, grip: material inverse mass with
                             static nodes zeroed, the grabbable test)         */
#define VS_B_BODIES 4     /* vs_body[]                                        */
#define VS_B_SCENES 5     /* vs_scene_params[] (the pass's uniforms)          */
#define VS_B_TILES 6      /* tile buffer                                      */
#define VS_B_ITEMS 7      /* elastic: u32 linear cell per item;
                             drive: 4 planes (u32 node, offset x, y, z);
                             contact: 9 planes (u32 node, u32 body, radius, u32 key,
                                      u32 rank, delta x, y, z, hits);
                             splats: float4 quaternion xyzw per splat, then 6 planes
                                     (mean x, y, z, scale x, y, z)                */
#define VS_B_GRID 8       /* contact: counters then slots                     */
#define VS_B_OUT 9        /* pick: {u32, i32 node (-1 = miss), f64 depth} then 16 B per
                                   workgroup;
                             grip: {u32 count, u32 done, u32 found, u32 kept, u32 status
                                   (0 ok, 1 anchor not grabbable, 2 over capacity,
                                   3 bad range), 3 pad}, then `pair_cap` u32 held nodes
                                   (anchor first, nearest first; padded to 8 B), then
                                   `pair_cap` f64 xyz offsets from the anchor;
                             splats: 9 planes (mean x, y, z, cov xx xy xz yy yz zz),
                                     16 with VS_SPLAT_AXES (then rotation x, y, z, w
                                     and axis lengths x, y, z);
                             contact collide: optional pair log, i32 count per
                             record then pair_cap u32 mate ids per record     */
#define VS_B_COUNT 10

typedef struct vs_launch {
    int32_t kernel;     /* VS_K_*                                             */
    uint32_t grid;      /* workgroups, at least 1 (one per tile, or fewer: persistent) */
    uint32_t block;     /* threads per workgroup; 0 = 64                      */
    uint32_t plane;     /* floats per plane of the node buffers               */
    uint32_t items;     /* floats per plane of the item buffer (contact, drive, splats) */
    uint32_t flags;     /* elastic: 1 = limit-only sweep; splats: VS_SPLAT_* outputs (0 = cov) */
    uint32_t buckets;   /* contact: buckets in the grid buffer (counters, then 32 slots each) */
    uint32_t pair_cap;  /* contact collide: pair log capacity per record (0 = off);
                           grip: most nodes held, anchor included (drive capacity) */
    uint64_t buffers[VS_B_COUNT];
} vs_launch;

/* Creates a backend on HIP device `device` with its own stream.
 * Mirrors `new ComputeBackend(device)` (backend.js:115). */
int vs_backend_create(int device, vs_backend **out);
/* Frees every allocation the backend owns; idempotent (backend.js:159-162). */
void vs_backend_destroy(vs_backend *b);
/* Last error message of this backend, or of the last failed create. */
const char *vs_last_error(const vs_backend *b);

/* `buffer(size)` (backend.js:116-123): rejects sizes beyond the device limit
 * with "Simulation buffer requires <n> bytes; device storage/buffer limit is <m>".
 * The handle is an opaque 64-bit id (the device address). */
int vs_buffer_alloc(vs_backend *b, uint64_t bytes, uint64_t *handle);
/* `Allocation.destroy()` (backend.js:58-61); freeing twice is a no-op. */
int vs_buffer_free(vs_backend *b, uint64_t handle);
/* `Allocation.write(offset, data)` (backend.js:39-41): stream-ordered; the
 * source may be reused as soon as the call returns. */
int vs_buffer_write(vs_backend *b, uint64_t handle, uint64_t offset, const void *src, uint64_t bytes);
/* `Allocation.read(offset, size, target)` (backend.js:42-57): waits for every
 * queued launch, then copies. A zero-size read is a no-op. */
int vs_buffer_read(vs_backend *b, uint64_t handle, uint64_t offset, uint64_t bytes, void *dst);
/* Kernel id for a reference pass label ("sim-predict", "sim-elastic-cells",
 * ...), or VS_E_INVALID. Stands in for `kernel(name, wgsl)` (backend.js:124-136);
 * the WGSL text is not needed. */
int vs_kernel_id(const char *name);
/* `dispatch(sequence, label)` (backend.js:138-158): validates every launch,
 * then enqueues them in order on the backend's stream with no host wait. */
int vs_dispatch(vs_backend *b, const vs_launch *launches, uint32_t count);
/* Waits until every enqueued launch has finished. */
int vs_synchronize(vs_backend *b);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The grip of VS_K_GRIP (withinRadius, grip.js:99-138)
 * with the question passed by value and its candidate queue and counters in
 * a static device array the library owns: 256 B + 16 B x
 * VS_GRIP_NOALLOC_MAX_NODES (4 MiB + 256 B), held on every GPU the library
 * runs on until the process ends. `pos` (3 planes of `plane` floats) and
 * `mass` (the grab mass, `plane` floats) are backend buffers, as is `out`,
 * laid out as VS_B_OUT of VS_K_GRIP for `limit` held nodes; `grid` is 1..1024
 * workgroups. A body range (params->end - params->begin) above
 * VS_GRIP_NOALLOC_MAX_NODES is rejected with VS_E_INVALID. Enqueued on the
 * backend's stream with no host wait; calls no allocating routine. Calls on
 * one GPU share the array, so calls that overlap on one GPU (two streams,
 * two backends) race. */
#define VS_GRIP_NOALLOC_MAX_NODES (1u << 18)
int vs_grip_noalloc(vs_backend *b, uint64_t pos, uint32_t plane, uint64_t mass, const vs_grab_params *params,
                    uint32_t limit, uint32_t grid, uint64_t out);

/* ------------------------------------------------------ layer 2: world */

typedef struct vs_world vs_world;

/* Capacities of the device pools, i.e. the HBM budget. */
typedef struct vs_world_config {
    int32_t device;
    uint32_t max_scenes;
    uint32_t max_bodies;
    uint64_t max_nodes;
    uint64_t max_cells;
    uint64_t max_surface;
    uint64_t max_buckets;
    uint64_t max_driven;
    uint64_t max_splats;
    uint32_t tile_nodes;    /* items per tile for node passes (0 = 256)        */
    uint32_t tile_cells;    /* items per tile for the colour passes (0 = 256)  */
    uint32_t block_cells;   /* threads per workgroup of the colour passes      */
    uint32_t persistent;    /* workgroups of a persistent launch (0 = 1 per tile) */
    uint32_t xcd_chunked;   /* 1: give each XCD a contiguous run of tiles      */
    uint32_t separate_node_passes; /* 1: keep every substep's finalize and the
                               next substep's predict as two node passes; 0 (the
                               default) runs them as one pass inside a frame
                               (same arithmetic, same results)                 */
    uint32_t separate_colours; /* 1: one launch per colour for every body; 0
                               (the default) runs the colours of a dense body in
                               pairs (2c, 2c+1) over whole x-rows of cells, one
                               launch per pair (same arithmetic, same results) */
    uint32_t colour_launches; /* 1: every body's colours go through the colour-pair
                               and colour launches; 0 (the default) runs the
                               eight colours of a small dense body (at most 11^3
                               nodes) of a scene whose node passes do not fold
                               into its colour passes (contact, tensile sweeps
                               or drive) in one launch, the body in LDS (same
                               arithmetic, same results)                        */
    uint32_t branches;        /* 2: a batched frame runs as two chains of launches
                               side by side (scenes split by work, contact scenes
                               all in the first), captured as two branches of the
                               frame's graph, so that one chain's launches fill
                               the other's tails; 0 or 1: one chain (same
                               results)                                          */
} vs_world_config;

/* One body of a scene, in the reference's host layout: the dense node grid
 * of buildLattice (lattice.js:166-239) and its colour-grouped occupied cells. */
typedef struct vs_body_desc {
    uint32_t cell_dims[3];
    uint32_t n_cells;            /* occupied cells                          */
    const uint32_t *cells;       /* linear cell ids grouped by colour        */
    uint32_t color_count[8];     /* cells of each colour, in order           */
    float origin[3];
    float h;
    const float *rest;           /* 3 * node_count, xyz per node            */
    const float *inv_mass;       /* node_count, prescribed nodes already 0  */
    float velocity[3];           /* initial velocity of every node (0 = at rest) */
    float alpha_dev, alpha_hyd, gamma, relaxation, max_stretch, friction, damping;
    uint32_t n_surface;          /* contact records of this body            */
    const uint32_t *surface;     /* body-local node ids, ascending          */
    float contact_radius;
    uint32_t n_splats;
    const float *splat_mean;     /* 3 * n_splats                            */
    const float *splat_quat;     /* 4 * n_splats, xyzw                      */
    const float *splat_scale;    /* 3 * n_splats                            */
    /* Lattice built on the device (buildLattice and surfaceNodes,
     * lattice.js:128-239): when `occupancy` is set, n_cells, cells,
     * color_count, origin, h, rest, inv_mass, n_surface and surface are
     * ignored and computed from the fields below. */
    const uint8_t *occupancy;    /* cell_dims product bytes, x fastest, non-zero = occupied */
    double origin64[3];          /* min corner of cell (0,0,0); rest = f32(origin + i*h) in f64 */
    double h64;                  /* cell edge                               */
    double density;              /* kg/m^3: corner mass density*h^3/8 per incident cell */
    uint32_t n_static;           /* body-local nodes prescribed static (w = 0, */
    const uint32_t *static_nodes;/*   staticPrescription.js:9-20)             */
} vs_body_desc;

typedef struct vs_scene_desc {
    uint32_t n_bodies;
    const vs_body_desc *bodies;
    float dt;
    float gravity[3];
    float tank_min[3];
    float tank_max[3];
    int32_t contact;             /* 0 off; 1 the four contact passes when the scene
                                    has two bodies or more; 2 always            */
    float bucket_size;           /* grid edge, 2 x the widest contact radius (bodyContact.js:87) */
    float grid_origin[3];        /* tank min - edge (broadphase.js:46-58)   */
    uint32_t grid_dims[3];       /* 0: derive origin and dims from the tank  */
    uint32_t n_driven;
    const uint32_t *driven;      /* scene-global node ids                   */
    const float *driven_offset;  /* 3 per record                            */
    uint32_t drive_law;
    float response_time;
    float sphere[4];             /* sphere collider centre xyz, radius (<= 0: none) */
} vs_scene_desc;

int vs_world_create(const vs_world_config *cfg, vs_world **out);
void vs_world_destroy(vs_world *w);
const char *vs_world_error(const vs_world *w);
/* Admits a scene: allocates its ranges in every pool and uploads its state.
 * VS_E_NOMEM when a pool (the HBM budget) cannot hold it. */
int vs_world_add_scene(vs_world *w, const vs_scene_desc *desc, uint32_t *scene_id);
/* Releases a scene's ranges. The host does not wait for the device: the
 * frames enqueued before the removal still run over the scene, and the next
 * admission writes into its ranges only after them (stream-ordered). */
int vs_world_remove_scene(vs_world *w, uint32_t scene_id);
/* Restores a scene to rest (particles.js:245-249). */
int vs_world_reset_scene(vs_world *w, uint32_t scene_id);
/* Per-frame grab target of a scene's drive (simulator.js:309-315). */
int vs_world_set_target(vs_world *w, uint32_t scene_id, const float target[3], const float velocity[3]);
/* One frame of every admitted scene: `substeps` substeps as one batch per
 * pass, then splat deformation when `splats` is non-zero: VS_SPLAT_COV,
 * VS_SPLAT_AXES or both (any other non-zero value: the covariances). `graph`
 * replays a captured hipGraph of the frame. Asynchronous. */
int vs_world_frame(vs_world *w, uint32_t substeps, int32_t splats, int32_t graph);
/* The plain path: one scene's frame with per-scene launches, as the
 * reference dispatches a scene (stepper.js:101-108). Asynchronous. */
int vs_world_frame_scene(vs_world *w, uint32_t scene_id, uint32_t substeps, int32_t splats);
int vs_world_synchronize(vs_world *w);
/* Records an event after the work enqueued so far; returns its slot. */
int vs_world_record(vs_world *w, uint32_t *slot);
/* Milliseconds between two recorded events (after they completed). */
int vs_world_elapsed(vs_world *w, uint32_t from, uint32_t to, float *ms);
/* Copies a scene's node positions (3 planes of node_count) and velocities. */
int vs_world_read_nodes(vs_world *w, uint32_t scene_id, float *pos, float *vel);
/* Copies a scene's deformed splats: means (3 planes) and covariances (6 planes). */
int vs_world_read_splats(vs_world *w, uint32_t scene_id, float *mean, float *cov);
/* Sizes of one body as admitted (after a device lattice build, the counts it
 * found), and its lattice arrays: rest (3 planes), GPU inverse mass, cells
 * grouped by colour, surface nodes (body-local, ascending). */
typedef struct vs_body_info {
    uint32_t node_count, n_cells, n_surface, pad0;
    uint32_t color_count[8];
} vs_body_info;
int vs_world_body_info(vs_world *w, uint32_t scene_id, uint32_t body, vs_body_info *out);
int vs_world_read_lattice(vs_world *w, uint32_t scene_id, uint32_t body, float *rest, float *inv_mass, uint32_t *cells,
                          uint32_t *surface);
/* The engine format of the last VS_SPLAT_AXES frame: means (3 planes),
 * rotations (4 planes, xyzw) and axis lengths (3 planes). */
int vs_world_read_splat_axes(vs_world *w, uint32_t scene_id, float *mean, float *rot, float *axis);
/* Contact results of the last substep: bucket of each record and, when the
 * world was asked to log pairs (vs_world_log_pairs), up to `cap` mates each,
 * as the scene's own record numbers (0 = its first record). */
int vs_world_log_pairs(vs_world *w, uint32_t cap);
int vs_world_read_contacts(vs_world *w, uint32_t scene_id, uint32_t *keys, int32_t *counts, uint32_t *pairs);
/* Counters for the measurement harness. */
typedef struct vs_world_stats {
    uint64_t scenes, bodies, nodes, cells, surface, buckets, splats, driven;
    uint64_t bytes_reserved;
    uint32_t graph_captures, pad0;
} vs_world_stats;
int vs_world_stats_get(const vs_world *w, vs_world_stats *out);
/* Pass kinds of vs_world_profile. */
#define VS_P_PREDICT 0
#define VS_P_DRIVE 1
#define VS_P_ELASTIC 2
#define VS_P_BIN 3
#define VS_P_SORT 4
#define VS_P_COLLIDE 5
#define VS_P_SETTLE 6
#define VS_P_FINALIZE 7
#define VS_P_SPLAT 8
#define VS_P_FINALIZE_PREDICT 9 /* a substep's finalize and the next one's predict */
/* Breakdown of VS_P_ELASTIC in vs_world_profile (the same launches, counted
 * again): colour-pair launches with no node pass on board, and colour-pair
 * launches asked to run one (the first pair of a frame predicts, the last
 * pair of a substep finalizes, in scenes flagged VS_SCENE_FOLD_PREDICT /
 * VS_SCENE_FOLD). Not accepted by vs_world_time_pass. */
#define VS_P_PAIR 10
#define VS_P_PAIR_NODES 11
#define VS_P_COUNT 12
/* One frame of every admitted scene, batched but not captured, with HIP
 * events around every launch: ms[k] is the summed duration of the launches of
 * kind k, launches[k] their number. Synchronous. */
int vs_world_profile(vs_world *w, uint32_t substeps, int32_t splats, float *ms, uint32_t *launches);
/* Times `rounds` back-to-back batched launches of one pass kind (elastic:
 * one round = the eight colour passes; splats: the outputs of the last frame) with a HIP event before the first and
 * after the last, so *ms / *launches is the kernel's average duration with no
 * per-launch event in between. Node state is saved first and restored after,
 * so the simulation is left where it was. The contact kinds (VS_P_BIN ..
 * VS_P_SETTLE) are refused with VS_E_INVALID: their bucket counters would not
 * be left where they were. Synchronous. */
int vs_world_time_pass(vs_world *w, int32_t kind, uint32_t rounds, float *ms, uint32_t *launches);
/* Re-times the colour-pass launch shapes on the admitted scenes and keeps the
 * fastest (block size x cells per tile x persistent grid x XCD mapping x one
 * or two chains), timing graph-replayed frames with the splat output of the
 * world's last frame; node state is put back afterwards. */
int vs_world_autotune(vs_world *w, uint32_t substeps, char *report, uint32_t report_bytes);

/* ------------------------------------------------------ the server */

typedef struct vs_server vs_server;

typedef struct vs_server_config {
    uint32_t substeps;     /* substeps per frame                                 */
    int32_t splats;        /* deform splats every frame                          */
    int32_t graph;         /* replay frames from a captured hipGraph             */
    int32_t keep_results;  /* keep each finished scene's final node positions     */
} vs_server_config;

typedef struct vs_server_report {
    double wall_s;
    uint64_t node_substeps, frames;
    uint32_t requests, completed, iterations, max_active;
    double p50_ms, p99_ms, mean_ms, max_ms;   /* frame latency                  */
    double mean_queue_s, max_queue_s;          /* arrival to admission           */
} vs_server_report;

/* A continuous-batching server over `world` (which it borrows). */
int vs_server_create(vs_world *world, const vs_server_config *cfg, vs_server **out);
void vs_server_destroy(vs_server *s);
const char *vs_server_error(const vs_server *s);
/* Queues a scene (copied) arriving `arrival_s` seconds after run() starts,
 * to run `frames` frames; `targets` holds frames + 1 grab-target positions
 * (xyz) when the scene has a drive, else NULL. Submit in arrival order. */
int vs_server_submit(vs_server *s, const vs_scene_desc *desc, double arrival_s, uint32_t frames,
                     const float *targets, uint32_t *request_id);
/* Serves every request: admits arrivals while the world's pools hold them,
 * steps all admitted scenes one batched frame per iteration, retires finished
 * scenes. Frames are pipelined (the next is enqueued before the host waits
 * for the current one); a frame's completion is the GPU time of an event
 * after it. `realtime` = 0 ignores arrival times (all requests are ready). */
int vs_server_run(vs_server *s, int32_t realtime, vs_server_report *report);
/* Every frame latency (ms) of the last run, in completion order. */
int vs_server_latencies(const vs_server *s, float *ms, uint64_t cap, uint64_t *count);
/* Final node positions (3 planes) of a finished request (keep_results). */
int vs_server_result(const vs_server *s, uint32_t request_id, float *pos);
/* Seconds since run() began: arrival, admission and completion of a request. */
int vs_server_times(const vs_server *s, uint32_t request_id, double *arrival, double *admitted, double *finished);

/* ------------------------------------------------------ ingest
 *
 * Capture formats into occupancy grids, the input of the device lattice build
 * (vs_body_desc.occupancy). The grids match the reference's voxelisers cell
 * for cell: every float expression is evaluated in double, in the order the
 * JS writes it, with the rotation matrix rounded to f32 as mat3FromQuat
 * stores it (mat3.js:71-86). Math.hypot is evaluated as the square root of
 * the sum of squares.
 *
 * These calls allocate their own device buffers for the call's duration and
 * return host results allocated with malloc; free them with the matching
 * *_free. vs_ingest_error() names the last failure on this thread in the
 * reference's words. */

/* OccupancyGrid (voxelize.js:21-29): dims, origin, cell edge h, one byte per
 * cell, x fastest, 1 = material. */
typedef struct vs_occupancy {
    uint32_t dims[3];
    uint32_t pad0;
    double origin[3];
    double h;
    uint8_t *cells;
} vs_occupancy;

#define VS_SHAPE_SPHERE 0 /* size[0] = radius (shapes.js:60-67)                */
#define VS_SHAPE_BOX 1    /* size = half extents (shapes.js:89-108)            */
#define VS_SHAPE_TORUS 2  /* size[0] major, size[1] minor radius (shapes.js:32-43) */

/* An analytic shape moved by `n_offsets` (at most 4) translations
 * (translateShape, shapes.js:148-156), offset[0] the first applied: its sdf
 * subtracts offset[n_offsets - 1] from the query point first, and its bounds
 * add offset[0] first, as nested translateShape calls evaluate. */
typedef struct vs_shape {
    int32_t kind;
    uint32_t n_offsets;
    double size[3];
    double offset[4][3];
} vs_shape;

/* voxelizeCloud (cloudVolume.js:95-121): the ellipsoids at `sigma` standard
 * deviations stamped into a grid over their bounds, then (fill != 0) sealed
 * by `seal` dilations, flooded from the faces, eroded back and re-stamped.
 * Reference defaults: padding 2, sigma 2, seal 2, fill true. Rotations are
 * xyzw unit quaternions; scales in metres. */
int vs_voxelize_cloud(int device, const float *positions, const float *rotations, const float *scales, uint64_t count,
                      double element_size, uint32_t padding, double sigma, uint32_t seal, int32_t fill,
                      vs_occupancy *out);
/* voxelizeMesh (voxelize.js:174-203): two of three axis-parity votes or a
 * surface crossing, then (close != 0) a morphological close. `bounds` is
 * min[3], max[3]; NULL uses the vertices' bounds (meshBounds,
 * restGeometry.js:29-40). Reference defaults: padding 2, close true. */
int vs_voxelize_mesh(int device, const float *positions, uint64_t vertices, const uint32_t *indices, uint64_t triangles,
                     const double *bounds, double element_size, uint32_t padding, int32_t close, vs_occupancy *out);
/* voxelizeSdf (voxelize.js:118-148) of an analytic shape. */
int vs_voxelize_shape(int device, const vs_shape *shape, double element_size, uint32_t padding, vs_occupancy *out);
void vs_occupancy_free(vs_occupancy *grid);
const char *vs_ingest_error(void);

/* readPly (apps/simulator/src/assets/splats/ply.js:65-104): the first element
 * of a binary little-endian PLY held in memory. `names` (n_names of them)
 * selects the properties to keep; each becomes one malloc'd float column,
 * columns[i] for names[i], whatever its stored type. *count receives the
 * record count. */
int vs_ply_read(const void *bytes, uint64_t size, const char *const *names, uint32_t n_names, uint64_t *count,
                float **columns);
/* The first element's property names, space-separated, into `list`. */
int vs_ply_properties(const void *bytes, uint64_t size, char *list, uint64_t list_bytes, uint64_t *count);

/* SplatCloud (apps/simulator/src/assets/splats/cloud.js:52-81): positions,
 * xyzw unit rotations with w >= 0, scales in metres, opacity in [0, 1],
 * colours from the degree-0 harmonic. Arrays are malloc'd. */
typedef struct vs_cloud {
    uint64_t count;
    float *positions, *rotations, *scales, *opacity, *colors;
} vs_cloud;
/* cloudFromPly over the columns of an element already read: columns[i] holds
 * the property REQUIRED_PROPERTIES[i] (cloud.js:28-34) names, in that order:
 * x y z f_dc_0 f_dc_1 f_dc_2 opacity scale_0 scale_1 scale_2 rot_0 rot_1
 * rot_2 rot_3. */
int vs_cloud_from_columns(uint64_t count, const float *const *columns, vs_cloud *out);
/* readPly with the fourteen required properties, then cloudFromPly. */
int vs_cloud_from_ply(const void *bytes, uint64_t size, vs_cloud *out);
void vs_cloud_free(vs_cloud *cloud);

#ifdef __cplusplus
}
#endif

#endif
