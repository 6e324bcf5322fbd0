// kernels.h -- argument views and host launchers of the gfx950 kernels.
//
// Every pass runs over a tile buffer: a 16-byte header whose first word is the
// tile count, then vs_tile records. A workgroup takes whole tiles, so a tile's
// body and scene are uniform across it and their constants come in through
// scalar loads. The count lives in device memory, so a launch captured in a
// hipGraph stays valid when the host rewrites the tables between replays.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "../../include/vsim.h"

namespace vsim {

struct NodeView {
    float *px, *py, *pz;  // position planes
    float *qx, *qy, *qz;  // position at substep start
    float *vx, *vy, *vz;  // velocity planes
    const float *w;       // inverse mass, zero where prescribed or inactive
};

struct TileView {
    const uint32_t *count;  // header word 0
    const vs_tile *tiles;   // header + 16 bytes
};

// Contact surface records as planes (struct Surface, contacts.wgsl:36-46).
struct SurfaceView {
    const uint32_t *node;
    const uint32_t *body;
    const float *radius;
    uint32_t *key;   // bucket of the record this substep
    uint32_t *rank;  // arrival slot returned by the bucket counter
    float *dx, *dy, *dz, *hits;
    // Optional: the record's node position and radius, written by bin for
    // collide (positions do not change between the two). Null: collide reads
    // them through the node id, as the reference does.
    float4 *pay = nullptr;
    // Optional with `pay`: the record's node prev position and inverse mass,
    // written by bin, so that collide reaches a mate's in one load.
    float4 *prev = nullptr;
};

// contactGrid (bodyContact.js:125-127): counters, then slots per bucket.
// Optional (not in the reference's grid): `ovf`, the records that arrive at a
// full bucket, as (record, grid bucket) in arrival order, `ovf_count` of them
// this substep (appends past `ovf_cap` are dropped: sort then finds fewer
// members than the bucket counted and rescans); sort finds an overfull
// bucket's members there instead of rescanning the scene's records.
// Optional too: `ovb`, the overfull buckets as (grid bucket, scene), one entry
// per bucket, appended by the record whose arrival makes it overfull; sort's
// waves take them from it by wave index, so that a few waves do not each sort
// several crowded buckets in turn. With `ovb`, `ovf_count` holds three words:
// [0] records listed in `ovf`, [1] buckets listed in `ovb`, [2] set when
// `ovb` ran out of room (sort then takes every overfull bucket in the wave of
// its first record).
// Optional too: `ext`, `ext_slots` more slots a bucket past its `slots`: the
// record that arrives at position k (slots <= k < slots + ext_slots) of a
// bucket writes itself to ext[bucket * ext_slots + k - slots], and only later
// arrivals go to `ovf`; sort then finds most overfull buckets' members in
// their slots and extension alone.
struct GridView {
    uint32_t *count;
    uint32_t *slots;
    uint2 *ovf = nullptr;
    uint32_t *ovf_count = nullptr;
    uint32_t ovf_cap = 0;
    uint2 *ovb = nullptr;
    uint32_t ovb_cap = 0;
    uint32_t *ext = nullptr;
    uint32_t ext_slots = 0;
    // Per bucket, for collide: the slots it holds (at most the slot count, 6
    // bits) | the body all its records belong to << 6 (SUMMARY_MIXED when
    // several, or past 8 records). Sort writes it for every occupied bucket,
    // settle clears it, so an empty bucket reads 0.
    uint32_t *summary = nullptr;
};

constexpr uint32_t SUMMARY_MIXED = 0x3ffffffu;
__device__ __forceinline__ uint32_t summary_held(uint32_t w) { return w & 63u; }
__device__ __forceinline__ uint32_t summary_sole(uint32_t w) { return w >> 6; }

// Optional log of the pairs collide accepts (record, mate), `cap` per record.
struct PairLog {
    int32_t *count;
    uint32_t *mates;
    uint32_t cap;
};

struct DriveView {
    const uint32_t *node;
    const float *ox, *oy, *oz;
};

struct SplatView {
    const float *mx, *my, *mz;
    const float4 *quat;
    const float *sx, *sy, *sz;
    float *ox, *oy, *oz;
    float *cov;        // 6 planes of `plane` floats: xx xy xz yy yz zz
    float *rot;        // engine format: 4 planes (xyzw), then
    float *axis;       // 3 planes of axis lengths
    uint32_t plane;
    uint32_t mode;     // VS_SPLAT_COV | VS_SPLAT_AXES
};

// One body's lattice build on the device (lattice.hip).
struct LatticeView {
    const uint8_t *occ;          // cell_dims product bytes, x fastest
    uint32_t cell_dims[3];
    uint32_t node_offset;        // world id of the body's node 0
    uint32_t body;               // world body id of the surface records
    float radius;                // contact radius of the surface records
    double origin[3], h;         // rest = f32(origin + i * h), in f64
    float velocity[3];
    float wtab[9];               // inverse mass by occupied incident cells
    uint32_t tile_base[9];       // first tile of each colour; [8] = first node tile
    uint32_t *tile_count;        // count pass: kept elements per tile
    const uint32_t *tile_offset; // write pass: exclusive offset per tile
    uint32_t *cells;             // the body's range of the cell pool
    uint32_t *s_node, *s_body;   // the body's range of the surface planes
    float *s_radius;
};

struct Shape {
    uint32_t grid;     // workgroups
    uint32_t block;    // 64, 128, 256 or 512
    uint32_t xcd;      // 1: contiguous tile runs per XCD
};

// The nodes of a dense body that row class `cls` leaves out, numbered in
// node-id order. The class writes node rows y = ylo..yhi and z = zlo..zhi,
// ylo = cls & 1, yhi = ylo + 2 nyr - 1 with nyr = (cdy + 1 - ylo) / 2 its
// rows along y (likewise z); the rest is: the planes z < zlo (`lead` nodes),
// then in each plane zlo..zhi the rows y < ylo (`rlo` nodes) and y > yhi,
// then the planes z > zhi. With no rows (nyr or nzr 0) it is the whole body.
// The node tiles of a colour-pair launch cover index ranges of this list, so
// a tile is a full workgroup of leftover nodes however short their runs are.
struct Leftover {
    uint32_t nxy, lead, rlo, per, mid, total, zlo, skip, tail_at;

    __host__ __device__ Leftover(const uint32_t cell_dims[3], uint32_t cls)
    {
        const uint32_t nx = cell_dims[0] + 1u, ny = cell_dims[1] + 1u, nz = cell_dims[2] + 1u;
        const uint32_t py = cls & 1u, pz = cls >> 1;
        const uint32_t nyr = (cell_dims[1] + 1u - py) / 2u, nzr = (cell_dims[2] + 1u - pz) / 2u;
        nxy = nx * ny;
        if (nyr == 0u || nzr == 0u) {
            lead = total = nxy * nz;
            rlo = per = mid = zlo = skip = tail_at = 0u;
            return;
        }
        const uint32_t yhi = py + 2u * nyr - 1u, zhi = pz + 2u * nzr - 1u;
        zlo = pz;
        lead = pz * nxy;
        rlo = py * nx;
        per = rlo + (ny - 1u - yhi) * nx;
        mid = (zhi - pz + 1u) * per;
        skip = 2u * nyr * nx;  // a mid plane's rows y = ylo..yhi
        tail_at = (zhi + 1u) * nxy;
        total = lead + mid + (nz - 1u - zhi) * nxy;
    }

    // The body-relative node id of leftover node i < total.
    __host__ __device__ uint32_t node(uint32_t i) const
    {
        if (i < lead) return i;
        const uint32_t j = i - lead;
        if (j >= mid) return tail_at + (j - mid);
        const uint32_t plane = j / per, r = j - plane * per;
        return (zlo + plane) * nxy + r + (r >= rlo ? skip : 0u);
    }
};

// Host launchers; each enqueues one kernel on `stream`. They return the HIP
// launch status.
hipError_t launch_predict(hipStream_t s, Shape sh, NodeView n, const vs_scene_params *scenes, TileView t);
hipError_t launch_finalize(hipStream_t s, Shape sh, NodeView n, const vs_body *bodies,
                           const vs_scene_params *scenes, TileView t);
// One substep's finalize and the next one's predict in one node pass.
hipError_t launch_finalize_predict(hipStream_t s, Shape sh, NodeView n, const vs_body *bodies,
                                  const vs_scene_params *scenes, TileView t);
// `stretch` false: no body of the launch has maxStretch above 1, so the
// code of the tensile limit is left out (the elastic launches below too).
hipError_t launch_elastic(hipStream_t s, Shape sh, NodeView n, const vs_body *bodies,
                          const vs_scene_params *scenes, const uint32_t *cells, TileView t, uint32_t limit_only,
                          bool stretch = true);
// Colours 2*cls and 2*cls + 1 over whole x-rows of dense lattices; a tile is
// `count` rows of the class from row `begin` (row = a + ny*b for y = (cls&1)
// + 2a, z = (cls>>1) + 2b, ny = the class's rows along y). `node_mode` 1
// finalizes and predicts, 2 finalizes, the nodes written, in scenes flagged
// VS_SCENE_FOLD; a VS_NODE_TILE tile is `count` entries of the body's
// Leftover list from entry `begin`.
hipError_t launch_elastic_rows(hipStream_t s, Shape sh, NodeView n, const vs_body *bodies,
                               const vs_scene_params *scenes, TileView t, uint32_t cls, uint32_t limit_only,
                               uint32_t node_mode, bool stretch);
// The eight colours of whole small dense bodies (at most 11^3 nodes), one
// workgroup a body, in LDS.
// `node_mode` (scenes flagged VS_SCENE_FOLD_PREDICT): 3 predicts the nodes
// first, 1 finalizes them (the previous substep) and predicts them first.
// A tile is {first contact record, contact records, body, scene}. With
// `contact` bit 1, the body's records get the previous substep's settle
// (k_settle's arithmetic and counter clearing, grid `g`) after the load and
// before the node passes, in place of a settle launch; with bit 2 they are
// binned (k_bin's, grid `gn`) after the colours, in place of a bin launch.
// With both, `g` and `gn` must not share bucket counters, summaries or list
// counters.
hipError_t launch_elastic_bodies(hipStream_t s, Shape sh, NodeView n, const vs_body *bodies,
                                 const vs_scene_params *scenes, TileView t, uint32_t limit_only, uint32_t node_mode,
                                 bool stretch, SurfaceView sv, GridView g, GridView gn, uint32_t contact);
hipError_t launch_contact_bin(hipStream_t s, Shape sh, NodeView n, const vs_scene_params *scenes,
                              SurfaceView sv, GridView g, TileView t);
hipError_t launch_contact_sort(hipStream_t s, Shape sh, const vs_scene_params *scenes, SurfaceView sv,
                               GridView g, TileView t);
// With sv.pay and g.summary, nine lanes a record over tiles of at most
// `tile_max` records; without, one lane a record through the node ids (the
// reference's form).
hipError_t launch_contact_collide(hipStream_t s, Shape sh, NodeView n, const vs_body *bodies,
                                  const vs_scene_params *scenes, SurfaceView sv, GridView g, TileView t,
                                  PairLog log, uint32_t tile_max);
hipError_t launch_contact_settle(hipStream_t s, Shape sh, NodeView n, const vs_scene_params *scenes,
                                 SurfaceView sv, GridView g, TileView t);
hipError_t launch_drive(hipStream_t s, Shape sh, NodeView n, const vs_scene_params *scenes, DriveView d,
                        TileView t, int compliant);
// Lattice build: tile layout of one body (returns all tiles, colour tiles
// first), the count / write passes, the node planes, and mass zeroing.
uint32_t lattice_tiles(const uint32_t cell_dims[3], uint32_t tile_base[9], uint32_t *node_tiles);
hipError_t launch_lattice_lists(hipStream_t s, const LatticeView &L, uint32_t tiles, uint32_t write);
hipError_t launch_lattice_nodes(hipStream_t s, const LatticeView &L, NodeView n, float *rx, float *ry, float *rz);
hipError_t launch_zero_mass(hipStream_t s, float *w, const uint32_t *nodes, uint32_t count);
// Grab queries (grab.hip): one launch each; `out` as VS_B_OUT documents.
hipError_t launch_pick(hipStream_t s, uint32_t grid, const float *px, const float *py, const float *pz,
                       const float *mass, uint32_t count, const vs_grab_params *params, void *out);
hipError_t launch_grip(hipStream_t s, uint32_t grid, const float *px, const float *py, const float *pz,
                       const float *mass, uint32_t count, const vs_grab_params *params, void *candidates,
                       uint32_t cap, uint32_t limit, void *out);
// The grip with its candidates and counters in the library's static device
// array (vs_grip_noalloc): grip_static_scratch() returns the array's address
// on `device`, which must be current, looking it up once per device;
// launch_grip_static() clears the counters with a kernel and runs the grip.
hipError_t grip_static_scratch(int device, void **ptr);
hipError_t launch_grip_static(hipStream_t s, void *scratch_ptr, uint32_t grid, const float *px, const float *py,
                              const float *pz, const float *mass, uint32_t count, const vs_grab_params &params,
                              uint32_t limit, void *out);
hipError_t launch_splats(hipStream_t s, Shape sh, const float *px, const float *py, const float *pz,
                         const vs_body *bodies, SplatView sp, TileView t);

// Records the message vs_ingest_error() returns on this thread; returns code.
int ingest_fail(int code, const std::string &message);

}  // namespace vsim
