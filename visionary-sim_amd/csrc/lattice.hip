// lattice.hip -- buildLattice and surfaceNodes on the device
// (packages/visionary-sim/src/sim/discretize/lattice/lattice.js:128-239).
//
// Input: one body's occupancy grid, one byte per cell (x fastest, non-zero =
// occupied). Outputs, written straight into the world's pools:
//
// * node planes: rest = f32(origin + i*h) evaluated in f64 (as the JS host
//   does before storing into a Float32Array), position = start-of-substep =
//   rest, velocity, and the inverse mass of a node with k occupied incident
//   cells, wtab[k] (the host forms the 9-entry table with buildLattice's f64
//   corner mass accumulated into an f32 slot);
// * the occupied cells grouped by colour, each colour in z, y, x scan order;
// * the surface records: nodes with 1..7 of their 8 incident cells occupied,
//   ascending.
//
// The two lists are stream compactions. A colour's cells are a regular
// sub-grid (every other cell on each axis, starting at the colour's parities),
// so colour c's list is the compaction of that sub-grid in its own scan
// order. Each pass runs twice: `count` writes each 256-element tile's number
// of kept elements, the host turns those into exclusive offsets (it needs the
// totals anyway to size the pool ranges), and `write` places every kept
// element at its tile's offset plus its rank in the tile, from wave ballots.
#include "kernels.h"

namespace vsim {
namespace {

constexpr uint32_t TILE = 256;  // elements per tile = threads per workgroup

__device__ __forceinline__ uint32_t incident(const uint8_t *__restrict__ occ, const uint32_t cd[3], uint32_t i,
                                             uint32_t j, uint32_t k)
{
    // Cells around node (i, j, k) have their min corner one step back on each
    // axis; outside the grid counts as empty (lattice.js:140-151).
    uint32_t filled = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int ci = (int)i + (c & 1) - 1, cj = (int)j + ((c >> 1) & 1) - 1, ck = (int)k + ((c >> 2) & 1) - 1;
        const bool inside = ci >= 0 && cj >= 0 && ck >= 0 && ci < (int)cd[0] && cj < (int)cd[1] && ck < (int)cd[2];
        if (inside) filled += occ[(uint32_t)ci + cd[0] * ((uint32_t)cj + cd[1] * (uint32_t)ck)] != 0;
    }
    return filled;
}

// Rank of this thread's flag among the workgroup's set flags, and the total.
__device__ __forceinline__ uint32_t block_rank(bool flag, uint32_t *total)
{
    __shared__ uint32_t wave_count[TILE / 64];
    const uint64_t mask = __ballot(flag);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t below = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) wave_count[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (uint32_t w = 0; w < TILE / 64; w++) {
        before += w < wave ? wave_count[w] : 0u;
        all += wave_count[w];
    }
    *total = all;
    return before + below;
}

// Colour sub-grid of colour c: dims and the cell of element e.
struct SubGrid {
    uint32_t s[3], par[3];
};

__device__ __forceinline__ SubGrid subgrid(const uint32_t cd[3], uint32_t c)
{
    SubGrid g;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        g.par[a] = (c >> a) & 1u;
        g.s[a] = cd[a] > g.par[a] ? (cd[a] - g.par[a] + 1u) / 2u : 0u;
    }
    return g;
}

// Workgroups [tile_base[c], tile_base[c+1]) handle colour c; the node tiles
// follow the eight colours. write == 0: count; 1: place.
__global__ __launch_bounds__(TILE) void k_lattice_lists(LatticeView L, uint32_t write)
{
    const uint32_t b = blockIdx.x;
    const uint32_t cd[3] = { L.cell_dims[0], L.cell_dims[1], L.cell_dims[2] };
    if (b < L.tile_base[8]) {
        uint32_t c = 0;
        while (b >= L.tile_base[c + 1]) c++;
        const SubGrid g = subgrid(cd, c);
        const uint32_t e = (b - L.tile_base[c]) * TILE + threadIdx.x;
        const uint32_t n = g.s[0] * g.s[1] * g.s[2];
        uint32_t lin = 0;
        bool keep = false;
        if (e < n) {
            const uint32_t a0 = e % g.s[0], a1 = (e / g.s[0]) % g.s[1], a2 = e / (g.s[0] * g.s[1]);
            const uint32_t i = g.par[0] + 2u * a0, j = g.par[1] + 2u * a1, k = g.par[2] + 2u * a2;
            lin = i + cd[0] * (j + cd[1] * k);
            keep = L.occ[lin] != 0;
        }
        uint32_t total = 0;
        const uint32_t rank = block_rank(keep, &total);
        if (!write) {
            if (threadIdx.x == 0) L.tile_count[b] = total;
        } else if (keep) {
            L.cells[L.tile_offset[b] + rank] = lin;
        }
        return;
    }
    const uint32_t nd[3] = { cd[0] + 1u, cd[1] + 1u, cd[2] + 1u };
    const uint32_t e = (b - L.tile_base[8]) * TILE + threadIdx.x;
    bool keep = false;
    if (e < nd[0] * nd[1] * nd[2]) {
        const uint32_t k = incident(L.occ, cd, e % nd[0], (e / nd[0]) % nd[1], e / (nd[0] * nd[1]));
        keep = k > 0u && k < 8u;
    }
    uint32_t total = 0;
    const uint32_t rank = block_rank(keep, &total);
    if (!write) {
        if (threadIdx.x == 0) L.tile_count[b] = total;
    } else if (keep) {
        const uint32_t at = L.tile_offset[b] + rank;
        L.s_node[at] = L.node_offset + e;
        L.s_body[at] = L.body;
        L.s_radius[at] = L.radius;
    }
}

// Node planes of the body: rest, position, start-of-substep position,
// velocity, inverse mass.
__global__ __launch_bounds__(TILE) void k_lattice_nodes(LatticeView L, NodeView n, float *rx, float *ry, float *rz)
{
    const uint32_t cd[3] = { L.cell_dims[0], L.cell_dims[1], L.cell_dims[2] };
    const uint32_t nd[3] = { cd[0] + 1u, cd[1] + 1u, cd[2] + 1u };
    const uint32_t count = nd[0] * nd[1] * nd[2];
    for (uint32_t e = blockIdx.x * TILE + threadIdx.x; e < count; e += gridDim.x * TILE) {
        const uint32_t i = e % nd[0], j = (e / nd[0]) % nd[1], k = e / (nd[0] * nd[1]);
        const float x = (float)(L.origin[0] + (double)i * L.h);
        const float y = (float)(L.origin[1] + (double)j * L.h);
        const float z = (float)(L.origin[2] + (double)k * L.h);
        const uint32_t g = L.node_offset + e;
        n.px[g] = x;
        n.py[g] = y;
        n.pz[g] = z;
        n.qx[g] = x;
        n.qy[g] = y;
        n.qz[g] = z;
        rx[g] = x;
        ry[g] = y;
        rz[g] = z;
        n.vx[g] = L.velocity[0];
        n.vy[g] = L.velocity[1];
        n.vz[g] = L.velocity[2];
        const_cast<float *>(n.w)[g] = L.wtab[incident(L.occ, cd, i, j, k)];
    }
}

// Prescribed nodes (static regions, hard-held drive nodes) lose their mass.
__global__ __launch_bounds__(TILE) void k_zero_mass(float *w, const uint32_t *__restrict__ nodes, uint32_t count)
{
    const uint32_t e = blockIdx.x * TILE + threadIdx.x;
    if (e < count) w[nodes[e]] = 0.0f;
}

}  // namespace

uint32_t lattice_tiles(const uint32_t cell_dims[3], uint32_t tile_base[9], uint32_t *node_tiles)
{
    uint32_t at = 0;
    for (uint32_t c = 0; c < 8; c++) {
        tile_base[c] = at;
        uint32_t n = 1;
        for (int a = 0; a < 3; a++) {
            const uint32_t par = (c >> a) & 1u;
            n *= cell_dims[a] > par ? (cell_dims[a] - par + 1u) / 2u : 0u;
        }
        at += (n + TILE - 1) / TILE;
    }
    tile_base[8] = at;
    const uint64_t nodes = (uint64_t)(cell_dims[0] + 1) * (cell_dims[1] + 1) * (cell_dims[2] + 1);
    *node_tiles = (uint32_t)((nodes + TILE - 1) / TILE);
    return at + *node_tiles;
}

hipError_t launch_lattice_lists(hipStream_t s, const LatticeView &L, uint32_t tiles, uint32_t write)
{
    if (tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(k_lattice_lists, dim3(tiles), dim3(TILE), 0, s, L, write);
    return hipGetLastError();
}

hipError_t launch_lattice_nodes(hipStream_t s, const LatticeView &L, NodeView n, float *rx, float *ry, float *rz)
{
    const uint64_t nodes = (uint64_t)(L.cell_dims[0] + 1) * (L.cell_dims[1] + 1) * (L.cell_dims[2] + 1);
    const uint32_t grid = (uint32_t)std::min<uint64_t>((nodes + TILE - 1) / TILE, 4096);
    hipLaunchKernelGGL(k_lattice_nodes, dim3(grid), dim3(TILE), 0, s, L, n, rx, ry, rz);
    return hipGetLastError();
}

hipError_t launch_zero_mass(hipStream_t s, float *w, const uint32_t *nodes, uint32_t count)
{
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_mass, dim3((count + TILE - 1) / TILE), dim3(TILE), 0, s, w, nodes, count);
    return hipGetLastError();
}

}  // namespace vsim
