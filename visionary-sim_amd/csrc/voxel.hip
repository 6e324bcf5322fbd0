// voxel.hip -- the ingest side of a capture on the device: occupancy grids
// from a cloud of Gaussians (voxelizeCloud, packages/visionary-sim/src/sim/
// discretize/lattice/cloudVolume.js), a triangle mesh (voxelizeMesh,
// voxelize.js:176-364) or an analytic shape (voxelizeSdf, voxelize.js:120-148,
// shapes from apps/simulator/src/assets/shapes.js), ready for the device
// lattice build.
//
// Arithmetic is the JS's, in doubles: f32 inputs widened, expressions
// evaluated left to right as the reference writes them (the build has no FMA
// contraction), rotation columns rounded to f32 as mat3FromQuat stores them,
// Math.hypot restated as the square root of the sum of squares. Every pass is
// order-independent (a byte set to 1, a min/max, a sort), so the grid is the
// same whatever order the lanes run in.
//
// Passes: a grid-stride stamp per Gaussian; 6-neighbourhood dilate/erode; a
// flood of the empty space from the faces by alternating line sweeps along
// the three axes in both directions until nothing changes; for meshes, per
// axis a count / scatter of ray crossings into per-column segments and a
// per-column sort and parity walk, then the edge walk and the 2-of-3 vote.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.h"

namespace vsim {
namespace {

struct Grid {
    uint32_t dims[3];
    double origin[3];
    double h;
};

__host__ __device__ __forceinline__ size_t cell_at(const Grid &g, int64_t i, int64_t j, int64_t k)
{
    return (size_t)i + (size_t)g.dims[0] * ((size_t)j + (size_t)g.dims[1] * (size_t)k);
}

__host__ __device__ __forceinline__ size_t cells_of(const Grid &g)
{
    return (size_t)g.dims[0] * g.dims[1] * g.dims[2];
}

// Math.min / Math.max: NaN in, NaN out.
__host__ __device__ __forceinline__ double js_min(double a, double b)
{
    return a != a || b != b ? NAN : a < b ? a : b;
}
__host__ __device__ __forceinline__ double js_max(double a, double b)
{
    return a != a || b != b ? NAN : a > b ? a : b;
}

// The largest grid built: 2^24 cells on an axis, 2^36 in all. The reference
// has no such limit and fails inside its typed-array allocation instead.
constexpr double GRID_AXIS_MAX = 16777216.0;
constexpr double GRID_CELLS_MAX = 68719476736.0;

// gridForBounds (voxelize.js:39-67). VS_OK, or VS_E_INVALID with the message.
int grid_for_bounds(const double mn[3], const double mx[3], double h, uint32_t padding, Grid &g);

__device__ __forceinline__ double hypot3(double a, double b, double c) { return sqrt(a * a + b * b + c * c); }
__device__ __forceinline__ double hypot2(double a, double b) { return sqrt(a * a + b * b); }

__device__ __forceinline__ int64_t clampi(double v, int64_t lo, int64_t hi)
{
    // Math.min(Math.max(v, lo), hi) for a finite v: hi wins when lo > hi.
    const double t = v < (double)lo ? (double)lo : v;
    return t > (double)hi ? hi : (int64_t)t;
}

// mat3FromQuat (mat3.js:71-86), stored in f32 like the reference's Float32Array.
__device__ void rotation_columns(const float *q, float out[9])
{
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    out[0] = (float)(1 - 2 * (y * y + z * z));
    out[1] = (float)(2 * (x * y + z * w));
    out[2] = (float)(2 * (x * z - y * w));
    out[3] = (float)(2 * (x * y - z * w));
    out[4] = (float)(1 - 2 * (x * x + z * z));
    out[5] = (float)(2 * (y * z + x * w));
    out[6] = (float)(2 * (x * z + y * w));
    out[7] = (float)(2 * (y * z - x * w));
    out[8] = (float)(1 - 2 * (x * x + y * y));
}

// ellipsoidOf (cloudVolume.js:182-210).
__device__ void ellipsoid_of(const float *rot, const float *scale, double sigma, float col[9], double radius[3],
                             double half[3])
{
    for (int a = 0; a < 3; a++) radius[a] = sigma * scale[a];
    rotation_columns(rot, col);
    for (int a = 0; a < 3; a++)
        half[a] = hypot3((double)col[a] * radius[0], (double)col[3 + a] * radius[1], (double)col[6 + a] * radius[2]);
}

constexpr uint32_t BLOCK = 256;

// materialBounds (cloudVolume.js:212-234): per-workgroup min / max, six
// doubles each; the host folds them (min and max do not depend on order).
__global__ __launch_bounds__(BLOCK) void k_cloud_bounds(const float *pos, const float *rot, const float *scale,
                                                        uint64_t count, double sigma, double *partial)
{
    __shared__ double s[6][BLOCK];
    double mn[3] = { INFINITY, INFINITY, INFINITY }, mx[3] = { -INFINITY, -INFINITY, -INFINITY };
    for (uint64_t i = blockIdx.x * (uint64_t)BLOCK + threadIdx.x; i < count; i += (uint64_t)gridDim.x * BLOCK) {
        float col[9];
        double radius[3], half[3];
        ellipsoid_of(rot + 4 * i, scale + 3 * i, sigma, col, radius, half);
        for (int a = 0; a < 3; a++) {
            const double lo = (double)pos[3 * i + a] - half[a], hi = (double)pos[3 * i + a] + half[a];
            mn[a] = js_min(mn[a], lo);
            mx[a] = js_max(mx[a], hi);
        }
    }
    for (int a = 0; a < 3; a++) {
        s[a][threadIdx.x] = mn[a];
        s[3 + a][threadIdx.x] = mx[a];
    }
    __syncthreads();
    for (uint32_t stride = BLOCK / 2; stride > 0; stride >>= 1) {
        if (threadIdx.x < stride) {
            for (int a = 0; a < 3; a++) {
                s[a][threadIdx.x] = js_min(s[a][threadIdx.x], s[a][threadIdx.x + stride]);
                s[3 + a][threadIdx.x] = js_max(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + stride]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// stampCloud (cloudVolume.js:138-180): the centre's cell unconditionally,
// then every cell of the ellipsoid's box (kept off the faces) one of whose
// nine sample points (corners, centre) lies inside. The result is a union,
// so the order the cells are visited in does not matter. A Gaussian's box
// can be one cell or the whole grid (a large floater in a capture), so the
// boxes are cut into units of UNIT cells and handed out like the mesh work.
constexpr uint64_t UNIT = 64;

struct Stamp {
    double c[3], radius[3];
    float col[9];
    int64_t lo[3], size[3];  // the box: lo + [0, size) on each axis
    uint64_t cells;
};

__device__ Stamp stamp_of(const float *pos, const float *rot, const float *scale, uint64_t i, double sigma,
                          const Grid &g)
{
    Stamp st;
    double half[3];
    for (int a = 0; a < 3; a++) st.c[a] = pos[3 * i + a];
    ellipsoid_of(rot + 4 * i, scale + 3 * i, sigma, st.col, st.radius, half);
    st.cells = 1;
    for (int a = 0; a < 3; a++) {
        st.lo[a] = clampi(floor((st.c[a] - half[a] - g.origin[a]) / g.h), 1, (int64_t)g.dims[a] - 2);
        const int64_t hi = clampi(floor((st.c[a] + half[a] - g.origin[a]) / g.h), 1, (int64_t)g.dims[a] - 2);
        st.size[a] = hi >= st.lo[a] ? hi - st.lo[a] + 1 : 0;
        st.cells *= (uint64_t)st.size[a];
    }
    return st;
}

// Units of Gaussian i: one per UNIT cells of its box and at least one, the
// first also stamping the centre. (A direct one-lane-per-Gaussian pass for
// the small boxes measured slower: a capture's scales spread wide enough
// that its lanes diverge.)
__device__ __forceinline__ uint64_t stamp_units(const float *pos, const float *rot, const float *scale, uint64_t i,
                                                double sigma, const Grid &g)
{
    const uint64_t cells = stamp_of(pos, rot, scale, i, sigma, g).cells;
    return cells > UNIT ? (cells + UNIT - 1) / UNIT : 1;
}

// The last item whose first unit is <= unit: the owner of `unit`.
__device__ __forceinline__ uint64_t owner(const uint64_t *start, uint64_t items, uint64_t unit)
{
    uint64_t lo = 0, hi = items;  // start[lo] <= unit < start[hi] (start[items] = total)
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (start[mid] <= unit) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Tests cells [first, last) of a Gaussian's box (x fastest), and stamps the
// centre's cell when `first` is 0.
__device__ void stamp_cells(const Stamp &st, const Grid &g, uint64_t first, uint64_t last, uint8_t *occ)
{
    const uint64_t n = cells_of(g);
    if (first == 0) {
        int64_t cc[3];
        for (int a = 0; a < 3; a++) cc[a] = clampi(floor((st.c[a] - g.origin[a]) / g.h), 0, (int64_t)g.dims[a] - 1);
        occ[cell_at(g, cc[0], cc[1], cc[2])] = 1;
    }
    for (uint64_t b = first; b < last; b++) {
        const int64_t ii = st.lo[0] + (int64_t)(b % (uint64_t)st.size[0]);
        const int64_t j = st.lo[1] + (int64_t)((b / (uint64_t)st.size[0]) % (uint64_t)st.size[1]);
        const int64_t k = st.lo[2] + (int64_t)(b / ((uint64_t)st.size[0] * (uint64_t)st.size[1]));
        // On a grid under three cells thick the box can reach index -1;
        // like a typed array, an index outside the grid is skipped.
        const int64_t lin = ii + (int64_t)g.dims[0] * (j + (int64_t)g.dims[1] * k);
        if (lin < 0 || (uint64_t)lin >= n) continue;
        const size_t cell = (size_t)lin;
        if (occ[cell]) continue;
        for (int s = 0; s < 9; s++) {
            const double ox = s == 8 ? 0.5 : (double)(s & 1), oy = s == 8 ? 0.5 : (double)((s >> 1) & 1),
                         oz = s == 8 ? 0.5 : (double)((s >> 2) & 1);
            const double dx = g.origin[0] + ((double)ii + ox) * g.h - st.c[0];
            const double dy = g.origin[1] + ((double)j + oy) * g.h - st.c[1];
            const double dz = g.origin[2] + ((double)k + oz) * g.h - st.c[2];
            double sum = 0;
            for (int a = 0; a < 3; a++) {
                const double t = ((double)st.col[a * 3] * dx + (double)st.col[a * 3 + 1] * dy +
                                  (double)st.col[a * 3 + 2] * dz) / st.radius[a];
                sum += t * t;
            }
            if (sum <= 1) {
                occ[cell] = 1;
                break;
            }
        }
    }
}

// UNIT cells of one Gaussian's box per lane (the owner found by binary
// search over the Gaussians' first units).
__global__ __launch_bounds__(BLOCK) void k_cloud_stamp(const float *pos, const float *rot, const float *scale,
                                                             uint64_t count, double sigma, Grid g,
                                                             const uint64_t *start, uint64_t units, uint8_t *occ)
{
    for (uint64_t unit = blockIdx.x * (uint64_t)BLOCK + threadIdx.x; unit < units;
         unit += (uint64_t)gridDim.x * BLOCK) {
        const uint64_t i = owner(start, count, unit);
        const Stamp st = stamp_of(pos, rot, scale, i, sigma, g);
        const uint64_t first = (unit - start[i]) * UNIT;
        stamp_cells(st, g, first, first + UNIT < st.cells ? first + UNIT : st.cells, occ);
    }
}

// Dilate (1) or erode (0) with the 6-neighbourhood, outside counting as
// empty (voxelize.js:396-435).
__global__ __launch_bounds__(BLOCK) void k_neighbourhood(Grid g, const uint8_t *src, uint8_t *dst, int dilate)
{
    const size_t n = cells_of(g);
    const int64_t nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
    for (size_t c = blockIdx.x * (size_t)BLOCK + threadIdx.x; c < n; c += (size_t)gridDim.x * BLOCK) {
        const int64_t i = (int64_t)(c % nx), j = (int64_t)((c / nx) % ny), k = (int64_t)(c / ((size_t)nx * ny));
        bool result = src[c] == 1;
        for (int axis = 0; axis < 3 && result != (dilate != 0); axis++)
            for (int step = -1; step <= 1; step += 2) {
                const int64_t ni = i + (axis == 0 ? step : 0), nj = j + (axis == 1 ? step : 0),
                              nk = k + (axis == 2 ? step : 0);
                const bool outside = ni < 0 || nj < 0 || nk < 0 || ni >= nx || nj >= ny || nk >= nz;
                const uint8_t nb = outside ? 0 : src[cell_at(g, ni, nj, nk)];
                if (dilate ? nb == 1 : nb == 0) {
                    result = dilate != 0;
                    break;
                }
            }
        dst[c] = result ? 1 : 0;
    }
}

__global__ __launch_bounds__(BLOCK) void k_or(uint8_t *dst, const uint8_t *src, size_t n)
{
    for (size_t c = blockIdx.x * (size_t)BLOCK + threadIdx.x; c < n; c += (size_t)gridDim.x * BLOCK) dst[c] |= src[c];
}

// fillEnclosed (cloudVolume.js:322-373): the empty cells on the faces start
// reached; sweeps carry "reached" along each line of empty cells; the cells
// never reached become material.
__global__ __launch_bounds__(BLOCK) void k_flood_seed(Grid g, const uint8_t *occ, uint8_t *reached)
{
    const size_t n = cells_of(g);
    const uint32_t nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
    for (size_t c = blockIdx.x * (size_t)BLOCK + threadIdx.x; c < n; c += (size_t)gridDim.x * BLOCK) {
        const uint32_t i = (uint32_t)(c % nx), j = (uint32_t)((c / nx) % ny), k = (uint32_t)(c / ((size_t)nx * ny));
        const bool face = i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1;
        reached[c] = face && occ[c] == 0 ? 1 : 0;
    }
}

__global__ __launch_bounds__(BLOCK) void k_flood_sweep(Grid g, const uint8_t *occ, uint8_t *reached, int axis,
                                                       int forward, uint32_t *changed)
{
    // Neighbouring lanes take neighbouring lines along the lower of the other
    // two axes (x unless the sweep runs along x), so their bytes share cache
    // lines.
    const int u = axis == 0 ? 1 : 0, v = axis == 2 ? 1 : 2;
    const size_t lines = (size_t)g.dims[u] * g.dims[v];
    for (size_t l = blockIdx.x * (size_t)BLOCK + threadIdx.x; l < lines; l += (size_t)gridDim.x * BLOCK) {
        int64_t ijk[3];
        ijk[u] = (int64_t)(l % g.dims[u]);
        ijk[v] = (int64_t)(l / g.dims[u]);
        const int64_t len = g.dims[axis];
        bool carry = false, any = false;
        for (int64_t s = 0; s < len; s++) {
            ijk[axis] = forward ? s : len - 1 - s;
            const size_t c = cell_at(g, ijk[0], ijk[1], ijk[2]);
            if (occ[c]) {
                carry = false;
                continue;
            }
            if (reached[c]) {
                carry = true;
            } else if (carry) {
                reached[c] = 1;
                any = true;
            }
        }
        if (any) atomicOr(changed, 1u);
    }
}

__global__ __launch_bounds__(BLOCK) void k_flood_finish(uint8_t *occ, const uint8_t *reached, size_t n)
{
    for (size_t c = blockIdx.x * (size_t)BLOCK + threadIdx.x; c < n; c += (size_t)gridDim.x * BLOCK)
        if (!occ[c] && !reached[c]) occ[c] = 1;
}

// ---------------------------------------------------------------- meshes
//
// voxelizeMesh casts one ray per cell column along each axis through every
// triangle's projected box of columns, and walks every edge in steps of 0.4
// cells. A triangle's box can be a handful of columns or the whole grid, so
// the work is cut into units of at most UNIT ray tests (or edge samples): a
// scan gives each triangle (edge) its first unit, and every lane takes units
// and finds their owner by binary search.

// A triangle seen along `axis` (voxelize.js:283-306): vertices in (u, v, w)
// order, 1/area2, and its box of columns [a0, a0 + width) x [b0, ...).
struct Tri {
    double au, av, aw, bu, bv, bw, cu, cv, cw, inv;
    int64_t a0, b0;
    uint64_t width, area;  // area 0: degenerate or no column
};

__device__ Tri triangle(const float *pos, const uint32_t *idx, uint64_t t, const Grid &g, int axis)
{
    const int u = (axis + 1) % 3, v = (axis + 2) % 3, w = axis;
    const float *A = pos + 3 * (size_t)idx[3 * t], *B = pos + 3 * (size_t)idx[3 * t + 1],
                *C = pos + 3 * (size_t)idx[3 * t + 2];
    Tri r;
    r.au = A[u], r.av = A[v], r.aw = A[w], r.bu = B[u], r.bv = B[v], r.bw = B[w];
    r.cu = C[u], r.cv = C[v], r.cw = C[w];
    r.a0 = r.b0 = 0;
    r.width = r.area = 0;
    const double area2 = (r.bv - r.cv) * (r.au - r.cu) + (r.cu - r.bu) * (r.av - r.cv);
    if (fabs(area2) < 1e-20) return r;  // degenerate when seen along this axis
    r.inv = 1 / area2;
    // The vertices are finite (checked on the host), so no NaN reaches the
    // Math.min / Math.max restated here.
    const double min_u = fmin(r.au, fmin(r.bu, r.cu)), max_u = fmax(r.au, fmax(r.bu, r.cu));
    const double min_v = fmin(r.av, fmin(r.bv, r.cv)), max_v = fmax(r.av, fmax(r.bv, r.cv));
    double a0 = ceil((min_u - g.origin[u]) / g.h - 0.5), a1 = floor((max_u - g.origin[u]) / g.h - 0.5);
    double b0 = ceil((min_v - g.origin[v]) / g.h - 0.5), b1 = floor((max_v - g.origin[v]) / g.h - 0.5);
    a0 = a0 > 0 ? a0 : 0;
    b0 = b0 > 0 ? b0 : 0;
    a1 = a1 < g.dims[u] - 1.0 ? a1 : g.dims[u] - 1.0;
    b1 = b1 < g.dims[v] - 1.0 ? b1 : g.dims[v] - 1.0;
    if (!(a0 <= a1 && b0 <= b1)) return r;  // no column: the loops run zero times
    r.a0 = (int64_t)a0, r.b0 = (int64_t)b0;
    r.width = (uint64_t)(a1 - a0) + 1;
    r.area = r.width * ((uint64_t)(b1 - b0) + 1);
    return r;
}

// Samples of edge e = 3 t + k, from vertex k to k + 1 (voxelize.js:233-245).
struct Edge {
    double ax, ay, az, dx, dy, dz, steps;
};

__device__ Edge edge(const float *pos, const uint32_t *idx, uint64_t e, const Grid &g)
{
    const uint64_t t = e / 3;
    const int k = (int)(e % 3);
    const float *A = pos + 3 * (size_t)idx[3 * t + k], *B = pos + 3 * (size_t)idx[3 * t + (k + 1) % 3];
    Edge r;
    r.ax = A[0], r.ay = A[1], r.az = A[2];
    r.dx = (double)B[0] - A[0], r.dy = (double)B[1] - A[1], r.dz = (double)B[2] - A[2];
    const double steps = ceil(hypot3(r.dx, r.dy, r.dz) / (g.h * 0.4));
    r.steps = steps > 1 ? steps : 1;
    return r;
}

// Work units of item e: triangles along `mode` (0..2) or edges (mode 3),
// capped one past the most a mesh may need so tile sums cannot overflow (the
// host refuses any total past that limit).
constexpr uint64_t ITEM_UNITS_CAP = (1ull << 30) + 1;

__device__ __forceinline__ uint64_t units_of(const float *pos, const uint32_t *idx, uint64_t e, const Grid &g, int mode)
{
    double units;
    if (mode < 3) {
        units = (double)((triangle(pos, idx, e, g, mode).area + UNIT - 1) / UNIT);
    } else {
        units = ceil((edge(pos, idx, e, g).steps + 1) / (double)UNIT);  // steps + 1 samples
    }
    return units < (double)ITEM_UNITS_CAP ? (uint64_t)units : ITEM_UNITS_CAP;
}

// What a units pass counts: mesh triangles along an axis or edges
// (mode 0..3, pos + idx), or the Gaussians of a cloud (mode 4: pos, rot,
// scale, sigma).
struct Items {
    const float *pos, *rot, *scale;
    const uint32_t *idx;
    double sigma;
    int mode;
};

// Per tile of BLOCK items: each item's first unit relative to its tile, and
// the tile's total (the host scans the totals).
__global__ __launch_bounds__(BLOCK) void k_units(Items it, uint64_t items, Grid g, uint64_t *start,
                                                 uint64_t *tile_total)
{
    __shared__ uint64_t s[BLOCK];
    const uint64_t e = blockIdx.x * (uint64_t)BLOCK + threadIdx.x;
    uint64_t mine = 0;
    if (e < items) {
        mine = it.mode < 4 ? units_of(it.pos, it.idx, e, g, it.mode)
                           : stamp_units(it.pos, it.rot, it.scale, e, it.sigma, g);
        mine = mine < ITEM_UNITS_CAP ? mine : ITEM_UNITS_CAP;
    }
    s[threadIdx.x] = mine;
    __syncthreads();
    // Hillis-Steele inclusive scan over the tile.
    for (uint32_t d = 1; d < BLOCK; d <<= 1) {
        const uint64_t add = threadIdx.x >= d ? s[threadIdx.x - d] : 0;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    if (e < items) start[e] = s[threadIdx.x] - mine;
    if (threadIdx.x == BLOCK - 1) tile_total[blockIdx.x] = s[BLOCK - 1];
}

__global__ __launch_bounds__(BLOCK) void k_add_offsets(uint64_t *start, uint64_t items, const uint64_t *tile_offset)
{
    for (uint64_t e = blockIdx.x * (uint64_t)BLOCK + threadIdx.x; e < items; e += (uint64_t)gridDim.x * BLOCK)
        start[e] += tile_offset[e / BLOCK];
}

// Where column (cu2, cv2) of `axis` keeps its crossings. Only memory layout:
// x runs fastest where x is one of the column's axes (along y the columns
// are (z, x)), so the lanes that walk neighbouring columns share cache lines.
__device__ __forceinline__ size_t column_of(const Grid &g, int axis, int64_t cu2, int64_t cv2)
{
    const int u = (axis + 1) % 3, v = (axis + 2) % 3;
    return axis == 1 ? (size_t)cv2 + (size_t)g.dims[v] * (size_t)cu2 : (size_t)cu2 + (size_t)g.dims[u] * (size_t)cv2;
}

// The ray tests of accumulateAxis (voxelize.js:308-326). offset == nullptr:
// count crossings per column; else write each into its column's segment
// (the column's offset plus a slot from the fill counter).
__global__ __launch_bounds__(BLOCK) void k_mesh_crossings(const float *pos, const uint32_t *idx, uint64_t tris, Grid g,
                                                          int axis, const uint64_t *start, uint64_t units,
                                                          uint32_t *count, const uint64_t *offset, double *hits)
{
    const int u = (axis + 1) % 3, v = (axis + 2) % 3;
    const double jitter_u = g.h * 1.7e-4, jitter_v = g.h * 0.9e-4;
    for (uint64_t unit = blockIdx.x * (uint64_t)BLOCK + threadIdx.x; unit < units;
         unit += (uint64_t)gridDim.x * BLOCK) {
        const uint64_t t = owner(start, tris, unit);
        const Tri r = triangle(pos, idx, t, g, axis);
        const uint64_t first = (unit - start[t]) * UNIT;
        const uint64_t last = first + UNIT < r.area ? first + UNIT : r.area;
        for (uint64_t c = first; c < last; c++) {
            const int64_t cu2 = r.a0 + (int64_t)(c % r.width), cv2 = r.b0 + (int64_t)(c / r.width);
            const double pv = g.origin[v] + ((double)cv2 + 0.5) * g.h + jitter_v;
            const double pu = g.origin[u] + ((double)cu2 + 0.5) * g.h + jitter_u;
            const double l0 = ((r.bv - r.cv) * (pu - r.cu) + (r.cu - r.bu) * (pv - r.cv)) * r.inv;
            if (l0 < 0) continue;
            const double l1 = ((r.cv - r.av) * (pu - r.cu) + (r.au - r.cu) * (pv - r.cv)) * r.inv;
            if (l1 < 0) continue;
            const double l2 = 1 - l0 - l1;
            if (l2 < 0) continue;
            const size_t col = column_of(g, axis, cu2, cv2);
            const uint32_t slot = atomicAdd(&count[col], 1u);
            if (offset) hits[offset[col] + slot] = l0 * r.aw + l1 * r.bw + l2 * r.cw;
        }
    }
}

// Sorts a column's crossings ascending: insertion for short columns, heap
// sort past that. Equal doubles are interchangeable here (+0 and -0 floor
// and compare alike), so stability does not matter.
__device__ void sort_hits(double *h, uint32_t n)
{
    if (n <= 32) {
        for (uint32_t a = 1; a < n; a++) {
            const double x = h[a];
            uint32_t b = a;
            while (b > 0 && h[b - 1] > x) {
                h[b] = h[b - 1];
                b--;
            }
            h[b] = x;
        }
        return;
    }
    auto sift = [&](uint32_t root, uint32_t end) {
        while (2 * root + 1 < end) {
            uint32_t child = 2 * root + 1;
            if (child + 1 < end && h[child] < h[child + 1]) child++;
            if (!(h[root] < h[child])) return;
            const double tmp = h[root];
            h[root] = h[child];
            h[child] = tmp;
            root = child;
        }
    };
    for (uint32_t i = n / 2; i-- > 0;) sift(i, n);
    for (uint32_t end = n - 1; end > 0; end--) {
        const double tmp = h[0];
        h[0] = h[end];
        h[end] = tmp;
        sift(0, end);
    }
}

// Per column: sort its crossings, mark the cells they land in as surface,
// and walk the column flipping parity at each crossing behind the centre;
// interior cells get a vote (voxelize.js:330-361).
__global__ __launch_bounds__(BLOCK) void k_mesh_columns(Grid g, int axis, const uint32_t *count, const uint64_t *offset,
                                                        double *hits, uint8_t *votes, uint8_t *surface)
{
    const int u = (axis + 1) % 3, v = (axis + 2) % 3, w = axis;
    const size_t columns = (size_t)g.dims[u] * g.dims[v];
    for (size_t col = blockIdx.x * (size_t)BLOCK + threadIdx.x; col < columns; col += (size_t)gridDim.x * BLOCK) {
        const uint32_t n = count[col];
        if (n == 0) continue;
        double *h = hits + offset[col];
        sort_hits(h, n);
        int64_t ijk[3];
        // column_of, inverted
        if (axis == 1) {
            ijk[v] = (int64_t)(col % g.dims[v]);
            ijk[u] = (int64_t)(col / g.dims[v]);
        } else {
            ijk[u] = (int64_t)(col % g.dims[u]);
            ijk[v] = (int64_t)(col / g.dims[u]);
        }
        for (uint32_t e = 0; e < n; e++) {
            const double cw2 = floor((h[e] - g.origin[w]) / g.h);
            if (cw2 >= 0 && cw2 < g.dims[w]) {
                ijk[w] = (int64_t)cw2;
                surface[cell_at(g, ijk[0], ijk[1], ijk[2])] = 1;
            }
        }
        uint32_t next = 0;
        bool inside = false;
        for (int64_t cw2 = 0; cw2 < g.dims[w]; cw2++) {
            const double centre = g.origin[w] + ((double)cw2 + 0.5) * g.h;
            while (next < n && h[next] <= centre) {
                inside = !inside;
                next++;
            }
            if (inside) {
                ijk[w] = cw2;
                votes[cell_at(g, ijk[0], ijk[1], ijk[2])]++;
            }
        }
    }
}

// markEdgeCells (voxelize.js:212-248), UNIT samples per lane.
__global__ __launch_bounds__(BLOCK) void k_mesh_edges(const float *pos, const uint32_t *idx, uint64_t edges, Grid g,
                                                      const uint64_t *start, uint64_t units, uint8_t *surface)
{
    for (uint64_t unit = blockIdx.x * (uint64_t)BLOCK + threadIdx.x; unit < units;
         unit += (uint64_t)gridDim.x * BLOCK) {
        const uint64_t e = owner(start, edges, unit);
        const Edge r = edge(pos, idx, e, g);
        const double first = (double)((unit - start[e]) * UNIT);
        for (double s = first; s < first + (double)UNIT && s <= r.steps; s++) {
            const double f = s / r.steps;
            const double x = r.ax + r.dx * f, y = r.ay + r.dy * f, z = r.az + r.dz * f;
            const double i = floor((x - g.origin[0]) / g.h), j = floor((y - g.origin[1]) / g.h),
                         k = floor((z - g.origin[2]) / g.h);
            if (i >= 0 && j >= 0 && k >= 0 && i < g.dims[0] && j < g.dims[1] && k < g.dims[2])
                surface[cell_at(g, (int64_t)i, (int64_t)j, (int64_t)k)] = 1;
        }
    }
}

__global__ __launch_bounds__(BLOCK) void k_mesh_vote(const uint8_t *votes, const uint8_t *surface, uint8_t *occ, size_t n)
{
    for (size_t c = blockIdx.x * (size_t)BLOCK + threadIdx.x; c < n; c += (size_t)gridDim.x * BLOCK)
        occ[c] = votes[c] >= 2 || surface[c] == 1 ? 1 : 0;
}

// Signed distances of the analytic shapes (shapes.js:32-108), translated.
__device__ double shape_sdf(const vs_shape &s, double x, double y, double z)
{
    for (uint32_t t = s.n_offsets; t-- > 0;) {
        x -= s.offset[t][0];
        y -= s.offset[t][1];
        z -= s.offset[t][2];
    }
    if (s.kind == VS_SHAPE_SPHERE) return hypot3(x, y, z) - s.size[0];
    if (s.kind == VS_SHAPE_BOX) {
        const double dx = fabs(x) - s.size[0], dy = fabs(y) - s.size[1], dz = fabs(z) - s.size[2];
        const double outside = hypot3(dx > 0 ? dx : 0, dy > 0 ? dy : 0, dz > 0 ? dz : 0);
        double m = dx > dy ? dx : dy;
        m = m > dz ? m : dz;
        const double inside = m < 0 ? m : 0;
        return outside + inside;
    }
    const double radial = hypot2(x, z) - s.size[0];
    return hypot2(radial, y) - s.size[1];
}

__global__ __launch_bounds__(BLOCK) void k_shape(vs_shape s, Grid g, uint8_t *occ)
{
    const size_t n = cells_of(g);
    const uint32_t nx = g.dims[0], ny = g.dims[1];
    const double h = g.h, threshold = h * 0.5;
    for (size_t c = blockIdx.x * (size_t)BLOCK + threadIdx.x; c < n; c += (size_t)gridDim.x * BLOCK) {
        const double i = (double)(c % nx), j = (double)((c / nx) % ny), k = (double)(c / ((size_t)nx * ny));
        const double x = g.origin[0] + i * h, y = g.origin[1] + j * h, z = g.origin[2] + k * h;
        bool in = shape_sdf(s, x + h * 0.5, y + h * 0.5, z + h * 0.5) < threshold;
        for (int q = 0; q < 8 && !in; q++) in = shape_sdf(s, x + (q & 1 ? h : 0), y + (q & 2 ? h : 0), z + (q & 4 ? h : 0)) < 0;
        occ[c] = in ? 1 : 0;
    }
}

// ------------------------------------------------------------ host side
int fail(int code, const std::string &m) { return ingest_fail(code, m); }

// A number as JS prints it, for the values an error message shows (NaN,
// Infinity, integers, short decimals).
std::string js_number(double v)
{
    if (std::isnan(v)) return "NaN";
    if (std::isinf(v)) return v > 0 ? "Infinity" : "-Infinity";
    if (v == 0) return "0";
    char buf[40];
    for (int digits = 1; digits <= 17; digits++) {
        std::snprintf(buf, sizeof buf, "%.*g", digits, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    return buf;
}

int grid_for_bounds(const double mn[3], const double mx[3], double h, uint32_t padding, Grid &g)
{
    if (!(h > 0 && std::isfinite(h)))
        return fail(VS_E_INVALID, "gridForBounds: elementSize must be a positive length in metres, got " + js_number(h));
    double total = 1;
    for (int a = 0; a < 3; a++) {
        const double extent = mx[a] - mn[a];
        double cells = std::ceil(extent / h - 1e-6);
        if (cells < 1) cells = 1;  // Math.max(1, NaN) is NaN: caught below
        cells += 2.0 * padding;
        g.origin[a] = mn[a] - h * padding;
        if (!(cells <= GRID_AXIS_MAX) || !std::isfinite(g.origin[a]))
            return fail(VS_E_INVALID, "gridForBounds: the bounds need more cells than one grid holds");
        total *= cells;
        g.dims[a] = (uint32_t)cells;
    }
    if (total > GRID_CELLS_MAX) return fail(VS_E_INVALID, "gridForBounds: the bounds need more cells than one grid holds");
    g.h = h;
    return VS_OK;
}


uint32_t grid_blocks(size_t items)
{
    const size_t b = (items + BLOCK - 1) / BLOCK;
    return (uint32_t)(b < 1 ? 1 : b > 65536 ? 65536 : b);
}

// Device buffers and a stream for one call, released on every exit.
struct Session {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<void *> buffers;
    hipError_t status = hipSuccess;
    ~Session()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : buffers) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    bool open(int dev)
    {
        device = dev;
        status = hipSetDevice(dev);
        if (status == hipSuccess) status = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
        return status == hipSuccess;
    }
    template <typename T>
    T *alloc(size_t count)
    {
        void *p = nullptr;
        if (status == hipSuccess) status = hipMalloc(&p, std::max<size_t>(count * sizeof(T), 4));
        if (status == hipSuccess) {
            buffers.push_back(p);
            status = hipMemsetAsync(p, 0, std::max<size_t>(count * sizeof(T), 4), stream);
        }
        return status == hipSuccess ? static_cast<T *>(p) : nullptr;
    }
    template <typename T>
    T *upload(const T *src, size_t count)
    {
        T *d = alloc<T>(count);
        if (d && count) status = hipMemcpyAsync(d, src, count * sizeof(T), hipMemcpyHostToDevice, stream);
        return status == hipSuccess ? d : nullptr;
    }
    bool check(hipError_t e)
    {
        if (status == hipSuccess) status = e;
        return status == hipSuccess;
    }
    // Frees one buffer early (the stream must be done with it).
    void release(void *p)
    {
        for (auto &b : buffers)
            if (b == p) {
                (void)hipFree(b);
                b = buffers.back();
                buffers.pop_back();
                return;
            }
    }
};

#define SESSION_OK(S)                                                                          \
    do {                                                                                       \
        if ((S).status != hipSuccess) return fail(VS_E_DEVICE, hipGetErrorString((S).status)); \
    } while (0)

// Hands the grid to the caller: layout plus a malloc'd copy of the bytes.
int deliver(Session &S, const Grid &g, const uint8_t *d_occ, vs_occupancy *out)
{
    const size_t n = cells_of(g);
    uint8_t *host = static_cast<uint8_t *>(std::malloc(n ? n : 1));
    if (!host) return fail(VS_E_NOMEM, "occupancy grid does not fit host memory");
    S.check(hipMemcpyAsync(host, d_occ, n, hipMemcpyDeviceToHost, S.stream));
    S.check(hipStreamSynchronize(S.stream));
    if (S.status != hipSuccess) {
        std::free(host);
        return fail(VS_E_DEVICE, hipGetErrorString(S.status));
    }
    for (int a = 0; a < 3; a++) {
        out->dims[a] = g.dims[a];
        out->origin[a] = g.origin[a];
    }
    out->h = g.h;
    out->cells = host;
    return VS_OK;
}

// The most work units one call may need (UNIT ray tests, edge samples or
// stamped cells each): 2^30 units are 6.9e10 of them, minutes of work for
// the reference's loops.
constexpr uint64_t UNITS_MAX = ITEM_UNITS_CAP - 1;

// First work unit of every item into `start`, and the total.
int plan_units(Session &S, const Items &it, uint64_t items, const Grid &g, uint64_t *start, uint64_t *tile,
               uint64_t *total)
{
    const uint64_t tiles = (items + BLOCK - 1) / BLOCK;
    hipLaunchKernelGGL(k_units, dim3((uint32_t)tiles), dim3(BLOCK), 0, S.stream, it, items, g, start, tile);
    std::vector<uint64_t> ht(tiles);
    S.check(hipMemcpyAsync(ht.data(), tile, tiles * 8, hipMemcpyDeviceToHost, S.stream));
    S.check(hipStreamSynchronize(S.stream));
    SESSION_OK(S);
    uint64_t at = 0;
    for (uint64_t t = 0; t < tiles; t++) {
        const uint64_t n = ht[t];
        ht[t] = at;
        at += n;
        if (n > UNITS_MAX || at > UNITS_MAX)
            return fail(VS_E_INVALID, it.mode < 3    ? "voxelizeMesh: the triangles cover more than 2^36 ray tests"
                                      : it.mode == 3 ? "voxelizeMesh: the edges need more than 2^36 samples"
                                                     : "voxelizeCloud: the Gaussians' boxes hold more than 2^36 cells");
    }
    S.check(hipMemcpyAsync(tile, ht.data(), tiles * 8, hipMemcpyHostToDevice, S.stream));
    hipLaunchKernelGGL(k_add_offsets, dim3(grid_blocks(items)), dim3(BLOCK), 0, S.stream, start, items,
                       (const uint64_t *)tile);
    // `ht` must outlive the copy.
    S.check(hipStreamSynchronize(S.stream));
    S.check(hipGetLastError());
    SESSION_OK(S);
    *total = at;
    return VS_OK;
}

// sealAndFill (cloudVolume.js:288-320) on the device.
void seal_and_fill(Session &S, const Grid &g, uint8_t *occ, uint32_t seal)
{
    const size_t n = cells_of(g);
    uint8_t *stamp = S.alloc<uint8_t>(n), *tmp = S.alloc<uint8_t>(n), *reached = S.alloc<uint8_t>(n);
    uint32_t *changed = S.alloc<uint32_t>(1);
    if (!changed) return;
    S.check(hipMemcpyAsync(stamp, occ, n, hipMemcpyDeviceToDevice, S.stream));
    const uint32_t blocks = grid_blocks(n);
    for (uint32_t p = 0; p < seal; p++) {
        hipLaunchKernelGGL(k_neighbourhood, dim3(blocks), dim3(BLOCK), 0, S.stream, g, occ, tmp, 1);
        S.check(hipMemcpyAsync(occ, tmp, n, hipMemcpyDeviceToDevice, S.stream));
    }
    hipLaunchKernelGGL(k_flood_seed, dim3(blocks), dim3(BLOCK), 0, S.stream, g, occ, reached);
    // Every round that changes something reaches one more cell, so this ends.
    while (S.status == hipSuccess) {
        uint32_t host_changed = 0;
        S.check(hipMemsetAsync(changed, 0, 4, S.stream));
        for (int axis = 0; axis < 3; axis++)
            for (int forward = 1; forward >= 0; forward--) {
                const int u = (axis + 1) % 3, v = (axis + 2) % 3;
                const size_t lines = (size_t)g.dims[u] * g.dims[v];
                hipLaunchKernelGGL(k_flood_sweep, dim3(grid_blocks(lines)), dim3(BLOCK), 0, S.stream, g, occ, reached,
                                   axis, forward, changed);
            }
        S.check(hipMemcpyAsync(&host_changed, changed, 4, hipMemcpyDeviceToHost, S.stream));
        S.check(hipStreamSynchronize(S.stream));
        if (!host_changed) break;
    }
    hipLaunchKernelGGL(k_flood_finish, dim3(blocks), dim3(BLOCK), 0, S.stream, occ, reached, n);
    for (uint32_t p = 0; p < seal; p++) {
        hipLaunchKernelGGL(k_neighbourhood, dim3(blocks), dim3(BLOCK), 0, S.stream, g, occ, tmp, 0);
        S.check(hipMemcpyAsync(occ, tmp, n, hipMemcpyDeviceToDevice, S.stream));
    }
    hipLaunchKernelGGL(k_or, dim3(blocks), dim3(BLOCK), 0, S.stream, occ, stamp, n);
    S.check(hipGetLastError());
}

}  // namespace
}  // namespace vsim

extern "C" {

using namespace vsim;

void vs_occupancy_free(vs_occupancy *o)
{
    if (!o) return;
    std::free(o->cells);
    o->cells = nullptr;
}

int vs_voxelize_cloud(int device, const float *positions, const float *rotations, const float *scales, uint64_t count,
                      double element_size, uint32_t padding, double sigma, uint32_t seal, int32_t fill,
                      vs_occupancy *out)
{
    if (!out || (count && (!positions || !rotations || !scales))) return fail(VS_E_INVALID, "missing argument");
    *out = vs_occupancy{};
    if (count == 0) return fail(VS_E_INVALID, "voxelizeCloud: a cloud with no Gaussians has no volume");
    Session S;
    if (!S.open(device)) return fail(VS_E_DEVICE, hipGetErrorString(S.status));
    const float *p = S.upload(positions, 3 * count), *r = S.upload(rotations, 4 * count), *s = S.upload(scales, 3 * count);
    const uint32_t blocks = std::min<uint32_t>(grid_blocks(count), 1024);
    double *partial = S.alloc<double>(6 * (size_t)blocks);
    SESSION_OK(S);
    hipLaunchKernelGGL(k_cloud_bounds, dim3(blocks), dim3(BLOCK), 0, S.stream, p, r, s, count, sigma, partial);
    std::vector<double> part(6 * (size_t)blocks);
    S.check(hipMemcpyAsync(part.data(), partial, part.size() * 8, hipMemcpyDeviceToHost, S.stream));
    S.check(hipStreamSynchronize(S.stream));
    SESSION_OK(S);
    double mn[3] = { INFINITY, INFINITY, INFINITY }, mx[3] = { -INFINITY, -INFINITY, -INFINITY };
    for (uint32_t b = 0; b < blocks; b++)
        for (int a = 0; a < 3; a++) {
            mn[a] = js_min(mn[a], part[6 * b + a]);
            mx[a] = js_max(mx[a], part[6 * b + 3 + a]);
        }
    // Math.max over the extents: NaN if any extent is NaN.
    double widest = -INFINITY;
    for (int a = 0; a < 3; a++) {
        const double e = mx[a] - mn[a];
        widest = js_max(widest, e);
    }
    if (!(widest > 0)) return fail(VS_E_INVALID, "voxelizeCloud: a cloud with no extent has no volume");
    Grid g;
    const int grid_rc = grid_for_bounds(mn, mx, element_size, padding, g);
    if (grid_rc != VS_OK) return grid_rc;
    uint8_t *occ = S.alloc<uint8_t>(cells_of(g));
    uint64_t *start = S.alloc<uint64_t>(count), *tile = S.alloc<uint64_t>((count + BLOCK - 1) / BLOCK);
    SESSION_OK(S);
    uint64_t units = 0;
    const int rc = plan_units(S, Items{ p, r, s, nullptr, sigma, 4 }, count, g, start, tile, &units);
    if (rc != VS_OK) return rc;
    hipLaunchKernelGGL(k_cloud_stamp, dim3(grid_blocks(units)), dim3(BLOCK), 0, S.stream, p, r, s, count, sigma, g,
                       (const uint64_t *)start, units, occ);
    S.check(hipGetLastError());
    if (fill) seal_and_fill(S, g, occ, seal);
    SESSION_OK(S);
    return deliver(S, g, occ, out);
}

int vs_voxelize_mesh(int device, const float *positions, uint64_t vertices, const uint32_t *indices, uint64_t triangles,
                     const double *bounds, double element_size, uint32_t padding, int32_t close, vs_occupancy *out)
{
    if (!out || !positions || (triangles && !indices)) return fail(VS_E_INVALID, "missing argument");
    *out = vs_occupancy{};
    // Inputs on which the reference's grid depends on its engine's sort order
    // (a NaN crossing) or on out-of-range reads are refused.
    for (uint64_t i = 0; i < 3 * vertices; i++)
        if (!std::isfinite(positions[i]))
            return fail(VS_E_INVALID, "voxelizeMesh: vertex coordinates must be finite and indices inside the vertices");
    for (uint64_t i = 0; i < 3 * triangles; i++)
        if (indices[i] >= vertices)
            return fail(VS_E_INVALID, "voxelizeMesh: vertex coordinates must be finite and indices inside the vertices");
    // meshBounds (restGeometry.js:29-40) on the host unless the caller
    // passes bounds: one pass over the floats.
    double mn[3] = { INFINITY, INFINITY, INFINITY }, mx[3] = { -INFINITY, -INFINITY, -INFINITY };
    if (bounds) {
        for (int a = 0; a < 3; a++) mn[a] = bounds[a], mx[a] = bounds[3 + a];
    } else {
        for (uint64_t i = 0; i < vertices; i++)
            for (int a = 0; a < 3; a++) {
                const double v = positions[3 * i + a];
                if (v < mn[a]) mn[a] = v;
                if (v > mx[a]) mx[a] = v;
            }
    }
    Grid g;
    const int grid_rc = grid_for_bounds(mn, mx, element_size, padding, g);
    if (grid_rc != VS_OK) return grid_rc;
    Session S;
    if (!S.open(device)) return fail(VS_E_DEVICE, hipGetErrorString(S.status));
    const size_t n = cells_of(g);
    const float *p = S.upload(positions, 3 * vertices);
    const uint32_t *ix = S.upload(indices, 3 * triangles);
    uint8_t *votes = S.alloc<uint8_t>(n), *surface = S.alloc<uint8_t>(n), *occ = S.alloc<uint8_t>(n);
    size_t most_columns = 0;
    for (int axis = 0; axis < 3; axis++)
        most_columns = std::max(most_columns, (size_t)g.dims[(axis + 1) % 3] * g.dims[(axis + 2) % 3]);
    uint64_t *start = S.alloc<uint64_t>(3 * triangles);
    uint64_t *tile = S.alloc<uint64_t>((3 * triangles + BLOCK - 1) / BLOCK);
    uint32_t *count = S.alloc<uint32_t>(most_columns);
    uint64_t *offset = S.alloc<uint64_t>(most_columns);
    SESSION_OK(S);
    for (int axis = 0; triangles && axis < 3; axis++) {
        const int u = (axis + 1) % 3, v = (axis + 2) % 3;
        const size_t columns = (size_t)g.dims[u] * g.dims[v];
        uint64_t units = 0;
        const int rc = plan_units(S, Items{ p, nullptr, nullptr, ix, 0.0, axis }, triangles, g, start, tile, &units);
        if (rc != VS_OK) return rc;
        S.check(hipMemsetAsync(count, 0, columns * 4, S.stream));
        SESSION_OK(S);
        const uint32_t blocks = grid_blocks(units);
        hipLaunchKernelGGL(k_mesh_crossings, dim3(blocks), dim3(BLOCK), 0, S.stream, p, ix, triangles, g, axis,
                           (const uint64_t *)start, units, count, (const uint64_t *)nullptr, (double *)nullptr);
        std::vector<uint32_t> hc(columns);
        std::vector<uint64_t> ho(columns);
        S.check(hipMemcpyAsync(hc.data(), count, columns * 4, hipMemcpyDeviceToHost, S.stream));
        S.check(hipStreamSynchronize(S.stream));
        SESSION_OK(S);
        uint64_t total = 0;
        for (size_t c = 0; c < columns; c++) {
            ho[c] = total;
            total += hc[c];
        }
        double *hits = S.alloc<double>(total);
        S.check(hipMemcpyAsync(offset, ho.data(), columns * 8, hipMemcpyHostToDevice, S.stream));
        S.check(hipMemsetAsync(count, 0, columns * 4, S.stream));
        SESSION_OK(S);
        hipLaunchKernelGGL(k_mesh_crossings, dim3(blocks), dim3(BLOCK), 0, S.stream, p, ix, triangles, g, axis,
                           (const uint64_t *)start, units, count, (const uint64_t *)offset, hits);
        hipLaunchKernelGGL(k_mesh_columns, dim3(grid_blocks(columns)), dim3(BLOCK), 0, S.stream, g, axis,
                           (const uint32_t *)count, (const uint64_t *)offset, hits, votes, surface);
        // The host copies must outlive the stream-ordered copy above.
        S.check(hipStreamSynchronize(S.stream));
        S.release(hits);
        SESSION_OK(S);
    }
    if (triangles) {
        uint64_t units = 0;
        const int rc = plan_units(S, Items{ p, nullptr, nullptr, ix, 0.0, 3 }, 3 * triangles, g, start, tile, &units);
        if (rc != VS_OK) return rc;
        hipLaunchKernelGGL(k_mesh_edges, dim3(grid_blocks(units)), dim3(BLOCK), 0, S.stream, p, ix, 3 * triangles, g,
                           (const uint64_t *)start, units, surface);
    }
    hipLaunchKernelGGL(k_mesh_vote, dim3(grid_blocks(n)), dim3(BLOCK), 0, S.stream, (const uint8_t *)votes,
                       (const uint8_t *)surface, occ, n);
    if (close) {
        hipLaunchKernelGGL(k_neighbourhood, dim3(grid_blocks(n)), dim3(BLOCK), 0, S.stream, g, (const uint8_t *)occ,
                           votes, 1);
        hipLaunchKernelGGL(k_neighbourhood, dim3(grid_blocks(n)), dim3(BLOCK), 0, S.stream, g, (const uint8_t *)votes,
                           occ, 0);
    }
    S.check(hipGetLastError());
    SESSION_OK(S);
    return deliver(S, g, occ, out);
}

int vs_voxelize_shape(int device, const vs_shape *shape, double element_size, uint32_t padding, vs_occupancy *out)
{
    if (!out || !shape || shape->kind < VS_SHAPE_SPHERE || shape->kind > VS_SHAPE_TORUS || shape->n_offsets > 4)
        return fail(VS_E_INVALID, "unknown shape");
    *out = vs_occupancy{};
    double lo[3], hi[3];
    if (shape->kind == VS_SHAPE_SPHERE) {
        for (int a = 0; a < 3; a++) lo[a] = -shape->size[0], hi[a] = shape->size[0];
    } else if (shape->kind == VS_SHAPE_BOX) {
        for (int a = 0; a < 3; a++) lo[a] = -shape->size[a], hi[a] = shape->size[a];
    } else {
        const double outer = shape->size[0] + shape->size[1];
        lo[0] = -outer, lo[1] = -shape->size[1], lo[2] = -outer;
        hi[0] = outer, hi[1] = shape->size[1], hi[2] = outer;
    }
    double mn[3], mx[3];
    for (int a = 0; a < 3; a++) {
        mn[a] = lo[a];
        mx[a] = hi[a];
        for (uint32_t t = 0; t < shape->n_offsets; t++) {
            mn[a] += shape->offset[t][a];
            mx[a] += shape->offset[t][a];
        }
    }
    Grid g;
    const int grid_rc = grid_for_bounds(mn, mx, element_size, padding, g);
    if (grid_rc != VS_OK) return grid_rc;
    for (int a = 0; a < 3; a++) g.origin[a] = g.origin[a] - element_size * 0.5;
    Session S;
    if (!S.open(device)) return fail(VS_E_DEVICE, hipGetErrorString(S.status));
    uint8_t *occ = S.alloc<uint8_t>(cells_of(g));
    SESSION_OK(S);
    hipLaunchKernelGGL(k_shape, dim3(grid_blocks(cells_of(g))), dim3(BLOCK), 0, S.stream, *shape, g, occ);
    S.check(hipGetLastError());
    SESSION_OK(S);
    return deliver(S, g, occ, out);
}

}  // extern "C"
