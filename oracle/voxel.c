/*
 * voxel.c -- the CPU oracle of the voxelisers: Gaussian clouds
 * (packages/visionary-sim/src/sim/discretize/lattice/cloudVolume.js) and
 * triangle meshes (voxelize.js:176-364), plus the grid layout
 * (gridForBounds, voxelize.js:40-67).
 *
 * TEST INFRASTRUCTURE, like oracle.c: only tests/ load it, as the checker of
 * the device voxelisers.
 *
 * Arithmetic is the JS's: doubles throughout, f32 inputs widened, every
 * expression evaluated left to right as written there (no contraction), and
 * the rotation columns rounded to f32 as the reference stores them in a
 * Float32Array (mat3.js:71-86). One deliberate difference: Math.hypot is
 * restated as sqrt(a*a + b*b + c*c), which can differ from an engine's hypot
 * in the last bit of a double.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint32_t dims[3];
    double origin[3];
    double h;
} o_grid;

static size_t cell_at(const o_grid *g, int64_t i, int64_t j, int64_t k)
{
    return (size_t)i + (size_t)g->dims[0] * ((size_t)j + (size_t)g->dims[1] * (size_t)k);
}

/* The largest grid either side builds: 2^24 cells on an axis, 2^36 in all.
 * The reference has no such limit and fails inside its typed-array
 * allocation instead. */
#define GRID_AXIS_MAX 16777216.0
#define GRID_CELLS_MAX 68719476736.0

/* gridForBounds (voxelize.js:40-67). Returns 0; -3 for a bad size, -4 for a
 * grid beyond the limit above (infinite or NaN bounds included). */
int o_grid_for_bounds(const double mn[3], const double mx[3], double h, uint32_t padding, o_grid *g)
{
    if (!(h > 0 && isfinite(h))) return -3;
    double total = 1;
    for (int a = 0; a < 3; a++) {
        const double extent = mx[a] - mn[a];
        double cells = ceil(extent / h - 1e-6);
        if (cells < 1) cells = 1;                  /* Math.max(1, NaN) is NaN: caught below */
        cells += 2.0 * padding;
        if (!(cells <= GRID_AXIS_MAX)) return -4;
        total *= cells;
        g->dims[a] = (uint32_t)cells;
        g->origin[a] = mn[a] - h * padding;
        if (!isfinite(g->origin[a])) return -4;
    }
    if (total > GRID_CELLS_MAX) return -4;
    g->h = h;
    return 0;
}

/* Dilate (1) or erode (0) with the 6-neighbourhood, outside counting as
 * empty (voxelize.js:396-435). */
void o_neighbourhood(const o_grid *g, const uint8_t *src, uint8_t *out, int dilate)
{
    const int64_t nx = g->dims[0], ny = g->dims[1], nz = g->dims[2];
    for (int64_t k = 0; k < nz; k++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t i = 0; i < nx; i++) {
                const size_t at = cell_at(g, i, j, k);
                int result = src[at] == 1;
                for (int axis = 0; axis < 3 && result != dilate; axis++)
                    for (int step = -1; step <= 1; step += 2) {
                        const int64_t ni = i + (axis == 0 ? step : 0), nj = j + (axis == 1 ? step : 0),
                                      nk = k + (axis == 2 ? step : 0);
                        const int outside = ni < 0 || nj < 0 || nk < 0 || ni >= nx || nj >= ny || nk >= nz;
                        const uint8_t nb = outside ? 0 : src[cell_at(g, ni, nj, nk)];
                        if (dilate ? nb == 1 : nb == 0) {
                            result = dilate;
                            break;
                        }
                    }
                out[at] = result ? 1 : 0;
            }
}

/* fillEnclosed (cloudVolume.js:322-373): empty cells the faces cannot reach
 * through face neighbours become material. Returns how many. */
uint64_t o_fill_enclosed(const o_grid *g, uint8_t *occ)
{
    const int64_t nx = g->dims[0], ny = g->dims[1], nz = g->dims[2];
    const size_t n = (size_t)nx * ny * nz;
    uint8_t *reached = calloc(n, 1);
    int64_t *stack = malloc(sizeof(int64_t) * (n + 1));
    size_t top = 0;
    for (int64_t k = 0; k < nz; k++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t i = 0; i < nx; i++)
                if (i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1) {
                    const size_t c = cell_at(g, i, j, k);
                    if (!occ[c] && !reached[c]) {
                        reached[c] = 1;
                        stack[top++] = (int64_t)c;
                    }
                }
    while (top) {
        const int64_t c = stack[--top];
        const int64_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
        const int64_t d[6][3] = { { 1, 0, 0 }, { -1, 0, 0 }, { 0, 1, 0 }, { 0, -1, 0 }, { 0, 0, 1 }, { 0, 0, -1 } };
        for (int e = 0; e < 6; e++) {
            const int64_t ni = i + d[e][0], nj = j + d[e][1], nk = k + d[e][2];
            if (ni < 0 || nj < 0 || nk < 0 || ni >= nx || nj >= ny || nk >= nz) continue;
            const size_t m = cell_at(g, ni, nj, nk);
            if (!occ[m] && !reached[m]) {
                reached[m] = 1;
                stack[top++] = (int64_t)m;
            }
        }
    }
    uint64_t filled = 0;
    for (size_t c = 0; c < n; c++)
        if (!occ[c] && !reached[c]) {
            occ[c] = 1;
            filled++;
        }
    free(stack);
    free(reached);
    return filled;
}

/* mat3FromQuat (mat3.js:71-86), stored in f32 as the reference stores it. */
static void rotation_columns(const float *q, float out[9])
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

static double hypot3(double a, double b, double c) { return sqrt(a * a + b * b + c * c); }

/* ellipsoidOf (cloudVolume.js:182-210). */
static void ellipsoid_of(const float *rot, const float *scale, double sigma, float col[9], double radius[3],
                         double half[3])
{
    for (int a = 0; a < 3; a++) radius[a] = sigma * scale[a];
    rotation_columns(rot, col);
    for (int a = 0; a < 3; a++)
        half[a] = hypot3((double)col[a] * radius[0], (double)col[3 + a] * radius[1], (double)col[6 + a] * radius[2]);
}

static int64_t clampi(double v, int64_t lo, int64_t hi)
{
    /* Math.min(Math.max(v, lo), hi) for a finite v: hi wins when lo > hi. */
    const double t = v < (double)lo ? (double)lo : v;
    return t > (double)hi ? hi : (int64_t)t;
}

/* voxelizeCloud (cloudVolume.js:95-121): the grid of a cloud of Gaussians,
 * stamped as solid ellipsoids at `sigma` standard deviations and, with
 * `fill`, sealed (dilate `seal` times), flooded from the faces, eroded back
 * and re-stamped. Returns 0; -1 no Gaussians; -2 no extent; a grid error.
 * `occ` (dims product bytes) is allocated with malloc when *occ is NULL. */
int o_voxelize_cloud(const float *pos, const float *rot, const float *scale, uint64_t count, double h,
                     uint32_t padding, double sigma, uint32_t seal, int fill, o_grid *g, uint8_t **occ)
{
    if (count == 0) return -1;
    double mn[3] = { INFINITY, INFINITY, INFINITY }, mx[3] = { -INFINITY, -INFINITY, -INFINITY };
    float col[9];
    double radius[3], half[3];
    for (uint64_t i = 0; i < count; i++) {
        ellipsoid_of(rot + 4 * i, scale + 3 * i, sigma, col, radius, half);
        /* Math.min / Math.max: a NaN anywhere makes the bound NaN. */
        for (int a = 0; a < 3; a++) {
            const double lo = (double)pos[3 * i + a] - half[a], hi = (double)pos[3 * i + a] + half[a];
            mn[a] = isnan(lo) || isnan(mn[a]) ? NAN : lo < mn[a] ? lo : mn[a];
            mx[a] = isnan(hi) || isnan(mx[a]) ? NAN : hi > mx[a] ? hi : mx[a];
        }
    }
    double widest = -INFINITY;
    for (int a = 0; a < 3; a++) {
        const double e = mx[a] - mn[a];
        widest = isnan(e) || isnan(widest) ? NAN : e > widest ? e : widest;
    }
    if (!(widest > 0)) return -2;
    const int rc = o_grid_for_bounds(mn, mx, h, padding, g);
    if (rc) return rc;
    const size_t n = (size_t)g->dims[0] * g->dims[1] * g->dims[2];
    if (!*occ) *occ = malloc(n);
    uint8_t *o = *occ;
    memset(o, 0, n);
    for (uint64_t i = 0; i < count; i++) {
        const double c[3] = { pos[3 * i], pos[3 * i + 1], pos[3 * i + 2] };
        int64_t cc[3];
        for (int a = 0; a < 3; a++) cc[a] = clampi(floor((c[a] - g->origin[a]) / g->h), 0, (int64_t)g->dims[a] - 1);
        o[cell_at(g, cc[0], cc[1], cc[2])] = 1;
        ellipsoid_of(rot + 4 * i, scale + 3 * i, sigma, col, radius, half);
        int64_t lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = clampi(floor((c[a] - half[a] - g->origin[a]) / g->h), 1, (int64_t)g->dims[a] - 2);
            hi[a] = clampi(floor((c[a] + half[a] - g->origin[a]) / g->h), 1, (int64_t)g->dims[a] - 2);
        }
        for (int64_t k = lo[2]; k <= hi[2]; k++)
            for (int64_t j = lo[1]; j <= hi[1]; j++)
                for (int64_t ii = lo[0]; ii <= hi[0]; ii++) {
                    /* On a grid under three cells thick the box can reach
                     * index -1; like a typed array, an index outside the
                     * grid is skipped. */
                    const int64_t lin = ii + (int64_t)g->dims[0] * (j + (int64_t)g->dims[1] * k);
                    if (lin < 0 || (uint64_t)lin >= n) continue;
                    const size_t cell = (size_t)lin;
                    if (o[cell]) continue;
                    for (int s = 0; s < 9; s++) {
                        const double ox = s == 8 ? 0.5 : (s & 1), oy = s == 8 ? 0.5 : ((s >> 1) & 1),
                                     oz = s == 8 ? 0.5 : ((s >> 2) & 1);
                        const double dx = g->origin[0] + (ii + ox) * g->h - c[0];
                        const double dy = g->origin[1] + (j + oy) * g->h - c[1];
                        const double dz = g->origin[2] + (k + oz) * g->h - c[2];
                        double sum = 0;
                        for (int a = 0; a < 3; a++) {
                            const double t = ((double)col[a * 3] * dx + (double)col[a * 3 + 1] * dy +
                                              (double)col[a * 3 + 2] * dz) / radius[a];
                            sum += t * t;
                        }
                        if (sum <= 1) {
                            o[cell] = 1;
                            break;
                        }
                    }
                }
    }
    if (fill) {
        uint8_t *stamp = malloc(n), *tmp = malloc(n);
        memcpy(stamp, o, n);
        for (uint32_t p = 0; p < seal; p++) {
            o_neighbourhood(g, o, tmp, 1);
            memcpy(o, tmp, n);
        }
        o_fill_enclosed(g, o);
        for (uint32_t p = 0; p < seal; p++) {
            o_neighbourhood(g, o, tmp, 0);
            memcpy(o, tmp, n);
        }
        for (size_t c = 0; c < n; c++) o[c] |= stamp[c];
        free(tmp);
        free(stamp);
    }
    return 0;
}

static int cmp_double(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y;
}

/* accumulateAxis (voxelize.js:258-362). */
static void accumulate_axis(const float *pos, const uint32_t *idx, uint64_t tris, const o_grid *g, int axis,
                            uint8_t *votes, uint8_t *surface)
{
    const int u = (axis + 1) % 3, v = (axis + 2) % 3, w = axis;
    const double h = g->h;
    const size_t columns = (size_t)g->dims[u] * g->dims[v];
    size_t *count = calloc(columns, sizeof(size_t)), *cap = calloc(columns, sizeof(size_t));
    double **hits = calloc(columns, sizeof(double *));
    const double jitter_u = h * 1.7e-4, jitter_v = h * 0.9e-4;
    for (uint64_t t = 0; t < tris; t++) {
        const float *A = pos + 3 * (size_t)idx[3 * t], *B = pos + 3 * (size_t)idx[3 * t + 1],
                    *C = pos + 3 * (size_t)idx[3 * t + 2];
        const double au = A[u], av = A[v], aw = A[w], bu = B[u], bv = B[v], bw = B[w], cu = C[u], cv = C[v],
                     cw = C[w];
        const double area2 = (bv - cv) * (au - cu) + (cu - bu) * (av - cv);
        if (fabs(area2) < 1e-20) continue;
        const double inv = 1 / area2;
        const double min_u = fmin(au, fmin(bu, cu)), max_u = fmax(au, fmax(bu, cu));
        const double min_v = fmin(av, fmin(bv, cv)), max_v = fmax(av, fmax(bv, cv));
        double a0 = ceil((min_u - g->origin[u]) / h - 0.5), a1 = floor((max_u - g->origin[u]) / h - 0.5);
        double b0 = ceil((min_v - g->origin[v]) / h - 0.5), b1 = floor((max_v - g->origin[v]) / h - 0.5);
        a0 = a0 > 0 ? a0 : 0;
        b0 = b0 > 0 ? b0 : 0;
        a1 = a1 < g->dims[u] - 1.0 ? a1 : g->dims[u] - 1.0;
        b1 = b1 < g->dims[v] - 1.0 ? b1 : g->dims[v] - 1.0;
        for (double cv2 = b0; cv2 <= b1; cv2++) {
            const double pv = g->origin[v] + (cv2 + 0.5) * h + jitter_v;
            for (double cu2 = a0; cu2 <= a1; cu2++) {
                const double pu = g->origin[u] + (cu2 + 0.5) * h + jitter_u;
                const double l0 = ((bv - cv) * (pu - cu) + (cu - bu) * (pv - cv)) * inv;
                if (l0 < 0) continue;
                const double l1 = ((cv - av) * (pu - cu) + (au - cu) * (pv - cv)) * inv;
                if (l1 < 0) continue;
                const double l2 = 1 - l0 - l1;
                if (l2 < 0) continue;
                const size_t col = (size_t)cu2 + (size_t)g->dims[u] * (size_t)cv2;
                if (count[col] == cap[col]) {
                    cap[col] = cap[col] ? 2 * cap[col] : 4;
                    hits[col] = realloc(hits[col], sizeof(double) * cap[col]);
                }
                hits[col][count[col]++] = l0 * aw + l1 * bw + l2 * cw;
            }
        }
    }
    int64_t ijk[3];
    for (size_t cv2 = 0; cv2 < g->dims[v]; cv2++)
        for (size_t cu2 = 0; cu2 < g->dims[u]; cu2++) {
            const size_t col = cu2 + g->dims[u] * cv2;
            if (!count[col]) continue;
            qsort(hits[col], count[col], sizeof(double), cmp_double);
            ijk[u] = (int64_t)cu2;
            ijk[v] = (int64_t)cv2;
            for (size_t e = 0; e < count[col]; e++) {
                const double cw2 = floor((hits[col][e] - g->origin[w]) / h);
                if (cw2 >= 0 && cw2 < g->dims[w]) {
                    ijk[w] = (int64_t)cw2;
                    surface[cell_at(g, ijk[0], ijk[1], ijk[2])] = 1;
                }
            }
            size_t next = 0;
            int inside = 0;
            for (int64_t cw2 = 0; cw2 < g->dims[w]; cw2++) {
                const double centre = g->origin[w] + (cw2 + 0.5) * h;
                while (next < count[col] && hits[col][next] <= centre) {
                    inside = !inside;
                    next++;
                }
                if (inside) {
                    ijk[w] = cw2;
                    votes[cell_at(g, ijk[0], ijk[1], ijk[2])]++;
                }
            }
            free(hits[col]);
        }
    free(hits);
    free(cap);
    free(count);
}

/* markEdgeCells (voxelize.js:205-256). */
static void mark_edges(const float *pos, const uint32_t *idx, uint64_t tris, const o_grid *g, uint8_t *surface)
{
    const double step = g->h * 0.4;
    for (uint64_t t = 0; t < tris; t++)
        for (int e = 0; e < 3; e++) {
            const float *A = pos + 3 * (size_t)idx[3 * t + e], *B = pos + 3 * (size_t)idx[3 * t + (e + 1) % 3];
            const double dx = (double)B[0] - A[0], dy = (double)B[1] - A[1], dz = (double)B[2] - A[2];
            double steps = ceil(hypot3(dx, dy, dz) / step);
            steps = steps > 1 ? steps : 1;
            for (double s = 0; s <= steps; s++) {
                const double f = s / steps;
                const double x = A[0] + dx * f, y = A[1] + dy * f, z = A[2] + dz * f;
                const double i = floor((x - g->origin[0]) / g->h), j = floor((y - g->origin[1]) / g->h),
                             k = floor((z - g->origin[2]) / g->h);
                if (i >= 0 && j >= 0 && k >= 0 && i < g->dims[0] && j < g->dims[1] && k < g->dims[2])
                    surface[cell_at(g, (int64_t)i, (int64_t)j, (int64_t)k)] = 1;
            }
        }
}

/* voxelizeMesh (voxelize.js:174-203) over `bounds` (min[3], max[3]) or, when
 * NULL, meshBounds (restGeometry.js:29-40). Returns 0, a grid error, or -5
 * for a non-finite vertex coordinate or an index past the vertices (inputs
 * on which the reference's result depends on its engine's sort). */
int o_voxelize_mesh(const float *pos, uint64_t verts, const uint32_t *idx, uint64_t tris, const double *bounds,
                    double h, uint32_t padding, int close, o_grid *g, uint8_t **occ)
{
    double mn[3] = { INFINITY, INFINITY, INFINITY }, mx[3] = { -INFINITY, -INFINITY, -INFINITY };
    for (uint64_t i = 0; i < 3 * verts; i++)
        if (!isfinite(pos[i])) return -5;
    for (uint64_t i = 0; i < 3 * tris; i++)
        if (idx[i] >= verts) return -5;
    if (bounds) {
        for (int a = 0; a < 3; a++) mn[a] = bounds[a], mx[a] = bounds[3 + a];
    } else {
        for (uint64_t i = 0; i < verts; i++)
            for (int a = 0; a < 3; a++) {
                const double v = pos[3 * i + a];
                if (v < mn[a]) mn[a] = v;
                if (v > mx[a]) mx[a] = v;
            }
    }
    const int rc = o_grid_for_bounds(mn, mx, h, padding, g);
    if (rc) return rc;
    const size_t n = (size_t)g->dims[0] * g->dims[1] * g->dims[2];
    uint8_t *votes = calloc(n, 1), *surface = calloc(n, 1);
    for (int axis = 0; axis < 3; axis++) accumulate_axis(pos, idx, tris, g, axis, votes, surface);
    mark_edges(pos, idx, tris, g, surface);
    if (!*occ) *occ = malloc(n);
    uint8_t *o = *occ;
    for (size_t c = 0; c < n; c++) o[c] = votes[c] >= 2 || surface[c] == 1 ? 1 : 0;
    if (close) {
        o_neighbourhood(g, o, votes, 1);
        o_neighbourhood(g, votes, o, 0);
    }
    free(surface);
    free(votes);
    return 0;
}

/* The analytic shapes of apps/simulator/src/assets/shapes.js as signed
 * distances: kind 0 sphere (size[0] = radius, :60-67), 1 box (size = half
 * extents, :89-108), 2 torus (size[0] major, size[1] minor radius, :32-43);
 * then `n_offsets` translations (translateShape, :148-156), offset[0] the
 * first one applied to the shape. A translated shape's sdf subtracts the
 * last translation from the query point first. */
typedef struct {
    int32_t kind;
    uint32_t n_offsets;
    double size[3];
    double offset[4][3];
} o_shape;

static double hypot2(double a, double b) { return sqrt(a * a + b * b); }

double o_shape_sdf(const o_shape *s, double x, double y, double z)
{
    for (uint32_t t = s->n_offsets; t-- > 0;) {
        x -= s->offset[t][0];
        y -= s->offset[t][1];
        z -= s->offset[t][2];
    }
    if (s->kind == 0) return hypot3(x, y, z) - s->size[0];
    if (s->kind == 1) {
        const double dx = fabs(x) - s->size[0], dy = fabs(y) - s->size[1], dz = fabs(z) - s->size[2];
        const double outside = hypot3(dx > 0 ? dx : 0, dy > 0 ? dy : 0, dz > 0 ? dz : 0);
        double m = dx > dy ? dx : dy;
        m = m > dz ? m : dz;
        const double inside = m < 0 ? m : 0;
        return outside + inside;
    }
    const double radial = hypot2(x, z) - s->size[0];
    return hypot2(radial, y) - s->size[1];
}

void o_shape_bounds(const o_shape *s, double mn[3], double mx[3])
{
    double lo[3], hi[3];
    if (s->kind == 0) {
        for (int a = 0; a < 3; a++) lo[a] = -s->size[0], hi[a] = s->size[0];
    } else if (s->kind == 1) {
        for (int a = 0; a < 3; a++) lo[a] = -s->size[a], hi[a] = s->size[a];
    } else {
        const double outer = s->size[0] + s->size[1];
        lo[0] = -outer, lo[1] = -s->size[1], lo[2] = -outer;
        hi[0] = outer, hi[1] = s->size[1], hi[2] = outer;
    }
    for (int a = 0; a < 3; a++) {
        mn[a] = lo[a];
        mx[a] = hi[a];
        for (uint32_t t = 0; t < s->n_offsets; t++) {
            mn[a] += s->offset[t][a];
            mx[a] += s->offset[t][a];
        }
    }
}

/* voxelizeSdf (voxelize.js:120-148): cell centres on the bounds (origin half a
 * cell back), material where the centre is within half a cell of the surface
 * or any corner is inside. */
int o_voxelize_shape(const o_shape *s, double h, uint32_t padding, o_grid *g, uint8_t **occ)
{
    if (s->n_offsets > 4) return -3;
    double mn[3], mx[3];
    o_shape_bounds(s, mn, mx);
    const int rc = o_grid_for_bounds(mn, mx, h, padding, g);
    if (rc) return rc;
    for (int a = 0; a < 3; a++) g->origin[a] = g->origin[a] - h * 0.5;
    const size_t n = (size_t)g->dims[0] * g->dims[1] * g->dims[2];
    if (!*occ) *occ = malloc(n);
    uint8_t *o = *occ;
    const double threshold = h * 0.5;
    for (int64_t k = 0; k < g->dims[2]; k++)
        for (int64_t j = 0; j < g->dims[1]; j++)
            for (int64_t i = 0; i < g->dims[0]; i++) {
                const double x = g->origin[0] + i * h, y = g->origin[1] + j * h, z = g->origin[2] + k * h;
                int in = o_shape_sdf(s, x + h * 0.5, y + h * 0.5, z + h * 0.5) < threshold;
                for (int c = 0; c < 8 && !in; c++)
                    in = o_shape_sdf(s, x + (c & 1 ? h : 0), y + (c & 2 ? h : 0), z + (c & 4 ? h : 0)) < 0;
                o[cell_at(g, i, j, k)] = in ? 1 : 0;
            }
    return 0;
}

void o_free(void *p) { free(p); }
