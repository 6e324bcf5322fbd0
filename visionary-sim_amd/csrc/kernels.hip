// kernels.hip -- the per-substep hot path as hand-written gfx950 kernels.
//
// Arithmetic follows the evaluation order the oracle defines, with
// -ffp-contract=off, IEEE divide and square root, and fused multiply-adds
// only where written (fmaf), so a substep is bit-for-bit the oracle's f32
// substep. The elastic pass uses the oracle's fused order (the Kuhn-path form
// of F = Ds Dm^-1, fmaf in every product sum, one divide per block); the
// other passes follow the WGSL formulas term by term, dropping products with
// a constant 0 and folding constant signs (the tank plane normals), which
// changes no finite result.
//
// Layout: nodes as SoA planes, one wavefront lane per cell / node / record,
// workgroups over tiles (kernels.h). No MFMA: the work is 3x3 algebra per
// cell, not a contraction. The colour passes of large batches are bound by
// VALU issue (about 1.4 k wave instructions a cell, 90% and more of the issue
// rate), the node and splat passes by HBM, and contact by chains of dependent
// loads.
#include "kernels.h"

namespace vsim {
namespace {

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 mk(float x, float y, float z) { return V3{ x, y, z }; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 mul(V3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 neg(V3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// a + l * d, per component, one rounding each.
__device__ __forceinline__ V3 lerpf(V3 a, V3 d, float l) { return mk(fmaf(l, d.x, a.x), fmaf(l, d.y, a.y), fmaf(l, d.z, a.z)); }

// Kuhn tetrahedra (kuhn.js:28-35). Each is a path of three axis steps from
// corner 0 to corner 7 (corner c = (c&1, (c>>1)&1, (c>>2)&1)), so
// F = Ds * Dm^-1 has the path's edge differences times 1/h as its columns,
// axis by axis, and (dC/dF) * B^T hands each path corner the difference of
// the per-axis gradients of the steps into and out of it. `path` lists the
// corners in path order, `axis[j]` the axis of step j; both computed at
// compile time.
struct Tet {
    int v[4];
    int path[4];
    int axis[3];
};

constexpr Tet make_tet(int a, int b, int c, int d)
{
    Tet t{ { a, b, c, d }, { 0, 0, 0, 0 }, { 0, 0, 0 } };
    int at = 0;
    for (int j = 0; j < 3; j++) {
        for (int ax = 0; ax < 3; ax++) {
            const int next = at | (1 << ax);
            if (next != at && (next == a || next == b || next == c || next == d)) {
                t.axis[j] = ax;
                at = next;
                break;
            }
        }
        t.path[j + 1] = at;
    }
    return t;
}

constexpr Tet TETS[6] = { make_tet(0, 1, 3, 7), make_tet(0, 5, 1, 7), make_tet(0, 3, 2, 7),
                          make_tet(0, 2, 6, 7), make_tet(0, 4, 5, 7), make_tet(0, 6, 4, 7) };

__device__ __forceinline__ float dotf(V3 a, V3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
__device__ __forceinline__ float dotacc(V3 a, V3 b, float acc) { return fmaf(a.z, b.z, fmaf(a.y, b.y, fmaf(a.x, b.x, acc))); }
__device__ __forceinline__ V3 crossf(V3 a, V3 b)
{
    return mk(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x)));
}

// One tetrahedron of project.wgsl:75-153 (the stable Neo-Hookean pair solved
// as the 2x2 block of xpbdCouple.wgsl:58-71), in the evaluation order the
// oracle's project_tet_fused defines. A_j = p(P_j+1) - p(P_j) are the path's
// edges (E: the same by axis, F = s E, s = 1/h), K = cof(E) (cof F = s^2 K),
// K_j the column of step j's axis. Corner sums become edge sums: with W_j the
// weight at path position j,
//     Ge = (W0+W1)|A0|^2 + (W1+W2)|A1|^2 + (W2+W3)|A2|^2 - 2(W1 A0.A1 + W2 A1.A2)
// (Gf the same over K), and the rates Re = sum_j A_j.(A_j - Q_j) = |E|^2 -
// sum_j A_j.Q_j, Rf = sum_j K_j.(A_j - Q_j) = 3 det E - sum_j K_j.Q_j over
// the edges at the substep's start Q_j = q(P_j+1) - q(P_j) (A_j.K_j = det E
// for each j). With c = s/|E|, the scaled multipliers u = c lambda_D, v = s^3
// lambda_H, the block's first row multiplied by 1/c^2 = |E|^2 h^2 and its
// second by h^6, the block is symmetric:
//     [Ge st + aD h^2 |E|^2,  Gef st        ] [u]   [-(|E|^2 + r Re)            ]
//     [Gef st,                Gf st + aH h^6] [v] = [-(det E - gamma h^3 + r Rf)]
// (st = 1 + r), and edge j moves by G_j = A_j u + K_j v, corner P_j by
// W_j (G_j-1 - G_j). Same mathematics as the WGSL form, with no square root,
// one division and no 3x3 scaling of F or of the gradients. |F| <= 1e-9
// leaves the deviatoric row out (u = 0), as the WGSL's zero gradient does;
// the singular-block test scales the determinant back by s^6.
struct TetConsts {
    float s6, gh3, aD, aHh6, aDh2, relax, devMin, detMin;
};

template <int T>
__device__ __forceinline__ void project_tet(V3 (&p)[8], const V3 (&q)[8], const float (&w)[8], const TetConsts &k)
{
    constexpr Tet tet = TETS[T];
    static_assert(tet.path[0] == 0 && tet.path[3] == 7, "a Kuhn tetrahedron runs from corner 0 to corner 7");
    const float W0 = w[tet.path[0]], W1 = w[tet.path[1]], W2 = w[tet.path[2]], W3 = w[tet.path[3]];
    // Inverse masses are never negative: this sum is 0 exactly when the
    // WGSL's is.
    const float wend = W0 + W3, wmid = W1 + W2;
    if (wend + wmid == 0.0f) return;

    V3 A[3], E[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        A[j] = sub(p[tet.path[j + 1]], p[tet.path[j]]);
        E[tet.axis[j]] = A[j];
    }
    const V3 K[3] = { crossf(E[1], E[2]), crossf(E[2], E[0]), crossf(E[0], E[1]) };
    const V3 Kp[3] = { K[tet.axis[0]], K[tet.axis[1]], K[tet.axis[2]] };
    float n[3];
#pragma unroll
    for (int j = 0; j < 3; j++) n[j] = dotf(A[j], A[j]);
    const float d01 = dotf(A[0], A[1]), d12 = dotf(A[1], A[2]), d02 = dotf(A[0], A[2]);
    const float e2 = (n[0] + n[1]) + n[2];
    const float detE = dotf(E[0], K[0]);
    const bool dev = e2 > k.devMin;
    const float W01 = W0 + W1, W23 = W2 + W3;
    float ge = fmaf(W23, n[2], fmaf(wmid, n[1], W01 * n[0]));
    ge = fmaf(-2.0f, fmaf(W2, d12, W1 * d01), ge);
    // The K_j are the path's edges crossed pairwise, all with one sign, so
    // (Binet-Cauchy) |K_0|^2 = |A1|^2 |A2|^2 - (A1.A2)^2, K_0.K_1 = (A1.A2)(A0.A2)
    // - (A0.A1)|A2|^2, and so on round the path.
    const float m0 = fmaf(n[1], n[2], -(d12 * d12)), m1 = fmaf(n[2], n[0], -(d02 * d02));
    const float m2 = fmaf(n[0], n[1], -(d01 * d01));
    const float k01 = fmaf(d12, d02, -(d01 * n[2])), k12 = fmaf(d02, d01, -(d12 * n[0]));
    float gf = fmaf(W23, m2, fmaf(wmid, m1, W01 * m0));
    gf = fmaf(-2.0f, fmaf(W2, k12, W1 * k01), gf);
    V3 Q0[3];
#pragma unroll
    for (int j = 0; j < 3; j++) Q0[j] = sub(q[tet.path[j + 1]], q[tet.path[j]]);
    const float re = e2 - dotacc(A[2], Q0[2], dotacc(A[1], Q0[1], dotf(A[0], Q0[0])));
    const float rf = fmaf(3.0f, detE, -dotacc(Kp[2], Q0[2], dotacc(Kp[1], Q0[1], dotf(Kp[0], Q0[0]))));
    // e_i . f_i is det E at the path's ends and 2 det E inside it (E_a . K_a
    // = det E, E_a . K_b = 0 for a != b): Gef = det E * (W0 + 2 W1 + 2 W2 + W3).
    const float gef = detE * fmaf(2.0f, wmid, wend);
    const float st = 1.0f + k.relax;
    const float Ab = dev ? fmaf(ge, st, k.aDh2 * e2) : k.aD;
    const float Bb = dev ? gef * st : 0.0f;
    const float Db = fmaf(gf, st, k.aHh6);
    const float det = fmaf(Ab, Db, -(Bb * Bb));
    if (det * k.s6 < (dev ? k.detMin * e2 : 1e-20f)) return;
    const float P = dev ? -fmaf(k.relax, re, e2) : 0.0f;
    const float Q = -fmaf(k.relax, rf, detE - k.gh3);
    const float inv = 1.0f / det;
    const float u = fmaf(Db, P, -(Bb * Q)) * inv;
    const float v = fmaf(Ab, Q, -(Bb * P)) * inv;
    V3 G[3];
#pragma unroll
    for (int j = 0; j < 3; j++)
        G[j] = mk(fmaf(Kp[j].x, v, A[j].x * u), fmaf(Kp[j].y, v, A[j].y * u), fmaf(Kp[j].z, v, A[j].z * u));
    const V3 c[4] = { neg(G[0]), sub(G[0], G[1]), sub(G[1], G[2]), G[2] };
    const float W[4] = { W0, W1, W2, W3 };
#pragma unroll
    for (int j = 0; j < 4; j++) {
        V3 &pj = p[tet.path[j]];
        pj = mk(fmaf(c[j].x, W[j], pj.x), fmaf(c[j].y, W[j], pj.y), fmaf(c[j].z, W[j], pj.z));
    }
}

// Tile scheduling. Interleaved: workgroup b takes tiles b, b+G, ... .
// Chunked: the 8 XCDs each take one contiguous run of tiles (hardware deals
// workgroup b to XCD b % 8), so every pass of a substep leaves a region's
// nodes in the same XCD's L2.
struct Range {
    uint32_t first, end, stride;
};

__device__ __forceinline__ Range tile_range(uint32_t count, uint32_t xcd)
{
    const uint32_t G = gridDim.x, b = blockIdx.x;
    if (xcd && G % 8u == 0u) {
        const uint32_t lb = (b % 8u) * (G / 8u) + b / 8u;
        return Range{ (uint32_t)(((uint64_t)lb * count) / G), (uint32_t)(((uint64_t)(lb + 1) * count) / G), 1u };
    }
    return Range{ b, count, G };
}

// STRETCH_EDGES (stretch.js:5-7) with f32 lengths 1, sqrt(2), sqrt(3).
__device__ __forceinline__ float edge_len(int a, int b)
{
    const int d = ((a ^ b) & 1) + (((a ^ b) >> 1) & 1) + (((a ^ b) >> 2) & 1);
    return d == 1 ? 1.0f : d == 2 ? 1.41421353816986083984375f : 1.73205077648162841796875f;
}

// A cell's corners c and c|1 are x-neighbours, adjacent in every node plane,
// so the eight corners are four 8-byte pairs per plane. Lanes of a wave hold
// consecutive cells of one colour (every other cell along x), so their pairs
// tile a node row without gaps: each pair access is one contiguous span per
// row instead of two half-used strided ones. Pairs start on odd nodes for
// odd x-parity colours; the vector type is declared 4-byte aligned, which
// the hardware serves with dword alignment.
typedef float F2 __attribute__((ext_vector_type(2), aligned(4)));

__device__ __forceinline__ F2 ld2(const float *plane, uint32_t off)
{
    return *reinterpret_cast<const F2 *>(reinterpret_cast<const char *>(plane) + off);
}
__device__ __forceinline__ void st2(float *plane, uint32_t off, float a, float b)
{
    F2 v;
    v.x = a;
    v.y = b;
    *reinterpret_cast<F2 *>(reinterpret_cast<char *>(plane) + off) = v;
}

__device__ __forceinline__ V3 ld3(const float *x, const float *y, const float *z, uint32_t n)
{
    return mk(x[n], y[n], z[n]);
}

// ------------------------------------------------------ predict/finalize
// predict.wgsl:19-36 for one node with w != 0: symplectic Euler.
__device__ __forceinline__ void predict_node(V3 &p, V3 &q, V3 &v, float gx, float gy, float gz, float dt)
{
    v = mk(v.x + gx * dt, v.y + gy * dt, v.z + gz * dt);
    q = p;
    p = mk(p.x + v.x * dt, p.y + v.y * dt, p.z + v.z * dt);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_predict(NodeView n, const vs_scene_params *__restrict__ scenes, TileView tv)
{
    const Range r = tile_range(*tv.count, 0);
    for (uint32_t t = r.first; t < r.end; t += r.stride) {
        const vs_tile tile = tv.tiles[t];
        const vs_scene_params &sc = scenes[tile.scene];
        const float dt = sc.dt, gx = sc.gravity[0], gy = sc.gravity[1], gz = sc.gravity[2];
        for (uint32_t i = threadIdx.x; i < tile.count; i += BLOCK) {
            const uint32_t k = tile.begin + i;
            if (n.w[k] == 0.0f) continue;
            V3 v = ld3(n.vx, n.vy, n.vz, k), p = ld3(n.px, n.py, n.pz, k), q;
            predict_node(p, q, v, gx, gy, gz, dt);
            n.qx[k] = q.x;
            n.qy[k] = q.y;
            n.qz[k] = q.z;
            n.px[k] = p.x;
            n.py[k] = p.y;
            n.pz[k] = p.z;
            n.vx[k] = v.x;
            n.vy[k] = v.y;
            n.vz[k] = v.z;
        }
    }
}

// resolvePlane (finalize.wgsl:29-42) for an axis-aligned normal +-e_A.
template <int A, int SIGN>
__device__ __forceinline__ V3 plane(V3 p, V3 prev, float depth, float friction)
{
    if (!(depth > 0.0f)) return p;
    float c[3] = { p.x, p.y, p.z };
    c[A] = SIGN > 0 ? c[A] + depth : c[A] - depth;
    V3 q = mk(c[0], c[1], c[2]);
    const V3 delta = sub(q, prev);
    // tangent = delta - n * dot(delta, n): the A component cancels exactly.
    float tg[3] = { delta.x, delta.y, delta.z };
    tg[A] = 0.0f;
    const V3 tangent = mk(tg[0], tg[1], tg[2]);
    const float len = sqrtf(dot(tangent, tangent));
    if (len > 1e-9f) {
        const float lim = (friction * depth) / len;
        q = sub(q, mul(tangent, lim < 1.0f ? lim : 1.0f));
    }
    return q;
}

// The sphere collider (no reference counterpart): a node inside the ball is
// moved out along the radius to the surface (a node on the centre leaves
// along +y), with resolvePlane's position-level Coulomb friction in the
// tangent plane there (finalize.wgsl:29-42), in the oracle's order.
__device__ __noinline__ V3 sphere(V3 p, V3 prev, float cx, float cy, float cz, float r, float friction)
{
    const V3 d = mk(p.x - cx, p.y - cy, p.z - cz);
    const float len = sqrtf(dot(d, d));
    if (!(len < r)) return p;
    const V3 n = len > 1e-9f ? mk(d.x / len, d.y / len, d.z / len) : mk(0.0f, 1.0f, 0.0f);
    const float depth = r - len;
    V3 q = add(p, mul(n, depth));
    const V3 delta = sub(q, prev);
    const V3 tangent = sub(delta, mul(n, dot(delta, n)));
    const float tl = sqrtf(dot(tangent, tangent));
    if (tl > 1e-9f) {
        const float lim = (friction * depth) / tl;
        q = sub(q, mul(tangent, lim < 1.0f ? lim : 1.0f));
    }
    return q;
}

// The uniforms finalize reads for one body of one scene.
// `rate` = exp(-damping dt) / dt, the f32 quotient of the two factors, formed
// once per tile: the damped velocity is then one product a component.
struct Tank {
    float lx, ly, lz, hx, hz, fr, rate;
    float sx, sy, sz, sr;  // sphere collider; sr <= 0: none
};

__device__ __forceinline__ Tank tank_of(const vs_scene_params &sc, const vs_body &b)
{
    return Tank{ sc.tank_min[0], sc.tank_min[1], sc.tank_min[2], sc.tank_max[0], sc.tank_max[2], b.friction,
                 b.damp / sc.dt, sc.sphere[0], sc.sphere[1], sc.sphere[2], sc.sphere[3] };
}

// finalize.wgsl:44-74 for one node with w != 0: the five tank planes (no
// ceiling), the sphere collider when the scene has one, then the damped
// velocity (p - prev) / dt * exp(-damping dt), as (p - prev) * rate (the
// oracle's o_finalize order).
__device__ __forceinline__ void finalize_node(V3 &p, V3 prev, V3 &v, const Tank &k)
{
    p = plane<1, 1>(p, prev, k.ly - p.y, k.fr);
    p = plane<0, 1>(p, prev, k.lx - p.x, k.fr);
    p = plane<0, -1>(p, prev, p.x - k.hx, k.fr);
    p = plane<2, 1>(p, prev, k.lz - p.z, k.fr);
    p = plane<2, -1>(p, prev, p.z - k.hz, k.fr);
    if (k.sr > 0.0f) p = sphere(p, prev, k.sx, k.sy, k.sz, k.sr, k.fr);
    v = mk((p.x - prev.x) * k.rate, (p.y - prev.y) * k.rate, (p.z - prev.z) * k.rate);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_finalize(NodeView n, const vs_body *__restrict__ bodies,
                                                    const vs_scene_params *__restrict__ scenes, TileView tv)
{
    const Range r = tile_range(*tv.count, 0);
    for (uint32_t t = r.first; t < r.end; t += r.stride) {
        const vs_tile tile = tv.tiles[t];
        const Tank tk = tank_of(scenes[tile.scene], bodies[tile.body]);
        for (uint32_t i = threadIdx.x; i < tile.count; i += BLOCK) {
            const uint32_t k = tile.begin + i;
            if (n.w[k] == 0.0f) continue;
            const V3 prev = ld3(n.qx, n.qy, n.qz, k);
            V3 p = ld3(n.px, n.py, n.pz, k), v;
            finalize_node(p, prev, v, tk);
            n.px[k] = p.x;
            n.py[k] = p.y;
            n.pz[k] = p.z;
            n.vx[k] = v.x;
            n.vy[k] = v.y;
            n.vz[k] = v.z;
        }
    }
}

// The last finalize of one substep and the predict of the next, per node in
// one pass (nothing runs between them in Stepper.sequence): the node passes'
// arithmetic in their order. The velocity is not stored: every pass up to the
// next finalize, which overwrites it, reads positions only (predict reads it
// only at a frame's start, after a plain finalize). 52 B a node instead of
// 52 + 64.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_finalize_predict(NodeView n, const vs_body *__restrict__ bodies,
                                                            const vs_scene_params *__restrict__ scenes, TileView tv)
{
    const Range r = tile_range(*tv.count, 0);
    for (uint32_t t = r.first; t < r.end; t += r.stride) {
        const vs_tile tile = tv.tiles[t];
        const vs_scene_params &sc = scenes[tile.scene];
        const Tank tk = tank_of(sc, bodies[tile.body]);
        const float dt = sc.dt, gx = sc.gravity[0], gy = sc.gravity[1], gz = sc.gravity[2];
        for (uint32_t i = threadIdx.x; i < tile.count; i += BLOCK) {
            const uint32_t k = tile.begin + i;
            if (n.w[k] == 0.0f) continue;
            V3 q = ld3(n.qx, n.qy, n.qz, k), p = ld3(n.px, n.py, n.pz, k), v;
            finalize_node(p, q, v, tk);
            predict_node(p, q, v, gx, gy, gz, dt);
            n.qx[k] = q.x;
            n.qy[k] = q.y;
            n.qz[k] = q.z;
            n.px[k] = p.x;
            n.py[k] = p.y;
            n.pz[k] = p.z;
        }
    }
}

// ------------------------------------------------------------- elastic
// The constants a cell projection reads, per body and dt (project.wgsl:35-60,
// xpbdCouple.wgsl:58-71).
struct CellConsts {
    TetConsts tet;
    float h, maxStretch;
};

// Per body and dt, in the oracle's order: s = 1/h and its powers rounded
// step by step, alphas over dt^2, the thresholds scaled by h^2.
__device__ __forceinline__ CellConsts cell_consts(const vs_body &bd, float dt)
{
    const float invDt2 = 1.0f / (dt * dt);
    const float h = bd.h, s = 1.0f / h, s2 = s * s, s3 = s2 * s, h2 = h * h, h3 = h2 * h;
    const float aD = bd.alpha_dev * invDt2, aH = bd.alpha_hyd * invDt2;
    return CellConsts{ TetConsts{ s3 * s3, bd.gamma * h3, aD, aH * (h3 * h3), aD * h2, bd.relaxation / dt,
                                  1e-18f * h2, 1e-20f * h2 },
                       h, bd.max_stretch };
}

// One cell: the six Kuhn tetrahedra in order, then the tensile limit
// (project.wgsl:162-179) when maxStretch > 1. `limit_only` skips the tets
// (the extra stretch sweeps of elasticCells.js:197-214). STRETCH false
// leaves the tensile limit out of the code: launches whose bodies have no
// maxStretch above 1 (the limit would never run) take that form, half the
// size, which runs the colour passes 5-7% faster.
template <bool STRETCH>
__device__ __forceinline__ void solve_cell(V3 (&p)[8], const V3 (&q)[8], const float (&w)[8], const CellConsts &k,
                                           uint32_t limit_only)
{
    if (!limit_only) {
        project_tet<0>(p, q, w, k.tet);
        project_tet<1>(p, q, w, k.tet);
        project_tet<2>(p, q, w, k.tet);
        project_tet<3>(p, q, w, k.tet);
        project_tet<4>(p, q, w, k.tet);
        project_tet<5>(p, q, w, k.tet);
    }
    if (STRETCH && k.maxStretch > 1.0f) {
        for (int sweep = 0; sweep < 2; sweep++) {
#pragma unroll
            for (int a = 0; a < 8; a++) {
#pragma unroll
                for (int b = a + 1; b < 8; b++) {
                    const V3 delta = sub(p[b], p[a]);
                    const float len = sqrtf(dot(delta, delta));
                    const float maximum = (edge_len(a, b) * k.h) * k.maxStretch;
                    const float weight = w[a] + w[b];
                    if (len > maximum && weight > 0.0f) {
                        const V3 corr = mul(delta, (len - maximum) / (len * weight));
                        p[a] = add(p[a], mul(corr, w[a]));
                        p[b] = sub(p[b], mul(corr, w[b]));
                    }
                }
            }
        }
    }
}

// One colour pass over a list of cells: one lane per cell, its eight corners
// gathered as four x-pairs per plane, solved, and scattered back.
template <int BLOCK, bool STRETCH>
__global__ __launch_bounds__(BLOCK) void k_elastic(NodeView n, const vs_body *__restrict__ bodies,
                                                   const vs_scene_params *__restrict__ scenes,
                                                   const uint32_t *__restrict__ cells, TileView tv,
                                                   uint32_t limit_only, uint32_t xcd)
{
    const Range r = tile_range(*tv.count, xcd);
    for (uint32_t t = r.first; t < r.end; t += r.stride) {
        const vs_tile tile = tv.tiles[t];
        const vs_body bd = bodies[tile.body];
        const vs_scene_params &sc = scenes[bd.scene];
        if (limit_only && !(bd.max_stretch > 1.0f)) continue;
        const CellConsts kc = cell_consts(bd, sc.dt);
        const uint32_t cdx = bd.cell_dims[0], cdy = bd.cell_dims[1];
        const uint32_t ndx = bd.node_dims[0], ndxy = bd.node_dims[0] * bd.node_dims[1];

        for (uint32_t i = threadIdx.x; i < tile.count; i += BLOCK) {
            const uint32_t lin = cells[tile.begin + i];
            const uint32_t cx = lin % cdx, cy = (lin / cdx) % cdy, cz = lin / (cdx * cdy);
            const uint32_t n0 = bd.node_offset + cx + ndx * cy + ndxy * cz;
            // Byte offsets of the four x-pairs: pair k holds corners 2k, 2k+1.
            uint32_t off[4];
            V3 p[8], q[8];
            float w[8];
#pragma unroll
            for (int k = 0; k < 4; k++) off[k] = 4u * (n0 + (k & 1 ? ndx : 0u) + (k & 2 ? ndxy : 0u));
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const F2 x = ld2(n.px, off[k]), y = ld2(n.py, off[k]), z = ld2(n.pz, off[k]);
                const F2 m = ld2(n.w, off[k]);
                const F2 u = ld2(n.qx, off[k]), v = ld2(n.qy, off[k]), s = ld2(n.qz, off[k]);
                p[2 * k] = mk(x.x, y.x, z.x);
                p[2 * k + 1] = mk(x.y, y.y, z.y);
                w[2 * k] = m.x;
                w[2 * k + 1] = m.y;
                q[2 * k] = mk(u.x, v.x, s.x);
                q[2 * k + 1] = mk(u.y, v.y, s.y);
            }
            solve_cell<STRETCH>(p, q, w, kc, limit_only);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                st2(n.px, off[k], p[2 * k].x, p[2 * k + 1].x);
                st2(n.py, off[k], p[2 * k].y, p[2 * k + 1].y);
                st2(n.pz, off[k], p[2 * k].z, p[2 * k + 1].z);
            }
        }
    }
}

// Two colours in one pass, over whole x-rows of cells of a dense lattice.
//
// Colours 2c and 2c+1 (c = 0..3) are the cells with y parity c&1 and z parity
// c>>1, even x then odd x: the cells of one such row y, z. Cells of one colour
// share no corner, and a row's odd cell shares its corners only with the two
// even cells beside it in the same row, so solving every row's even cells and
// then its odd cells is exactly the two colour passes in order.
//
// Lane m of a row holds the node pair x = 2m, 2m+1 of the row's four node
// rows (y or y+1, z or z+1): an 8-byte access, the row's lanes together one
// contiguous span per node row. It solves cell 2m (colour 2c), hands node 2m
// to lane m-1 through LDS, takes node 2m+2 from lane m+1, solves cell 2m+1
// (colour 2c+1) and hands node 2m+2 back. Every node of the lattice is read
// and written once per pass instead of once per colour. Rows never span
// workgroups: a tile is `count` whole rows starting at row `begin` of the
// class, rows of L = cdx/2 + 1 lanes laid end to end.
//
// `node_mode` (class 3, the last colours, in a scene flagged VS_SCENE_FOLD):
// after the colours, each node the lane writes is finalized (1, 2) and then
// predicted for the next substep (1), as the node passes would do next.
// `node_mode` 3 (class 0, the first colours, at a frame's first substep, in
// a scene flagged VS_SCENE_FOLD_PREDICT): each node the lane holds is
// predicted before the colours, as the predict pass would do first.
//
// A tile with VS_NODE_TILE set in its count holds nodes instead of rows:
// entries begin .. begin + count - 1 of the body's Leftover list (the nodes
// of a folded body the class's rows leave out). It gets the node arithmetic
// of `node_mode` alone (nothing when the mode is 0).
// Four waves per SIMD: at most 128 VGPRs.
// MODES: the node-pass code the form holds: 0 none (launches with
// `node_mode` 0; the colour-pair kernel less than half the size), 1 the
// finalize and predict of modes 1 and 2, 3 the predict of mode 3.
template <int BLOCK, bool STRETCH, int MODES>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_elastic_rows(NodeView n, const vs_body *__restrict__ bodies,
                                                        const vs_scene_params *__restrict__ scenes, TileView tv,
                                                        uint32_t cls, uint32_t limit_only, uint32_t mode,
                                                        uint32_t xcd)
{
    const uint32_t node_mode = MODES == 0 ? 0u : MODES == 3 ? 3u : mode;
    // Per lane: start-of-substep position (12) and inverse mass (4) of its
    // node 2m in the four node rows, then that node's position (12).
    __shared__ float lds[28 * BLOCK];
    const uint32_t py = cls & 1u, pz = cls >> 1;
    const uint32_t tid = threadIdx.x;
    const Range r = tile_range(*tv.count, xcd);
    for (uint32_t t = r.first; t < r.end; t += r.stride) {
        const vs_tile tile = tv.tiles[t];
        const vs_body bd = bodies[tile.body];
        const vs_scene_params &sc = scenes[bd.scene];
        if (tile.count & VS_NODE_TILE) {
            const bool fin = (node_mode == 1u || node_mode == 2u) && (sc.flags & VS_SCENE_FOLD) != 0u;
            const bool pre = node_mode == 3u && (sc.flags & VS_SCENE_FOLD_PREDICT) != 0u;
            if (limit_only || !(fin || pre)) continue;
            const uint32_t count = tile.count & ~VS_NODE_TILE;
            const Leftover rest(bd.cell_dims, cls);
            const Tank tk = tank_of(sc, bd);
            const float dt = sc.dt, gx = sc.gravity[0], gy = sc.gravity[1], gz = sc.gravity[2];
            for (uint32_t i = tid; i < count; i += BLOCK) {
                const uint32_t k = bd.node_offset + rest.node(tile.begin + i);
                if (n.w[k] == 0.0f) continue;
                V3 p = ld3(n.px, n.py, n.pz, k), q, v;
                if (pre) {
                    // k_predict's arithmetic; the velocity is dead in the frame.
                    v = ld3(n.vx, n.vy, n.vz, k);
                    predict_node(p, q, v, gx, gy, gz, dt);
                } else {
                    q = ld3(n.qx, n.qy, n.qz, k);
                    finalize_node(p, q, v, tk);
                    if (node_mode == 1u) predict_node(p, q, v, gx, gy, gz, dt);
                }
                if (pre || node_mode == 1u) {
                    n.qx[k] = q.x;
                    n.qy[k] = q.y;
                    n.qz[k] = q.z;
                } else {
                    n.vx[k] = v.x;
                    n.vy[k] = v.y;
                    n.vz[k] = v.z;
                }
                n.px[k] = p.x;
                n.py[k] = p.y;
                n.pz[k] = p.z;
            }
            continue;
        }
        if (limit_only && !(bd.max_stretch > 1.0f)) continue;
        const CellConsts kc = cell_consts(bd, sc.dt);
        const uint32_t cdx = bd.cell_dims[0];
        const uint32_t ndx = bd.node_dims[0], ndxy = bd.node_dims[0] * bd.node_dims[1];
        const uint32_t L = cdx / 2u + 1u;
        const uint32_t nyr = (bd.cell_dims[1] + 1u - py) / 2u;
        const uint32_t row = tid / L, m = tid - row * L;
        const bool on = row < tile.count;
        const uint32_t x0 = 2u * m;
        const bool pair = x0 + 1u <= cdx;      // node x0 + 1 is in the row
        const bool cellA = on && x0 < cdx;     // cell x0, colour 2c
        const bool cellB = on && x0 + 1u < cdx;  // cell x0 + 1, colour 2c + 1
        uint32_t off[4] = { 0u, 0u, 0u, 0u };
        V3 p[8], q[8];
        float w[8];
        if (on) {
            const uint32_t rho = tile.begin + row;
            const uint32_t y = py + 2u * (rho % nyr), z = pz + 2u * (rho / nyr);
            const uint32_t n0 = bd.node_offset + x0 + ndx * y + ndxy * z;
#pragma unroll
            for (int k = 0; k < 4; k++) off[k] = 4u * (n0 + (k & 1 ? ndx : 0u) + (k & 2 ? ndxy : 0u));
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (on && pair) {
                const F2 x = ld2(n.px, off[k]), y = ld2(n.py, off[k]), z = ld2(n.pz, off[k]);
                const F2 mm = ld2(n.w, off[k]);
                const F2 u = ld2(n.qx, off[k]), v = ld2(n.qy, off[k]), s = ld2(n.qz, off[k]);
                p[2 * k] = mk(x.x, y.x, z.x);
                p[2 * k + 1] = mk(x.y, y.y, z.y);
                w[2 * k] = mm.x;
                w[2 * k + 1] = mm.y;
                q[2 * k] = mk(u.x, v.x, s.x);
                q[2 * k + 1] = mk(u.y, v.y, s.y);
            } else if (on) {
                const uint32_t e = off[k] / 4u;
                p[2 * k] = ld3(n.px, n.py, n.pz, e);
                q[2 * k] = ld3(n.qx, n.qy, n.qz, e);
                w[2 * k] = n.w[e];
                p[2 * k + 1] = q[2 * k + 1] = mk(0.f, 0.f, 0.f);
                w[2 * k + 1] = 0.f;
            } else {
                p[2 * k] = q[2 * k] = p[2 * k + 1] = q[2 * k + 1] = mk(0.f, 0.f, 0.f);
                w[2 * k] = w[2 * k + 1] = 0.f;
            }
        }
        const bool pre = node_mode == 3u && (sc.flags & VS_SCENE_FOLD_PREDICT) != 0u;
        if (on && pre) {
            // predict.wgsl:19-36 per node with w != 0 (k_predict's
            // arithmetic); the velocity it leaves is dead inside the frame.
            const float dt = sc.dt, gx = sc.gravity[0], gy = sc.gravity[1], gz = sc.gravity[2];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                V3 v0, v1 = mk(0.f, 0.f, 0.f);
                if (pair) {
                    const F2 a = ld2(n.vx, off[k]), b = ld2(n.vy, off[k]), c = ld2(n.vz, off[k]);
                    v0 = mk(a.x, b.x, c.x);
                    v1 = mk(a.y, b.y, c.y);
                } else {
                    v0 = ld3(n.vx, n.vy, n.vz, off[k] / 4u);
                }
                if (w[2 * k] != 0.0f) predict_node(p[2 * k], q[2 * k], v0, gx, gy, gz, dt);
                if (pair && w[2 * k + 1] != 0.0f) predict_node(p[2 * k + 1], q[2 * k + 1], v1, gx, gy, gz, dt);
            }
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            lds[(3 * k + 0) * BLOCK + tid] = q[2 * k].x;
            lds[(3 * k + 1) * BLOCK + tid] = q[2 * k].y;
            lds[(3 * k + 2) * BLOCK + tid] = q[2 * k].z;
            lds[(12 + k) * BLOCK + tid] = w[2 * k];
        }
        if (cellA) solve_cell<STRETCH>(p, q, w, kc, limit_only);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            lds[(16 + 3 * k + 0) * BLOCK + tid] = p[2 * k].x;
            lds[(16 + 3 * k + 1) * BLOCK + tid] = p[2 * k].y;
            lds[(16 + 3 * k + 2) * BLOCK + tid] = p[2 * k].z;
        }
        __syncthreads();
        // Cell x0 + 1: its low corners are this lane's node x0 + 1, its high
        // corners lane m+1's node x0 + 2 (same row whenever the cell exists).
#pragma unroll
        for (int k = 0; k < 4; k++) {
            p[2 * k] = p[2 * k + 1];
            q[2 * k] = q[2 * k + 1];
            w[2 * k] = w[2 * k + 1];
        }
        if (cellB) {
            const uint32_t nb = tid + 1u;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                q[2 * k + 1] = mk(lds[(3 * k + 0) * BLOCK + nb], lds[(3 * k + 1) * BLOCK + nb], lds[(3 * k + 2) * BLOCK + nb]);
                w[2 * k + 1] = lds[(12 + k) * BLOCK + nb];
                p[2 * k + 1] = mk(lds[(16 + 3 * k + 0) * BLOCK + nb], lds[(16 + 3 * k + 1) * BLOCK + nb],
                                  lds[(16 + 3 * k + 2) * BLOCK + nb]);
            }
            solve_cell<STRETCH>(p, q, w, kc, limit_only);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                lds[(16 + 3 * k + 0) * BLOCK + nb] = p[2 * k + 1].x;
                lds[(16 + 3 * k + 1) * BLOCK + nb] = p[2 * k + 1].y;
                lds[(16 + 3 * k + 2) * BLOCK + nb] = p[2 * k + 1].z;
            }
        }
        __syncthreads();
        // Node x0 is final in this lane's slot: lane m-1 wrote it back after
        // cell x0 - 1, or, at the row's start, this lane's own cell x0 did.
        // Node x0 + 1 is this lane's low corners.
        const bool fold = (node_mode == 1u || node_mode == 2u) && (sc.flags & VS_SCENE_FOLD) != 0u;
        if (on && !fold) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float lx = lds[(16 + 3 * k + 0) * BLOCK + tid];
                const float ly = lds[(16 + 3 * k + 1) * BLOCK + tid];
                const float lz = lds[(16 + 3 * k + 2) * BLOCK + tid];
                if (pair) {
                    st2(n.px, off[k], lx, p[2 * k].x);
                    st2(n.py, off[k], ly, p[2 * k].y);
                    st2(n.pz, off[k], lz, p[2 * k].z);
                } else {
                    const uint32_t e = off[k] / 4u;
                    n.px[e] = lx;
                    n.py[e] = ly;
                    n.pz[e] = lz;
                }
                if (pre) {
                    // The start-of-substep positions the predict set: node x0's
                    // in this lane's slot, node x0 + 1's in the low corners.
                    const float ux = lds[(3 * k + 0) * BLOCK + tid];
                    const float uy = lds[(3 * k + 1) * BLOCK + tid];
                    const float uz = lds[(3 * k + 2) * BLOCK + tid];
                    if (pair) {
                        st2(n.qx, off[k], ux, q[2 * k].x);
                        st2(n.qy, off[k], uy, q[2 * k].y);
                        st2(n.qz, off[k], uz, q[2 * k].z);
                    } else {
                        const uint32_t e = off[k] / 4u;
                        n.qx[e] = ux;
                        n.qy[e] = uy;
                        n.qz[e] = uz;
                    }
                }
            }
        } else if (on) {
            // finalize.wgsl:44-74 and then predict.wgsl:19-36 per node with
            // w != 0, as k_finalize_predict runs them; a node with w = 0
            // keeps its prev and velocity (its prev is rewritten unchanged).
            // Inside a frame the velocity is not stored: the next finalize
            // overwrites it before anything reads it.
            const Tank tk = tank_of(sc, bd);
            const float dt = sc.dt, gx = sc.gravity[0], gy = sc.gravity[1], gz = sc.gravity[2];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                V3 lo = mk(lds[(16 + 3 * k + 0) * BLOCK + tid], lds[(16 + 3 * k + 1) * BLOCK + tid],
                           lds[(16 + 3 * k + 2) * BLOCK + tid]);
                V3 qlo = mk(lds[(3 * k + 0) * BLOCK + tid], lds[(3 * k + 1) * BLOCK + tid],
                            lds[(3 * k + 2) * BLOCK + tid]);
                const float wlo = lds[(12 + k) * BLOCK + tid];
                V3 vlo = mk(0.f, 0.f, 0.f), vhi = mk(0.f, 0.f, 0.f);
                V3 &hi = p[2 * k], &qhi = q[2 * k];
                const bool mlo = wlo != 0.0f, mhi = pair && w[2 * k] != 0.0f;
                if (mlo) {
                    finalize_node(lo, qlo, vlo, tk);
                    if (node_mode == 1u) predict_node(lo, qlo, vlo, gx, gy, gz, dt);
                }
                if (mhi) {
                    finalize_node(hi, qhi, vhi, tk);
                    if (node_mode == 1u) predict_node(hi, qhi, vhi, gx, gy, gz, dt);
                }
                const uint32_t e = off[k] / 4u;
                if (pair) {
                    st2(n.px, off[k], lo.x, hi.x);
                    st2(n.py, off[k], lo.y, hi.y);
                    st2(n.pz, off[k], lo.z, hi.z);
                    if (node_mode == 1u) {
                        st2(n.qx, off[k], qlo.x, qhi.x);
                        st2(n.qy, off[k], qlo.y, qhi.y);
                        st2(n.qz, off[k], qlo.z, qhi.z);
                    }
                    if (node_mode == 2u && mlo && mhi) {
                        st2(n.vx, off[k], vlo.x, vhi.x);
                        st2(n.vy, off[k], vlo.y, vhi.y);
                        st2(n.vz, off[k], vlo.z, vhi.z);
                    } else if (node_mode == 2u && mhi) {
                        n.vx[e + 1] = vhi.x;
                        n.vy[e + 1] = vhi.y;
                        n.vz[e + 1] = vhi.z;
                    }
                } else {
                    n.px[e] = lo.x;
                    n.py[e] = lo.y;
                    n.pz[e] = lo.z;
                    if (node_mode == 1u) {
                        n.qx[e] = qlo.x;
                        n.qy[e] = qlo.y;
                        n.qz[e] = qlo.z;
                    }
                }
                if (node_mode == 2u && mlo && !(pair && mhi)) {
                    n.vx[e] = vlo.x;
                    n.vy[e] = vlo.y;
                    n.vz[e] = vlo.z;
                }
            }
        }
        __syncthreads();
    }
}

// gridCoord (contacts.wgsl:84-87)
__device__ __forceinline__ void grid_coord(const vs_scene_params &sc, V3 p, uint32_t &x, uint32_t &y, uint32_t &z)
{
    const float c[3] = { p.x, p.y, p.z };
    uint32_t o[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float last = (float)sc.grid_dims[a] - 1.0f;
        float g = floorf((c[a] - sc.grid_origin[a]) / sc.bucket_size);
        g = g < 0.0f ? 0.0f : g;
        g = g > last ? last : g;
        o[a] = (uint32_t)g;
    }
    x = o[0];
    y = o[1];
    z = o[2];
}

// Bin (bin.wgsl:17) in two parts, around the caller's atomicAdd on the
// bucket's counter (the arrival rank): bin_key writes record s's bucket key
// and the payloads beside it (its position and radius; its node's prev
// position and inverse mass, `prev`, for collide) and returns the bucket;
// bin_place stores the rank and puts the record in its slot, extension slot
// or the overflow list.
__device__ __forceinline__ uint32_t bin_key(const vs_scene_params &sc, SurfaceView sv, uint32_t s, V3 p, float4 prev)
{
    uint32_t x, y, z;
    grid_coord(sc, p, x, y, z);
    const uint32_t b = x + sc.grid_dims[0] * (y + sc.grid_dims[1] * z);
    sv.key[s] = b;
    if (sv.pay) sv.pay[s] = make_float4(p.x, p.y, p.z, sv.radius[s]);
    if (sv.prev) sv.prev[s] = prev;
    return b;
}

__device__ __forceinline__ void bin_place(const vs_scene_params &sc, uint32_t scene, SurfaceView sv, GridView g,
                                          uint32_t s, uint32_t b, uint32_t slot)
{
    const uint32_t slots = sc.bucket_slots, gb = sc.grid_base;
    sv.rank[s] = slot;
    if (slot < slots) g.slots[(size_t)(gb + b) * slots + slot] = s;
    else if (slot - slots < g.ext_slots) g.ext[(size_t)(gb + b) * g.ext_slots + (slot - slots)] = s;
    else if (g.ovf) {
        const uint32_t at = atomicAdd(g.ovf_count, 1u);
        if (at < g.ovf_cap) g.ovf[at] = make_uint2(s, gb + b);
    }
    if (slot == slots && g.ovb) {
        // This arrival made the bucket overfull: list the bucket once.
        const uint32_t at = atomicAdd(&g.ovf_count[1], 1u);
        if (at < g.ovb_cap) g.ovb[at] = make_uint2(gb + b, scene);
        else atomicOr(&g.ovf_count[2], 1u);
    }
}

// The colour passes of whole small dense bodies, one workgroup a body: its
// nodes' positions, start-of-substep positions and inverse masses are
// loaded into LDS once, colours 0..7 run from LDS in order (a barrier
// between colours), and the positions go back once. Colour k's cells are
// the body's cells of x, y, z parities (k & 1, (k >> 1) & 1, k >> 2)
// (cellColor, lattice.js:44-46); cells of one colour share no corner, so
// their order within the colour does not matter. Bodies share no node, so
// a body's eight colours in its own launch give the colour-by-colour
// launches' results (elasticCells.js:210-214 order), bit for bit.
// `limit_only` runs the tensile-limit sweep alone (elasticCells.js:197-214).
// In a scene flagged VS_SCENE_FOLD_PREDICT (no drive between predict and the
// first colour), `node_mode` 3 predicts the body's nodes before the colours
// (a frame's first substep) and `node_mode` 1 runs the previous substep's
// finalize and then this one's predict (k_finalize_predict's arithmetic):
// nothing runs between them but this pass's load.
// `contact` bit 1 runs the previous substep's settle on the body's contact
// records (tile.begin .. + tile.count) against grid `g`, between the load and
// the node passes; bit 2 bins the records against grid `gn` after the
// colours, from the positions in LDS (k_bin's arithmetic; nothing runs
// between the sweep and the bin). With both, `g` and `gn` are different
// grids (bucket counters, summaries and list counters by substep parity),
// since one body's bin may run before another body's settle.
constexpr uint32_t BODY_SWEEP_NODES = 11 * 11 * 11;  // bodies of up to 10^3 cells

template <int BLOCK, bool STRETCH>
__global__ __launch_bounds__(BLOCK) void k_elastic_bodies(NodeView n, const vs_body *__restrict__ bodies,
                                                          const vs_scene_params *__restrict__ scenes, TileView tv,
                                                          uint32_t limit_only, uint32_t node_mode, SurfaceView sv,
                                                          GridView g, GridView gn, uint32_t contact)
{
    __shared__ float sp[3][BODY_SWEEP_NODES], sq[3][BODY_SWEEP_NODES], sw[BODY_SWEEP_NODES];
    const uint32_t tid = threadIdx.x;
    const Range r = tile_range(*tv.count, 0u);
    for (uint32_t t = r.first; t < r.end; t += r.stride) {
        const vs_tile tile = tv.tiles[t];
        const vs_body bd = bodies[tile.body];
        const vs_scene_params &sc = scenes[bd.scene];
        if (limit_only && !(bd.max_stretch > 1.0f)) continue;
        const uint32_t nx = bd.node_dims[0], nxy = bd.node_dims[0] * bd.node_dims[1];
        const uint32_t N = nxy * bd.node_dims[2];
        if (N > BODY_SWEEP_NODES) continue;  // the host lists no such body
        const uint32_t o = bd.node_offset;
        const uint32_t mode = !limit_only && (sc.flags & VS_SCENE_FOLD_PREDICT) ? node_mode : 0u;
        const Tank tk = tank_of(sc, bd);
        const float dt = sc.dt, gx = sc.gravity[0], gy = sc.gravity[1], gz = sc.gravity[2];
        if (!(contact & 1u)) {
            for (uint32_t i = tid; i < N; i += BLOCK) {
                V3 p = ld3(n.px, n.py, n.pz, o + i), q = ld3(n.qx, n.qy, n.qz, o + i);
                const float w = n.w[o + i];
                if (mode != 0u && w != 0.0f) {
                    V3 v;
                    if (mode == 3u) v = ld3(n.vx, n.vy, n.vz, o + i);
                    else finalize_node(p, q, v, tk);
                    predict_node(p, q, v, gx, gy, gz, dt);
                }
                sp[0][i] = p.x;
                sp[1][i] = p.y;
                sp[2][i] = p.z;
                sq[0][i] = q.x;
                sq[1][i] = q.y;
                sq[2][i] = q.z;
                sw[i] = w;
            }
        } else {
            // The previous substep's settle (k_settle's arithmetic) on the
            // body's contact records tile.begin .. + tile.count, between the
            // load and the node passes: each record's bucket counter and
            // summary cleared (arrival rank 0), each record with hits moved
            // by its corrections' sum times gain / hits.
            for (uint32_t i = tid; i < N; i += BLOCK) {
                sp[0][i] = n.px[o + i];
                sp[1][i] = n.py[o + i];
                sp[2][i] = n.pz[o + i];
                sq[0][i] = n.qx[o + i];
                sq[1][i] = n.qy[o + i];
                sq[2][i] = n.qz[o + i];
                sw[i] = n.w[o + i];
            }
            if (g.ovf_count && blockIdx.x == 0 && tid == 0) {
                g.ovf_count[0] = 0u;
                if (g.ovb) g.ovf_count[1] = g.ovf_count[2] = 0u;
            }
            __syncthreads();
            const float gain = sc.gain;
            const uint32_t gb = sc.grid_base;
            for (uint32_t i = tid; i < tile.count; i += BLOCK) {
                const uint32_t s = tile.begin + i;
                if (sv.rank[s] == 0u) {
                    g.count[gb + sv.key[s]] = 0u;
                    if (g.summary) g.summary[gb + sv.key[s]] = 0u;
                }
                const float h = sv.hits[s];
                if (h > 0.0f) {
                    const uint32_t k = sv.node[s] - o;
                    const float f = gain / h;
                    sp[0][k] = sp[0][k] + sv.dx[s] * f;
                    sp[1][k] = sp[1][k] + sv.dy[s] * f;
                    sp[2][k] = sp[2][k] + sv.dz[s] * f;
                }
            }
            __syncthreads();
            if (mode != 0u) {
                for (uint32_t i = tid; i < N; i += BLOCK) {
                    if (sw[i] == 0.0f) continue;
                    V3 p = mk(sp[0][i], sp[1][i], sp[2][i]), q = mk(sq[0][i], sq[1][i], sq[2][i]), v;
                    if (mode == 3u) v = ld3(n.vx, n.vy, n.vz, o + i);
                    else finalize_node(p, q, v, tk);
                    predict_node(p, q, v, gx, gy, gz, dt);
                    sp[0][i] = p.x;
                    sp[1][i] = p.y;
                    sp[2][i] = p.z;
                    sq[0][i] = q.x;
                    sq[1][i] = q.y;
                    sq[2][i] = q.z;
                }
            }
        }
        __syncthreads();
        const CellConsts kc = cell_consts(bd, sc.dt);
        for (uint32_t col = 0; col < 8u; col++) {
            const uint32_t ox = col & 1u, oy = (col >> 1) & 1u, oz = col >> 2;
            const uint32_t mx = (bd.cell_dims[0] + 1u - ox) / 2u, my = (bd.cell_dims[1] + 1u - oy) / 2u,
                           mz = (bd.cell_dims[2] + 1u - oz) / 2u;
            const uint32_t cells = mx * my * mz;
            for (uint32_t i = tid; i < cells; i += BLOCK) {
                const uint32_t cx = ox + 2u * (i % mx), cy = oy + 2u * ((i / mx) % my), cz = oz + 2u * (i / (mx * my));
                const uint32_t n0 = cx + nx * cy + nxy * cz;
                V3 p[8], q[8];
                float w[8];
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    const uint32_t k = n0 + (uint32_t)(c & 1) + (c & 2 ? nx : 0u) + (c & 4 ? nxy : 0u);
                    p[c] = mk(sp[0][k], sp[1][k], sp[2][k]);
                    q[c] = mk(sq[0][k], sq[1][k], sq[2][k]);
                    w[c] = sw[k];
                }
                solve_cell<STRETCH>(p, q, w, kc, limit_only);
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    const uint32_t k = n0 + (uint32_t)(c & 1) + (c & 2 ? nx : 0u) + (c & 4 ? nxy : 0u);
                    sp[0][k] = p[c].x;
                    sp[1][k] = p[c].y;
                    sp[2][k] = p[c].z;
                }
            }
            __syncthreads();
        }
        for (uint32_t i = tid; i < N; i += BLOCK) {
            n.px[o + i] = sp[0][i];
            n.py[o + i] = sp[1][i];
            n.pz[o + i] = sp[2][i];
            if (mode != 0u) {
                n.qx[o + i] = sq[0][i];
                n.qy[o + i] = sq[1][i];
                n.qz[o + i] = sq[2][i];
            }
        }
        if (contact & 2u)
            for (uint32_t i = tid; i < tile.count; i += BLOCK) {
                const uint32_t s = tile.begin + i, k = sv.node[s] - o;
                const uint32_t b = bin_key(sc, sv, s, mk(sp[0][k], sp[1][k], sp[2][k]),
                                           make_float4(sq[0][k], sq[1][k], sq[2][k], sw[k]));
                bin_place(sc, bd.scene, sv, gn, s, b, atomicAdd(&gn.count[sc.grid_base + b], 1u));
            }
        __syncthreads();  // the next body reuses the LDS
    }
}

// -------------------------------------------------------------- contact
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_bin(NodeView n, const vs_scene_params *__restrict__ scenes, SurfaceView sv,
                                               GridView g, TileView tv)
{
    const Range r = tile_range(*tv.count, 0);
    for (uint32_t t = r.first; t < r.end; t += r.stride) {
        const vs_tile tile = tv.tiles[t];
        const vs_scene_params &sc = scenes[tile.scene];
        for (uint32_t i = threadIdx.x; i < tile.count; i += BLOCK) {
            const uint32_t s = tile.begin + i;
            const uint32_t node = sv.node[s];
            const float4 prev =
                sv.prev ? make_float4(n.qx[node], n.qy[node], n.qz[node], n.w[node]) : make_float4(0.f, 0.f, 0.f, 0.f);
            const uint32_t b = bin_key(sc, sv, s, ld3(n.px, n.py, n.pz, node), prev);
            bin_place(sc, tile.scene, sv, g, s, b, atomicAdd(&g.count[sc.grid_base + b], 1u));
        }
    }
}

// Overfull buckets keep the lowest `slots` record ids of the scene's records
// in them, ascending (sort.wgsl:27-37): the reference's scan from the scene's
// first record upwards, taken by a whole wave for every overfull bucket its
// lanes hold at once (lane j's bucket b, for each bit j of `pending`), 16 x 64
// keys a round (16 loads in flight a lane). Lane l reads record o0 + 64u + l;
// per bucket a ballot marks the matches and each match's place is the count
// of matches before it in record order, so the slots and their order are
// those of a one-record-at-a-time scan. A bucket leaves the scan once it holds
// `slots`. `grid` is the scene's slot array; every lane of the wave calls this
// with the same `pending` (the loop bounds are wave-uniform).
__device__ __forceinline__ void overflow_scan(const vs_scene_params &sc, const uint32_t *__restrict__ key,
                                              uint32_t *grid, uint32_t b, uint64_t pending, uint32_t slots)
{
    constexpr uint32_t U = 16;
    const uint32_t lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t end = sc.surf_base + sc.surf_count;
    uint32_t kept = 0;  // lane j: bucket b's matches so far
    for (uint32_t o0 = sc.surf_base; o0 < end && pending; o0 += 64u * U) {
        uint32_t k[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t o = o0 + 64u * u + lane;
            k[u] = o < end ? key[o] : ~0u;  // bucket keys are grid indices, never ~0
        }
        for (uint64_t left = pending; left; left &= left - 1ull) {
            const int j = __builtin_ctzll(left);
            const uint32_t bb = (uint32_t)__shfl((int)b, j);
            uint32_t kj = (uint32_t)__shfl((int)kept, j);
            uint32_t *base = grid + (size_t)bb * slots;
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                const bool m = k[u] == bb;
                const uint64_t mask = __ballot(m);
                const uint32_t pos = kj + (uint32_t)__popcll(mask & below);
                if (m && pos < slots) base[pos] = o0 + 64u * u + lane;
                kj += (uint32_t)__popcll(mask);
            }
            if (lane == (uint32_t)j) kept = kj;
            if (kj >= slots) pending &= ~(1ull << j);
        }
    }
}

// The wave-wide minimum of v.
// The wave's minimum, in every lane: within each row of 16 lanes by DPP
// (quad swaps, then rotations by 4 and 8), then the four rows' minima read
// from their first lanes.
__device__ __forceinline__ uint32_t wave_min(uint32_t v)
{
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xb1, 0xf, 0xf, false));   // quad_perm 1,0,3,2
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4e, 0xf, 0xf, false));   // quad_perm 2,3,0,1
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xf, 0xf, false));  // row_ror 4
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xf, 0xf, false));  // row_ror 8
    const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)v, 0), b = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
    const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)v, 32), d = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
    return min(min(a, b), min(c, d));
}

// An overfull bucket (grid bucket `gbk`, `held` records) from its members:
// the `slots` that arrived first (in `base`), the next ones in the bucket's
// extension slots (g.ext), and the records bin listed in g.ovf, gathered in
// that order into the wave's LDS list `mem` (64 K entries: a ballot gives
// each list match its place), so that lane l then holds members l, l + 64,
// ... (at most K). The list is scanned only when the bucket holds more than
// its slots and extension. Then `slots` rounds of wave minimum write the
// lowest ids ascending: the slots the scene-wide scan would give. Returns
// false, having written nothing, when the members found are not `held` or
// more than 64 K (the caller then scans). Called by every lane of the wave
// with the same bucket.
constexpr uint32_t MEMBERS_K = 8;

__device__ __forceinline__ bool overflow_members(GridView g, uint32_t *base, uint32_t gbk, uint32_t held,
                                                 uint32_t slots, uint32_t *mem_lds)
{
    constexpr uint32_t K = MEMBERS_K, U = 8;
    const uint32_t lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    if (slots > 64u || held > 64u * K) return false;
    if (lane < slots) mem_lds[lane] = base[lane];
    uint32_t n = slots;  // wave-uniform
    const uint32_t in_ext = min(held - slots, g.ext_slots);
    for (uint32_t k = lane; k < in_ext; k += 64u) mem_lds[n + k] = g.ext[(size_t)gbk * g.ext_slots + k];
    n += in_ext;
    const uint32_t total = n < held ? min(*g.ovf_count, g.ovf_cap) : 0u;
    for (uint32_t o0 = 0; o0 < total && n <= held; o0 += 64u * U) {
        uint2 e[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t o = o0 + 64u * u + lane;
            e[u] = o < total ? g.ovf[o] : make_uint2(0u, ~0u);
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const bool m = e[u].y == gbk;
            const uint64_t mask = __ballot(m);
            const uint32_t at = n + (uint32_t)__popcll(mask & below);
            if (m && at < 64u * K) mem_lds[at] = e[u].x;
            n += (uint32_t)__popcll(mask);
        }
    }
    if (n != held) return false;
    // The list is written by other lanes of this wave: order its reads after
    // those writes.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint32_t mem[K];
#pragma unroll
    for (uint32_t k = 0; k < K; k++) mem[k] = lane + 64u * k < n ? mem_lds[lane + 64u * k] : ~0u;
    uint32_t prev = 0u;
    for (uint32_t r = 0; r < slots; r++) {
        uint32_t best = ~0u;
#pragma unroll
        for (uint32_t k = 0; k < K; k++)
            if (mem[k] != ~0u && (r == 0u || mem[k] > prev) && mem[k] < best) best = mem[k];
        prev = wave_min(best);
        if (lane == 0u) base[r] = prev;
    }
    // The next bucket this wave takes rewrites the list.
    __builtin_amdgcn_wave_barrier();
    return true;
}

// Whether the first `held` (2..8) records of bucket slots `base` all belong
// to body `own`: every load issued before any is compared (slots past
// `held` read slot 0 again).
__device__ __forceinline__ bool bucket_of_one_body(const uint32_t *base, uint32_t held, uint32_t own,
                                                   const uint32_t *__restrict__ body)
{
    uint32_t id[8], b[8];
#pragma unroll
    for (int k = 0; k < 8; k++) id[k] = base[(uint32_t)k < held ? k : 0];
#pragma unroll
    for (int k = 0; k < 8; k++) b[k] = body[id[k]];
    bool same = true;
#pragma unroll
    for (int k = 0; k < 8; k++) same = same && b[k] == own;
    return same;
}

// sort.wgsl:14-53, run by the record that arrived first in each bucket; the
// overfull buckets are taken from bin's list of them by whichever wave asks
// next (or, without the list, by their first record's wave), from their
// members when bin listed them, else by the scan.
// Sorts the first `held` (<= N) ids of `base` ascending: a bitonic network
// over N registers, unrolled.
template <int N>
__device__ __forceinline__ void sort_slots(uint32_t *base, uint32_t held)
{
    uint32_t v[N];
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = (uint32_t)k < held ? base[k] : ~0u;
#pragma unroll
    for (int size = 2; size <= N; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
            for (int k = 0; k < N; k++) {
                const int o = k ^ stride;
                if (o > k) {
                    const bool up = (k & size) == 0;
                    const uint32_t a = v[k], c = v[o];
                    const bool swap = up ? a > c : a < c;
                    v[k] = swap ? c : a;
                    v[o] = swap ? a : c;
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < N; k++)
        if ((uint32_t)k < held) base[k] = v[k];
}

// Workgroups a sort launch adds past its tiles for the overfull buckets.
constexpr uint32_t SORT_EXTRA_GROUPS = 32;

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sort(const vs_scene_params *__restrict__ scenes, SurfaceView sv, GridView g,
                                                TileView tv)
{
    __shared__ uint32_t s_members[BLOCK / 64][64 * MEMBERS_K];
    uint32_t *const members = s_members[threadIdx.x >> 6];
    // Overfull buckets go through bin's list unless it ran out of room.
    const bool queued = g.ovb && g.ovf_count[2] == 0u;
    const Range r = tile_range(*tv.count, 0);
    for (uint32_t t = r.first; t < r.end; t += r.stride) {
        const vs_tile tile = tv.tiles[t];
        const vs_scene_params &sc = scenes[tile.scene];
        const uint32_t slots = sc.bucket_slots, gb = sc.grid_base;
        // Every lane runs every round (the rounds are uniform over the
        // workgroup), so the ballots below see the whole wave.
        for (uint32_t i0 = 0; i0 < tile.count; i0 += BLOCK) {
            const uint32_t i = i0 + threadIdx.x;
            uint32_t b = 0u;
            bool over = false;
            if (i < tile.count && sv.rank[tile.begin + i] == 0u) {
                b = sv.key[tile.begin + i];
                const uint32_t held = g.count[gb + b];
                uint32_t *base = g.slots + (size_t)(gb + b) * slots;
                over = held > slots;
                // The bucket's ids ascend (sort.wgsl:17-26's insertion sort:
                // ids are distinct, so any sort gives its order): loaded
                // together, sorted in registers by a network of 8, 16 or 32
                // (ids past the bucket's count pad with ~0 and sort last),
                // written back together.
                if (g.summary) {
                    // The bucket's summary for collide: slots held and the
                    // body all its records belong to. This record is one of
                    // them; past 8 records the bucket counts as mixed.
                    const uint32_t own = sv.body[tile.begin + i];
                    uint32_t sole = over || held > 8u ? SUMMARY_MIXED : own;
                    if (!over && held > 1u && held <= 8u && !bucket_of_one_body(base, held, own, sv.body))
                        sole = SUMMARY_MIXED;
                    g.summary[gb + b] = (held < slots ? held : slots) | sole << 6;
                }
                if (!over && held > 1u && held <= 32u) {
                    if (held <= 8u) sort_slots<8>(base, held);
                    else if (held <= 16u) sort_slots<16>(base, held);
                    else sort_slots<32>(base, held);
                }
                for (uint32_t a = 1; a < held && !over && held > 32u; a++) {
                    // slot arrays wider than 32 (a layer-1 caller's choice)
                    const uint32_t value = base[a];
                    uint32_t at = a;
                    while (at > 0u) {
                        const uint32_t previous = base[at - 1u];
                        if (previous <= value) break;
                        base[at] = previous;
                        at--;
                    }
                    base[at] = value;
                }
            }
            uint64_t pending = queued ? 0ull : __ballot(over);
            if (pending && g.ovf) {
                for (uint64_t left = pending; left; left &= left - 1ull) {
                    const int j = __builtin_ctzll(left);
                    const uint32_t bb = (uint32_t)__shfl((int)b, j);
                    if (overflow_members(g, g.slots + (size_t)(gb + bb) * slots, gb + bb, g.count[gb + bb], slots,
                                         members))
                        pending &= ~(1ull << j);
                }
            }
            if (pending) overflow_scan(sc, sv.key, g.slots + (size_t)gb * slots, b, pending, slots);
        }
    }
    if (!queued) return;
    // Wave w of the grid, counted from its end, takes listed buckets w, w +
    // W, ... (W waves in all): one each while there are fewer than waves.
    // The launch adds SORT_EXTRA_GROUPS workgroups past the tiles, which
    // come here at once and take the first of them beside the tile work.
    // Every lane of the wave is here (the tile loop is uniform over the
    // workgroup).
    const uint32_t listed = min(g.ovf_count[1], g.ovb_cap);
    const uint32_t W = gridDim.x * (BLOCK / 64u);
    for (uint32_t j = (gridDim.x - 1u - blockIdx.x) * (BLOCK / 64u) + (threadIdx.x >> 6); j < listed; j += W) {
        const uint2 e = g.ovb[j];
        const vs_scene_params &so = scenes[e.y];
        const uint32_t sl = so.bucket_slots;
        if (!overflow_members(g, g.slots + (size_t)e.x * sl, e.x, g.count[e.x], sl, members))
            overflow_scan(so, sv.key, g.slots + (size_t)so.grid_base * sl, e.x - so.grid_base, 1ull, sl);
    }
}

// The record a lane resolves against its mates (collide.wgsl:16-125).
struct Collider {
    uint32_t s, node, body;
    float radius, wi, frictionI;
    V3 pi, movedI;
    V3 sum;
    float hits;
    int32_t logged;
};

// One mate record `other` of `c`, with its node position `pj` and radius
// `rj`: the reach test and, on contact, the correction (collide.wgsl:52-118).
__device__ __forceinline__ void collide_mate(Collider &c, uint32_t other, uint32_t mateBody, V3 pj, float rj,
                                             NodeView n, const vs_body *__restrict__ bodies, SurfaceView sv,
                                             PairLog log)
{
    if (mateBody == c.body) return;
    const V3 d = sub(c.pi, pj);
    const float dist2 = dot(d, d);
    const float reach = c.radius + rj;
    if (dist2 >= reach * reach) return;
    if (log.cap) {
        if ((uint32_t)c.logged < log.cap) log.mates[(size_t)c.s * log.cap + c.logged] = other;
        c.logged++;
    }
    uint32_t mate = 0u;
    float wj;
    V3 qj;
    if (sv.prev) {
        const float4 m = sv.prev[other];
        wj = m.w;
        qj = mk(m.x, m.y, m.z);
    } else {
        mate = sv.node[other];
        wj = n.w[mate];
        qj = ld3(n.qx, n.qy, n.qz, mate);
    }
    const float total = c.wi + wj;
    if (total <= 0.0f) return;
    const float share = c.wi / total;
    const float dist = sqrtf(dist2);
    V3 normal;
    if (dist > 1e-9f) {
        normal = mk(d.x / dist, d.y / dist, d.z / dist);
    } else {
        // Coincident: +y or -y by node index (collide.wgsl's tie-break).
        if (sv.prev) mate = sv.node[other];
        normal = mk(0.f, c.node < mate ? 1.0f : -1.0f, 0.f);
    }
    const float depth = reach - dist;
    V3 corr = mul(normal, depth * share);
    const V3 sliding = sub(c.movedI, sub(pj, qj));
    const V3 tangent = sub(sliding, mul(normal, dot(sliding, normal)));
    const float slid = sqrtf(dot(tangent, tangent));
    if (slid > 1e-9f) {
        const float fj = bodies[mateBody].friction;
        const float mu = c.frictionI < fj ? c.frictionI : fj;
        const float lim = (mu * depth) / slid;
        corr = sub(corr, mul(tangent, share * (lim < 1.0f ? lim : 1.0f)));
    }
    c.sum = add(c.sum, corr);
    c.hits += 1.0f;
}

// The reach test of collide.wgsl:52-57 (another body, closer than the two
// radii), in collide_mate's arithmetic.
__device__ __forceinline__ bool reaches(const Collider &c, uint32_t mateBody, V3 pj, float rj)
{
    if (mateBody == c.body) return false;
    const V3 d = sub(c.pi, pj);
    const float reach = c.radius + rj;
    return dot(d, d) < reach * reach;
}

// A record's own data (collide.wgsl:20-40) and its bucket's coordinates.
__device__ __forceinline__ void load_collider(Collider &c, uint32_t s, NodeView n, const vs_body *__restrict__ bodies,
                                              SurfaceView sv, int dx, int dy, int &cx, int &cy, int &cz)
{
    c.s = s;
    c.node = sv.node[s];
    c.body = sv.body[s];
    c.radius = sv.radius[s];
    c.wi = n.w[c.node];
    c.pi = ld3(n.px, n.py, n.pz, c.node);
    c.movedI = sub(c.pi, ld3(n.qx, n.qy, n.qz, c.node));
    c.frictionI = bodies[c.body].friction;
    c.sum = mk(0.f, 0.f, 0.f);
    c.hits = 0.f;
    c.logged = 0;
    const uint32_t key = sv.key[s];
    cx = (int)(key % (uint32_t)dx);
    cy = (int)((key / (uint32_t)dx) % (uint32_t)dy);
    cz = (int)(key / (uint32_t)(dx * dy));
}

__device__ __forceinline__ void store_collider(const Collider &c, SurfaceView sv, PairLog log)
{
    if (log.cap) log.count[c.s] = c.logged;
    sv.dx[c.s] = c.sum.x;
    sv.dy[c.s] = c.sum.y;
    sv.dz[c.s] = c.sum.z;
    sv.hits[c.s] = c.hits;
}

// collide.wgsl:16-125 as the reference dispatches it: one lane per record,
// the 3x3x3 neighbour buckets in dz, dy, dx order, their slots ascending,
// each mate's position read through its node; Jacobi (a record writes only
// its own correction). The layer-1 dispatch runs this form.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_collide(NodeView n, const vs_body *__restrict__ bodies,
                                                   const vs_scene_params *__restrict__ scenes, SurfaceView sv,
                                                   GridView g, TileView tv, PairLog log)
{
    const Range r = tile_range(*tv.count, 0);
    for (uint32_t t = r.first; t < r.end; t += r.stride) {
        const vs_tile tile = tv.tiles[t];
        const vs_scene_params &sc = scenes[tile.scene];
        const uint32_t slots = sc.bucket_slots, gb = sc.grid_base;
        const int dx = (int)sc.grid_dims[0], dy = (int)sc.grid_dims[1], dz = (int)sc.grid_dims[2];
        for (uint32_t i = threadIdx.x; i < tile.count; i += BLOCK) {
            Collider c;
            int cx, cy, cz;
            load_collider(c, tile.begin + i, n, bodies, sv, dx, dy, cx, cy, cz);
            for (int oz = -1; oz <= 1; oz++) {
                const int z = cz + oz;
                if (z < 0 || z >= dz) continue;
                for (int oy = -1; oy <= 1; oy++) {
                    const int y = cy + oy;
                    if (y < 0 || y >= dy) continue;
                    // The row's corrections into a row sum, added to the
                    // record's in row order.
                    const V3 before = c.sum;
                    c.sum = mk(0.f, 0.f, 0.f);
                    for (int ox = -1; ox <= 1; ox++) {
                        const int x = cx + ox;
                        if (x < 0 || x >= dx) continue;
                        const uint32_t bk = gb + (uint32_t)(x + dx * (y + dy * z));
                        const uint32_t cnt = g.count[bk];
                        const uint32_t held = cnt < slots ? cnt : slots;
                        const uint32_t *base = g.slots + (size_t)bk * slots;
                        for (uint32_t k = 0; k < held; k++) {
                            const uint32_t other = base[k];
                            const uint32_t mate = sv.node[other];
                            collide_mate(c, other, sv.body[other], ld3(n.px, n.py, n.pz, mate), sv.radius[other], n,
                                         bodies, sv, log);
                        }
                    }
                    c.sum = add(before, c.sum);
                }
            }
            store_collider(c, sv, log);
        }
    }
}

// The same pass over bin's payload (sv.pay: a record's position and radius
// beside it) and sort's bucket summaries (slots held, sole body), nine lanes
// a record. A record's 27 neighbour buckets are nine rows of three along x,
// (dz, dy) = row / 3 - 1, row % 3 - 1; lane `row` of the record's group
// walks its row (loads four slots at a time, ids then bodies and payloads),
// resolves each mate in reach as it meets it, and sums the row's
// corrections; the group's first lane adds the nine row sums in row order
// (the oracle's order: the row sums of collide.wgsl's mates). A bucket whose
// records all belong to the record's own body is passed by, as is one beyond
// reach: every mate there fails the reach test. With a pair log, a first
// walk counts each row's mates in reach, so that every row logs its mates at
// its place in the record's mate order. A wave holds seven records (lane 63
// idles); a workgroup takes WAVES x 7 records of a tile at a time, so a tile
// is `parts` work items.
constexpr uint32_t ROW_LANES = 9;
constexpr uint32_t ROW_GROUPS = 64 / ROW_LANES;

template <int BLOCK>
constexpr uint32_t collide_parts(uint32_t tile_max)
{
    return (tile_max + (BLOCK / 64) * ROW_GROUPS - 1) / ((BLOCK / 64) * ROW_GROUPS);
}

// Walks bucket bk's first `held` slots in mate order, four at a time, and
// calls f(other, mateBody, payload) for each.
template <typename F>
__device__ __forceinline__ void walk_bucket(uint32_t bk, uint32_t held, uint32_t slots, uint32_t ownBody, GridView g,
                                            SurfaceView sv, F f)
{
    const uint32_t *base = g.slots + (size_t)bk * slots;
    for (uint32_t k0 = 0; k0 < held; k0 += 4u) {
        uint32_t o[4], mb[4];
        float4 q[4];
#pragma unroll
        for (int u = 0; u < 4; u++) o[u] = k0 + u < held ? base[k0 + u] : 0u;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            mb[u] = ownBody;  // past the bucket's end: the record's own body, never in reach
            q[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (k0 + u < held) {
                mb[u] = sv.body[o[u]];
                q[u] = sv.pay[o[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; u++) f(o[u], mb[u], q[u]);
    }
}

// The row's buckets bk[0..2] (held[ox] slots each) in mate order.
template <typename F>
__device__ __forceinline__ void walk_row(const uint32_t (&bk)[3], const uint32_t (&held)[3], uint32_t slots,
                                         uint32_t ownBody, GridView g, SurfaceView sv, F f)
{
#pragma unroll
    for (int ox = 0; ox < 3; ox++) walk_bucket(bk[ox], held[ox], slots, ownBody, g, sv, f);
}

// A record's neighbour buckets out of reach, per axis: the bucket below
// when the record sits farther than the largest reach (plus the rounding
// margin) above its bucket's lower face, the bucket above likewise. Mates
// there fail the reach test, so passing them by changes nothing (with the
// reference's default grid, buckets twice the widest radius, nothing is
// ever passed by).
__device__ __forceinline__ void reach_flags(const vs_scene_params &sc, V3 p, int cx, int cy, int cz, bool (&lo)[3],
                                            bool (&hi)[3])
{
#pragma unroll
    for (int a = 0; a < 3; a++) lo[a] = hi[a] = false;
    if (sc.reach_max > 0.0f) {
        const float bound = sc.reach_max + sc.skip_margin, size = sc.bucket_size;
        const float pc[3] = { p.x, p.y, p.z };
        const int cell[3] = { cx, cy, cz };
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float f = (pc[a] - sc.grid_origin[a]) / size - (float)cell[a];
            lo[a] = f * size > bound;
            hi[a] = (1.0f - f) * size > bound;
        }
    }
}

// Neighbour bucket (ox, oy, oz) of a record in bucket (cx, cy, cz): its
// index (grid base added), and the slots the record walks there: 0 outside
// the grid, beyond reach, or when the bucket's records are all of `body`.
__device__ __forceinline__ uint32_t row_bucket(const vs_scene_params &sc, GridView g, int cx, int cy, int cz, int ox,
                                               int oy, int oz, const bool (&lo)[3], const bool (&hi)[3], uint32_t body,
                                               uint32_t &held)
{
    const int x = cx + ox, y = cy + oy, z = cz + oz;
    const int dx = (int)sc.grid_dims[0], dy = (int)sc.grid_dims[1], dz = (int)sc.grid_dims[2];
    const bool in = x >= 0 && x < dx && y >= 0 && y < dy && z >= 0 && z < dz && !(ox < 0 && lo[0]) &&
                    !(ox > 0 && hi[0]) && !(oy < 0 && lo[1]) && !(oy > 0 && hi[1]) && !(oz < 0 && lo[2]) &&
                    !(oz > 0 && hi[2]);
    const uint32_t bk = in ? sc.grid_base + (uint32_t)(x + dx * (y + dy * z)) : 0u;
    const uint32_t w = in ? g.summary[bk] : 0u;
    held = summary_sole(w) == body ? 0u : summary_held(w);
    return bk;
}

// Six waves a SIMD (at most 80 VGPRs): the pass is latency-bound, and at
// five it runs slower. LOG false: the form for launches without a pair log,
// whose logging code is left out (with it, the compiler spills two registers
// to stay at six waves).
template <int BLOCK, bool LOG>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6))) void k_collide_rows(NodeView n, const vs_body *__restrict__ bodies,
                                                        const vs_scene_params *__restrict__ scenes, SurfaceView sv,
                                                        GridView g, TileView tv, PairLog log, uint32_t parts)
{
    if (!LOG) log.cap = 0u;
    constexpr uint32_t WAVES = BLOCK / 64;
    constexpr uint32_t PER_PART = WAVES * ROW_GROUPS;
    const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t group = lane / ROW_LANES, row = lane - group * ROW_LANES;
    const uint32_t lead = group * ROW_LANES;
    const uint32_t items = *tv.count * parts;
    for (uint32_t it = blockIdx.x; it < items; it += gridDim.x) {
        const uint32_t t = it / parts, part = it - t * parts;
        const vs_tile tile = tv.tiles[t];
        const uint32_t first = part * PER_PART;
        if (first >= tile.count) continue;  // the whole workgroup: `it` is its own
        const vs_scene_params &sc = scenes[tile.scene];
        const uint32_t slots = sc.bucket_slots;
        const int dx = (int)sc.grid_dims[0], dy = (int)sc.grid_dims[1];
        const uint32_t i = first + wv * ROW_GROUPS + group;
        const bool active = group < ROW_GROUPS && i < tile.count;
        Collider c;
        uint32_t held[3] = { 0u, 0u, 0u }, bk[3] = { 0u, 0u, 0u };
        if (active) {
            int cx, cy, cz;
            load_collider(c, tile.begin + i, n, bodies, sv, dx, dy, cx, cy, cz);
            bool lo[3], hi[3];
            reach_flags(sc, c.pi, cx, cy, cz, lo, hi);
            const int oz = (int)(row / 3u) - 1, oy = (int)(row % 3u) - 1;
#pragma unroll
            for (int ox = 0; ox < 3; ox++) bk[ox] = row_bucket(sc, g, cx, cy, cz, ox - 1, oy, oz, lo, hi, c.body, held[ox]);
        }
        if (log.cap) {
            // Each row's mates in reach, so that the rows log in mate order.
            uint32_t mine = 0;
            if (active)
                walk_row(bk, held, slots, c.body, g, sv, [&](uint32_t, uint32_t mb, float4 q) {
                    if (reaches(c, mb, mk(q.x, q.y, q.z), q.w)) mine++;
                });
            uint32_t before = 0, total = 0;
#pragma unroll
            for (uint32_t k = 0; k < ROW_LANES; k++) {
                const uint32_t m = (uint32_t)__shfl((int)mine, (int)min(lead + k, 63u));
                before += k < row ? m : 0u;
                total += m;
            }
            if (active) {
                c.logged = (int32_t)before;
                if (row == 0u) log.count[c.s] = (int32_t)total;
            }
        }
        if (active)
            walk_row(bk, held, slots, c.body, g, sv, [&](uint32_t other, uint32_t mb, float4 q) {
                collide_mate(c, other, mb, mk(q.x, q.y, q.z), q.w, n, bodies, sv, log);
            });
        // The row sums in row order (shuffles take the whole wave).
        V3 sum = mk(0.f, 0.f, 0.f);
        float hits = 0.f;
#pragma unroll
        for (uint32_t k = 0; k < ROW_LANES; k++) {
            const int from = (int)min(lead + k, 63u);
            const V3 r = mk(__shfl(active ? c.sum.x : 0.f, from), __shfl(active ? c.sum.y : 0.f, from),
                            __shfl(active ? c.sum.z : 0.f, from));
            sum = add(sum, r);
            hits += __shfl(active ? c.hits : 0.f, from);
        }
        if (active && row == 0u) {
            sv.dx[c.s] = sum.x;
            sv.dy[c.s] = sum.y;
            sv.dz[c.s] = sum.z;
            sv.hits[c.s] = hits;
        }
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_settle(NodeView n, const vs_scene_params *__restrict__ scenes,
                                                  SurfaceView sv, GridView g, TileView tv)
{
    // The overflow list is empty again for the next substep's bin (which
    // runs after this launch).
    if (g.ovf_count && blockIdx.x == 0 && threadIdx.x == 0) {
        g.ovf_count[0] = 0u;
        if (g.ovb) g.ovf_count[1] = g.ovf_count[2] = 0u;
    }
    const Range r = tile_range(*tv.count, 0);
    for (uint32_t t = r.first; t < r.end; t += r.stride) {
        const vs_tile tile = tv.tiles[t];
        const vs_scene_params &sc = scenes[tile.scene];
        const float gain = sc.gain;
        const uint32_t gb = sc.grid_base;
        for (uint32_t i = threadIdx.x; i < tile.count; i += BLOCK) {
            const uint32_t s = tile.begin + i;
            if (sv.rank[s] == 0u) {
                g.count[gb + sv.key[s]] = 0u;
                if (g.summary) g.summary[gb + sv.key[s]] = 0u;
            }
            const float h = sv.hits[s];
            if (h > 0.0f) {
                const uint32_t node = sv.node[s];
                const float f = gain / h;
                n.px[node] = n.px[node] + sv.dx[s] * f;
                n.py[node] = n.py[node] + sv.dy[s] * f;
                n.pz[node] = n.pz[node] + sv.dz[s] * f;
            }
        }
    }
}

// ---------------------------------------------------------------- drive
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_drive(NodeView n, const vs_scene_params *__restrict__ scenes, DriveView d,
                                                 TileView tv, int law)
{
    const Range r = tile_range(*tv.count, 0);
    for (uint32_t t = r.first; t < r.end; t += r.stride) {
        const vs_tile tile = tv.tiles[t];
        const vs_scene_params &sc = scenes[tile.scene];
        const float dt = sc.dt;
        const V3 target = mk(sc.target[0], sc.target[1], sc.target[2]);
        const V3 tvel = mk(sc.target_vel[0], sc.target_vel[1], sc.target_vel[2]);
        const int compliant = law < 0 ? (int)sc.drive_law : law;
        const float rt = sc.response_time;
        for (uint32_t i = threadIdx.x; i < tile.count; i += BLOCK) {
            const uint32_t k = tile.begin + i;
            const uint32_t node = d.node[k];
            const V3 goal = add(target, mk(d.ox[k], d.oy[k], d.oz[k]));
            const V3 here = ld3(n.px, n.py, n.pz, node);
            if (!compliant) {
                const V3 step = mul(tvel, dt);
                const V3 toGoal = sub(goal, here);
                const V3 next = add(here, dot(step, step) >= dot(toGoal, toGoal) ? toGoal : step);
                n.qx[node] = here.x;
                n.qy[node] = here.y;
                n.qz[node] = here.z;
                n.px[node] = next.x;
                n.py[node] = next.y;
                n.pz[node] = next.z;
                const V3 v = sub(next, here);
                n.vx[node] = v.x / dt;
                n.vy[node] = v.y / dt;
                n.vz[node] = v.z / dt;
            } else {
                const float dt2 = dt * dt;
                const float tau2 = rt * rt;
                const float frac = dt2 / (dt2 + tau2);
                const V3 np = add(here, mul(sub(goal, here), frac));
                n.px[node] = np.x;
                n.py[node] = np.y;
                n.pz[node] = np.z;
            }
        }
    }
}

// --------------------------------------------------------------- splats
// A covariance taken apart into the engine's splat format, a unit rotation
// (xyzw) and three axis lengths, by four sweeps of cyclic Jacobi
// (ellipsoid.wgsl:32-91) in the oracle's o_ellipsoid_axes order: planes
// (p, p+1, other = p+2), an exactly-zero off-diagonal skipped, the smaller
// angle, the half-angle tangent -sin/(1+cos) about +other joined onto the
// quaternion (only its non-zero terms), renormalised by the reciprocal
// length; lengths sqrt(max(eigenvalue, 0)). d = diagonal, o[k] = entry
// (k, k+1 mod 3).
__device__ __forceinline__ void ellipsoid_axes(float d[3], float o[3], float q[4])
{
    q[0] = 0.f;
    q[1] = 0.f;
    q[2] = 0.f;
    q[3] = 1.f;
#pragma unroll
    for (int sweep = 0; sweep < 4; sweep++) {
#pragma unroll
        for (int p = 0; p < 3; p++) {
            const int r = (p + 1) % 3, other = (p + 2) % 3;
            const float apr = o[p];
            if (apr == 0.0f) continue;
            const float app = d[p], arr = d[r];
            const float theta = (arr - app) / (2.0f * apr);
            const float sign = theta >= 0.0f ? 1.0f : -1.0f;
            const float tangent = sign / (fabsf(theta) + sqrtf(theta * theta + 1.0f));
            const float cosine = 1.0f / sqrtf(tangent * tangent + 1.0f);
            const float sine = tangent * cosine;
            const float apo = o[other], aro = o[r];
            o[other] = cosine * apo - sine * aro;
            o[r] = sine * apo + cosine * aro;
            d[p] = app - tangent * apr;
            d[r] = arr + tangent * apr;
            o[p] = 0.0f;
            const float half = -(sine / (1.0f + cosine));
            const float x = q[0], y = q[1], z = q[2], w = q[3];
            float nx, ny, nz, nw;
            if (other == 0) {
                nx = w * half + x;
                ny = y + z * half;
                nz = z - y * half;
                nw = w - x * half;
            } else if (other == 1) {
                nx = x - z * half;
                ny = w * half + y;
                nz = x * half + z;
                nw = w - y * half;
            } else {
                nx = x + y * half;
                ny = y - x * half;
                nz = w * half + z;
                nw = w - z * half;
            }
            const float inv = 1.0f / sqrtf(((nx * nx + ny * ny) + nz * nz) + nw * nw);
            q[0] = nx * inv;
            q[1] = ny * inv;
            q[2] = nz * inv;
            q[3] = nw * inv;
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) d[k] = sqrtf(d[k] > 0.0f ? d[k] : 0.0f);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_splats(const float *__restrict__ px, const float *__restrict__ py,
                                                  const float *__restrict__ pz, const vs_body *__restrict__ bodies,
                                                  SplatView sp, TileView tv, uint32_t xcd)
{
    // With `xcd`, each XCD takes a contiguous run of tiles: a body's splats
    // (stored in cell order) then gather its corner nodes through one L2.
    const Range r = tile_range(*tv.count, xcd);
    for (uint32_t t = r.first; t < r.end; t += r.stride) {
        const vs_tile tile = tv.tiles[t];
        const vs_body bd = bodies[tile.body];
        const float inv = 1.0f / bd.h;
        const float last[3] = { (float)bd.cell_dims[0] - 1.0f, (float)bd.cell_dims[1] - 1.0f,
                                (float)bd.cell_dims[2] - 1.0f };
        const uint32_t ndx = bd.node_dims[0], ndxy = bd.node_dims[0] * bd.node_dims[1];
        for (uint32_t i = threadIdx.x; i < tile.count; i += BLOCK) {
            const uint32_t k = tile.begin + i;
            // Every input of the splat up front, so that its loads are in
            // flight together (the outputs may alias them for the compiler).
            const float rest[3] = { sp.mx[k], sp.my[k], sp.mz[k] };
            const float4 qv = sp.quat[k];
            const float sx = sp.sx[k], sy = sp.sy[k], sz = sp.sz[k];
            float local[3];
            uint32_t base[3];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const float grid = (rest[a] - bd.origin[a]) * inv;
                float cell = floorf(grid);
                cell = cell < 0.0f ? 0.0f : cell;
                cell = cell > last[a] ? last[a] : cell;
                local[a] = grid - cell;
                base[a] = (uint32_t)cell;
            }
            const uint32_t n0 = bd.node_offset + base[0] + ndx * base[1] + ndxy * base[2];
            // The trilinear sum and its gradient axis by axis (the oracle's
            // o_deform_splats order).
            V3 Pj[4], Dj[4];  // j = fy + 2 fz
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t e = n0 + (j & 1 ? ndx : 0u) + (j & 2 ? ndxy : 0u);
                const V3 p0 = ld3(px, py, pz, e), p1 = ld3(px, py, pz, e + 1u);
                Dj[j] = sub(p1, p0);
                Pj[j] = lerpf(p0, Dj[j], local[0]);
            }
            V3 E[2], Q[2], Dy[2];
#pragma unroll
            for (int fz = 0; fz < 2; fz++) {
                E[fz] = sub(Pj[1 + 2 * fz], Pj[2 * fz]);
                Q[fz] = lerpf(Pj[2 * fz], E[fz], local[1]);
                Dy[fz] = lerpf(Dj[2 * fz], sub(Dj[1 + 2 * fz], Dj[2 * fz]), local[1]);
            }
            const V3 Z = sub(Q[1], Q[0]);
            const V3 pos = lerpf(Q[0], Z, local[2]);
            const V3 c0 = mul(lerpf(Dy[0], sub(Dy[1], Dy[0]), local[2]), inv);
            const V3 c1 = mul(lerpf(E[0], sub(E[1], E[0]), local[2]), inv);
            const V3 c2 = mul(Z, inv);
            sp.ox[k] = pos.x;
            sp.oy[k] = pos.y;
            sp.oz[k] = pos.z;
            const float x = qv.x, y = qv.y, z = qv.z, w = qv.w;
            const V3 r0 = mk(fmaf(-2.0f, fmaf(y, y, z * z), 1.0f), 2.0f * fmaf(x, y, z * w), 2.0f * fmaf(x, z, -(y * w)));
            const V3 r1 = mk(2.0f * fmaf(x, y, -(z * w)), fmaf(-2.0f, fmaf(x, x, z * z), 1.0f), 2.0f * fmaf(y, z, x * w));
            const V3 r2 = mk(2.0f * fmaf(x, z, y * w), 2.0f * fmaf(y, z, -(x * w)), fmaf(-2.0f, fmaf(x, x, y * y), 1.0f));
            const V3 a0 = mul(lerpf(lerpf(mul(c0, r0.x), c1, r0.y), c2, r0.z), sx);
            const V3 a1 = mul(lerpf(lerpf(mul(c0, r1.x), c1, r1.y), c2, r1.z), sy);
            const V3 a2 = mul(lerpf(lerpf(mul(c0, r2.x), c1, r2.y), c2, r2.z), sz);
            const size_t P = sp.plane;
            float d[3] = { fmaf(a2.x, a2.x, fmaf(a1.x, a1.x, a0.x * a0.x)), fmaf(a2.y, a2.y, fmaf(a1.y, a1.y, a0.y * a0.y)),
                           fmaf(a2.z, a2.z, fmaf(a1.z, a1.z, a0.z * a0.z)) };
            float o[3] = { fmaf(a2.x, a2.y, fmaf(a1.x, a1.y, a0.x * a0.y)), fmaf(a2.y, a2.z, fmaf(a1.y, a1.z, a0.y * a0.z)),
                           fmaf(a2.x, a2.z, fmaf(a1.x, a1.z, a0.x * a0.z)) };
            if (sp.mode & VS_SPLAT_COV) {
                sp.cov[0 * P + k] = d[0];
                sp.cov[1 * P + k] = o[0];
                sp.cov[2 * P + k] = o[2];
                sp.cov[3 * P + k] = d[1];
                sp.cov[4 * P + k] = o[1];
                sp.cov[5 * P + k] = d[2];
            }
            if (sp.mode & VS_SPLAT_AXES) {
                float q[4];
                ellipsoid_axes(d, o, q);
#pragma unroll
                for (int a = 0; a < 4; a++) sp.rot[a * P + k] = q[a];
#pragma unroll
                for (int a = 0; a < 3; a++) sp.axis[a * P + k] = d[a];
            }
        }
    }
}

}  // namespace

#define VSIM_DISPATCH_BLOCK(KERNEL, SH, STREAM, ...)                                                           \
    do {                                                                                                       \
        switch ((SH).block) {                                                                                  \
        case 64: hipLaunchKernelGGL(KERNEL<64>, dim3((SH).grid), dim3(64), 0, STREAM, __VA_ARGS__); break;    \
        case 128: hipLaunchKernelGGL(KERNEL<128>, dim3((SH).grid), dim3(128), 0, STREAM, __VA_ARGS__); break; \
        case 256: hipLaunchKernelGGL(KERNEL<256>, dim3((SH).grid), dim3(256), 0, STREAM, __VA_ARGS__); break; \
        case 512: hipLaunchKernelGGL(KERNEL<512>, dim3((SH).grid), dim3(512), 0, STREAM, __VA_ARGS__); break; \
        default: return hipErrorInvalidConfiguration;                                                          \
        }                                                                                                      \
        return hipGetLastError();                                                                              \
    } while (0)

// The same for a kernel with a second template argument ARG.
#define VSIM_DISPATCH_BLOCK2(KERNEL, ARG, SH, STREAM, ...)                                                         \
    do {                                                                                                           \
        switch ((SH).block) {                                                                                      \
        case 64: hipLaunchKernelGGL((KERNEL<64, ARG>), dim3((SH).grid), dim3(64), 0, STREAM, __VA_ARGS__); break;    \
        case 128: hipLaunchKernelGGL((KERNEL<128, ARG>), dim3((SH).grid), dim3(128), 0, STREAM, __VA_ARGS__); break; \
        case 256: hipLaunchKernelGGL((KERNEL<256, ARG>), dim3((SH).grid), dim3(256), 0, STREAM, __VA_ARGS__); break; \
        case 512: hipLaunchKernelGGL((KERNEL<512, ARG>), dim3((SH).grid), dim3(512), 0, STREAM, __VA_ARGS__); break; \
        default: return hipErrorInvalidConfiguration;                                                              \
        }                                                                                                          \
        return hipGetLastError();                                                                                  \
    } while (0)

// The same for a kernel with two more template arguments A and B.
#define VSIM_DISPATCH_BLOCK3(KERNEL, A, B, SH, STREAM, ...)                                                          \
    do {                                                                                                             \
        switch ((SH).block) {                                                                                        \
        case 64: hipLaunchKernelGGL((KERNEL<64, A, B>), dim3((SH).grid), dim3(64), 0, STREAM, __VA_ARGS__); break;    \
        case 128: hipLaunchKernelGGL((KERNEL<128, A, B>), dim3((SH).grid), dim3(128), 0, STREAM, __VA_ARGS__); break; \
        case 256: hipLaunchKernelGGL((KERNEL<256, A, B>), dim3((SH).grid), dim3(256), 0, STREAM, __VA_ARGS__); break; \
        case 512: hipLaunchKernelGGL((KERNEL<512, A, B>), dim3((SH).grid), dim3(512), 0, STREAM, __VA_ARGS__); break; \
        default: return hipErrorInvalidConfiguration;                                                                \
        }                                                                                                            \
        return hipGetLastError();                                                                                    \
    } while (0)

hipError_t launch_predict(hipStream_t s, Shape sh, NodeView n, const vs_scene_params *scenes, TileView t)
{
    VSIM_DISPATCH_BLOCK(k_predict, sh, s, n, scenes, t);
}

hipError_t launch_finalize(hipStream_t s, Shape sh, NodeView n, const vs_body *bodies, const vs_scene_params *scenes,
                           TileView t)
{
    VSIM_DISPATCH_BLOCK(k_finalize, sh, s, n, bodies, scenes, t);
}

hipError_t launch_finalize_predict(hipStream_t s, Shape sh, NodeView n, const vs_body *bodies,
                                  const vs_scene_params *scenes, TileView t)
{
    VSIM_DISPATCH_BLOCK(k_finalize_predict, sh, s, n, bodies, scenes, t);
}

hipError_t launch_elastic(hipStream_t s, Shape sh, NodeView n, const vs_body *bodies, const vs_scene_params *scenes,
                          const uint32_t *cells, TileView t, uint32_t limit_only, bool stretch)
{
    if (stretch) VSIM_DISPATCH_BLOCK2(k_elastic, true, sh, s, n, bodies, scenes, cells, t, limit_only, sh.xcd);
    VSIM_DISPATCH_BLOCK2(k_elastic, false, sh, s, n, bodies, scenes, cells, t, limit_only, sh.xcd);
}

hipError_t launch_elastic_rows(hipStream_t s, Shape sh, NodeView n, const vs_body *bodies,
                               const vs_scene_params *scenes, TileView t, uint32_t cls, uint32_t limit_only,
                               uint32_t node_mode, bool stretch)
{
#define VSIM_ROWS(S, M) \
    VSIM_DISPATCH_BLOCK3(k_elastic_rows, S, M, sh, s, n, bodies, scenes, t, cls, limit_only, node_mode, sh.xcd)
    if (node_mode == 0u) {
        if (stretch) VSIM_ROWS(true, 0);
        VSIM_ROWS(false, 0);
    }
    if (node_mode == 3u) {
        if (stretch) VSIM_ROWS(true, 3);
        VSIM_ROWS(false, 3);
    }
    if (stretch) VSIM_ROWS(true, 1);
    VSIM_ROWS(false, 1);
#undef VSIM_ROWS
}

hipError_t launch_elastic_bodies(hipStream_t s, Shape sh, NodeView n, const vs_body *bodies,
                                 const vs_scene_params *scenes, TileView t, uint32_t limit_only, uint32_t node_mode,
                                 bool stretch, SurfaceView sv, GridView g, GridView gn, uint32_t contact)
{
    if (stretch)
        VSIM_DISPATCH_BLOCK2(k_elastic_bodies, true, sh, s, n, bodies, scenes, t, limit_only, node_mode, sv, g, gn,
                             contact);
    VSIM_DISPATCH_BLOCK2(k_elastic_bodies, false, sh, s, n, bodies, scenes, t, limit_only, node_mode, sv, g, gn,
                         contact);
}

hipError_t launch_contact_bin(hipStream_t s, Shape sh, NodeView n, const vs_scene_params *scenes, SurfaceView sv,
                              GridView g, TileView t)
{
    VSIM_DISPATCH_BLOCK(k_bin, sh, s, n, scenes, sv, g, t);
}

hipError_t launch_contact_sort(hipStream_t s, Shape sh, const vs_scene_params *scenes, SurfaceView sv, GridView g,
                               TileView t)
{
    if (g.ovb) sh.grid += SORT_EXTRA_GROUPS;
    VSIM_DISPATCH_BLOCK(k_sort, sh, s, scenes, sv, g, t);
}

hipError_t launch_contact_collide(hipStream_t s, Shape sh, NodeView n, const vs_body *bodies,
                                  const vs_scene_params *scenes, SurfaceView sv, GridView g, TileView t, PairLog log,
                                  uint32_t tile_max)
{
    if (!sv.pay || !g.summary) VSIM_DISPATCH_BLOCK(k_collide, sh, s, n, bodies, scenes, sv, g, t, log);
    // Row lanes: each tile is `parts` work items of WAVES x 7 records.
    const uint32_t parts = sh.block == 64    ? collide_parts<64>(tile_max)
                           : sh.block == 128 ? collide_parts<128>(tile_max)
                           : sh.block == 256 ? collide_parts<256>(tile_max)
                                             : collide_parts<512>(tile_max);
    sh.grid *= parts;
    if (log.cap) VSIM_DISPATCH_BLOCK2(k_collide_rows, true, sh, s, n, bodies, scenes, sv, g, t, log, parts);
    VSIM_DISPATCH_BLOCK2(k_collide_rows, false, sh, s, n, bodies, scenes, sv, g, t, log, parts);
}

hipError_t launch_contact_settle(hipStream_t s, Shape sh, NodeView n, const vs_scene_params *scenes, SurfaceView sv,
                                 GridView g, TileView t)
{
    VSIM_DISPATCH_BLOCK(k_settle, sh, s, n, scenes, sv, g, t);
}

hipError_t launch_drive(hipStream_t s, Shape sh, NodeView n, const vs_scene_params *scenes, DriveView d, TileView t,
                        int law)
{
    VSIM_DISPATCH_BLOCK(k_drive, sh, s, n, scenes, d, t, law);
}

hipError_t launch_splats(hipStream_t s, Shape sh, const float *px, const float *py, const float *pz,
                         const vs_body *bodies, SplatView sp, TileView t)
{
    VSIM_DISPATCH_BLOCK(k_splats, sh, s, px, py, pz, bodies, sp, t, sh.xcd);
}

}  // namespace vsim
