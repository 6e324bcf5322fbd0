// grab.hip -- what a grab needs from the device, answered on the device:
// the node a ray picks (nearestNodeToRay, packages/visionary-sim/src/sim/
// pick.js:41-68) and the held set around it (withinRadius, grip.js:99-138),
// so that a grab reads back a few hundred bytes instead of every position.
//
// Both follow the reference's f64 arithmetic over the f32 positions: every
// difference, product and sum is a double, in the order the JS evaluates it
// (no contraction), so the answers are the JS answers exactly.
//
// Eligibility is the reference's isGrabbableNode: material inverse mass > 0
// and not statically prescribed. The caller passes it as one f32 plane, the
// material inverse mass with static nodes zeroed ("grab mass"); a node the
// drive already holds stays eligible, as pick.js:14-24 requires.
//
// Each kernel is one launch. Workgroups scan a share of the nodes; the last
// one to finish (a device-scope counter after a release fence) reduces or
// selects, writes the answer and clears the counters it used, so the output
// buffer is ready for the next launch without a host-side clear.
#include <mutex>

#include "kernels.h"

namespace vsim {
namespace {

constexpr uint32_t PICK_BLOCK = 256;
constexpr uint32_t GRIP_BLOCK = 1024;
constexpr uint32_t SELECT_CAP = 4096;  // held nodes a grip can sort in LDS

// (depth, node) lexicographic minimum; node < 0 is "none".
__device__ __forceinline__ bool nearer(double d, int32_t n, double bd, int32_t bn)
{
    if (n < 0) return false;
    if (bn < 0) return true;
    return d < bd || (d == bd && n < bn);
}

struct PickOut {
    uint32_t done;   // workgroups finished (reset by the last)
    int32_t node;    // answer: nearest node, or -1
    double depth;    // its depth along the ray
    // then one {depth, node, pad} partial per workgroup
};

__global__ __launch_bounds__(PICK_BLOCK) void k_pick(const float *__restrict__ px, const float *__restrict__ py,
                                                     const float *__restrict__ pz, const float *__restrict__ mass,
                                                     uint32_t count, const vs_grab_params *__restrict__ gp,
                                                     PickOut *out)
{
    __shared__ double s_depth[PICK_BLOCK / 64];
    __shared__ int32_t s_node[PICK_BLOCK / 64];
    __shared__ bool s_last;
    const vs_grab_params g = *gp;
    const double r2 = g.radius * g.radius;
    double best = 0.0;
    int32_t bn = -1;
    for (uint32_t n = blockIdx.x * PICK_BLOCK + threadIdx.x; n < count; n += gridDim.x * PICK_BLOCK) {
        if (!(mass[n] > 0.0f)) continue;
        const double dx = (double)px[n] - g.origin[0];
        const double dy = (double)py[n] - g.origin[1];
        const double dz = (double)pz[n] - g.origin[2];
        const double depth = (dx * g.direction[0] + dy * g.direction[1]) + dz * g.direction[2];
        if (!(depth > 0.0)) continue;
        const double off = ((dx * dx + dy * dy) + dz * dz) - depth * depth;
        if (off > r2) continue;
        if (nearer(depth, (int32_t)n, best, bn)) {
            best = depth;
            bn = (int32_t)n;
        }
    }
    auto wave_min = [&]() {
        for (int s = 32; s > 0; s >>= 1) {
            const double od = __shfl_xor(best, s);
            const int32_t on = __shfl_xor(bn, s);
            if (nearer(od, on, best, bn)) {
                best = od;
                bn = on;
            }
        }
    };
    wave_min();
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_depth[wave] = best;
        s_node[wave] = bn;
    }
    __syncthreads();
    double *partial = reinterpret_cast<double *>(out + 1);
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < PICK_BLOCK / 64; w++) {
            if (nearer(s_depth[w], s_node[w], best, bn)) {
                best = s_depth[w];
                bn = s_node[w];
            }
        }
        partial[2 * blockIdx.x] = best;
        reinterpret_cast<int32_t *>(partial)[4 * blockIdx.x + 2] = bn;
        __threadfence();
        s_last = atomicAdd(&out->done, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    best = 0.0;
    bn = -1;
    for (uint32_t b = threadIdx.x; b < gridDim.x; b += PICK_BLOCK) {
        const double d = __hip_atomic_load(&partial[2 * b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int32_t n = __hip_atomic_load(&reinterpret_cast<int32_t *>(partial)[4 * b + 2], __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
        if (nearer(d, n, best, bn)) {
            best = d;
            bn = n;
        }
    }
    wave_min();
    __syncthreads();
    if (lane == 0) {
        s_depth[wave] = best;
        s_node[wave] = bn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < PICK_BLOCK / 64; w++) {
            if (nearer(s_depth[w], s_node[w], best, bn)) {
                best = s_depth[w];
                bn = s_node[w];
            }
        }
        out->node = bn;
        out->depth = bn < 0 ? 0.0 : best;
        out->done = 0u;
    }
}

struct Candidate {
    double d2;
    uint32_t node, pad;
};

__device__ __forceinline__ bool before(double a, uint32_t an, double b, uint32_t bn)
{
    return a < b || (a == b && an < bn);
}

__device__ __forceinline__ uint64_t key_bits(double d) { return __double_as_longlong(d); }

// out: header then `limit` node ids then `limit` offsets (x, y, z doubles).
struct GripOut {
    uint32_t count;   // candidates appended (reset by the last workgroup)
    uint32_t done;    // workgroups finished (reset by the last)
    uint32_t found;   // eligible nodes inside the hand, anchor included
    uint32_t kept;    // entries written, anchor first
    uint32_t status;  // 0 ok; 1 the anchor is not eligible; 2 candidates overflowed; 3 bad range
    uint32_t pad[3];
};

// The grip of one question `g`. `ctr` holds the launch's two counters
// (candidates appended, workgroups finished), zero at launch; the last
// workgroup clears them again.
__device__ __forceinline__ void grip(const float *__restrict__ px, const float *__restrict__ py,
                                     const float *__restrict__ pz, const float *__restrict__ mass, uint32_t count,
                                     const vs_grab_params &g, Candidate *cand, uint32_t cap, uint32_t limit,
                                     uint32_t *ctr, GripOut *out)
{
    __shared__ double s_key[SELECT_CAP];
    __shared__ uint32_t s_node[SELECT_CAP];
    __shared__ uint32_t s_hist[256];
    __shared__ uint32_t s_at, s_digit, s_before;
    __shared__ bool s_last;
    const bool range_ok = g.begin <= g.end && g.end <= count && g.anchor >= g.begin && g.anchor < g.end;
    const bool eligible = range_ok && mass[g.anchor] > 0.0f;
    const double ax = range_ok ? (double)px[g.anchor] : 0.0, ay = range_ok ? (double)py[g.anchor] : 0.0,
                 az = range_ok ? (double)pz[g.anchor] : 0.0;
    const double r2 = g.radius * g.radius;
    if (eligible) {
        const uint32_t lane = threadIdx.x & 63u;
        for (uint32_t n = g.begin + blockIdx.x * GRIP_BLOCK + threadIdx.x; n < g.end; n += gridDim.x * GRIP_BLOCK) {
            double d2 = 0.0;
            bool take = false;
            if (n != g.anchor && mass[n] > 0.0f) {
                const double dx = (double)px[n] - ax, dy = (double)py[n] - ay, dz = (double)pz[n] - az;
                d2 = (dx * dx + dy * dy) + dz * dz;
                take = d2 <= r2;
            }
            // One counter update per wave: the lowest taking lane reserves
            // the wave's slots.
            const uint64_t m = __ballot(take);
            if (m == 0) continue;
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1u;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&ctr[0], (uint32_t)__popcll(m));
            base = __shfl(base, (int)leader);
            const uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (take && slot < cap) cand[slot] = Candidate{ d2, n, 0u };
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(&ctr[1], 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    const uint32_t c = __hip_atomic_load(&ctr[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t *nodes = reinterpret_cast<uint32_t *>(out + 1);
    double *offs = reinterpret_cast<double *>(nodes + limit + (limit & 1u));
    auto finish = [&](uint32_t found, uint32_t kept, uint32_t status) {
        if (threadIdx.x == 0) {
            out->found = found;
            out->kept = kept;
            out->status = status;
            out->count = 0u;
            out->done = 0u;
            ctr[0] = 0u;
            ctr[1] = 0u;
        }
    };
    if (!eligible) {
        finish(0u, 0u, range_ok ? 1u : 3u);
        return;
    }
    // The anchor first with an exact zero offset; then the `want` nearest
    // candidates by (distance^2, node).
    const uint32_t want = min(c, limit > 0u ? limit - 1u : 0u);
    if (c > cap || want > SELECT_CAP) {
        finish(c + 1u, 0u, 2u);
        return;
    }
    if (threadIdx.x == 0) {
        nodes[0] = g.anchor;
        offs[0] = 0.0;
        offs[1] = 0.0;
        offs[2] = 0.0;
    }
    auto load = [&](uint32_t i) -> Candidate {
        Candidate k;
        k.d2 = __hip_atomic_load(&cand[i].d2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        k.node = __hip_atomic_load(&cand[i].node, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return k;
    };
    // Gather the entries to sort into LDS: all of them when they fit, else
    // the `want` smallest found by radix selection (the 64-bit patterns of
    // non-negative doubles order like the doubles), ties at the threshold
    // broken by node index with a second selection over the ids.
    if (threadIdx.x == 0) s_at = 0u;
    __syncthreads();
    if (c <= SELECT_CAP) {
        for (uint32_t i = threadIdx.x; i < c; i += GRIP_BLOCK) {
            const Candidate k = load(i);
            s_key[i] = k.d2;
            s_node[i] = k.node;
        }
        if (threadIdx.x == 0) s_at = c;
    } else if (want > 0u) {
        uint64_t prefix = 0, mask = 0;
        uint32_t rank = want - 1u;  // rank of the threshold among keys matching the prefix
        for (int shift = 56; shift >= 0; shift -= 8) {
            for (uint32_t i = threadIdx.x; i < 256u; i += GRIP_BLOCK) s_hist[i] = 0u;
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < c; i += GRIP_BLOCK) {
                const uint64_t k = key_bits(load(i).d2);
                if ((k & mask) == prefix) atomicAdd(&s_hist[(k >> shift) & 255u], 1u);
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t acc = 0, dgt = 0;
                while (acc + s_hist[dgt] <= rank) acc += s_hist[dgt++];
                s_digit = dgt;
                s_before = acc;
            }
            __syncthreads();
            prefix |= (uint64_t)s_digit << shift;
            mask |= 255ull << shift;
            rank -= s_before;
            __syncthreads();
        }
        // prefix is the threshold key T; `rank` = how many keys equal to T
        // precede the one of rank want-1, so want - (rank + 1) keys are < T
        // and the rank+1 lowest node ids among keys == T complete the set.
        const uint64_t T = prefix;
        uint32_t iprefix = 0, imask = 0, irank = rank;
        for (int shift = 24; shift >= 0; shift -= 8) {
            for (uint32_t i = threadIdx.x; i < 256u; i += GRIP_BLOCK) s_hist[i] = 0u;
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < c; i += GRIP_BLOCK) {
                const Candidate k = load(i);
                if (key_bits(k.d2) == T && (k.node & imask) == iprefix) atomicAdd(&s_hist[(k.node >> shift) & 255u], 1u);
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t acc = 0, dgt = 0;
                while (acc + s_hist[dgt] <= irank) acc += s_hist[dgt++];
                s_digit = dgt;
                s_before = acc;
            }
            __syncthreads();
            iprefix |= s_digit << shift;
            imask |= 255u << shift;
            irank -= s_before;
            __syncthreads();
        }
        const uint32_t I = iprefix;  // the largest node id kept among keys == T
        for (uint32_t i = threadIdx.x; i < c; i += GRIP_BLOCK) {
            const Candidate k = load(i);
            const uint64_t b = key_bits(k.d2);
            if (b < T || (b == T && k.node <= I)) {
                const uint32_t at = atomicAdd(&s_at, 1u);
                if (at < SELECT_CAP) {
                    s_key[at] = k.d2;
                    s_node[at] = k.node;
                }
            }
        }
    }
    __syncthreads();
    const uint32_t m = min(s_at, SELECT_CAP);
    uint32_t N = 1;
    while (N < m) N <<= 1;
    for (uint32_t i = m + threadIdx.x; i < N; i += GRIP_BLOCK) {
        s_key[i] = __longlong_as_double(0x7ff0000000000000ll);  // +inf sorts last
        s_node[i] = 0xffffffffu;
    }
    __syncthreads();
    for (uint32_t k = 2; k <= N; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < N; i += GRIP_BLOCK) {
                const uint32_t o = i ^ j;
                if (o > i) {
                    const bool up = (i & k) == 0;
                    const bool swap = up ? before(s_key[o], s_node[o], s_key[i], s_node[i])
                                         : before(s_key[i], s_node[i], s_key[o], s_node[o]);
                    if (swap) {
                        const double tk = s_key[i];
                        s_key[i] = s_key[o];
                        s_key[o] = tk;
                        const uint32_t tn = s_node[i];
                        s_node[i] = s_node[o];
                        s_node[o] = tn;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = threadIdx.x; i < want; i += GRIP_BLOCK) {
        const uint32_t n = s_node[i];
        nodes[1 + i] = n;
        offs[3 * (1 + i)] = (double)px[n] - ax;
        offs[3 * (1 + i) + 1] = (double)py[n] - ay;
        offs[3 * (1 + i) + 2] = (double)pz[n] - az;
    }
    __syncthreads();
    finish(c + 1u, want + 1u, 0u);
}

// VS_K_GRIP: the question in device memory, the counters in the output header.
__global__ __launch_bounds__(GRIP_BLOCK) void k_grip(const float *__restrict__ px, const float *__restrict__ py,
                                                     const float *__restrict__ pz, const float *__restrict__ mass,
                                                     uint32_t count, const vs_grab_params *__restrict__ gp,
                                                     Candidate *cand, uint32_t cap, uint32_t limit, GripOut *out)
{
    const vs_grab_params g = *gp;
    grip(px, py, pz, mass, count, g, cand, cap, limit, &out->count, out);
}

// vs_grip_noalloc: the question by value; counters and candidates in the
// static array below.
__global__ __launch_bounds__(GRIP_BLOCK) void k_grip_static(const float *__restrict__ px,
                                                            const float *__restrict__ py,
                                                            const float *__restrict__ pz,
                                                            const float *__restrict__ mass, uint32_t count,
                                                            vs_grab_params g, uint32_t *ctr, Candidate *cand,
                                                            uint32_t cap, uint32_t limit, GripOut *out)
{
    grip(px, py, pz, mass, count, g, cand, cap, limit, ctr, out);
}

__global__ void k_clear_counters(uint32_t *ctr)
{
    if (threadIdx.x < 2) ctr[threadIdx.x] = 0u;
}

// The device memory of vs_grip_noalloc: 256 bytes of counters, then
// VS_GRIP_NOALLOC_MAX_NODES candidates of 16 bytes (4 MiB + 256 B). Every
// device the library runs on holds its own copy until the process ends. In
// host code `scratch` names a host-side stand-in, not the device array: host
// code uses only the address grip_static_scratch() returns for the device.
constexpr size_t SCRATCH_COUNTERS = 256;
static __device__ __align__(256) unsigned char scratch[SCRATCH_COUNTERS + sizeof(Candidate) * (size_t)VS_GRIP_NOALLOC_MAX_NODES];
static_assert(sizeof(Candidate) == 16, "candidate records are 16 bytes");

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_lock;
void *g_scratch[MAX_DEVICES];

}  // namespace

hipError_t grip_static_scratch(int device, void **ptr)
{
    *ptr = nullptr;
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> hold(g_scratch_lock);
    if (!g_scratch[device]) {
        // The caller has made `device` current: the symbol resolves to that
        // device's copy of the array.
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = p;
    }
    *ptr = g_scratch[device];
    return hipSuccess;
}

hipError_t launch_grip_static(hipStream_t s, void *scratch_ptr, uint32_t grid, const float *px, const float *py,
                              const float *pz, const float *mass, uint32_t count, const vs_grab_params &params,
                              uint32_t limit, void *out)
{
    uint32_t *ctr = reinterpret_cast<uint32_t *>(scratch_ptr);
    Candidate *cand = reinterpret_cast<Candidate *>(static_cast<unsigned char *>(scratch_ptr) + SCRATCH_COUNTERS);
    hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(64), 0, s, ctr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_grip_static, dim3(grid), dim3(GRIP_BLOCK), 0, s, px, py, pz, mass, count, params, ctr, cand,
                       (uint32_t)VS_GRIP_NOALLOC_MAX_NODES, limit, reinterpret_cast<GripOut *>(out));
    return hipGetLastError();
}

hipError_t launch_pick(hipStream_t s, uint32_t grid, const float *px, const float *py, const float *pz,
                       const float *mass, uint32_t count, const vs_grab_params *params, void *out)
{
    hipLaunchKernelGGL(k_pick, dim3(grid), dim3(PICK_BLOCK), 0, s, px, py, pz, mass, count, params,
                       reinterpret_cast<PickOut *>(out));
    return hipGetLastError();
}

hipError_t launch_grip(hipStream_t s, uint32_t grid, const float *px, const float *py, const float *pz,
                       const float *mass, uint32_t count, const vs_grab_params *params, void *candidates,
                       uint32_t cap, uint32_t limit, void *out)
{
    hipLaunchKernelGGL(k_grip, dim3(grid), dim3(GRIP_BLOCK), 0, s, px, py, pz, mass, count, params,
                       reinterpret_cast<Candidate *>(candidates), cap, limit, reinterpret_cast<GripOut *>(out));
    return hipGetLastError();
}

}  // namespace vsim
