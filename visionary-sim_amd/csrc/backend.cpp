// backend.cpp -- layer 1 of vsim.h: device buffers and named kernel launches,
// the drop-in for the kernel package's ComputeBackend
// (packages/visionary-sim/src/sim/compute/backend.js:114-163).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "kernels.h"

struct vs_backend {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t written = nullptr;              // recorded after the latest write
    std::vector<std::vector<char>> pending;   // host copies of writes in flight
    std::map<uint64_t, uint64_t> allocations;  // device address -> bytes
    std::string error;
    bool destroyed = false;
    uint64_t limit = 0;
    void *grip_scratch = nullptr;            // this device's static grip array
    hipError_t grip_scratch_status = hipSuccess;
};

namespace {

std::string g_create_error;

int fail(vs_backend *b, int code, const std::string &message)
{
    if (b) b->error = message;
    return code;
}

int hip_fail(vs_backend *b, hipError_t e, const char *what)
{
    return fail(b, VS_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

const char *const KERNEL_NAMES[VS_K_COUNT] = {
    "sim-predict",         "sim-finalize",        "sim-elastic-cells", "sim-contact-bin", "sim-contact-sort",
    "sim-contact-collide", "sim-contact-settle", "sim-drive-hard",    "sim-drive-compliant", "sim-splat-deform",
    "sim-pick-ray",        "sim-grip-radius",
};

const char *const SLOT_NAMES[VS_B_COUNT] = { "nodePos", "nodePrev", "nodeVel", "nodeInvMass", "bodies",
                                             "params",  "tiles",    "items",   "contactGrid", "out" };

// Buffers each kernel reads or writes, as a bit mask over VS_B_*.
uint32_t required_slots(int kernel)
{
    const uint32_t node = (1u << VS_B_POS) | (1u << VS_B_PREV) | (1u << VS_B_INVMASS);
    const uint32_t common = (1u << VS_B_SCENES) | (1u << VS_B_TILES);
    switch (kernel) {
    case VS_K_PREDICT: return node | (1u << VS_B_VEL) | common;
    case VS_K_FINALIZE: return node | (1u << VS_B_VEL) | (1u << VS_B_BODIES) | common;
    case VS_K_ELASTIC: return node | (1u << VS_B_BODIES) | common | (1u << VS_B_ITEMS);
    case VS_K_CONTACT_BIN:
    case VS_K_CONTACT_SORT:
    case VS_K_CONTACT_SETTLE: return node | common | (1u << VS_B_ITEMS) | (1u << VS_B_GRID);
    case VS_K_CONTACT_COLLIDE: return node | (1u << VS_B_BODIES) | common | (1u << VS_B_ITEMS) | (1u << VS_B_GRID);
    case VS_K_DRIVE_HARD:
    case VS_K_DRIVE_COMPLIANT: return node | (1u << VS_B_VEL) | common | (1u << VS_B_ITEMS);
    case VS_K_SPLAT: return (1u << VS_B_POS) | (1u << VS_B_BODIES) | (1u << VS_B_TILES) | (1u << VS_B_ITEMS) |
                            (1u << VS_B_OUT);
    case VS_K_PICK: return (1u << VS_B_POS) | (1u << VS_B_INVMASS) | (1u << VS_B_SCENES) | (1u << VS_B_OUT);
    case VS_K_GRIP:
        return (1u << VS_B_POS) | (1u << VS_B_INVMASS) | (1u << VS_B_SCENES) | (1u << VS_B_ITEMS) | (1u << VS_B_OUT);
    default: return 0;
    }
}

// Bytes a slot must hold for the launch's plane sizes.
uint64_t slot_bytes(const vs_launch &l, int slot)
{
    const uint64_t P = l.plane, I = l.items;
    switch (slot) {
    case VS_B_POS:
    case VS_B_PREV:
    case VS_B_VEL: return 12 * P;
    case VS_B_INVMASS: return 4 * P;
    case VS_B_BODIES: return sizeof(vs_body);
    case VS_B_SCENES: return l.kernel == VS_K_PICK || l.kernel == VS_K_GRIP ? sizeof(vs_grab_params) : sizeof(vs_scene_params);
    case VS_B_TILES: return 16;
    case VS_B_ITEMS:
        if (l.kernel == VS_K_ELASTIC) return 4;
        if (l.kernel == VS_K_DRIVE_HARD || l.kernel == VS_K_DRIVE_COMPLIANT) return 16 * I;
        if (l.kernel == VS_K_SPLAT) return 40 * I;
        if (l.kernel == VS_K_GRIP) return 16 * I;  // candidates: f64 distance^2, u32 node, pad
        return 36 * I;  // nine surface planes
    case VS_B_GRID: return 4ull * l.buckets * (1 + 32);
    case VS_B_OUT:
        if (l.kernel == VS_K_SPLAT) return (l.flags & VS_SPLAT_AXES ? 64 : 36) * I;
        if (l.kernel == VS_K_PICK) return 16ull + 16ull * l.grid;
        if (l.kernel == VS_K_GRIP) return 32ull + 4ull * (l.pair_cap + (l.pair_cap & 1u)) + 24ull * l.pair_cap;
        return l.pair_cap ? 4ull * I * (1 + l.pair_cap) : 0;
    default: return 0;
    }
}

}  // namespace

extern "C" {

int vs_backend_create(int device, vs_backend **out)
{
    if (!out) return VS_E_INVALID;
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || device < 0 || device >= count) {
        g_create_error = "no HIP device " + std::to_string(device);
        return VS_E_DEVICE;
    }
    vs_backend *b = new vs_backend();
    b->device = device;
    if ((e = hipSetDevice(device)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->written, hipEventDisableTiming)) != hipSuccess) {
        g_create_error = std::string("stream creation failed: ") + hipGetErrorString(e);
        delete b;
        return VS_E_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) b->limit = prop.totalGlobalMem;
    // The static array of vs_grip_noalloc, looked up once per device with the
    // device current; a failure is reported by that call.
    b->grip_scratch_status = vsim::grip_static_scratch(device, &b->grip_scratch);
    *out = b;
    return VS_OK;
}

void vs_backend_destroy(vs_backend *b)
{
    if (!b) return;
    if (!b->destroyed) {
        (void)hipSetDevice(b->device);
        (void)hipStreamSynchronize(b->stream);
        for (auto &kv : b->allocations) (void)hipFree(reinterpret_cast<void *>(kv.first));
        b->allocations.clear();
        (void)hipEventDestroy(b->written);
        (void)hipStreamDestroy(b->stream);
        b->pending.clear();
        b->destroyed = true;
    }
    delete b;
}

const char *vs_last_error(const vs_backend *b) { return b ? b->error.c_str() : g_create_error.c_str(); }

int vs_buffer_alloc(vs_backend *b, uint64_t bytes, uint64_t *handle)
{
    if (!b || !handle) return VS_E_INVALID;
    if (b->destroyed) return fail(b, VS_E_DESTROYED, "Compute backend is destroyed");
    const uint64_t padded = bytes < 4 ? 4 : (bytes + 3) / 4 * 4;
    if (padded > b->limit) {
        return fail(b, VS_E_NOMEM, "Simulation buffer requires " + std::to_string(bytes) +
                                       " bytes; device storage/buffer limit is " + std::to_string(b->limit));
    }
    (void)hipSetDevice(b->device);
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, padded);
    if (e != hipSuccess) {
        return fail(b, VS_E_NOMEM, "Simulation buffer requires " + std::to_string(bytes) + " bytes; " +
                                       hipGetErrorString(e));
    }
    e = hipMemsetAsync(p, 0, padded, b->stream);
    if (e != hipSuccess) return hip_fail(b, e, "buffer clear");
    *handle = reinterpret_cast<uint64_t>(p);
    b->allocations[*handle] = padded;
    return VS_OK;
}

int vs_buffer_free(vs_backend *b, uint64_t handle)
{
    if (!b) return VS_E_INVALID;
    auto it = b->allocations.find(handle);
    if (it == b->allocations.end()) return VS_OK;  // freeing twice is a no-op
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    (void)hipFree(reinterpret_cast<void *>(handle));
    b->allocations.erase(it);
    return VS_OK;
}

static int check_range(vs_backend *b, uint64_t handle, uint64_t offset, uint64_t bytes)
{
    auto it = b->allocations.find(handle);
    if (it == b->allocations.end()) return fail(b, VS_E_INVALID, "invalid buffer handle");
    if (offset > it->second || bytes > it->second - offset) {
        return fail(b, VS_E_INVALID, "range [" + std::to_string(offset) + ", +" + std::to_string(bytes) +
                                         ") exceeds buffer of " + std::to_string(it->second) + " bytes");
    }
    return VS_OK;
}

int vs_buffer_write(vs_backend *b, uint64_t handle, uint64_t offset, const void *src, uint64_t bytes)
{
    if (!b || (!src && bytes)) return VS_E_INVALID;
    if (b->destroyed) return fail(b, VS_E_DESTROYED, "Compute backend is destroyed");
    if (!bytes) return VS_OK;
    if (int rc = check_range(b, handle, offset, bytes)) return rc;
    (void)hipSetDevice(b->device);
    // The caller may reuse `src` at once: copy it, and keep the copy until
    // the stream has consumed it.
    if (hipEventQuery(b->written) == hipSuccess) b->pending.clear();
    b->pending.emplace_back(static_cast<const char *>(src), static_cast<const char *>(src) + bytes);
    hipError_t e = hipMemcpyAsync(reinterpret_cast<char *>(handle) + offset, b->pending.back().data(), bytes,
                                  hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipEventRecord(b->written, b->stream);
    if (e != hipSuccess) return hip_fail(b, e, "buffer write");
    return VS_OK;
}

int vs_buffer_read(vs_backend *b, uint64_t handle, uint64_t offset, uint64_t bytes, void *dst)
{
    if (!b || (!dst && bytes)) return VS_E_INVALID;
    if (b->destroyed) return fail(b, VS_E_DESTROYED, "Compute backend is destroyed");
    if (!bytes) return VS_OK;
    if (int rc = check_range(b, handle, offset, bytes)) return rc;
    (void)hipSetDevice(b->device);
    hipError_t e = hipMemcpyAsync(dst, reinterpret_cast<char *>(handle) + offset, bytes, hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return hip_fail(b, e, "buffer read");
    return VS_OK;
}

int vs_kernel_id(const char *name)
{
    if (!name) return VS_E_INVALID;
    for (int k = 0; k < VS_K_COUNT; k++) {
        if (std::strcmp(name, KERNEL_NAMES[k]) == 0) return k;
    }
    return VS_E_INVALID;
}

int vs_synchronize(vs_backend *b)
{
    if (!b) return VS_E_INVALID;
    if (b->destroyed) return fail(b, VS_E_DESTROYED, "Compute backend is destroyed");
    (void)hipSetDevice(b->device);
    hipError_t e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? VS_OK : hip_fail(b, e, "synchronize");
}

int vs_dispatch(vs_backend *b, const vs_launch *launches, uint32_t count)
{
    using namespace vsim;
    if (!b || (!launches && count)) return VS_E_INVALID;
    if (b->destroyed) return fail(b, VS_E_DESTROYED, "Compute backend is destroyed");
    // Validate everything before enqueuing anything (backend.js:142-148).
    for (uint32_t i = 0; i < count; i++) {
        const vs_launch &l = launches[i];
        if (l.kernel < 0 || l.kernel >= VS_K_COUNT) return fail(b, VS_E_INVALID, "unknown kernel " + std::to_string(l.kernel));
        const char *name = KERNEL_NAMES[l.kernel];
        const uint32_t block = l.block ? l.block : 64;
        if (l.grid < 1 || l.grid > 0x7fffffffu || (block != 64 && block != 128 && block != 256 && block != 512)) {
            return fail(b, VS_E_INVALID, std::string(name) + ": invalid dispatch " + std::to_string(l.grid) + "," +
                                             std::to_string(block));
        }
        if (l.plane >= (1u << 29)) return fail(b, VS_E_INVALID, std::string(name) + ": plane exceeds 2^29 floats");
        const uint32_t need = required_slots(l.kernel);
        for (int s = 0; s < VS_B_COUNT; s++) {
            const bool wanted = (need >> s) & 1u;
            const bool log = s == VS_B_OUT && l.kernel == VS_K_CONTACT_COLLIDE && l.pair_cap;
            if (!wanted && !log) continue;
            auto it = b->allocations.find(l.buffers[s]);
            if (it == b->allocations.end() || it->second < slot_bytes(l, s)) {
                return fail(b, VS_E_INVALID, std::string(name) + ": invalid buffer " + SLOT_NAMES[s]);
            }
        }
    }
    (void)hipSetDevice(b->device);
    for (uint32_t i = 0; i < count; i++) {
        const vs_launch &l = launches[i];
        auto F = [&](int slot) { return reinterpret_cast<float *>(l.buffers[slot]); };
        auto U = [&](int slot) { return reinterpret_cast<uint32_t *>(l.buffers[slot]); };
        const uint64_t P = l.plane, I = l.items;
        NodeView n{ F(VS_B_POS),  F(VS_B_POS) + P,  F(VS_B_POS) + 2 * P, F(VS_B_PREV), F(VS_B_PREV) + P,
                    F(VS_B_PREV) + 2 * P, F(VS_B_VEL), F(VS_B_VEL) + P, F(VS_B_VEL) + 2 * P, F(VS_B_INVMASS) };
        TileView t{ U(VS_B_TILES), reinterpret_cast<const vs_tile *>(U(VS_B_TILES) + 4) };
        const vs_body *bodies = reinterpret_cast<const vs_body *>(l.buffers[VS_B_BODIES]);
        const vs_scene_params *scenes = reinterpret_cast<const vs_scene_params *>(l.buffers[VS_B_SCENES]);
        Shape sh{ l.grid, l.block ? l.block : 64u, 0u };
        SurfaceView sv{};
        GridView g{};
        if (l.kernel >= VS_K_CONTACT_BIN && l.kernel <= VS_K_CONTACT_SETTLE) {
            uint32_t *it = U(VS_B_ITEMS);
            sv = SurfaceView{ it, it + I, reinterpret_cast<float *>(it + 2 * I), it + 3 * I, it + 4 * I,
                              reinterpret_cast<float *>(it + 5 * I), reinterpret_cast<float *>(it + 6 * I),
                              reinterpret_cast<float *>(it + 7 * I), reinterpret_cast<float *>(it + 8 * I) };
            g = GridView{ U(VS_B_GRID), U(VS_B_GRID) + l.buckets };
        }
        hipError_t e = hipSuccess;
        switch (l.kernel) {
        case VS_K_PREDICT: e = launch_predict(b->stream, sh, n, scenes, t); break;
        case VS_K_FINALIZE: e = launch_finalize(b->stream, sh, n, bodies, scenes, t); break;
        case VS_K_ELASTIC: e = launch_elastic(b->stream, sh, n, bodies, scenes, U(VS_B_ITEMS), t, l.flags & 1u); break;
        case VS_K_CONTACT_BIN: e = launch_contact_bin(b->stream, sh, n, scenes, sv, g, t); break;
        case VS_K_CONTACT_SORT: e = launch_contact_sort(b->stream, sh, scenes, sv, g, t); break;
        case VS_K_CONTACT_COLLIDE: {
            PairLog log{ nullptr, nullptr, 0 };
            if (l.pair_cap) {
                int32_t *c = reinterpret_cast<int32_t *>(l.buffers[VS_B_OUT]);
                log = PairLog{ c, reinterpret_cast<uint32_t *>(c + I), l.pair_cap };
            }
            e = launch_contact_collide(b->stream, sh, n, bodies, scenes, sv, g, t, log, 256u);
            break;
        }
        case VS_K_CONTACT_SETTLE: e = launch_contact_settle(b->stream, sh, n, scenes, sv, g, t); break;
        case VS_K_DRIVE_HARD:
        case VS_K_DRIVE_COMPLIANT: {
            uint32_t *it = U(VS_B_ITEMS);
            DriveView d{ it, reinterpret_cast<float *>(it + I), reinterpret_cast<float *>(it + 2 * I),
                         reinterpret_cast<float *>(it + 3 * I) };
            e = launch_drive(b->stream, sh, n, scenes, d, t, l.kernel == VS_K_DRIVE_COMPLIANT ? 1 : 0);
            break;
        }
        case VS_K_PICK:
            e = launch_pick(b->stream, l.grid, n.px, n.py, n.pz, n.w, l.plane,
                            reinterpret_cast<const vs_grab_params *>(l.buffers[VS_B_SCENES]),
                            reinterpret_cast<void *>(l.buffers[VS_B_OUT]));
            break;
        case VS_K_GRIP:
            e = launch_grip(b->stream, l.grid, n.px, n.py, n.pz, n.w, l.plane,
                            reinterpret_cast<const vs_grab_params *>(l.buffers[VS_B_SCENES]),
                            reinterpret_cast<void *>(l.buffers[VS_B_ITEMS]), l.items, l.pair_cap,
                            reinterpret_cast<void *>(l.buffers[VS_B_OUT]));
            break;
        case VS_K_SPLAT: {
            float *in = F(VS_B_ITEMS), *out = F(VS_B_OUT);
            const uint32_t mode = l.flags & (VS_SPLAT_COV | VS_SPLAT_AXES) ? l.flags & (VS_SPLAT_COV | VS_SPLAT_AXES)
                                                                            : VS_SPLAT_COV;
            SplatView sp{ in + 4 * I, in + 5 * I,  in + 6 * I,  reinterpret_cast<const float4 *>(in),
                          in + 7 * I, in + 8 * I,  in + 9 * I,  out,
                          out + I,    out + 2 * I, out + 3 * I, out + 9 * I,
                          out + 13 * I, (uint32_t)I, mode };
            e = launch_splats(b->stream, sh, n.px, n.py, n.pz, bodies, sp, t);
            break;
        }
        }
        if (e != hipSuccess) return hip_fail(b, e, KERNEL_NAMES[l.kernel]);
    }
    return VS_OK;
}

int vs_grip_noalloc(vs_backend *b, uint64_t pos, uint32_t plane, uint64_t mass, const vs_grab_params *params,
                    uint32_t limit, uint32_t grid, uint64_t out)
{
    if (!b || !params) return VS_E_INVALID;
    if (b->destroyed) return fail(b, VS_E_DESTROYED, "Compute backend is destroyed");
    const char *name = "sim-grip-radius";
    if (grid < 1 || grid > 1024) {
        return fail(b, VS_E_INVALID, std::string(name) + ": invalid dispatch " + std::to_string(grid));
    }
    if (plane >= (1u << 29)) return fail(b, VS_E_INVALID, std::string(name) + ": plane exceeds 2^29 floats");
    if (params->end > params->begin && params->end - params->begin > VS_GRIP_NOALLOC_MAX_NODES) {
        return fail(b, VS_E_INVALID, std::string(name) + ": body of " + std::to_string(params->end - params->begin) +
                                         " nodes exceeds the static candidate array of " +
                                         std::to_string(VS_GRIP_NOALLOC_MAX_NODES) + " nodes");
    }
    vs_launch l{};
    l.kernel = VS_K_GRIP;
    l.grid = grid;
    l.plane = plane;
    l.pair_cap = limit;
    const std::pair<int, uint64_t> slots[] = { { VS_B_POS, pos }, { VS_B_INVMASS, mass }, { VS_B_OUT, out } };
    for (const auto &sl : slots) {
        auto it = b->allocations.find(sl.second);
        if (it == b->allocations.end() || it->second < slot_bytes(l, sl.first)) {
            return fail(b, VS_E_INVALID, std::string(name) + ": invalid buffer " + SLOT_NAMES[sl.first]);
        }
    }
    if (b->grip_scratch_status != hipSuccess || !b->grip_scratch) {
        return hip_fail(b, b->grip_scratch_status != hipSuccess ? b->grip_scratch_status : hipErrorInvalidSymbol,
                        "sim-grip-radius static array");
    }
    (void)hipSetDevice(b->device);
    const float *px = reinterpret_cast<const float *>(pos);
    const hipError_t e =
        vsim::launch_grip_static(b->stream, b->grip_scratch, grid, px, px + plane, px + 2 * (uint64_t)plane,
                                 reinterpret_cast<const float *>(mass), plane, *params, limit,
                                 reinterpret_cast<void *>(out));
    return e == hipSuccess ? VS_OK : hip_fail(b, e, name);
}

}  // extern "C"
