// addon.cpp -- the Node-API binding of the compute backend in include/vsim.h.
//
// This is the piece a maintainer of the kernel package adds so that its
// ComputeBackend (packages/visionary-sim/src/sim/compute/backend.js:114-163)
// runs on HIP instead of WebGPU: every export forwards to one vs_* entry
// point, and every failure becomes a JS Error carrying vs_last_error(). The
// ingest exports (voxelisers, PLY) do the same over vs_ingest_error().
#include <node_api.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vsim.h"

namespace {

#define CALL(env, expr)                                                  \
    do {                                                                 \
        if ((expr) != napi_ok) {                                         \
            napi_throw_error((env), nullptr, "napi call failed: " #expr); \
            return nullptr;                                              \
        }                                                                \
    } while (0)

napi_value fail(napi_env env, const vs_backend *b)
{
    napi_throw_error(env, nullptr, vs_last_error(b));
    return nullptr;
}

bool get_args(napi_env env, napi_callback_info info, size_t want, napi_value *argv)
{
    size_t argc = want;
    if (napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr) != napi_ok || argc < want) {
        napi_throw_type_error(env, nullptr, "missing arguments");
        return false;
    }
    return true;
}

vs_backend *backend_of(napi_env env, napi_value v)
{
    void *p = nullptr;
    if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
        napi_throw_type_error(env, nullptr, "not a backend handle");
        return nullptr;
    }
    return static_cast<vs_backend *>(p);
}

bool u64_of(napi_env env, napi_value v, uint64_t *out)
{
    bool lossless = false;
    if (napi_get_value_bigint_uint64(env, v, out, &lossless) == napi_ok) return true;
    double d = 0;
    if (napi_get_value_double(env, v, &d) != napi_ok || d < 0) {
        napi_throw_type_error(env, nullptr, "expected a non-negative integer");
        return false;
    }
    *out = (uint64_t)d;
    return true;
}

bool bytes_of(napi_env env, napi_value v, void **data, size_t *length)
{
    bool is_ta = false;
    napi_is_typedarray(env, v, &is_ta);
    if (is_ta) {
        napi_typedarray_type type;
        size_t count = 0, offset = 0;
        napi_value ab;
        if (napi_get_typedarray_info(env, v, &type, &count, data, &ab, &offset) != napi_ok) return false;
        size_t elem = 1;
        switch (type) {
        case napi_int16_array:
        case napi_uint16_array: elem = 2; break;
        case napi_int32_array:
        case napi_uint32_array:
        case napi_float32_array: elem = 4; break;
        case napi_float64_array:
        case napi_bigint64_array:
        case napi_biguint64_array: elem = 8; break;
        default: elem = 1;
        }
        *length = count * elem;
        return true;
    }
    bool is_ab = false;
    napi_is_arraybuffer(env, v, &is_ab);
    if (is_ab) return napi_get_arraybuffer_info(env, v, data, length) == napi_ok;
    napi_throw_type_error(env, nullptr, "expected a TypedArray or ArrayBuffer");
    return false;
}

napi_value Create(napi_env env, napi_callback_info info)
{
    napi_value argv[1];
    if (!get_args(env, info, 1, argv)) return nullptr;
    int32_t device = 0;
    CALL(env, napi_get_value_int32(env, argv[0], &device));
    vs_backend *b = nullptr;
    if (vs_backend_create(device, &b) != VS_OK) return fail(env, nullptr);
    napi_value out;
    CALL(env, napi_create_external(env, b, nullptr, nullptr, &out));
    return out;
}

napi_value Destroy(napi_env env, napi_callback_info info)
{
    napi_value argv[1];
    if (!get_args(env, info, 1, argv)) return nullptr;
    vs_backend *b = backend_of(env, argv[0]);
    if (!b) return nullptr;
    vs_backend_destroy(b);
    return nullptr;
}

napi_value Alloc(napi_env env, napi_callback_info info)
{
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    vs_backend *b = backend_of(env, argv[0]);
    uint64_t bytes = 0, handle = 0;
    if (!b || !u64_of(env, argv[1], &bytes)) return nullptr;
    if (vs_buffer_alloc(b, bytes, &handle) != VS_OK) return fail(env, b);
    napi_value out;
    CALL(env, napi_create_bigint_uint64(env, handle, &out));
    return out;
}

napi_value Free(napi_env env, napi_callback_info info)
{
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    vs_backend *b = backend_of(env, argv[0]);
    uint64_t handle = 0;
    if (!b || !u64_of(env, argv[1], &handle)) return nullptr;
    if (vs_buffer_free(b, handle) != VS_OK) return fail(env, b);
    return nullptr;
}

napi_value Write(napi_env env, napi_callback_info info)
{
    napi_value argv[4];
    if (!get_args(env, info, 4, argv)) return nullptr;
    vs_backend *b = backend_of(env, argv[0]);
    uint64_t handle = 0, offset = 0;
    void *data = nullptr;
    size_t length = 0;
    if (!b || !u64_of(env, argv[1], &handle) || !u64_of(env, argv[2], &offset) || !bytes_of(env, argv[3], &data, &length))
        return nullptr;
    if (vs_buffer_write(b, handle, offset, data, length) != VS_OK) return fail(env, b);
    return nullptr;
}

napi_value Read(napi_env env, napi_callback_info info)
{
    napi_value argv[4];
    if (!get_args(env, info, 4, argv)) return nullptr;
    vs_backend *b = backend_of(env, argv[0]);
    uint64_t handle = 0, offset = 0;
    void *data = nullptr;
    size_t length = 0;
    if (!b || !u64_of(env, argv[1], &handle) || !u64_of(env, argv[2], &offset) || !bytes_of(env, argv[3], &data, &length))
        return nullptr;
    if (vs_buffer_read(b, handle, offset, length, data) != VS_OK) return fail(env, b);
    return nullptr;
}

napi_value KernelId(napi_env env, napi_callback_info info)
{
    napi_value argv[1];
    if (!get_args(env, info, 1, argv)) return nullptr;
    char name[128];
    size_t len = 0;
    CALL(env, napi_get_value_string_utf8(env, argv[0], name, sizeof name, &len));
    napi_value out;
    CALL(env, napi_create_int32(env, vs_kernel_id(name), &out));
    return out;
}

napi_value Dispatch(napi_env env, napi_callback_info info)
{
    napi_value argv[3];
    if (!get_args(env, info, 3, argv)) return nullptr;
    vs_backend *b = backend_of(env, argv[0]);
    void *data = nullptr;
    size_t length = 0;
    uint32_t count = 0;
    if (!b || !bytes_of(env, argv[1], &data, &length)) return nullptr;
    CALL(env, napi_get_value_uint32(env, argv[2], &count));
    if ((size_t)count * sizeof(vs_launch) > length) {
        napi_throw_range_error(env, nullptr, "launch buffer shorter than the launch count");
        return nullptr;
    }
    if (vs_dispatch(b, static_cast<const vs_launch *>(data), count) != VS_OK) return fail(env, b);
    return nullptr;
}

napi_value Sync(napi_env env, napi_callback_info info)
{
    napi_value argv[1];
    if (!get_args(env, info, 1, argv)) return nullptr;
    vs_backend *b = backend_of(env, argv[0]);
    if (!b) return nullptr;
    if (vs_synchronize(b) != VS_OK) return fail(env, b);
    return nullptr;
}

napi_value LaunchSize(napi_env env, napi_callback_info)
{
    napi_value out;
    CALL(env, napi_create_uint32(env, (uint32_t)sizeof(vs_launch), &out));
    return out;
}

// ------------------------------------------------------------ ingest
//
// The voxelisers and the PLY reader: plain objects and typed arrays in,
// { dims, origin, h, occupancy } grids and Float32Array clouds out. Every
// failure throws an Error carrying vs_ingest_error().

napi_value ingest_fail(napi_env env)
{
    napi_throw_error(env, nullptr, vs_ingest_error());
    return nullptr;
}

bool typed_of(napi_env env, napi_value v, napi_typedarray_type want, void **data, size_t *count, const char *what)
{
    bool is_ta = false;
    napi_is_typedarray(env, v, &is_ta);
    napi_typedarray_type type;
    napi_value ab;
    size_t offset = 0;
    if (!is_ta || napi_get_typedarray_info(env, v, &type, count, data, &ab, &offset) != napi_ok || type != want) {
        napi_throw_type_error(env, nullptr, what);
        return false;
    }
    return true;
}

bool number_of(napi_env env, napi_value v, double *out)
{
    if (napi_get_value_double(env, v, out) != napi_ok) {
        napi_throw_type_error(env, nullptr, "expected a number");
        return false;
    }
    return true;
}

napi_value typed_copy(napi_env env, napi_typedarray_type type, const void *src, size_t count, size_t elem)
{
    void *data = nullptr;
    napi_value ab, out;
    if (napi_create_arraybuffer(env, count * elem, &data, &ab) != napi_ok) return nullptr;
    if (count) std::memcpy(data, src, count * elem);
    if (napi_create_typedarray(env, type, count, ab, 0, &out) != napi_ok) return nullptr;
    return out;
}

napi_value number_array(napi_env env, const double *v, uint32_t n)
{
    napi_value arr, x;
    napi_create_array_with_length(env, n, &arr);
    for (uint32_t i = 0; i < n; i++) {
        napi_create_double(env, v[i], &x);
        napi_set_element(env, arr, i, x);
    }
    return arr;
}

// OccupancyGrid (voxelize.js:21-29); frees the native grid.
napi_value grid_object(napi_env env, vs_occupancy &g)
{
    const size_t n = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
    const double dims[3] = { (double)g.dims[0], (double)g.dims[1], (double)g.dims[2] };
    napi_value obj, h, occ = typed_copy(env, napi_uint8_array, g.cells, n, 1);
    vs_occupancy_free(&g);
    if (!occ) return nullptr;
    napi_create_object(env, &obj);
    napi_create_double(env, g.h, &h);
    napi_set_named_property(env, obj, "dims", number_array(env, dims, 3));
    napi_set_named_property(env, obj, "origin", number_array(env, g.origin, 3));
    napi_set_named_property(env, obj, "h", h);
    napi_set_named_property(env, obj, "occupancy", occ);
    return obj;
}

// voxelizeCloud(device, positions, rotations, scales, elementSize, padding, sigma, seal, fill)
napi_value VoxelizeCloud(napi_env env, napi_callback_info info)
{
    napi_value argv[9];
    if (!get_args(env, info, 9, argv)) return nullptr;
    int32_t device = 0;
    uint32_t padding = 0, seal = 0;
    bool fill = true;
    double element = 0, sigma = 0;
    void *p = nullptr, *r = nullptr, *sc = nullptr;
    size_t np = 0, nr = 0, ns = 0;
    CALL(env, napi_get_value_int32(env, argv[0], &device));
    if (!typed_of(env, argv[1], napi_float32_array, &p, &np, "positions must be a Float32Array") ||
        !typed_of(env, argv[2], napi_float32_array, &r, &nr, "rotations must be a Float32Array") ||
        !typed_of(env, argv[3], napi_float32_array, &sc, &ns, "scales must be a Float32Array") ||
        !number_of(env, argv[4], &element) || !number_of(env, argv[6], &sigma))
        return nullptr;
    CALL(env, napi_get_value_uint32(env, argv[5], &padding));
    CALL(env, napi_get_value_uint32(env, argv[7], &seal));
    CALL(env, napi_get_value_bool(env, argv[8], &fill));
    const uint64_t count = np / 3;
    if (nr < 4 * count || ns < 3 * count) {
        napi_throw_range_error(env, nullptr, "rotations and scales are shorter than the positions");
        return nullptr;
    }
    vs_occupancy g{};
    if (vs_voxelize_cloud(device, static_cast<float *>(p), static_cast<float *>(r), static_cast<float *>(sc), count,
                          element, padding, sigma, seal, fill ? 1 : 0, &g) != VS_OK)
        return ingest_fail(env);
    return grid_object(env, g);
}

// voxelizeMesh(device, positions, indices, bounds | null, elementSize, padding, close)
napi_value VoxelizeMesh(napi_env env, napi_callback_info info)
{
    napi_value argv[7];
    if (!get_args(env, info, 7, argv)) return nullptr;
    int32_t device = 0;
    uint32_t padding = 0;
    bool close = true;
    double element = 0;
    void *p = nullptr, *ix = nullptr, *b = nullptr;
    size_t np = 0, ni = 0, nb = 0;
    CALL(env, napi_get_value_int32(env, argv[0], &device));
    if (!typed_of(env, argv[1], napi_float32_array, &p, &np, "positions must be a Float32Array") ||
        !typed_of(env, argv[2], napi_uint32_array, &ix, &ni, "indices must be a Uint32Array"))
        return nullptr;
    napi_valuetype bt;
    CALL(env, napi_typeof(env, argv[3], &bt));
    if (bt != napi_null && bt != napi_undefined) {
        if (!typed_of(env, argv[3], napi_float64_array, &b, &nb, "bounds must be a Float64Array of min, max")) return nullptr;
        if (nb < 6) {
            napi_throw_range_error(env, nullptr, "bounds must hold min[3] and max[3]");
            return nullptr;
        }
    }
    if (!number_of(env, argv[4], &element)) return nullptr;
    CALL(env, napi_get_value_uint32(env, argv[5], &padding));
    CALL(env, napi_get_value_bool(env, argv[6], &close));
    vs_occupancy g{};
    if (vs_voxelize_mesh(device, static_cast<float *>(p), np / 3, static_cast<uint32_t *>(ix), ni / 3,
                         static_cast<double *>(b), element, padding, close ? 1 : 0, &g) != VS_OK)
        return ingest_fail(env);
    return grid_object(env, g);
}

// voxelizeShape(device, kind, size Float64Array(3), offsets Float64Array(3 n), elementSize, padding)
napi_value VoxelizeShape(napi_env env, napi_callback_info info)
{
    napi_value argv[6];
    if (!get_args(env, info, 6, argv)) return nullptr;
    int32_t device = 0, kind = 0;
    uint32_t padding = 0;
    double element = 0;
    void *size = nullptr, *offs = nullptr;
    size_t nsize = 0, noffs = 0;
    CALL(env, napi_get_value_int32(env, argv[0], &device));
    CALL(env, napi_get_value_int32(env, argv[1], &kind));
    if (!typed_of(env, argv[2], napi_float64_array, &size, &nsize, "size must be a Float64Array") ||
        !typed_of(env, argv[3], napi_float64_array, &offs, &noffs, "offsets must be a Float64Array") ||
        !number_of(env, argv[4], &element))
        return nullptr;
    CALL(env, napi_get_value_uint32(env, argv[5], &padding));
    if (nsize < 3 || noffs % 3 || noffs / 3 > 4) {
        napi_throw_range_error(env, nullptr, "size holds 3 numbers and offsets at most 4 translations");
        return nullptr;
    }
    vs_shape sh{};
    sh.kind = kind;
    sh.n_offsets = (uint32_t)(noffs / 3);
    std::memcpy(sh.size, size, 3 * sizeof(double));
    std::memcpy(sh.offset, offs, noffs * sizeof(double));
    vs_occupancy g{};
    if (vs_voxelize_shape(device, &sh, element, padding, &g) != VS_OK) return ingest_fail(env);
    return grid_object(env, g);
}

// plyProperties(buffer) -> { count, names }
napi_value PlyProperties(napi_env env, napi_callback_info info)
{
    napi_value argv[1];
    if (!get_args(env, info, 1, argv)) return nullptr;
    void *data = nullptr;
    size_t length = 0;
    if (!bytes_of(env, argv[0], &data, &length)) return nullptr;
    std::string list(length + 1, '\0');
    uint64_t count = 0;
    if (vs_ply_properties(data, length, &list[0], list.size(), &count) != VS_OK) return ingest_fail(env);
    list.resize(std::strlen(list.c_str()));
    napi_value obj, c, names, name;
    napi_create_object(env, &obj);
    napi_create_double(env, (double)count, &c);
    napi_create_array(env, &names);
    uint32_t at = 0;
    size_t i = 0;
    while (i < list.size()) {
        size_t j = list.find(' ', i);
        if (j == std::string::npos) j = list.size();
        napi_create_string_utf8(env, list.data() + i, j - i, &name);
        napi_set_element(env, names, at++, name);
        i = j + 1;
    }
    napi_set_named_property(env, obj, "count", c);
    napi_set_named_property(env, obj, "names", names);
    return obj;
}

// readPly(buffer, names) -> { count, properties: { name: Float32Array } }
napi_value ReadPly(napi_env env, napi_callback_info info)
{
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    void *data = nullptr;
    size_t length = 0;
    if (!bytes_of(env, argv[0], &data, &length)) return nullptr;
    uint32_t n = 0;
    CALL(env, napi_get_array_length(env, argv[1], &n));
    std::vector<std::string> names(n);
    std::vector<const char *> ptrs(n);
    for (uint32_t i = 0; i < n; i++) {
        napi_value v;
        size_t len = 0;
        CALL(env, napi_get_element(env, argv[1], i, &v));
        CALL(env, napi_get_value_string_utf8(env, v, nullptr, 0, &len));
        names[i].resize(len + 1);
        CALL(env, napi_get_value_string_utf8(env, v, &names[i][0], len + 1, &len));
        names[i].resize(len);
        ptrs[i] = names[i].c_str();
    }
    std::vector<float *> cols(n, nullptr);
    uint64_t count = 0;
    if (vs_ply_read(data, length, ptrs.data(), n, &count, cols.data()) != VS_OK) return ingest_fail(env);
    napi_value obj, props, c;
    napi_create_object(env, &obj);
    napi_create_object(env, &props);
    bool ok = true;
    for (uint32_t i = 0; i < n; i++) {
        napi_value col = ok ? typed_copy(env, napi_float32_array, cols[i], count, 4) : nullptr;
        ok = ok && col && napi_set_named_property(env, props, ptrs[i], col) == napi_ok;
        std::free(cols[i]);
    }
    if (!ok) {
        napi_throw_error(env, nullptr, "PLY columns do not fit");
        return nullptr;
    }
    napi_create_double(env, (double)count, &c);
    napi_set_named_property(env, obj, "count", c);
    napi_set_named_property(env, obj, "properties", props);
    return obj;
}

// cloudFromColumns(count, [14 Float32Array in REQUIRED_PROPERTIES order])
napi_value CloudFromColumns(napi_env env, napi_callback_info info)
{
    napi_value argv[2];
    if (!get_args(env, info, 2, argv)) return nullptr;
    double count_d = 0;
    if (!number_of(env, argv[0], &count_d)) return nullptr;
    const uint64_t count = (uint64_t)count_d;
    uint32_t n = 0;
    CALL(env, napi_get_array_length(env, argv[1], &n));
    if (n != 14) {
        napi_throw_range_error(env, nullptr, "cloudFromPly needs the 14 required columns");
        return nullptr;
    }
    const float *cols[14];
    for (uint32_t i = 0; i < 14; i++) {
        napi_value v;
        void *d = nullptr;
        size_t len = 0;
        CALL(env, napi_get_element(env, argv[1], i, &v));
        if (!typed_of(env, v, napi_float32_array, &d, &len, "columns must be Float32Arrays")) return nullptr;
        if (len < count) {
            napi_throw_range_error(env, nullptr, "a column is shorter than the record count");
            return nullptr;
        }
        cols[i] = static_cast<const float *>(d);
    }
    vs_cloud cloud{};
    if (vs_cloud_from_columns(count, cols, &cloud) != VS_OK) return ingest_fail(env);
    napi_value obj, c;
    napi_create_object(env, &obj);
    napi_create_double(env, (double)count, &c);
    napi_set_named_property(env, obj, "count", c);
    const struct { const char *name; const float *data; size_t width; } fields[] = {
        { "positions", cloud.positions, 3 }, { "rotations", cloud.rotations, 4 }, { "scales", cloud.scales, 3 },
        { "opacity", cloud.opacity, 1 },     { "colors", cloud.colors, 3 },
    };
    for (const auto &f : fields) {
        napi_value arr = typed_copy(env, napi_float32_array, f.data, count * f.width, 4);
        if (!arr) {
            vs_cloud_free(&cloud);
            napi_throw_error(env, nullptr, "cloud does not fit");
            return nullptr;
        }
        napi_set_named_property(env, obj, f.name, arr);
    }
    vs_cloud_free(&cloud);
    return obj;
}

napi_value Init(napi_env env, napi_value exports)
{
    napi_property_descriptor props[] = {
        { "create", nullptr, Create, nullptr, nullptr, nullptr, napi_default, nullptr },
        { "destroy", nullptr, Destroy, nullptr, nullptr, nullptr, napi_default, nullptr },
        { "alloc", nullptr, Alloc, nullptr, nullptr, nullptr, napi_default, nullptr },
        { "free", nullptr, Free, nullptr, nullptr, nullptr, napi_default, nullptr },
        { "write", nullptr, Write, nullptr, nullptr, nullptr, napi_default, nullptr },
        { "read", nullptr, Read, nullptr, nullptr, nullptr, napi_default, nullptr },
        { "kernelId", nullptr, KernelId, nullptr, nullptr, nullptr, napi_default, nullptr },
        { "dispatch", nullptr, Dispatch, nullptr, nullptr, nullptr, napi_default, nullptr },
        { "synchronize", nullptr, Sync, nullptr, nullptr, nullptr, napi_default, nullptr },
        { "launchSize", nullptr, LaunchSize, nullptr, nullptr, nullptr, napi_default, nullptr },
        { "voxelizeCloud", nullptr, VoxelizeCloud, nullptr, nullptr, nullptr, napi_default, nullptr },
        { "voxelizeMesh", nullptr, VoxelizeMesh, nullptr, nullptr, nullptr, napi_default, nullptr },
        { "voxelizeShape", nullptr, VoxelizeShape, nullptr, nullptr, nullptr, napi_default, nullptr },
        { "plyProperties", nullptr, PlyProperties, nullptr, nullptr, nullptr, napi_default, nullptr },
        { "readPly", nullptr, ReadPly, nullptr, nullptr, nullptr, napi_default, nullptr },
        { "cloudFromColumns", nullptr, CloudFromColumns, nullptr, nullptr, nullptr, napi_default, nullptr },
    };
    if (napi_define_properties(env, exports, sizeof props / sizeof props[0], props) != napi_ok) return nullptr;
    return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
