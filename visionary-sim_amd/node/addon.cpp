// addon.cpp -- the Node-API binding of the compute backend in include/vsim.h.
//
// This is the piece a maintainer of the kernel package adds so that its
// ComputeBackend (packages/visionary-sim/src/sim/compute/backend.js:114-163)
// runs on HIP instead of WebGPU: every export forwards to one vs_* entry
// point, and every failure becomes a JS Error carrying vs_last_error().
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>

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
    };
    if (napi_define_properties(env, exports, sizeof props / sizeof props[0], props) != napi_ok) return nullptr;
    return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
