// ply.cpp -- the PLY container and the Gaussian cloud read out of it, on the
// host: readPly (apps/simulator/src/assets/splats/ply.js:65-152) and
// cloudFromPly (apps/simulator/src/assets/splats/cloud.js:52-120).
//
// Only binary little-endian is read, the first element only, as the
// reference does; the error messages are the reference's. Values of every
// scalar type become f32, as the reference's Float32Array columns hold them;
// the activations (exp scale, sigmoid opacity, degree-0 colour, unit
// quaternion with a non-negative real part) are evaluated in double and
// rounded to f32 on store, like the JS.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.h"

namespace vsim {

namespace {
thread_local std::string g_ingest_error;
}

int ingest_fail(int code, const std::string &message)
{
    g_ingest_error = message;
    return code;
}

namespace {

struct Property {
    std::string type, name;
    uint32_t size;
};

struct Element {
    std::string name;
    uint64_t count = 0;
    std::vector<Property> properties;
};

uint32_t type_size(const std::string &t)
{
    if (t == "char" || t == "int8" || t == "uchar" || t == "uint8") return 1;
    if (t == "short" || t == "int16" || t == "ushort" || t == "uint16") return 2;
    if (t == "int" || t == "int32" || t == "uint" || t == "uint32" || t == "float" || t == "float32") return 4;
    if (t == "double" || t == "float64") return 8;
    return 0;
}

// One stored value, little-endian, as the f32 a Float32Array keeps.
float read_value(const std::string &t, const uint8_t *at)
{
    uint8_t b[8];
    std::memcpy(b, at, type_size(t));
    auto u16 = [&] { return (uint16_t)(b[0] | (b[1] << 8)); };
    auto u32 = [&] { return (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24); };
    if (t == "char" || t == "int8") return (float)(int8_t)b[0];
    if (t == "uchar" || t == "uint8") return (float)b[0];
    if (t == "short" || t == "int16") return (float)(int16_t)u16();
    if (t == "ushort" || t == "uint16") return (float)u16();
    if (t == "int" || t == "int32") return (float)(int32_t)u32();
    if (t == "uint" || t == "uint32") return (float)u32();
    if (t == "float" || t == "float32") {
        const uint32_t bits = u32();
        float f;
        std::memcpy(&f, &bits, 4);
        return f;
    }
    const uint64_t bits = (uint64_t)u32() | ((uint64_t)((uint32_t)b[4] | ((uint32_t)b[5] << 8) | ((uint32_t)b[6] << 16) |
                                                         ((uint32_t)b[7] << 24))
                                             << 32);
    double d;
    std::memcpy(&d, &bits, 8);
    return (float)d;
}

std::vector<std::string> split_ws(const std::string &s)
{
    std::vector<std::string> out;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
        size_t j = i;
        while (j < s.size() && !std::isspace((unsigned char)s[j])) j++;
        if (j > i) out.push_back(s.substr(i, j - i));
        i = j;
    }
    return out;
}

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

// parseHeader (ply.js:111-152). Returns VS_OK and the element and the offset
// of the binary body.
int parse_header(const uint8_t *bytes, uint64_t size, Element &element, uint64_t &header_length)
{
    const std::string text(reinterpret_cast<const char *>(bytes), (size_t)std::min<uint64_t>(size, 64 * 1024));
    const size_t end = text.find("end_header");
    if (text.compare(0, 3, "ply") != 0 || end == std::string::npos)
        return ingest_fail(VS_E_INVALID, "not a PLY file: no \"ply\" magic or no end_header in the first 64 KB");
    const size_t newline = text.find('\n', end);
    header_length = newline == std::string::npos ? 0 : newline + 1;

    std::vector<std::string> lines;
    size_t at = 0;
    const std::string head = text.substr(0, end);
    while (at <= head.size()) {
        size_t nl = head.find('\n', at);
        if (nl == std::string::npos) nl = head.size();
        std::string line = head.substr(at, nl - at);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        line = trim(line);
        if (!line.empty()) lines.push_back(line);
        at = nl + 1;
    }
    const std::string *format = nullptr;
    for (const auto &l : lines)
        if (l.compare(0, 7, "format ") == 0) {
            format = &l;
            break;
        }
    if (!format || *format != "format binary_little_endian 1.0")
        return ingest_fail(VS_E_INVALID, "unsupported PLY format \"" + (format ? *format : std::string("undefined")) +
                                             "\"; only binary_little_endian 1.0 is read");
    bool have = false, usable = false;
    for (const auto &l : lines) {
        const std::vector<std::string> w = split_ws(l);
        if (w.empty()) continue;
        if (w[0] == "element") {
            if (have) break;  // only the first element is read
            have = true;
            element.name = w.size() > 1 ? w[1] : "undefined";
            if (w.size() > 2) {
                char *stop = nullptr;
                const double n = std::strtod(w[2].c_str(), &stop);
                usable = *stop == 0 && n >= 0 && n == std::floor(n) && n < 18446744073709551616.0;
                element.count = usable ? (uint64_t)n : 0;
            }
        } else if (w[0] == "property" && have) {
            const std::string type = w.size() > 1 ? w[1] : "undefined";
            if (type == "list")
                return ingest_fail(VS_E_INVALID, "PLY element \"" + element.name +
                                                     "\" has a list property, which this reader cannot stride past");
            const uint32_t bytes_of = type_size(type);
            if (!bytes_of) return ingest_fail(VS_E_INVALID, "unknown PLY property type \"" + type + "\"");
            element.properties.push_back({ type, w.size() > 2 ? w[2] : "undefined", bytes_of });
        }
    }
    if (!have || !usable) return ingest_fail(VS_E_INVALID, "PLY header declares no usable element");
    return VS_OK;
}

int read_columns(const uint8_t *bytes, uint64_t size, const char *const *names, uint32_t n_names, uint64_t *count,
                 float **columns)
{
    Element element;
    uint64_t header = 0;
    int rc = parse_header(bytes, size, element, header);
    if (rc != VS_OK) return rc;
    std::string missing;
    std::vector<int> which(n_names, -1);
    for (uint32_t i = 0; i < n_names; i++) {
        // A name declared twice keeps its last column, as the reference's
        // name-keyed record does.
        for (size_t p = 0; p < element.properties.size(); p++)
            if (element.properties[p].name == names[i]) which[i] = (int)p;
        if (which[i] < 0) missing += (missing.empty() ? "" : ", ") + std::string(names[i]);
    }
    if (!missing.empty())
        return ingest_fail(VS_E_INVALID, "PLY element \"" + element.name + "\" has no property " + missing);
    uint64_t stride = 0;
    std::vector<uint64_t> offset(element.properties.size());
    for (size_t p = 0; p < element.properties.size(); p++) {
        offset[p] = stride;
        stride += element.properties[p].size;
    }
    // The reference sums in doubles; in integers the product can wrap, so
    // compare without forming it.
    const bool fits = size >= header && (element.count == 0 || stride <= (size - header) / element.count);
    if (!fits) {
        char needed[32];
        std::snprintf(needed, sizeof needed, "%.17g", (double)header + (double)stride * (double)element.count);
        return ingest_fail(VS_E_INVALID, "PLY is truncated: " + std::to_string(element.count) + " records of " +
                                             std::to_string(stride) + " bytes need " + needed + " bytes, file has " +
                                             std::to_string(size));
    }
    for (uint32_t i = 0; i < n_names; i++) columns[i] = nullptr;
    for (uint32_t i = 0; i < n_names; i++) {
        columns[i] = static_cast<float *>(std::malloc(std::max<uint64_t>(element.count, 1) * sizeof(float)));
        if (!columns[i]) {
            for (uint32_t k = 0; k < i; k++) std::free(columns[k]), columns[k] = nullptr;
            return ingest_fail(VS_E_NOMEM, "PLY column does not fit host memory");
        }
        const Property &p = element.properties[which[i]];
        const uint8_t *base = bytes + header + offset[which[i]];
        for (uint64_t r = 0; r < element.count; r++) columns[i][r] = read_value(p.type, base + r * stride);
    }
    *count = element.count;
    return VS_OK;
}

// REQUIRED_PROPERTIES (cloud.js:28-34).
const char *const REQUIRED[14] = { "x",       "y",       "z",       "f_dc_0", "f_dc_1", "f_dc_2", "opacity",
                                   "scale_0", "scale_1", "scale_2", "rot_0",  "rot_1",  "rot_2",  "rot_3" };

// Degree-0 spherical harmonic (cloud.js:25).
constexpr double SH_C0 = 0.28209479177387814;

double sigmoid(double v)
{
    if (v >= 0) return 1 / (1 + std::exp(-v));
    const double t = std::exp(v);
    return t / (1 + t);
}

}  // namespace
}  // namespace vsim

extern "C" {

using namespace vsim;

const char *vs_ingest_error(void) { return g_ingest_error.c_str(); }

int vs_ply_read(const void *bytes, uint64_t size, const char *const *names, uint32_t n_names, uint64_t *count,
                float **columns)
{
    if ((!bytes && size) || !count || (n_names && (!names || !columns))) return ingest_fail(VS_E_INVALID, "missing argument");
    for (uint32_t i = 0; i < n_names; i++)
        if (!names[i]) return ingest_fail(VS_E_INVALID, "missing property name");
    return read_columns(static_cast<const uint8_t *>(bytes), size, names, n_names, count, columns);
}

int vs_ply_properties(const void *bytes, uint64_t size, char *list, uint64_t list_bytes, uint64_t *count)
{
    if ((!bytes && size) || !count) return ingest_fail(VS_E_INVALID, "missing argument");
    Element element;
    uint64_t header = 0;
    const int rc = parse_header(static_cast<const uint8_t *>(bytes), size, element, header);
    if (rc != VS_OK) return rc;
    std::string names;
    for (const auto &p : element.properties) names += (names.empty() ? "" : " ") + p.name;
    if (list) {
        if (list_bytes < names.size() + 1) return ingest_fail(VS_E_INVALID, "property list does not fit");
        std::memcpy(list, names.c_str(), names.size() + 1);
    }
    *count = element.count;
    return VS_OK;
}

void vs_cloud_free(vs_cloud *c)
{
    if (!c) return;
    std::free(c->positions);
    std::free(c->rotations);
    std::free(c->scales);
    std::free(c->opacity);
    std::free(c->colors);
    *c = vs_cloud{};
}

int vs_cloud_from_columns(uint64_t n, const float *const *columns, vs_cloud *out)
{
    if (!out || (n && !columns)) return ingest_fail(VS_E_INVALID, "missing argument");
    *out = vs_cloud{};
    for (int c = 0; n && c < 14; c++)
        if (!columns[c]) return ingest_fail(VS_E_INVALID, std::string("missing column ") + REQUIRED[c]);
    const size_t m = std::max<uint64_t>(n, 1);
    out->count = n;
    out->positions = static_cast<float *>(std::malloc(12 * m));
    out->rotations = static_cast<float *>(std::malloc(16 * m));
    out->scales = static_cast<float *>(std::malloc(12 * m));
    out->opacity = static_cast<float *>(std::malloc(4 * m));
    out->colors = static_cast<float *>(std::malloc(12 * m));
    if (!out->positions || !out->rotations || !out->scales || !out->opacity || !out->colors) {
        vs_cloud_free(out);
        return ingest_fail(VS_E_NOMEM, "cloud does not fit host memory");
    }
    if (n == 0) return VS_OK;
    const float *x = columns[0], *y = columns[1], *z = columns[2], *dc0 = columns[3], *dc1 = columns[4],
                *dc2 = columns[5], *op = columns[6], *s0 = columns[7], *s1 = columns[8], *s2 = columns[9],
                *r0 = columns[10], *r1 = columns[11], *r2 = columns[12], *r3 = columns[13];
    for (uint64_t i = 0; i < n; i++) {
        out->positions[3 * i] = x[i];
        out->positions[3 * i + 1] = y[i];
        out->positions[3 * i + 2] = z[i];
        out->scales[3 * i] = (float)std::exp((double)s0[i]);
        out->scales[3 * i + 1] = (float)std::exp((double)s1[i]);
        out->scales[3 * i + 2] = (float)std::exp((double)s2[i]);
        out->opacity[i] = (float)sigmoid(op[i]);
        out->colors[3 * i] = (float)((double)dc0[i] * SH_C0 + 0.5);
        out->colors[3 * i + 1] = (float)((double)dc1[i] * SH_C0 + 0.5);
        out->colors[3 * i + 2] = (float)((double)dc2[i] * SH_C0 + 0.5);
        // writeUnitQuat (cloud.js:99-108): rot_0 is the real part; the result
        // is xyzw, normalised, real part >= 0, the identity when zero.
        const double qx = r1[i], qy = r2[i], qz = r3[i], qw = r0[i];
        const double length = std::sqrt(qx * qx + qy * qy + qz * qz + qw * qw);
        const double sign = qw < 0 ? -1 : 1;
        const double scale = length > 0 ? sign / length : 0;
        out->rotations[4 * i] = (float)(qx * scale);
        out->rotations[4 * i + 1] = (float)(qy * scale);
        out->rotations[4 * i + 2] = (float)(qz * scale);
        out->rotations[4 * i + 3] = (float)(length > 0 ? qw * scale : 1);
    }
    return VS_OK;
}

int vs_cloud_from_ply(const void *bytes, uint64_t size, vs_cloud *out)
{
    if (!out || (!bytes && size)) return ingest_fail(VS_E_INVALID, "missing argument");
    *out = vs_cloud{};
    float *col[14];
    uint64_t n = 0;
    int rc = read_columns(static_cast<const uint8_t *>(bytes), size, REQUIRED, 14, &n, col);
    if (rc != VS_OK) return rc;
    rc = vs_cloud_from_columns(n, col, out);
    for (float *c : col) std::free(c);
    return rc;
}

}  // extern "C"
