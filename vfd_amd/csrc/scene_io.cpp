// scene_io.cpp - scene-description I/O: the DFSPH description as VFD's scene files store it.
//
// Reference: serialize(Archive&, DFSPHSimulationDescription&) (VFD/Source/Scene/Components/
// DFSPHSimulationComponent.h:7-39) writes the 26 fields under cereal names ("timeStepSize", ..., "gravity" {x,y,z})
// inside a "description" object (DFSPHSimulationComponent::save/load, :57-70); scene files hold one per simulation
// component (e.g. Resources/Scenes/New/DFSPH/default.json). vfd_read_descriptions finds every "description" object
// in a JSON text and fills one vfd_desc per object, starting from the reference defaults for fields it does not
// name; vfd_write_description writes one in the same form. A small JSON reader of its own (objects, arrays, strings,
// numbers, literals) walks the text; everything outside the description objects is skipped.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/vfd.h"

namespace vfd {
int set_error(int code, const std::string& msg);  // sim.cpp: vfd_last_error's message
}

namespace {

struct Reader {
    const char* p;
    const char* end;
    std::string error;

    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
    }
    bool fail(const char* what) {
        if (error.empty()) error = std::string(what) + " at offset " + std::to_string((long)(end - p)) + " from the end";
        return false;
    }
    bool lit(const char* s) {
        const size_t n = std::strlen(s);
        if ((size_t)(end - p) < n || std::strncmp(p, s, n) != 0) return false;
        p += n;
        return true;
    }
    bool string(std::string* out) {
        ws();
        if (p >= end || *p != '"') return fail("expected a string");
        ++p;
        std::string s;
        while (p < end && *p != '"') {
            if (*p == '\\') {
                if (++p >= end) break;
                const char c = *p;
                s.push_back(c == 'n' ? '\n' : c == 't' ? '\t' : c == 'r' ? '\r' : c == 'b' ? '\b' : c == 'f' ? '\f' : c);
                if (c == 'u') p = std::min(end, p + 4);  // \uXXXX: kept as 'u' (names here are ASCII)
                ++p;
            } else {
                s.push_back(*p++);
            }
        }
        if (p >= end) return fail("unterminated string");
        ++p;
        if (out) *out = s;
        return true;
    }
    bool number(double* out) {
        ws();
        char buf[64];
        size_t n = 0;
        while (p < end && n + 1 < sizeof(buf) && (std::strchr("+-0123456789.eE", *p) != nullptr)) buf[n++] = *p++;
        buf[n] = 0;
        char* e = nullptr;
        const double v = std::strtod(buf, &e);
        if (n == 0 || e != buf + n) return fail("expected a number");
        *out = v;
        return true;
    }
    // any value; with `d` non-null a number or literal is stored there (true = 1, false = 0)
    bool value(double* d, int depth = 0) {
        ws();
        if (p >= end) return fail("unexpected end of text");
        if (depth > 256) return fail("nesting too deep");
        const char c = *p;
        if (c == '{' || c == '[') {
            const char close = c == '{' ? '}' : ']';
            ++p;
            ws();
            if (p < end && *p == close) {
                ++p;
                return true;
            }
            for (;;) {
                if (c == '{') {
                    if (!string(nullptr)) return false;
                    ws();
                    if (p >= end || *p != ':') return fail("expected ':'");
                    ++p;
                }
                if (!value(nullptr, depth + 1)) return false;
                ws();
                if (p < end && *p == ',') {
                    ++p;
                    continue;
                }
                if (p < end && *p == close) {
                    ++p;
                    return true;
                }
                return fail("expected ',' or a closing bracket");
            }
        }
        if (c == '"') return string(nullptr);
        if (lit("true")) {
            if (d) *d = 1.0;
            return true;
        }
        if (lit("false")) {
            if (d) *d = 0.0;
            return true;
        }
        if (lit("null")) return true;
        double v = 0.0;
        if (!number(&v)) return false;
        if (d) *d = v;
        return true;
    }
};

// the reference's cereal names, in serialization order
enum Kind { F32, U32, I32, BOOL };
struct Field {
    const char* name;
    Kind kind;
    size_t offset;
};
#define VFD_FIELD(n, k, m) {n, k, offsetof(vfd_desc, m)}
const Field kFields[] = {
    VFD_FIELD("timeStepSize", F32, TimeStepSize),
    VFD_FIELD("minTimeStepSize", F32, MinTimeStepSize),
    VFD_FIELD("maxTimeStepSize", F32, MaxTimeStepSize),
    VFD_FIELD("frameLength", F32, FrameLength),
    VFD_FIELD("frameCount", U32, FrameCount),
    VFD_FIELD("minPressureSolverIterations", U32, MinPressureSolverIterations),
    VFD_FIELD("maxPressureSolverIterations", U32, MaxPressureSolverIterations),
    VFD_FIELD("maxPressureSolverError", F32, MaxPressureSolverError),
    VFD_FIELD("enableDivergenceSolverError", BOOL, EnableDivergenceSolverError),
    VFD_FIELD("minDivergenceSolverIterations", U32, MinDivergenceSolverIterations),
    VFD_FIELD("maxDivergenceSolverIterations", U32, MaxDivergenceSolverIterations),
    VFD_FIELD("maxDivergenceSolverError", F32, MaxDivergenceSolverError),
    VFD_FIELD("enableViscositySolver", BOOL, EnableViscositySolver),
    VFD_FIELD("minViscositySolverIterations", U32, MinViscositySolverIterations),
    VFD_FIELD("maxViscositySolverIterations", U32, MaxViscositySolverIterations),
    VFD_FIELD("maxViscositySolverError", F32, MaxViscositySolverError),
    VFD_FIELD("viscosity", F32, Viscosity),
    VFD_FIELD("boundaryViscosity", F32, BoundaryViscosity),
    VFD_FIELD("tangentialDistanceFactor", F32, TangentialDistanceFactor),
    VFD_FIELD("enableSurfaceTensionSolver", BOOL, EnableSurfaceTensionSolver),
    VFD_FIELD("surfaceTensionSmoothPassCount", U32, SurfaceTensionSmoothPassCount),
    VFD_FIELD("surfaceTension", F32, SurfaceTension),
    VFD_FIELD("temporalSmoothing", BOOL, TemporalSmoothing),
    VFD_FIELD("CSDFix", I32, CSDFix),
    VFD_FIELD("CSD", I32, CSD),
    VFD_FIELD("particleRadius", F32, ParticleRadius),
};
#undef VFD_FIELD

bool store(vfd_desc& d, const Field& f, double v, std::string& why) {
    char* base = reinterpret_cast<char*>(&d) + f.offset;
    switch (f.kind) {
        case F32: *reinterpret_cast<float*>(base) = (float)v; return true;
        case BOOL: *reinterpret_cast<int32_t*>(base) = v != 0.0 ? 1 : 0; return true;
        case U32:
            if (!(v >= 0.0 && v <= 4294967295.0) || v != std::floor(v)) break;
            *reinterpret_cast<uint32_t*>(base) = (uint32_t)v;
            return true;
        case I32:
            if (!(v >= -2147483648.0 && v <= 2147483647.0) || v != std::floor(v)) break;
            *reinterpret_cast<int32_t*>(base) = (int32_t)v;
            return true;
    }
    why = std::string("field ") + f.name + ": not a valid integer";
    return false;
}

// members of one "description" object into d (starting from the defaults)
bool read_description(Reader& r, vfd_desc& d) {
    vfd_desc_default(&d);
    r.ws();
    if (r.p >= r.end || *r.p != '{') return r.fail("a description must be an object");
    ++r.p;
    r.ws();
    if (r.p < r.end && *r.p == '}') {
        ++r.p;
        return true;
    }
    for (;;) {
        std::string name;
        if (!r.string(&name)) return false;
        r.ws();
        if (r.p >= r.end || *r.p != ':') return r.fail("expected ':'");
        ++r.p;
        const Field* f = nullptr;
        for (const Field& k : kFields)
            if (name == k.name) f = &k;
        if (name == "gravity") {
            r.ws();
            if (r.p >= r.end || *r.p != '{') return r.fail("gravity must be an object {x, y, z}");
            ++r.p;
            for (;;) {
                std::string axis;
                if (!r.string(&axis)) return false;
                r.ws();
                if (r.p >= r.end || *r.p != ':') return r.fail("expected ':'");
                ++r.p;
                double v = 0.0;
                if (!r.number(&v)) return false;
                if (axis == "x") d.Gravity[0] = (float)v;
                else if (axis == "y") d.Gravity[1] = (float)v;
                else if (axis == "z") d.Gravity[2] = (float)v;
                r.ws();
                if (r.p < r.end && *r.p == ',') {
                    ++r.p;
                    continue;
                }
                if (r.p < r.end && *r.p == '}') {
                    ++r.p;
                    break;
                }
                return r.fail("expected ',' or '}' in gravity");
            }
        } else if (f) {
            double v = 0.0;
            if (!r.value(&v)) return false;
            std::string why;
            if (!store(d, *f, v, why)) {
                r.error = why;
                return false;
            }
        } else if (!r.value(nullptr)) {  // a field this parser does not know: skipped
            return false;
        }
        r.ws();
        if (r.p < r.end && *r.p == ',') {
            ++r.p;
            continue;
        }
        if (r.p < r.end && *r.p == '}') {
            ++r.p;
            return true;
        }
        return r.fail("expected ',' or '}' in a description");
    }
}

// walks every value; each member named "description" whose value is an object is read as one
bool scan(Reader& r, vfd_desc* out, uint32_t cap, uint32_t& n, int depth) {
    r.ws();
    if (r.p >= r.end) return r.fail("unexpected end of text");
    if (depth > 256) return r.fail("nesting too deep");
    const char c = *r.p;
    if (c != '{' && c != '[') return r.value(nullptr);
    const char close = c == '{' ? '}' : ']';
    ++r.p;
    r.ws();
    if (r.p < r.end && *r.p == close) {
        ++r.p;
        return true;
    }
    for (;;) {
        bool isDesc = false;
        if (c == '{') {
            std::string name;
            if (!r.string(&name)) return false;
            r.ws();
            if (r.p >= r.end || *r.p != ':') return r.fail("expected ':'");
            ++r.p;
            r.ws();
            isDesc = name == "description" && r.p < r.end && *r.p == '{';
        }
        if (isDesc) {
            vfd_desc d;
            if (!read_description(r, d)) return false;
            if (n < cap && out) out[n] = d;
            ++n;
        } else if (!scan(r, out, cap, n, depth + 1)) {
            return false;
        }
        r.ws();
        if (r.p < r.end && *r.p == ',') {
            ++r.p;
            continue;
        }
        if (r.p < r.end && *r.p == close) {
            ++r.p;
            return true;
        }
        return r.fail("expected ',' or a closing bracket");
    }
}

}  // namespace

extern "C" {

int vfd_read_descriptions(const char* json, size_t length, vfd_desc* out, uint32_t capacity, uint32_t* count) {
    if (!json || !count) return vfd::set_error(VFD_ERROR_INVALID_ARGUMENT, "json and count must not be NULL");
    Reader r{json, json + length, {}};
    uint32_t n = 0;
    if (!scan(r, out, capacity, n, 0))
        return vfd::set_error(VFD_ERROR_INVALID_ARGUMENT, "scene JSON: " + r.error);
    r.ws();
    if (r.p != r.end) return vfd::set_error(VFD_ERROR_INVALID_ARGUMENT, "scene JSON: text after the top-level value");
    *count = n;
    return VFD_OK;
}

int vfd_write_description(const vfd_desc* desc, char* buffer, size_t capacity, size_t* length) {
    if (!desc || !length) return vfd::set_error(VFD_ERROR_INVALID_ARGUMENT, "desc and length must not be NULL");
    std::string s = "{\n    \"description\": {\n";
    char num[64];
    for (const Field& f : kFields) {
        const char* base = reinterpret_cast<const char*>(desc) + f.offset;
        switch (f.kind) {
            case F32: std::snprintf(num, sizeof(num), "%.17g", (double)*reinterpret_cast<const float*>(base)); break;
            case U32: std::snprintf(num, sizeof(num), "%u", *reinterpret_cast<const uint32_t*>(base)); break;
            case I32: std::snprintf(num, sizeof(num), "%d", *reinterpret_cast<const int32_t*>(base)); break;
            case BOOL: std::snprintf(num, sizeof(num), "%s", *reinterpret_cast<const int32_t*>(base) ? "true" : "false");
        }
        s += std::string("        \"") + f.name + "\": " + num + ",\n";
    }
    std::snprintf(num, sizeof(num), "%.17g", (double)desc->Gravity[0]);
    s += std::string("        \"gravity\": {\n            \"x\": ") + num;
    std::snprintf(num, sizeof(num), "%.17g", (double)desc->Gravity[1]);
    s += std::string(",\n            \"y\": ") + num;
    std::snprintf(num, sizeof(num), "%.17g", (double)desc->Gravity[2]);
    s += std::string(",\n            \"z\": ") + num + "\n        }\n    }\n}\n";
    *length = s.size();
    if (buffer) {
        if (capacity < s.size() + 1) return vfd::set_error(VFD_ERROR_INVALID_ARGUMENT, "buffer too small");
        std::memcpy(buffer, s.c_str(), s.size() + 1);
    }
    return VFD_OK;
}

}  // extern "C"
