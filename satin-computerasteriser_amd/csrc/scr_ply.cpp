// scr_ply.cpp — PLY input on the host: the header and vertex-body parsing of
// PLYPointCloudLoader.parse (PLYPointCloudLoader.swift:38-333), producing the
// float4 positions and colours that scr_pack takes.
//
// Behaviour followed from the reference:
//   * the header ends at the first "end_header" anywhere in the image, plus
//     one optional CR and one optional LF (:173-198); it must be UTF-8;
//   * header lines split on every newline character, fields on single
//     spaces with empty fields dropped (:205-207); "format" (ascii or
//     binary_little_endian), "element vertex N" and the vertex element's
//     scalar "property" lines are read, everything else is skipped
//     (:209-238); a list property or unknown type is unsupported;
//   * x, y and z must be present (:150-154); red/r/diffuse_red,
//     green/g/diffuse_green and blue/b/diffuse_blue are colour channels, any
//     other property is skipped (:161-171); a later field of the same role
//     overwrites an earlier one;
//   * ASCII: non-empty body lines, one vertex each, whitespace split as in
//     the header, every field a float (:245-282); binary: packed
//     little-endian records of the properties in order (:284-314, :335-369);
//   * colours start at (1, 1, 1, 1); integer channels are divided by the
//     type's maximum, all channels clamped to [0, 1] (:316-333).
// Deviation: a negative vertex count is a header error (the reference traps
// on the empty range it makes).
#include <locale.h>
#include <stdlib.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/scr.h"

namespace {

enum class Scalar { kI8, kU8, kI16, kU16, kI32, kU32, kF32, kF64 };
enum class Role { kX, kY, kZ, kRed, kGreen, kBlue, kSkip };

struct Field {
    Role role;
    std::string name;
    Scalar type;
    size_t offset;
};

struct Header {
    bool ascii = false;
    long long vertices = 0;
    std::vector<Field> fields;
    size_t stride = 0;
    size_t body = 0;
};

struct Error {
    int code;
    std::string msg;
};

bool scalar_of(const std::string& s, Scalar* t) {
    static const struct {
        const char* name;
        Scalar type;
    } kNames[] = {{"char", Scalar::kI8},    {"int8", Scalar::kI8},     {"uchar", Scalar::kU8},
                  {"uint8", Scalar::kU8},   {"short", Scalar::kI16},   {"int16", Scalar::kI16},
                  {"ushort", Scalar::kU16}, {"uint16", Scalar::kU16},  {"int", Scalar::kI32},
                  {"int32", Scalar::kI32},  {"uint", Scalar::kU32},    {"uint32", Scalar::kU32},
                  {"float", Scalar::kF32},  {"float32", Scalar::kF32}, {"double", Scalar::kF64},
                  {"float64", Scalar::kF64}};
    for (const auto& n : kNames)
        if (s == n.name) {
            *t = n.type;
            return true;
        }
    return false;
}

size_t scalar_bytes(Scalar t) {
    switch (t) {
        case Scalar::kI8:
        case Scalar::kU8: return 1;
        case Scalar::kI16:
        case Scalar::kU16: return 2;
        case Scalar::kF64: return 8;
        default: return 4;
    }
}

// Largest value of an integer channel type; 1 for floating-point channels.
float channel_max(Scalar t) {
    switch (t) {
        case Scalar::kI8: return 127.0f;
        case Scalar::kU8: return 255.0f;
        case Scalar::kI16: return 32767.0f;
        case Scalar::kU16: return 65535.0f;
        case Scalar::kI32: return 2147483647.0f;   // rounds to 2^31 in binary32, as in the reference
        case Scalar::kU32: return 4294967295.0f;   // rounds to 2^32
        default: return 1.0f;
    }
}

Role role_of(const std::string& n) {
    if (n == "x") return Role::kX;
    if (n == "y") return Role::kY;
    if (n == "z") return Role::kZ;
    if (n == "red" || n == "r" || n == "diffuse_red") return Role::kRed;
    if (n == "green" || n == "g" || n == "diffuse_green") return Role::kGreen;
    if (n == "blue" || n == "b" || n == "diffuse_blue") return Role::kBlue;
    return Role::kSkip;
}

// Length of the newline sequence at p (0 if none): LF, VT, FF, CR, CR LF,
// NEL (U+0085), LS (U+2028), PS (U+2029) — the characters Swift's
// Character.isNewline accepts.
size_t newline_at(const uint8_t* p, const uint8_t* end) {
    const uint8_t c = *p;
    if (c == '\r') return (p + 1 < end && p[1] == '\n') ? 2 : 1;
    if (c == '\n' || c == 0x0b || c == 0x0c) return 1;
    if (c == 0xc2 && p + 1 < end && p[1] == 0x85) return 2;
    if (c == 0xe2 && p + 2 < end && p[1] == 0x80 && (p[2] == 0xa8 || p[2] == 0xa9)) return 3;
    return 0;
}

// Non-empty runs between newline sequences.
std::vector<std::string> lines_of(const uint8_t* b, size_t n) {
    std::vector<std::string> out;
    const uint8_t* end = b + n;
    const uint8_t* start = b;
    const uint8_t* p = b;
    while (p < end) {
        const size_t k = newline_at(p, end);
        if (k == 0) {
            p++;
            continue;
        }
        if (p > start) out.emplace_back((const char*)start, (size_t)(p - start));
        p += k;
        start = p;
    }
    if (end > start) out.emplace_back((const char*)start, (size_t)(end - start));
    return out;
}

// Non-empty fields between single spaces.
std::vector<std::string> fields_of(const std::string& line) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && line[i] == ' ') i++;
        size_t j = i;
        while (j < line.size() && line[j] != ' ') j++;
        if (j > i) out.emplace_back(line, i, j - i);
        i = j;
    }
    return out;
}

bool valid_utf8(const uint8_t* s, size_t n) {
    size_t i = 0;
    while (i < n) {
        const uint8_t c = s[i];
        size_t len;
        uint32_t cp;
        if (c < 0x80) {
            i++;
            continue;
        } else if ((c & 0xe0) == 0xc0) {
            len = 2;
            cp = c & 0x1f;
        } else if ((c & 0xf0) == 0xe0) {
            len = 3;
            cp = c & 0x0f;
        } else if ((c & 0xf8) == 0xf0) {
            len = 4;
            cp = c & 0x07;
        } else {
            return false;
        }
        if (i + len > n) return false;
        for (size_t k = 1; k < len; k++) {
            if ((s[i + k] & 0xc0) != 0x80) return false;
            cp = (cp << 6) | (s[i + k] & 0x3f);
        }
        // overlong forms, surrogates and values past U+10FFFF are not UTF-8
        if ((len == 2 && cp < 0x80) || (len == 3 && cp < 0x800) || (len == 4 && cp < 0x10000) ||
            (cp >= 0xd800 && cp <= 0xdfff) || cp > 0x10ffff)
            return false;
        i += len;
    }
    return true;
}

// Swift Int(String): optional sign, then decimal digits only, in range.
bool parse_int(const std::string& s, long long* v) {
    size_t i = 0;
    bool neg = false;
    if (i < s.size() && (s[i] == '+' || s[i] == '-')) neg = s[i++] == '-';
    if (i == s.size()) return false;
    unsigned long long acc = 0;
    const unsigned long long lim = neg ? 9223372036854775808ull : 9223372036854775807ull;
    for (; i < s.size(); i++) {
        if (s[i] < '0' || s[i] > '9') return false;
        const unsigned d = (unsigned)(s[i] - '0');
        if (acc > (lim - d) / 10) return false;
        acc = acc * 10 + d;
    }
    *v = neg ? (long long)(0ull - acc) : (long long)acc;
    return true;
}

// Swift Float(String): the C-locale strtof reading of the whole field,
// correctly rounded; leading whitespace is not accepted.
bool parse_float(const char* s, size_t n, float* v) {
    static locale_t c_locale = newlocale(LC_ALL_MASK, "C", (locale_t)0);
    if (n == 0 || s[0] == ' ' || (s[0] >= '\t' && s[0] <= '\r')) return false;
    char* end = nullptr;
    const float r = strtof_l(s, &end, c_locale);
    if (end != s + n) return false;
    *v = r;
    return true;
}

uint32_t read_u32(const uint8_t* p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

float read_scalar(Scalar t, const uint8_t* p) {
    switch (t) {
        case Scalar::kI8: return (float)(int8_t)p[0];
        case Scalar::kU8: return (float)p[0];
        case Scalar::kI16: return (float)(int16_t)(uint16_t)(p[0] | (p[1] << 8));
        case Scalar::kU16: return (float)(uint16_t)(p[0] | (p[1] << 8));
        case Scalar::kI32: return (float)(int32_t)read_u32(p);
        case Scalar::kU32: return (float)read_u32(p);
        case Scalar::kF32: {
            const uint32_t b = read_u32(p);
            float f;
            memcpy(&f, &b, 4);
            return f;
        }
        case Scalar::kF64: {
            const uint64_t b = (uint64_t)read_u32(p) | ((uint64_t)read_u32(p + 4) << 32);
            double d;
            memcpy(&d, &b, 8);
            return (float)d;
        }
    }
    return 0.0f;
}

// min(max(x, 0), 1) with NaN -> 0 (simd_clamp)
float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

float channel(float v, Scalar t) {
    if (t == Scalar::kF32 || t == Scalar::kF64) return clamp01(v);
    return clamp01(v / channel_max(t));
}

void assign(const Field& f, float v, float* pos, float* col) {
    switch (f.role) {
        case Role::kX: pos[0] = v; break;
        case Role::kY: pos[1] = v; break;
        case Role::kZ: pos[2] = v; break;
        case Role::kRed: col[0] = channel(v, f.type); break;
        case Role::kGreen: col[1] = channel(v, f.type); break;
        case Role::kBlue: col[2] = channel(v, f.type); break;
        case Role::kSkip: break;
    }
}

bool parse_header(const uint8_t* d, size_t n, Header* h, Error* err) {
    *err = {SCR_ERR_PLY_INVALID_HEADER, "The file is not a valid PLY file."};
    if (n < 4) return false;
    static const char kMarker[] = "end_header";
    const size_t m = sizeof(kMarker) - 1;
    size_t body = 0;
    bool found = false;
    for (size_t i = 0; !found && i + m <= n; i++) {
        if (memcmp(d + i, kMarker, m) != 0) continue;
        body = i + m;
        if (body < n && d[body] == '\r') body++;
        if (body < n && d[body] == '\n') body++;
        found = true;
    }
    if (!found || !valid_utf8(d, body)) return false;

    bool haveFormat = false, haveVertices = false, inVertex = false;
    std::vector<std::pair<std::string, Scalar>> props;
    for (const std::string& line : lines_of(d, body)) {
        const std::vector<std::string> w = fields_of(line);
        if (w.empty()) continue;
        if (w[0] == "format") {
            if (w.size() < 3) return false;
            if (w[1] == "ascii") {
                h->ascii = true;
            } else if (w[1] == "binary_little_endian") {
                h->ascii = false;
            } else {
                *err = {SCR_ERR_PLY_UNSUPPORTED_FORMAT, "Unsupported PLY format: " + w[1] + "."};
                return false;
            }
            haveFormat = true;
        } else if (w[0] == "element") {
            if (w.size() < 3) return false;
            inVertex = w[1] == "vertex";
            if (inVertex) {
                long long count;
                if (!parse_int(w[2], &count)) return false;
                if (count < 0) {
                    err->msg = "The file is not a valid PLY file (negative vertex count).";
                    return false;
                }
                h->vertices = count;
                haveVertices = true;
            }
        } else if (w[0] == "property" && inVertex) {
            if (w.size() < 3) return false;
            if (w[1] == "list") {
                std::string rest;
                for (size_t k = 2; k < w.size(); k++) rest += (k > 2 ? " " : "") + w[k];
                *err = {SCR_ERR_PLY_UNSUPPORTED_PROPERTY, "Unsupported PLY vertex property 'list " + rest + "'."};
                return false;
            }
            Scalar t;
            if (!scalar_of(w[1], &t)) {
                *err = {SCR_ERR_PLY_UNSUPPORTED_PROPERTY, "Unsupported PLY vertex property '" + w[1] + "'."};
                return false;
            }
            props.emplace_back(w[2], t);
        }
    }
    if (!haveFormat) return false;
    if (!haveVertices) {
        *err = {SCR_ERR_PLY_MISSING_VERTEX_ELEMENT, "The PLY file does not contain a vertex element."};
        return false;
    }
    bool hx = false, hy = false, hz = false;
    size_t offset = 0;
    for (const auto& pr : props) {
        const Field f{role_of(pr.first), pr.first, pr.second, offset};
        hx |= f.role == Role::kX;
        hy |= f.role == Role::kY;
        hz |= f.role == Role::kZ;
        offset += scalar_bytes(pr.second);
        h->fields.push_back(f);
    }
    if (!(hx && hy && hz)) {
        *err = {SCR_ERR_PLY_MISSING_PROPERTY, "The PLY file is missing required property 'x, y, z'."};
        return false;
    }
    h->stride = offset;
    h->body = body;
    return true;
}

const Error kTruncated{SCR_ERR_PLY_TRUNCATED_BODY, "The PLY file ended before all vertex data could be read."};

struct Span {
    size_t off, len;
};

// The first `limit` non-empty lines of b[0, n) (same separators as lines_of).
std::vector<Span> line_spans(const uint8_t* b, size_t n, size_t limit) {
    std::vector<Span> out;
    out.reserve(std::min(limit, n / 2 + 1));
    const uint8_t* end = b + n;
    size_t start = 0, i = 0;
    while (i < n && out.size() < limit) {
        const size_t k = newline_at(b + i, end);
        if (k == 0) {
            i++;
            continue;
        }
        if (i > start) out.push_back({start, i - start});
        i += k;
        start = i;
    }
    if (out.size() < limit && n > start) out.push_back({start, n - start});
    return out;
}

// One ASCII vertex line: 0 parsed; 1 fewer values than properties (checked
// before any value is read, as the reference counts them first); 2 the value
// of property *bad does not parse.
int parse_vertex_line(const char* s, size_t len, const Header& h, float* p, float* c, size_t* bad) {
    const size_t nf = h.fields.size();
    Span tok[64];
    std::vector<Span> more;
    Span* t = nf <= 64 ? tok : (more.resize(nf), more.data());
    size_t have = 0, i = 0;
    while (i < len && have < nf) {
        while (i < len && s[i] == ' ') i++;
        size_t j = i;
        while (j < len && s[j] != ' ') j++;
        if (j > i) t[have++] = {i, j - i};
        i = j;
    }
    if (have < nf) return 1;
    char buf[72];
    for (size_t k = 0; k < nf; k++) {
        float v;
        bool ok;
        if (t[k].len < sizeof(buf)) {
            memcpy(buf, s + t[k].off, t[k].len);
            buf[t[k].len] = '\0';
            ok = parse_float(buf, t[k].len, &v);
        } else {
            const std::string w(s + t[k].off, t[k].len);
            ok = parse_float(w.c_str(), w.size(), &v);
        }
        if (!ok) {
            *bad = k;
            return 2;
        }
        assign(h.fields[k], v, p, c);
    }
    return 0;
}

// Parses the body; writes the vertices when pos and col are given.
bool parse_body(const uint8_t* d, size_t n, const Header& h, float* pos, float* col, Error* err) {
    const size_t count = (size_t)h.vertices;
    if (h.ascii) {
        if (!valid_utf8(d + h.body, n - h.body)) {
            *err = kTruncated;
            return false;
        }
        const char* text = (const char*)(d + h.body);
        const std::vector<Span> lines = line_spans(d + h.body, n - h.body, count);
        if (lines.size() < count) {
            *err = kTruncated;
            return false;
        }
        // lines parse independently: ranges on up to 16 threads; the error
        // reported is the one of the first failing line, as a sequential parse
        // would report it
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const size_t nt = std::min<size_t>(std::min<size_t>(16, hw), std::max<size_t>(1, count / 32768));
        std::vector<size_t> badLine(nt, (size_t)-1), badField(nt, 0);
        std::vector<int> badKind(nt, 0);
        auto work = [&](size_t t) {
            const size_t a = count * t / nt, b = count * (t + 1) / nt;
            for (size_t i = a; i < b; i++) {
                float p[3] = {0.0f, 0.0f, 0.0f}, c[4] = {1.0f, 1.0f, 1.0f, 1.0f};
                size_t bf = 0;
                const int r = parse_vertex_line(text + lines[i].off, lines[i].len, h, p, c, &bf);
                if (r != 0) {
                    badLine[t] = i, badKind[t] = r, badField[t] = bf;
                    return;
                }
                if (pos) {
                    float* q = pos + 4 * i;
                    q[0] = p[0], q[1] = p[1], q[2] = p[2], q[3] = 0.0f;
                    memcpy(col + 4 * i, c, sizeof(c));
                }
            }
        };
        if (nt == 1) {
            work(0);
        } else {
            std::vector<std::thread> pool;
            for (size_t t = 0; t < nt; t++) pool.emplace_back(work, t);
            for (auto& th : pool) th.join();
        }
        for (size_t t = 0; t < nt; t++) {
            if (badLine[t] == (size_t)-1) continue;
            if (badKind[t] == 1) *err = kTruncated;
            else
                *err = {SCR_ERR_PLY_INVALID_ASCII_VALUE, "Could not parse property '" + h.fields[badField[t]].name +
                                                             "' on vertex line " + std::to_string(badLine[t] + 1) + "."};
            return false;
        }
        return true;
    }
    if (h.stride != 0 && count > (n - h.body) / h.stride) {
        *err = kTruncated;
        return false;
    }
    if (!pos) return true;
    // records convert independently: ranges on up to 16 threads
    auto work = [&](size_t a, size_t b) {
        for (size_t i = a; i < b; i++) {
            const uint8_t* rec = d + h.body + i * h.stride;
            float p[3] = {0.0f, 0.0f, 0.0f}, c[4] = {1.0f, 1.0f, 1.0f, 1.0f};
            for (const Field& f : h.fields) assign(f, read_scalar(f.type, rec + f.offset), p, c);
            float* q = pos + 4 * i;
            q[0] = p[0], q[1] = p[1], q[2] = p[2], q[3] = 0.0f;
            memcpy(col + 4 * i, c, sizeof(c));
        }
    };
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>(std::min<size_t>(16, hw), std::max<size_t>(1, count / 262144));
    if (nt == 1) {
        work(0, count);
    } else {
        std::vector<std::thread> pool;
        for (size_t t = 0; t < nt; t++) pool.emplace_back(work, count * t / nt, count * (t + 1) / nt);
        for (auto& th : pool) th.join();
    }
    return true;
}

int fail(const Error& e, char* msg, size_t msgBytes) {
    if (msg && msgBytes > 0) snprintf(msg, msgBytes, "%s", e.msg.c_str());
    return e.code;
}

}  // namespace

extern "C" int scr_ply_parse(const void* data, size_t bytes, size_t capacity, float* positions4, float* colors4,
                             size_t* vertexCount, char* msg, size_t msgBytes) {
    if (msg && msgBytes > 0) msg[0] = '\0';
    if (!vertexCount || (!data && bytes > 0) || ((positions4 == nullptr) != (colors4 == nullptr)))
        return fail({SCR_ERR_INVALID_ARGUMENT, "bad arguments"}, msg, msgBytes);
    *vertexCount = 0;
    const uint8_t* d = static_cast<const uint8_t*>(data);
    Header h;
    Error err{SCR_OK, ""};
    if (!parse_header(d, bytes, &h, &err)) return fail(err, msg, msgBytes);
    if (positions4 && (unsigned long long)h.vertices > (unsigned long long)capacity)
        return fail({SCR_ERR_INVALID_ARGUMENT, "capacity is smaller than the vertex count"}, msg, msgBytes);
    if (!parse_body(d, bytes, h, positions4, colors4, &err)) return fail(err, msg, msgBytes);
    *vertexCount = (size_t)h.vertices;
    return SCR_OK;
}
