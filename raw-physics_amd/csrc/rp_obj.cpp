// rp_obj.cpp — the OBJ asset path: what obj_parse (src/render/obj.cpp:7-151)
// returns for a Wavefront OBJ file, so that the reference's own assets
// (res/*.obj) load into a world with the values the reference sees.
//
// obj_parse reads the file with the tinyobjloader the reference bundles
// (include/tiny_obj_loader.h, the 2.0.0 release candidate: LoadObj with no
// material list, triangulation on, vertex-colour fallback on) and flattens
// its triangles into one vertex per triangle corner. The parts of that
// loader that decide the values are restated here:
//   * line splitting at "\n", "\r\n" or "\r" (safeGetline, tiny_obj_loader.h:699-731);
//   * its own decimal parser (tryParseDouble, :805-931), which is not
//     correctly rounded: digits accumulate in a double, fraction digits are
//     scaled by tenths (a table for the first seven, pow(10, -k) after), the
//     exponent is applied as ldexp(m * pow(5, e), e), and the double is then
//     rounded to float;
//   * face corners "v", "v/t", "v//n", "v/t/n" with 1-based or negative
//     (relative) indices (parseTriple and fixIndex, :739-760, :1071-1123);
//   * ear-clipping triangulation of polygons in the projection plane picked
//     from the first non-degenerate corner, in float arithmetic, with the
//     crossing-number point-in-triangle test (exportGroupsToShape and pnpoly,
//     :1315-1560). Faces are triangulated when a group ('g'), object ('o') or
//     the file ends, against the positions read so far, as the loader does.
// and obj_parse's own flattening: corners in face order, positions and
// texture coordinates copied from the file, normals copied from the file or,
// when it has none, each corner's normal is 0.0f plus its triangle's
// unnormalised float cross product (obj.cpp:112-148).
//
// Differences, all where the reference's behaviour is undefined or ends the
// process: a parse failure or an index outside the file's arrays returns an
// error (obj_parse calls exit(1), or reads out of bounds); texture
// coordinates of a file without any are 0 (obj_parse leaves them
// uninitialised). Materials ("mtllib", "usemtl") are not read: obj_parse
// passes no material list, so no material is ever found and "usemtl" never
// splits a shape.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/rawphys.h"

namespace rp {
void rp_set_last_error(uint32_t code, const std::string& msg);  // rp_world.cpp
}

namespace {

inline bool is_blank(char c) { return c == ' ' || c == '\t'; }
inline bool is_digit(char c) { return (unsigned)(c - '0') < 10u; }
inline bool is_line_end(char c) { return c == '\r' || c == '\n' || c == '\0'; }

// The loader's decimal parser (tryParseDouble). Grammar: [sign] digits
// [. digits] [(e|E) [sign] digits], or a fraction without integer digits
// (".5", "-.5"). Returns false (leaving *out) when nothing parses.
bool parse_decimal(const char* s, const char* end, double* out) {
    if (s >= end) return false;
    const char* c = s;
    bool negative = false, fraction_first = false;
    if (*c == '+' || *c == '-') {
        negative = *c == '-';
        ++c;
        fraction_first = c != end && *c == '.';
    } else if (*c == '.') {
        fraction_first = true;
    } else if (!is_digit(*c)) {
        return false;
    }
    double m = 0.0;
    if (!fraction_first) {
        int digits = 0;
        for (; c != end && is_digit(*c); ++c, ++digits) {
            m *= 10;
            m += (int)(*c - '0');
        }
        if (digits == 0) return false;
    }
    int e = 0;
    if (c != end) {
        bool more = true;
        if (*c == '.') {
            // tenths 10^-k for k < 8 are the double literals 0.1 ... 0.0000001
            static const double tenth[8] = {1.0, 0.1, 0.01, 0.001, 0.0001, 0.00001, 0.000001, 0.0000001};
            ++c;
            for (int k = 1; c != end && is_digit(*c); ++c, ++k)
                m += (int)(*c - '0') * (k < 8 ? tenth[k] : std::pow(10.0, -k));
            more = c != end;
        } else if (*c != 'e' && *c != 'E') {
            more = false;
        }
        if (more && (*c == 'e' || *c == 'E')) {
            ++c;
            bool eneg = false;
            if (c != end && (*c == '+' || *c == '-')) {
                eneg = *c == '-';
                ++c;
            } else if (!is_digit(*c)) {  // *end is a delimiter, never a digit
                return false;
            }
            int digits = 0;
            for (; c != end && is_digit(*c); ++c, ++digits) {
                if (e < 100000) {  // saturates where the loader's int would overflow
                    e *= 10;
                    e += (int)(*c - '0');
                }
            }
            if (eneg) e = -e;
            if (digits == 0) return false;
        }
    }
    *out = (negative ? -1 : 1) * (e ? std::ldexp(m * std::pow(5.0, e), e) : m);
    return true;
}

// One number of a 'v' / 'vn' / 'vt' line (parseReal): the next
// blank-delimited token, `fallback` when it does not parse; rounded to float.
float next_real(const char*& t, double fallback) {
    t += strspn(t, " \t");
    const char* end = t + strcspn(t, " \t\r");
    double val = fallback;
    parse_decimal(t, end, &val);
    t = end;
    return (float)val;
}

struct Corner {
    int v = -1, vt = -1, vn = -1;
};

// 1-based index, or negative relative to the n elements read so far; 0 is invalid.
bool obj_index(int idx, int n, int* ret) {
    if (idx > 0) { *ret = idx - 1; return true; }
    if (idx < 0) { *ret = n + idx; return true; }
    return false;
}

// One face corner: "v", "v/t", "v//n" or "v/t/n".
bool next_corner(const char*& t, int nv, int nvn, int nvt, Corner* out) {
    Corner r;
    if (!obj_index(atoi(t), nv, &r.v)) return false;
    t += strcspn(t, "/ \t\r");
    if (*t == '/') {
        ++t;
        if (*t == '/') {
            ++t;
            if (!obj_index(atoi(t), nvn, &r.vn)) return false;
            t += strcspn(t, "/ \t\r");
        } else {
            if (!obj_index(atoi(t), nvt, &r.vt)) return false;
            t += strcspn(t, "/ \t\r");
            if (*t == '/') {
                ++t;
                if (!obj_index(atoi(t), nvn, &r.vn)) return false;
                t += strcspn(t, "/ \t\r");
            }
        }
    }
    *out = r;
    return true;
}

// Crossing-number test of (px, py) against a triangle.
bool in_triangle(const float* xs, const float* ys, float px, float py) {
    bool inside = false;
    for (int i = 0, j = 2; i < 3; j = i++) {
        if (((ys[i] > py) != (ys[j] > py)) && (px < (xs[j] - xs[i]) * (py - ys[i]) / (ys[j] - ys[i]) + xs[i]))
            inside = !inside;
    }
    return inside;
}

struct Loader {
    std::vector<float> v, vn, vt;
    std::vector<std::vector<Corner>> pending;  // faces of the open group
    std::vector<Corner> tri;                   // triangle corners, in order

    bool pos_ok(int vi) const { return vi >= 0 && 3 * (size_t)vi + 2 < v.size(); }
    float coord(int vi, size_t axis) const {
        size_t k = 3 * (size_t)(unsigned)vi + axis;
        return k < v.size() ? v[k] : 0.0f;
    }

    // Triangulates one polygon with the positions read so far.
    void triangulate(const std::vector<Corner>& face) {
        size_t n = face.size();
        if (n < 3) return;
        // projection plane: drop the axis of largest |normal| at the first
        // corner whose float cross product exceeds float epsilon
        size_t ax0 = 1, ax1 = 2;
        for (size_t k = 0; k < n; ++k) {
            int a = face[k].v, b = face[(k + 1) % n].v, c = face[(k + 2) % n].v;
            if (!pos_ok(a) || !pos_ok(b) || !pos_ok(c)) continue;
            float e0x = v[3 * b] - v[3 * a], e0y = v[3 * b + 1] - v[3 * a + 1], e0z = v[3 * b + 2] - v[3 * a + 2];
            float e1x = v[3 * c] - v[3 * b], e1y = v[3 * c + 1] - v[3 * b + 1], e1z = v[3 * c + 2] - v[3 * b + 2];
            float cx = std::fabs(e0y * e1z - e0z * e1y);
            float cy = std::fabs(e0z * e1x - e0x * e1z);
            float cz = std::fabs(e0x * e1y - e0y * e1x);
            const float eps = std::numeric_limits<float>::epsilon();
            if (cx > eps || cy > eps || cz > eps) {
                if (!(cx > cy && cx > cz)) {
                    ax0 = 0;
                    if (cz > cx && cz > cy) ax1 = 1;
                }
                break;
            }
        }
        // signed area in that plane (float, halved per edge)
        float area = 0;
        for (size_t k = 0; k < n; ++k) {
            int a = face[k].v, b = face[(k + 1) % n].v;
            size_t ka0 = 3 * (size_t)(unsigned)a + ax0, ka1 = 3 * (size_t)(unsigned)a + ax1;
            size_t kb0 = 3 * (size_t)(unsigned)b + ax0, kb1 = 3 * (size_t)(unsigned)b + ax1;
            if (ka0 >= v.size() || ka1 >= v.size() || kb0 >= v.size() || kb1 >= v.size()) continue;
            area += (v[ka0] * v[kb1] - v[ka1] * v[kb0]) * static_cast<float>(0.5);
        }
        // ear clipping: try the triangle at `guess`; clip its middle corner
        // when it turns the polygon's way and holds no other corner
        std::vector<Corner> rest = face;
        size_t guess = 0;
        size_t tries = rest.size(), last_size = rest.size();
        while (rest.size() > 3 && tries > 0) {
            size_t m = rest.size();
            if (guess >= m) guess -= m;
            if (last_size != m) {
                last_size = m;
                tries = m;
            } else {
                --tries;
            }
            Corner c3[3];
            float xs[3], ys[3];
            for (size_t k = 0; k < 3; ++k) {
                c3[k] = rest[(guess + k) % m];
                xs[k] = coord(c3[k].v, ax0);
                ys[k] = coord(c3[k].v, ax1);
                if (3 * (size_t)(unsigned)c3[k].v + ax0 >= v.size() || 3 * (size_t)(unsigned)c3[k].v + ax1 >= v.size())
                    xs[k] = ys[k] = 0.0f;
            }
            float turn = (xs[1] - xs[0]) * (ys[2] - ys[1]) - (ys[1] - ys[0]) * (xs[2] - xs[1]);
            if (turn * area < static_cast<float>(0.0)) {
                guess += 1;
                continue;
            }
            bool blocked = false;
            for (size_t o = 3; o < m; ++o) {
                size_t idx = (guess + o) % m;
                int ov = rest[idx].v;
                size_t k0 = 3 * (size_t)(unsigned)ov + ax0, k1 = 3 * (size_t)(unsigned)ov + ax1;
                if (k0 >= v.size() || k1 >= v.size()) continue;
                if (in_triangle(xs, ys, v[k0], v[k1])) {
                    blocked = true;
                    break;
                }
            }
            if (blocked) {
                guess += 1;
                continue;
            }
            tri.push_back(c3[0]);
            tri.push_back(c3[1]);
            tri.push_back(c3[2]);
            rest.erase(rest.begin() + (long)((guess + 1) % m));
        }
        if (rest.size() == 3) tri.insert(tri.end(), rest.begin(), rest.end());
    }

    void flush() {
        for (const auto& f : pending) triangulate(f);
        pending.clear();
    }

    bool line(const char* t, std::string& err, size_t line_no) {
        t += strspn(t, " \t");
        if (t[0] == '\0' || t[0] == '#') return true;
        if (t[0] == 'v' && is_blank(t[1])) {
            t += 2;
            for (int k = 0; k < 3; ++k) v.push_back(next_real(t, 0.0));
            return true;
        }
        if (t[0] == 'v' && t[1] == 'n' && is_blank(t[2])) {
            t += 3;
            for (int k = 0; k < 3; ++k) vn.push_back(next_real(t, 0.0));
            return true;
        }
        if (t[0] == 'v' && t[1] == 't' && is_blank(t[2])) {
            t += 3;
            for (int k = 0; k < 2; ++k) vt.push_back(next_real(t, 0.0));
            return true;
        }
        bool face = t[0] == 'f' && is_blank(t[1]);
        if (face || ((t[0] == 'l' || t[0] == 'p') && is_blank(t[1]))) {
            t += 2;
            if (face) t += strspn(t, " \t");
            std::vector<Corner> poly;
            while (!is_line_end(t[0])) {
                Corner c;
                if (!next_corner(t, (int)(v.size() / 3), (int)(vn.size() / 3), (int)(vt.size() / 2), &c)) {
                    err = "invalid index (zero) on line " + std::to_string(line_no);
                    return false;
                }
                poly.push_back(c);
                t += strspn(t, " \t\r");
            }
            if (face) pending.push_back(poly);  // lines and points make no triangles
            return true;
        }
        if ((t[0] == 'g' || t[0] == 'o') && is_blank(t[1])) flush();
        return true;  // other statements do not change the geometry
    }
};

uint32_t fail(const std::string& msg) {
    rp::rp_set_last_error(RP_ERR_OBJ, "obj: " + msg);
    return RP_ERR_OBJ;
}

uint32_t parse_text(const std::string& text, rp_vertex* out, uint32_t cap, uint32_t* count) {
    Loader L;
    std::string err;
    size_t pos = 0, line_no = 0;
    std::string buf;
    while (pos < text.size()) {
        size_t e = pos;
        while (e < text.size() && text[e] != '\n' && text[e] != '\r') ++e;
        buf.assign(text, pos, e - pos);
        if (e < text.size() && text[e] == '\r' && e + 1 < text.size() && text[e + 1] == '\n') ++e;
        pos = e + 1;
        ++line_no;
        if (!L.line(buf.c_str(), err, line_no)) return fail(err);
    }
    L.flush();
    const size_t nv = L.v.size() / 3, nvn = L.vn.size() / 3, nvt = L.vt.size() / 2;
    const bool make_normals = L.vn.empty();
    for (const Corner& c : L.tri) {
        if (c.v < 0 || (size_t)c.v >= nv) return fail("vertex index out of range");
        if (!make_normals && (c.vn < 0 || (size_t)c.vn >= nvn)) return fail("normal index out of range");
        if (!L.vt.empty() && (c.vt < 0 || (size_t)c.vt >= nvt)) return fail("texture coordinate index out of range");
    }
    if (L.tri.size() > UINT32_MAX) return fail("too many corners");
    *count = (uint32_t)L.tri.size();
    if (!out) return 0;
    for (size_t i = 0; i < L.tri.size() && i < cap; ++i) {
        const Corner& c = L.tri[i];
        rp_vertex& o = out[i];
        memset(&o, 0, sizeof(o));
        for (int k = 0; k < 3; ++k) o.position[k] = L.v[3 * (size_t)c.v + k];
        if (!L.vt.empty())
            for (int k = 0; k < 2; ++k) o.texture_coordinates[k] = L.vt[2 * (size_t)c.vt + k];
        if (!make_normals)
            for (int k = 0; k < 3; ++k) o.normal[k] = L.vn[3 * (size_t)c.vn + k];
    }
    if (make_normals) {
        for (size_t i = 0; i + 2 < L.tri.size() && i + 2 < cap; i += 3) {
            const float* A = out[i].position;
            const float* B = out[i + 1].position;
            const float* C = out[i + 2].position;
            float e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
            float e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
            float nx = e1[1] * e2[2] - e1[2] * e2[1];
            float ny = e1[2] * e2[0] - e1[0] * e2[2];
            float nz = e1[0] * e2[1] - e1[1] * e2[0];
            for (size_t k = i; k < i + 3; ++k) {
                out[k].normal[0] = out[k].normal[0] + nx;
                out[k].normal[1] = out[k].normal[1] + ny;
                out[k].normal[2] = out[k].normal[2] + nz;
            }
        }
    }
    return 0;
}

}  // namespace

extern "C" {

uint32_t rp_obj_parse(const char* obj_path, rp_vertex* vertices, uint32_t capacity, uint32_t* num_vertices) {
    if (!obj_path || !num_vertices) return fail("null argument");
    std::ifstream f(obj_path, std::ios::binary);
    if (!f) return fail(std::string("cannot open ") + obj_path);
    std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    return parse_text(text, vertices, capacity, num_vertices);
}

uint32_t rp_obj_parse_memory(const char* text, size_t length, rp_vertex* vertices, uint32_t capacity,
                             uint32_t* num_vertices) {
    if (!text || !num_vertices) return fail("null argument");
    return parse_text(std::string(text, length), vertices, capacity, num_vertices);
}

}  // extern "C"
