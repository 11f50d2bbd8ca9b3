// Mesh readers of the app (js/meshio.js:19-214, dispatch js/meshio.js:322-340):
// binary and ASCII STL, OBJ (fan-triangulated) and binary glTF 2.0 (node
// transforms applied) into a float32 triangle soup, 9 values per triangle,
// in the file's own coordinates. Host code: parsing is byte work that ends
// in one upload for Voxel.build.
//
// The conversions follow the reference's JavaScript: numbers through the
// rules of Number() and parseInt(s, 10), doubles rounded once to float32,
// ASCII STL through its vertex pattern, glTF integer components stored as
// unsigned 32-bit values, and its error messages.

#include "wt_internal.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <functional>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace {

struct ParseError {
    std::string msg;
};

[[noreturn]] void fail(const std::string &m) { throw ParseError{m}; }

// ---- text: UTF-8 code points and JavaScript whitespace -------------------

// Code point at s[i] (invalid bytes read as U+FFFD), advancing i.
uint32_t next_cp(const std::string &s, size_t &i)
{
    const unsigned char c = (unsigned char)s[i];
    if (c < 0x80) { i++; return c; }
    const int n = c >= 0xF0 ? 3 : c >= 0xE0 ? 2 : c >= 0xC0 ? 1 : -1;
    if (n < 0 || i + n >= s.size()) { i++; return 0xFFFD; }
    uint32_t cp = c & (0x3F >> n);
    for (int k = 1; k <= n; k++) {
        const unsigned char d = (unsigned char)s[i + k];
        if ((d & 0xC0) != 0x80) { i++; return 0xFFFD; }
        cp = (cp << 6) | (d & 0x3F);
    }
    i += n + 1;
    return cp;
}

// WhiteSpace and LineTerminator of ECMAScript (the \s class, trim()).
bool js_space(uint32_t c)
{
    return c == 9 || c == 10 || c == 11 || c == 12 || c == 13 || c == 32 || c == 0xA0 || c == 0x1680 ||
           (c >= 0x2000 && c <= 0x200A) || c == 0x2028 || c == 0x2029 || c == 0x202F || c == 0x205F ||
           c == 0x3000 || c == 0xFEFF;
}

size_t skip_space(const std::string &s, size_t i)
{
    while (i < s.size()) {
        size_t j = i;
        if (!js_space(next_cp(s, j))) break;
        i = j;
    }
    return i;
}

std::string js_trim(const std::string &s)
{
    size_t a = skip_space(s, 0);
    size_t b = s.size();
    // trailing: walk code points from a, remember the end of the last non-space
    size_t i = a, last = a;
    while (i < b) {
        size_t j = i;
        const uint32_t c = next_cp(s, j);
        if (!js_space(c)) last = j;
        i = j;
    }
    return s.substr(a, last - a);
}

// Number(str): StringToNumber of ECMAScript.
double js_number(const std::string &raw)
{
    const std::string s = js_trim(raw);
    if (s.empty()) return 0;
    const char *p = s.c_str();
    const size_t n = s.size();
    if (s == "Infinity" || s == "+Infinity") return INFINITY;
    if (s == "-Infinity") return -INFINITY;
    if (n > 2 && p[0] == '0' && (p[1] | 32) != 'e' && !(p[1] >= '0' && p[1] <= '9') && p[1] != '.') {
        int base = (p[1] | 32) == 'x' ? 16 : (p[1] | 32) == 'o' ? 8 : (p[1] | 32) == 'b' ? 2 : 0;
        if (!base) return NAN;
        double v = 0;
        for (size_t k = 2; k < n; k++) {
            const char c = p[k];
            int d = c >= '0' && c <= '9' ? c - '0' : (c | 32) >= 'a' && (c | 32) <= 'f' ? (c | 32) - 'a' + 10 : 99;
            if (d >= base) return NAN;
            v = v * base + d;
        }
        return v;
    }
    // StrDecimalLiteral: [+-] (digits [. digits] | . digits) [(e|E) [+-] digits]
    size_t k = 0;
    if (p[k] == '+' || p[k] == '-') k++;
    size_t d0 = k;
    while (k < n && p[k] >= '0' && p[k] <= '9') k++;
    size_t nd = k - d0;
    if (k < n && p[k] == '.') {
        k++;
        size_t f0 = k;
        while (k < n && p[k] >= '0' && p[k] <= '9') k++;
        nd += k - f0;
    }
    if (nd == 0) return NAN;
    if (k < n && (p[k] | 32) == 'e') {
        k++;
        if (k < n && (p[k] == '+' || p[k] == '-')) k++;
        size_t e0 = k;
        while (k < n && p[k] >= '0' && p[k] <= '9') k++;
        if (k == e0) return NAN;
    }
    if (k != n) return NAN;
    return std::strtod(p, nullptr);  // correctly rounded, as Number() is
}

// parseInt(str, 10): leading whitespace, sign, decimal digits; NaN without digits.
double js_parse_int(const std::string &s)
{
    size_t i = skip_space(s, 0);
    double sign = 1;
    if (i < s.size() && (s[i] == '+' || s[i] == '-')) { sign = s[i] == '-' ? -1 : 1; i++; }
    size_t d0 = i;
    while (i < s.size() && s[i] >= '0' && s[i] <= '9') i++;
    if (i == d0) return NAN;
    return sign * std::strtod(s.substr(d0, i - d0).c_str(), nullptr);
}

// ---- STL -------------------------------------------------------------------

uint32_t u32(const uint8_t *b) { uint32_t v; std::memcpy(&v, b, 4); return v; }

bool is_binary_stl(const uint8_t *b, size_t len)
{
    if (len < 84) return false;
    return 84 + (double)u32(b + 80) * 50 == (double)len;
}

std::vector<float> stl_binary(const uint8_t *b, size_t len)
{
    (void)len;
    const uint32_t n = u32(b + 80);
    std::vector<float> out((size_t)n * 9);
    size_t o = 84, p = 0;
    for (uint32_t i = 0; i < n; i++) {
        o += 12;
        for (int k = 0; k < 9; k++, o += 4) std::memcpy(&out[p++], b + o, 4);
        o += 2;
    }
    return out;
}

bool stl_token_char(char c) { return (c >= '0' && c <= '9') || c == '.' || c == 'e' || c == 'E' || c == '+' || c == '-'; }

// The vertex pattern: "vertex", whitespace, then three runs of [-0-9.eE+]
// separated by whitespace, searched left to right without overlap.
std::vector<float> stl_ascii(const std::string &t)
{
    std::vector<double> nums;
    size_t i = 0;
    while ((i = t.find("vertex", i)) != std::string::npos) {
        size_t j = i + 6;
        double v[3];
        bool ok = true;
        for (int c = 0; c < 3 && ok; c++) {
            const size_t s0 = j;
            j = skip_space(t, j);
            if (j == s0) { ok = false; break; }
            const size_t a = j;
            while (j < t.size() && stl_token_char(t[j])) j++;
            if (j == a) { ok = false; break; }
            v[c] = js_number(t.substr(a, j - a));
        }
        if (!ok) { i++; continue; }
        nums.insert(nums.end(), v, v + 3);
        i = j;
    }
    if (nums.empty()) fail("No vertices found in ASCII STL.");
    return std::vector<float>(nums.begin(), nums.end());
}

// ---- OBJ -------------------------------------------------------------------

std::vector<std::string> split_space(const std::string &s)
{
    // String.prototype.split(/\s+/) on a trimmed line
    std::vector<std::string> out;
    size_t i = 0, a = 0;
    while (i < s.size()) {
        size_t j = i;
        if (js_space(next_cp(s, j))) {
            out.push_back(s.substr(a, i - a));
            i = skip_space(s, j);
            a = i;
        } else {
            i = j;
        }
    }
    out.push_back(s.substr(a));
    return out;
}

std::vector<float> obj(const std::string &t)
{
    std::vector<double> vx, vy, vz;
    std::vector<double> idx;
    size_t a = 0;
    while (a <= t.size()) {
        size_t b = t.find('\n', a);
        if (b == std::string::npos) b = t.size();
        const std::string line = js_trim(t.substr(a, b - a));
        a = b + 1;
        // line.length counts UTF-16 units; two ASCII bytes decide every case used below
        size_t units = 0;
        for (size_t i = 0; i < line.size() && units < 2;) {
            const uint32_t c = next_cp(line, i);
            units += c >= 0x10000 ? 2 : 1;
        }
        if (units < 2) continue;
        if (line[0] == 'v' && line[1] == ' ') {
            const auto p = split_space(line);
            vx.push_back(p.size() > 1 ? js_number(p[1]) : NAN);
            vy.push_back(p.size() > 2 ? js_number(p[2]) : NAN);
            vz.push_back(p.size() > 3 ? js_number(p[3]) : NAN);
        } else if (line[0] == 'f' && line[1] == ' ') {
            const auto p = split_space(line);
            std::vector<double> face;
            for (size_t i = 1; i < p.size(); i++) {
                const double v = js_parse_int(p[i]);
                if (!std::isnan(v)) face.push_back(v < 0 ? (double)vx.size() + v : v - 1);
            }
            for (size_t i = 1; i + 1 < face.size(); i++) {
                idx.push_back(face[0]); idx.push_back(face[i]); idx.push_back(face[i + 1]);
            }
        }
        if (b == t.size()) break;
    }
    if (idx.empty()) fail("No faces found in OBJ.");
    std::vector<float> out(idx.size() * 3);
    for (size_t i = 0; i < idx.size(); i++) {
        const double v = idx[i];
        const bool in = v >= 0 && v < (double)vx.size() && v == std::floor(v);
        const size_t k = in ? (size_t)v : 0;
        out[i * 3] = in ? (float)vx[k] : NAN;
        out[i * 3 + 1] = in ? (float)vy[k] : NAN;
        out[i * 3 + 2] = in ? (float)vz[k] : NAN;
    }
    return out;
}

// ---- JSON (the glTF chunk) --------------------------------------------------

struct Json {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    double num = 0;
    bool b = false;
    std::string str;
    std::vector<Json> arr;
    std::map<std::string, Json> obj;
    const Json *get(const char *k) const
    {
        if (kind != OBJ) return nullptr;
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
    const Json *at(double i) const
    {
        if (kind != ARR || !(i >= 0) || i != std::floor(i) || i >= (double)arr.size()) return nullptr;
        return &arr[(size_t)i];
    }
};

struct JsonReader {
    const std::string &s;
    size_t i = 0;
    void ws() { while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) i++; }
    [[noreturn]] void bad() { fail("GLB JSON chunk is not valid JSON."); }
    void put_utf8(std::string &o, uint32_t c)
    {
        if (c < 0x80) o += (char)c;
        else if (c < 0x800) { o += (char)(0xC0 | (c >> 6)); o += (char)(0x80 | (c & 63)); }
        else if (c < 0x10000) { o += (char)(0xE0 | (c >> 12)); o += (char)(0x80 | ((c >> 6) & 63)); o += (char)(0x80 | (c & 63)); }
        else { o += (char)(0xF0 | (c >> 18)); o += (char)(0x80 | ((c >> 12) & 63)); o += (char)(0x80 | ((c >> 6) & 63)); o += (char)(0x80 | (c & 63)); }
    }
    uint32_t hex4()
    {
        if (i + 4 > s.size()) bad();
        uint32_t v = 0;
        for (int k = 0; k < 4; k++) {
            const char c = s[i++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= c - '0';
            else if ((c | 32) >= 'a' && (c | 32) <= 'f') v |= (c | 32) - 'a' + 10;
            else bad();
        }
        return v;
    }
    std::string string()
    {
        std::string o;
        i++;  // opening quote
        while (true) {
            if (i >= s.size()) bad();
            const char c = s[i++];
            if (c == '"') return o;
            if ((unsigned char)c < 0x20) bad();
            if (c != '\\') { o += c; continue; }
            if (i >= s.size()) bad();
            const char e = s[i++];
            switch (e) {
            case '"': o += '"'; break;
            case '\\': o += '\\'; break;
            case '/': o += '/'; break;
            case 'b': o += '\b'; break;
            case 'f': o += '\f'; break;
            case 'n': o += '\n'; break;
            case 'r': o += '\r'; break;
            case 't': o += '\t'; break;
            case 'u': {
                uint32_t c1 = hex4();
                if (c1 >= 0xD800 && c1 < 0xDC00 && i + 6 <= s.size() && s[i] == '\\' && s[i + 1] == 'u') {
                    const size_t save = i;
                    i += 2;
                    const uint32_t c2 = hex4();
                    if (c2 >= 0xDC00 && c2 < 0xE000) c1 = 0x10000 + ((c1 - 0xD800) << 10) + (c2 - 0xDC00);
                    else i = save;
                }
                put_utf8(o, (c1 >= 0xD800 && c1 < 0xE000) ? 0xFFFD : c1);
                break;
            }
            default: bad();
            }
        }
    }
    Json value(int depth)
    {
        if (depth > 512) bad();
        ws();
        if (i >= s.size()) bad();
        Json v;
        const char c = s[i];
        if (c == '{') {
            v.kind = Json::OBJ;
            i++;
            ws();
            if (i < s.size() && s[i] == '}') { i++; return v; }
            while (true) {
                ws();
                if (i >= s.size() || s[i] != '"') bad();
                std::string k = string();
                ws();
                if (i >= s.size() || s[i] != ':') bad();
                i++;
                v.obj[k] = value(depth + 1);  // a repeated key keeps its last value, as JSON.parse does
                ws();
                if (i < s.size() && s[i] == ',') { i++; continue; }
                if (i < s.size() && s[i] == '}') { i++; return v; }
                bad();
            }
        }
        if (c == '[') {
            v.kind = Json::ARR;
            i++;
            ws();
            if (i < s.size() && s[i] == ']') { i++; return v; }
            while (true) {
                v.arr.push_back(value(depth + 1));
                ws();
                if (i < s.size() && s[i] == ',') { i++; continue; }
                if (i < s.size() && s[i] == ']') { i++; return v; }
                bad();
            }
        }
        if (c == '"') { v.kind = Json::STR; v.str = string(); return v; }
        if (s.compare(i, 4, "true") == 0) { i += 4; v.kind = Json::BOOL; v.b = true; return v; }
        if (s.compare(i, 5, "false") == 0) { i += 5; v.kind = Json::BOOL; return v; }
        if (s.compare(i, 4, "null") == 0) { i += 4; return v; }
        // number: -? (0 | [1-9][0-9]*) (. [0-9]+)? ([eE] [+-]? [0-9]+)?
        const size_t a = i;
        if (s[i] == '-') i++;
        if (i < s.size() && s[i] == '0') i++;
        else if (i < s.size() && s[i] >= '1' && s[i] <= '9') { while (i < s.size() && isdigit((unsigned char)s[i])) i++; }
        else bad();
        if (i < s.size() && s[i] == '.') {
            i++;
            if (!(i < s.size() && isdigit((unsigned char)s[i]))) bad();
            while (i < s.size() && isdigit((unsigned char)s[i])) i++;
        }
        if (i < s.size() && (s[i] | 32) == 'e') {
            i++;
            if (i < s.size() && (s[i] == '+' || s[i] == '-')) i++;
            if (!(i < s.size() && isdigit((unsigned char)s[i]))) bad();
            while (i < s.size() && isdigit((unsigned char)s[i])) i++;
        }
        v.kind = Json::NUM;
        v.num = std::strtod(s.substr(a, i - a).c_str(), nullptr);
        return v;
    }
};

// ---- glTF 2.0 binary ----------------------------------------------------------

double num_or(const Json *j, double d) { return j && j->kind == Json::NUM ? j->num : d; }

// A JSON value used as a byte offset or count must be a whole number.
size_t whole(double v, const char *what)
{
    if (!(v >= 0) || v != std::floor(v) || v > 9.0e15) fail(std::string("GLB ") + what + " is not a whole number.");
    return (size_t)v;
}

struct Accessor {
    bool is_float = false;
    std::vector<float> f;
    std::vector<uint32_t> u;
    size_t size() const { return is_float ? f.size() : u.size(); }
    double at(double k) const
    {
        if (!(k >= 0) || k != std::floor(k) || k >= (double)size()) return NAN;  // undefined reads as NaN
        return is_float ? (double)f[(size_t)k] : (double)u[(size_t)k];
    }
};

// readAccessor (js/meshio.js:87-114): false when the accessor has no bufferView.
bool read_accessor(const Json &g, const uint8_t *file, size_t flen, size_t bin_off, double index, Accessor &out)
{
    const Json *accs = g.get("accessors");
    const Json *a = accs ? accs->at(index) : nullptr;
    if (!a || a->kind != Json::OBJ) fail("GLB accessor " + std::to_string((long long)index) + " does not exist.");
    const Json *bvi = a->get("bufferView");
    if (!bvi) return false;
    const Json *bvs = g.get("bufferViews");
    const Json *bv = bvs ? bvs->at(num_or(bvi, -1)) : nullptr;
    if (!bv || bv->kind != Json::OBJ) fail("GLB bufferView does not exist.");
    static const std::map<std::string, int> NC = {{"SCALAR", 1}, {"VEC2", 2}, {"VEC3", 3}, {"VEC4", 4}, {"MAT4", 16}};
    const Json *ty = a->get("type");
    auto it = ty && ty->kind == Json::STR ? NC.find(ty->str) : NC.end();
    if (it == NC.end()) fail("GLB accessor type is not supported.");
    const int nc = it->second;
    const double ctd = num_or(a->get("componentType"), -1);
    const int ct = (int)ctd;
    const int cs = ct == 5120 || ct == 5121 ? 1 : ct == 5122 || ct == 5123 ? 2 : ct == 5125 || ct == 5126 ? 4 : 0;
    if (!cs || ctd != ct) fail("GLB accessor componentType is not supported.");
    const double bst = num_or(bv->get("byteStride"), 0);
    const size_t stride = bst ? whole(bst, "byteStride") : (size_t)(cs * nc);
    const size_t base = bin_off + whole(num_or(bv->get("byteOffset"), 0), "byteOffset") +
                        whole(num_or(a->get("byteOffset"), 0), "byteOffset");
    const size_t count = whole(num_or(a->get("count"), NAN), "accessor count");
    out.is_float = ct == 5126;
    if (out.is_float) out.f.assign(count * nc, 0.0f);
    else out.u.assign(count * nc, 0u);
    for (size_t k = 0; k < count; k++) {
        size_t p = base + k * stride;
        for (int c = 0; c < nc; c++, p += cs) {
            if (p + cs > flen) fail("GLB accessor reads past the end of the file.");
            const uint8_t *q = file + p;
            const size_t o = k * nc + c;
            switch (ct) {
            case 5126: std::memcpy(&out.f[o], q, 4); break;
            case 5125: out.u[o] = u32(q); break;
            case 5123: out.u[o] = (uint32_t)(q[0] | (q[1] << 8)); break;
            case 5122: out.u[o] = (uint32_t)(int32_t)(int16_t)(q[0] | (q[1] << 8)); break;
            case 5121: out.u[o] = q[0]; break;
            default: out.u[o] = (uint32_t)(int32_t)(int8_t)q[0]; break;
            }
        }
    }
    return true;
}

using M4 = std::vector<double>;

// column-major product, as glTF stores its matrices (js/meshio.js:118-127)
M4 mul4(const M4 &a, const M4 &b)
{
    M4 o(16);
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++)
            o[c * 4 + r] = a[r] * b[c * 4] + a[4 + r] * b[c * 4 + 1] + a[8 + r] * b[c * 4 + 2] + a[12 + r] * b[c * 4 + 3];
    return o;
}

// node transform: matrix, or translation * rotation (x, y, z, w) * scale (js/meshio.js:129-146)
M4 node_matrix(const Json &n)
{
    auto vec = [&](const char *k, std::vector<double> d) {
        const Json *j = n.get(k);
        if (!j || j->kind != Json::ARR) return d;
        std::vector<double> v(d.size(), NAN);
        for (size_t i = 0; i < v.size(); i++) v[i] = i < j->arr.size() ? num_or(&j->arr[i], NAN) : NAN;
        return v;
    };
    const Json *m = n.get("matrix");
    if (m && m->kind == Json::ARR) {
        M4 o(16, NAN);
        for (size_t i = 0; i < 16 && i < m->arr.size(); i++) o[i] = num_or(&m->arr[i], NAN);
        return o;
    }
    const auto t = vec("translation", {0, 0, 0});
    const auto r = vec("rotation", {0, 0, 0, 1});
    const auto s = vec("scale", {1, 1, 1});
    const double x = r[0], y = r[1], z = r[2], w = r[3];
    const double x2 = x + x, y2 = y + y, z2 = z + z;
    const double xx = x * x2, xy = x * y2, xz = x * z2;
    const double yy = y * y2, yz = y * z2, zz = z * z2;
    const double wx = w * x2, wy = w * y2, wz = w * z2;
    return {(1 - (yy + zz)) * s[0], (xy + wz) * s[0], (xz - wy) * s[0], 0,
            (xy - wz) * s[1], (1 - (xx + zz)) * s[1], (yz + wx) * s[1], 0,
            (xz + wy) * s[2], (yz - wx) * s[2], (1 - (xx + yy)) * s[2], 0,
            t[0], t[1], t[2], 1};
}

std::vector<float> glb(const uint8_t *b, size_t len)
{
    if (len < 4) fail("GLB header is truncated.");
    if (u32(b) != 0x46546C67u) fail("Not a GLB file (bad magic).");
    if (len < 12) fail("GLB header is truncated.");
    const size_t total = std::min<size_t>(u32(b + 8), len);
    size_t o = 12;
    std::unique_ptr<Json> g;
    bool has_bin = false;
    size_t bin_off = 0;
    while (o + 8 <= total) {
        const size_t clen = u32(b + o), type = u32(b + o + 4);
        const size_t start = o + 8;
        if ((type == 0x4E4F534Au || type == 0x004E4942u) && start + clen > len) fail("GLB chunk runs past the end of the file.");
        if (type == 0x4E4F534Au) {
            const std::string txt((const char *)b + start, clen);
            JsonReader r{txt};
            // TextDecoder drops a leading byte-order mark
            if (txt.compare(0, 3, "\xEF\xBB\xBF") == 0) r.i = 3;
            g.reset(new Json(r.value(0)));
            r.ws();
            if (r.i != txt.size()) r.bad();
        } else if (type == 0x004E4942u) {
            has_bin = true;
            bin_off = start;
        }
        o = start + clen + ((4 - (clen & 3)) & 3);
    }
    if (!g) fail("GLB contains no JSON chunk.");
    if (!has_bin) fail("GLB has no binary chunk (external .bin files are not supported).");
    const Json *ext = g->get("extensionsRequired");
    if (ext && ext->kind == Json::ARR && !ext->arr.empty()) {
        std::string names;
        for (size_t i = 0; i < ext->arr.size(); i++) {
            if (i) names += ", ";
            const Json &e = ext->arr[i];
            names += e.kind == Json::STR ? e.str : e.kind == Json::NUL ? "" : "?";
        }
        fail("This GLB needs " + names + ". Re-export without compression (uncheck Draco / meshopt) and try again.");
    }
    const Json *meshes = g->get("meshes");
    if (!meshes || meshes->kind != Json::ARR || meshes->arr.empty()) fail("GLB contains no meshes.");

    std::vector<float> out;
    auto emit = [&](const Json &prim, const M4 &M) {
        const Json *mode = prim.get("mode");
        if (mode && !(mode->kind == Json::NUM && mode->num == 4)) return;
        const Json *attrs = prim.get("attributes");
        const Json *pa = attrs ? attrs->get("POSITION") : nullptr;
        if (!pa) return;
        Accessor pos, ind;
        if (!read_accessor(*g, b, len, bin_off, num_or(pa, NAN), pos)) return;
        const Json *pi = prim.get("indices");
        const bool indexed = pi && read_accessor(*g, b, len, bin_off, num_or(pi, NAN), ind);
        const double ntri = indexed ? (double)ind.size() / 3 : (double)pos.size() / 9;
        if (ntri != std::floor(ntri)) fail("Invalid typed array length: " + std::to_string(ntri * 9));
        const size_t base = out.size();
        out.resize(base + (size_t)ntri * 9);
        for (size_t t = 0; t < (size_t)ntri * 3; t++) {
            const double v = (indexed ? ind.at((double)t) : (double)t) * 3;
            const double x = pos.at(v), y = pos.at(v + 1), z = pos.at(v + 2);
            out[base + t * 3] = (float)(M[0] * x + M[4] * y + M[8] * z + M[12]);
            out[base + t * 3 + 1] = (float)(M[1] * x + M[5] * y + M[9] * z + M[13]);
            out[base + t * 3 + 2] = (float)(M[2] * x + M[6] * y + M[10] * z + M[14]);
        }
    };
    auto prims = [&](const Json &mesh, const M4 &M) {
        const Json *ps = mesh.get("primitives");
        if (!ps || ps->kind != Json::ARR) fail("GLB mesh has no primitives list.");
        for (const Json &p : ps->arr) emit(p, M);
    };
    const M4 IDENT = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    const Json *nodes = g->get("nodes");
    std::function<void(double, const M4 &, int)> walk = [&](double ni, const M4 &parent, int depth) {
        if (depth > 64) return;
        const Json *n = nodes ? nodes->at(ni) : nullptr;
        if (!n || n->kind != Json::OBJ) return;
        const M4 M = mul4(parent, node_matrix(*n));
        const Json *mi = n->get("mesh");
        if (mi) {
            const Json *mesh = meshes->at(num_or(mi, NAN));
            if (mesh && mesh->kind == Json::OBJ) prims(*mesh, M);
        }
        const Json *ch = n->get("children");
        if (ch && ch->kind == Json::ARR)
            for (const Json &c : ch->arr) walk(num_or(&c, NAN), M, depth + 1);
    };
    const Json *scenes = g->get("scenes");
    // "scene" absent selects scene 0; a value that is not an index selects none
    const Json *sc = g->get("scene");
    const Json *scene = scenes ? scenes->at(sc ? num_or(sc, NAN) : 0) : nullptr;
    const Json *roots = scene ? scene->get("nodes") : nullptr;
    if (roots && roots->kind == Json::ARR) {
        for (const Json &r : roots->arr) walk(num_or(&r, NAN), IDENT, 0);
    } else if (nodes && nodes->kind == Json::ARR) {
        for (size_t i = 0; i < nodes->arr.size(); i++) walk((double)i, IDENT, 0);
    } else {
        for (const Json &m : meshes->arr) prims(m, IDENT);
    }
    if (out.empty()) fail("GLB has meshes but no readable triangles.");
    return out;
}

// ---- dispatch (js/meshio.js:322-340) --------------------------------------------

std::string lower_ext(const char *name)
{
    std::string n = name ? name : "";
    const size_t d = n.rfind('.');
    std::string e = d == std::string::npos ? n : n.substr(d + 1);
    for (char &c : e) c = (char)std::tolower((unsigned char)c);
    return e;
}

bool starts_solid(const std::string &t)
{
    const size_t i = skip_space(t, 0);
    return t.compare(i, 5, "solid") == 0;
}

// /^\s*(v|vn|vt|f|o|g|mtllib)\s/m on the first 4096 bytes
bool looks_obj(const std::string &t)
{
    size_t ls = 0;
    while (ls <= t.size()) {
        size_t i = skip_space(t, ls);
        static const char *keys[] = {"mtllib", "vn", "vt", "v", "f", "o", "g"};
        for (const char *k : keys) {
            const size_t kl = std::strlen(k);
            if (t.compare(i, kl, k) == 0 && i + kl < t.size()) {
                size_t j = i + kl;
                if (js_space(next_cp(t, j))) return true;
            }
        }
        // \s* can cross line ends, so every position after a line terminator starts a line
        size_t e = ls;
        bool found = false;
        while (e < t.size()) {
            size_t j = e;
            const uint32_t c = next_cp(t, j);
            e = j;
            if (c == '\n' || c == '\r' || c == 0x2028 || c == 0x2029) { found = true; break; }
        }
        if (!found) break;
        ls = e;
    }
    return false;
}

std::vector<float> parse(const char *name, const uint8_t *b, size_t len)
{
    const std::string ext = lower_ext(name);
    const std::string all((const char *)b, len);
    if (ext == "glb") return glb(b, len);
    if (ext == "gltf") fail("Plain .gltf references external files. Export as .glb (binary glTF) instead.");
    if (ext == "stl") return is_binary_stl(b, len) ? stl_binary(b, len) : stl_ascii(all);
    if (ext == "obj") return obj(all);
    if (len > 12 && u32(b) == 0x46546C67u) return glb(b, len);
    if (is_binary_stl(b, len)) return stl_binary(b, len);
    const std::string head = all.substr(0, std::min<size_t>(4096, len));
    if (starts_solid(head)) return stl_ascii(all);
    if (looks_obj(head)) return obj(all);
    fail("Unrecognised file. Supported: .obj, .stl, .glb");
}

}  // namespace

extern "C" int wt_mesh_parse(const char *name, const void *buf, int64_t len, float *out, int64_t *count)
{
    if (!count || (!buf && len > 0) || len < 0) return wt_fail(WT_E_ARG, "bad mesh parse request");
    try {
        const std::vector<float> soup = parse(name, (const uint8_t *)buf, (size_t)len);
        if (!out) {
            *count = (int64_t)soup.size();
            return 0;
        }
        if (*count < (int64_t)soup.size())
            return wt_fail(WT_E_ARG, "output holds %lld floats, the mesh needs %zu", (long long)*count, soup.size());
        std::memcpy(out, soup.data(), sizeof(float) * soup.size());
        *count = (int64_t)soup.size();
        return 0;
    } catch (const ParseError &e) {
        return wt_fail(WT_E_ARG, "%s", e.msg.c_str());
    } catch (const std::bad_alloc &) {
        return wt_fail(WT_E_NOMEM, "out of host memory while parsing the mesh");
    }
}
