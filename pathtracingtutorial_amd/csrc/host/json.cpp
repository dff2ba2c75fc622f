#include "json.h"

#include <cstdlib>
#include <stdexcept>

namespace ptamd {

static const Json& json_null() {
    static const Json n;
    return n;
}

const Json& Json::operator[](const std::string& k) const {
    if (type != Object) throw std::runtime_error("json: not an object (key '" + k + "')");
    auto it = obj.find(k);
    if (it == obj.end()) throw std::runtime_error("json: missing key '" + k + "'");
    return it->second;
}

const Json& Json::operator[](size_t i) const {
    if (type != Array) throw std::runtime_error("json: not an array");
    if (i >= arr.size()) throw std::runtime_error("json: index out of range");
    return arr[i];
}

float Json::as_float() const {
    if (type == Number) return static_cast<float>(num);
    if (type == Bool) return b ? 1.0f : 0.0f;
    throw std::runtime_error("json: expected a number");
}

int Json::as_int() const {
    if (type == Number) return static_cast<int>(num);
    if (type == Bool) return b ? 1 : 0;
    throw std::runtime_error("json: expected an integer");
}

bool Json::as_bool() const {
    if (type == Bool) return b;
    if (type == Number) return num != 0.0;
    throw std::runtime_error("json: expected a boolean");
}

namespace {

struct Parser {
    const std::string& s;
    size_t p = 0;
    std::string err;

    explicit Parser(const std::string& t) : s(t) {}

    void ws() {
        while (p < s.size()) {
            char c = s[p];
            if (c == ' ' || c == '\t' || c == '\n' || c == '\r') { ++p; continue; }
            // tolerate // line comments, which hand-written scene files sometimes carry
            if (c == '/' && p + 1 < s.size() && s[p + 1] == '/') {
                while (p < s.size() && s[p] != '\n') ++p;
                continue;
            }
            break;
        }
    }
    bool fail(const char* what) {
        if (err.empty()) err = std::string(what) + " at offset " + std::to_string(p);
        return false;
    }
    bool lit(const char* w) {
        size_t n = 0;
        while (w[n]) ++n;
        if (s.compare(p, n, w) != 0) return fail("unexpected token");
        p += n;
        return true;
    }
    static void put_utf8(std::string& o, unsigned cp) {
        if (cp < 0x80) o += static_cast<char>(cp);
        else if (cp < 0x800) { o += static_cast<char>(0xC0 | (cp >> 6)); o += static_cast<char>(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) {
            o += static_cast<char>(0xE0 | (cp >> 12)); o += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            o += static_cast<char>(0x80 | (cp & 0x3F));
        } else {
            o += static_cast<char>(0xF0 | (cp >> 18)); o += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
            o += static_cast<char>(0x80 | ((cp >> 6) & 0x3F)); o += static_cast<char>(0x80 | (cp & 0x3F));
        }
    }
    bool hex4(unsigned& v) {
        if (p + 4 > s.size()) return fail("bad \\u escape");
        v = 0;
        for (int i = 0; i < 4; ++i) {
            char c = s[p++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= c - '0';
            else if (c >= 'a' && c <= 'f') v |= c - 'a' + 10;
            else if (c >= 'A' && c <= 'F') v |= c - 'A' + 10;
            else return fail("bad \\u escape");
        }
        return true;
    }
    bool string(std::string& out) {
        if (p >= s.size() || s[p] != '"') return fail("expected string");
        ++p;
        out.clear();
        while (p < s.size()) {
            char c = s[p++];
            if (c == '"') return true;
            if (c == '\\') {
                if (p >= s.size()) break;
                char e = s[p++];
                switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    unsigned cp = 0;
                    if (!hex4(cp)) return false;
                    if (cp >= 0xD800 && cp < 0xDC00 && p + 6 <= s.size() && s[p] == '\\' && s[p + 1] == 'u') {
                        p += 2;
                        unsigned lo = 0;
                        if (!hex4(lo)) return false;
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    put_utf8(out, cp);
                    break;
                }
                default: return fail("bad escape");
                }
            } else {
                out += c;
            }
        }
        return fail("unterminated string");
    }
    bool number(Json& v) {
        size_t st = p;
        if (p < s.size() && (s[p] == '-' || s[p] == '+')) ++p;
        while (p < s.size() && ((s[p] >= '0' && s[p] <= '9') || s[p] == '.' || s[p] == 'e' || s[p] == 'E' ||
                                s[p] == '-' || s[p] == '+'))
            ++p;
        std::string tok = s.substr(st, p - st);
        char* end = nullptr;
        double d = std::strtod(tok.c_str(), &end);
        if (tok.empty() || end == tok.c_str() || *end != '\0') { p = st; return fail("bad number"); }
        v.type = Json::Number;
        v.num = d;
        return true;
    }
    bool value(Json& v, int depth) {
        if (depth > 256) return fail("nesting too deep");
        ws();
        if (p >= s.size()) return fail("unexpected end");
        char c = s[p];
        if (c == '{') {
            ++p;
            v.type = Json::Object;
            ws();
            if (p < s.size() && s[p] == '}') { ++p; return true; }
            while (true) {
                ws();
                std::string key;
                if (!string(key)) return false;
                ws();
                if (p >= s.size() || s[p] != ':') return fail("expected ':'");
                ++p;
                Json child;
                if (!value(child, depth + 1)) return false;
                v.obj[key] = std::move(child);
                ws();
                if (p < s.size() && s[p] == ',') { ++p; continue; }
                if (p < s.size() && s[p] == '}') { ++p; return true; }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++p;
            v.type = Json::Array;
            ws();
            if (p < s.size() && s[p] == ']') { ++p; return true; }
            while (true) {
                Json child;
                if (!value(child, depth + 1)) return false;
                v.arr.push_back(std::move(child));
                ws();
                if (p < s.size() && s[p] == ',') { ++p; continue; }
                if (p < s.size() && s[p] == ']') { ++p; return true; }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') { v.type = Json::String; return string(v.str); }
        if (c == 't') { v.type = Json::Bool; v.b = true; return lit("true"); }
        if (c == 'f') { v.type = Json::Bool; v.b = false; return lit("false"); }
        if (c == 'n') { v.type = Json::Null; return lit("null"); }
        return number(v);
    }
};

}  // namespace

bool json_parse(const std::string& text, Json& out, std::string* err) {
    Parser ps(text);
    out = Json();
    bool ok = ps.value(out, 0);
    if (ok) {
        ps.ws();
        if (ps.p != text.size()) ok = ps.fail("trailing characters");
    }
    if (!ok && err) *err = ps.err;
    return ok;
}

}  // namespace ptamd
