// Json.cpp -- recursive-descent JSON reader (RFC 8259 subset: no \u
// surrogate pairs beyond the BMP) and a compact writer.
#include "Json.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace sph {

class JsonReader {
public:
    explicit JsonReader(const std::string& s) : s_(s) {}

    Json document() {
        Json v = value();
        ws();
        if (p_ != s_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& s_;
    size_t p_ = 0;

    [[noreturn]] void fail(const std::string& msg) {
        size_t line = 1, col = 1;
        for (size_t i = 0; i < p_ && i < s_.size(); ++i) {
            if (s_[i] == '\n') { ++line; col = 1; } else ++col;
        }
        throw JsonError("JSON parse error at line " + std::to_string(line) + ", column " + std::to_string(col) +
                        ": " + msg);
    }
    void ws() {
        while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\n' || s_[p_] == '\r')) ++p_;
    }
    bool lit(const char* w) {
        size_t n = std::char_traits<char>::length(w);
        if (s_.compare(p_, n, w) == 0) { p_ += n; return true; }
        return false;
    }
    Json value() {
        ws();
        if (p_ >= s_.size()) fail("unexpected end of input");
        char c = s_[p_];
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') return Json::string(string());
        if (lit("true")) return Json::boolean(true);
        if (lit("false")) return Json::boolean(false);
        if (lit("null")) return Json();
        if (c == '-' || (c >= '0' && c <= '9')) return number();
        fail(std::string("unexpected character '") + c + "'");
    }
    Json number() {
        const char* b = s_.c_str() + p_;
        char* e = nullptr;
        double v = std::strtod(b, &e);
        if (e == b) fail("bad number");
        p_ += (size_t)(e - b);
        return Json::number(v);
    }
    std::string string() {
        ++p_;   // opening quote
        std::string out;
        while (true) {
            if (p_ >= s_.size()) fail("unterminated string");
            char c = s_[p_++];
            if (c == '"') break;
            if (c != '\\') { out += c; continue; }
            if (p_ >= s_.size()) fail("bad escape");
            char e = s_[p_++];
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
                    if (p_ + 4 > s_.size()) fail("bad \\u escape");
                    unsigned cp = (unsigned)std::strtoul(s_.substr(p_, 4).c_str(), nullptr, 16);
                    p_ += 4;
                    if (cp < 0x80) out += (char)cp;
                    else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
                    else {
                        out += (char)(0xE0 | (cp >> 12));
                        out += (char)(0x80 | ((cp >> 6) & 0x3F));
                        out += (char)(0x80 | (cp & 0x3F));
                    }
                    break;
                }
                default: fail("bad escape");
            }
        }
        return out;
    }
    Json array() {
        ++p_;
        Json a = Json::array();
        ws();
        if (p_ < s_.size() && s_[p_] == ']') { ++p_; return a; }
        while (true) {
            a.arr_.push_back(value());
            ws();
            if (p_ >= s_.size()) fail("unterminated array");
            if (s_[p_] == ',') { ++p_; continue; }
            if (s_[p_] == ']') { ++p_; return a; }
            fail("expected ',' or ']'");
        }
    }
    Json object() {
        ++p_;
        Json o = Json::object();
        ws();
        if (p_ < s_.size() && s_[p_] == '}') { ++p_; return o; }
        while (true) {
            ws();
            if (p_ >= s_.size() || s_[p_] != '"') fail("expected a key");
            std::string k = string();
            ws();
            if (p_ >= s_.size() || s_[p_] != ':') fail("expected ':'");
            ++p_;
            Json v = value();
            bool replaced = false;
            for (auto& kv : o.obj_)
                if (kv.first == k) { kv.second = v; replaced = true; }
            if (!replaced) o.obj_.emplace_back(k, v);
            ws();
            if (p_ >= s_.size()) fail("unterminated object");
            if (s_[p_] == ',') { ++p_; continue; }
            if (s_[p_] == '}') { ++p_; return o; }
            fail("expected ',' or '}'");
        }
    }
};

Json Json::parse(const std::string& text) { return JsonReader(text).document(); }

Json Json::parseFile(const std::string& path) {
    std::ifstream f(path);
    if (!f.is_open()) throw JsonError("cannot read " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

static const char* typeName(Json::Type t) {
    switch (t) {
        case Json::Null: return "null";
        case Json::Bool: return "boolean";
        case Json::Number: return "number";
        case Json::String: return "string";
        case Json::Array: return "array";
        default: return "object";
    }
}

double Json::num(const char* what) const {
    if (type_ == Number) return num_;
    if (type_ == Bool) return b_ ? 1.0 : 0.0;
    throw JsonError(std::string(what) + ": expected a number, found " + typeName(type_));
}
int Json::integer(const char* what) const {
    double v = num(what);
    if (v != std::floor(v)) throw JsonError(std::string(what) + ": expected an integer");
    return (int)v;
}
bool Json::flag(const char* what) const {
    if (type_ == Bool) return b_;
    if (type_ == Number) return num_ != 0.0;
    throw JsonError(std::string(what) + ": expected a boolean, found " + typeName(type_));
}
const std::string& Json::str(const char* what) const {
    if (type_ != String) throw JsonError(std::string(what) + ": expected a string, found " + typeName(type_));
    return str_;
}
bool Json::has(const std::string& key) const {
    if (type_ != Object) return false;
    for (auto& kv : obj_)
        if (kv.first == key) return true;
    return false;
}
const Json& Json::at(const std::string& key) const {
    if (type_ != Object) throw JsonError("'" + key + "': parent is not an object");
    for (auto& kv : obj_)
        if (kv.first == key) return kv.second;
    throw JsonError("missing key '" + key + "'");
}
const Json& Json::get(const std::string& key) const {
    static const Json null;
    if (type_ != Object) return null;
    for (auto& kv : obj_)
        if (kv.first == key) return kv.second;
    return null;
}
Json& Json::operator[](const std::string& key) {
    if (type_ == Null) type_ = Object;
    if (type_ != Object) throw JsonError("'" + key + "': not an object");
    for (auto& kv : obj_)
        if (kv.first == key) return kv.second;
    obj_.emplace_back(key, Json());
    return obj_.back().second;
}
size_t Json::size() const {
    if (type_ == Array) return arr_.size();
    if (type_ == Object) return obj_.size();
    return 0;   // null has no elements (a missing list reads as empty)
}
const Json& Json::at(size_t i) const {
    if (type_ != Array) throw JsonError("index into a non-array");
    if (i >= arr_.size()) throw JsonError("array index " + std::to_string(i) + " out of range");
    return arr_[i];
}
void Json::push(const Json& v) {
    if (type_ == Null) type_ = Array;
    if (type_ != Array) throw JsonError("push onto a non-array");
    arr_.push_back(v);
}

void Json::dumpTo(std::string& out) const {
    switch (type_) {
        case Null: out += "null"; break;
        case Bool: out += b_ ? "true" : "false"; break;
        case Number: {
            if (std::isfinite(num_)) {
                char buf[32];
                std::snprintf(buf, sizeof buf, "%.17g", num_);
                out += buf;
            } else {
                out += "null";
            }
            break;
        }
        case String: {
            out += '"';
            for (char c : str_) {
                if (c == '"' || c == '\\') { out += '\\'; out += c; }
                else if (c == '\n') out += "\\n";
                else out += c;
            }
            out += '"';
            break;
        }
        case Array:
            out += '[';
            for (size_t i = 0; i < arr_.size(); ++i) {
                if (i) out += ',';
                arr_[i].dumpTo(out);
            }
            out += ']';
            break;
        case Object:
            out += '{';
            for (size_t i = 0; i < obj_.size(); ++i) {
                if (i) out += ',';
                out += '"';
                out += obj_[i].first;
                out += "\":";
                obj_[i].second.dumpTo(out);
            }
            out += '}';
            break;
    }
}
std::string Json::dump() const {
    std::string out;
    dumpTo(out);
    return out;
}

}  // namespace sph
