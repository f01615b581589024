#include "json_lite.hpp"

#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace mdx_json {

double Value::as_double() const {
    if (t_ == Type::Double) return d_;
    if (t_ == Type::Int) return static_cast<double>(i_);
    if (t_ == Type::Bool) return b_ ? 1.0 : 0.0;
    throw std::runtime_error("[json.exception.type_error.302] type must be number");
}

int64_t Value::as_int() const {
    if (t_ == Type::Int) return i_;
    if (t_ == Type::Double) return static_cast<int64_t>(d_);
    if (t_ == Type::Bool) return b_ ? 1 : 0;
    throw std::runtime_error("[json.exception.type_error.302] type must be number");
}

bool Value::as_bool() const {
    if (t_ == Type::Bool) return b_;
    throw std::runtime_error("[json.exception.type_error.302] type must be boolean");
}

const std::string &Value::as_string() const {
    if (t_ == Type::String) return s_;
    throw std::runtime_error("[json.exception.type_error.302] type must be string");
}

const Value &Value::at(const std::string &k) const {
    auto it = obj_.find(k);
    if (t_ != Type::Object || it == obj_.end()) throw std::runtime_error("[json.exception.out_of_range.403] key '" + k + "' not found");
    return it->second;
}

Value &Value::operator[](const std::string &k) {
    if (t_ == Type::Null) t_ = Type::Object;
    if (t_ != Type::Object) throw std::runtime_error("[json.exception.type_error.305] cannot use operator[] with a string argument");
    return obj_[k];
}

// Shortest round-trip digits, laid out as nlohmann's dtoa format_buffer does:
// plain decimal for exponents in (-5, 15], exponent form otherwise, ".0" for
// integral plain values, at least two exponent digits.
std::string format_double(double v) {
    if (!std::isfinite(v)) return "null";
    if (v == 0.0) return std::signbit(v) ? "-0.0" : "0.0";
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof(buf), v, std::chars_format::scientific);
    std::string sci(buf, r.ptr);
    // sci = [-]d[.ddd]e[+-]XX
    bool neg = sci[0] == '-';
    std::string body = neg ? sci.substr(1) : sci;
    const size_t epos = body.find('e');
    std::string mant = body.substr(0, epos);
    const int e10 = std::stoi(body.substr(epos + 1));
    std::string digits;
    for (char c : mant)
        if (c != '.') digits.push_back(c);
    const int k = static_cast<int>(digits.size());
    const int n = e10 + 1;   // decimal point position relative to digits
    std::string out;
    if (k <= n && n <= 15) {
        out = digits + std::string(n - k, '0') + ".0";
    } else if (0 < n && n <= 15) {
        out = digits.substr(0, n) + "." + digits.substr(n);
    } else if (-4 < n && n <= 0) {
        out = "0." + std::string(-n, '0') + digits;
    } else {
        out = digits.substr(0, 1);
        if (k > 1) out += "." + digits.substr(1);
        const int ee = n - 1;
        char eb[16];
        std::snprintf(eb, sizeof(eb), "e%c%02d", ee < 0 ? '-' : '+', ee < 0 ? -ee : ee);
        out += eb;
    }
    return neg ? "-" + out : out;
}

namespace {
void escape(std::string &out, const std::string &s) {
    out.push_back('"');
    for (unsigned char c : s) {
        switch (c) {
        case '"': out += "\\\""; break;
        case '\\': out += "\\\\"; break;
        case '\n': out += "\\n"; break;
        case '\r': out += "\\r"; break;
        case '\t': out += "\\t"; break;
        case '\b': out += "\\b"; break;
        case '\f': out += "\\f"; break;
        default:
            if (c < 0x20) {
                char b[8];
                std::snprintf(b, sizeof(b), "\\u%04x", c);
                out += b;
            } else {
                out.push_back(static_cast<char>(c));
            }
        }
    }
    out.push_back('"');
}
}  // namespace

void Value::dump_to(std::string &out, int indent, int depth) const {
    switch (t_) {
    case Type::Null: out += "null"; break;
    case Type::Bool: out += b_ ? "true" : "false"; break;
    case Type::Int: out += std::to_string(i_); break;
    case Type::Double: out += format_double(d_); break;
    case Type::String: escape(out, s_); break;
    case Type::Array: {
        if (arr_.empty()) { out += "[]"; break; }
        out += "[";
        for (size_t i = 0; i < arr_.size(); ++i) {
            if (indent >= 0) { out += "\n"; out += std::string((depth + 1) * indent, ' '); }
            arr_[i].dump_to(out, indent, depth + 1);
            if (i + 1 < arr_.size()) out += ",";
        }
        if (indent >= 0) { out += "\n"; out += std::string(depth * indent, ' '); }
        out += "]";
        break;
    }
    case Type::Object: {
        if (obj_.empty()) { out += "{}"; break; }
        out += "{";
        size_t i = 0;
        for (const auto &kv : obj_) {
            if (indent >= 0) { out += "\n"; out += std::string((depth + 1) * indent, ' '); }
            escape(out, kv.first);
            out += indent >= 0 ? ": " : ":";
            kv.second.dump_to(out, indent, depth + 1);
            if (++i < obj_.size()) out += ",";
        }
        if (indent >= 0) { out += "\n"; out += std::string(depth * indent, ' '); }
        out += "}";
        break;
    }
    }
}

std::string Value::dump(int indent) const {
    std::string out;
    dump_to(out, indent, 0);
    return out;
}

namespace {
struct Parser {
    const std::string &s;
    size_t p = 0;
    [[noreturn]] void error(const char *what) {
        throw std::runtime_error(std::string("[json.exception.parse_error.101] parse error at byte ") +
                                 std::to_string(p + 1) + ": " + what);
    }
    void ws() {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\n' || s[p] == '\r' || s[p] == '\t')) ++p;
    }
    Value value() {
        ws();
        if (p >= s.size()) error("unexpected end of input");
        const char c = s[p];
        if (c == '{') return object();
        if (c == '[') return array();
        if (c == '"') return Value::from_string(string());
        if (c == 't' && s.compare(p, 4, "true") == 0) { p += 4; return Value::from_bool(true); }
        if (c == 'f' && s.compare(p, 5, "false") == 0) { p += 5; return Value::from_bool(false); }
        if (c == 'n' && s.compare(p, 4, "null") == 0) { p += 4; return Value(); }
        return number();
    }
    Value object() {
        Value v = Value::object();
        ++p;
        ws();
        if (p < s.size() && s[p] == '}') { ++p; return v; }
        while (true) {
            ws();
            if (p >= s.size() || s[p] != '"') error("expected string key");
            std::string k = string();
            ws();
            if (p >= s.size() || s[p] != ':') error("expected ':'");
            ++p;
            v[k] = value();
            ws();
            if (p < s.size() && s[p] == ',') { ++p; continue; }
            if (p < s.size() && s[p] == '}') { ++p; return v; }
            error("expected ',' or '}'");
        }
    }
    Value array() {
        Value v = Value::array();
        ++p;
        ws();
        if (p < s.size() && s[p] == ']') { ++p; return v; }
        while (true) {
            v.elements().push_back(value());
            ws();
            if (p < s.size() && s[p] == ',') { ++p; continue; }
            if (p < s.size() && s[p] == ']') { ++p; return v; }
            error("expected ',' or ']'");
        }
    }
    std::string string() {
        ++p;
        std::string out;
        while (p < s.size() && s[p] != '"') {
            char c = s[p++];
            if (c == '\\') {
                if (p >= s.size()) error("bad escape");
                char e = s[p++];
                switch (e) {
                case '"': out.push_back('"'); break;
                case '\\': out.push_back('\\'); break;
                case '/': out.push_back('/'); break;
                case 'b': out.push_back('\b'); break;
                case 'f': out.push_back('\f'); break;
                case 'n': out.push_back('\n'); break;
                case 'r': out.push_back('\r'); break;
                case 't': out.push_back('\t'); break;
                case 'u': {
                    if (p + 4 > s.size()) error("bad \\u escape");
                    const unsigned cp = static_cast<unsigned>(std::stoul(s.substr(p, 4), nullptr, 16));
                    p += 4;
                    if (cp < 0x80) out.push_back(static_cast<char>(cp));
                    else if (cp < 0x800) { out.push_back(static_cast<char>(0xC0 | (cp >> 6))); out.push_back(static_cast<char>(0x80 | (cp & 0x3F))); }
                    else { out.push_back(static_cast<char>(0xE0 | (cp >> 12))); out.push_back(static_cast<char>(0x80 | ((cp >> 6) & 0x3F))); out.push_back(static_cast<char>(0x80 | (cp & 0x3F))); }
                    break;
                }
                default: error("bad escape");
                }
            } else {
                out.push_back(c);
            }
        }
        if (p >= s.size()) error("unterminated string");
        ++p;
        return out;
    }
    Value number() {
        const size_t b = p;
        if (p < s.size() && s[p] == '-') ++p;
        bool is_float = false;
        while (p < s.size()) {
            const char c = s[p];
            if (c >= '0' && c <= '9') { ++p; continue; }
            if (c == '.' || c == 'e' || c == 'E' || c == '+' || c == '-') { is_float = true; ++p; continue; }
            break;
        }
        const std::string tok = s.substr(b, p - b);
        if (tok.empty() || tok == "-") error("syntax error while parsing value");
        if (!is_float) {
            int64_t v = 0;
            auto r = std::from_chars(tok.data(), tok.data() + tok.size(), v);
            if (r.ec == std::errc()) return Value::from_int(v);
        }
        double d = 0.0;
        auto r = std::from_chars(tok.data(), tok.data() + tok.size(), d);
        if (r.ec != std::errc() || r.ptr != tok.data() + tok.size()) error("invalid number");
        return Value::from_double(d);
    }
};
}  // namespace

Value parse(const std::string &text) {
    Parser ps{text};
    Value v = ps.value();
    ps.ws();
    if (ps.p != text.size()) ps.error("syntax error: unexpected trailing input");
    return v;
}

}  // namespace mdx_json
