// Minimal JSON value, parser and writer for the flat MDConfig documents.
//
// The writer reproduces the layout nlohmann::json::dump(indent) gives the
// reference's config files (src/md_config.cpp:315): keys in sorted order,
// "key": value pairs, integers without a fraction, doubles in shortest
// round-trip form with ".0" appended to integral values and exponents written
// as e+XX / e-XX, exponent form outside 1e-5 < |x| < 1e15.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace mdx_json {

class Value {
public:
    enum class Type { Null, Bool, Int, Double, String, Object, Array };

    Value() = default;
    static Value from_int(int64_t v) { Value x; x.t_ = Type::Int; x.i_ = v; return x; }
    static Value from_double(double v) { Value x; x.t_ = Type::Double; x.d_ = v; return x; }
    static Value from_bool(bool v) { Value x; x.t_ = Type::Bool; x.b_ = v; return x; }
    static Value from_string(std::string v) { Value x; x.t_ = Type::String; x.s_ = std::move(v); return x; }
    static Value object() { Value x; x.t_ = Type::Object; return x; }
    static Value array() { Value x; x.t_ = Type::Array; return x; }

    Type type() const { return t_; }
    bool is_number() const { return t_ == Type::Int || t_ == Type::Double; }

    // nlohmann-style numeric conversion (integers truncate doubles)
    double as_double() const;
    int64_t as_int() const;
    bool as_bool() const;
    const std::string &as_string() const;

    bool contains(const std::string &k) const { return t_ == Type::Object && obj_.count(k) != 0; }
    const Value &at(const std::string &k) const;
    Value &operator[](const std::string &k);
    const std::map<std::string, Value> &items() const { return obj_; }
    std::vector<Value> &elements() { return arr_; }

    std::string dump(int indent = -1) const;

private:
    void dump_to(std::string &out, int indent, int depth) const;
    Type t_ = Type::Null;
    bool b_ = false;
    int64_t i_ = 0;
    double d_ = 0.0;
    std::string s_;
    std::map<std::string, Value> obj_;
    std::vector<Value> arr_;
};

Value parse(const std::string &text);
std::string format_double(double v);   // nlohmann number formatting

}  // namespace mdx_json
