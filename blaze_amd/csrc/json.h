// json.h — minimal JSON reader for .blz scene files (objects, arrays,
// strings, numbers, true/false/null). Numbers are read with strtod, which
// rounds correctly to double.
#pragma once

#include <cstdlib>
#include <string>
#include <utility>
#include <vector>

namespace blz {
namespace json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;

    bool is_number() const { return kind == Number; }
    bool is_string() const { return kind == String; }
    bool is_array() const { return kind == Array; }
    bool is_object() const { return kind == Object; }
    const Value* find(const char* key) const {
        if (kind != Object) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

class Parser {
  public:
    explicit Parser(const std::string& t) : s_(t) {}
    bool parse(Value& out, std::string& err) {
        ws();
        if (!value(out, 0)) {
            err = err_.empty() ? "syntax error" : err_;
            err += " at offset " + std::to_string(p_);
            return false;
        }
        ws();
        if (p_ != s_.size()) {
            err = "trailing characters at offset " + std::to_string(p_);
            return false;
        }
        return true;
    }

  private:
    const std::string& s_;
    size_t p_ = 0;
    std::string err_;

    void ws() {
        while (p_ < s_.size() && (s_[p_] == ' ' || s_[p_] == '\n' || s_[p_] == '\r' || s_[p_] == '\t')) ++p_;
    }
    bool lit(const char* w) {
        size_t n = 0;
        while (w[n]) ++n;
        if (s_.compare(p_, n, w) != 0) return false;
        p_ += n;
        return true;
    }
    bool value(Value& v, int depth) {
        if (depth > 64) {
            err_ = "nesting too deep";
            return false;
        }
        if (p_ >= s_.size()) return false;
        const char c = s_[p_];
        if (c == '{') return object(v, depth);
        if (c == '[') return array(v, depth);
        if (c == '"') {
            v.kind = Value::String;
            return string(v.str);
        }
        if (c == 't') { v.kind = Value::Bool; v.b = true; return lit("true"); }
        if (c == 'f') { v.kind = Value::Bool; v.b = false; return lit("false"); }
        if (c == 'n') { v.kind = Value::Null; return lit("null"); }
        return number(v);
    }
    bool number(Value& v) {
        const char* b = s_.c_str() + p_;
        char* e = nullptr;
        const double d = std::strtod(b, &e);
        if (e == b) {
            err_ = "expected a value";
            return false;
        }
        p_ += (size_t)(e - b);
        v.kind = Value::Number;
        v.num = d;
        return true;
    }
    bool string(std::string& out) {
        ++p_; // opening quote
        out.clear();
        while (p_ < s_.size()) {
            const char c = s_[p_++];
            if (c == '"') return true;
            if (c == '\\') {
                if (p_ >= s_.size()) return false;
                const char e = s_[p_++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (p_ + 4 > s_.size()) return false;
                    const unsigned cp = (unsigned)std::strtoul(s_.substr(p_, 4).c_str(), nullptr, 16);
                    p_ += 4;
                    if (cp < 0x80) out += (char)cp;
                    else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
                    else { out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F)); }
                    break;
                }
                default: out += e; break;
                }
            } else {
                out += c;
            }
        }
        err_ = "unterminated string";
        return false;
    }
    bool array(Value& v, int depth) {
        ++p_;
        v.kind = Value::Array;
        ws();
        if (p_ < s_.size() && s_[p_] == ']') { ++p_; return true; }
        for (;;) {
            Value e;
            ws();
            if (!value(e, depth + 1)) return false;
            v.arr.push_back(std::move(e));
            ws();
            if (p_ >= s_.size()) return false;
            if (s_[p_] == ',') { ++p_; continue; }
            if (s_[p_] == ']') { ++p_; return true; }
            err_ = "expected , or ]";
            return false;
        }
    }
    bool object(Value& v, int depth) {
        ++p_;
        v.kind = Value::Object;
        ws();
        if (p_ < s_.size() && s_[p_] == '}') { ++p_; return true; }
        for (;;) {
            ws();
            if (p_ >= s_.size() || s_[p_] != '"') { err_ = "expected a key"; return false; }
            std::string key;
            if (!string(key)) return false;
            ws();
            if (p_ >= s_.size() || s_[p_] != ':') { err_ = "expected :"; return false; }
            ++p_;
            ws();
            Value e;
            if (!value(e, depth + 1)) return false;
            v.obj.emplace_back(std::move(key), std::move(e));
            ws();
            if (p_ >= s_.size()) return false;
            if (s_[p_] == ',') { ++p_; continue; }
            if (s_[p_] == '}') { ++p_; return true; }
            err_ = "expected , or }";
            return false;
        }
    }
};

inline bool parse(const std::string& text, Value& out, std::string& err) {
    Parser p(text);
    return p.parse(out, err);
}

} // namespace json
} // namespace blz
