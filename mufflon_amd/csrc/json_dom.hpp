// A small JSON document model and strict parser (RFC 8259: no comments, no
// trailing commas) for the scene loader. Objects keep their members in file
// order, and a lookup returns the first member of a name, which is the
// iteration and lookup order the scene format relies on (the first scenario is
// the default one, materials are declared in file order).
#pragma once

#include <charconv>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace mffjson {

struct Value {
    enum Type { Null, Bool, Number, String, Array, Object };
    Type type = Null;
    bool boolean = false;
    double number = 0.0;
    bool integral = false;  // a number written without fraction or exponent
    std::string str;
    std::vector<Value> items;
    std::vector<std::pair<std::string, Value>> members;

    bool is_null() const { return type == Null; }
    bool is_bool() const { return type == Bool; }
    bool is_number() const { return type == Number; }
    bool is_string() const { return type == String; }
    bool is_array() const { return type == Array; }
    bool is_object() const { return type == Object; }
    // a non-negative integer that fits 32 bits
    bool is_uint() const { return type == Number && integral && number >= 0.0 && number <= 4294967295.0; }
    bool is_int() const { return type == Number && integral && number >= -2147483648.0 && number <= 2147483647.0; }

    const Value* find(const char* key) const {
        for (const auto& m : members)
            if (m.first == key) return &m.second;
        return nullptr;
    }
    Value* find(const char* key) {
        for (auto& m : members)
            if (m.first == key) return &m.second;
        return nullptr;
    }
};

struct ParseError : std::runtime_error {
    size_t offset, line;
    ParseError(const std::string& what, size_t offset, size_t line)
        : std::runtime_error(what), offset(offset), line(line) {}
};

class Parser {
public:
    explicit Parser(const std::string& text) : s_(text.c_str()), n_(text.size()) {}

    Value parse_document() {
        ws();
        Value v = value(0);
        ws();
        if (p_ != n_) fail("The document root must not be followed by other values");
        return v;
    }

private:
    static constexpr int kMaxDepth = 512;
    const char* s_;
    size_t n_, p_ = 0;

    [[noreturn]] void fail(const char* what) const {
        size_t line = 0;
        for (size_t i = 0; i < p_ && i < n_; ++i)
            if (s_[i] == '\n') ++line;
        throw ParseError(std::string("Parser error: ") + what + " at offset " + std::to_string(p_) + " (line " +
                             std::to_string(line) + ")",
                         p_, line);
    }
    void ws() {
        while (p_ < n_ && (s_[p_] == ' ' || s_[p_] == '\t' || s_[p_] == '\n' || s_[p_] == '\r')) ++p_;
    }
    bool literal(const char* word) {
        const size_t k = std::strlen(word);
        if (n_ - p_ >= k && std::memcmp(s_ + p_, word, k) == 0) {
            p_ += k;
            return true;
        }
        return false;
    }
    Value value(int depth) {
        if (depth > kMaxDepth) fail("Nesting too deep");
        if (p_ >= n_) fail("The document is empty or ends early");
        Value v;
        const char c = s_[p_];
        if (c == '{') {
            v.type = Value::Object;
            ++p_;
            ws();
            if (p_ < n_ && s_[p_] == '}') {
                ++p_;
                return v;
            }
            for (;;) {
                ws();
                if (p_ >= n_ || s_[p_] != '"') fail("Missing a name for object member");
                std::string key = string();
                ws();
                if (p_ >= n_ || s_[p_] != ':') fail("Missing a colon after a name of object member");
                ++p_;
                ws();
                Value m = value(depth + 1);
                v.members.emplace_back(std::move(key), std::move(m));
                ws();
                if (p_ < n_ && s_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < n_ && s_[p_] == '}') {
                    ++p_;
                    return v;
                }
                fail("Missing a comma or '}' after an object member");
            }
        }
        if (c == '[') {
            v.type = Value::Array;
            ++p_;
            ws();
            if (p_ < n_ && s_[p_] == ']') {
                ++p_;
                return v;
            }
            for (;;) {
                ws();
                v.items.push_back(value(depth + 1));
                ws();
                if (p_ < n_ && s_[p_] == ',') {
                    ++p_;
                    continue;
                }
                if (p_ < n_ && s_[p_] == ']') {
                    ++p_;
                    return v;
                }
                fail("Missing a comma or ']' after an array element");
            }
        }
        if (c == '"') {
            v.type = Value::String;
            v.str = string();
            return v;
        }
        if (literal("true")) {
            v.type = Value::Bool;
            v.boolean = true;
            return v;
        }
        if (literal("false")) {
            v.type = Value::Bool;
            return v;
        }
        if (literal("null")) return v;
        if (c == '-' || (c >= '0' && c <= '9')) return number();
        fail("Invalid value");
    }
    static void put_utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4() {
        if (n_ - p_ < 4) fail("Incorrect hex digit after \\u escape in string");
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) {
            const char h = s_[p_++];
            v <<= 4;
            if (h >= '0' && h <= '9') v |= (uint32_t)(h - '0');
            else if (h >= 'a' && h <= 'f') v |= (uint32_t)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') v |= (uint32_t)(h - 'A' + 10);
            else fail("Incorrect hex digit after \\u escape in string");
        }
        return v;
    }
    std::string string() {
        ++p_;  // opening quote
        std::string out;
        for (;;) {
            if (p_ >= n_) fail("Missing a closing quotation mark in string");
            const unsigned char c = (unsigned char)s_[p_++];
            if (c == '"') return out;
            if (c < 0x20) fail("Invalid encoding in string");
            if (c != '\\') {
                out += (char)c;
                continue;
            }
            if (p_ >= n_) fail("Invalid escape character in string");
            const char e = s_[p_++];
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
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp <= 0xDBFF) {
                        if (n_ - p_ < 2 || s_[p_] != '\\' || s_[p_ + 1] != 'u')
                            fail("The surrogate pair in string is invalid");
                        p_ += 2;
                        const uint32_t lo = hex4();
                        if (lo < 0xDC00 || lo > 0xDFFF) fail("The surrogate pair in string is invalid");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp <= 0xDFFF) {
                        fail("The surrogate pair in string is invalid");
                    }
                    put_utf8(out, cp);
                    break;
                }
                default: fail("Invalid escape character in string");
            }
        }
    }
    Value number() {
        const size_t start = p_;
        bool integral = true;
        if (s_[p_] == '-') ++p_;
        if (p_ >= n_) fail("Missing a digit in number");
        if (s_[p_] == '0') {
            ++p_;
        } else if (s_[p_] >= '1' && s_[p_] <= '9') {
            while (p_ < n_ && s_[p_] >= '0' && s_[p_] <= '9') ++p_;
        } else {
            fail("Missing a digit in number");
        }
        if (p_ < n_ && s_[p_] == '.') {
            integral = false;
            ++p_;
            if (p_ >= n_ || s_[p_] < '0' || s_[p_] > '9') fail("Missing fraction part in number");
            while (p_ < n_ && s_[p_] >= '0' && s_[p_] <= '9') ++p_;
        }
        if (p_ < n_ && (s_[p_] == 'e' || s_[p_] == 'E')) {
            integral = false;
            ++p_;
            if (p_ < n_ && (s_[p_] == '+' || s_[p_] == '-')) ++p_;
            if (p_ >= n_ || s_[p_] < '0' || s_[p_] > '9') fail("Missing exponent in number");
            while (p_ < n_ && s_[p_] >= '0' && s_[p_] <= '9') ++p_;
        }
        Value v;
        v.type = Value::Number;
        v.integral = integral;
        // the nearest double to the decimal text, independent of the locale
        const auto r = std::from_chars(s_ + start, s_ + p_, v.number);
        if (r.ec == std::errc::result_out_of_range) fail("Number too big to be stored in double");
        if (r.ec != std::errc() || r.ptr != s_ + p_) fail("Invalid number");
        return v;
    }
};

inline Value parse(const std::string& text) { return Parser(text).parse_document(); }

}  // namespace mffjson
