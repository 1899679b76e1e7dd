// json.hpp — small JSON DOM for scene files (serde_json-compatible subset:
// objects, arrays, strings with escapes, numbers via correctly-rounded strtod,
// true/false/null). Host glue for the trace path; not on the device path.
#pragma once

#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace rtj {

struct ParseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Value {
    enum Kind { kNull, kBool, kNumber, kString, kArray, kObject };
    Kind kind = kNull;
    bool boolean = false;
    double number = 0.0;
    std::string text;
    std::vector<Value> items;
    std::vector<std::string> keys;  // insertion order
    std::vector<Value> vals;

    const Value* find(const std::string& k) const {
        for (size_t i = 0; i < keys.size(); ++i)
            if (keys[i] == k) return &vals[i];
        return nullptr;
    }
    bool is_null() const { return kind == kNull; }
};

class Reader {
  public:
    Reader(const char* b, size_t n) : cur_(b), end_(b + n), begin_(b) {}

    Value document() {
        Value v = value();
        skip();
        if (cur_ != end_) error("trailing characters");
        return v;
    }

  private:
    const char* cur_;
    const char* end_;
    const char* begin_;

    [[noreturn]] void error(const char* what) {
        size_t line = 1, col = 1;
        for (const char* p = begin_; p < cur_; ++p) {
            if (*p == '\n') { ++line; col = 1; } else { ++col; }
        }
        throw ParseError(std::string(what) + " at line " + std::to_string(line) + " column " + std::to_string(col));
    }
    void skip() {
        while (cur_ < end_ && (*cur_ == ' ' || *cur_ == '\n' || *cur_ == '\r' || *cur_ == '\t')) ++cur_;
    }
    bool lit(const char* w) {
        size_t n = strlen(w);
        if ((size_t)(end_ - cur_) >= n && memcmp(cur_, w, n) == 0) { cur_ += n; return true; }
        return false;
    }
    Value value() {
        skip();
        if (cur_ >= end_) error("EOF while parsing a value");
        Value v;
        switch (*cur_) {
            case '{': object(v); break;
            case '[': array(v); break;
            case '"': v.kind = Value::kString; v.text = string(); break;
            case 't': if (!lit("true")) error("expected value"); v.kind = Value::kBool; v.boolean = true; break;
            case 'f': if (!lit("false")) error("expected value"); v.kind = Value::kBool; break;
            case 'n': if (!lit("null")) error("expected value"); break;
            default: number(v);
        }
        return v;
    }
    void object(Value& v) {
        v.kind = Value::kObject;
        ++cur_;
        skip();
        if (cur_ < end_ && *cur_ == '}') { ++cur_; return; }
        for (;;) {
            skip();
            if (cur_ >= end_ || *cur_ != '"') error("key must be a string");
            std::string k = string();
            skip();
            if (cur_ >= end_ || *cur_ != ':') error("expected `:`");
            ++cur_;
            v.keys.push_back(std::move(k));
            v.vals.push_back(value());
            skip();
            if (cur_ < end_ && *cur_ == ',') { ++cur_; continue; }
            if (cur_ < end_ && *cur_ == '}') { ++cur_; return; }
            error("expected `,` or `}`");
        }
    }
    void array(Value& v) {
        v.kind = Value::kArray;
        ++cur_;
        skip();
        if (cur_ < end_ && *cur_ == ']') { ++cur_; return; }
        for (;;) {
            v.items.push_back(value());
            skip();
            if (cur_ < end_ && *cur_ == ',') { ++cur_; continue; }
            if (cur_ < end_ && *cur_ == ']') { ++cur_; return; }
            error("expected `,` or `]`");
        }
    }
    void number(Value& v) {
        const char* s = cur_;
        if (cur_ < end_ && *cur_ == '-') ++cur_;
        if (cur_ >= end_ || !(*cur_ >= '0' && *cur_ <= '9')) error("expected value");
        while (cur_ < end_ && ((*cur_ >= '0' && *cur_ <= '9') || *cur_ == '.' || *cur_ == 'e' || *cur_ == 'E' ||
                               *cur_ == '+' || *cur_ == '-'))
            ++cur_;
        std::string tok(s, cur_);
        char* e = nullptr;
        v.number = strtod(tok.c_str(), &e);
        if (!e || *e) error("invalid number");
        v.kind = Value::kNumber;
    }
    static void utf8(std::string& out, unsigned cp) {
        if (cp < 0x80) out += (char)cp;
        else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
        else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18)); out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F));
        }
    }
    unsigned hex4() {
        if (end_ - cur_ < 4) error("invalid escape");
        unsigned cp = 0;
        for (int i = 0; i < 4; ++i) {
            char c = *cur_++;
            cp <<= 4;
            if (c >= '0' && c <= '9') cp |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') cp |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') cp |= (unsigned)(c - 'A' + 10);
            else error("invalid escape");
        }
        return cp;
    }
    std::string string() {
        ++cur_;
        std::string out;
        for (;;) {
            if (cur_ >= end_) error("EOF while parsing a string");
            char c = *cur_++;
            if (c == '"') return out;
            if (c != '\\') { out += c; continue; }
            if (cur_ >= end_) error("EOF while parsing a string");
            char e = *cur_++;
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
                    unsigned cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && end_ - cur_ >= 6 && cur_[0] == '\\' && cur_[1] == 'u') {
                        cur_ += 2;
                        unsigned lo = hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    utf8(out, cp);
                    break;
                }
                default: error("invalid escape");
            }
        }
    }
};

inline Value parse(const std::string& s) { return Reader(s.data(), s.size()).document(); }

}  // namespace rtj
