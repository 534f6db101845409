// fs_json.hpp -- the small JSON reader and writer behind the checkpoint metadata.
//
// A checkpoint's metadata is a JSON object (FluidSim/FluidCheckpoint.h:9-15 keeps it as
// bl::JSON::json); the library writes and reads that text itself so that the C ABI stays free of
// any JSON type.  Numbers are written with 17 significant digits and read with strtod, so every
// finite double survives the text round trip bit for bit.  Non-finite numbers are written as null.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace fsjson {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object };
    Kind kind = Null;
    bool boolean = false;
    double number = 0.0;
    std::string text;
    std::vector<Value> items;
    std::vector<std::pair<std::string, Value>> members;

    const Value* find(const std::string& key) const
    {
        for (const auto& m : members)
            if (m.first == key)
                return &m.second;
        return nullptr;
    }
};

class Reader {
public:
    Reader(const char* text, size_t length) : s_(text), end_(text + length) {}

    Value document()
    {
        Value v = value(0);
        space();
        if (s_ != end_)
            error("trailing characters after the JSON value");
        return v;
    }

private:
    const char* s_;
    const char* end_;

    [[noreturn]] static void error(const char* what) { throw std::runtime_error(std::string("Checkpoint metadata: ") + what); }

    void space()
    {
        while (s_ != end_ && (*s_ == ' ' || *s_ == '\t' || *s_ == '\n' || *s_ == '\r'))
            ++s_;
    }

    bool literal(const char* word)
    {
        const char* p = s_;
        for (const char* w = word; *w; ++w, ++p)
            if (p == end_ || *p != *w)
                return false;
        s_ = p;
        return true;
    }

    Value value(int depth)
    {
        if (depth > 32)
            error("nesting too deep");
        space();
        if (s_ == end_)
            error("unexpected end of text");
        Value v;
        const char c = *s_;
        if (c == '{') {
            v.kind = Value::Object;
            ++s_;
            space();
            if (s_ != end_ && *s_ == '}') {
                ++s_;
                return v;
            }
            for (;;) {
                space();
                if (s_ == end_ || *s_ != '"')
                    error("expected a member name");
                std::string key = string();
                space();
                if (s_ == end_ || *s_ != ':')
                    error("expected ':'");
                ++s_;
                for (const auto& m : v.members)
                    if (m.first == key)
                        error("duplicate member name");
                v.members.emplace_back(std::move(key), value(depth + 1));
                space();
                if (s_ != end_ && *s_ == ',') {
                    ++s_;
                    continue;
                }
                if (s_ != end_ && *s_ == '}') {
                    ++s_;
                    return v;
                }
                error("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = Value::Array;
            ++s_;
            space();
            if (s_ != end_ && *s_ == ']') {
                ++s_;
                return v;
            }
            for (;;) {
                v.items.push_back(value(depth + 1));
                space();
                if (s_ != end_ && *s_ == ',') {
                    ++s_;
                    continue;
                }
                if (s_ != end_ && *s_ == ']') {
                    ++s_;
                    return v;
                }
                error("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = Value::String;
            v.text = string();
            return v;
        }
        if (literal("true")) {
            v.kind = Value::Bool;
            v.boolean = true;
            return v;
        }
        if (literal("false")) {
            v.kind = Value::Bool;
            return v;
        }
        if (literal("null"))
            return v;
        return number();
    }

    std::string string()
    {
        ++s_;  // opening quote
        std::string out;
        while (s_ != end_ && *s_ != '"') {
            char c = *s_++;
            if ((unsigned char)c < 0x20)
                error("control character in a string");
            if (c == '\\') {
                if (s_ == end_)
                    error("unterminated escape");
                const char e = *s_++;
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
                    if (end_ - s_ < 4)
                        error("short \\u escape");
                    unsigned code = 0;
                    for (int i = 0; i < 4; ++i) {
                        const char h = *s_++;
                        code <<= 4;
                        if (h >= '0' && h <= '9')
                            code |= unsigned(h - '0');
                        else if (h >= 'a' && h <= 'f')
                            code |= unsigned(h - 'a' + 10);
                        else if (h >= 'A' && h <= 'F')
                            code |= unsigned(h - 'A' + 10);
                        else
                            error("bad \\u escape");
                    }
                    // Names the library reads are ASCII; other code points are kept as UTF-8.
                    if (code < 0x80) {
                        out += char(code);
                    } else if (code < 0x800) {
                        out += char(0xC0 | (code >> 6));
                        out += char(0x80 | (code & 0x3F));
                    } else {
                        out += char(0xE0 | (code >> 12));
                        out += char(0x80 | ((code >> 6) & 0x3F));
                        out += char(0x80 | (code & 0x3F));
                    }
                    break;
                }
                default: error("unknown escape");
                }
            } else {
                out += c;
            }
        }
        if (s_ == end_)
            error("unterminated string");
        ++s_;
        return out;
    }

    Value number()
    {
        // JSON grammar: -?(0|[1-9][0-9]*)(\.[0-9]+)?([eE][+-]?[0-9]+)?
        const char* p = s_;
        auto digits = [&] {
            const char* q = p;
            while (p != end_ && *p >= '0' && *p <= '9')
                ++p;
            return p != q;
        };
        if (p != end_ && *p == '-')
            ++p;
        if (p != end_ && *p == '0')
            ++p;
        else if (!digits())
            error("malformed value");
        if (p != end_ && *p == '.') {
            ++p;
            if (!digits())
                error("malformed number");
        }
        if (p != end_ && (*p == 'e' || *p == 'E')) {
            ++p;
            if (p != end_ && (*p == '+' || *p == '-'))
                ++p;
            if (!digits())
                error("malformed number");
        }
        const std::string token(s_, p);
        s_ = p;
        Value v;
        v.kind = Value::Number;
        v.number = std::strtod(token.c_str(), nullptr);
        return v;
    }
};

inline Value parse(const char* text, size_t length) { return Reader(text, length).document(); }

// Writer helpers: append JSON tokens to a string.
inline void put_number(std::string& out, double x)
{
    if (!std::isfinite(x)) {
        out += "null";
        return;
    }
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", x);
    out += buf;
}

inline void put_int(std::string& out, long long x) { out += std::to_string(x); }

inline void put_string(std::string& out, const std::string& s)
{
    out += '"';
    for (char c : s) {
        if (c == '"' || c == '\\') {
            out += '\\';
            out += c;
        } else if ((unsigned char)c < 0x20) {
            char buf[8];
            std::snprintf(buf, sizeof buf, "\\u%04x", (unsigned)(unsigned char)c);
            out += buf;
        } else {
            out += c;
        }
    }
    out += '"';
}

inline void put_key(std::string& out, const char* key, bool& first)
{
    if (!first)
        out += ',';
    first = false;
    put_string(out, key);
    out += ':';
}

}  // namespace fsjson
