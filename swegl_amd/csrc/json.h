// json.h -- a small JSON document reader for glTF files (RFC 8259 grammar:
// objects, arrays, strings with escapes, numbers, true/false/null).  Numbers
// are kept as double, which is how a JSON library hands a non-integer
// number to get<float>() (double, then rounded to float); integers below
// 2^53 are exact in double, so reading them as float rounds once either way.
#pragma once
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

namespace sj {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object };
    Kind kind = Null;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;

    // member of an object, or nullptr when absent (or not an object)
    const Value *get(const char *key) const
    {
        if (kind != Object)
            return nullptr;
        for (const auto &kv : obj)
            if (kv.first == key)
                return &kv.second;
        return nullptr;
    }
    bool has(const char *key) const { return get(key) != nullptr; }
    size_t size() const { return kind == Array ? arr.size() : kind == Object ? obj.size() : 0; }
    bool is_num() const { return kind == Number; }
    bool is_str() const { return kind == String; }
    bool is_arr() const { return kind == Array; }
    bool is_obj() const { return kind == Object; }
};

class Parser
{
public:
    Parser(const char *b, const char *e) : p_(b), e_(e) {}

    bool parse(Value &out, std::string &err)
    {
        ws();
        if (!value(out, 0)) {
            err = err_.empty() ? "malformed JSON" : err_;
            err += " at byte " + std::to_string(off());
            return false;
        }
        ws();
        if (p_ != e_) {   // a GLB JSON chunk is padded with spaces, which ws() took
            err = "trailing bytes after the JSON document at byte " + std::to_string(off());
            return false;
        }
        return true;
    }

private:
    const char *p_, *e_, *b0_ = p_;
    std::string err_;

    size_t off() const { return (size_t)(p_ - b0_); }
    void ws()
    {
        while (p_ < e_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r'))
            p_++;
    }
    bool lit(const char *s)
    {
        size_t n = strlen(s);
        if ((size_t)(e_ - p_) < n || memcmp(p_, s, n) != 0)
            return false;
        p_ += n;
        return true;
    }
    bool value(Value &v, int depth)
    {
        if (depth > 256) {
            err_ = "JSON nested too deeply";
            return false;
        }
        if (p_ >= e_)
            return false;
        switch (*p_) {
        case '{': return object(v, depth);
        case '[': return array(v, depth);
        case '"': v.kind = Value::String; return string(v.str);
        case 't': v.kind = Value::Bool; v.b = true; return lit("true");
        case 'f': v.kind = Value::Bool; v.b = false; return lit("false");
        case 'n': v.kind = Value::Null; return lit("null");
        default: v.kind = Value::Number; return number(v.num);
        }
    }
    bool object(Value &v, int depth)
    {
        v.kind = Value::Object;
        p_++;
        ws();
        if (p_ < e_ && *p_ == '}') {
            p_++;
            return true;
        }
        for (;;) {
            ws();
            std::string key;
            if (p_ >= e_ || *p_ != '"' || !string(key))
                return false;
            ws();
            if (p_ >= e_ || *p_ != ':')
                return false;
            p_++;
            ws();
            v.obj.emplace_back(std::move(key), Value{});
            if (!value(v.obj.back().second, depth + 1))
                return false;
            ws();
            if (p_ < e_ && *p_ == ',') {
                p_++;
                continue;
            }
            if (p_ < e_ && *p_ == '}') {
                p_++;
                return true;
            }
            return false;
        }
    }
    bool array(Value &v, int depth)
    {
        v.kind = Value::Array;
        p_++;
        ws();
        if (p_ < e_ && *p_ == ']') {
            p_++;
            return true;
        }
        for (;;) {
            ws();
            v.arr.emplace_back();
            if (!value(v.arr.back(), depth + 1))
                return false;
            ws();
            if (p_ < e_ && *p_ == ',') {
                p_++;
                continue;
            }
            if (p_ < e_ && *p_ == ']') {
                p_++;
                return true;
            }
            return false;
        }
    }
    static void utf8(std::string &s, unsigned cp)
    {
        if (cp < 0x80) {
            s += (char)cp;
        } else if (cp < 0x800) {
            s += (char)(0xC0 | (cp >> 6));
            s += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            s += (char)(0xE0 | (cp >> 12));
            s += (char)(0x80 | ((cp >> 6) & 0x3F));
            s += (char)(0x80 | (cp & 0x3F));
        } else {
            s += (char)(0xF0 | (cp >> 18));
            s += (char)(0x80 | ((cp >> 12) & 0x3F));
            s += (char)(0x80 | ((cp >> 6) & 0x3F));
            s += (char)(0x80 | (cp & 0x3F));
        }
    }
    bool hex4(unsigned &cp)
    {
        if (e_ - p_ < 4)
            return false;
        cp = 0;
        for (int i = 0; i < 4; i++) {
            char c = *p_++;
            cp <<= 4;
            if (c >= '0' && c <= '9')
                cp |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f')
                cp |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F')
                cp |= (unsigned)(c - 'A' + 10);
            else
                return false;
        }
        return true;
    }
    bool string(std::string &s)
    {
        p_++;   // opening quote
        while (p_ < e_) {
            char c = *p_++;
            if (c == '"')
                return true;
            if ((unsigned char)c < 0x20)
                return false;
            if (c != '\\') {
                s += c;
                continue;
            }
            if (p_ >= e_)
                return false;
            char x = *p_++;
            switch (x) {
            case '"': s += '"'; break;
            case '\\': s += '\\'; break;
            case '/': s += '/'; break;
            case 'b': s += '\b'; break;
            case 'f': s += '\f'; break;
            case 'n': s += '\n'; break;
            case 'r': s += '\r'; break;
            case 't': s += '\t'; break;
            case 'u': {
                unsigned cp;
                if (!hex4(cp))
                    return false;
                if (cp >= 0xD800 && cp < 0xDC00) {   // surrogate pair
                    unsigned lo;
                    if (e_ - p_ < 6 || p_[0] != '\\' || p_[1] != 'u')
                        return false;
                    p_ += 2;
                    if (!hex4(lo) || lo < 0xDC00 || lo >= 0xE000)
                        return false;
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                } else if (cp >= 0xDC00 && cp < 0xE000) {
                    return false;
                }
                utf8(s, cp);
                break;
            }
            default: return false;
            }
        }
        return false;
    }
    bool number(double &d)
    {
        // validate the JSON number grammar, then convert with strtod
        const char *s = p_;
        const char *q = p_;
        if (q < e_ && *q == '-')
            q++;
        if (q >= e_)
            return false;
        if (*q == '0') {
            q++;
        } else if (*q >= '1' && *q <= '9') {
            while (q < e_ && *q >= '0' && *q <= '9')
                q++;
        } else {
            return false;
        }
        if (q < e_ && *q == '.') {
            q++;
            if (q >= e_ || !(*q >= '0' && *q <= '9'))
                return false;
            while (q < e_ && *q >= '0' && *q <= '9')
                q++;
        }
        if (q < e_ && (*q == 'e' || *q == 'E')) {
            q++;
            if (q < e_ && (*q == '+' || *q == '-'))
                q++;
            if (q >= e_ || !(*q >= '0' && *q <= '9'))
                return false;
            while (q < e_ && *q >= '0' && *q <= '9')
                q++;
        }
        std::string tmp(s, q);
        d = strtod(tmp.c_str(), nullptr);
        p_ = q;
        return true;
    }
};

inline bool parse(const char *b, const char *e, Value &out, std::string &err)
{
    Parser P(b, e);
    return P.parse(out, err);
}

}  // namespace sj
