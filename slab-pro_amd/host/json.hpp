// json.hpp -- a small JSON reader for the configuration front end
// (objects, arrays, strings, numbers, true/false/null).  Lookups that fail
// throw std::runtime_error, as the reference's j.at(...) does
// (src/simulation.cpp:118-208).
#pragma once
#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace slabhost {

class Json {
  public:
    enum Kind { Null, Bool, Number, String, Array, Object };

    Json() = default;

    static Json parse(const std::string& text) {
        size_t pos = 0;
        Json v = parse_value(text, pos);
        skip_ws(text, pos);
        if (pos != text.size()) throw std::runtime_error("JSON: trailing characters at offset " + std::to_string(pos));
        return v;
    }

    Kind kind() const { return kind_; }

    const Json& at(const std::string& key) const {
        if (kind_ != Object) throw std::runtime_error("JSON: '" + key + "' looked up in a non-object");
        auto it = obj_.find(key);
        if (it == obj_.end()) throw std::runtime_error("JSON: missing key '" + key + "'");
        return it->second;
    }
    bool contains(const std::string& key) const { return kind_ == Object && obj_.count(key) > 0; }
    const Json& operator[](size_t i) const {
        if (kind_ != Array || i >= arr_.size()) throw std::runtime_error("JSON: array index out of range");
        return arr_[i];
    }
    size_t size() const { return kind_ == Array ? arr_.size() : obj_.size(); }

    double num() const {
        if (kind_ != Number) throw std::runtime_error("JSON: number expected");
        return num_;
    }
    const std::string& str() const {
        if (kind_ != String) throw std::runtime_error("JSON: string expected");
        return str_;
    }
    bool boolean() const {
        if (kind_ != Bool) throw std::runtime_error("JSON: boolean expected");
        return b_;
    }

  private:
    Kind kind_ = Null;
    bool b_ = false;
    double num_ = 0.0;
    std::string str_;
    std::vector<Json> arr_;
    std::map<std::string, Json> obj_;

    static void skip_ws(const std::string& t, size_t& p) {
        while (p < t.size() && std::isspace((unsigned char)t[p])) ++p;
    }
    static void expect(const std::string& t, size_t& p, char c) {
        skip_ws(t, p);
        if (p >= t.size() || t[p] != c)
            throw std::runtime_error(std::string("JSON: expected '") + c + "' at offset " + std::to_string(p));
        ++p;
    }
    static std::string parse_string(const std::string& t, size_t& p) {
        expect(t, p, '"');
        std::string out;
        while (p < t.size() && t[p] != '"') {
            char c = t[p++];
            if (c == '\\') {
                if (p >= t.size()) break;
                char e = t[p++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': {
                        if (p + 4 > t.size()) throw std::runtime_error("JSON: bad \\u escape");
                        unsigned cp = (unsigned)std::strtoul(t.substr(p, 4).c_str(), nullptr, 16);
                        p += 4;
                        if (cp < 0x80) {
                            out += (char)cp;
                        } else if (cp < 0x800) {
                            out += (char)(0xC0 | (cp >> 6));
                            out += (char)(0x80 | (cp & 0x3F));
                        } else {
                            out += (char)(0xE0 | (cp >> 12));
                            out += (char)(0x80 | ((cp >> 6) & 0x3F));
                            out += (char)(0x80 | (cp & 0x3F));
                        }
                        break;
                    }
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (p >= t.size()) throw std::runtime_error("JSON: unterminated string");
        ++p;
        return out;
    }
    static Json parse_value(const std::string& t, size_t& p) {
        skip_ws(t, p);
        if (p >= t.size()) throw std::runtime_error("JSON: unexpected end of input");
        Json v;
        const char c = t[p];
        if (c == '{') {
            v.kind_ = Object;
            ++p;
            skip_ws(t, p);
            if (p < t.size() && t[p] == '}') {
                ++p;
                return v;
            }
            for (;;) {
                std::string k = parse_string(t, p);
                expect(t, p, ':');
                v.obj_[k] = parse_value(t, p);
                skip_ws(t, p);
                if (p < t.size() && t[p] == ',') {
                    ++p;
                    continue;
                }
                expect(t, p, '}');
                return v;
            }
        }
        if (c == '[') {
            v.kind_ = Array;
            ++p;
            skip_ws(t, p);
            if (p < t.size() && t[p] == ']') {
                ++p;
                return v;
            }
            for (;;) {
                v.arr_.push_back(parse_value(t, p));
                skip_ws(t, p);
                if (p < t.size() && t[p] == ',') {
                    ++p;
                    continue;
                }
                expect(t, p, ']');
                return v;
            }
        }
        if (c == '"') {
            v.kind_ = String;
            v.str_ = parse_string(t, p);
            return v;
        }
        if (t.compare(p, 4, "true") == 0) {
            v.kind_ = Bool;
            v.b_ = true;
            p += 4;
            return v;
        }
        if (t.compare(p, 5, "false") == 0) {
            v.kind_ = Bool;
            p += 5;
            return v;
        }
        if (t.compare(p, 4, "null") == 0) {
            p += 4;
            return v;
        }
        const char* s = t.c_str() + p;
        char* end = nullptr;
        const double d = std::strtod(s, &end);
        if (end == s) throw std::runtime_error("JSON: unexpected character at offset " + std::to_string(p));
        v.kind_ = Number;
        v.num_ = d;
        p += (size_t)(end - s);
        return v;
    }
};

} // namespace slabhost
