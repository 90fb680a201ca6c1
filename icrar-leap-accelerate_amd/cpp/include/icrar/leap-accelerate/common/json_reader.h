// A small JSON reader (RFC 8259 values; numbers as double) for the
// configuration file, the command-line direction and slice strings, and the
// raw measurement-set header.  The reference parses the same inputs with
// RapidJSON (leap-accelerate-cli/Arguments.cc:311-482).
#pragma once

#include <cctype>
#include <cstdint>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "icrar/leap-accelerate/exception/exception.h"

namespace icrar {

struct JsonValue {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    bool integral = false;  // a number written without fraction or exponent
    std::string str;
    std::vector<JsonValue> arr;
    std::vector<std::pair<std::string, JsonValue>> obj;

    bool IsNumber() const { return kind == Number; }
    // RapidJSON's IsInt / IsDouble: an integer literal in int range / a literal with fraction or exponent.
    bool IsInt() const { return kind == Number && integral && num >= -2147483648.0 && num <= 2147483647.0; }
    bool IsDouble() const { return kind == Number && !integral; }
    const JsonValue* Find(const std::string& key) const {
        for (const auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

class JsonReader {
    const char* m_p;
    const char* m_begin;

    [[noreturn]] void Fail(const std::string& what) const {
        throw json_exception(what + " at offset " + std::to_string(m_p - m_begin), __FILE__, __LINE__);
    }
    void Ws() {
        while (*m_p && std::isspace(static_cast<unsigned char>(*m_p))) ++m_p;
    }
    void Expect(char c) {
        Ws();
        if (*m_p != c) Fail(std::string("expected '") + c + "'");
        ++m_p;
    }
    // Four hex digits after "\\u" (m_p at the 'u'); leaves m_p after them.
    uint32_t ParseHex4() {
        uint32_t v = 0;
        for (int i = 1; i <= 4; ++i) {
            const char c = m_p[i];
            uint32_t d;
            if (c >= '0' && c <= '9') d = uint32_t(c - '0');
            else if (c >= 'a' && c <= 'f') d = uint32_t(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') d = uint32_t(c - 'A' + 10);
            else Fail("invalid \\u escape");
            v = v * 16 + d;
        }
        m_p += 5;
        return v;
    }
    // A \\u escape (m_p at the 'u'), with a UTF-16 surrogate pair joined into
    // one code point; a lone surrogate is an error.
    uint32_t ParseEscapedCodePoint() {
        const uint32_t hi = ParseHex4();
        if (hi >= 0xDC00 && hi <= 0xDFFF) Fail("lone low surrogate in \\u escape");
        if (hi < 0xD800 || hi > 0xDBFF) return hi;
        if (m_p[0] != '\\' || m_p[1] != 'u') Fail("high surrogate without a low surrogate");
        ++m_p;
        const uint32_t lo = ParseHex4();
        if (lo < 0xDC00 || lo > 0xDFFF) Fail("high surrogate without a low surrogate");
        return 0x10000 + ((hi - 0xD800) << 10) + (lo - 0xDC00);
    }
    static void AppendUtf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) {
            out += char(cp);
        } else if (cp < 0x800) {
            out += char(0xC0 | (cp >> 6));
            out += char(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += char(0xE0 | (cp >> 12));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        } else {
            out += char(0xF0 | (cp >> 18));
            out += char(0x80 | ((cp >> 12) & 0x3F));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        }
    }
    std::string ParseString() {
        Expect('"');
        std::string out;
        while (*m_p && *m_p != '"') {
            if (*m_p == '\\') {
                ++m_p;
                switch (*m_p) {
                    case '"': out += '"'; break;
                    case '\\': out += '\\'; break;
                    case '/': out += '/'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'n': out += '\n'; break;
                    case 'r': out += '\r'; break;
                    case 't': out += '\t'; break;
                    case 'u': AppendUtf8(out, ParseEscapedCodePoint()); continue;
                    default: Fail("unsupported escape");
                }
                ++m_p;
            } else if (static_cast<unsigned char>(*m_p) < 0x20) {
                Fail("unescaped control character in string");
            } else {
                out += *m_p++;
            }
        }
        if (*m_p != '"') Fail("unterminated string");
        ++m_p;
        return out;
    }
    JsonValue ParseValue(int depth) {
        if (depth > 64) Fail("nesting too deep");
        Ws();
        JsonValue v;
        if (*m_p == '{') {
            ++m_p;
            v.kind = JsonValue::Object;
            Ws();
            if (*m_p == '}') {
                ++m_p;
                return v;
            }
            for (;;) {
                std::string k = ParseString();
                Expect(':');
                v.obj.emplace_back(std::move(k), ParseValue(depth + 1));
                Ws();
                if (*m_p == ',') {
                    ++m_p;
                    continue;
                }
                Expect('}');
                return v;
            }
        }
        if (*m_p == '[') {
            ++m_p;
            v.kind = JsonValue::Array;
            Ws();
            if (*m_p == ']') {
                ++m_p;
                return v;
            }
            for (;;) {
                v.arr.push_back(ParseValue(depth + 1));
                Ws();
                if (*m_p == ',') {
                    ++m_p;
                    continue;
                }
                Expect(']');
                return v;
            }
        }
        if (*m_p == '"') {
            v.kind = JsonValue::String;
            v.str = ParseString();
            return v;
        }
        if (std::strncmp(m_p, "true", 4) == 0) {
            m_p += 4;
            v.kind = JsonValue::Bool;
            v.b = true;
            return v;
        }
        if (std::strncmp(m_p, "false", 5) == 0) {
            m_p += 5;
            v.kind = JsonValue::Bool;
            return v;
        }
        if (std::strncmp(m_p, "null", 4) == 0) {
            m_p += 4;
            return v;
        }
        if (*m_p == '-' || std::isdigit(static_cast<unsigned char>(*m_p))) {
            // RFC 8259 number grammar: -?(0|[1-9][0-9]*)(.[0-9]+)?([eE][+-]?[0-9]+)?; the text
            // is checked before strtod, which would also take hex, "inf" or a bare "5.".
            auto digit = [](char c) { return std::isdigit(static_cast<unsigned char>(c)) != 0; };
            const char* q = m_p;
            if (*q == '-') ++q;
            if (*q == '0') {
                ++q;
            } else if (digit(*q)) {
                while (digit(*q)) ++q;
            } else {
                Fail("bad number");
            }
            v.integral = true;
            if (*q == '.') {
                ++q;
                if (!digit(*q)) Fail("bad number");
                while (digit(*q)) ++q;
                v.integral = false;
            }
            if (*q == 'e' || *q == 'E') {
                ++q;
                if (*q == '+' || *q == '-') ++q;
                if (!digit(*q)) Fail("bad number");
                while (digit(*q)) ++q;
                v.integral = false;
            }
            const std::string text(m_p, q);
            v.num = std::strtod(text.c_str(), nullptr);
            if (!std::isfinite(v.num)) Fail("number too big");  // as rapidjson's kParseErrorNumberTooBig
            m_p = q;
            v.kind = JsonValue::Number;
            return v;
        }
        Fail("unexpected character");
    }

public:
    // Parses a whole document; trailing characters other than whitespace are an error.
    static JsonValue Parse(const std::string& text) {
        JsonReader r;
        r.m_begin = r.m_p = text.c_str();
        JsonValue v = r.ParseValue(0);
        r.Ws();
        if (*r.m_p) r.Fail("trailing characters");
        return v;
    }
};

}  // namespace icrar
