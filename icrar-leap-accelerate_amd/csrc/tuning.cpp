// The autotuner's choices as JSON, so a deployment tunes a shape once and
// reloads the choice in later processes.
//
// {"format": "leap-hip-tuning", "version": 1,
//  "batched": [{"mode": m, "vis_dtype": v, "all_pols": a, "directions": D, "timesteps": T,
//               "channels": C, "pols": P, "tile": t}, ...],
//  "multi": [{"key": [k0, k1, ...], "tile": t, "workgroups_per_cu": w}, ...]}
//
// "mode" is the batched tuner's key code (precision, FP32 blocks, lane-direction
// request, packed path and NUFFT channel sums); "key" the multi-set tuner's key (requested tile and
// grid depth, then nbl, T, C, D and storage type of each set).
#include <cctype>
#include <cstdint>
#include <cstring>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "leap_internal.h"

namespace leap {

namespace {

// A minimal JSON reader for the documents written below: objects, arrays,
// integers, strings, true / false / null.
struct JValue {
    enum Kind { Null, Bool, Int, Str, Arr, Obj } kind = Null;
    int64_t i = 0;
    std::string s;
    std::vector<JValue> a;
    std::vector<std::pair<std::string, JValue>> o;
    const JValue* get(const char* k) const {
        for (const auto& kv : o)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const char* p;
    void fail(const char* what) const {
        throw Error(LEAP_ERR_INVALID_ARGUMENT, std::string("tuning JSON: ") + what);
    }
    void ws() {
        while (*p && std::isspace(static_cast<unsigned char>(*p))) ++p;
    }
    void expect(char c) {
        ws();
        if (*p != c) fail("unexpected character");
        ++p;
    }
    std::string str() {
        expect('"');
        std::string out;
        while (*p && *p != '"') {
            if (*p == '\\') fail("escapes are not used in tuning documents");
            out += *p++;
        }
        if (*p != '"') fail("unterminated string");
        ++p;
        return out;
    }
    JValue value(int depth = 0) {
        if (depth > 16) fail("nesting too deep");
        ws();
        JValue v;
        if (*p == '{') {
            ++p;
            v.kind = JValue::Obj;
            ws();
            if (*p == '}') {
                ++p;
                return v;
            }
            for (;;) {
                std::string k = str();
                expect(':');
                v.o.emplace_back(std::move(k), value(depth + 1));
                ws();
                if (*p == ',') {
                    ++p;
                    continue;
                }
                expect('}');
                return v;
            }
        }
        if (*p == '[') {
            ++p;
            v.kind = JValue::Arr;
            ws();
            if (*p == ']') {
                ++p;
                return v;
            }
            for (;;) {
                v.a.push_back(value(depth + 1));
                ws();
                if (*p == ',') {
                    ++p;
                    continue;
                }
                expect(']');
                return v;
            }
        }
        if (*p == '"') {
            v.kind = JValue::Str;
            v.s = str();
            return v;
        }
        if (std::strncmp(p, "true", 4) == 0 || std::strncmp(p, "false", 5) == 0) {
            v.kind = JValue::Bool;
            v.i = (*p == 't');
            p += v.i ? 4 : 5;
            return v;
        }
        if (std::strncmp(p, "null", 4) == 0) {
            p += 4;
            return v;
        }
        if (*p == '-' || std::isdigit(static_cast<unsigned char>(*p))) {
            const bool neg = *p == '-';
            if (neg) ++p;
            if (!std::isdigit(static_cast<unsigned char>(*p))) fail("bad number");
            int64_t x = 0;
            while (std::isdigit(static_cast<unsigned char>(*p))) {
                if (x > (int64_t(1) << 53)) fail("number out of range");
                x = x * 10 + (*p++ - '0');
            }
            if (*p == '.' || *p == 'e' || *p == 'E') fail("tuning documents hold integers only");
            v.kind = JValue::Int;
            v.i = neg ? -x : x;
            return v;
        }
        fail("unexpected character");
        return v;
    }
};

int64_t int_field(const JValue& obj, const char* k) {
    const JValue* v = obj.get(k);
    if (!v || v->kind != JValue::Int)
        throw Error(LEAP_ERR_INVALID_ARGUMENT, std::string("tuning JSON: missing integer field \"") + k + "\"");
    return v->i;
}

bool batched_tile_known(int32_t t) {
    switch (t) {
        case 321: case 641: case 642: case 1281: case 1282: case 1284: case 2561: case 2562: case 2564: return true;
        default: return false;
    }
}

}  // namespace

std::string tuning_to_json(const leap_ctx& ctx) {
    std::ostringstream js;
    js << "{\"format\": \"leap-hip-tuning\", \"version\": 1, \"batched\": [";
    bool first = true;
    for (const auto& kv : ctx.tuned) {
        const TuneKey& k = kv.first;
        js << (first ? "" : ", ") << "{\"mode\": " << k.precision << ", \"vis_dtype\": " << k.vis_dtype
           << ", \"all_pols\": " << k.all_pols << ", \"directions\": " << k.D << ", \"timesteps\": " << k.T
           << ", \"channels\": " << k.C << ", \"pols\": " << k.P << ", \"tile\": " << kv.second << "}";
        first = false;
    }
    js << "], \"multi\": [";
    first = true;
    for (const auto& kv : ctx.tuned_multi) {
        js << (first ? "" : ", ") << "{\"key\": [";
        for (size_t i = 0; i < kv.first.size(); ++i) js << (i ? ", " : "") << kv.first[i];
        js << "], \"tile\": " << kv.second.first << ", \"workgroups_per_cu\": " << kv.second.second << "}";
        first = false;
    }
    js << "]}";
    return js.str();
}

void tuning_from_json(leap_ctx& ctx, const char* text) {
    require(text != nullptr, LEAP_ERR_INVALID_ARGUMENT, "null tuning document");
    JParser ps{text};
    const JValue doc = ps.value();
    ps.ws();
    require(*ps.p == '\0', LEAP_ERR_INVALID_ARGUMENT, "tuning JSON: trailing characters");
    require(doc.kind == JValue::Obj, LEAP_ERR_INVALID_ARGUMENT, "tuning JSON: not an object");
    const JValue* fmt = doc.get("format");
    require(fmt && fmt->kind == JValue::Str && fmt->s == "leap-hip-tuning", LEAP_ERR_INVALID_ARGUMENT,
            "tuning JSON: not a leap-hip-tuning document");
    require(int_field(doc, "version") == 1, LEAP_ERR_INVALID_ARGUMENT, "tuning JSON: unknown version");
    // Validate everything before changing the context.
    std::map<TuneKey, int32_t> batched;
    std::map<std::vector<uint32_t>, std::pair<int32_t, int32_t>> multi;
    if (const JValue* arr = doc.get("batched")) {
        require(arr->kind == JValue::Arr, LEAP_ERR_INVALID_ARGUMENT, "tuning JSON: \"batched\" is not an array");
        for (const JValue& e : arr->a) {
            require(e.kind == JValue::Obj, LEAP_ERR_INVALID_ARGUMENT, "tuning JSON: batched entry is not an object");
            const TuneKey k{int32_t(int_field(e, "mode")), int32_t(int_field(e, "vis_dtype")),
                            int32_t(int_field(e, "all_pols")), uint32_t(int_field(e, "directions")),
                            uint32_t(int_field(e, "timesteps")), uint32_t(int_field(e, "channels")),
                            uint32_t(int_field(e, "pols"))};
            const int32_t tile = int32_t(int_field(e, "tile"));
            require(batched_tile_known(tile), LEAP_ERR_INVALID_ARGUMENT,
                    "tuning JSON: unknown batched tile " + std::to_string(tile));
            require(!(k.precision & kTunePackedBit) || packed_tile_supported(tile), LEAP_ERR_INVALID_ARGUMENT,
                    "tuning JSON: tile " + std::to_string(tile) + " is not a packed-path tile");
            require(!(k.precision & kTuneNufftBit) || packed_tile_supported(tile), LEAP_ERR_INVALID_ARGUMENT,
                    "tuning JSON: tile " + std::to_string(tile) + " is not a NUFFT tile");
            require(!(k.precision & 2) || tile % 10 == 1, LEAP_ERR_INVALID_ARGUMENT,
                    "tuning JSON: the FP32 blocked kernel runs one direction per lane");
            batched[k] = tile;
        }
    }
    if (const JValue* arr = doc.get("multi")) {
        require(arr->kind == JValue::Arr, LEAP_ERR_INVALID_ARGUMENT, "tuning JSON: \"multi\" is not an array");
        for (const JValue& e : arr->a) {
            require(e.kind == JValue::Obj, LEAP_ERR_INVALID_ARGUMENT, "tuning JSON: multi entry is not an object");
            const JValue* key = e.get("key");
            require(key && key->kind == JValue::Arr, LEAP_ERR_INVALID_ARGUMENT, "tuning JSON: multi entry has no key");
            std::vector<uint32_t> k;
            for (const JValue& x : key->a) {
                require(x.kind == JValue::Int && x.i >= 0 && x.i <= int64_t(UINT32_MAX), LEAP_ERR_INVALID_ARGUMENT,
                        "tuning JSON: bad multi key");
                k.push_back(uint32_t(x.i));
            }
            const int32_t tile = int32_t(int_field(e, "tile"));
            const int32_t wpc = int32_t(int_field(e, "workgroups_per_cu"));
            require(multi_tile_supported(tile), LEAP_ERR_INVALID_ARGUMENT,
                    "tuning JSON: unknown multi-set tile " + std::to_string(tile));
            require(wpc == -1 || (wpc >= 1 && wpc <= 512), LEAP_ERR_INVALID_ARGUMENT,
                    "tuning JSON: workgroups_per_cu out of range");
            multi[k] = {tile, wpc};
        }
    }
    for (const auto& kv : batched) ctx.tuned[kv.first] = kv.second;
    for (const auto& kv : multi) ctx.tuned_multi[kv.first] = kv.second;
}

}  // namespace leap
