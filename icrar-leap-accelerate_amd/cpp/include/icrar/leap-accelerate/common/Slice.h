// icrar::Slice / icrar::Range: solution intervals over timesteps
// (src/icrar/leap-accelerate/common/Slice.h:36-55, Slice.cc:28-63, Range.h:34-64).
// Same constructors, wildcard (-1) semantics and validation messages.
#pragma once

#include <cstdint>

#include <string>

#include "icrar/leap-accelerate/common/json_reader.h"
#include "icrar/leap-accelerate/exception/exception.h"

namespace icrar {

struct Range {
    std::int32_t start, interval, end;
    Range(int s, int i, int e) {
        if (s < 0) throw icrar::exception("expected a positive integer", __FILE__, __LINE__);
        if (i < 1) throw icrar::exception("expected a positive integer", __FILE__, __LINE__);
        if (e < 0) throw icrar::exception("expected a positive integer", __FILE__, __LINE__);
        start = s;
        interval = i;
        end = e;
    }
    int GetSize() const { return (end - start) / interval; }
};

struct Slice {
    std::int32_t start = 0, interval = -1, end = -1;

    Slice() = default;
    explicit Slice(int interval_) : Slice(0, interval_, -1) {}
    Slice(int start_, int end_) : Slice(start_, end_ == -1 ? -1 : end_ - start_, end_) {}
    Slice(int start_, int interval_, int end_) {
        if (start_ < 0) throw icrar::exception("expected a positive integer", __FILE__, __LINE__);
        if (interval_ < -1) throw icrar::exception("expected a positive integer or -1", __FILE__, __LINE__);
        if (interval_ == 0) throw icrar::exception("expected a non zero integer", __FILE__, __LINE__);
        if (end_ < -1) throw icrar::exception("expected a positive integer or -1", __FILE__, __LINE__);
        if (end_ != -1 && start_ >= end_)
            throw icrar::exception("range start must be greater than end", __FILE__, __LINE__);
        if (end_ != -1) {
            if (interval_ == -1) interval_ = end_ - start_;
            if (interval_ > end_ - start_) throw icrar::exception("range increment out of bounds", __FILE__, __LINE__);
        }
        start = start_;
        interval = interval_;
        end = end_;
    }

    Range Evaluate(int collectionSize) const {
        return Range(start == -1 ? collectionSize : start, interval == -1 ? collectionSize : interval,
                     end == -1 ? collectionSize : end);
    }
};

// ParseSlice (common/Slice.cc:65-103): an integer is Slice(interval); an array
// of 3 integers Slice(start, interval, end).  A 2-element array is rejected as
// the reference's control flow does ("expected 3 integers").
inline Slice ParseSlice(const JsonValue& doc) {
    if (doc.IsInt()) return Slice(int(doc.num));
    if (doc.kind == JsonValue::Array) {
        if (doc.arr.size() == 3) {
            for (const auto& e : doc.arr)
                if (!e.IsInt()) throw icrar::exception("expected 3 integers", __FILE__, __LINE__);
            return Slice(int(doc.arr[0].num), int(doc.arr[1].num), int(doc.arr[2].num));
        }
        throw icrar::exception("expected 3 integers", __FILE__, __LINE__);
    }
    throw icrar::exception("expected an integer or array of integers", __FILE__, __LINE__);
}

inline Slice ParseSlice(const std::string& json) { return ParseSlice(JsonReader::Parse(json)); }

}  // namespace icrar
