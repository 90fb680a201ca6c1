// The raw measurement-set directory: the main-table columns and sub-table
// values the calibration reads from a casacore MeasurementSet
// (ms/MeasurementSet.cc:33-294, model/cpu/MetaData.cc:46-89,
// ms/utils.h:81-146), stored as little-endian binary files so the calibrator
// can run where casacore is not installed.  icrar-leap-accelerate_amd/ms_to_raw.py
// writes one from a MeasurementSet; leap_amd.write_raw_ms from numpy columns.
//
//   <dir>/header.json   {"format": "leap-raw-ms", "version": 1,
//                        "rows": R, "channels": C, "pols": P,
//                        "antenna_table_rows": N,           ANTENNA table rows
//                        "ref_frequency_hz": f0,            SPECTRAL_WINDOW REF_FREQUENCY[0]
//                        "chan_width_hz": df,               SPECTRAL_WINDOW CHAN_WIDTH[0][0]
//                        "phase_dir_rad": [ra, dec],        FIELD PHASE_DIR[0]
//                        "data_column": "DATA",
//                        "data_type": "complex64" | "complex128"}
//   <dir>/ANTENNA1.i32  R int32
//   <dir>/ANTENNA2.i32  R int32
//   <dir>/TIME.f64      R float64
//   <dir>/FLAG_C0P0.u8  R uint8: FLAG[row][channel 0][correlation 0]
//   <dir>/UVW.f64       R x 3 float64
//   <dir>/DATA.c64 | DATA.c128   R x C x P complex, correlation fastest (the MS's own order)
//
// DATA is memory-mapped read-only and handed to the calibrator as host memory
// (it is staged to the device in chunks); the other columns are read whole.
#pragma once

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdint>
#include <cstring>
#include <fstream>
#include <memory>
#include <optional>
#include <sstream>
#include <string>
#include <vector>

#include "icrar/leap-accelerate/common/json_reader.h"
#include "icrar/leap-accelerate/exception/exception.h"
#include "icrar/leap-accelerate/ms/MeasurementSet.h"

namespace icrar {

namespace raw_ms_detail {

inline std::string Join(const std::string& dir, const std::string& name) {
    return dir.empty() || dir.back() == '/' ? dir + name : dir + "/" + name;
}

template <typename T>
std::vector<T> ReadColumn(const std::string& path, size_t count) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw file_exception("cannot open column", path, __FILE__, __LINE__);
    std::vector<T> v(count);
    f.read(reinterpret_cast<char*>(v.data()), std::streamsize(count * sizeof(T)));
    if (size_t(f.gcount()) != count * sizeof(T))
        throw file_exception("column holds fewer than " + std::to_string(count) + " values", path, __FILE__, __LINE__);
    return v;
}

// A read-only mapping of a whole file, unmapped when the last owner goes.
inline std::shared_ptr<const void> MapFile(const std::string& path, size_t expected_bytes) {
    const int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) throw file_exception("cannot open data column", path, __FILE__, __LINE__);
    struct stat st {};
    if (::fstat(fd, &st) != 0 || size_t(st.st_size) < expected_bytes) {
        ::close(fd);
        throw file_exception("data column holds fewer than " + std::to_string(expected_bytes) + " bytes", path,
                             __FILE__, __LINE__);
    }
    void* p = expected_bytes ? ::mmap(nullptr, expected_bytes, PROT_READ, MAP_SHARED, fd, 0) : nullptr;
    ::close(fd);
    if (p == MAP_FAILED) throw file_exception("cannot map data column", path, __FILE__, __LINE__);
    return std::shared_ptr<const void>(p, [expected_bytes](const void* q) {
        if (q) ::munmap(const_cast<void*>(q), expected_bytes);
    });
}

inline double Number(const JsonValue& h, const char* key, const std::string& path) {
    const JsonValue* v = h.Find(key);
    if (!v || !v->IsNumber()) throw file_exception(std::string("header lacks numeric \"") + key + "\"", path, __FILE__, __LINE__);
    return v->num;
}

}  // namespace raw_ms_detail

// Loads a raw measurement-set directory as icrar::MeasurementSet, with the
// reference constructor's station rule, autocorrelation switch and validation
// (MeasurementSet.cc:33-103).
inline MeasurementSet LoadRawMeasurementSet(const std::string& dir, std::optional<int> overrideNStations = std::nullopt,
                                            bool readAutocorrelations = true) {
    using namespace raw_ms_detail;
    const std::string hpath = Join(dir, "header.json");
    std::ifstream hf(hpath);
    if (!hf) throw file_exception("cannot open header", hpath, __FILE__, __LINE__);
    std::stringstream ss;
    ss << hf.rdbuf();
    const JsonValue h = JsonReader::Parse(ss.str());
    const JsonValue* fmt = h.Find("format");
    if (h.kind != JsonValue::Object || !fmt || fmt->str != "leap-raw-ms")
        throw file_exception("not a leap-raw-ms header", hpath, __FILE__, __LINE__);
    if (Number(h, "version", hpath) != 1.0) throw file_exception("unknown version", hpath, __FILE__, __LINE__);
    const JsonValue* dtype = h.Find("data_type");
    if (!dtype || (dtype->str != "complex64" && dtype->str != "complex128"))
        throw file_exception("data_type must be complex64 or complex128", hpath, __FILE__, __LINE__);
    const JsonValue* dir_rad = h.Find("phase_dir_rad");
    if (!dir_rad || dir_rad->kind != JsonValue::Array || dir_rad->arr.size() != 2 || !dir_rad->arr[0].IsNumber() ||
        !dir_rad->arr[1].IsNumber())
        throw file_exception("phase_dir_rad must be [ra, dec]", hpath, __FILE__, __LINE__);

    MeasurementSetColumns c;
    const double rows = Number(h, "rows", hpath), chans = Number(h, "channels", hpath), pols = Number(h, "pols", hpath);
    const double ants = Number(h, "antenna_table_rows", hpath);
    if (rows < 0 || rows > 4294967295.0 || chans < 0 || chans > 1e9 || pols < 1 || pols > 4 || ants < 0)
        throw file_exception("row, channel, polarisation or antenna count out of range", hpath, __FILE__, __LINE__);
    c.num_rows = uint32_t(rows);
    c.num_channels = uint32_t(chans);
    c.num_pols = uint32_t(pols);
    c.total_antennas = uint32_t(ants);
    c.ref_frequency_hz = Number(h, "ref_frequency_hz", hpath);
    c.chan_width_hz = Number(h, "chan_width_hz", hpath);
    c.phase_centre_ra_rad = dir_rad->arr[0].num;
    c.phase_centre_dec_rad = dir_rad->arr[1].num;
    const size_t R = c.num_rows;
    c.antenna1 = ReadColumn<int32_t>(Join(dir, "ANTENNA1.i32"), R);
    c.antenna2 = ReadColumn<int32_t>(Join(dir, "ANTENNA2.i32"), R);
    c.time = ReadColumn<double>(Join(dir, "TIME.f64"), R);
    c.flag_c0p0 = ReadColumn<uint8_t>(Join(dir, "FLAG_C0P0.u8"), R);
    c.uvw = ReadColumn<double>(Join(dir, "UVW.f64"), R * 3);
    const bool c64 = dtype->str == "complex64";
    c.data_dtype = c64 ? LEAP_VIS_C64 : LEAP_VIS_C128;
    const size_t data_bytes = R * c.num_channels * c.num_pols * (c64 ? 8 : 16);
    c.data_owner = MapFile(Join(dir, c64 ? "DATA.c64" : "DATA.c128"), data_bytes);
    c.data = c.data_owner.get();
    return MeasurementSet(std::move(c), overrideNStations, readAutocorrelations);
}

}  // namespace icrar
