// Driver for the C++ API (ILeapCalibrator / LeapCalibratorFactory / HipLeapCalibrator).
//
//   hip_calibrator_main selftest
//       host-only checks of the JSON number format and Slice semantics (no GPU).
//   hip_calibrator_main calibrate <antennas> <timesteps> <channels> <seed> <directions>
//                                 <slice_start> <slice_interval> <slice_end> <ref_ant|-1>
//                                 <c64|c128> <plain|batched|mixed> [vis_out_file]
//       builds a synthetic measurement set through the C ABI, wraps it in
//       icrar::MeasurementSet, calibrates with the hip implementation from the
//       factory and prints the CalibrationCollection JSON.  When vis_out_file is
//       given, the visibilities are also written there (raw, [row][chan][pol])
//       so a checker can recompute the same calibration.
//   hip_calibrator_main msinfo <raw-ms-dir> [stations|-1] [autocorrelations 0|1]
//       loads a raw measurement-set directory (host only) and prints its
//       accessors (stations, baselines, epochs, ...) as JSON.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <random>
#include <string>
#include <vector>

#include "icrar/leap-accelerate/algorithm/LeapCalibratorFactory.h"
#include "icrar/leap-accelerate/common/SphericalDirection.h"
#include "icrar/leap-accelerate/ms/MeasurementSet.h"
#include "icrar/leap-accelerate/ms/RawMeasurementSet.h"

using namespace icrar;

static int selftest() {
    struct Case {
        double v;
        const char* s;
    } cases[] = {
        {5020320156.0, "5020320156.0"}, {2.58353835690993, "2.58353835690993"},
        {6.25039512024925e-15, "6.25039512024925e-15"}, {-0.252472710942695, "-0.252472710942695"},
        {1e-7, "1e-7"}, {0.000125, "0.000125"}, {1e22, "1e22"}, {0.0, "0.0"}, {-3.0, "-3.0"},
        {123456789012345680000.0, "123456789012345680000.0"}, {1.5e-6, "0.0000015"},
    };
    int bad = 0;
    for (const auto& c : cases) {
        const std::string got = FormatJsonDouble(c.v);
        if (got != c.s) {
            std::cerr << "format " << c.s << " -> " << got << "\n";
            ++bad;
        }
        if (std::strtod(got.c_str(), nullptr) != c.v) {
            std::cerr << "round trip " << got << "\n";
            ++bad;
        }
    }
    // Slice semantics (common/Slice.cc:28-63)
    if (Slice(0, 1).Evaluate(14).GetSize() != 1) ++bad;
    if (Slice(0, -1).Evaluate(14).GetSize() != 1) ++bad;
    if (Slice(1).Evaluate(14).GetSize() != 14) ++bad;
    try {
        Slice(0, 0, -1);
        ++bad;
    } catch (const icrar::exception&) {
    }
    try {
        LeapCalibratorFactory::Create(ComputeImplementation::cuda);
        ++bad;
    } catch (const invalid_argument_exception&) {
    }
    // ParseDirections accept / reject cases of leap-accelerate-cli/tests/JSONHelperTests.cc:59-90.
    // The one-direction case there compares against a casacore MVDirection round trip of
    // the same two numbers; here the parsed values are compared with the numbers themselves.
    if (!ParseDirections("[]").empty()) ++bad;
    // JSON number grammar as rapidjson reads it: no hex, no bare "5." or ".5", no leading zeros,
    // and no overflow to infinity
    for (const char* reject : {"{}", "[[]]", "[[1.0]]", "[[1.0,1.0,1.0]]", "[1.0,2.0]", "[[\"a\",1.0]]",
                               "[[0x10,2]]", "[[5.,2]]", "[[.5,2]]", "[[01,2]]", "[[1e999,2]]", "[[-,2]]",
                               "[[1e,2]]", "[[inf,2]]"}) {
        try {
            ParseDirections(std::string(reject));
            std::cerr << "ParseDirections accepted " << reject << "\n";
            ++bad;
        } catch (const icrar::exception&) {
        }
    }
    {
        const auto one = ParseDirections("[[-0.4606549305661674,-0.29719233792392513]]");
        if (one.size() != 1 || !(one[0] == SphericalDirection(-0.4606549305661674, -0.29719233792392513))) ++bad;
        const auto num = ParseDirections("[[-0.5e-3,2E+2],[0,-0.0],[1e-400,12.5E-1]]");
        if (num.size() != 3 || !(num[0] == SphericalDirection(-0.0005, 200.0)) ||
            !(num[1] == SphericalDirection(0.0, 0.0)) || !(num[2] == SphericalDirection(0.0, 1.25)))
            ++bad;
        const auto five = ParseDirections("[[0.0,0.0],[1.0,1.0],[2.0,2.0],[3.0,3.0],[4.0,4.0]]");
        if (five.size() != 5) ++bad;
        for (size_t i = 0; i < five.size(); ++i)
            if (!(five[i] == SphericalDirection(double(i), double(i)))) ++bad;
    }
    // String escapes as RapidJSON reads them: \uXXXX (surrogate pairs joined) decoded to UTF-8;
    // lone surrogates, bad hex digits and raw control characters rejected.
    {
        const std::pair<const char*, std::string> accept[] = {
            {"\"a\\u0041\"", "aA"},
            {"\"\\u00e9t\\u00C9\"", "\xC3\xA9t\xC3\x89"},
            {"\"\\u20AC\"", "\xE2\x82\xAC"},
            {"\"\\ud83d\\ude00\"", "\xF0\x9F\x98\x80"},
            {"\"/data/\\u6e2c\\u5b9a.ms\"", "/data/\xE6\xB8\xAC\xE5\xAE\x9A.ms"},
            {"\"\\u0000x\"", std::string("\0x", 2)},
        };
        for (const auto& c : accept) {
            try {
                const JsonValue v = JsonReader::Parse(c.first);
                if (v.kind != JsonValue::String || v.str != c.second) {
                    std::cerr << "string escape " << c.first << " decoded wrongly\n";
                    ++bad;
                }
            } catch (const icrar::exception&) {
                std::cerr << "string escape " << c.first << " rejected\n";
                ++bad;
            }
        }
        for (const char* reject : {"\"\\u12\"", "\"\\u12G4\"", "\"\\ud83d\"", "\"\\ud83dx\"",
                                   "\"\\ud83d\\u0041\"", "\"\\ude00\"", "\"a\tb\"", "\"\\x41\""}) {
            try {
                JsonReader::Parse(reject);
                std::cerr << "string escape accepted " << reject << "\n";
                ++bad;
            } catch (const icrar::exception&) {
            }
        }
    }
    cpu::CalibrationCollection cc({cpu::Calibration(1.0, 2.0, {cpu::BeamCalibration({0.5, -0.25}, {0.0, 1.5})})});
    std::ostringstream os;
    cc.Serialize(os);
    if (os.str() != "[{\"epoch\":{\"start\":1.0,\"end\":2.0},\"calibration\":[{\"direction\":[0.5,-0.25],"
                    "\"beamCalibration\":[0.0,1.5]}]}]\n") {
        std::cerr << "collection json " << os.str();
        ++bad;
    }
    std::cout << (bad ? "selftest FAILED\n" : "selftest ok\n");
    return bad ? 1 : 0;
}

static int calibrate(int argc, char** argv) {
    if (argc < 13) {
        std::cerr << "usage: calibrate <antennas> <timesteps> <channels> <seed> <directions> <slice_start> "
                     "<slice_interval> <slice_end> <ref_ant|-1> <c64|c128> <plain|batched|mixed> [vis_out]\n";
        return 2;
    }
    leap_synth_config cfg{};
    cfg.n_antennas = uint32_t(std::atoi(argv[2]));
    cfg.layout = 0;
    cfg.array_radius_m = 750.0;
    cfg.autocorrelations = 1;
    cfg.n_timesteps = uint32_t(std::atoi(argv[3]));
    cfg.integration_s = 2.0;
    cfg.n_channels = uint32_t(std::atoi(argv[4]));
    cfg.freq_start_hz = 169.6e6;
    cfg.freq_inc_hz = 40e3;
    cfg.n_pols = 4;
    cfg.phase_centre_ra_rad = 0.5;
    cfg.phase_centre_dec_rad = -27.0 * (M_PI / 180.0);  // as numpy.deg2rad
    cfg.latitude_rad = -26.7033 * (M_PI / 180.0);
    cfg.n_sources = 5;
    cfg.source_flux_min = 1.0;
    cfg.source_flux_max = 10.0;
    cfg.source_offset_max_rad = 0.05;
    cfg.noise_sigma = 0.1;
    cfg.flag_fraction = 0.01;
    cfg.seed = uint64_t(std::atoll(argv[5]));
    const int n_dirs = std::atoi(argv[6]);
    const Slice slice(std::atoi(argv[7]), std::atoi(argv[8]), std::atoi(argv[9]));
    const int ref = std::atoi(argv[10]);
    const bool c64 = std::string(argv[11]) == "c64";
    const std::string mode = argv[12];

    leap_synth_ms* synth = nullptr;
    if (leap_synth_create(&cfg, &synth) != LEAP_OK) {
        std::cerr << leap_hip_last_error(nullptr) << "\n";
        return 1;
    }
    leap_ms_view sv;
    leap_synth_view(synth, &sv);
    MeasurementSetColumns cols;
    cols.total_antennas = sv.total_antennas;
    cols.num_rows = sv.num_rows;
    cols.num_channels = sv.num_channels;
    cols.num_pols = sv.num_pols;
    cols.ref_frequency_hz = sv.freq_start_hz;
    cols.chan_width_hz = sv.freq_inc_hz;
    cols.phase_centre_ra_rad = sv.phase_centre_ra_rad;
    cols.phase_centre_dec_rad = sv.phase_centre_dec_rad;
    // ANTENNA1/2 and FLAG repeat every epoch in the MS.
    for (uint32_t r = 0; r < sv.num_rows; ++r) {
        cols.antenna1.push_back(sv.antenna1[r % sv.num_baselines]);
        cols.antenna2.push_back(sv.antenna2[r % sv.num_baselines]);
        cols.flag_c0p0.push_back(sv.flag[r % sv.num_baselines]);
    }
    cols.time.assign(sv.time, sv.time + sv.num_rows);
    cols.uvw.assign(sv.uvw, sv.uvw + size_t(sv.num_rows) * 3);
    const size_t n_vis = size_t(sv.num_rows) * sv.num_channels * sv.num_pols;
    std::vector<char> vis(n_vis * (c64 ? 8 : 16));
    leap_synth_fill_vis_host(synth, 0, sv.num_rows, c64 ? LEAP_VIS_C64 : LEAP_VIS_C128, vis.data());
    if (argc > 13) {
        std::ofstream f(argv[13], std::ios::binary);
        f.write(vis.data(), std::streamsize(vis.size()));
    }
    cols.data = vis.data();
    cols.data_dtype = c64 ? LEAP_VIS_C64 : LEAP_VIS_C128;
    // The synthetic set lists every antenna in ANTENNA1 (autocorrelations on),
    // so the unique-antenna station count is already right.
    MeasurementSet ms(std::move(cols), std::nullopt, true);

    std::vector<SphericalDirection> dirs;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> u(-0.02, 0.02);
    for (int i = 0; i < n_dirs; ++i) dirs.emplace_back(cfg.phase_centre_ra_rad + u(rng), cfg.phase_centre_dec_rad + u(rng));

    auto calibrator = LeapCalibratorFactory::Create(ComputeImplementation::hip);
    auto* hipcal = dynamic_cast<hip::HipLeapCalibrator*>(calibrator.get());
    leap_options o = hipcal->GetOptions();
    if (mode != "plain") o.strategy = LEAP_STRATEGY_BATCHED;
    if (mode == "mixed") o.precision = LEAP_PRECISION_MIXED;
    hipcal->SetOptions(o);
    const std::optional<unsigned int> refAnt = ref >= 0 ? std::optional<unsigned int>(unsigned(ref)) : std::nullopt;
    const cpu::CalibrationCollection result = calibrator->Calibrate(ms, dirs, slice, 0.0, refAnt, false);
    result.Serialize(std::cout);
    leap_synth_destroy(synth);
    return 0;
}

// msinfo <raw-ms-dir> [stations|-1] [autocorrelations 0|1]: the accessors of a
// loaded raw measurement set as JSON (host only).
static int msinfo(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "usage: msinfo <raw-ms-dir> [stations|-1] [autocorrelations 0|1]\n";
        return 2;
    }
    std::optional<int> stations;
    if (argc > 3 && std::atoi(argv[3]) >= 0) stations = std::atoi(argv[3]);
    const bool autos = argc > 4 ? std::atoi(argv[4]) != 0 : true;
    const MeasurementSet ms = LoadRawMeasurementSet(argv[2], stations, autos);
    const leap_ms_view v = ms.View();
    std::ostringstream os;
    os << "{\"stations\": " << ms.GetNumStations() << ", \"total_antennas\": " << ms.GetTotalAntennas()
       << ", \"baselines\": " << ms.GetNumBaselines() << ", \"rows\": " << ms.GetNumRows()
       << ", \"timesteps\": " << ms.GetNumTimesteps() << ", \"channels\": " << ms.GetNumChannels()
       << ", \"pols\": " << ms.GetNumPols() << ", \"flagged_baselines\": " << ms.GetNumFlaggedBaselines()
       << ", \"freq_start_hz\": " << FormatJsonDouble(v.freq_start_hz)
       << ", \"freq_inc_hz\": " << FormatJsonDouble(v.freq_inc_hz)
       << ", \"phase_centre\": [" << FormatJsonDouble(v.phase_centre_ra_rad) << ", "
       << FormatJsonDouble(v.phase_centre_dec_rad) << "], \"vis_dtype\": " << v.vis_dtype << ", \"epochs\": [";
    const std::vector<double> ep = ms.GetEpochs();
    for (size_t i = 0; i < ep.size(); ++i) os << (i ? ", " : "") << FormatJsonDouble(ep[i]);
    // checksum of the mapped visibilities: sum of all values as doubles
    double sum = 0.0;
    const size_t n = size_t(v.num_rows) * v.num_channels * v.num_pols * 2;
    if (v.vis_dtype == LEAP_VIS_C64) {
        const float* d = static_cast<const float*>(v.vis);
        for (size_t i = 0; i < n; ++i) sum += std::isfinite(d[i]) ? double(d[i]) : 0.0;
    } else {
        const double* d = static_cast<const double*>(v.vis);
        for (size_t i = 0; i < n; ++i) sum += std::isfinite(d[i]) ? d[i] : 0.0;
    }
    os << "], \"vis_sum\": " << FormatJsonDouble(sum) << "}";
    std::cout << os.str() << std::endl;
    return 0;
}

int main(int argc, char** argv) {
    try {
        if (argc > 1 && std::string(argv[1]) == "selftest") return selftest();
        if (argc > 1 && std::string(argv[1]) == "calibrate") return calibrate(argc, argv);
        if (argc > 1 && std::string(argv[1]) == "msinfo") return msinfo(argc, argv);
        std::cerr << "usage: hip_calibrator_main selftest | calibrate ...\n";
        return 2;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
