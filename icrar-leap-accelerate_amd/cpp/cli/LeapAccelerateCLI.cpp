// LeapAccelerateCLI over the gfx950 calibrator: the reference's command line
// (leap-accelerate-cli/main.cc:60-133) and argument layering
// (leap-accelerate-cli/Arguments.cc:37-238, 311-482): defaults, then the JSON
// config file, then the command-line options; the calibration collection is
// written as JSON to the output file or stdout.
//
// Differences: the measurement set is a raw measurement-set directory
// (icrar/leap-accelerate/ms/RawMeasurementSet.h; ms_to_raw.py converts a
// casacore MeasurementSet); the default and only available implementation is
// "hip" ("cpu" and "cuda" are refused as not built, as the reference's factory
// refuses an implementation it was built without); --precision and --strategy
// select the library's schedules.
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <optional>
#include <sstream>
#include <string>
#include <vector>

#include "icrar/leap-accelerate/algorithm/LeapCalibratorFactory.h"
#include "icrar/leap-accelerate/algorithm/hip/HipLeapCalibrator.h"
#include "icrar/leap-accelerate/common/Slice.h"
#include "icrar/leap-accelerate/common/SphericalDirection.h"
#include "icrar/leap-accelerate/common/json_reader.h"
#include "icrar/leap-accelerate/ms/RawMeasurementSet.h"

using namespace icrar;

namespace {

// Arguments as given (unset = not given), Arguments.h of the reference.
struct Arguments {
    std::optional<std::string> filePath, outputFilePath, directions, solutionInterval, computeImplementation;
    std::optional<std::vector<SphericalDirection>> parsedDirections;
    std::optional<Slice> parsedSlice;
    std::optional<int> stations;
    std::optional<unsigned int> referenceAntenna;
    std::optional<bool> readAutocorrelations, useFileSystemCache, mwaSupport;
    std::optional<double> minimumBaselineThreshold;
    std::optional<int> verbosity;
    std::optional<std::string> precision, strategy;
};

// Arguments.cc:37-55, with hip as the implementation.
Arguments Defaults() {
    Arguments a;
    a.computeImplementation = "hip";
    a.solutionInterval = "[0,-1,-1]";
    a.readAutocorrelations = true;
    a.minimumBaselineThreshold = 0.0;
    a.mwaSupport = false;
    a.useFileSystemCache = true;
    a.verbosity = 3;
    a.precision = "fp64";
    a.strategy = "per-direction";
    return a;
}

bool ParseVerbosityName(const std::string& s, int& out) {
    static const std::map<std::string, int> names{{"fatal", 0}, {"error", 1}, {"warn", 2},
                                                  {"info", 3},  {"debug", 4}, {"trace", 5}};
    auto it = names.find(s);
    if (it == names.end()) return false;
    out = it->second;
    return true;
}

// ParseConfig (Arguments.cc:311-482), key by key with the reference's type checks and messages.
Arguments ParseConfig(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw file_exception("cannot open config", path, __FILE__, __LINE__);
    std::stringstream ss;
    ss << f.rdbuf();
    const JsonValue doc = JsonReader::Parse(ss.str());
    if (doc.kind != JsonValue::Object) throw json_exception("expected config to be an object", __FILE__, __LINE__);
    Arguments a;
    for (const auto& kv : doc.obj) {
        const std::string& key = kv.first;
        const JsonValue& v = kv.second;
        if (key == "source") {
        } else if (key == "filePath") {
            if (v.kind != JsonValue::String) throw json_exception("filePath must be of type string", __FILE__, __LINE__);
            a.filePath = v.str;
        } else if (key == "configFilePath") {
            throw json_exception("recursive config detected", __FILE__, __LINE__);
        } else if (key == "outputFilePath") {
            if (v.kind != JsonValue::String)
                throw json_exception("outFilePath must be of type string", __FILE__, __LINE__);
            a.outputFilePath = v.str;
        } else if (key == "stations") {
            if (!v.IsInt()) throw json_exception("outFilePath must be of type int", __FILE__, __LINE__);
            a.stations = int(v.num);
        } else if (key == "solutionInterval") {
            if (v.kind == JsonValue::Array) a.parsedSlice = ParseSlice(v);
        } else if (key == "referenceAntenna") {
            if (!v.IsInt())
                throw json_exception("referenceAntenna must be of type unsigned int", __FILE__, __LINE__);
            a.referenceAntenna = static_cast<unsigned int>(int(v.num));
        } else if (key == "directions") {
            a.parsedDirections = ParseDirections(v);
        } else if (key == "computeImplementation") {
            ComputeImplementation e;
            if (v.kind != JsonValue::String || !TryParseComputeImplementation(v.str, e))
                throw json_exception("invalid compute implementation string", __FILE__, __LINE__);
            a.computeImplementation = v.str;
        } else if (key == "minimumBaselineThreshold") {
            if (!v.IsDouble())
                throw json_exception("minimumBaselineThreshold must be of type double", __FILE__, __LINE__);
            a.minimumBaselineThreshold = v.num;
        } else if (key == "useFileSystemCache") {
            if (v.kind != JsonValue::Bool)
                throw json_exception("useFileSystemCache must be of type bool", __FILE__, __LINE__);
            a.useFileSystemCache = v.b;
        } else if (key == "mwaSupport") {
            if (v.kind != JsonValue::Bool) throw json_exception("mwaSupport must be of type bool", __FILE__, __LINE__);
            a.mwaSupport = v.b;
        } else if (key == "autoCorrelations") {
            if (v.kind != JsonValue::Bool)
                throw json_exception("readAutoCorrelations must be of type bool", __FILE__, __LINE__);
            a.readAutocorrelations = v.b;
        } else if (key == "verbosity") {
            // Arguments.cc:451-470: an integer is taken, then anything but a string is refused.
            if (v.IsInt()) a.verbosity = int(v.num);
            if (v.kind == JsonValue::String) {
                int e = 0;
                if (!ParseVerbosityName(v.str, e)) throw json_exception("invalid verbosity string", __FILE__, __LINE__);
                a.verbosity = e;
            } else {
                throw json_exception("verbosity must be of type int or string", __FILE__, __LINE__);
            }
        } else {
            throw json_exception("invalid config key: " + key, __FILE__, __LINE__);
        }
    }
    return a;
}

template <typename T>
void Over(std::optional<T>& dst, const std::optional<T>& src) {
    if (src) dst = src;
}

void Apply(Arguments& dst, const Arguments& src) {
    Over(dst.filePath, src.filePath);
    Over(dst.outputFilePath, src.outputFilePath);
    if (src.directions) {
        dst.directions = src.directions;
        dst.parsedDirections.reset();
    }
    Over(dst.parsedDirections, src.parsedDirections);
    if (src.solutionInterval) {
        dst.solutionInterval = src.solutionInterval;
        dst.parsedSlice.reset();
    }
    Over(dst.parsedSlice, src.parsedSlice);
    Over(dst.computeImplementation, src.computeImplementation);
    Over(dst.stations, src.stations);
    Over(dst.referenceAntenna, src.referenceAntenna);
    Over(dst.readAutocorrelations, src.readAutocorrelations);
    Over(dst.useFileSystemCache, src.useFileSystemCache);
    Over(dst.mwaSupport, src.mwaSupport);
    Over(dst.minimumBaselineThreshold, src.minimumBaselineThreshold);
    Over(dst.verbosity, src.verbosity);
    Over(dst.precision, src.precision);
    Over(dst.strategy, src.strategy);
}

// boost::program_options' bool spellings.
bool ParseBool(const std::string& s, const std::string& name) {
    if (s == "1" || s == "true" || s == "yes" || s == "on") return true;
    if (s == "0" || s == "false" || s == "no" || s == "off") return false;
    throw invalid_argument_exception("expected a boolean, got '" + s + "'", name, __FILE__, __LINE__);
}

long ParseInteger(const std::string& s, const std::string& name) {
    char* end = nullptr;
    const long v = std::strtol(s.c_str(), &end, 10);
    if (s.empty() || *end) throw invalid_argument_exception("expected an integer, got '" + s + "'", name, __FILE__, __LINE__);
    return v;
}

const char* kHelp =
    "LeapAccelerateCLI (gfx950):\n"
    "  -h [ --help ]                   display help message\n"
    "  -v [ --version ]                display version information\n"
    "  -c [ --config ] arg             Configuration file relative path\n"
    "  -f [ --filepath ] arg           Measurement set file path (raw measurement-set directory)\n"
    "  -o [ --output ] arg             Calibration output file path\n"
    "  -d [ --directions ] arg         Directions to calibrations\n"
    "  -s [ --stations ] arg           Overrides number of stations in measurement set\n"
    "  -r [ --referenceAntenna ] arg   Specifies the reference antenna, defaults to the last antenna\n"
    "  -i [ --implementation ] arg     Compute implementation type (hip)\n"
    "  -n [ --solutionInterval ] arg   Sets the intervals to generate solutions for, [start, interval, end]\n"
    "  -a [ --autoCorrelations ] arg   Set to true if measurement set rows store autocorrelations\n"
    "  -m [ --minimumBaselineThreshold ] arg  Minimum baseline length in meters\n"
    "  -u [ --useFileSystemCache ] arg Use filesystem caching between calls\n"
    "  --verbosity arg                 Verbosity (0=fatal, 1=error, 2=warn, 3=info, 4=debug, 5=trace)\n"
    "  --precision arg                 fp64 (default) or mixed\n"
    "  --strategy arg                  per-direction (default) or batched\n";

}  // namespace

int main(int argc, char** argv) {
    try {
        // main.cc:66-81 option table (short and long forms; "--name value" or "--name=value").
        static const std::map<std::string, std::string> shortNames{
            {"-h", "help"},     {"-v", "version"},          {"-c", "config"},         {"-f", "filepath"},
            {"-o", "output"},   {"-d", "directions"},       {"-s", "stations"},       {"-r", "referenceAntenna"},
            {"-i", "implementation"}, {"-n", "solutionInterval"}, {"-a", "autoCorrelations"},
            {"-m", "minimumBaselineThreshold"}, {"-u", "useFileSystemCache"}};
        std::map<std::string, std::string> given;
        for (int i = 1; i < argc; ++i) {
            std::string tok = argv[i], name, value;
            bool has_value = false;
            if (tok.rfind("--", 0) == 0) {
                const size_t eq = tok.find('=');
                name = tok.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
                if (eq != std::string::npos) {
                    value = tok.substr(eq + 1);
                    has_value = true;
                }
            } else {
                auto it = shortNames.find(tok);
                if (it == shortNames.end()) throw invalid_argument_exception("unrecognised option", tok, __FILE__, __LINE__);
                name = it->second;
            }
            if (name != "help" && name != "version" && !has_value) {
                if (i + 1 >= argc) throw invalid_argument_exception("missing value", name, __FILE__, __LINE__);
                value = argv[++i];
            }
            given[name] = value;
        }
        static const char* known[] = {"help", "version", "config", "filepath", "output", "directions", "stations",
                                      "referenceAntenna", "implementation", "solutionInterval", "autoCorrelations",
                                      "minimumBaselineThreshold", "useFileSystemCache", "verbosity", "precision",
                                      "strategy"};
        for (const auto& kv : given) {
            bool ok = false;
            for (const char* k : known) ok = ok || kv.first == k;
            if (!ok) throw invalid_argument_exception("unrecognised option", kv.first, __FILE__, __LINE__);
        }
        if (given.count("help")) {
            std::cout << kHelp << std::endl;
            return 0;
        }
        if (given.count("version")) {
            std::cout << "LeapAccelerateCLI version " << leap_hip_version() << std::endl;
            return 0;
        }

        Arguments cli;
        auto opt = [&](const char* k) -> const std::string* {
            auto it = given.find(k);
            return it == given.end() ? nullptr : &it->second;
        };
        if (auto v = opt("filepath")) cli.filePath = *v;
        if (auto v = opt("output")) cli.outputFilePath = *v;
        if (auto v = opt("directions")) cli.directions = *v;
        if (auto v = opt("stations")) cli.stations = int(ParseInteger(*v, "stations"));
        if (auto v = opt("referenceAntenna")) cli.referenceAntenna = static_cast<unsigned int>(ParseInteger(*v, "referenceAntenna"));
        if (auto v = opt("implementation")) cli.computeImplementation = *v;
        if (auto v = opt("solutionInterval")) cli.solutionInterval = *v;
        if (auto v = opt("autoCorrelations")) cli.readAutocorrelations = ParseBool(*v, "autoCorrelations");
        if (auto v = opt("minimumBaselineThreshold")) cli.minimumBaselineThreshold = std::strtod(v->c_str(), nullptr);
        if (auto v = opt("useFileSystemCache")) cli.useFileSystemCache = ParseBool(*v, "useFileSystemCache");
        if (auto v = opt("verbosity")) cli.verbosity = int(ParseInteger(*v, "verbosity"));
        if (auto v = opt("precision")) cli.precision = *v;
        if (auto v = opt("strategy")) cli.strategy = *v;

        // Arguments.cc:106-118: defaults, then the config file, then the command line.
        Arguments args = Defaults();
        if (auto v = opt("config")) Apply(args, ParseConfig(*v));
        Apply(args, cli);

        const std::vector<SphericalDirection> directions =
            args.parsedDirections ? *args.parsedDirections
                                  : (args.directions ? ParseDirections(*args.directions) : std::vector<SphericalDirection>{});
        const Slice slice = args.parsedSlice ? *args.parsedSlice : ParseSlice(*args.solutionInterval);
        ComputeImplementation impl;
        if (!TryParseComputeImplementation(*args.computeImplementation, impl))
            throw invalid_argument_exception("invalid compute implementation string", "implementation", __FILE__, __LINE__);
        if (directions.empty()) throw std::invalid_argument("directions argument not provided");  // Arguments.cc:240-246
        if (!args.filePath) throw std::invalid_argument("measurement set filename not provided");

        leap_options o{};
        if (*args.precision == "mixed") o.precision = LEAP_PRECISION_MIXED;
        else if (*args.precision != "fp64") throw invalid_argument_exception("fp64 or mixed", "precision", __FILE__, __LINE__);
        if (*args.strategy == "batched") o.strategy = LEAP_STRATEGY_BATCHED;
        else if (*args.strategy != "per-direction")
            throw invalid_argument_exception("per-direction or batched", "strategy", __FILE__, __LINE__);
        if (o.precision == LEAP_PRECISION_MIXED) o.strategy = LEAP_STRATEGY_BATCHED;

        const MeasurementSet ms = LoadRawMeasurementSet(*args.filePath, args.stations, *args.readAutocorrelations);
        std::unique_ptr<ILeapCalibrator> calibrator = LeapCalibratorFactory::Create(impl);
        if (auto* h = dynamic_cast<hip::HipLeapCalibrator*>(calibrator.get())) h->SetOptions(o);
        const cpu::CalibrationCollection result =
            calibrator->Calibrate(ms, directions, slice, *args.minimumBaselineThreshold, args.referenceAntenna,
                                  *args.useFileSystemCache);

        std::ofstream file;
        std::ostream* out = &std::cout;
        if (args.outputFilePath) {
            file.open(*args.outputFilePath);
            if (!file.is_open()) throw icrar::exception("failed to open file " + *args.outputFilePath + "\n", __FILE__, __LINE__);
            out = &file;
        }
        result.Serialize(*out);  // cpu::PrintResult (model/cpu/CalibrateResult.h)
        return 0;
    } catch (const std::exception& e) {
        std::cerr << e.what() << '\n';
        return -1;
    }
}
