#!/usr/bin/env python3
"""Writes mwa_calibrations.json: the expected calibrations that the reference's
end-to-end tests hold for the MWA measurement set 1197638568-split.ms
(src/icrar/leap-accelerate/tests/algorithm/PhaseRotateTestCaseData.h:35-1150),
with the call each expectation belongs to (PhaseRotateTests.cc:79-83, 94-114,
420-428: 102 stations with autocorrelations, reference antenna 0, threshold
0.0, four directions, Eigen isApprox at 1e-11).

The numbers are copied as data; nothing of the reference is run.  Run here, where
the reference tree is mounted:  python3 tests/golden/extract_mwa_goldens.py
"""
import json
import os
import re
import sys

SRC = "/root/reference/src/icrar/leap-accelerate/tests/algorithm/PhaseRotateTestCaseData.h"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mwa_calibrations.json")

NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"


def parse_function(text, name):
    start = text.index(name + "()")
    nxt = text.find("cpu::CalibrationCollection Get", start + len(name))
    body = text[start: nxt if nxt >= 0 else len(text)]
    # named calibrations: auto <var> = cpu::Calibration(start, end, { BeamCalibration(...) ... });
    named = {}
    for m in re.finditer(r"auto\s+(\w+)\s*=\s*cpu::Calibration\(\s*(" + NUM + r")\s*,\s*(" + NUM + r")\s*,", body):
        var, s0, s1 = m.group(1), float(m.group(2)), float(m.group(3))
        end = body.find("});", m.end())
        chunk = body[m.end(): end]
        beams = []
        for b in re.finditer(r"SphericalDirection\(\s*(" + NUM + r")\s*,\s*(" + NUM + r")\s*\)\s*,\s*ToVector\(std::vector<double>\s*\{([^}]*)\}",
                             chunk):
            vals = [float(x) for x in re.findall(NUM, b.group(3))]
            beams.append({"direction": [float(b.group(1)), float(b.group(2))], "calibration": vals})
        named[var] = {"epoch": [s0, s1], "beams": beams}
    # the collection: return cpu::CalibrationCollection(...{ var, ..., cpu::Calibration(0.0, 0.0, {}) ... })
    r0 = body.index("return cpu::CalibrationCollection")
    ret = body[r0: body.index(";", r0)]
    items = []
    for tok in re.finditer(r"cpu::Calibration\(\s*(" + NUM + r")\s*,\s*(" + NUM + r")\s*,\s*\{\s*\}\s*\)|\b(calibration\w*)\b", ret):
        if tok.group(3):
            items.append(named[tok.group(3)])
        else:
            items.append({"epoch": [float(tok.group(1)), float(tok.group(2))], "beams": []})
    return items


def main():
    text = open(SRC).read()
    out = {
        "_about": "Expected calibrations of the reference's MWA end-to-end tests (PhaseRotateTestCaseData.h, "
                  "ska-telescope/icrar-leap-accelerate v0.8.0), copied as data by extract_mwa_goldens.py. "
                  "A collection may list fewer beams than directions (only those are compared).",
        "measurement_set": "testdata/mwa/1197638568-split.ms (downloaded by the reference's testdata/install.sh)",
        "stations": 102,
        "read_autocorrelations": True,
        "reference_antenna": 0,
        "min_baseline_threshold": 0.0,
        "directions": [[-0.4606549305661674, -0.29719233792392513], [-0.753231018062671, -0.44387635324622354],
                       [-0.4606549305661674, -0.29719233792392513], [-0.753231018062671, -0.44387635324622354]],
        "is_approx_threshold": 1e-11,
        "cases": [
            # Slice(0, 1), Slice(0, 14), Slice(0, -1), Slice(1) as (start, interval, end) (Slice.cc:27-61)
            {"test": "PhaseRotateFirstTimestepTestCpu", "slice": [0, 1, 1], "expected": "first_timestep"},
            {"test": "PhaseRotateAllTimesteps0TestCpu", "slice": [0, 14, 14], "expected": "all_timesteps"},
            {"test": "PhaseRotateAllTimesteps1TestCpu", "slice": [0, -1, -1], "expected": "all_timesteps"},
            {"test": "PhaseRotateEachTimestepTestCpu", "slice": [0, 1, -1], "expected": "each_timestep"},
        ],
        "collections": {
            "first_timestep": parse_function(text, "GetFirstTimestepMWACalibration"),
            "all_timesteps": parse_function(text, "GetAllTimestepsMWACalibration"),
            "each_timestep": parse_function(text, "GetEachTimestepMWACalibration"),
        },
    }
    for k, v in out["collections"].items():
        print(k, len(v), "calibrations;", [len(c["beams"]) for c in v][:4], "beams;",
              [len(b["calibration"]) for c in v for b in c["beams"]][:4], "antennas", file=sys.stderr)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
