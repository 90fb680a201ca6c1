"""The reference's end-to-end MWA goldens (PhaseRotateTests.cc:94-143, 420-428;
expected values in tests/golden/mwa_calibrations.json, copied as data from
PhaseRotateTestCaseData.h by tests/golden/extract_mwa_goldens.py).

The measurement set they belong to (testdata/mwa/1197638568-split.ms) is
downloaded by the reference's testdata/install.sh and is not present here.  The
CPU tests check the fixture against the call it belongs to; the GPU test runs
the reference's four cases through the raw measurement-set path and
LeapAccelerateCLI when a converted copy of that set is available:

    python3 icrar-leap-accelerate_amd/ms_to_raw.py 1197638568-split.ms mwa.raw
    LEAP_MWA_RAW=$PWD/mwa.raw python -m pytest tests/test_mwa_goldens.py -m gpu

and is skipped otherwise.  Comparison as the reference's: Eigen isApprox at
1e-11, ||got - want|| <= 1e-11 min(||got||, ||want||) per beam.
"""
import json
import os
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIXTURE = os.path.join(REPO, "tests", "golden", "mwa_calibrations.json")
CLI = os.path.join(REPO, "icrar-leap-accelerate_amd", "build", "LeapAccelerateCLI")


@pytest.fixture(scope="module")
def goldens():
    with open(FIXTURE) as f:
        return json.load(f)


def test_fixture_matches_the_reference_call(goldens, oracle):
    assert goldens["stations"] == 102 and goldens["read_autocorrelations"] and goldens["reference_antenna"] == 0
    dirs = goldens["directions"]
    assert len(dirs) == 4
    T = 14  # timesteps of 1197638568-split.ms (each_timestep holds one solution per timestep)
    for case in goldens["cases"]:
        coll = goldens["collections"][case["expected"]]
        n_solutions = oracle.slice_solutions(*case["slice"], T)
        assert len(coll) == n_solutions, case
        for cal in coll:
            for d, beam in enumerate(cal["beams"]):
                assert beam["direction"] == dirs[d]
                # 128 antennas: max(antenna1, antenna2) + 1 columns of A
                assert len(beam["calibration"]) == 128
                assert all(np.isfinite(beam["calibration"]))
    # the reference antenna's column is 0 up to rounding in every expected beam
    for coll in goldens["collections"].values():
        for cal in coll:
            for beam in cal["beams"]:
                assert abs(beam["calibration"][0]) < 1e-12


def is_approx(got, want, prec):
    got, want = np.asarray(got), np.asarray(want)
    return np.linalg.norm(got - want) <= prec * min(np.linalg.norm(got), np.linalg.norm(want))


@pytest.mark.gpu
@pytest.mark.skipif(not os.environ.get("LEAP_MWA_RAW"),
                    reason="the MWA measurement set (testdata/mwa/1197638568-split.ms) is not present; "
                           "convert it with ms_to_raw.py and set LEAP_MWA_RAW")
@pytest.mark.parametrize("case", range(4), ids=["first-timestep", "all-timesteps-0", "all-timesteps-1", "each-timestep"])
def test_mwa_goldens_through_the_cli(goldens, tmp_path, case):
    c = goldens["cases"][case]
    out_file = tmp_path / "cal.json"
    args = [CLI, "-f", os.environ["LEAP_MWA_RAW"], "-s", str(goldens["stations"]), "-a", "true",
            "-r", str(goldens["reference_antenna"]), "-d", json.dumps(goldens["directions"]),
            "-n", json.dumps(c["slice"]), "-o", str(out_file)]
    out = subprocess.run(args, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr
    got = json.loads(out_file.read_text())
    want = goldens["collections"][c["expected"]]
    assert len(got) == len(want)
    for sol, exp in zip(got, want):
        assert len(sol["calibration"]) == len(goldens["directions"])
        for d, beam in enumerate(exp["beams"]):  # expected collections may list fewer beams
            assert sol["calibration"][d]["direction"] == beam["direction"]
            assert is_approx(sol["calibration"][d]["beamCalibration"], beam["calibration"],
                             goldens["is_approx_threshold"]), (c["test"], d)
