"""The raw measurement-set directory, its loader (icrar::LoadRawMeasurementSet,
driven through hip_calibrator_main msinfo), the casacore converter
(ms_to_raw.py, against a stand-in for python-casacore's table API) and the
command line's argument handling up to the point where the device is used.

The loader follows the reference's MeasurementSet constructor: stations from
the unique ANTENNA1 values (MeasurementSet.cc:287-294 reads ANTENNA1 twice),
baselines from stations and the autocorrelation switch (:154-167), epochs
(:120-130), first-epoch flags (:181-197) and the validation of :64-103.
"""
import json
import os
import subprocess
import sys
import types

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "icrar-leap-accelerate_amd")
MAIN = os.path.join(PKG, "build", "hip_calibrator_main")
CLI = os.path.join(PKG, "build", "LeapAccelerateCLI")


def msinfo(path, stations=-1, autos=1):
    out = subprocess.run([MAIN, "msinfo", str(path), str(stations), str(autos)], capture_output=True, text=True,
                         timeout=60)
    return out.returncode, (json.loads(out.stdout) if out.returncode == 0 else out.stderr)


@pytest.fixture(scope="module")
def synth_ms(leap):
    s = leap.SynthMS(n_antennas=10, n_timesteps=4, n_channels=6, nan_fraction=0.01, flag_fraction=0.2, seed=5)
    s.fill_host(np.complex64)
    return s, s.host_dict()


def test_raw_ms_round_trip(leap, synth_ms, tmp_path):
    s, ms = synth_ms
    leap.write_raw_ms_from_dict(tmp_path / "a", ms)
    rc, info = msinfo(tmp_path / "a")
    assert rc == 0, info
    assert info["stations"] == 10 and info["total_antennas"] == 10
    assert info["baselines"] == 55 and info["rows"] == 220 and info["timesteps"] == 4
    assert info["channels"] == 6 and info["pols"] == 4 and info["vis_dtype"] == leap.LEAP_VIS_C64
    assert info["flagged_baselines"] == int(np.count_nonzero(ms["flag"]))
    assert info["freq_start_hz"] == ms["freq_start_hz"] and info["freq_inc_hz"] == ms["freq_inc_hz"]
    assert info["phase_centre"] == list(ms["phase_centre"])
    assert info["epochs"] == [ms["time"][t * 55] for t in range(4)]
    v = ms["vis"].view(np.float32).astype(np.float64)
    assert info["vis_sum"] == pytest.approx(np.nansum(np.where(np.isfinite(v), v, 0.0)), rel=1e-12)


def test_raw_ms_stations_rule_and_validation(leap, tmp_path):
    """Without autocorrelation rows ANTENNA1 never holds the last antenna, so the reference counts one
    station fewer and its validation refuses the set unless the station count is given (-s)."""
    s = leap.SynthMS(n_antennas=8, autocorrelations=0, n_timesteps=3, n_channels=4, seed=6)
    s.fill_host(np.complex64)
    ms = s.host_dict()
    leap.write_raw_ms_from_dict(tmp_path / "b", ms)
    rc, err = msinfo(tmp_path / "b", -1, 0)
    assert rc != 0 and "visibilities at first epoch does not match number of baselines" in err
    rc, info = msinfo(tmp_path / "b", 8, 0)
    assert rc == 0, info
    assert info["stations"] == 8 and info["baselines"] == 28 and info["timesteps"] == 3
    # rows that are not a whole number of epochs
    d = dict(ms)
    keep = ms["num_rows"] - 5
    leap.write_raw_ms(tmp_path / "c", antenna1=np.tile(ms["antenna1"], 3)[:keep],
                      antenna2=np.tile(ms["antenna2"], 3)[:keep], time=ms["time"][:keep],
                      flag_c0p0=np.tile(ms["flag"], 3)[:keep], uvw=ms["uvw"][:keep], data=ms["vis"][:keep],
                      antenna_table_rows=8, ref_frequency_hz=d["freq_start_hz"], chan_width_hz=d["freq_inc_hz"],
                      phase_dir_rad=d["phase_centre"])
    rc, err = msinfo(tmp_path / "c", 8, 0)
    assert rc != 0 and "number of rows not an integer multiple of baselines" in err
    # a truncated data file is refused
    with open(tmp_path / "b" / "DATA.c64", "r+b") as f:
        f.truncate(100)
    rc, err = msinfo(tmp_path / "b", 8, 0)
    assert rc != 0 and "data column holds fewer than" in err


class _StandInTable:
    """The subset of python-casacore's table API that ms_to_raw.py uses."""

    def __init__(self, cols, nrows):
        self.cols, self.n = cols, nrows

    def nrows(self):
        return self.n

    def getcol(self, name, startrow=0, nrow=-1):
        c = self.cols[name]
        return c[startrow:] if nrow < 0 else c[startrow:startrow + nrow]

    def close(self):
        pass


def test_converter_maps_the_measurement_set_columns(leap, synth_ms, tmp_path, monkeypatch):
    s, ms = synth_ms
    R, C, P = ms["vis"].shape
    nbl = ms["num_baselines"]
    flags = np.zeros((R, C, P), dtype=bool)
    flags[:, 0, 0] = np.tile(ms["flag"].astype(bool), R // nbl)
    flags[:, 1:, :] = True  # only [channel 0][correlation 0] is read
    tabs = {
        "x.ms": _StandInTable({"ANTENNA1": np.tile(ms["antenna1"], R // nbl), "ANTENNA2": np.tile(ms["antenna2"], R // nbl),
                               "TIME": ms["time"], "UVW": ms["uvw"].reshape(R, 3), "FLAG": flags, "DATA": ms["vis"]}, R),
        os.path.join("x.ms", "ANTENNA"): _StandInTable({}, 10),
        os.path.join("x.ms", "SPECTRAL_WINDOW"): _StandInTable(
            {"NUM_CHAN": np.array([C]), "REF_FREQUENCY": np.array([ms["freq_start_hz"]]),
             "CHAN_WIDTH": np.array([[ms["freq_inc_hz"]] * C])}, 1),
        os.path.join("x.ms", "POLARIZATION"): _StandInTable({"NUM_CORR": np.array([P])}, 1),
        os.path.join("x.ms", "FIELD"): _StandInTable({"PHASE_DIR": np.array([[list(ms["phase_centre"])]])}, 1),
    }
    fake_tables = types.ModuleType("casacore.tables")
    fake_tables.table = lambda path, readonly=True, ack=False: tabs[path]
    fake_casacore = types.ModuleType("casacore")
    fake_casacore.tables = fake_tables
    monkeypatch.setitem(sys.modules, "casacore", fake_casacore)
    monkeypatch.setitem(sys.modules, "casacore.tables", fake_tables)
    monkeypatch.syspath_prepend(PKG)
    import ms_to_raw

    h = ms_to_raw.convert("x.ms", str(tmp_path / "conv"), chunk_rows=37)
    assert h["rows"] == R and h["channels"] == C and h["pols"] == P and h["data_type"] == "complex64"
    rc, info = msinfo(tmp_path / "conv")
    assert rc == 0, info
    leap.write_raw_ms_from_dict(tmp_path / "direct", ms)
    rc2, info2 = msinfo(tmp_path / "direct")
    assert info == info2
    for name in ("ANTENNA1.i32", "ANTENNA2.i32", "TIME.f64", "FLAG_C0P0.u8", "UVW.f64", "DATA.c64"):
        assert open(tmp_path / "conv" / name, "rb").read() == open(tmp_path / "direct" / name, "rb").read(), name


def run_cli(*args, timeout=60):
    out = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    return out.returncode, out.stdout, out.stderr


def test_cli_arguments_before_the_device(leap, synth_ms, tmp_path):
    """Argument and configuration errors of the reference's CLI, with its messages (host only)."""
    s, ms = synth_ms
    leap.write_raw_ms_from_dict(tmp_path / "m", ms)
    rc, out, err = run_cli("--help")
    assert rc == 0 and "--solutionInterval" in out
    rc, out, err = run_cli("-f", tmp_path / "m")
    assert rc == 255 and "directions argument not provided" in err
    rc, out, err = run_cli("-d", "[[0.5,-0.4]]")
    assert rc == 255 and "measurement set filename not provided" in err
    rc, out, err = run_cli("-f", tmp_path / "m", "-d", "[[0.5,-0.4]]", "-i", "cpu")
    assert rc == 255 and "cpu implementation is not part of this library" in err
    rc, out, err = run_cli("-f", tmp_path / "m", "-d", "[[0.5,-0.4]]", "-i", "cuda")
    assert rc == 255 and "cuda build option not enabled" in err
    rc, out, err = run_cli("-f", tmp_path / "m", "-d", "[0.5,-0.4]")
    assert rc == 255 and "expected an array of 2 numbers" in err
    rc, out, err = run_cli("-f", tmp_path / "m", "-d", "[[0.5,-0.4]]", "-n", "[0,1]")
    assert rc == 255 and "expected 3 integers" in err
    rc, out, err = run_cli("-f", tmp_path / "m", "-d", "[[0.5,-0.4]]", "-n", "[0,0,-1]")
    assert rc == 255 and "expected a non zero integer" in err
    cfg = tmp_path / "c.json"
    for body, msg in [({"filePath": str(tmp_path / "m"), "bogus": 1}, "invalid config key: bogus"),
                      ({"configFilePath": "x"}, "recursive config detected"),
                      ({"minimumBaselineThreshold": 0}, "minimumBaselineThreshold must be of type double"),
                      ({"verbosity": 3}, "verbosity must be of type int or string"),
                      ({"computeImplementation": "gpu"}, "invalid compute implementation string"),
                      ({"autoCorrelations": 1}, "readAutoCorrelations must be of type bool")]:
        cfg.write_text(json.dumps(body))
        rc, out, err = run_cli("-c", cfg, "-d", "[[0.5,-0.4]]")
        assert rc == 255 and msg in err, (body, err)
    rc, out, err = run_cli("--bogus", "1")
    assert rc == 255 and "unrecognised option" in err
