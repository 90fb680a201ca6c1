"""The autotuner's choices saved as JSON and reloaded into another context."""
import json

import numpy as np
import pytest

import np_restatement as npr

pytestmark = pytest.mark.gpu


def test_tuning_round_trip_and_use(leap, ctx, oracle):
    s = leap.SynthMS(n_antennas=24, n_timesteps=6, n_channels=40, seed=97)
    s.fill_host()
    ms = s.host_dict()
    view = leap.view_from_arrays(ms)
    rng = np.random.default_rng(1)
    c = ms["phase_centre"]
    dirs = np.column_stack([c[0] + rng.uniform(-.02, .02, 150), c[1] + rng.uniform(-.02, .02, 150)])
    fresh = leap.Context(0)
    try:
        md = fresh.metadata(view)
        fresh.calibrate(view, dirs, opts=leap.options(strategy=1), md=md)
        chosen = fresh.last_direction_block()
        doc = json.loads(fresh.tuning_json())
        assert doc["format"] == "leap-hip-tuning" and doc["version"] == 1
        entries = [e for e in doc["batched"] if e["directions"] == 150 and e["channels"] == 40]
        assert len(entries) == 1 and entries[0]["tile"] == chosen
        # another context takes a (different, valid) choice from the document without tuning
        other = {641: 1282, 1282: 2564, 2564: 641}[chosen]
        entries[0]["tile"] = other
        second = leap.Context(0)
        try:
            second.load_tuning(json.dumps(doc))
            md2 = second.metadata(view)
            cal, _ = second.calibrate(view, dirs, opts=leap.options(strategy=1), md=md2)
            assert second.last_direction_block() == other
            want, _ = oracle.calibrate(ms, dirs)
            assert np.abs(npr.wrap(cal - want)).max() < 1e-9
            assert json.loads(second.tuning_json())["batched"][0]["tile"] == other
            # a broken or inconsistent document is refused whole
            with pytest.raises(leap.LeapError, match="tuning JSON"):
                second.load_tuning('{"format": "leap-hip-tuning", "version": 1, "batched": [{"mode": 1')
            bad = json.loads(json.dumps(doc))
            bad["batched"][0]["tile"] = 2562  # not a packed-path tile
            with pytest.raises(leap.LeapError, match="not a packed-path tile"):
                second.load_tuning(json.dumps(bad))
            assert json.loads(second.tuning_json())["batched"][0]["tile"] == other
        finally:
            second.close()
    finally:
        fresh.close()
