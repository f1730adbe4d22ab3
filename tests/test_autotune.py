"""The autotuner's selection and its result check, with a stand-in job (no
GPU): the fastest candidate is kept and saved under the workload's key; a
candidate whose result digest differs from the first one's is an error,
because launch parameters may change only the time."""
import json
from types import SimpleNamespace

import pytest

from orthoroute_amd.autotune import load_tuned, tune
from orthoroute_amd.netgen import plain_64


class _Job:
    """Route time as a function of the candidate; a digest that changes for
    one bucket width when ``bad_delta`` is set."""

    def __init__(self, nl, cfg, bad_delta=None):
        self.cfg = cfg
        self.bad_delta = bad_delta
        self.restored = 0

    def run_iteration(self):
        c = self.cfg
        t = abs(c.delta - 16.0) + (1024 - c.block) / 256 + (16384 - c.lds_nodes) / 8192 \
            + abs(c.small_wave_nodes - 16384) / 4096
        return SimpleNamespace(route_ms=10.0 + t, waves=7)

    def snapshot(self):
        return "snap"

    def restore(self, snap):
        assert snap == "snap"
        self.restored += 1

    def digest(self):
        return "changed" if self.cfg.delta == self.bad_delta else "same"


def test_tuner_keeps_the_fastest_and_saves_it(tmp_path):
    nl = plain_64()
    path = tmp_path / "tuned.json"
    best = tune(nl, deltas=(8.0, 16.0, 64.0), blocks=(512, 1024), lds_nodes=(8192, 16384),
                small_wave_nodes=(4096, 16384), runner=lambda n, c: _Job(n, c), path=path, log=lambda m: None)
    assert best == {"delta": 16.0, "block": 1024, "lds_nodes": 16384, "small_wave_nodes": 16384}
    assert load_tuned(nl, path) == best
    assert list(json.loads(path.read_text())) == ["gfx950:64x64x6:200"]


def test_tuner_rejects_a_candidate_that_changes_the_result(tmp_path):
    with pytest.raises(RuntimeError, match="changed the result"):
        tune(plain_64(), deltas=(16.0, 64.0), blocks=(1024,), lds_nodes=(16384,), small_wave_nodes=(16384,),
             runner=lambda n, c: _Job(n, c, bad_delta=64.0), path=tmp_path / "t.json", log=lambda m: None)
    assert not (tmp_path / "t.json").exists()
