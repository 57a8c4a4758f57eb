"""The batch autotuner's cache logic, on the CPU: keys name the device and
the library build, a hit returns the stored winner without timing, a miss
times every candidate once and stores the fastest. The timer and the plans
are stand-ins here (no GPU); the GPU test of a tuned plan is in
tests/test_gpu_batch.py."""

import json

import numpy as np
import pytest

import igl
from igl import _autotune, batch


class FakePlan:
    made = []

    def __init__(self, counts, queries, **config):
        self.config = config
        FakePlan.made.append(config)

    def run(self, *args, **kwargs):
        return None

    def close(self):
        pass


@pytest.fixture
def tuner(tmp_path, monkeypatch):
    monkeypatch.setenv("WN_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(_autotune, "device_name", lambda: "devA/gfx950")
    monkeypatch.setattr(_autotune, "library_build", lambda: "abi2.123.456")
    monkeypatch.setattr(batch, "WindingBatch", FakePlan)
    monkeypatch.setattr(batch, "StripBatch", FakePlan)
    timed = []

    def fake_time(fn, repeats):
        fn()
        cfg = FakePlan.made[-1]
        timed.append(cfg)
        # fastest: qpl 2 with chunk 4096; chunk_tiles 8
        if "chunk_faces" in cfg:
            return abs(cfg["chunk_faces"] - 4096) / 1024 + (0 if cfg["qpl"] == 2 else 0.5)
        return abs(cfg["chunk_tiles"] - 8) + 1.0

    monkeypatch.setattr(_autotune, "_time", fake_time)
    FakePlan.made = []
    return timed, tmp_path / "tune.json"


def test_batch_tuner_caches_by_device_and_build(tuner, monkeypatch):
    timed, path = tuner
    fc, qc = [1000, 5000, 200], [64, 4096, 300]
    best = _autotune.tune_batch(fc, qc, None, None, 1)
    assert best == (2, 4096)
    assert len(timed) == 2 * 4  # every (qpl, chunk) candidate once
    data = json.loads(path.read_text())
    (key,) = data
    assert key.startswith("devA/gfx950|abi2.123.456|batch:p1:")
    assert data[key]["config"] == [2, 4096] and len(data[key]["tried"]) == 8
    # a hit: no timing
    assert _autotune.tune_batch(fc, qc, None, None, 1) == (2, 4096)
    assert len(timed) == 8
    # another precision, another device, another build: misses
    _autotune.tune_batch(fc, qc, None, None, 0)
    assert len(timed) == 16
    monkeypatch.setattr(_autotune, "device_name", lambda: "devB/gfx942")
    _autotune.tune_batch(fc, qc, None, None, 1)
    assert len(timed) == 24
    monkeypatch.setattr(_autotune, "library_build", lambda: "abi2.999.1")
    _autotune.tune_batch(fc, qc, None, None, 1)
    assert len(timed) == 32
    assert len(json.loads(path.read_text())) == 4


def test_strip_tuner_and_no_cache(tuner):
    timed, path = tuner
    rc, qc = [255 * 4, 255 * 40], [100, 1000]
    assert _autotune.tune_batch_strips(rc, qc, None, None) == 8
    assert len(timed) == 5
    assert _autotune.tune_batch_strips(rc, qc, None, None) == 8
    assert len(timed) == 5
    assert _autotune.tune_batch_strips(rc, qc, None, None, precision="fp64") == 8
    assert len(timed) == 10
    keys = list(json.loads(path.read_text()))
    assert any(":p1:" in k for k in keys) and any(":p0:" in k for k in keys)
    # use_cache=False times again and writes nothing new
    before = path.read_text()
    assert _autotune.tune_batch_strips(rc, qc, None, None, use_cache=False) == 8
    assert len(timed) == 15 and path.read_text() == before


def test_ties_keep_the_first_candidate(tuner, monkeypatch):
    monkeypatch.setattr(_autotune, "_time", lambda fn, repeats: 1.0)
    assert _autotune.tune_batch_strips([255], [1], None, None, chunks=(16, 4, 2)) == 16


def test_real_key_names_the_library(monkeypatch):
    """library_build() reads the ABI version through the C ABI (no GPU)."""
    build = _autotune.library_build()
    assert build.startswith(f"abi{igl.lib().wn_version()}.")
    assert build.split(".")[1] != "nofile"
    assert isinstance(_autotune.device_name(), str)
    assert np.isfinite(_autotune._bucket(0))
