"""On-disk formats: flax-style msgpack bytes (netket/vqs/mc/mc_state/state.py:963-1016)
and the JsonLog / RuntimeLog files (netket/_src/logging/json_log.py:28-237,
runtime_log.py:30-201).  CPU only."""

import json

import numpy as np
import pytest

from netket_amd import logging as nklog
from netket_amd import serialization as ser
from netket_amd.stats import Stats


def test_msgpack_known_bytes():
    # msgpack spec: fixmap(1) 0x81, fixstr "a" 0xa1 0x61, ext8 0xc7 <len> <type 1>, payload =
    # fixarray(3) 0x93 [fixarray(1) 0x91 0x02, fixstr(4) "int8", bin8 0xc4 0x02 0x01 0x02]
    b = ser.msgpack_serialize({"a": np.array([1, 2], dtype=np.int8)})
    expect = bytes([0x81, 0xA1, 0x61, 0xC7, 0x0C, 0x01, 0x93, 0x91, 0x02, 0xA4]) + b"int8" + bytes(
        [0xC4, 0x02, 0x01, 0x02])
    assert b == expect
    # a numpy scalar is extension type 3 with a 0-d array payload: shape [] is 0x90
    b = ser.msgpack_serialize({"n": np.int64(5)})
    payload = bytes([0x93, 0x90, 0xA5]) + b"int64" + bytes([0xC4, 0x08, 5, 0, 0, 0, 0, 0, 0, 0])
    assert b == bytes([0x81, 0xA1, 0x6E, 0xC7, len(payload), 0x03]) + payload


def test_msgpack_round_trip():
    rng = np.random.default_rng(0)
    tree = {"params": {"Dense": {"kernel": rng.normal(size=(5, 7)).astype(np.float32),
                                 "bias": rng.normal(size=7)},
                       "visible_bias": np.arange(5, dtype=np.float64)},
            "n": np.int64(12), "c": complex(1.5, -2.0), "none": None, "s": "text", "i": 3}
    back = ser.msgpack_restore(ser.msgpack_serialize(tree))
    k = back["params"]["Dense"]["kernel"]
    assert k.dtype == np.float32 and k.shape == (5, 7)
    np.testing.assert_array_equal(k, tree["params"]["Dense"]["kernel"])
    np.testing.assert_array_equal(back["params"]["Dense"]["bias"], tree["params"]["Dense"]["bias"])
    np.testing.assert_array_equal(back["params"]["visible_bias"], tree["params"]["visible_bias"])
    assert back["n"] == 12 and back["c"] == complex(1.5, -2.0) and back["none"] is None
    assert back["s"] == "text" and back["i"] == 3


def _stats(m):
    return Stats(mean=m, error_of_mean=0.1, variance=2.0, tau_corr=0.5, R_hat=1.01)


def test_json_log_files(tmp_path):
    prefix = tmp_path / "run"
    log = nklog.JsonLog(prefix, write_every=1, save_params=False)
    for step in range(3):
        log(step, {"Energy": _stats(-1.0 - step), "acc": 0.5 + 0.1 * step, "z": complex(step, 1.0)})
    log.flush()
    data = json.load(open(str(prefix) + ".log"))
    assert data["Energy"]["iters"] == [0, 1, 2]
    assert data["Energy"]["Mean"] == [-1.0, -2.0, -3.0]
    assert set(data["Energy"]) == {"iters", "Mean", "Variance", "Sigma", "R_hat", "TauCorr"}
    assert data["acc"]["iters"] == [0, 1, 2] and np.allclose(data["acc"]["value"], [0.5, 0.6, 0.7])
    assert data["z"]["value"] == {"real": [0.0, 1.0, 2.0], "imag": [1.0, 1.0, 1.0]}
    # fail mode refuses an existing prefix; append continues from the file
    with pytest.raises(ValueError):
        nklog.JsonLog(prefix, mode="x")
    log2 = nklog.JsonLog(prefix, mode="a", write_every=1, save_params=False)
    log2(3, {"Energy": _stats(-4.0)})
    log2.flush()
    data = json.load(open(str(prefix) + ".log"))
    assert data["Energy"]["iters"] == [0, 1, 2, 3]
    assert data["Energy"]["Mean"][-1] == -4.0
    with pytest.raises(ValueError):
        nklog.JsonLog(prefix, mode="bad")


def test_runtime_log_serialize_round_trip(tmp_path):
    log = nklog.RuntimeLog()
    for step in range(4):
        log(step, {"Energy": _stats(0.5 * step)})
    log.serialize(tmp_path / "out")          # ".json" is appended
    back = nklog.RuntimeLog.deserialize(tmp_path / "out.json")
    np.testing.assert_array_equal(back["Energy"]["Mean"], log["Energy"]["Mean"])
    np.testing.assert_array_equal(back["Energy"]["iters"], [0, 1, 2, 3])
    assert back._old_step == 3


class _StubState:
    """Only what JsonLog reads from a variational state: its variables."""

    def __init__(self):
        self.step = -1

    @property
    def variables(self):
        return {"params": {"step": np.array([self.step], dtype=np.int64)}}


def test_json_log_write_cadence(tmp_path):
    # with no time budget for early writes, each file is written at the first
    # step and then whenever `every` logged steps are missing from it
    prefix = str(tmp_path / "cad")
    log = nklog.JsonLog(prefix, write_every=2, save_params_every=3, autoflush_cost=0.0)
    vs = _StubState()
    seen_log, seen_params = [], []
    for step in range(10):
        vs.step = step
        log(step, {"Energy": _stats(float(step))}, vs)
        seen_log.append(json.load(open(prefix + ".log"))["Energy"]["iters"][-1])
        seen_params.append(int(ser.msgpack_restore(open(prefix + ".mpack", "rb").read())["params"]["step"][0]))
    assert seen_log == [0, 0, 2, 2, 4, 4, 6, 6, 8, 8]
    assert seen_params == [0, 0, 0, 3, 3, 3, 6, 6, 6, 9]
    vs.step = 10
    log.flush(vs)
    assert json.load(open(prefix + ".log"))["Energy"]["iters"][-1] == 9
    assert int(ser.msgpack_restore(open(prefix + ".mpack", "rb").read())["params"]["step"][0]) == 10
    # a generous budget keeps both files current at every step
    log2 = nklog.JsonLog(str(tmp_path / "eager"), write_every=1000, save_params_every=1000,
                         autoflush_cost=1e9)
    for step in range(4):
        vs.step = step
        log2(step, {"Energy": _stats(float(step))}, vs)
        assert json.load(open(str(tmp_path / "eager") + ".log"))["Energy"]["iters"][-1] == step
    assert not list(tmp_path.glob("*.partial"))
