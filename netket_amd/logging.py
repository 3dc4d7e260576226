"""Loggers with the data layout and files of netket/_src/logging:

* RuntimeLog (runtime_log.py:30-190): ``log.data["Energy"]["Mean"]`` etc. in
  memory, one history per key with its ``iters``; ``serialize`` writes it as
  JSON and ``deserialize`` reads it back.
* JsonLog (json_log.py:28-237): a RuntimeLog that writes ``<prefix>.log``
  (the same JSON) every ``write_every`` steps and ``<prefix>.mpack`` (the
  variational state's variables in flax msgpack encoding) every
  ``save_params_every`` steps, both also when the time since the last write
  keeps the writing cost under ``autoflush_cost`` of the run time, and at the
  end of a run.
"""

from __future__ import annotations

import json
import os
import time
from pathlib import Path

import numpy as np


def _jsonable(x):
    """Plain JSON values; complex numbers as {"real": .., "imag": ..}
    (runtime_log.py:185-201)."""
    if hasattr(x, "detach"):
        x = x.detach().cpu().numpy()
    if isinstance(x, dict):
        return {k: _jsonable(v) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        arr = np.asarray(x)
        if arr.dtype != object:
            return _jsonable(arr)
        return [_jsonable(v) for v in x]
    if isinstance(x, np.ndarray):
        if np.iscomplexobj(x):
            return {"real": x.real.tolist(), "imag": x.imag.tolist()}
        return x.tolist()
    if isinstance(x, (complex, np.complexfloating)):
        return {"real": float(x.real), "imag": float(x.imag)}
    if isinstance(x, np.generic):
        return x.item()
    return x


class RuntimeLog:
    def __init__(self):
        self._data = {}
        self._old_step = 0

    def __call__(self, step: int, item: dict, variational_state=None):
        for key, val in item.items():
            if hasattr(val, "to_compound"):
                _, fields = val.to_compound()
            elif hasattr(val, "to_dict"):
                fields = val.to_dict()
            elif isinstance(val, dict):
                fields = val
            else:
                fields = {"value": val}
            d = self._data.setdefault(key, {"iters": []})
            d["iters"].append(step)
            for k, v in fields.items():
                d.setdefault(k, []).append(v)
        self._old_step = step

    @property
    def data(self) -> dict:
        out = {}
        for key, d in self._data.items():
            out[key] = {k: np.asarray(v) for k, v in d.items()}
        return out

    def __getitem__(self, key):
        return self.data[key]

    def flush(self, variational_state=None):
        pass

    # -- files -------------------------------------------------------------
    def serialize(self, path):
        """Write the data as JSON; ``.json`` is appended unless the name ends in
        ``.log`` or ``.json`` (runtime_log.py:124-150)."""
        if isinstance(path, (str, Path)):
            path = Path(path)
            if not path.name.endswith((".log", ".json")):
                path = path.parent / (path.name + ".json")
            with open(path, "w") as f:
                json.dump(_jsonable(self._data), f)
        else:
            path.write(json.dumps(_jsonable(self._data)))

    @classmethod
    def deserialize(cls, path) -> "RuntimeLog":
        """Read a file written by ``serialize`` (runtime_log.py:76-122)."""
        path = Path(path)
        if not path.name.endswith((".log", ".json")):
            path = path.parent / (path.name + ".json")
        with open(path) as f:
            raw = json.load(f)
        log = cls.__new__(cls)
        log._data = {}
        last = 0
        for key, d in raw.items():
            log._data[key] = {k: list(v) if isinstance(v, list) else v for k, v in d.items()}
            iters = d.get("iters", [])
            if iters:
                last = max(last, int(iters[-1]))
        log._old_step = last
        return log

    def __repr__(self):
        return f"RuntimeLog():\n keys = {list(self._data.keys())}"


class JsonLog(RuntimeLog):
    def __init__(self, output_prefix, mode: str = "write", save_params_every: int = 50,
                 write_every: int = 50, save_params: bool = True, autoflush_cost: float = 0.005):
        super().__init__()
        mode = {"w": "write", "a": "append", "x": "fail"}.get(mode, mode)
        if mode not in ("write", "append", "fail"):
            raise ValueError("Mode not recognized: should be one of `[w]rite`, `[a]ppend` or`[x]`(fail).")
        output_prefix = str(output_prefix)
        exists = os.path.exists(output_prefix + ".log") or os.path.exists(output_prefix + ".mpack")
        if exists and mode == "fail":
            raise ValueError("Output file already exists. Either delete it manually orchange `output_prefix`.")
        dir_name = os.path.dirname(output_prefix)
        if dir_name:
            os.makedirs(dir_name, exist_ok=True)
        self._prefix = output_prefix
        self._file_mode = mode
        self._write_every = write_every
        self._save_params_every = save_params_every
        self._save_params = save_params
        self._steps_notflushed_write = 0
        self._steps_notflushed_pars = 0
        self._autoflush_cost = autoflush_cost
        self._last_flush_time = time.time()
        self._last_flush_runtime = 0.0
        self._last_flush_pars_time = time.time()
        self._last_flush_pars_runtime = 0.0
        if mode == "append" and os.path.exists(output_prefix + ".log"):
            existing = RuntimeLog.deserialize(output_prefix + ".log")
            self._data = existing._data
            self._old_step = existing._old_step

    def __call__(self, step: int, item: dict, variational_state=None):
        old_step = self._old_step
        super().__call__(step, item, variational_state)
        elapsed = time.time() - self._last_flush_time
        flush_anyway = self._last_flush_runtime / (elapsed + 1e-7) < self._autoflush_cost
        if self._steps_notflushed_write % self._write_every == 0 or step == old_step - 1 or flush_anyway:
            self._flush_log()
        elapsed = time.time() - self._last_flush_pars_time
        flush_anyway = self._last_flush_pars_runtime / (elapsed + 1e-7) < self._autoflush_cost
        if self._steps_notflushed_pars % self._save_params_every == 0 or step == old_step - 1 or flush_anyway:
            self._flush_params(variational_state)
        self._old_step = step
        self._steps_notflushed_write += 1
        self._steps_notflushed_pars += 1

    def _flush_log(self):
        self._last_flush_time = time.time()
        self.serialize(self._prefix + ".log")
        self._last_flush_runtime = time.time() - self._last_flush_time
        self._steps_notflushed_write = 0

    def _flush_params(self, variational_state):
        if not self._save_params or variational_state is None:
            return
        from . import serialization

        self._last_flush_pars_time = time.time()
        data = serialization.to_bytes(variational_state.variables)
        with open(self._prefix + ".mpack", "wb") as f:
            f.write(data)
        self._last_flush_pars_runtime = time.time() - self._last_flush_pars_time

    def flush(self, variational_state=None):
        """Write both files now (end of a run)."""
        self._flush_log()
        if variational_state is not None:
            self._flush_params(variational_state)

    def __repr__(self):
        return f"JsonLog('{self._prefix}', mode={self._file_mode})"
