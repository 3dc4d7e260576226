"""Loggers with the data layout and files of netket/_src/logging:

* RuntimeLog (runtime_log.py:30-190): ``log.data["Energy"]["Mean"]`` etc. in
  memory, one history per key with its ``iters``; ``serialize`` writes it as
  JSON and ``deserialize`` reads it back.
* JsonLog (the files and arguments of json_log.py:28-237): a RuntimeLog
  that keeps ``<prefix>.log`` (the same JSON) and ``<prefix>.mpack`` (the
  variational state's variables in flax msgpack encoding) on disk.  When to
  rewrite each file is decided by ``_WriteCadence`` below; both are written
  at the end of a run.
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


class _WriteCadence:
    """When to rewrite one output file of a JsonLog.

    A file is always rewritten once ``every`` logged steps are missing from
    it.  Before that it is rewritten whenever the total time spent writing
    it, plus one more write as long as the last one, stays within
    ``cost_fraction`` of the wall time since the logger started.  A file that
    is cheap to write is therefore kept current at every step, while one that
    is slow to write (large parameters) falls back to every ``every`` steps,
    and the writes never take more than about ``cost_fraction`` of a run that
    is not limited by ``every``.  The first step is always written.
    """

    def __init__(self, every: int, cost_fraction: float):
        self.every = max(int(every), 1)
        self.cost_fraction = float(cost_fraction)
        self.pending = 0
        self.spent_s = 0.0
        self.last_s = 0.0
        self.started = time.perf_counter()

    def note_step(self) -> bool:
        """Count one logged step; True when the file should be written now."""
        self.pending += 1
        if self.pending >= self.every:
            return True
        budget = self.cost_fraction * (time.perf_counter() - self.started)
        return self.spent_s + self.last_s <= budget

    def timed(self, write) -> None:
        t = time.perf_counter()
        write()
        self.last_s = time.perf_counter() - t
        self.spent_s += self.last_s
        self.pending = 0


def _replace_file(path: str, payload, binary: bool) -> None:
    """Write ``payload`` next to ``path`` and move it into place, so a reader
    (or a crash during the write) never sees a half-written file."""
    tmp = f"{path}.partial"
    with open(tmp, "wb" if binary else "w") as f:
        f.write(payload)
    os.replace(tmp, path)


_MODES = {"w": "write", "write": "write", "a": "append", "append": "append", "x": "fail",
          "fail": "fail"}


class JsonLog(RuntimeLog):
    """RuntimeLog that mirrors its histories to ``<output_prefix>.log`` and the
    variational state's variables to ``<output_prefix>.mpack``.

    ``mode``: ``"write"``/``"w"`` replaces existing files, ``"append"``/``"a"``
    continues the histories of an existing ``.log``, ``"fail"``/``"x"``
    refuses a prefix whose files exist.  ``write_every`` and
    ``save_params_every`` bound how many logged steps each file may lag
    behind; ``autoflush_cost`` is the fraction of the run time the writes of
    each file may take when they are made earlier than that."""

    def __init__(self, output_prefix, mode: str = "write", save_params_every: int = 50,
                 write_every: int = 50, save_params: bool = True, autoflush_cost: float = 0.005):
        super().__init__()
        if mode not in _MODES:
            raise ValueError(f"JsonLog mode {mode!r} is not one of 'write' ('w'), 'append' ('a') "
                             f"or 'fail' ('x')")
        self.mode = _MODES[mode]
        self.output_prefix = str(output_prefix)
        self.log_path = self.output_prefix + ".log"
        self.params_path = self.output_prefix + ".mpack"
        existing = [p for p in (self.log_path, self.params_path) if os.path.exists(p)]
        if existing and self.mode == "fail":
            raise ValueError(f"JsonLog in 'fail' mode: {', '.join(existing)} already exist(s); "
                             f"remove the file(s) or choose another output_prefix")
        folder = os.path.dirname(self.output_prefix)
        if folder:
            os.makedirs(folder, exist_ok=True)
        self.save_params = bool(save_params)
        self._log_cadence = _WriteCadence(write_every, autoflush_cost)
        self._params_cadence = _WriteCadence(save_params_every, autoflush_cost)
        if self.mode == "append" and os.path.exists(self.log_path):
            previous = RuntimeLog.deserialize(self.log_path)
            self._data, self._old_step = previous._data, previous._old_step

    def __call__(self, step: int, item: dict, variational_state=None):
        super().__call__(step, item, variational_state)
        if self._log_cadence.note_step():
            self._write_log()
        if self.save_params and variational_state is not None and self._params_cadence.note_step():
            self._write_params(variational_state)

    def _write_log(self):
        self._log_cadence.timed(
            lambda: _replace_file(self.log_path, json.dumps(_jsonable(self._data)), binary=False))

    def _write_params(self, variational_state):
        if not self.save_params or variational_state is None:
            return
        from . import serialization

        self._params_cadence.timed(lambda: _replace_file(
            self.params_path, serialization.to_bytes(variational_state.variables), binary=True))

    def flush(self, variational_state=None):
        """Write both files now (the end of a run)."""
        self._write_log()
        self._write_params(variational_state)

    def __repr__(self):
        return f"JsonLog('{self.output_prefix}', mode={self.mode})"
