"""In-memory logger with the RuntimeLog data layout
(netket/_src/logging/runtime_log.py:30): ``log.data["Energy"]["Mean"]`` etc."""

from __future__ import annotations

import numpy as np


class RuntimeLog:
    def __init__(self):
        self._data = {}

    def __call__(self, step: int, item: dict, variational_state=None):
        for key, val in item.items():
            if hasattr(val, "to_dict"):
                d = self._data.setdefault(key, {"iters": []})
                d["iters"].append(step)
                for k, v in val.to_dict().items():
                    d.setdefault(k, []).append(v)
            else:
                d = self._data.setdefault(key, {"iters": [], "value": []})
                d["iters"].append(step)
                d["value"].append(val)

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
