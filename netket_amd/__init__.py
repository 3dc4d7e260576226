"""netket_amd — NetKet's VMC step on AMD MI355X (gfx950).

A NetKet-compatible host API (hilbert, graph, operator, models, sampler, vqs,
optimizer, driver, stats, logging) over the C ABI of hand-written gfx950 HIP
kernels in ``libnetket_amd.so`` (``include/netket_amd.h``).  PyTorch provides
device memory and the current stream only.
"""

from . import graph, hilbert, logging, models, operator, optimizer, sampler, serialization, stats, vqs
from . import driver
from .driver import VMC, VMC_SR

__all__ = ["graph", "hilbert", "logging", "models", "operator", "optimizer", "sampler", "serialization", "stats",
           "vqs", "driver", "VMC", "VMC_SR"]
__version__ = "0.1.0"
