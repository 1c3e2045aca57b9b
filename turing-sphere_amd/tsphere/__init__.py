"""tsphere: MI355X (gfx950) spectral IMEX step behind the C ABI in
include/tsphere.h.

Host-side mirror of the reference's hot-path interfaces:

* :mod:`tsphere.layout`  -- src/sht/layout.ts
* :mod:`tsphere.sht`     -- src/sht/sht.ts ``ShtPlan``
* :mod:`tsphere.ir`      -- the numbl IR surface of src/mgpu/numbl.d.ts
* :mod:`tsphere.hipgen`  -- src/mgpu/wgsl.ts ``buildKernel`` as a HIP emitter
* :mod:`tsphere.plan`    -- src/mgpu/plan.ts ``ModelPlan`` / ``HostBuffers``
* :mod:`tsphere.model`   -- src/mgpu/model.ts ``GpuModel`` / ``eigenvalues``

All compute runs in libtsphere.so; this package only plans and records.
"""
from . import _lib
from ._lib import TsphereError
from .device import DeviceBuffer, Event, Stream, device_count
from .layout import ShtConfig, grid_for_lmax, lm_index, nlm_calc, validate_config
from .sht import ShtBinding, ShtPlan, plan_tables

__all__ = [
    "TsphereError", "DeviceBuffer", "Event", "Stream", "device_count", "ShtConfig",
    "grid_for_lmax", "lm_index", "nlm_calc", "validate_config", "ShtBinding", "ShtPlan",
    "plan_tables", "load_library",
]


def load_library():
    """Load libtsphere.so (raises if it is not built)."""
    return _lib.load()
