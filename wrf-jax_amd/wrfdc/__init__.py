"""MI355X-native RK3 split-explicit dycore (WRF-JAX step API on HIP kernels).

Public entry points:

* ``wrfdc.device.Domain`` -- device-resident domain advanced through the C ABI
  (include/wrfdc.h, libwrfdc.so);
* ``wrfdc.step_em`` -- the reference's typed step API (``step_em_core_once``,
  ``build_step_em_core_context``) over ``Domain``;
* ``wrfdc.cases`` -- seeded synthetic ideal cases of the benchmark workloads.
"""

from .config import DycoreConfig
from .fields import FIELD_SPECS, Dims

__all__ = ["DycoreConfig", "Dims", "FIELD_SPECS"]
