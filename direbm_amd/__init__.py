"""MI355X-native DiReBM step engine (gfx950 HIP kernels behind a C-ABI).

Drop-in for the ``direbm.warp`` surface of berfandil/DiReBM (direbm/warp/__init__.py:9-24):
collide, equilibrium, default_device, disperse, create_control_points, refine_control_points,
resample, GpuSimulator, GpuHexLBM.  There is no CPU path: compute calls raise NoDeviceError
without a HIP GPU and ImportError-style errors if libdirebm_hip.so is not built.
"""

from ._device import DeviceArray, default_device
from .batch import BatchSimulator
from ._lib import DrbError, NoDeviceError
from .lbm import GpuHexLBM
from .physics import collide, equilibrium
from .propagation import create_control_points, disperse, refine_control_points, resample
from .sim3d import GpuSimulator3D
from .simulator import GpuSimulator

__all__ = [
    "collide",
    "equilibrium",
    "default_device",
    "disperse",
    "create_control_points",
    "refine_control_points",
    "resample",
    "GpuSimulator",
    "GpuHexLBM",
    "BatchSimulator",
    "GpuSimulator3D",
    "DeviceArray",
    "DrbError",
    "NoDeviceError",
]
