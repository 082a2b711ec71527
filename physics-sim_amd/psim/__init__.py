"""MI355X-native DEM time step: host mirror of the reference's State / Settings /
WGPUComputeProg over the C ABI of libpsim.so (include/dem.h)."""
from .settings import Settings, SetupSettings, SimulationSettings, PhysicsSettings  # noqa: F401
from .state import State, PREALLOC, DATA_SIZE  # noqa: F401

__all__ = ["Settings", "State", "ComputeProg", "load_library"]


def __getattr__(name):
    if name == "ComputeProg":
        from .compute import ComputeProg
        return ComputeProg
    if name == "load_library":
        from .abi import load_library
        return load_library
    raise AttributeError(name)
