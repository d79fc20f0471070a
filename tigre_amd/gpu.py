"""Device selection with the interface of Python/tigre/utilities/gpu.py:12-70."""
from __future__ import annotations

import copy
import ctypes

from . import _native


def getGpuNames():
    lib = _native.lib()
    names = []
    for i in range(lib.tigre_get_gpu_count()):
        buf = ctypes.create_string_buffer(256)
        if lib.tigre_get_gpu_name(i, buf, 256) == 0:
            names.append(buf.value.decode())
    return names


class GpuIds(object):
    """IDs of the devices whose name matches ``nameGPU`` (all devices if empty)."""

    def __init__(self, nameGPU=None):
        self.name = "" if nameGPU is None else nameGPU
        names = getGpuNames()
        self.devices = [i for i, n in enumerate(names) if not self.name or n == self.name]

    def __len__(self):
        return len(self.devices)

    def __str__(self):
        return str({"name": self.name, "devices": self.devices})

    def __getitem__(self, key):
        single = copy.deepcopy(self)
        single.devices = [self.devices[key]]
        return single


def getGpuIds(gpuName=None):
    return GpuIds(gpuName)
