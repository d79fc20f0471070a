"""The C ABI: the library loads, exports what include/tigre_amd.h declares, keeps
the reference's Geometry layout, and rejects bad arguments before touching a device."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import tigre_amd
from tigre_amd import _native
from tigre_amd.geometry import cone_geometry, prepared

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tigre_amd.h")
REF_TYPES = "/root/reference/Common/CUDA/types_TIGRE.hpp"


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(tigre_\w+)\s*\(", text, re.M)))


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 12
    for name in syms:
        assert hasattr(lib, name), name
    assert set(syms) == set(_native.SIGNATURES)


def test_struct_layout_matches_header(tmp_path):
    fields = [f[0] for f in _native.TigreGeometry._fields_]
    src = tmp_path / "layout.c"
    body = "\n".join('printf("%s %%zu\\n", offsetof(TigreGeometry, %s));' % (f, f) for f in fields)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "%s"\nint main(void){printf("size %%zu\\n", '
                   'sizeof(TigreGeometry));\n%s\nreturn 0;}\n' % (HEADER, body))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split("\n")
               if line)
    assert int(out["size"]) == ctypes.sizeof(_native.TigreGeometry)
    for f in fields:
        assert int(out[f]) == getattr(_native.TigreGeometry, f).offset, f


@pytest.mark.skipif(not os.path.exists(REF_TYPES), reason="reference checkout not present")
def test_struct_is_byte_compatible_with_reference_geometry(tmp_path):
    """Drop-in: TigreGeometry has the offsets of the reference's struct Geometry."""
    fields = [f[0] for f in _native.TigreGeometry._fields_]
    checks = "\n".join("static_assert(offsetof(TigreGeometry, %s) == offsetof(Geometry, %s), \"%s\");" % (f, f, f)
                       for f in fields)
    src = tmp_path / "compat.cpp"
    src.write_text('#include <cstddef>\n#include "%s"\n#include "%s"\nstatic_assert(sizeof(TigreGeometry) == '
                   'sizeof(Geometry), "size");\n%s\nint main(){return 0;}\n' % (HEADER, REF_TYPES, checks))
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", str(src)], check=True)


def test_invalid_arguments_are_rejected_without_a_device():
    lib = _native.lib()
    geo = prepared(cone_geometry((8, 8, 8), (8.0, 8.0, 8.0), (8, 8), (1.0, 1.0), 100.0, 150.0), np.zeros(2))
    cg = _native.CGeometry(geo, 2)
    img = np.zeros((8, 8, 8), dtype=np.float32)
    out = np.zeros((2, 8, 8), dtype=np.float32)
    ang = np.zeros((2, 3), dtype=np.float32)
    ptrs = (ctypes.POINTER(ctypes.c_float) * 2)(*[ctypes.cast(out.ctypes.data + 256 * i,
                                                             ctypes.POINTER(ctypes.c_float)) for i in range(2)])
    st = lib.tigre_ax(img.ctypes.data_as(ctypes.c_void_p), cg.ptr(), ptrs, ang.ctypes.data_as(ctypes.c_void_p), 0,
                      0, 0, None, 0)
    assert st == 6 and "nangles" in _native.last_error()
    cg.struct.dVoxelX = 0.0
    st = lib.tigre_atb(out.ctypes.data_as(ctypes.c_void_p), cg.ptr(), img.ctypes.data_as(ctypes.c_void_p),
                       ang.ctypes.data_as(ctypes.c_void_p), 2, 0, 0, None, 0)
    assert st == 6 and "dVoxel" in _native.last_error()
    sizes = (ctypes.c_long * 3)(8, 8, 0)
    assert lib.tigre_pocs_tv(img.ctypes.data_as(ctypes.c_void_p), img.ctypes.data_as(ctypes.c_void_p), 1.0, sizes,
                             5, None, 0) == 6


def test_front_end_checks():
    geo = cone_geometry((8, 8, 8), (8.0, 8.0, 8.0), (8, 8), (1.0, 1.0), 100.0, 150.0)
    with pytest.raises(TypeError):
        tigre_amd.Ax(np.zeros((8, 8, 8)), geo, np.zeros(2))
    with pytest.raises(ValueError):
        tigre_amd.Ax(np.zeros((8, 8, 9), dtype=np.float32), geo, np.zeros(2))
    with pytest.raises(ValueError):
        tigre_amd.Atb(np.zeros((3, 8, 8), dtype=np.float32), geo, np.zeros(2))
    with pytest.raises(TypeError):
        tigre_amd.Atb(np.zeros((2, 8, 8), dtype=np.float64), geo, np.zeros(2))
