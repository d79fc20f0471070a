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


# ---- the reference's C++ entry points ------------------------------------------
REF_CUDA = "/root/reference/Common/CUDA"
# One gateway per header, as the reference's own gateways include them (GD_TV.hpp
# and PICCS.hpp share an include guard, so no unit can include both), with the
# C++ expressions that name every function and GpuIds member it can reach.
REF_GATEWAYS = {
    "Siddon_projection.hpp": ["&siddon_ray_projection"],
    "ray_interpolated_projection.hpp": ["&interpolation_projection"],
    "Siddon_projection_parallel.hpp": ["&siddon_ray_projection_parallel"],
    "ray_interpolated_projection_parallel.hpp": ["&interpolation_projection_parallel"],
    "voxel_backprojection.hpp": ["&voxel_backprojection"],
    "voxel_backprojection2.hpp": ["&voxel_backprojection2"],
    "voxel_backprojection_parallel.hpp": ["&voxel_backprojection_parallel"],
    "GD_TV.hpp": ["&pocs_tv"],
    "GD_AwTV.hpp": ["&aw_pocs_tv"],
    "PICCS.hpp": ["&piccs_tv"],
    "tv_proximal.hpp": ["&tvdenoising"],
    "gpuUtils.hpp": ["&GetGpuCount", "&GetGpuIdArray", "&GetGpuName"],
    "GpuIds.hpp": ["&GpuIds::SetIds", "&GpuIds::GetLength", "&GpuIds::SetAllGpus", "&GpuIds::AreEqualDevices"],
}


@pytest.mark.parametrize("path", [_native.LIB_PATH, _native.LIB_PATH_MATLAB])
def test_exports_reference_entry_points(path):
    lib = ctypes.CDLL(path)
    for name, mangled in _native.REFERENCE_ABI.items():
        assert hasattr(lib, mangled), name


@pytest.mark.skipif(not os.path.isdir(REF_CUDA), reason="reference checkout not present")
def test_gateway_built_on_reference_headers_links(tmp_path):
    """A caller compiled against the reference's own headers (as the Cython and
    MEX gateways are, Python build: IS_FOR_PYTIGRE) links against the library
    with no undefined symbol, and the symbols it needs are the mangled names
    the library's table lists."""
    objs, undefined = [], set()
    for k, (header, exprs) in enumerate(REF_GATEWAYS.items()):
        uses = "\n".join("    { auto p = %s; keep((const void*)&p); }" % e for e in exprs)
        extra = ""
        if header == "GpuIds.hpp":  # constructor, destructor and both subscripts
            extra = "    GpuIds ids;\n    (void)ids[0];\n    const GpuIds& c = ids;\n    (void)c[0];\n"
        src = tmp_path / ("gateway%d.cpp" % k)
        src.write_text('#include <cstddef>\n#include "%s"\nvoid keep(const void*);\nvoid use%d() {\n%s\n%s}\n'
                       % (header, k, uses, extra))
        obj = tmp_path / ("gateway%d.o" % k)
        subprocess.run(["g++", "-std=c++17", "-DIS_FOR_PYTIGRE", "-I", REF_CUDA, "-c", str(src), "-o", str(obj)],
                       check=True)
        objs.append(str(obj))
        undefined |= set(subprocess.run(["nm", "-u", str(obj)], capture_output=True, text=True,
                                        check=True).stdout.split())
    # what the gateways need beyond the C++ runtime and the helper
    runtime = {"U", "_Z4keepPKv", "_GLOBAL_OFFSET_TABLE_", "__stack_chk_fail", "_Unwind_Resume",
               "__gxx_personality_v0"}
    assert undefined - runtime == set(_native.REFERENCE_ABI.values())
    keep = tmp_path / "keep.cpp"
    keep.write_text("void keep(const void*) {}\n")
    libdir = os.path.dirname(_native.LIB_PATH)
    subprocess.run(["g++", "-shared", "-fPIC", "-Wl,--no-undefined", *objs, str(keep), "-o",
                    str(tmp_path / "libgateway.so"), "-L", libdir, "-ltigre_amd"], check=True)


def test_gpu_ids_members():
    """GpuIds members behave as GpuIds.cpp:6-60 describes (SetAllGpus lists
    devices 0 .. n-1), with malloc'd ids released by the destructor."""
    abi = _native.reference_abi()
    g = _native.GpuIds()
    abi["GpuIds::GpuIds"](ctypes.byref(g))
    assert g.m_iCount == 0 and not g.m_piDeviceIds
    ids = (ctypes.c_int * 3)(4, 1, 7)
    abi["GpuIds::SetIds"](ctypes.byref(g), 3, ids)
    assert abi["GpuIds::GetLength"](ctypes.byref(g)) == 3
    assert [abi["GpuIds::operator[] const"](ctypes.byref(g), i) for i in range(3)] == [4, 1, 7]
    abi["GpuIds::operator[]"](ctypes.byref(g), 1).contents.value = 9
    assert g.m_piDeviceIds[1] == 9
    abi["GpuIds::SetAllGpus"](ctypes.byref(g), 2)
    assert g.m_iCount == 2 and [g.m_piDeviceIds[i] for i in range(2)] == [0, 1]
    abi["GpuIds::~GpuIds"](ctypes.byref(g))
    assert g.m_iCount == 0 and not g.m_piDeviceIds


def test_reference_entry_points_reject_bad_arguments():
    """Status codes instead of process exit: no views, and a zero voxel size."""
    abi = _native.reference_abi()
    geo = prepared(cone_geometry((8, 8, 8), (8.0, 8.0, 8.0), (8, 8), (1.0, 1.0), 100.0, 150.0), np.zeros(2))
    cg = _native.CGeometry(geo, 2)
    img = np.zeros((8, 8, 8), dtype=np.float32)
    out = np.zeros((2, 8, 8), dtype=np.float32)
    ang = np.zeros((2, 3), dtype=np.float32)
    ptrs = (ctypes.c_void_p * 2)(out.ctypes.data, out.ctypes.data + 256)
    ids = _native.GpuIds()
    st = abi["siddon_ray_projection"](img.ctypes.data, cg.struct, ptrs, ang.ctypes.data, 0, ctypes.byref(ids))
    assert st == 6 and "nangles" in _native.last_error()
    cg.struct.dVoxelX = 0.0
    st = abi["voxel_backprojection2"](out.ctypes.data, cg.struct, img.ctypes.data, ang.ctypes.data, 2,
                                      ctypes.byref(ids))
    assert st == 6 and "dVoxel" in _native.last_error()
