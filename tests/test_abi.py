"""The C ABI (include/vfd.h) without a GPU: the HIP library loads, exports every declared entry point, and the
ctypes mirrors have the header's layout (checked against gcc's offsetof). No compute call is made here."""
import ctypes
import os
import re
import subprocess

import pytest

import vfd_amd
from vfd_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vfd.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(vfd_\w+)\s*\(", text, re.M)))


def test_library_loads_and_exports_every_declared_symbol(vfd_lib):
    names = _declared()
    assert len(names) >= 36
    for n in names:
        assert hasattr(vfd_lib, n), n
    assert set(names) == set(_abi.EXPORTED)
    assert vfd_lib.vfd_abi_version() == 3


def test_default_description_matches_reference(vfd_lib):
    d = _abi.Desc()
    assert vfd_lib.vfd_desc_default(ctypes.byref(d)) == 0
    ref = _abi.make_desc()
    for name, _ in _abi.Desc._fields_:
        a, b = getattr(d, name), getattr(ref, name)
        if name == "Gravity":
            assert list(a) == pytest.approx(list(b))
        else:
            assert a == pytest.approx(b), name


def test_null_handles_are_rejected(vfd_lib):
    assert vfd_lib.vfd_step(None) == 1
    assert b"NULL" in vfd_lib.vfd_last_error()
    assert vfd_lib.vfd_get_debug_info(None, 0, None) == 1


def test_struct_layout_matches_header(tmp_path):
    structs = {"vfd_desc": _abi.Desc, "vfd_info": _abi.Info, "vfd_debug_info": _abi.DebugInfo,
               "vfd_boundary_desc": _abi.BoundaryDesc, "vfd_kernel_stat": _abi.KernelStat,
               "vfd_mesh": _abi.Mesh, "vfd_mesh_boundary_desc": _abi.MeshBoundaryDesc, "vfd_tuning": _abi.Tuning}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "vfd.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in cls._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = subprocess.check_output([str(exe)], text=True).split("\n")
    got = dict(l.rsplit(" ", 1) for l in out if l.strip())
    for cname, cls in structs.items():
        assert int(got["%s size" % cname]) == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(got["%s.%s" % (cname, fname)]) == getattr(cls, fname).offset, (cname, fname)


def test_package_exposes_reference_interface():
    for name in ["DFSPHSimulation", "DFSPHSimulationDescription", "FluidObject", "RigidBodyDescription", "DFSPHBatch"]:
        assert hasattr(vfd_amd, name)
    for m in ["SetDescription", "GetDescription", "GetInfo", "SetFluidObjects", "SetRigidBodies", "OnUpdate",
              "Simulate", "GetParticleCount", "GetMaxVelocityMagnitude", "GetCurrentTimeStepSize", "GetDebugInfo",
              "GetParticleFrameBuffer", "GetKernel", "GetRigidBodyCount", "GetParticleRadius"]:
        assert hasattr(vfd_amd.DFSPHSimulation, m), m


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "vfd_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                text = open(os.path.join(dirpath, f), errors="replace").read()
                assert "oracle" not in re.sub(r"(#|//).*", "", text).lower() or f == "__init__.py", f


def test_tuning_struct_and_defaults_without_gpu(vfd_lib):
    t = _abi.Tuning()
    assert vfd_lib.vfd_get_tuning(None, ctypes.byref(t)) == 1  # NULL handle rejected before any GPU call
    assert b"NULL" in vfd_lib.vfd_last_error()


def test_abi_version_agrees_between_header_library_and_binding(vfd_lib):
    m = re.search(r"#define VFD_ABI_VERSION (\d+)", open(HEADER).read())
    assert m, "VFD_ABI_VERSION missing from include/vfd.h"
    assert int(m.group(1)) == vfd_lib.vfd_abi_version() == _abi.ABI_VERSION
