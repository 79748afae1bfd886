"""The C ABI libraries load and export every symbol their headers declare
(no compute calls here: those need a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(g2h?_[a-z_0-9]+)\s*\(", txt)))


def test_hip_library_exports(g2):
    names = declared("g2hip.h")
    assert "g2_mce_run" in names and "g2_gm_run" in names
    lib = ctypes.CDLL(os.path.join(ROOT, "g2-aimd_amd", "libg2hip.so"))
    for n in names:
        assert hasattr(lib, n), n
    from importlib import import_module
    dev = import_module("g2-aimd_amd.device")
    assert set(dev.EXPORTS) == set(names)


def test_host_library_exports(g2):
    names = declared("g2host.h")
    lib = ctypes.CDLL(os.path.join(ROOT, "g2-aimd_amd", "libg2host.so"))
    for n in names:
        assert hasattr(lib, n), n


def test_plan_struct_layouts(g2, orc):
    # g2_plan: 3 scalars + tables of include/g2plan.h, all u32
    Q = 8
    words = 3 + 2 * Q + Q + Q * Q + Q + 2 * Q * Q + Q + Q + Q * Q + Q + Q * Q + Q + 2 * Q * Q + Q + 2 * Q * Q + (Q + 1) + Q
    assert ctypes.sizeof(g2.G2Plan) == 4 * words
    assert ctypes.sizeof(orc.OrcPlan) == ctypes.sizeof(g2.G2Plan)
    for (a, _), (b, _) in zip(g2.G2Plan._fields_, orc.OrcPlan._fields_):
        assert getattr(g2.G2Plan, a).offset == getattr(orc.OrcPlan, b).offset
