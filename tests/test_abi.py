"""The C-ABI library loads without a GPU and exports every entry point that
include/*.h declares; the ctypes layouts match the ABI sizes."""
import ctypes as C
import os
import re
import subprocess

from helpers import ROOT
from reactive_amd import LIBRARY_PATH, abi, lib as rlib

HEADERS = ["reactive_transport.h", "reactive_gpu_hem.h", "reactive_amd.h"]


def declared():
    names = set()
    for h in HEADERS:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(reactive_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIBRARY_PATH), "run spuma-multiphysics_amd/build.sh"
    lib = C.CDLL(LIBRARY_PATH)
    names = declared()
    assert len(names) >= 60
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIBRARY_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (reactive_\w+)", out))
    assert names <= exported
    assert rlib.load() is not None


def test_abi_layouts():
    assert C.sizeof(abi.ThermoState) == 176 and abi.STATE_DTYPE.itemsize == 176
    assert C.sizeof(abi.Face) == 80 and abi.FACE_DTYPE.itemsize == 80
    assert C.sizeof(abi.TransportState) == 56 and abi.TSTATE_DTYPE.itemsize == 56
    assert C.sizeof(abi.Config) == 88 and C.sizeof(abi.OptionsV2) == 88 and C.sizeof(abi.OptionsV21) == 120
    assert C.sizeof(abi.HemProfile) == 144 and C.sizeof(abi.HemOutput) == 184 and C.sizeof(abi.CapillaryInput) == 24
    assert C.sizeof(abi.Token) == 24 and C.sizeof(abi.Primitive) == 32


def test_policy_string_is_the_legacy_fp64_policy():
    assert rlib.load().reactive_gpu_hem_numerical_policy_v1() == b""


def test_step_report_layout_matches_the_header(tmp_path):
    """ReactiveAmdStepReport as gcc lays it out from include/reactive_amd.h
    against the ctypes mirror: size and every field offset."""
    fields = [f[0] for f in abi.StepReport._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "reactive_amd.h"\nint main(void) {\n'
                   '  printf("%zu\\n", sizeof(ReactiveAmdStepReport));\n' +
                   "".join(f'  printf("%zu\\n", offsetof(ReactiveAmdStepReport, {f}));\n' for f in fields) +
                   "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    values = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert values[0] == C.sizeof(abi.StepReport)
    assert values[1:] == [getattr(abi.StepReport, f).offset for f in fields]


def test_every_declared_entry_point_has_a_ctypes_signature():
    """A ctypes call without argtypes passes a 64-bit handle as a C int and
    truncates it; the binding table must cover every declared entry point."""
    missing = sorted(declared() - set(rlib._SIGNATURES))
    assert not missing, missing
    lib = rlib.load()
    for name in declared():
        assert getattr(lib, name).argtypes is not None, name
