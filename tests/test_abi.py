"""The C ABI library (libmdy_hip.so): it loads without a GPU, exports every
entry point that include/mdy.h declares, and the ctypes mirrors in
medyan_amd/_abi.py have the C layouts. No compute call is made here."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from medyan_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mdy.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set(re.findall(r"\b(mdy_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("mdy_create", "mdy_destroy", "mdy_set_network", "mdy_build_neighbors", "mdy_get_pairs",
                 "mdy_energy", "mdy_forces", "mdy_minimize", "mdy_minimize_dev", "mdy_hvp_dev", "mdy_energy_jvp_dev", "mdy_minimize_batch", "mdy_stretch_forces", "mdy_load_forces",
                 "mdy_get_search_dir"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version_and_defaults(lib):
    assert lib.mdy_abi_version() == _abi.MDY_ABI_VERSION
    o = _abi.Opts()
    lib.mdy_default_opts(C.byref(o))
    assert o.precision == _abi.MDY_PREC_FP64 and o.coplanar_mode == _abi.MDY_COPLANAR_REFERENCE


def test_create_without_device_fails_cleanly(lib):
    if lib.mdy_device_count() > 0:
        pytest.skip("a HIP device is present")
    h = C.c_void_p()
    rc = lib.mdy_create(C.byref(h), 0, None)
    assert rc == _abi.MDY_ERR_NOGPU and not h.value


STRUCTS = ["mdy_cylinders", "mdy_ff_tables", "mdy_boundary", "mdy_system", "mdy_cg_params", "mdy_min_result",
           "mdy_opts", "mdy_gen_params", "mdy_network"]
MIRROR = {"mdy_cylinders": _abi.Cylinders, "mdy_ff_tables": _abi.FFTables, "mdy_boundary": _abi.Boundary,
          "mdy_system": _abi.System, "mdy_cg_params": _abi.CGParams, "mdy_min_result": _abi.MinResult,
          "mdy_opts": _abi.Opts, "mdy_gen_params": _abi.GenParams, "mdy_network": _abi.Network}


def test_ctypes_layouts_match_header():
    prog = "#include <stdio.h>\n#include <stddef.h>\n#include \"mdy.h\"\nint main(void){\n"
    for s in STRUCTS:
        prog += f'printf("{s} %zu\\n", sizeof({s}));\n'
    prog += "return 0;}\n"
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "sz.c")
        exe = os.path.join(d, "sz")
        open(c, "w").write(prog)
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    sizes = dict(line.split() for line in out.strip().splitlines())
    for s in STRUCTS:
        assert int(sizes[s]) == C.sizeof(MIRROR[s]), s


def test_ctypes_field_offsets_match_header():
    """Every field of every mirrored struct at the C offset (equal sizes alone
    would not catch two fields in swapped order)."""
    prog = "#include <stdio.h>\n#include <stddef.h>\n#include \"mdy.h\"\nint main(void){\n"
    expect = {}
    for s in STRUCTS:
        for name, _ in MIRROR[s]._fields_:
            prog += f'printf("{s}.{name} %zu\\n", offsetof({s}, {name}));\n'
            expect[f"{s}.{name}"] = getattr(MIRROR[s], name).offset
    prog += "return 0;}\n"
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "off.c")
        exe = os.path.join(d, "off")
        open(c, "w").write(prog)
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    got = {k: int(v) for k, v in (line.split() for line in out.strip().splitlines())}
    bad = {k: (got[k], expect[k]) for k in expect if got[k] != expect[k]}
    assert not bad, bad
