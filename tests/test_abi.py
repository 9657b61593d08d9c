"""The C-ABI library loads on the CPU and exports every entry point include/rcmps.h declares;
the Python binding covers exactly those symbols.  No compute call is made without a GPU."""
import ctypes
import re
from pathlib import Path

import pytest

from conftest import ROOT

HEADER = ROOT / "include" / "rcmps.h"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rcmps_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_symbols()
    for expected in ("rcmps_context_create", "rcmps_state_create", "rcmps_fg_phi4",
                     "rcmps_boundary_environment", "rcmps_trajectory_samples"):
        assert expected in names


def test_library_exports_every_declared_symbol():
    import rcmps_gpu._lib as L
    if not L.LIB_PATH.exists():
        pytest.fail(f"{L.LIB_PATH} not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(str(L.LIB_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert set(L.SIGNATURES) == set(declared_symbols())


def test_status_strings_without_device():
    import rcmps_gpu._lib as L
    lib = L.load()
    for code, name in L.STATUS.items():
        assert lib.rcmps_status_string(code).decode() == name
    assert lib.rcmps_max_bond_dimension() >= 64


def test_package_import_does_not_touch_the_device():
    import rcmps_gpu
    assert callable(rcmps_gpu.fg_phi4)


def test_flag_values_match_header():
    import rcmps_gpu._lib as L
    flags = dict(re.findall(r"#define RCMPS_FLAG_([A-Z_]+)\s+(\d+)", HEADER.read_text()))
    assert flags == {"NO_BOUNDARY": "1", "SPLIT": "2", "KINETIC_KRYLOV": "4", "PHASE_TIMERS": "8", "DENSE": "16"}
    assert (L.FLAG_NO_BOUNDARY, L.FLAG_SPLIT, L.FLAG_KINETIC_KRYLOV, L.FLAG_PHASE_TIMERS, L.FLAG_DENSE) == \
        (1, 2, 4, 8, 16)


def test_struct_sizes_match_header():
    """The ctypes mirrors of rcmps_integrator, rcmps_kinetic_options and rcmps_stats have the C sizes
    (compiled against include/rcmps.h)."""
    import subprocess, tempfile, os
    import rcmps_gpu._lib as L
    src = '#include <stdio.h>\n#include "rcmps.h"\nint main(void){printf("%zu %zu %zu\\n", sizeof(rcmps_integrator), ' \
          'sizeof(rcmps_kinetic_options), sizeof(rcmps_stats)); return 0;}\n'
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        exe = os.path.join(d, "s")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", str(ROOT / "include"), c, "-o", exe], check=True)
        sizes = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert sizes == [ctypes.sizeof(L.Integrator), ctypes.sizeof(L.KineticOptions), ctypes.sizeof(L.Stats)]
