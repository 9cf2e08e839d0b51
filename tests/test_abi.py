"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "flowsolversourcecode_amd", "libfsamd.so")


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(fs_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for need in ("fs_mesh_create", "fs_solve_equation", "fs_solve_equation_device", "fs_last_error"):
        assert need in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfsamd.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfsamd.so not built")
def test_invalid_topology_is_rejected_on_the_host():
    import numpy as np
    import flowsolversourcecode_amd as fs
    nbr = np.array([[2, 5]], np.int32)  # neighbour 5 does not exist in a 2-cell mesh
    with pytest.raises(fs.FsError):
        fs.Mesh(nbr, np.array([1, 1], np.int32))
