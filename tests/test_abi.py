"""The C-ABI library loads, exports every symbol include/*.h declares, and its
host-only helpers (mesh generation, colouring) behave.  No GPU needed."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from alya_sma_amd import _lib, problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", txt, flags=re.M):
            if not m.group(0).lstrip().startswith(("#", "typedef")):
                names.add(m.group(1))
    return sorted(names)


def test_library_exports_header_symbols():
    L = _lib.lib()
    names = declared_functions()
    assert "sld_dev_step" in names and "sma160_eval" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_mesh_box_matches_numpy():
    c1, l1 = problems.box_mesh(5, 3, 2, 5.0, 3.0, 2.0)
    c2, l2 = problems.box_mesh_np(5, 3, 2, 5.0, 3.0, 2.0)
    assert np.array_equal(l1, l2)
    np.testing.assert_allclose(c1, c2, rtol=0, atol=1e-15)
    # positive reference Jacobian (node order of mod_lagrange.f90:946-953)
    e = c1[l1[0]]
    assert np.linalg.det(np.stack([e[1] - e[0], e[3] - e[0], e[4] - e[0]])) > 0


@pytest.mark.parametrize("dims", [(10, 2, 2), (7, 5, 3), (1, 1, 1)])
def test_coloring_is_valid(dims):
    coord, lnods = problems.box_mesh_np(*dims, 1.0, 1.0, 1.0)
    color = np.zeros(lnods.shape[0], dtype=np.int32)
    nc = _lib.lib().sld_color_elements(lnods.shape[0], coord.shape[0],
                                       lnods.ctypes.data, color.ctypes.data)
    assert nc >= 1 and nc <= 8
    for c in range(nc):
        nodes = lnods[color == c].ravel()
        assert len(nodes) == len(np.unique(nodes)), "two elements of one colour share a node"


def test_coloring_rejects_bad_index():
    lnods = np.array([[0, 1, 2, 3, 4, 5, 6, 99]], dtype=np.int32)
    color = np.zeros(1, dtype=np.int32)
    nc = _lib.lib().sld_color_elements(1, 8, lnods.ctypes.data, color.ctypes.data)
    assert nc < 0
    assert b"out of range" in _lib.lib().sma_last_error()


def test_create_rejects_bad_input_before_touching_the_gpu():
    from alya_sma_amd.solid import Solidz
    p = problems.bar_problem(nx=2, ny=1, nz=1, lx=2.0, ly=1.0, lz=1.0)
    p["lnods"] = p["lnods"].copy()
    p["lnods"][0, 0] = 10_000
    with pytest.raises(_lib.HipError, match="out of range"):
        Solidz(p)


@pytest.mark.parametrize("case, message", [
    ("unused_node", "zero mass"),
    ("bad_material", "material index out of range"),
    ("bad_law", "law must be"),
    ("bad_function", "func_of_node out of range"),
    ("bad_fext_function", "fext_func out of range"),
    ("bad_npts", "func_npts out of range"),
    ("no_material", "no material"),
])
def test_create_error_paths(case, message):
    """sld_dev_create validates every index and table on the host and reports
    a message (sma_last_error) without touching the GPU."""
    from alya_sma_amd.solid import Solidz
    p = problems.bar_problem(nx=2, ny=1, nz=1, lx=2.0, ly=1.0, lz=1.0)
    if case == "unused_node":
        p["coord"] = np.vstack([p["coord"], [[9.0, 9.0, 9.0]]])
        p["fixmask"] = np.append(p["fixmask"], np.uint8(0))
        p["func_of_node"] = np.append(p["func_of_node"], np.int32(-1))
    elif case == "bad_material":
        p["mat_of_elem"] = np.array([0, 1], np.int32)
    elif case == "bad_law":
        p["law_of_mat"] = np.array([154], np.int32)
    elif case == "bad_function":
        p["func_of_node"] = p["func_of_node"].copy()
        p["func_of_node"][0] = 3
    elif case == "bad_fext_function":
        p["fext0"] = np.zeros_like(p["coord"])
        p["fext_func"] = 5
    elif case == "bad_npts":
        p["func_npts"] = np.array([7], np.int32)
    elif case == "no_material":
        p["props"] = np.zeros((0, 18))
    with pytest.raises(_lib.HipError, match=message):
        Solidz(p)
