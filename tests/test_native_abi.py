"""CPU-side checks of the C ABI: the library loads and exports every declared symbol, and
the host-side validation of the drop-in classes raises the reference's errors before any
device work.  No compute call is made here (there is no GPU on this side)."""

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "celltree_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int64_t|size_t|int)\s+(ct_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert len(syms) >= 30
    assert "ct_build" in syms and "ct_locate_points" in syms


def test_library_exports_every_declared_symbol():
    from numba_celltree_amd import _native

    L = _native.lib()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert set(declared_symbols()) == set(_native.SIGNATURES), "binding and header diverge"


def test_host_only_entry_points():
    from numba_celltree_amd import _native

    L = _native.lib()
    assert L.ct_version() == 1
    # creation.py:216-230 pessimistic_n_nodes, against the oracle's restatement.
    from oracle import refcpu as R

    for n in (1, 2, 3, 7, 20_000, 999_941):
        assert L.ct_max_nodes(n) == R.lib().ref_pessimistic_n_nodes(n)
    assert L.ct_build_workspace_bytes(1_000_000, 4) > 0
    assert L.ct_scan_workspace_bytes(10) > 0


def test_argument_errors_without_device():
    from numba_celltree_amd import _native

    L = _native.lib()
    rc = L.ct_locate_points(None, None, 0, 0.0, None, None)
    assert rc == _native.CT_ERR_ARG
    assert "incomplete tree" in _native.last_error()
    with pytest.raises(ValueError):
        _native.check(_native.CT_ERR_ARG, "x")
    with pytest.raises(IndexError):
        _native.check(_native.CT_ERR_DEGENERATE, "x")


def test_constructor_validation_is_host_side():
    import numba_celltree_amd as M

    with pytest.raises(ValueError):
        M.CellTree2d([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]], [[0, 1, 2]], -1, n_buckets=1)
    with pytest.raises(ValueError):
        M.CellTree2d([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]], [[0, 1, 2]], -1, cells_per_leaf=0)
    with pytest.raises(ValueError):
        M.CellTree2d([(1, 2, 3), (3, 4, 5)], [[0, 1, 2]], -1)
    with pytest.raises(ValueError):
        M.CellTree2d([[0.0, 0.0], [1.0, 0.0]], [0, 1, 2], -1)
    with pytest.raises(ValueError):
        M.CellTree2d([[0.0, 0.0]], np.zeros((1, 33), dtype=int), -1)


def test_cast_matches_reference_contract():
    from numba_celltree_amd.cast import cast_faces, cast_vertices

    f = cast_faces(np.array([[0, 1, 2, -999]]), -999)
    assert f.dtype == np.intp and f[0, -1] == -1
    v = cast_vertices(np.array([[0, 1]], dtype=np.float32), copy=True)
    assert v.dtype == np.float64


@pytest.mark.skipif(__import__("torch").cuda.is_available(), reason="checks the no-GPU error path")
def test_no_gpu_fails_loudly():
    import numba_celltree_amd as M

    with pytest.raises(RuntimeError, match="ROCm GPU"):
        M.CellTree2d([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]], [[0, 1, 2]], -1)


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes mirror of a header struct has the C compiler's size and field offsets."""
    import shutil
    import subprocess

    from numba_celltree_amd import _native

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    mirrors = {"ct_node": _native.CtNode, "ct_tree": _native.CtTree, "ct_batch_tile": _native.CtBatchTile,
               "ct_batch": _native.CtBatch, "ct_regrid_topology": _native.CtRegridTopology}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'  printf("{cname} @sizeof %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in mirrors.items():
        assert got[(cname, "@sizeof")] == ctypes.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, (cname, field)


def test_start_grid_size_rule():
    from numba_celltree_amd import celltree_base as CB

    assert CB.START_GRID_N is None
    assert CB.start_grid_size(10) == 0 and CB.start_grid_size(63) == 0
    assert CB.start_grid_size(64) == 8
    assert CB.start_grid_size(1000) == 32  # 31.6 -> 32
    assert CB.start_grid_size(1024) == 32
    assert CB.start_grid_size(1025) == 64
    assert CB.start_grid_size(1_201_919) == CB.START_GRID_MAX == 1024


def test_element_indices_are_checked_before_any_launch():
    """Out-of-range vertex indices, and -1 in edge rows (which have no fill value), raise
    ValueError on the host: the kernels read vertices without bounds checks."""
    from numba_celltree_amd import _device as D

    faces = np.array([[0, 1, 2], [1, 3, 2]])
    with pytest.raises(ValueError, match="out of bounds"):
        D.faces_to_dev(faces, 3)
    with pytest.raises(ValueError, match=r"\[-1, n_vertices\)"):
        D.faces_to_dev(np.array([[0, 1, -2]]), 4)
    with pytest.raises(ValueError, match=r"\[0, n_vertices\)"):
        D.faces_to_dev(np.array([[0, 1], [1, -1]]), 4, allow_fill=False)
    with pytest.raises(ValueError, match="indices must lie"):
        D.index_to_dev(np.array([0, 5]), 5)
    with pytest.raises(ValueError, match="indices must lie"):
        D.index_to_dev(np.array([0, -1]), 5)
    with pytest.raises(ValueError, match="indices must lie"):
        D.index_to_dev(np.array([0, -2]), 5, allow_missing=True)
