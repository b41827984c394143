"""The C-ABI library: loads, exports every symbol include/*.h declares, and
reports errors as status codes with a message (no compute calls: CPU only)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols():
    syms = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b(nodals_[a-z0-9_]+)\s*\(", text):
            syms.add(m.group(1))
    return syms


def test_library_exports_every_declared_symbol():
    from nodals_amd import _lib

    L = _lib.lib()
    declared = _declared_symbols()
    assert len(declared) >= 25
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    assert set(_lib.EXPORTS) <= declared


def test_version_and_empty_error():
    from nodals_amd import _lib

    L = _lib.lib()
    assert b"gfx950" in L.nodals_version()


def test_usage_errors_are_status_codes(mesh_dir):
    from nodals_amd._lib import H8, NodalsError

    with pytest.raises(NodalsError) as e:
        H8(["--mesh", mesh_dir["small"], "--bogus", "1"], device=False)
    assert e.value.code == 1 and "usage error" in str(e.value)
    with pytest.raises(NodalsError, match="fixed10 requires maxOuter=10"):
        H8(["--mesh", mesh_dir["small"]], device=False)
    with pytest.raises(NodalsError, match="not supported by this build"):
        H8(["--mesh", mesh_dir["small"], "--max-outer", "10", "--weak-wall", "1"], device=False)
    with pytest.raises(NodalsError, match="cannot open"):
        H8(["--mesh", "/nonexistent/polyMesh", "--max-outer", "10"], device=False)
    for n in ("0", "61"):
        with pytest.raises(NodalsError, match="mom-sweeps must be in"):
            H8(["--mesh", mesh_dir["small"], "--max-outer", "10", "--mom-sweeps", n], device=False)
    H8(["--mesh", mesh_dir["small"], "--max-outer", "10", "--mom-sweeps", "20"], device=False).close()


def test_host_only_handle_rejects_device_calls(mesh_dir):
    from nodals_amd._lib import H8, NodalsError

    h = H8(["--mesh", mesh_dir["small"], "--max-outer", "10"], device=False)
    for call in (lambda: h.assemble(0.5), lambda: h.continuity(), lambda: h.pressure_refresh(),
                 lambda: h.run(), lambda: h.get("u0", h.info["nv"])):
        with pytest.raises(NodalsError, match="without a device"):
            call()
    with pytest.raises(NodalsError, match="count mismatch"):
        h.get("mom_col", 3)
    with pytest.raises(NodalsError, match="unknown field"):
        h.get("nope", 1)
    h.close()


def test_null_handles():
    from nodals_amd import _lib

    L = _lib.lib()
    assert L.nodals_h8_assemble(None, 0.5, 1.0) == 1
    assert b"null" in L.nodals_last_error()
    assert L.nodals_sys_spmv(None, None, None) == 1


def test_sys_rejects_bad_input():
    from nodals_amd import _lib

    L = _lib.lib()
    h = C.c_void_p()
    row = np.array([0, 1, 2], dtype=np.int64)
    col = np.array([0, 5], dtype=np.int32)   # column out of range
    val = np.array([1.0, 1.0])
    rc = L.nodals_sys_create(2, row.ctypes.data, col.ctypes.data, val.ctypes.data, b"", C.byref(h))
    assert rc == 1 and b"column out of range" in L.nodals_last_error()
    col = np.array([0, 1], dtype=np.int32)
    rc = L.nodals_sys_create(2, row.ctypes.data, col.ctypes.data, val.ctypes.data, b"precision=fp16", C.byref(h))
    assert rc == 1 and b"precision" in L.nodals_last_error()


def test_cell_graph_is_the_face_adjacency(mesh_dir):
    """The learned warm start's graph: cells sharing a face, symmetric."""
    import numpy as np
    from nodals_amd._lib import H8

    h = H8(["--mesh", mesh_dir["small"], "--max-outer", "10"], device=False)
    off, nbr = h.gnn_graph()
    nc = h.info["cells"]
    ref = h.get("cell_ref", nc * 8).reshape(nc, 8)
    pairs = set()
    for i in range(nc):
        for k in range(off[i], off[i + 1]):
            pairs.add((i, int(nbr[k])))
    assert all((j, i) in pairs for (i, j) in pairs)
    assert np.diff(off).max() <= 4 and np.diff(off).min() >= 1
    # every shared free face bubble (DOFs 4..7 of a cell) links exactly two cells
    faces = {}
    for c in range(nc):
        for a in range(4, 8):
            if ref[c, a] >= 0:
                faces.setdefault(int(ref[c, a]), []).append(c)
    for cs in faces.values():
        if len(cs) == 2:
            assert (cs[0], cs[1]) in pairs
    h.close()


def test_noalloc_profile5_rejects_oversize_before_device_work():
    """More than 65536 faces is refused with an error before any HIP call
    (so this runs without a GPU as well)."""
    from nodals_amd import _lib

    n = 65537
    z = np.zeros((n, 5))
    with pytest.raises(_lib.NodalsError, match="65536"):
        _lib.wall_profile5_noalloc(z, z, z + 1.0, 1e-5)
