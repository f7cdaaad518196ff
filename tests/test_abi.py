"""The C-ABI library loads on a CPU-only host and exports every entry point
declared in include/spk.h (no compute calls are made)."""
import ctypes
import os
import re

from spkmod import spk

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")


def declared_symbols():
    names = set()
    for fn in sorted(os.listdir(HDR)):
        if fn.endswith(".h"):
            txt = open(os.path.join(HDR, fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(spk_\w+)\s*\(", txt, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(spk.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30, names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding covers exactly the declared surface
    assert set(spk.SIGNATURES) == names


def test_abi_version_and_param_layout():
    import pyoracle as orc
    L = spk.lib()
    assert L.spk_abi_version() == 1
    assert L.spk_params_size() == ctypes.sizeof(spk.SpkParams)
    assert orc.lib().orc_params_size() == ctypes.sizeof(spk.SpkParams)
    assert [f[0] for f in spk.SpkParams._fields_] == [f[0] for f in orc.OrcParams._fields_]


def test_snapshot_library_exports_spk_io_h():
    """libspkio.so (HDF5 snapshots) exports every entry point of include/spk_io.h."""
    import importlib

    import pytest

    snapshot = importlib.import_module("spooky-git_amd.snapshot")
    if not os.path.exists(snapshot.LIBPATH):
        pytest.skip("libspkio.so not built (no HDF5 C library)")
    txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(HDR, "spk_io.h")).read(), flags=re.S)
    names = set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(spkio_\w+)\s*\(", txt, flags=re.M))
    assert names == {"spkio_last_error", "spkio_write_snapshot", "spkio_read_header", "spkio_read_field",
                     "spkio_last_snapshot"}
    L = snapshot.lib()
    assert all(hasattr(L, n) for n in names)
