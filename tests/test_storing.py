"""Snapshot storing on the host (storing.c:40-843): the files this library
writes against those of the reference build for the same System and status
(h5diff: same groups, datasets, types and values), and
finalize_storing_param's directory handling against the reference's."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import _cases as K
import _h5 as H
import _oracle as O
from hydrofv_amd import abi, driver

needs_h5 = pytest.mark.skipif(not H.available() or H.h5diff() is None, reason="HDF5 library or h5diff missing")
needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) missing")


@pytest.fixture(scope="module")
def libs():
    import hydrofv_amd
    ours = hydrofv_amd.load()
    ref = abi.bind(C.CDLL(O.REF_SO)) if O.ref_available() else None
    return ours, ref


def write(lib, sim, out_dir, count=3):
    ip = lib.get_new_integrator_param()
    st = abi.SimulationStatus(num_steps=7, t=0.125, dt=0.0045)
    sp = lib.get_new_storing_param()
    keep = out_dir.encode()
    sp.output_dir = keep
    sp.is_storing = True
    sp.storing_interval = 0.1
    sp.store_count_ = count
    driver.check(lib, lib.store_snapshot(C.byref(sim.bc), C.byref(sim.system), C.byref(ip), C.byref(st), C.byref(sp)))
    assert sp.store_count_ == count + 1
    return os.path.join(out_dir, f"snapshot_{count}.h5")


@needs_h5
@needs_ref
@pytest.mark.parametrize("case", ["2d_mixed", "2d_none", "1d_sod", "1d_spherical"])
def test_snapshot_file_equals_reference_writer(libs, tmp_path, case):
    ours, ref = libs
    if case.startswith("2d"):
        bcs = K.BC_SETS["mixed" if case == "2d_mixed" else "none"]
        mk = lambda L: K.setup_random_2d(L, 13, 7, 2, bcs, 21)
    elif case == "1d_sod":
        mk = lambda L: K.setup_sod(L, 50)
    else:
        mk = lambda L: K.setup_sod(L, 40, coord="spherical_1d", bc=("reflective", "transmissive"))
    sims = [mk(ours), mk(ref)]
    files = []
    for lib, sim, sub in zip((ours, ref), sims, ("ours/", "ref/")):
        d = str(tmp_path / sub)
        os.makedirs(d)
        files.append(write(lib, sim, d + "/"))
    same, text = H.same_files(*files)
    assert same, text
    # and the content is the System's state
    W = np.stack(sims[0].primitive())
    assert np.array_equal(H.read(files[0], "/fields/density"), W[0].reshape(-1))
    assert H.read(files[0], "/simulation_status/num_steps", "i8")[0] == 7
    assert H.read(files[0], "/parameters/num_ghost_cells_side", "i4")[0] == sims[0].g


def finalize(lib, output_dir, interval=0.1):
    sp = lib.get_new_storing_param()
    sp.is_storing = True
    sp.storing_interval = interval
    keep = None if output_dir is None else output_dir.encode()
    sp.output_dir = keep
    st = lib.finalize_storing_param(C.byref(sp))
    text = abi.traceback_text(lib, st) if st.return_code != abi.SUCCESS else ""
    return st.return_code, text, sp.output_dir


@pytest.mark.parametrize("which", ["ours", "ref"])
def test_finalize_storing_param(libs, tmp_path, monkeypatch, which):
    lib = libs[0] if which == "ours" else libs[1]
    if lib is None:
        pytest.skip("reference build (oracle/_ref) missing")
    monkeypatch.chdir(tmp_path)
    code, text, _ = finalize(lib, "no_slash")
    assert code == 2 and 'must end with a trailing slash ("/"). Got: "no_slash".' in text
    code, text, _ = finalize(lib, "out/", interval=0.0)
    assert code == 2 and "Storing interval must be positive." in text
    code, text, d = finalize(lib, None)
    assert code == 0 and re.fullmatch(rb"snapshots_\d{8}_\d{6}/", d) and os.path.isdir(d.decode())
    code, text, _ = finalize(lib, "fresh/")
    assert code == 0 and os.path.isdir("fresh")
    code, text, _ = finalize(lib, "fresh/")  # exists: a warning, no error
    assert code == 0
    open("a_file", "w").close()
    code, text, _ = finalize(lib, "a_file/")
    assert code == 5 and "storing snapshots" in text
    code, text, _ = finalize(lib, "missing/parent/")
    assert code == 5 and 'Failed to access path for storing snapshots: "missing/parent/".' in text
