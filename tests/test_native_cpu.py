"""CPU-side checks of the C ABI library: it is built, loads, and exports exactly the entry
points include/rapidnj_amd.h declares; calls that need no device behave. No kernel runs here."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rapidnj_amd.h")
LIB = os.path.join(ROOT, "rapidnj_amd", "librapidnj_amd.so")


def declared_functions():
    text = open(HEADER).read()
    return re.findall(r"RNJ_EXPORT\s+[\w\s\*]+?\b(\w+)\s*\(", text)


@pytest.fixture(scope="module")
def native():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-C", os.path.join(ROOT, "rapidnj_amd", "csrc")], check=True)
    from rapidnj_amd import native
    native.lib()
    return native


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], check=True, capture_output=True,
                         text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_header_declares_the_reference_boundary():
    names = set(declared_functions())
    for ref in ("initialiseDNA2_gpu", "storeDataDNA2_gpu", "computeDistancesDNA2_gpu",
                "initialiseDNA_gpu", "storeDataDNA_gpu", "computeDistancesDNA_gpu",
                "getResultsDNA_gpu", "initialiseProtein_gpu", "computeDistancesProtein_gpu",
                "getResultsProtein_gpu"):
        assert ref in names


def test_library_exports_exactly_the_declared_entry_points(native):
    declared = set(declared_functions())
    assert len(declared) >= 30
    assert exported_symbols() == declared


def test_binding_covers_every_entry_point(native):
    assert set(native.SIGNATURES) == set(declared_functions())
    for name in native.SIGNATURES:
        assert getattr(native.lib(), name)


def test_calls_without_a_device(native):
    L = native.lib()
    assert L.rnj_api_version() == 2
    assert L.rnj_status_string(native.RNJ_E_NOT_DNA) == b"RNJ_E_NOT_DNA"
    assert L.rnj_tile_config_count() >= 1
    assert all(native.tile_configs())
    assert L.rnj_tile_config_name(-1) is None
    # argument checks come before any device call
    assert L.rnj_distances_device(None, None, 1, 0, None, 0, 0, None, None) == native.RNJ_E_INVALID
    assert b"null" in L.rnj_last_error()
    assert L.rnj_distances_device_noalloc(None, None, 1, 0, None, 0, 0, None,
                                          None) == native.RNJ_E_INVALID
    assert b"null" in L.rnj_last_error()
    assert L.rnj_set_tile_config(None, 0) == native.RNJ_E_INVALID
    assert L.rnj_set_char_rules(None, 1) == native.RNJ_E_INVALID
    assert L.rnj_server_submit(None, None, 0, 0, 0, 0, None) == native.RNJ_E_INVALID


def test_context_needs_a_device(native):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(native.RnjError) as e:
        native.Context(0)
    assert e.value.status == native.RNJ_E_NO_DEVICE


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from rapidnj_amd import native as N
    monkeypatch.setattr(N, "_lib", None)
    monkeypatch.setattr(N, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(ImportError):
        N.lib()
