"""Host-side pieces of the serving loop that need no GPU: the device-memory
arena (compiled from its header with g++ and exercised by a randomised
test), and the request validation of the C ABI's argument structs."""
import ctypes as C
import shutil
import subprocess
from pathlib import Path

import pytest

import randomfield_amd as R

ROOT = Path(__file__).resolve().parent.parent


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_arena_allocator(tmp_path):
    exe = tmp_path / "arena_test"
    subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-o", str(exe),
                    str(ROOT / "tests" / "cpp" / "arena_test.cpp")], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True)
    assert "arena ok" in out.stdout


def test_server_structs_match_header():
    """Field order and sizes of the ctypes mirrors against the header."""
    hdr = (ROOT / "include" / "rf_capi.h").read_text()
    for name in ("rf_server_config", "rf_request", "rf_request_result", "rf_server_stats"):
        assert f"}} {name};" in hdr
    from randomfield_amd import _capi as K
    assert C.sizeof(K.ServerConfig) == 8 + 4 * 4 + 24 + 8 + 32
    assert C.sizeof(K.Request) == 16 + 8 + 8 + 24 + 16 + 8 + 24
    assert C.sizeof(K.RequestResult) == 8 + 8 + 8 + 24 + 8
    assert C.sizeof(K.ServerStats) == 7 * 8 + 8 + 8


def test_server_null_arguments():
    L = R.lib()
    rid = C.c_int64(0)
    assert L.rf_server_submit(None, None, C.byref(rid)) == R.RF_E_ARG
    assert L.rf_server_destroy(None) == R.RF_OK
