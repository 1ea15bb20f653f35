"""CPU checks of the drop-in boundary: the C ABI library loads and exports
every entry point declared in include/gpu4snn.h, the host connectivity
builder equals the oracle's bit for bit, invalid arguments are rejected with
G4S_ERR_ARG, and the CLI / gpu4snn.py driver keep the reference's contract.
No GPU compute is called here."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "gpu4snn.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s+\**\s*(g4s_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol(g4s):
    syms = declared_symbols()
    assert len(syms) >= 15, syms
    L = g4s.lib()
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    nm = subprocess.run(["nm", "-D", "--defined-only", g4s.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (g4s_\w+)", nm))
    assert set(syms) <= exported


@pytest.mark.parametrize("n,nsyn,seed", [(2, 1, 1), (3, 2, 9), (97, 96, 4), (2500, 1000, 20221), (4099, 33, 77)])
def test_host_builder_equals_oracle(g4s, oracle, n, nsyn, seed):
    post, w = g4s.build_synapses_host(n, nsyn, seed)
    onet = oracle.Network(n, nsyn, seed)
    assert np.array_equal(post, onet.post)
    assert np.array_equal(w.view(np.uint32), onet.w.view(np.uint32))


def test_host_builder_thread_count_invariant(g4s):
    a = g4s.build_synapses_host(20000, 50, 3, threads=1)
    b = g4s.build_synapses_host(20000, 50, 3, threads=7)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("n,nsyn", [(1, 1), (10, 10), (10, 0), (5, -3)])
def test_invalid_spec_rejected(g4s, n, nsyn):
    spec = g4s.make_spec(n, nsyn, 1)
    post = np.zeros(max(1, n * max(nsyn, 1)), np.uint32)
    w = np.zeros_like(post, dtype=np.float32)
    rc = g4s.lib().g4s_build_synapses_host(C.byref(spec), post.ctypes.data, w.ctypes.data, 1)
    assert rc == -1
    assert b"N" in g4s.lib().g4s_last_error() or b"Nsyn" in g4s.lib().g4s_last_error()


def test_invalid_regime_rejected(g4s):
    spec = g4s.make_spec(100, 10, 1, regime=7)
    h = C.c_void_p()
    assert g4s.lib().g4s_create(C.byref(spec), C.byref(h)) == -1


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_create_without_device_fails_cleanly(g4s):
    with pytest.raises(g4s.G4SError) as e:
        g4s.Network(100, 10, 1)
    assert e.value.code == -2


def test_cli_usage_and_validation(g4s):
    if not os.path.exists(g4s.CLI_PATH):
        pytest.skip("CLI not built")
    assert subprocess.run([g4s.CLI_PATH], capture_output=True).returncode == 2
    r = subprocess.run([g4s.CLI_PATH, "10", "10", "5"], capture_output=True, text=True)
    assert r.returncode == 2 and "Nsyn" in r.stderr
    r = subprocess.run([g4s.CLI_PATH, "10", "3", "5", "--alg", "XYZ"], capture_output=True, text=True)
    assert r.returncode == 2


def test_driver_module_has_no_import_side_effects(tmp_path):
    code = f"import sys; sys.path.insert(0, {ROOT!r}); import gpu4snn; print(gpu4snn.N, gpu4snn.Nsyn)"
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0 and r.stdout.split() == ["2500", "1000"]
    assert os.listdir(tmp_path) == []


def test_driver_usage(tmp_path):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "gpu4snn.py")], capture_output=True, cwd=tmp_path)
    assert r.returncode == 2
