"""GPU: result streaming (g4s_run_stream) and the reference-style front
ends writing the reference's result files (CLI gpu4snn, driver gpu4snn.py)."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("alg,accum", [("AB", "det"), ("SKL", "det"), ("N", "det"), ("AP", "atomic"), ("S", "atomic")])
def test_stream_equals_run(g4s, alg, accum):
    n, nsyn, seed, steps = 3000, 300, 17, 350
    with g4s.Network(n, nsyn, seed) as net:
        ref, cref = net.run(steps, alg=alg, accum=accum, chunk=64)
        net.reset()
        got, cnt, t0s = [], [], []

        def on_chunk(t0, r, c):
            t0s.append(t0)
            got.append(r)
            cnt.append(c)
            return False

        net.run_stream(steps, on_chunk, alg=alg, accum=accum, chunk=64)
        assert t0s == list(range(0, steps, 64))
        words = np.concatenate(got)
        if accum == "det":
            assert np.array_equal(words, ref) and np.array_equal(np.concatenate(cnt), cref)
        else:
            assert np.array_equal(words[:100], ref[:100])
        assert net.t == steps


def test_stream_stop(g4s):
    with g4s.Network(2000, 200, 5) as net:
        seen = []
        with pytest.raises(g4s.G4SError) as e:
            net.run_stream(500, lambda t0, r, c: seen.append(t0) or t0 >= 100, alg="AB", accum="det", chunk=50)
        assert e.value.code == -4 and seen[-1] == 100


def read_spiketimes(path, steps, n):
    r = np.zeros((steps, n), np.uint8)
    with open(path) as f:
        assert f.readline().strip() == "SpkTime NeuronID"
        for line in f:
            t, i = line.split()
            r[int(t), int(i)] = 1
    return r


def test_cli_result_files_det(g4s, oracle, tmp_path):
    n, nsyn, steps = 2500, 1000, 300
    out = tmp_path / "Results"
    p = subprocess.run([g4s.CLI_PATH, str(n), str(nsyn), str(steps), "--alg", "AB", "--regime", "1", "--mode", "det",
                        "--seed", "20221", "--outdir", str(out)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "Total number of spikes:" in p.stdout
    d = out / "N2500Nsyn1000Regime1Alg3"
    for f in ("neuroninfo.csv", "spiketimes.csv", "computetime.csv", "sim_overview.csv", "neuroninfoafter.csv"):
        assert (d / f).exists(), f
    onet = oracle.Network(n, nsyn, 20221)
    oraster, ocounts, _ = onet.run(steps)
    assert np.array_equal(read_spiketimes(d / "spiketimes.csv", steps, n), oraster)
    ct = np.loadtxt(d / "computetime.csv", skiprows=1)
    assert ct.shape == (steps, 4) and np.array_equal(ct[:, 1].astype(np.int64), ocounts)
    ov = (d / "sim_overview.csv").read_text()
    assert ov.startswith("N elapsed(ms) elapsedwithdata(ms) totalspkcount\n") and not ov.endswith("\n")
    assert int(ov.split("\n")[1].split()[3]) == int(ocounts.sum())
    after = np.loadtxt(d / "neuroninfoafter.csv", skiprows=1)
    assert after.shape == (n, 9) and np.array_equal(after[:, 8].astype(np.int64), onet.nospks)


def test_driver_layout(g4s, tmp_path):
    """The driver in a child process.  It runs the non-cooperative schemes
    here: this test process has already made cooperative launches (AP, SKL),
    and the runtime refuses them to a second process on the same device."""
    p = subprocess.run([sys.executable, os.path.join(ROOT, "gpu4snn.py"), "MI355X", "120", "--alg", "N"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    base = tmp_path / "MI355X" / "120" / "Results"
    runs = sorted(x.name for x in base.iterdir())
    assert runs == [f"N2500Nsyn1000Regime{r}Alg1" for r in range(3)]
    for r in runs:
        assert len(list((base / r).iterdir())) == 5
