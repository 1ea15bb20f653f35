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


def test_driver_default_sweep(g4s, tmp_path):
    """The drop-in driver with no flags, in a child process: the reference's
    default run (py:42, cu:1082-1083), 3 regimes x 5 schemes -- AP and SKL
    included, whose persistent grids are ordinary launches sized to the
    device's residency -- laid out as GPU_NAME/ITERATIONS/Results with the
    five files of every run."""
    steps = 200
    p = subprocess.run([sys.executable, os.path.join(ROOT, "gpu4snn.py"), "MI355X", str(steps)],
                       cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr[-2000:]
    assert p.stdout.count("Launched ") == 15
    base = tmp_path / "MI355X" / str(steps) / "Results"
    runs = sorted(x.name for x in base.iterdir())
    assert runs == sorted(f"N2500Nsyn1000Regime{r}Alg{a}" for r in range(3) for a in range(5))
    totals = {}
    for r in runs:
        d = base / r
        assert sorted(x.name for x in d.iterdir()) == sorted(
            ["neuroninfo.csv", "spiketimes.csv", "computetime.csv", "sim_overview.csv", "neuroninfoafter.csv"])
        ct = np.loadtxt(d / "computetime.csv", skiprows=1)
        assert ct.shape == (steps, 4)
        # per-step device times, not one average repeated
        assert (ct[:, 2] > 0).all() and (ct[:, 3] >= ct[:, 2]).all() and np.unique(ct[:, 2]).size > 1
        ov = (d / "sim_overview.csv").read_text()
        totals[r] = int(ov.split("\n")[1].split()[3])
        assert totals[r] == int(ct[:, 1].sum())
    for regime in (1, 2):  # the balanced and irregular runs fire under every scheme
        for a in range(5):
            name = f"N2500Nsyn1000Regime{regime}Alg{a}"
            assert totals[name] > 0, name
            spk = (base / name / "spiketimes.csv").read_text().splitlines()
            assert len(spk) == totals[name] + 1
