"""GPU: one network of N=1,000,000 neurons, Nsyn=1000
(1e9 synapses, 8 GB), thalamic gains x1.5 (the irregular regime, 7.5/3),
1000 steps with the persistent schemes AB and SKL in atomic mode.

FP32 voltage error against the FP64 oracle on a 10,000-neuron subsample
(8,000 excitatory and 2,000 inhibitory neurons around the E/I boundary),
teacher-forced by the GPU's own raster: the oracle integrates the subsample
in double from the same initial state with the presynaptic spikes taken from
the GPU run, so the comparison isolates the fp32 arithmetic.  Measured on MI355X (AB,
atomic): 88.9% of the subsample keep an identical spike train over the 1000
steps; for those |v_fp32 - v_fp64| at the last step has mean 1.5e-2 mV and
99th percentile 0.26 mV.  A shifted threshold crossing changes the rest of a
neuron's train (it is reset at a different step), which is what the
remaining 11% show.  The assertions are sanity bounds around those figures:
at least 80% identical trains, 99th percentile below 1 mV.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, NSYN, SEED, REGIME, STEPS = 1_000_000, 1000, 20221, 2, 1000
LO, HI = 4 * N // 5 - 8000, 4 * N // 5 + 2000


@pytest.fixture(scope="module")
def big_net(g4s):
    net = g4s.Network(N, NSYN, SEED, regime=REGIME)
    yield net
    net.close()


@pytest.fixture(scope="module")
def f64_oracle(oracle, big_net):
    post, w = big_net.synapses()
    return oracle.Network(N, NSYN, SEED, regime=REGIME, table=(post, w))


@pytest.mark.parametrize("alg", ["AB", "SKL"])
def test_n1e6_fp32_voltage_error(g4s, big_net, f64_oracle, alg):
    big_net.reset()
    # 256 slices requested (255 of 3936 neurons): SKL needs every slice co-resident
    words, counts = big_net.run(STEPS, alg=alg, accum="atomic", slices=256, block=1024)
    v, _, _, _ = big_net.state()
    rate = counts.sum() / N / (STEPS / 1000.0)
    o = f64_oracle
    tr, own = o.run_f64_forced_ranges(words, [(LO, HI)], 0, o.v[LO:HI], o.u[LO:HI])
    gpu_sub = g4s.raster_to_dense(words, N)[:, LO:HI]
    same = np.all(own == gpu_sub, axis=0)
    err = np.abs(tr[-1] - v[LO:HI].astype(np.float64))
    p99 = float(np.percentile(err[same], 99))
    print(f"\nN=1e6 {alg}: rate {rate:.2f} Hz, {counts.mean():.0f} spikes/step; identical spike trains "
          f"{same.mean():.4f}; |dv| mean {err[same].mean():.2e} p99 {p99:.2e} max {err[same].max():.2e} mV")
    assert 5.0 < rate < 60.0
    assert same.mean() >= 0.80
    assert p99 < 1.0


DET_STEPS = 100


@pytest.fixture(scope="module")
def det_oracle(oracle, big_net):
    """The fp32 oracle's first DET_STEPS steps on the device's synapse table."""
    post, w = big_net.synapses()
    onet = oracle.Network(N, NSYN, SEED, regime=REGIME, table=(post, w))
    oraster, ocounts, _ = onet.run(DET_STEPS, threads=16)
    v, u = onet.v.copy(), onet.u.copy()
    del onet, post, w
    return oraster, ocounts, v, u


@pytest.mark.parametrize("alg", ["AB", "SKL"])
def test_n1e6_det_bitwise(g4s, big_net, det_oracle, alg):
    """Det mode at N=10^6 (255 slices of 3936 neurons, 1024 threads: the large
    plan, about 16k spikes per step in two passes) is bit-identical to the fp32
    oracle run on the same synapse table for the first 100 steps."""
    oraster, ocounts, ov, ou = det_oracle
    big_net.reset()
    words, counts = big_net.run(DET_STEPS, alg=alg, accum="det", slices=256, block=1024)
    assert big_net.timing()["grid"] == 255
    assert np.array_equal(counts, ocounts), alg
    assert np.array_equal(g4s.raster_to_dense(words, N), oraster), alg
    v, u, _, _ = big_net.state()
    assert np.array_equal(v.view(np.uint32), ov.view(np.uint32)), alg
    assert np.array_equal(u.view(np.uint32), ou.view(np.uint32)), alg
