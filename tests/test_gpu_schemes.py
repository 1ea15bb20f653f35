"""GPU: scheme switching, chunking, edge sizes and larger networks.

Deterministic mode is compared with the oracle bit for bit wherever the
oracle runs in seconds; at N=100,000 the five schemes must agree with each
other bit for bit (and with the oracle over a short horizon), and atomic mode
must keep the firing rate within 1% of the deterministic run.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ALGS = ["AP", "N", "S", "AB", "SKL"]


def oracle_raster(oracle, n, nsyn, steps, seed, regime=1, scale=1.0, v_init=None):
    net = oracle.Network(n, nsyn, seed, regime=regime, thal_scale=scale)
    if v_init is not None:
        net.v[:] = v_init
    r, c, last = net.run(steps)
    return net, r, c


def dense(g4s, words, n):
    return g4s.raster_to_dense(words, n)


@pytest.mark.parametrize("n,nsyn", [(2, 1), (65, 64), (130, 7), (1000, 999), (4099, 100)])
def test_det_edge_sizes(g4s, oracle, n, nsyn):
    onet, oraster, ocounts = oracle_raster(oracle, n, nsyn, 300, seed=n + nsyn, regime=2)
    for alg in ALGS:
        with g4s.Network(n, nsyn, n + nsyn, regime=2) as net:
            words, counts = net.run(300, alg=alg, accum="det")
            assert np.array_equal(dense(g4s, words, n), oraster), (alg, n, nsyn)
            v, u, _, nos = net.state()
            assert np.array_equal(v.view(np.uint32), onet.v.view(np.uint32)), alg
            assert np.array_equal(nos, onet.nospks), alg


def test_det_switching_schemes_and_chunks(g4s, oracle):
    n, nsyn, seed = 3000, 300, 5
    onet, oraster, _ = oracle_raster(oracle, n, nsyn, 700, seed)
    plan = [("SKL", 100, 0), ("N", 50, 7), ("AP", 150, 0), ("S", 100, 33), ("AB", 120, 0), ("SKL", 180, 11)]
    with g4s.Network(n, nsyn, seed) as net:
        rows = []
        for alg, steps, chunk in plan:
            w, _ = net.run(steps, alg=alg, accum="det", chunk=chunk)
            rows.append(dense(g4s, w, n))
        assert np.array_equal(np.concatenate(rows), oraster)
        v, u, _, _ = net.state()
        assert np.array_equal(u.view(np.uint32), onet.u.view(np.uint32))


def test_atomic_switching_keeps_pending_input(g4s, oracle):
    """Switching among schemes that keep the pending input in HBM (N, S, AP)
    and those that derive it from the queue (AB, SKL) or bitmap (det) must
    not lose or double the input of the step at the switch."""
    n, nsyn, seed = 2500, 1000, 9
    _, oraster, _ = oracle_raster(oracle, n, nsyn, 100, seed)
    plan = [("N", 10), ("SKL", 10), ("S", 10), ("AB", 10), ("AP", 10), ("SKL", 10), ("N", 10), ("AP", 10),
            ("S", 10), ("AB", 10)]
    with g4s.Network(n, nsyn, seed) as net:
        rows = [dense(g4s, net.run(k, alg=a, accum="atomic")[0], n) for a, k in plan]
    got = np.concatenate(rows)
    assert np.array_equal(got, oraster)


@pytest.mark.parametrize("graph,chunk", [(-1, 0), (0, 37), (0, 0)])
def test_ab_graph_replay_and_chunks(g4s, oracle, graph, chunk):
    """AB per-step launches replayed from a captured HIP graph (captured again
    for a new chunk length) or launched directly give the same dynamics."""
    n, nsyn, seed = 2500, 1000, 31
    _, oraster, ocounts = oracle_raster(oracle, n, nsyn, 300, seed)
    with g4s.Network(n, nsyn, seed) as net:
        a = dense(g4s, net.run(100, alg="AB", graph=graph, chunk=chunk)[0], n)
        w, c = net.run(200, alg="AB", graph=graph, chunk=chunk)
    assert np.array_equal(a, oraster[:100])
    total = a.sum() + c.sum()
    assert abs(total - ocounts.sum()) <= 0.01 * ocounts.sum()


def test_det_then_atomic_rebuilds_input(g4s, oracle):
    n, nsyn, seed = 2500, 1000, 13
    _, oraster, _ = oracle_raster(oracle, n, nsyn, 80, seed)
    with g4s.Network(n, nsyn, seed) as net:
        a = dense(g4s, net.run(40, alg="SKL", accum="det")[0], n)
        b = dense(g4s, net.run(40, alg="N", accum="atomic")[0], n)
    assert np.array_equal(np.concatenate([a, b]), oraster)


def test_nan_injection_spikes_and_resets(g4s, oracle):
    n, nsyn, seed = 256, 10, 3
    v0 = np.full(n, -65.0, np.float32)
    v0[[0, 77, 255]] = np.nan
    onet, oraster, _ = oracle_raster(oracle, n, nsyn, 20, seed, v_init=v0)
    for alg in ALGS:
        for accum in ("det", "atomic"):
            with g4s.Network(n, nsyn, seed) as net:
                net.set_state(v=v0)
                words, _ = net.run(20, alg=alg, accum=accum)
                d = dense(g4s, words, n)
                assert d[0, [0, 77, 255]].all(), (alg, accum)
                if accum == "det":
                    assert np.array_equal(d, oraster), alg


def test_quiet_regime_and_zero_steps(g4s, oracle):
    n, nsyn, seed = 2500, 1000, 21
    _, oraster, ocounts = oracle_raster(oracle, n, nsyn, 500, seed, regime=0)
    with g4s.Network(n, nsyn, seed, regime=0) as net:
        w, c = net.run(0, alg="SKL")
        assert w.shape[0] == 0 and net.t == 0
        w, c = net.run(500, alg="SKL", accum="det")
        assert np.array_equal(dense(g4s, w, n), oraster) and c.sum() == ocounts.sum()


@pytest.fixture(scope="module")
def big():
    """N=100,000, Nsyn=1000, balanced regime: the headline network."""
    return dict(n=100_000, nsyn=1000, seed=20221)


BIG_STEPS = 1000


@pytest.fixture(scope="module")
def big_oracle(g4s, oracle, big):
    """The oracle's whole 1000-step run of the N=100,000 network (threaded
    form, bit-identical to the single-thread oracle), on the host builder's
    table (equal to the oracle's own builder; checked against the device
    table below)."""
    post, w = g4s.build_synapses_host(big["n"], big["nsyn"], big["seed"])
    onet = oracle.Network(big["n"], big["nsyn"], big["seed"], table=(post, w))
    r, c, _ = onet.run(BIG_STEPS, threads=16)
    return onet, r, c


def test_n100k_det_all_schemes_identical_and_oracle(g4s, big, big_oracle):
    """Every scheme's det run of the whole configuration (1000 steps) is
    bit-identical to the oracle: raster, counts, v, u, spike counts."""
    n, nsyn, seed = big["n"], big["nsyn"], big["seed"]
    onet, oraster, ocounts = big_oracle
    with g4s.Network(n, nsyn, seed) as net:
        post, w = net.synapses()
        assert np.array_equal(post, onet.post) and np.array_equal(w.view(np.uint32), onet.w.view(np.uint32))
        del post, w
        for alg in ALGS:
            net.reset()
            words, counts = net.run(BIG_STEPS, alg=alg, accum="det")
            assert np.array_equal(counts, ocounts), (alg, int(np.argmax(counts != ocounts)))
            assert np.array_equal(dense(g4s, words, n), oraster), alg
            v, u, _, nos = net.state()
            assert np.array_equal(v.view(np.uint32), onet.v.view(np.uint32)), alg
            assert np.array_equal(u.view(np.uint32), onet.u.view(np.uint32)), alg
            assert np.array_equal(nos, onet.nospks), alg


@pytest.mark.parametrize("alg", ["AB", "SKL"])
def test_n100k_bench_plan_det_bitwise(g4s, big, big_oracle, alg):
    """The tiling the bench times (256 slices requested, 1024 threads per
    workgroup) is bit-identical to the oracle in det mode over the whole
    1000-step configuration."""
    n, nsyn, seed = big["n"], big["nsyn"], big["seed"]
    onet, oraster, _ = big_oracle
    with g4s.Network(n, nsyn, seed) as net:
        words, _ = net.run(BIG_STEPS, alg=alg, accum="det", slices=256, block=1024)
        assert net.timing()["grid"] == 241  # slices of 416 neurons
        assert np.array_equal(dense(g4s, words, n), oraster)
        v, u, _, _ = net.state()
        assert np.array_equal(v.view(np.uint32), onet.v.view(np.uint32))
        assert np.array_equal(u.view(np.uint32), onet.u.view(np.uint32))


# Atomic mode at N=100,000 leaves the oracle's raster once a reordered
# binary32 addition moves a near-threshold neuron across 30 mV.  Over five
# seeds and every scheme (three runs each) that happened after 36 to 167
# steps, at the same steps at which a sequential CPU run in descending
# presynaptic order leaves it (test_oracle_model.py); for this seed after
# 48, 58 or 114 steps.  The bar asserted here is the lower bound that holds
# for this seed, with margin; 100 steps holds at N=2500 (test_gpu_parity).
N100K_ATOMIC_IDENTICAL_STEPS = 40


@pytest.mark.parametrize("plan", ["default", "bench"])
def test_n100k_atomic_identity_and_rate_vs_oracle(g4s, big, big_oracle, plan):
    """Each scheme in atomic mode: the oracle's raster for the first
    N100K_ATOMIC_IDENTICAL_STEPS steps, and its 1000-step spike total within
    1% of the oracle's."""
    n, nsyn, seed = big["n"], big["nsyn"], big["seed"]
    _, oraster, ocounts = big_oracle
    ref = int(ocounts.sum())
    with g4s.Network(n, nsyn, seed) as net:
        for alg in ALGS:
            if plan == "bench" and alg not in ("AB", "SKL"):
                continue
            kw = dict(slices=256, block=1024) if plan == "bench" else {}
            net.reset()
            words, c = net.run(BIG_STEPS, alg=alg, accum="atomic", **kw)
            k = N100K_ATOMIC_IDENTICAL_STEPS
            d = dense(g4s, words[:k], n)
            assert np.array_equal(d, oraster[:k]), (alg, plan, int(np.argmax(np.any(d != oraster[:k], axis=1))))
            assert abs(int(c.sum()) - ref) <= 0.01 * ref, (alg, plan, int(c.sum()), ref)


@pytest.mark.parametrize("accum", ["det", "atomic"])
def test_n100k_skl_two_slices_per_cu(g4s, big, big_oracle, accum):
    """SKL's grid is sized from the compiler's occupancy of the kernel it
    launches: at 512 threads the slice kernel holds two workgroups per CU, so
    447 slices of 224 neurons (more than the 256 CUs) are all resident and
    the grid barrier completes.  Det: the oracle bit for bit over the whole
    configuration; atomic: the single-network atomic bar."""
    n, nsyn, seed = big["n"], big["nsyn"], big["seed"]
    onet, oraster, ocounts = big_oracle
    with g4s.Network(n, nsyn, seed) as net:
        words, c = net.run(BIG_STEPS, alg="SKL", accum=accum, slices=512, block=512)
        li = net.launch_info()
        assert li["grid"] == 447 and li["persistent"] == 1, li
        assert li["resident_per_cu"] >= 2 and li["grid"] <= li["resident_per_cu"] * li["cus"], li
        assert li["grid"] > li["cus"], li
        if accum == "det":
            assert np.array_equal(dense(g4s, words, n), oraster)
            v, _, _, _ = net.state()
            assert np.array_equal(v.view(np.uint32), onet.v.view(np.uint32))
        else:
            k = N100K_ATOMIC_IDENTICAL_STEPS
            assert np.array_equal(dense(g4s, words[:k], n), oraster[:k])
            assert abs(int(c.sum()) - int(ocounts.sum())) <= 0.01 * int(ocounts.sum())


def test_step_times_and_launch_info(g4s):
    """Per-step device times (the reference's computetime.csv rows) sum to
    the run's device time, and the launch report names the kernel run."""
    with g4s.Network(20000, 500, 3) as net:
        for alg in ALGS:
            net.reset()
            net.run(300, alg=alg, accum="atomic", chunk=128)
            st = net.step_times()
            tm = net.timing()
            assert st.shape == (300,) and (st > 0).all(), alg
            # stamps start at each step's first kernel: the chunk's sum is
            # within a few microseconds per chunk of the event-timed total
            assert abs(float(st.sum()) - tm["device_ms"]) <= 0.05 * tm["device_ms"] + 0.05, (alg, st.sum(), tm)
            li = net.launch_info()
            want = {"AP": "k_ap", "N": "k_push_rows", "S": "k_push_flat", "AB": "k_slice_step<",
                    "SKL": "k_skl_slice<"}[alg]
            assert li["kernel"].startswith(want) and li["waves_per_simd"] >= 1, (alg, li)


@pytest.mark.parametrize("alg", ALGS)
def test_kernel_events_do_not_change_results(g4s, oracle, alg):
    """Event-bracketed launches (roofline timing) simulate the same steps."""
    n, nsyn, seed = 3000, 300, 77
    _, oraster, _ = oracle_raster(oracle, n, nsyn, 200, seed)
    with g4s.Network(n, nsyn, seed) as net:
        words, _ = net.run(200, alg=alg, accum="det", kernel_events=True, chunk=64)
        assert np.array_equal(dense(g4s, words, n), oraster), alg
        tm = net.timing()
        assert tm["kernel_launches"] >= 1 and tm["kernel_ms"] > 0.0
        assert tm["kernel_ms"] <= tm["device_ms"] * 1.05 + 0.05
        net.reset()
        _, c = net.run(200, alg=alg, accum="atomic", kernel_events=True)
        assert net.timing()["kernel_launches"] >= 1


def test_autotune_plan_runs_and_is_cached(g4s, oracle):
    n, nsyn, seed = 4000, 200, 5
    _, oraster, _ = oracle_raster(oracle, n, nsyn, 100, seed)
    with g4s.Network(n, nsyn, seed) as net:
        p = net.autotune("det", trial_steps=30)
        assert p.accum == 1 and 0 <= p.alg <= 4 and p.us_per_step > 0
        assert net.t == 0
        words, _ = net.run_plan(p, 100)
        assert np.array_equal(dense(g4s, words, n), oraster)
        pa = net.autotune("atomic", trial_steps=30)
        assert pa.accum == 0 and pa.us_per_step > 0
        q = net.autotune("atomic", trial_steps=30)   # bucket hit: cached plan
        assert (q.alg, q.opts.block, q.opts.slices) == (pa.alg, pa.opts.block, pa.opts.slices)
        assert q.us_per_step == pa.us_per_step
        net.reset()
        _, c = net.run_plan(pa, 300, raster=False)
        assert c.sum() > 0


def test_default_tiling_fits_det_workspace(g4s, oracle):
    """Default AB tiling at N=4M: one slice per CU would need more than the
    160 KB of LDS of one workgroup for det's workspace, so the runtime doubles
    the slice count until it fits; the run stays bit-identical to the oracle,
    and atomic mode keeps one slice per CU."""
    n, nsyn, seed, steps = 4_000_000, 50, 5, 20
    onet = oracle.Network(n, nsyn, seed)
    oraster, ocounts, _ = onet.run(steps)
    with g4s.Network(n, nsyn, seed) as net:
        words, counts = net.run(steps, alg="AB", accum="det")
        p_det = net.timing()["grid"]
        assert np.array_equal(counts, ocounts)
        assert np.array_equal(g4s.raster_to_dense(words, n), oraster)
        v, u, _, _ = net.state()
        assert np.array_equal(v.view(np.uint32), onet.v.view(np.uint32))
        net.reset()
        net.run(steps, alg="AB", accum="atomic")
        p_atomic = net.timing()["grid"]
    # lengths round up to multiples of 32: 511 slices of 7840 vs 256 of 15648
    assert p_det > 3 * p_atomic // 2 > 0, (p_det, p_atomic)


@pytest.mark.parametrize("block,slices", [(64, 11), (256, 11), (64, 32)])
def test_slices_of_odd_word_pairs(g4s, oracle, block, slices):
    """Slice lengths are multiples of 32 neurons (whole bitmap words), not of
    64: at N=3000 with 11 slices requested each slice holds 288 neurons, so
    the last wave of a slice covers 32 neurons of the next one and must leave
    that slice's bitmap word alone; with 32 slices each holds 96 neurons, and
    at 64 threads every slice's second wave spans two slices.  Det is bit-identical to the oracle; the
    atomic run keeps the oracle's raster for its first 100 steps."""
    n, nsyn, seed, steps = 3000, 300, 41, 300
    onet, oraster, _ = oracle_raster(oracle, n, nsyn, steps, seed, regime=2)
    for alg in ("AB", "SKL"):
        with g4s.Network(n, nsyn, seed, regime=2) as net:
            words, _ = net.run(steps, alg=alg, accum="det", slices=slices, block=block)
            assert net.timing()["grid"] == slices
            assert np.array_equal(dense(g4s, words, n), oraster), (alg, block)
            v, _, _, _ = net.state()
            assert np.array_equal(v.view(np.uint32), onet.v.view(np.uint32)), (alg, block)
            net.reset()
            words, _ = net.run(100, alg=alg, accum="atomic", slices=slices, block=block)
            assert np.array_equal(dense(g4s, words, n), oraster[:100]), (alg, block)


@pytest.mark.parametrize("block", [512, 1024])
def test_det_slices_around_the_wave_local_limit(g4s, oracle, block):
    """Det slices just under 1024 neurons (960: the wave-local stable
    placement with its largest per-target rows) and over it (1504: ordered
    windows), at N=30,000 with 500 synapses per row (about 16 entries per
    spike and slice).  With the thalamic drive at 2.5x, busy steps overflow a
    wave's 384 register entries and take the windows while quiet ones stay
    wave-local, so one run mixes both.  Bit-identical to the oracle."""
    n, nsyn, seed, steps = 30_000, 500, 17, 200
    for thal in (1.0, 2.5):
        onet = oracle.Network(n, nsyn, seed, regime=1, thal_scale=thal)
        oraster, ocounts, _ = onet.run(steps, threads=8)
        for slices in (32, 20):
            for alg in ("AB", "SKL"):
                with g4s.Network(n, nsyn, seed, regime=1, thal_scale=thal) as net:
                    words, counts = net.run(steps, alg=alg, accum="det", slices=slices, block=block)
                    assert net.timing()["grid"] == slices, (alg, slices)
                    assert np.array_equal(counts, ocounts), (alg, block, slices, thal)
                    assert np.array_equal(g4s.raster_to_dense(words, n), oraster), (alg, block, slices, thal)
                    v, _, _, _ = net.state()
                    assert np.array_equal(v.view(np.uint32), onet.v.view(np.uint32)), (alg, block, slices, thal)


@pytest.mark.parametrize("block", [256, 1024])
def test_large_slices_both_modes(g4s, oracle, block):
    """16 slices of 3136 neurons at N=50,000.  Atomic mode runs the large
    plan (slices of 2048 neurons or more: 8 spikes per thread per pass, 8
    entry loads in flight per thread); det mode at 1024 threads needs 86 KB of
    LDS per workgroup.  Det is bit-identical to the oracle in the balanced and
    irregular regimes.  Atomic mode keeps the oracle's raster for the first 10
    steps (a lost or misplaced segment changes it at once) and the spike total
    within 1%; in the irregular regime (about 1200 spikes per step) a step
    needs two spike passes at 256 threads."""
    n, nsyn, seed, steps = 50_000, 1000, 7, 300
    for regime in (1, 2):
        onet = oracle.Network(n, nsyn, seed, regime=regime)
        oraster, ocounts, _ = onet.run(steps, threads=8)
        for alg in ("AB", "SKL"):
            with g4s.Network(n, nsyn, seed, regime=regime) as net:
                words, counts = net.run(steps, alg=alg, accum="det", slices=16, block=block)
                assert net.timing()["grid"] == 16
                assert np.array_equal(g4s.raster_to_dense(words, n), oraster), (alg, block, regime)
                v, _, _, _ = net.state()
                assert np.array_equal(v.view(np.uint32), onet.v.view(np.uint32)), (alg, block, regime)
                net.reset()
                words, counts = net.run(steps, alg=alg, accum="atomic", slices=16, block=block)
                assert net.timing()["grid"] == 16
                d = g4s.raster_to_dense(words[:10], n)
                assert np.array_equal(d, oraster[:10]), (alg, block, regime)
                assert abs(int(counts.sum()) - int(ocounts.sum())) <= 0.01 * ocounts.sum(), (alg, block, regime)


def test_det_nsyn_above_device_builder(g4s, oracle):
    """Nsyn above the device builder's 1024 (the table is built on the host
    and uploaded) through every scheme: det bit-identical to the oracle."""
    n, nsyn, seed, steps = 3000, 2000, 17, 150
    onet, oraster, ocounts = oracle_raster(oracle, n, nsyn, steps, seed)
    for alg in ALGS:
        with g4s.Network(n, nsyn, seed) as net:
            words, counts = net.run(steps, alg=alg, accum="det")
            assert np.array_equal(counts, ocounts), alg
            assert np.array_equal(dense(g4s, words, n), oraster), alg
            v, _, _, _ = net.state()
            assert np.array_equal(v.view(np.uint32), onet.v.view(np.uint32)), alg
