"""GPU: the batched form (g4s_batch_*).

det mode: every instance of a batch -- whatever else shares its launches,
whenever it joined -- gives the oracle's raster and final state bit for bit.
atomic mode: the first 100 steps equal the deterministic raster and the
spike total stays within 1% of it (the single-network bar).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (N, Nsyn, seed, thalamic scale, steps): uneven sizes and lengths
SPECS = [
    (300, 100, 11, 1.2, 250),
    (1000, 999, 12, 0.9, 200),
    (2500, 1000, 13, 1.0, 300),
    (4099, 100, 14, 1.4, 150),
    (777, 500, 15, 1.1, 260),
    (64, 63, 16, 1.5, 90),
]


def oracle_run(oracle, n, nsyn, seed, scale, steps):
    net = oracle.Network(n, nsyn, seed, regime=1, thal_scale=scale)
    r, c, _ = net.run(steps)
    return net, r, c


def check_instance(g4s, oracle, b, slot, n, nsyn, seed, scale, steps):
    onet, oraster, ocounts = oracle_run(oracle, n, nsyn, seed, scale, steps)
    info = b.info(slot)
    assert info["t"] == steps and info["done"] == 1
    words = b.raster(slot)
    assert np.array_equal(g4s.raster_to_dense(words, n), oraster), (n, nsyn, seed)
    assert np.array_equal(b.counts(slot), ocounts)
    v, u, _, nos = b.state(slot)
    assert np.array_equal(v.view(np.uint32), onet.v.view(np.uint32))
    assert np.array_equal(u.view(np.uint32), onet.u.view(np.uint32))
    assert np.array_equal(nos, onet.nospks)


@pytest.mark.parametrize("chunk,graph,schedule", [(64, 0, "queue"), (7, 0, "queue"), (64, 0, "step"),
                                                  (100, -1, "step"), (7, 0, "step")])
def test_batch_det_matches_oracle(g4s, oracle, chunk, graph, schedule):
    with g4s.Batch("det", chunk=chunk, graph=graph, schedule=schedule) as b:
        slots = [b.add(n, ns, sd, st, thal_scale=sc) for (n, ns, sd, sc, st) in SPECS]
        b.run()
        tm = b.timing()
        assert tm["steps"] == sum(s[4] for s in SPECS)
        for slot, (n, ns, sd, sc, st) in zip(slots, SPECS):
            check_instance(g4s, oracle, b, slot, n, ns, sd, sc, st)


@pytest.mark.parametrize("streams,graph", [(1, 0), (2, -1), (3, 0), (4, 0), (4, -1)])
def test_batch_stream_groups_det(g4s, oracle, streams, graph):
    """Instances split into groups, each group's launches on its own stream
    (graph-replayed or launched directly): every instance still matches the
    oracle bit for bit, also when a group is left with no instance (more
    groups than the last chunk's unfinished instances)."""
    with g4s.Batch("det", chunk=40, graph=graph, streams=streams) as b:
        slots = [b.add(n, ns, sd, st, thal_scale=sc) for (n, ns, sd, sc, st) in SPECS]
        b.run()
        assert b.timing()["launches"] >= 1
        for slot, (n, ns, sd, sc, st) in zip(slots, SPECS):
            check_instance(g4s, oracle, b, slot, n, ns, sd, sc, st)


def test_batch_stream_groups_invalid(g4s):
    with pytest.raises(g4s.G4SError):
        g4s.Batch("det", streams=5)


@pytest.mark.parametrize("slice_len", [64, 192, 1024, 2048])
def test_batch_det_slice_lengths(g4s, oracle, slice_len):
    specs = [(3000, 300, 21, 1.0, 120), (5000, 1000, 22, 1.3, 100)]
    with g4s.Batch("det", chunk=50, slice_len=slice_len) as b:
        slots = [b.add(n, ns, sd, st, thal_scale=sc) for (n, ns, sd, sc, st) in specs]
        b.run()
        for slot, (n, ns, sd, sc, st) in zip(slots, specs):
            assert b.info(slot)["slice_len"] == slice_len
            check_instance(g4s, oracle, b, slot, n, ns, sd, sc, st)


def test_batch_slice_length_beyond_lds(g4s):
    """det needs more LDS per slice than atomic: 4096-neuron slices at
    1024 threads (about 100 KB of det workspace) are refused in det mode
    (G4S_ERR_UNSUPPORTED, a batched workgroup may use 64 KB); at 256 threads
    they run in both modes (62.5 KB det, 34 KB atomic)."""
    with g4s.Batch("det", slice_len=4096, block=1024) as b:
        with pytest.raises(g4s.G4SError) as e:
            b.add(5000, 100, 1, 10)
        assert e.value.code == -5
    for accum, block in (("det", 256), ("atomic", 256)):
        with g4s.Batch(accum, slice_len=4096, block=block) as b:
            s = b.add(5000, 100, 1, 10)
            b.run()
            assert b.info(s)["done"] == 1 and b.info(s)["slices"] == 2, (accum, block)


def test_batch_continuous_join_and_slot_reuse(g4s, oracle):
    """Instances joining at chunk boundaries, and a freed slot reused, give
    the same results as alone."""
    a = (2000, 500, 31, 1.0, 300)
    bb = (1500, 200, 32, 1.2, 120)
    c = (900, 300, 33, 0.8, 80)
    with g4s.Batch("det", chunk=50) as b:
        sa = b.add(a[0], a[1], a[2], a[4], thal_scale=a[3])
        assert b.step() == 1 and b.step() == 1
        assert b.info(sa)["t"] == 100
        sb = b.add(bb[0], bb[1], bb[2], bb[4], thal_scale=bb[3])
        while b.info(sb)["done"] == 0:
            b.step()
        check_instance(g4s, oracle, b, sb, bb[0], bb[1], bb[2], bb[3], bb[4])
        b.remove(sb)
        sc = b.add(c[0], c[1], c[2], c[4], thal_scale=c[3])
        assert sc == sb
        b.run()
        check_instance(g4s, oracle, b, sa, a[0], a[1], a[2], a[3], a[4])
        check_instance(g4s, oracle, b, sc, c[0], c[1], c[2], c[3], c[4])
        assert b.step() == 0


def test_batch_queue_many_instances_det(g4s, oracle):
    """The queue schedule with more units than resident workgroups: 40
    instances, instances of 1 to 40 slices, chunk boundaries inside runs."""
    rng = np.random.default_rng(5)
    specs = []
    for i in range(40):
        n = int(rng.integers(200, 20000))
        nsyn = int(min(n - 1, rng.choice([50, 200, 400])))
        specs.append((n, nsyn, 1000 + i, float(rng.uniform(0.8, 1.5)), int(rng.integers(30, 130))))
    with g4s.Batch("det", chunk=50, slice_len=512) as b:
        slots = [b.add(n, ns, sd, st, thal_scale=sc) for (n, ns, sd, sc, st) in specs]
        b.run()
        for slot, (n, ns, sd, sc, st) in list(zip(slots, specs))[::4]:
            check_instance(g4s, oracle, b, slot, n, ns, sd, sc, st)
        for slot, (n, ns, sd, sc, st) in zip(slots, specs):
            assert b.info(slot)["done"] == 1


def test_batch_equals_single_network(g4s):
    """A batched instance and the same network run alone (AB det) agree."""
    n, nsyn, seed = 6000, 700, 41
    with g4s.Network(n, nsyn, seed, thal_scale=1.1) as net:
        words_alone, _ = net.run(400, alg="AB", accum="det")
        v_alone = net.state()[0]
    with g4s.Batch("det", chunk=100) as b:
        b.add(3000, 300, 42, 400, thal_scale=1.3)
        s = b.add(n, nsyn, seed, 400, thal_scale=1.1)
        b.add(700, 100, 43, 250)
        b.run()
        assert np.array_equal(b.raster(s), words_alone)
        assert np.array_equal(b.state(s)[0].view(np.uint32), v_alone.view(np.uint32))


@pytest.mark.parametrize("schedule", ["queue", "step"])
def test_batch_atomic_first_steps_and_rates(g4s, oracle, schedule):
    """First 100 steps identical per instance; the spike total of the three
    instances over 8000 steps (~460k spikes, so decorrelated runs differ by
    ~0.2% rms and the 1% bar is over 4 sigma) within 1% of the oracle's."""
    specs = [(2500, 1000, 51, 1.0, 8000), (2500, 1000, 52, 1.0, 8000), (4000, 500, 53, 1.2, 8000)]
    with g4s.Batch("atomic", chunk=500, schedule=schedule) as b:
        slots = [b.add(n, ns, sd, st, thal_scale=sc) for (n, ns, sd, sc, st) in specs]
        b.run()
        tot_gpu = tot_ref = 0
        for slot, (n, ns, sd, sc, st) in zip(slots, specs):
            _, oraster, ocounts = oracle_run(oracle, n, ns, sd, sc, st)
            dense = g4s.raster_to_dense(b.raster(slot)[:100], n)
            assert np.array_equal(dense, oraster[:100]), (n, sd)
            tot_gpu += int(b.counts(slot).sum())
            tot_ref += int(ocounts.sum())
        assert abs(tot_gpu - tot_ref) <= 0.01 * tot_ref, (tot_gpu, tot_ref)


def test_batch_budget_admission(g4s):
    fp = g4s.footprint(3000, 300, chunk=50)
    with g4s.Batch("det", chunk=50, budget=fp + fp // 2) as b:
        s0 = b.add(3000, 300, 61, 50)
        with pytest.raises(g4s.G4SError) as e:
            b.add(3000, 300, 62, 50)
        assert e.value.code == -7
        b.remove(s0)
        s1 = b.add(3000, 300, 62, 50)
        b.run()
        assert b.info(s1)["done"] == 1


def test_batch_counts_only_and_invalid(g4s):
    with g4s.Batch("det", chunk=64) as b:
        s = b.add(1000, 100, 71, 100, keep_raster=False)
        b.run()
        assert b.counts(s).shape == (100,)
        with pytest.raises(g4s.G4SError):
            b.raster(s)
        with pytest.raises(g4s.G4SError):
            b.info(99)
        with pytest.raises(g4s.G4SError):
            b.add(10, 10, 1, 5)  # Nsyn >= N


def test_batch_tune_sets_bucket_plan(g4s, oracle):
    """g4s_batch_tune picks a slice length for the (N, Nsyn) bucket; later
    instances of the bucket use it, with unchanged det results."""
    plan = g4s.batch_tune(3000, 300, accum="det", trial_steps=40, force=True)
    assert plan["slice_len"] in (128, 256, 512, 768, 1024, 2048) and plan["copies"] >= 4
    assert plan["us_per_inst_step"] > 0
    again = g4s.batch_tune(3000, 300, accum="det", trial_steps=40)
    assert again["slice_len"] == plan["slice_len"] and again["copies"] == 0  # cached
    with g4s.Batch("det", chunk=50) as b:
        s = b.add(3500, 300, 77, 120)  # same bucket: floor(log2 N) = 11
        b.run()
        assert b.info(s)["slice_len"] == min(plan["slice_len"], 3520)
        check_instance(g4s, oracle, b, s, 3500, 300, 77, 1.0, 120)


def test_batched_workload_at_its_size(g4s, oracle):
    """The benchmark's batched workload itself: 256 networks, N log-uniform on
    [1e3, 5e4], Nsyn in {100, 500, 1000}, thalamic x U(0.8, 1.5) (seed
    20221), at the benchmark's chunk and default tiling, 200 steps.  Det: one
    instance of every (floor(log2 N), Nsyn) bucket -- at least 16 -- matches
    the oracle's per-step counts and final state bit for bit.  Atomic: the
    whole batch's spike total is within 1% of det's."""
    from bench import batch_specs
    specs = batch_specs(256, 20221)
    steps = 200
    first = {}
    for k, (n, ns, _, _) in enumerate(specs):
        first.setdefault((n.bit_length() - 1, ns), k)
    checked = sorted(first.values())
    assert len(checked) >= 16, len(checked)
    totals = {}
    for accum in ("det", "atomic"):
        with g4s.Batch(accum, chunk=100) as b:
            slots = [b.add(n, ns, sd, steps, thal_scale=sc, keep_raster=False) for (n, ns, sd, sc) in specs]
            b.run()
            totals[accum] = sum(int(b.counts(s).sum()) for s in slots)
            if accum == "det":
                for k in checked:
                    n, ns, sd, sc = specs[k]
                    post, w = g4s.build_synapses_host(n, ns, sd)
                    onet = oracle.Network(n, ns, sd, regime=1, thal_scale=sc, table=(post, w))
                    _, ocounts, _ = onet.run(steps, raster=False)
                    assert np.array_equal(b.counts(slots[k]), ocounts), (k, n, ns)
                    v, u, _, nos = b.state(slots[k])
                    assert np.array_equal(v.view(np.uint32), onet.v.view(np.uint32)), (k, n, ns)
                    assert np.array_equal(u.view(np.uint32), onet.u.view(np.uint32)), (k, n, ns)
                    assert np.array_equal(nos, onet.nospks), (k, n, ns)
    assert abs(totals["atomic"] - totals["det"]) <= 0.01 * totals["det"], totals


def test_batch_tune_mix_plan_is_used_and_exact(g4s, oracle):
    """g4s_batch_tune_mix times the mix itself, keeps its plan only when it
    beats every instance at min(768, N), records it per bucket, and later
    batches use it; det results do not depend on the slicing."""
    rng = np.random.default_rng(9)
    specs = []
    for i in range(24):
        n = int(np.exp(rng.uniform(np.log(1500), np.log(30000))))
        specs.append((n, int(rng.choice([100, 500])), 500 + i, float(rng.uniform(0.8, 1.5))))
    plan = g4s.batch_tune_mix(specs, accum="det", trial_steps=40, max_trials=8)
    assert plan["trials"] >= 3 and plan["buckets"] >= 1
    assert plan["us_per_step_tuned"] <= plan["us_per_step_default"] + 1e-6
    n, ns, sd, sc = specs[0]
    with g4s.Batch("det", chunk=50) as b:
        s = b.add(n, ns, sd, 90, thal_scale=sc)
        b.run()
        assert b.info(s)["slice_len"] % 64 == 0
        check_instance(g4s, oracle, b, s, n, ns, sd, sc, 90)
