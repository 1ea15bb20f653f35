"""GPU: the device connectivity builder equals the host builder bit for bit
(and therefore the oracle's sampler), including dense rows where most draws
collide (Nsyn = N-1) and the largest supported row (Nsyn = 1024)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,nsyn,seed", [(2, 1, 3), (65, 64, 4), (130, 7, 5), (1000, 999, 6), (1025, 1024, 7),
                                         (4099, 100, 8), (2500, 1000, 20221), (20000, 37, 9), (3000, 1024, 10)])
def test_device_table_equals_host(g4s, n, nsyn, seed):
    hp, hw = g4s.build_synapses_host(n, nsyn, seed)
    dp, dw = g4s.build_synapses_device(n, nsyn, seed)
    assert np.array_equal(dp, hp), (n, nsyn)
    assert np.array_equal(dw.view(np.uint32), hw.view(np.uint32)), (n, nsyn)


def test_device_table_equals_oracle(g4s, oracle):
    n, nsyn, seed = 777, 500, 31
    onet = oracle.Network(n, nsyn, seed)
    dp, dw = g4s.build_synapses_device(n, nsyn, seed)
    assert np.array_equal(dp, onet.post) and np.array_equal(dw.view(np.uint32), onet.w.view(np.uint32))


def test_device_table_large_rows_sorted_distinct(g4s):
    """N=100,000, Nsyn=1000: equal to the host table; rows strictly
    ascending, no self-connection."""
    n, nsyn, seed = 100_000, 1000, 20221
    dp, dw = g4s.build_synapses_device(n, nsyn, seed)
    hp, hw = g4s.build_synapses_host(n, nsyn, seed)
    assert np.array_equal(dp, hp) and np.array_equal(dw.view(np.uint32), hw.view(np.uint32))
    rows = dp.reshape(n, nsyn)
    assert np.all(np.diff(rows.astype(np.int64), axis=1) > 0)
    assert not np.any(rows == np.arange(n, dtype=np.uint32)[:, None])


def test_device_build_rejects_wide_rows(g4s):
    with pytest.raises(g4s.G4SError) as e:
        g4s.build_synapses_device(3000, 1025, 1)
    assert e.value.code == -5
