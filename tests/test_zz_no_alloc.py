"""GPU: g4s_run_noalloc, the entry point that allocates nothing, against
g4s_run on the same networks.

The caller owns device buffers for the raster rows and counts (torch tensors
here); the run's scratch comes from the library's static device array.  Two
networks are run back to back, the second smaller, so stale contents of the
shared array from the first call would show in the second.  Det runs are
bit-identical to g4s_run; atomic runs keep the deterministic raster for
their first 30 steps.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# (N, Nsyn, seed, steps, chunk): the second call is the smaller input
CALLS = [(5000, 500, 11, 420, 150), (2500, 1000, 12, 130, 0)]


def run_noalloc(g4s, net, alg, accum, steps, chunk=0, slices=0, block=0):
    words = g4s.words_of(net.n)
    raster = torch.full((steps * words,), -1, dtype=torch.int32, device="cuda")
    counts = torch.full((steps,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    opts = g4s.RunOpts(block, slices, chunk, 0, 0)
    g4s._check(g4s.lib().g4s_run_noalloc(net._h, g4s.ALG[alg], g4s.ACCUM[accum], int(steps), raster.data_ptr(),
                                         counts.data_ptr(), C.byref(opts)))
    r = raster.cpu().numpy().view(np.uint32).reshape(steps, words)
    return r, counts.cpu().numpy()


@pytest.mark.parametrize("alg,accum,kw", [("SKL", "det", {}), ("AB", "det", {}), ("SKL", "det", dict(slices=64)),
                                          ("N", "atomic", {}), ("AP", "atomic", {}), ("S", "det", {})])
def test_noalloc_equals_run(g4s, alg, accum, kw):
    for (n, nsyn, seed, steps, chunk) in CALLS:
        with g4s.Network(n, nsyn, seed) as ref:
            rw, rc = ref.run(steps, alg=alg, accum="det" if accum == "atomic" else accum, chunk=chunk, **kw)
            rv = ref.state()[0]
        with g4s.Network(n, nsyn, seed) as net:
            net.run(0, alg=alg, accum=accum, chunk=chunk, **kw)  # prepares slices and staging
            w, c = run_noalloc(g4s, net, alg, accum, steps, chunk=chunk, **kw)
            assert net.t == steps
            v = net.state()[0]
            dense = g4s.raster_to_dense(w, n)
            assert np.array_equal(c, dense.sum(axis=1)), (alg, n)
            if accum == "det":
                assert np.array_equal(w, rw), (alg, n)
                assert np.array_equal(c, rc), (alg, n)
                assert np.array_equal(v.view(np.uint32), rv.view(np.uint32)), (alg, n)
            else:
                k = 30  # a reordered addition moves spikes only later at these sizes
                assert np.array_equal(dense[:k], g4s.raster_to_dense(rw[:k], n)), (alg, n)
            # the handle's own buffers are back: g4s_run continues the same run
            net.run(20, alg=alg, accum=accum)


def test_noalloc_requires_prepared_handle(g4s):
    with g4s.Network(3000, 300, 5) as net:
        with pytest.raises(g4s.G4SError) as e:
            run_noalloc(g4s, net, "SKL", "det", 10)
        assert e.value.code == -4  # no slice partition yet
        net.run(0, alg="SKL", accum="det")
        with pytest.raises(g4s.G4SError) as e:
            words = g4s.words_of(net.n)
            raster = torch.zeros(10 * words, dtype=torch.int32, device="cuda")
            opts = g4s.RunOpts(0, 0, 0, 0, 1)  # kernel_events
            g4s._check(g4s.lib().g4s_run_noalloc(net._h, 4, 1, 10, raster.data_ptr(), None, C.byref(opts)))
        assert e.value.code == -1
        w, c = run_noalloc(g4s, net, "SKL", "det", 10)
        assert c.shape == (10,)
