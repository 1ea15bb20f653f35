"""CPU checks of the build-time kernel resource table (no GPU compute).

The runtime sizes every grid whose workgroups wait on each other (SKL's grid
barrier, AP's work queue) from the compiler's resource report of the exact
kernel instantiation it launches.  These tests check that the table the
build generated covers every such instantiation, that its occupancy figures
agree with the register limits of gfx950, and that the report parser reads
the compiler's remark format.
"""
import pytest

from gpu4snn_amd import build as gbuild

RS = (1, 2, 4, 8, 16)


def slice_targs():
    for r in RS:
        for det in ("false", "true"):
            for large in (("false", "true") if r >= 2 else ("false",)):
                yield f"{r},{det},{large}"


def vgpr_waves(vgprs, agprs):
    """Waves per SIMD that gfx950's 512 unified vector registers per lane
    hold (allocation granule 8)."""
    regs = vgprs + agprs
    return min(8, 512 // max(8, (regs + 7) // 8 * 8))


def sgpr_waves(sgprs):
    """Waves per SIMD the scalar register file holds on gfx9-family SIMDs
    (the AMDGPU backend's published table: <= 80 SGPRs: 10, <= 88: 9,
    <= 100: 8, otherwise 7), capped at 8 waves per SIMD on gfx950."""
    w = 10 if sgprs <= 80 else 9 if sgprs <= 88 else 8 if sgprs <= 100 else 7
    return min(8, w)


@pytest.mark.parametrize("name", ["k_skl_slice", "k_slice_step"])
def test_table_covers_slice_kernels(g4s, name):
    for targs in slice_targs():
        r = g4s.kernel_resources(name, targs)
        assert r["waves_per_simd"] >= 1 and r["vgprs"] > 0, (name, targs, r)
        # __launch_bounds__(1024): four waves per SIMD must fit
        assert r["waves_per_simd"] >= 4, (name, targs, r)


def test_one_neuron_per_thread_slice_kernels_do_not_spill(g4s):
    """The slice kernels that the default single-network plans run (one
    neuron per thread: slices up to the workgroup size) and the batched
    kernels keep every value in registers: no scratch memory.  A build whose
    per-step det kernel spilled gave wrong det results."""
    names = [(n, f"1,{det},false") for n in ("k_skl_slice", "k_slice_step") for det in ("false", "true")]
    names += [("k_batch_step", f"{det},{b}") for det in ("false", "true") for b in (256, 512, 1024)]
    names += [("k_batch_chunk", det) for det in ("false", "true")]
    for name, targs in names:
        r = g4s.kernel_resources(name, targs)
        assert r["scratch"] == 0, (name, targs, r)


def test_table_covers_persistent_ap(g4s):
    r = g4s.kernel_resources("k_ap", "")
    assert 1 <= r["waves_per_simd"] <= 8


def test_occupancy_agrees_with_register_limits(g4s):
    names = [("k_ap", ""), ("k_update", ""), ("k_push_rows", "")]
    names += [("k_skl_slice", t) for t in slice_targs()] + [("k_slice_step", t) for t in slice_targs()]
    for name, targs in names:
        r = g4s.kernel_resources(name, targs)
        limit = min(vgpr_waves(r["vgprs"], r["agprs"]), sgpr_waves(r["sgprs"]))
        assert r["waves_per_simd"] <= limit, (name, targs, r, limit)


def test_unknown_kernel_rejected(g4s):
    with pytest.raises(g4s.G4SError) as e:
        g4s.kernel_resources("k_skl_slice", "3,false,false")
    assert e.value.code == -1


def test_parse_resource_remarks():
    text = """\
x.hip:10:1: remark: Function Name: _ZN3g4s12_GLOBAL__N_14k_apENS_6DevNetEiiii [-Rpass-analysis=kernel-resource-usage]
   10 | {
x.hip:10:1: remark:     TotalSGPRs: 106 [-Rpass-analysis=kernel-resource-usage]
x.hip:10:1: remark:     VGPRs: 35 [-Rpass-analysis=kernel-resource-usage]
x.hip:10:1: remark:     AGPRs: 0 [-Rpass-analysis=kernel-resource-usage]
x.hip:10:1: remark:     ScratchSize [bytes/lane]: 0 [-Rpass-analysis=kernel-resource-usage]
x.hip:10:1: remark:     Dynamic Stack: False [-Rpass-analysis=kernel-resource-usage]
x.hip:10:1: remark:     Occupancy [waves/SIMD]: 7 [-Rpass-analysis=kernel-resource-usage]
x.hip:10:1: remark:     SGPRs Spill: 0 [-Rpass-analysis=kernel-resource-usage]
x.hip:10:1: remark:     VGPRs Spill: 0 [-Rpass-analysis=kernel-resource-usage]
x.hip:10:1: remark:     LDS Size [bytes/block]: 68 [-Rpass-analysis=kernel-resource-usage]
"""
    recs = gbuild.parse_resource_remarks(text)
    assert recs == [{"mangled": "_ZN3g4s12_GLOBAL__N_14k_apENS_6DevNetEiiii", "sgprs": 106, "vgprs": 35,
                     "agprs": 0, "scratch": 0, "occupancy": 7, "lds": 68}]
    assert gbuild.kernel_key("g4s::(anonymous namespace)::k_ap(g4s::DevNet, int, int, int, int)") == ("k_ap", "")
    assert gbuild.kernel_key("void g4s::(anonymous namespace)::k_skl_slice<1, false, true>(g4s::DevNet, int, int, "
                             "int)") == ("k_skl_slice", "1,false,true")
    assert gbuild.kernel_key("std::vector<int>::size() const") is None


def test_queue_schedule_rejects_unpackable_chunk(g4s):
    """The queue schedule packs a work unit's chunk step in 16 bits."""
    with pytest.raises(g4s.G4SError) as e:
        g4s.Batch("det", chunk=70000, schedule="queue")
    assert e.value.code == -5
