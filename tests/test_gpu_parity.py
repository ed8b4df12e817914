"""GPU parity: the HIP path against the CPU oracle on identical seeded
inputs (the reference's own GPU tests compare CUDA against CpuSim the same
way: sim/tests/gpu_vs_cpu.rs:25-177).

Bar: the int ring gives a bit-exact raster (neuron id, step) and
bit-exact v/u; the float ring agrees within the stated tolerances.
"""
import numpy as np
import pytest

from neuro_edit_v2_amd import Sim, SimConfig
from neuro_edit_v2_amd import net as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# BASELINE config-0 dynamics at reduced size: dt 0.5, tau 5, Gaussian sd 1
# plus the reference's Poisson kicks (`--noise`), here 20 Hz for activity.
CFG0 = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_sd=1.0, noise_rate_hz=20.0, noise_kick_mv=20.0,
            seed=7)


def gpu_run(net, steps, **kw):
    sim = Sim(net, SimConfig(raster_steps=max(steps, 1), **kw))
    sim.run_steps(steps)
    _, cnt, ids = sim.read_spikes()
    return sim, cnt, ids


def oracle_run(net, steps, **kw):
    ref = O.OracleSim(net, O.default_cfg(**kw))
    rc, ri = ref.run_steps(steps, record=True)
    return ref, rc, ri


@pytest.fixture(scope="module")
def cube10k():
    return N.cube_net(10000, mean_deg=100.0, seed=1, w_scale=0.7)


@pytest.mark.parametrize("shape", [0, 1], ids=["two_kernel", "fused"])
@pytest.mark.parametrize("exec_mode", [1, 2], ids=["direct", "graph"])
@pytest.mark.parametrize("scale", [4096.0, 65536.0], ids=["q12", "q16"])
def test_int_raster_bit_exact(cube10k, shape, exec_mode, scale):
    steps = 300
    kw = dict(CFG0, int_weights=True, int_scale=scale)
    sim, cnt, ids = gpu_run(cube10k, steps, step_shape=shape, exec_mode=exec_mode, **kw)
    ref, rc, ri = oracle_run(cube10k, steps, **kw)
    print(f"spikes gpu={cnt.sum()} oracle={rc.sum()}")
    assert rc.sum() > 500, "workload too quiet to test the ring"
    assert np.array_equal(cnt, rc)
    assert np.array_equal(ids, ri)
    assert np.array_equal(sim.v(), ref.v().astype(np.float32))
    assert np.array_equal(sim.state(1), ref.state(1).astype(np.float32))
    assert sim.spikes_total() == ref.spikes_total()
    assert np.array_equal(sim.pn_total(), ref.pn_total())


def test_int_raster_reference_bench_net():
    # the reference benchmark network (sim/src/bench.rs:10-53), dt 0.1,
    # tau 2 ms, Poisson kicks 500 Hz (NOTES-for-data.md:2391 regime)
    net = N.bench_net(3000, 100, 1, 0.7, 5.0, 0.1)
    kw = dict(dt_ms=0.1, noise_rate_hz=500.0, noise_kick_mv=20.0, seed=1, int_weights=True)
    sim, cnt, ids = gpu_run(net, 1000, **kw)
    ref, rc, ri = oracle_run(net, 1000, **kw)
    print(f"spikes {cnt.sum()}")
    assert rc.sum() > 1000
    assert np.array_equal(cnt, rc) and np.array_equal(ids, ri)
    assert np.array_equal(sim.v(), ref.v().astype(np.float32))


def test_int_runs_bitwise_deterministic(cube10k):
    # sim/tests/gpu_vs_cpu.rs:117-141
    kw = dict(CFG0, int_weights=True)
    a, ca, ia = gpu_run(cube10k, 200, **kw)
    b, cb, ib = gpu_run(cube10k, 200, **kw)
    assert np.array_equal(ca, cb) and np.array_equal(ia, ib)
    assert np.array_equal(a.v(), b.v())


@pytest.mark.parametrize("shape", [0, 1], ids=["two_kernel", "fused"])
def test_float_ring_traces_and_counts(cube10k, shape):
    """Float ring: atomic order varies, so the bar is a tolerance -- v/u of
    every neuron within 1e-4 relative (denominator max(|x|, 1)) over the
    first 200 steps, per-neuron spike counts over 1 s within 2 % mean
    absolute deviation; the deviation from the FP64 oracle is printed."""
    kw = dict(CFG0)
    sim = Sim(cube10k, SimConfig(step_shape=shape, exec_mode=1, **kw))
    ref = O.OracleSim(cube10k, O.default_cfg(**kw))
    ref64 = O.OracleSim(cube10k, O.default_cfg(precision=64, **kw))
    worst = 0.0
    for blk in range(10):
        sim.run_steps(20)
        ref.run_steps(20)
        ref64.run_steps(20)
        for slot in (0, 1):
            g = sim.state(slot).astype(np.float64)
            r = ref.state(slot)
            rel = np.abs(g - r) / np.maximum(np.abs(r), 1.0)
            worst = max(worst, float(rel.max()))
    print(f"float ring, 200 steps: max relative v/u error vs f32 oracle {worst:.2e}")
    assert worst < 1e-4
    sim.run_steps(1800)
    ref.run_steps(1800)
    ref64.run_steps(1800)
    g = sim.pn_total().astype(np.float64)
    r = ref.pn_total().astype(np.float64)
    r64 = ref64.pn_total().astype(np.float64)
    mad = np.abs(g - r).mean() / max(r.mean(), 1e-9)
    mad64 = np.abs(r - r64).mean() / max(r64.mean(), 1e-9)
    print(f"1 s spike counts: MAD gpu-vs-f32 oracle {mad:.3%}, f32-vs-f64 oracle {mad64:.3%}, "
          f"totals gpu {g.sum():.0f} f32 {r.sum():.0f} f64 {r64.sum():.0f}")
    assert mad < 0.02


def test_smoke_entry():
    import __graft_entry__ as ge
    ge.smoke()
