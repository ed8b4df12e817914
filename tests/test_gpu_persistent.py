"""Many steps per launch (step_shape=3): one persistent kernel per chunk of
nh_run_steps, the delay-1 arrivals pulled from published spike words, the
other deliveries gated on the launch's own progress tags.

Int ring: raster, v, u and I are the oracle's bit for bit -- random cube
(few delay-1 synapses), all synapses of delay 1 (more than 64 into a
wave), random networks (sizes, degrees, delay mixes, Izhikevich/AdEx,
noise, scales, workgroup sizes) -- for runs cut into chunks of every length
(one-step launches, two-step launches, long launches) and across raster
drains.  Float ring: within the fused shape's tolerances.  Where the shape
cannot run (stimulus inputs, watched neurons) the step falls back to one
launch per step.  The bounded waits' failure path, forced from the host,
ends the launch and reports; a reset recovers.
"""
import numpy as np
import pytest

from neuro_edit_v2_amd import Sim, SimConfig
from neuro_edit_v2_amd import _lib as L
from neuro_edit_v2_amd import net as N
from oracle import oracle as O
from test_gpu_paired import all_delay1_net, grid_case, random_net

pytestmark = pytest.mark.gpu

CFG0 = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_sd=1.0, noise_rate_hz=20.0, noise_kick_mv=20.0,
            seed=7)


def run_chunks(sim, chunks):
    for c in chunks:
        sim.run_steps(c)


def check_int(net, chunks, kw, block=0, exp_mode=0):
    steps = int(sum(chunks))
    sim = Sim(net, SimConfig(raster_steps=steps, step_shape=3, block=block, **kw))
    assert sim.launch_shape() == 3
    run_chunks(sim, chunks)
    _, cnt, ids = sim.read_spikes()
    ref = O.OracleSim(net, O.default_cfg(exp_mode=exp_mode, **kw))
    rc, ri = ref.run_steps(steps, record=True)
    print(f"spikes gpu={cnt.sum()} oracle={rc.sum()}")
    assert np.array_equal(cnt, rc) and np.array_equal(ids, ri)
    assert np.array_equal(sim.v(), ref.v().astype(np.float32))
    assert np.array_equal(sim.state(1), ref.state(1).astype(np.float32))
    assert np.array_equal(sim.syn_current(), ref.syn_current().astype(np.float32))
    assert np.array_equal(sim.pn_total(), ref.pn_total())
    return int(rc.sum())


@pytest.mark.parametrize("chunks", [[301], [1, 2, 3, 1, 7, 64, 1, 222], [20] * 15 + [1]],
                         ids=["one_launch", "mixed_chunks", "twenty_step_launches"])
def test_persistent_cube_int_bit_exact(chunks):
    net = N.cube_net(10000, mean_deg=100.0, seed=1, w_scale=0.7)
    kw = dict(CFG0, int_weights=True, int_scale=65536.0)
    print("case: persistent cube 10k", chunks, flush=True)
    assert check_int(net, chunks, kw) > 500


@pytest.mark.parametrize("scale", [4096.0, 65536.0], ids=["q12", "q16"])
def test_persistent_all_delay1_bit_exact(scale):
    net = all_delay1_net(3000, 40, seed=5)
    print("case: persistent all delay 1", scale, flush=True)
    kw = dict(CFG0, int_weights=True, int_scale=scale)
    assert check_int(net, [257], kw) > 500


@pytest.mark.parametrize("seed", [101, 202, 303, 404, 505, 606])
def test_persistent_random_networks_bit_exact(seed):
    r = np.random.default_rng(seed + 1)
    net = random_net(seed)
    kw = dict(dt_ms=float(r.choice([0.1, 0.25, 0.5])), synapse_tau_ms=float(r.uniform(1, 10)),
              noise_rate_hz=float(r.uniform(5, 60)), noise_kick_mv=20.0,
              noise_sd=float(r.choice([0.0, 0.5, 1.0])), seed=int(r.integers(1, 1 << 30)),
              int_weights=True, int_scale=float(r.choice([4096.0, 65536.0])))
    steps = int(r.integers(50, 400))
    block = int(r.choice([64, 128, 256, 512]))
    cut = int(r.integers(1, steps))
    print(f"case: persistent random net seed {seed}: n={net['n']} E={len(net['tgt'])} "
          f"steps={steps} cut={cut} block={block} {kw}", flush=True)
    check_int(net, [cut, steps - cut], kw, block=block, exp_mode=1)


def test_persistent_headline_100k_full_run():
    """The 100k headline network, Q16.16, 2,000 steps in launches of 20."""
    net = N.cube_net(100000, mean_deg=100.0, seed=1, w_scale=0.7)
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_sd=1.0, noise_rate_hz=1.0, noise_kick_mv=20.0,
              seed=1, int_weights=True, int_scale=65536.0)
    print("case: persistent 100k, 2000 steps", flush=True)
    assert check_int(net, [20] * 100, kw) > 50000


def test_persistent_float_ring_tolerance():
    net = N.cube_net(10000, mean_deg=100.0, seed=1, w_scale=0.7)
    kw = dict(CFG0, int_weights=False)
    sim = Sim(net, SimConfig(step_shape=3, **kw))
    assert sim.launch_shape() == 3
    ref = O.OracleSim(net, O.default_cfg(**kw))
    worst = 0.0
    for _ in range(10):
        sim.run_steps(20)
        ref.run_steps(20)
        for g, r in ((sim.state(0), ref.state(0)), (sim.state(1), ref.state(1)),
                     (sim.syn_current(), ref.syn_current())):
            rel = np.abs(g.astype(np.float64) - r) / np.maximum(np.abs(r), 1.0)
            worst = max(worst, float(rel.max()))
    print(f"max rel error {worst:.2e}")
    assert worst < 1e-4
    sim.run_steps(1800)
    ref.run_steps(1800)
    g, r = sim.pn_total().astype(np.float64), ref.pn_total().astype(np.float64)
    assert np.abs(g - r).mean() / r.mean() < 0.02


def test_persistent_with_raster_drains_mid_run():
    net = N.cube_net(8000, mean_deg=80.0, seed=9, w_scale=0.7)
    kw = dict(CFG0, int_weights=True)
    sim = Sim(net, SimConfig(raster_steps=37, step_shape=3, **kw))
    assert sim.launch_shape() == 3
    for n in (101, 64, 136):
        sim.run_steps(n)
    _, cnt, ids = sim.read_spikes()
    rc, ri = O.OracleSim(net, O.default_cfg(**kw)).run_steps(301, record=True)
    assert np.array_equal(cnt, rc) and np.array_equal(ids, ri)


def test_persistent_falls_back_with_inputs():
    """Stimulus inputs are outside the persistent kernel: one step per
    launch, the oracle's raster."""
    net, bits = grid_case(0)
    kw = dict(CFG0, int_weights=True)
    sim = Sim(net, SimConfig(raster_steps=300, step_shape=3, **kw))
    assert sim.launch_shape() == 1
    ref = O.OracleSim(net, O.default_cfg(**kw))
    for q in range(bits.shape[1]):
        sim.play_input(q, bits[:, q])
        ref.play_input(q, bits[:, q])
    sim.run_steps(300)
    _, cnt, ids = sim.read_spikes()
    rc, ri = ref.run_steps(300, record=True)
    assert np.array_equal(cnt, rc) and np.array_equal(ids, ri)


def test_persistent_reset_and_rerun():
    net = N.cube_net(5000, mean_deg=60.0, seed=2, w_scale=0.7)
    kw = dict(CFG0, int_weights=True)
    sim = Sim(net, SimConfig(raster_steps=400, step_shape=3, **kw))
    sim.run_steps(150)
    sim.read_spikes()
    sim.reset_state()
    sim.run_steps(250)
    _, cnt, ids = sim.read_spikes()
    rc, ri = O.OracleSim(net, O.default_cfg(**kw)).run_steps(250, record=True)
    assert np.array_equal(cnt, rc) and np.array_equal(ids, ri)


def test_persistent_wait_gives_up_and_reports():
    """Failure path, forced from the host: wave 0 withholds its published
    words and the poll limit is 64, so the waves it feeds give up, and the
    others' progress waits then give up too; the launch ends, the next
    synchronisation reports the failure, and a reset recovers."""
    net = all_delay1_net(3000, 40, seed=5)
    kw = dict(CFG0, int_weights=True)
    sim = Sim(net, SimConfig(raster_steps=200, step_shape=3, **kw))
    assert sim.launch_shape() == 3
    L.check(L.lib().nh_test_pair_wait(sim._h, 64, 1, None))
    print("case: persistent wait, wave 0 muted, 64 polls", flush=True)
    sim.run_steps(20)
    with pytest.raises(L.NHError) as e:
        sim.synchronize()
    assert e.value.code == L.NH_ERR_HIP and "reset" in str(e.value)
    L.check(L.lib().nh_test_pair_wait(sim._h, 0, 0, None))
    sim.reset_state()
    sim.run_steps(200)
    _, cnt, ids = sim.read_spikes()
    rc, ri = O.OracleSim(net, O.default_cfg(**kw)).run_steps(200, record=True)
    assert np.array_equal(cnt, rc) and np.array_equal(ids, ri)
