"""LDS-binned push (nh_config.lds_bins, one step per launch): the
workgroup queues its spiking rows in LDS, walks them with all its lanes,
combines deliveries to the same ring word in an LDS hash table and sends
one HBM atomic per word (entries that find no bin go straight to HBM).

Int ring: raster, v and u bit-exact against the oracle -- random cube,
grid liquids with inputs and both models (local wiring, so deliveries of a
workgroup share targets), a small table that overflows (heavy firing, all
synapses of delay 1), every workgroup size -- and a packed batch equals its
networks alone.  Float ring: within the one-step shape's tolerances."""
import numpy as np
import pytest

from neuro_edit_v2_amd import Sim, SimConfig
from neuro_edit_v2_amd import net as N
from neuro_edit_v2_amd import workloads as W
from neuro_edit_v2_amd.batch import Batch
from oracle import oracle as O
from test_gpu_paired import all_delay1_net, grid_case

pytestmark = pytest.mark.gpu

CFG0 = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_sd=1.0, noise_rate_hz=20.0, noise_kick_mv=20.0,
            seed=7)


def check_int(net, steps, kw, bins, block=256, series=None, exp_mode=0):
    sim = Sim(net, SimConfig(raster_steps=steps, step_shape=1, lds_bins=bins, block=block, **kw))
    ref = O.OracleSim(net, O.default_cfg(exp_mode=exp_mode, **kw))
    if series is not None:
        for q in range(series.shape[1]):
            sim.play_input(q, series[:, q])
            ref.play_input(q, series[:, q])
    sim.run_steps(steps)
    _, cnt, ids = sim.read_spikes()
    rc, ri = ref.run_steps(steps, record=True)
    print(f"bins {bins} block {block}: spikes gpu={cnt.sum()} oracle={rc.sum()}")
    assert np.array_equal(cnt, rc) and np.array_equal(ids, ri)
    assert np.array_equal(sim.v(), ref.v().astype(np.float32))
    assert np.array_equal(sim.state(1), ref.state(1).astype(np.float32))
    assert np.array_equal(sim.syn_current(), ref.syn_current().astype(np.float32))
    return int(rc.sum())


@pytest.mark.parametrize("bins", [256, 1024, 8192])
def test_binned_cube_int_bit_exact(bins):
    net = N.cube_net(10000, mean_deg=100.0, seed=1, w_scale=0.7)
    kw = dict(CFG0, int_weights=True, int_scale=65536.0)
    print("case: binned cube 10k", bins, flush=True)
    assert check_int(net, 301, kw, bins) > 500


@pytest.mark.parametrize("model", [0, 1], ids=["izh", "adex"])
@pytest.mark.parametrize("block", [256, 512])
def test_binned_grid_liquid_bit_exact(model, block):
    net, bits = grid_case(model)
    kw = dict(CFG0, int_weights=True, int_scale=65536.0)
    print("case: binned grid liquid", model, block, flush=True)
    check_int(net, 300, kw, 1024, block=block, series=bits, exp_mode=1)


def test_binned_table_overflow_bit_exact():
    """Heavy firing into a 256-bin table: most deliveries find no bin and
    take the direct path; the raster is still the oracle's."""
    net = all_delay1_net(3000, 40, seed=5)
    kw = dict(CFG0, noise_rate_hz=400.0, int_weights=True, int_scale=4096.0)
    print("case: binned overflow", flush=True)
    assert check_int(net, 257, kw, 256) > 5000


def test_binned_batch_equals_alone():
    steps = 300
    sizes = [257, 1000, 64, 3000, 1331, 700]
    nets, series, cfgs = [], [], []
    for k, n in enumerate(sizes):
        nets.append(N.grid_net(0, model=k % 2, seed=k + 3, dt_ms=0.5, dims=W.grid_dims(n)))
        series.append(W.stream_window(W.rng_stream(100 + k, 8, 0.05), 8, steps))
        cfgs.append(SimConfig(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=15.0 + k, noise_sd=0.5,
                              seed=11 * k + 1, int_weights=True, raster_steps=steps, step_shape=1,
                              lds_bins=2048))
    print("case: binned batch", flush=True)
    b = Batch(nets, cfgs)
    for k in range(len(nets)):
        for q in range(8):
            b.play_input(k, q, series[k][:, q])
    b.run_steps(steps)
    for k, net in enumerate(nets):
        _, bc, bi = b.read_spikes(k)
        o = O.OracleSim(net, O.default_cfg(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=15.0 + k,
                                           noise_sd=0.5, seed=11 * k + 1, int_weights=1, exp_mode=1))
        for q in range(8):
            o.play_input(q, series[k][:, q])
        oc, oi = o.run_steps(steps, record=True)
        assert np.array_equal(oc, bc) and np.array_equal(oi, bi), f"net {k}"


def test_binned_float_ring_tolerance():
    net = N.cube_net(10000, mean_deg=100.0, seed=1, w_scale=0.7)
    kw = dict(CFG0, int_weights=False)
    sim = Sim(net, SimConfig(step_shape=1, lds_bins=1024, **kw))
    ref = O.OracleSim(net, O.default_cfg(**kw))
    worst = 0.0
    for _ in range(10):
        sim.run_steps(20)
        ref.run_steps(20)
        for g, r in ((sim.state(0), ref.state(0)), (sim.state(1), ref.state(1))):
            rel = np.abs(g.astype(np.float64) - r) / np.maximum(np.abs(r), 1.0)
            worst = max(worst, float(rel.max()))
    print(f"max rel error {worst:.2e}")
    assert worst < 1e-4
    sim.run_steps(1800)
    ref.run_steps(1800)
    g, r = sim.pn_total().astype(np.float64), ref.pn_total().astype(np.float64)
    assert np.abs(g - r).mean() / r.mean() < 0.02


def test_bad_bin_counts_refused():
    net = N.cube_net(500, mean_deg=20.0, seed=2)
    from neuro_edit_v2_amd._lib import NHError
    for bad in (100, 128, 16384):
        with pytest.raises(NHError, match="lds_bins"):
            Sim(net, SimConfig(step_shape=1, lds_bins=bad))
