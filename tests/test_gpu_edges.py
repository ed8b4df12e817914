"""Edge cases of the step kernels against the oracle, int ring bit for bit:
a ring scale that is not a power of two (the read divides instead of
multiplying by the exact inverse), networks whose size is not a multiple of
64 (a partial last wave) down to one neuron, and hub neurons whose rows
hold thousands of synapses (the push walks many 256-entry spans, several
spiking lanes of one wave at once), in the one-step and paired shapes."""
import numpy as np
import pytest

from neuro_edit_v2_amd import Sim, SimConfig
from neuro_edit_v2_amd import net as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CFG0 = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_sd=1.0, noise_rate_hz=20.0, noise_kick_mv=20.0,
            seed=7)


def izh_net(n, deg_of, seed, maxd=8):
    """Izhikevich network: out-degree deg_of[i], random targets, delays
    1..maxd (a third of them 1), signed weights."""
    r = np.random.default_rng(seed)
    deg = np.asarray(deg_of, np.int64)
    E = int(deg.sum())
    src_ptr = np.zeros(n + 1, np.uint32)
    src_ptr[1:] = np.cumsum(deg)
    w = r.uniform(0.2, 3.0, E).astype(np.float32)
    w[r.random(E) < 0.2] *= -1.0
    delay = np.where(r.random(E) < 0.33, 1, r.integers(1, maxd + 1, E)).astype(np.uint16)
    return dict(n=n, src_ptr=src_ptr, tgt=r.integers(0, n, E, dtype=np.uint32), w=w, delay=delay,
                model_ids=np.zeros(1, np.uint8), model=np.zeros(n, np.uint8), n_params=4, n_state=2,
                p=np.tile(np.float32([0.02, 0.2, -65.0, 8.0]), n),
                x0=np.tile(np.float32([-65.0, -13.0]), n))


def run_both(net, steps, kw, shape, exp_mode=0):
    sim = Sim(net, SimConfig(raster_steps=steps, step_shape=shape, **kw))
    got = sim.launch_shape()
    sim.run_steps(steps)
    _, cnt, ids = sim.read_spikes()
    ref = O.OracleSim(net, O.default_cfg(exp_mode=exp_mode, **kw))
    rc, ri = ref.run_steps(steps, record=True)
    print(f"shape {got}: spikes gpu={cnt.sum()} oracle={rc.sum()}")
    assert np.array_equal(cnt, rc) and np.array_equal(ids, ri)
    assert np.array_equal(sim.v(), ref.v().astype(np.float32))
    assert np.array_equal(sim.state(1), ref.state(1).astype(np.float32))
    assert np.array_equal(sim.syn_current(), ref.syn_current().astype(np.float32))
    assert np.array_equal(sim.pn_total(), ref.pn_total())
    return got, int(rc.sum())


@pytest.mark.parametrize("shape", [1, 2], ids=["fused", "paired"])
@pytest.mark.parametrize("scale", [1000.0, 30000.0], ids=["s1000", "s30000"])
def test_int_scale_not_power_of_two(shape, scale):
    net = N.cube_net(10000, mean_deg=100.0, seed=1, w_scale=0.7)
    kw = dict(CFG0, int_weights=True, int_scale=scale)
    print(f"case: cube 10k, int scale {scale}, shape {shape}", flush=True)
    got, spikes = run_both(net, 301, kw, shape)
    assert got == shape and spikes > 500


@pytest.mark.parametrize("shape", [1, 2], ids=["fused", "paired"])
@pytest.mark.parametrize("n", [1, 63, 65, 130, 1000])
def test_sizes_not_multiple_of_64(shape, n):
    r = np.random.default_rng(n)
    net = izh_net(n, r.poisson(12, n) if n > 1 else [1], seed=n)
    kw = dict(CFG0, noise_rate_hz=80.0, int_weights=True, int_scale=65536.0)
    print(f"case: n {n}, shape {shape}", flush=True)
    got, spikes = run_both(net, 257, kw, shape)
    assert got == shape and spikes > 0


@pytest.mark.parametrize("shape", [1, 2], ids=["fused", "paired"])
def test_hub_rows_span_many_push_batches(shape):
    """20 hubs with 1,500-5,000 out-synapses among 4,000 neurons of degree
    ~20; hubs sit in two waves, so one wave often pushes several long rows
    together."""
    n = 4000
    r = np.random.default_rng(11)
    deg = r.poisson(20, n)
    hubs = np.r_[np.arange(5, 15), np.arange(70, 80)]
    deg[hubs] = r.integers(1500, 5000, len(hubs))
    net = izh_net(n, deg, seed=12, maxd=12)
    kw = dict(CFG0, noise_rate_hz=60.0, int_weights=True, int_scale=65536.0)
    print(f"case: hubs, shape {shape}, {int(deg.sum())} synapses", flush=True)
    got, spikes = run_both(net, 301, kw, shape)
    assert got == shape and spikes > 1000
