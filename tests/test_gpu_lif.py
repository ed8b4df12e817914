"""The reference's third model, its test-only leaky integrate-and-fire
plug-in (sim/tests/lif.rs:14-129; registry id 2 here): row
[tau, e, r, vth, vreset], state [v], entry reset, kick gain 1.

GPU against the oracle: the hand-computed trace of lif.rs:131-195; the
mixed Izhikevich/LIF net of lif.rs:197-273 (spike counts exact as there,
and here v bit-exact with both rings); and a noisy random-cube network with
a fifth of its neurons LIF in every launch shape and parameter mode, int
ring raster bit-exact."""
import numpy as np
import pytest

from neuro_edit_v2_amd import Sim, SimConfig
from neuro_edit_v2_amd import net as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LIF_DEF = [20.0, -65.0, 1.0, -50.0, -70.0]     # sim/tests/lif.rs:67
# a LIF row that fires in the noisy cube (input resistance 10, threshold
# 7 mV above rest: a 20 mV kick moves v by 10 mV)
LIF_CUBE = [20.0, -65.0, 10.0, -58.0, -70.0]


def line(n):
    b = O.NetBuilder()
    for i in range(n):
        b.add_neuron((i * 50.0, 0.0, 0.0), O.NetBuilder.EXC)
    return b


def with_lif(net, frac, seed, rows=None):
    """The network with a fraction of its neurons switched to LIF (model
    table [izh, lif], rows widened to 5 params); `rows` distinct LIF rows
    (varied tau) instead of the default one."""
    rng = np.random.default_rng(seed)
    n = int(net["n"])
    pick = rng.random(n) < frac
    p4 = net["p"].reshape(n, net["n_params"])
    p = np.zeros((n, 5), np.float32)
    p[:, :p4.shape[1]] = p4[:, :5]
    p[pick] = LIF_CUBE
    if rows:
        p[pick, 0] = (10.0 + 20.0 * (np.arange(pick.sum()) % rows) / max(rows - 1, 1)).astype(np.float32)
    x0 = net["x0"].reshape(n, 2).copy()
    x0[pick] = [LIF_CUBE[1], 0.0]
    out = dict(net)
    out.update(model_ids=np.array([0, 2], np.uint8), model=pick.astype(np.uint8), n_params=5,
               p=p.ravel(), x0=x0.ravel().astype(np.float32))
    return out, int(pick.sum())


def test_lif_hand_trace_on_gpu():
    """lif.rs:131-195 on the GPU: -65 at rest, -64.5 / -64.025 under drive
    10, then under drive 100 the post-reset value -64.75 -- and every value
    equal to the oracle's."""
    b = line(1)
    b.set_model(0, 2, LIF_DEF)
    net = b.build(1.0)
    g = Sim(net, SimConfig(dt_ms=1.0, record_spikes=True))
    o = O.OracleSim(net, O.default_cfg(dt_ms=1.0, record_spikes=1))
    g.run_steps(3)
    o.run_steps(3)
    assert float(g.v()[0]) == -65.0
    g.set_external(0, 10.0)
    o.set_external(0, 10.0)
    for want in (-64.5, -64.025):
        g.run_steps(1)
        o.run_steps(1)
        assert abs(np.float32(g.v()[0]) - np.float32(want)) < 1e-6
        assert g.v()[0] == np.float32(o.v()[0])
    g.set_external(0, 100.0)
    o.set_external(0, 100.0)
    for _ in range(40):
        g.run_steps(1)
        o.run_steps(1)
        if g.spike_row(0)[0] & 1:
            break
    else:
        raise AssertionError("LIF must spike under drive 100")
    assert o.spike_row(0)[0] & 1
    assert float(g.v()[0]) >= -50.0
    g.run_steps(1)
    assert abs(np.float32(g.v()[0]) - np.float32(-64.75)) < 1e-6


@pytest.mark.parametrize("ring", [False, True], ids=["float", "int"])
def test_lif_mixed_net_matches_oracle(ring):
    """lif.rs:197-273: 8 neurons, LIF at 1, 3, 4, 6, synapses izh->lif,
    lif->izh, lif->lif, drives 15 + (0, 25, 18, 30); 300 steps, then 50
    watched.  Spike counts equal (as lif.rs asks), every LIF fires at least
    twice, and v equal to the oracle's bit for bit (same operation order,
    no contraction), then the watched rows too."""
    b = line(8)
    for i in (1, 3, 4, 6):
        b.set_model(i, 2, LIF_DEF)
    b.connect(0, 1, 8.0)
    b.connect(1, 2, 8.0)
    b.connect(3, 4, 5.0)
    net = b.build(0.1)
    g = Sim(net, SimConfig(record_spikes=True, int_weights=ring))
    o = O.OracleSim(net, O.default_cfg(record_spikes=1, int_weights=int(ring)))
    for i, s in [(0, 0.0), (1, 25.0), (3, 18.0), (6, 30.0)]:
        g.set_external(i, 15.0 + s)
        o.set_external(i, 15.0 + s)
    g.run_steps(300)
    o.run_steps(300)
    gc, oc = g.spike_counts(1 << 62), o.spike_counts(1 << 62)
    assert np.array_equal(gc, oc) and gc.sum() > 0
    assert all(gc[i] >= 2 for i in (1, 3, 6))
    assert np.array_equal(g.v(), o.v().astype(np.float32))
    ids = [1, 4]
    g.set_v_watch(ids)
    want_v, want_u = [], []
    for _ in range(50):                 # the oracle stepped one at a time
        o.run_steps(1)
        want_v.append(o.v()[ids].astype(np.float32))
        want_u.append(o.state(1)[ids].astype(np.float32))
    g.run_steps(50)
    dv, du = g.drain_watch()
    for c in range(2):
        assert np.array_equal(dv[c], np.array(want_v)[:, c])
        assert np.array_equal(du[c], np.array(want_u)[:, c])


@pytest.mark.parametrize("shape,rows", [(1, 0), (2, 0), (0, 0), (2, 300), (1, 40)],
                         ids=["fused", "paired", "two-kernel", "paired-per-neuron-rows", "fused-table"])
def test_lif_in_cube_network_bit_exact(shape, rows):
    """A 6,000-neuron random cube with a fifth of its neurons LIF, Gaussian
    current and kicks, int ring: raster, v and I equal to the oracle's in
    each launch shape, with the parameter rows in kernel arguments (two
    rows), a table (40 LIF rows) or per neuron (300)."""
    base = N.cube_net(6000, mean_deg=60.0, seed=21, w_scale=0.7)
    net, n_lif = with_lif(base, 0.2, seed=5, rows=rows)
    assert n_lif > 1000
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=30.0, noise_kick_mv=20.0, noise_sd=1.0,
              seed=9, int_weights=True, int_scale=65536.0)
    steps = 301
    print(f"case: lif cube shape {shape} rows {rows}", flush=True)
    g = Sim(net, SimConfig(raster_steps=steps, step_shape=shape, **kw))
    assert g.launch_shape() == shape
    g.run_steps(steps)
    _, cnt, ids = g.read_spikes()
    o = O.OracleSim(net, O.default_cfg(**kw))
    rc, ri = o.run_steps(steps, record=True)
    lif_ids = np.nonzero(net["model"] == 1)[0]
    assert np.isin(ri, lif_ids).sum() > 100        # the LIF neurons do fire
    assert np.array_equal(cnt, rc) and np.array_equal(ids, ri)
    assert np.array_equal(g.v(), o.v().astype(np.float32))
    assert np.array_equal(g.syn_current(), o.syn_current().astype(np.float32))
