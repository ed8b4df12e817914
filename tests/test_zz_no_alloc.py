"""nh_spike_counts_noalloc: the window counted into the library's static
scratch instead of a device buffer the library allocates -- the same counts
as nh_spike_counts, call after call (the second call on a smaller network
reads nothing left over from the first), and a clear error past the
scratch's size."""
import numpy as np
import pytest

from neuro_edit_v2_amd import Sim, SimConfig
from neuro_edit_v2_amd import net as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def driven(n, rows, seed):
    net = N.cube_net(n, mean_deg=50.0, seed=seed)
    kw = dict(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=60.0, seed=seed, int_weights=True,
              record_spikes=True, spike_record_steps=rows)
    return net, kw


def test_noalloc_counts_equal_allocating_counts_twice():
    print("case: noalloc counts 6000 then 700 neurons", flush=True)
    big_net, big_kw = driven(6000, 100, 3)
    small_net, small_kw = driven(700, 64, 4)
    big = Sim(big_net, SimConfig(**big_kw))
    big.run_steps(257)
    small = Sim(small_net, SimConfig(**small_kw))
    small.run_steps(150)
    for sim, rows in ((big, 100), (big, 37), (small, 64), (small, 9), (big, 1)):
        a = sim.spike_counts(rows, no_alloc=True)
        b = sim.spike_counts(rows)
        assert a.shape == b.shape and np.array_equal(a, b), rows
    assert int(big.spike_counts(100, no_alloc=True).sum()) > 100
    # and against the oracle for the smaller network
    o = O.OracleSim(small_net, O.default_cfg(**{k: v for k, v in small_kw.items()}))
    o.run_steps(150)
    assert np.array_equal(small.spike_counts(64, no_alloc=True), o.spike_counts(64))
    # window error and lifetime counts as the allocating entry point
    with pytest.raises(ValueError, match="exceeds"):
        small.spike_counts(65, no_alloc=True)
    assert np.array_equal(small.spike_counts(1 << 40, no_alloc=True), o.spike_counts(1 << 40))


def test_noalloc_refuses_networks_past_the_scratch():
    n = (1 << 22) + 1
    print(f"case: noalloc refuses {n} neurons", flush=True)
    net = dict(n=n, src_ptr=np.zeros(n + 1, np.uint32), tgt=np.zeros(0, np.uint32),
               w=np.zeros(0, np.float32), delay=np.zeros(0, np.uint16), tau=None,
               model_ids=np.zeros(1, np.uint8), model=np.zeros(n, np.uint8), n_params=4, n_state=2,
               p=np.tile(np.array([0.02, 0.2, -65.0, 8.0], np.float32), n),
               x0=np.tile(np.array([-65.0, -13.0], np.float32), n), n_inputs=0,
               stim_ptr=np.zeros(n + 1, np.uint32), stim_input=np.zeros(0, np.uint32),
               stim_w=np.zeros(0, np.float32))
    sim = Sim(net, SimConfig(dt_ms=0.5, record_spikes=True, spike_record_steps=8))
    sim.run_steps(3)
    with pytest.raises(RuntimeError, match="exceed the scratch"):
        sim.spike_counts(2, no_alloc=True)
    assert sim.spike_counts(2).shape == (n,)
    sim.close()
