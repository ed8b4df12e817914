"""Host-side checks made before any device call (nh_check_net, run first
by nh_create, nh_batch_create and nh_server_submit): a valid network
passes, and each planted defect is refused with NH_ERR_ARG and a reason
(the checks of NeuronParams::check, sim/src/state.rs:197-217, and of
Network::validate's synapse and input rules, net-format/src/network.rs:
1074-1150, on the flat layout)."""
import ctypes as C

import numpy as np
import pytest

from neuro_edit_v2_amd import _lib as L
from neuro_edit_v2_amd import net as N
from neuro_edit_v2_amd.sim import SimConfig


def check(net, **cfg):
    d, keep = L.make_desc(net)
    c = SimConfig(**cfg).to_c()
    rc = L.lib().nh_check_net(C.byref(d), C.byref(c))
    return rc, L.lib().nh_last_error().decode()


@pytest.fixture(params=[True, False], ids=["tau", "no_tau"])
def net(request):
    """With a per-synapse tau array the synapse checks run one synapse at a
    time; without it they run as reductions first (and one synapse at a
    time only to name a failure), so every defect is planted in both."""
    n = N.grid_net(5, n_inputs=2, per_input=4, seed=3)
    if request.param:
        n["tau"] = np.zeros(len(n["tgt"]), np.float32)
    return n


def test_valid_network_passes(net):
    assert check(net)[0] == L.NH_OK
    assert check(net, int_weights=True, int_scale=65536.0)[0] == L.NH_OK


def _mut(net, key, fn):
    v = net[key].copy()
    fn(v)
    net[key] = v
    return net


@pytest.mark.parametrize("defect,cfg,reason", [
    (lambda n: _mut(n, "tgt", lambda a: a.__setitem__(3, n["n"])), {}, "target out of range"),
    (lambda n: _mut(n, "delay", lambda a: a.__setitem__(0, 0)), {}, "delay must be >= 1"),
    (lambda n: _mut(n, "w", lambda a: a.__setitem__(1, np.nan)), {}, "non-finite synapse weight"),
    (lambda n: _mut(n, "tau", lambda a: a.__setitem__(2, -1.0)), {}, "tau must be finite"),
    (lambda n: _mut(n, "src_ptr", lambda a: a.__setitem__(5, a[6] + 1)), {}, "src_ptr not monotone"),
    (lambda n: _mut(n, "model", lambda a: a.__setitem__(0, 3)), {}, "outside the model table"),
    (lambda n: _mut(n, "model_ids", lambda a: a.__setitem__(0, 7)), {}, "unknown model id"),
    (lambda n: _mut(n, "p", lambda a: a.__setitem__(4, np.inf)), {}, "non-finite parameter"),
    (lambda n: _mut(n, "x0", lambda a: a.__setitem__(0, np.nan)), {}, "non-finite initial state"),
    (lambda n: _mut(n, "stim_input", lambda a: a.__setitem__(0, 9)), {}, "stimulus input index"),
    (lambda n: n, dict(dt_ms=0.0), "dt_ms outside"),
    (lambda n: n, dict(dt_ms=1000.0), "dt_ms outside"),
    (lambda n: n, dict(noise_sd=-1.0), "noise rate, kick and sd"),
    (lambda n: n, dict(int_weights=True, int_scale=0.0), "int_scale must be positive"),
    (lambda n: _mut(n, "w", lambda a: a.__setitem__(0, 1e6)), dict(int_weights=True, int_scale=65536.0),
     "exceeds the int32 ring"),
    (lambda n: _mut(n, "w", lambda a: a.__setitem__(len(a) - 1, -np.inf)), {}, "non-finite synapse weight"),
    (lambda n: _mut(n, "w", lambda a: a.__setitem__(7, -np.float32(np.nan))), {}, "non-finite synapse weight"),
    (lambda n: _mut(n, "tgt", lambda a: a.__setitem__(len(a) - 1, 0xFFFFFFFF)), {}, "target out of range"),
    (lambda n: _mut(n, "x0", lambda a: a.__setitem__(len(a) - 1, -np.inf)), {}, "non-finite initial state"),
    (lambda n: dict(_mut(n, "model_ids", lambda a: a.__setitem__(0, 1)), n_params=4), {},
     "adex rows need 10 params"),
], ids=["tgt", "delay0", "w_nan", "tau_neg", "src_ptr", "model", "model_id", "param_inf", "x0_nan",
        "stim_input", "dt0", "dt_big", "sd_neg", "scale0", "int_overflow", "w_neg_inf", "w_neg_nan",
        "tgt_last", "x0_last_inf", "adex_short_rows"])
def test_planted_defect_is_refused(net, defect, cfg, reason):
    if reason.startswith("tau") and net.get("tau") is None:
        pytest.skip("no tau array to plant the defect in")
    rc, msg = check(defect(net), **cfg)
    assert rc == L.NH_ERR_ARG, msg
    assert reason in msg, msg


def test_n_syn_must_match_src_ptr(net):
    d, keep = L.make_desc(net)
    d.n_syn = d.n_syn + 1
    c = SimConfig().to_c()
    assert L.lib().nh_check_net(C.byref(d), C.byref(c)) == L.NH_ERR_ARG
    assert "n_syn" in L.lib().nh_last_error().decode()


def test_create_checks_before_touching_the_device(net):
    """Without a GPU, a defective network is still reported as defective
    (the checks run first), a valid one as 'no device'."""
    from conftest import gpu_present
    if gpu_present():
        pytest.skip("a GPU is present")
    from neuro_edit_v2_amd import Sim
    bad = _mut(dict(net), "delay", lambda a: a.__setitem__(0, 0))
    with pytest.raises(L.NHError) as e:
        Sim(bad)
    assert e.value.code == L.NH_ERR_ARG
    with pytest.raises(L.NHError) as e:
        Sim(net)
    assert e.value.code == L.NH_ERR_NO_DEVICE


def test_int_ring_fan_in_overflow_refused():
    """Int ring: a slot of a neuron sums at most one arrival per in-synapse,
    so the check refuses a network whose in-synapse weights x int_scale sum
    past 2^31 - 1 for some neuron (the ring would wrap silently), accepts it
    at a scale where the sum fits, and accepts it with the float ring."""
    n = N.grid_net(5, n_inputs=2, per_input=4, seed=3)
    w = n["w"].copy()
    hub = int(np.bincount(n["tgt"]).argmax())
    into = np.flatnonzero(n["tgt"] == hub)
    assert len(into) >= 3
    w[into] = np.sign(w[into]) * 15000.0     # each |w| * 65536 < 2^31, their sum is not
    n["w"] = w
    rc, msg = check(n, int_weights=True, int_scale=65536.0)
    assert rc == L.NH_ERR_NET and "sum past the int32" in msg and f"neuron {hub}" in msg
    assert check(n, int_weights=True, int_scale=4096.0)[0] == L.NH_OK
    assert check(n, int_weights=False)[0] == L.NH_OK
