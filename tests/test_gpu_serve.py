"""Served form: requests admitted while others run, batched into the same
step launches (two steps per launch); each request's result is the one it
gets alone with one step per launch (int ring: bit-exact per-window spike
counts)."""
import time

import numpy as np
import pytest

from neuro_edit_v2_amd import Sim, SimConfig
from neuro_edit_v2_amd import workloads as W
from neuro_edit_v2_amd.serve import Server, window_counts

pytestmark = pytest.mark.gpu

WIN, NWIN = 40, 5
TOTAL = WIN * NWIN


def request(seed, edge, ring=True):
    net, series, _ = W.lsm_request(seed, dt=0.5, windows=NWIN, win_ms=WIN // 2, edge=edge)
    cfg = SimConfig(dt_ms=0.5, synapse_tau_ms=5.0, noise_rate_hz=10.0, noise_sd=0.5, seed=seed,
                    int_weights=ring, int_scale=65536.0, raster_steps=TOTAL)
    return net, cfg, series


def alone(net, cfg, series):
    cfg.step_shape = 1                    # one step per launch: the other shape
    s = Sim(net, cfg)
    for q in range(series.shape[1]):
        s.play_input(q, series[:, q])
    s.run_steps(TOTAL)
    _, c, ids = s.read_spikes()
    return window_counts(c, ids, int(net["n"]), WIN, NWIN)


def test_served_results_equal_alone_in_the_mix():
    srv = Server(cap_neurons=64 * 1024, cap_synapses=2_000_000, max_requests=8, n_windows=NWIN,
                 chunk_steps=16, int_weights=1, int_scale=65536.0)
    reqs = [request(s, e) for s, e in [(1, 8), (2, 11), (3, 6), (4, 13), (5, 9), (6, 7), (7, 12)]]
    ids = []
    for k, (net, cfg, ser) in enumerate(reqs):
        ids.append(srv.submit(net, cfg, ser, TOTAL, WIN))
        time.sleep(0.002 * (k % 3))       # arrivals while others are in flight
    got = [srv.wait(r) for r in ids]
    st = srv.stats()
    assert st["completed"] == len(reqs)
    assert st["peak_active"] >= 2, "requests never overlapped on the device"
    assert st["paired_rounds"] > 0, "the served arena never ran two steps per launch"
    t0s = {g[1]["t0"] for g in got}
    assert len(t0s) > 1, "every request started at the same step"
    for (net, cfg, ser), (counts, info) in zip(reqs, got):
        want = alone(net, cfg, ser)
        assert np.array_equal(counts, want)
        assert info["spikes"] == int(want.sum())
    assert sum(int(c.sum()) for c, _ in got) > 0
    srv.close()


def test_admission_waits_for_capacity():
    # room for about two of these requests at a time
    srv = Server(cap_neurons=2 * 1024 + 64, cap_synapses=200_000, max_requests=4, n_windows=NWIN,
                 chunk_steps=16, int_weights=1, int_scale=65536.0)
    reqs = [request(10 + k, 10) for k in range(5)]     # 1000 neurons each
    ids = [srv.submit(net, cfg, ser, TOTAL, WIN) for net, cfg, ser in reqs]
    got = [srv.wait(r) for r in ids]
    st = srv.stats()
    assert st["completed"] == 5 and st["peak_active"] <= 2
    admits = sorted(g[1]["t0"] for g in got)
    assert admits[-1] > admits[0]
    for (net, cfg, ser), (counts, _) in zip(reqs, got):
        assert np.array_equal(counts, alone(net, cfg, ser))
    srv.close()


def test_request_too_large_is_refused():
    srv = Server(cap_neurons=1024, cap_synapses=10_000, max_requests=2, n_windows=NWIN,
                 int_weights=1)
    net, cfg, ser = request(3, 12)
    from neuro_edit_v2_amd._lib import NHError
    with pytest.raises(NHError):
        srv.submit(net, cfg, ser, TOTAL, WIN)
    srv.close()


def test_served_float_ring_close_to_alone():
    """Float ring: the order of float atomic adds differs between the mix
    and a request alone, so the bar is a tolerance -- per-request spike
    totals within 2 % and per-neuron window counts within 2 % mean absolute
    deviation (relative to the mean count)."""
    srv = Server(cap_neurons=64 * 1024, cap_synapses=2_000_000, max_requests=8, n_windows=NWIN,
                 chunk_steps=16, int_weights=0)
    reqs = [request(s, e, ring=False) for s, e in [(21, 9), (22, 12), (23, 10), (24, 14)]]
    ids = [srv.submit(net, cfg, ser, TOTAL, WIN) for net, cfg, ser in reqs]
    got = [srv.wait(r) for r in ids]
    assert srv.stats()["peak_active"] >= 2
    for (net, cfg, ser), (counts, info) in zip(reqs, got):
        want = alone(net, cfg, ser).astype(np.float64)
        g = counts.astype(np.float64)
        assert want.sum() > 0
        assert abs(g.sum() - want.sum()) <= 0.02 * want.sum() + 1
        assert np.abs(g - want).mean() <= 0.02 * max(want.mean(), 1e-9) + 1e-3
    srv.close()


def oracle_windows(net, cfg, series, total, win, n_win, ring=True):
    """Per-window spike counts of one request from the oracle's raster."""
    from oracle import oracle as O
    ref = O.OracleSim(net, O.default_cfg(dt_ms=cfg.dt_ms, synapse_tau_ms=cfg.synapse_tau_ms,
                                         noise_rate_hz=cfg.noise_rate_hz, noise_kick_mv=cfg.noise_kick_mv,
                                         noise_sd=cfg.noise_sd, seed=cfg.seed, int_weights=int(ring),
                                         int_scale=cfg.int_scale, exp_mode=1))
    for q in range(series.shape[1]):
        ref.play_input(q, series[:, q])
    rc, ri = ref.run_steps(total, record=True)
    return window_counts(rc, ri, int(net["n"]), win, n_win)


def test_served_requests_match_oracle():
    """Window counts of served requests (admitted while others run) equal
    the counts derived from the oracle's raster of each request alone."""
    srv = Server(cap_neurons=64 * 1024, cap_synapses=2_000_000, max_requests=8, n_windows=NWIN,
                 chunk_steps=16, int_weights=1, int_scale=65536.0)
    reqs = [request(s, e) for s, e in [(31, 9), (32, 12), (33, 7)]]
    ids = [srv.submit(net, cfg, ser, TOTAL, WIN) for net, cfg, ser in reqs]
    got = [srv.wait(r) for r in ids]
    assert srv.stats()["peak_active"] >= 2
    for (net, cfg, ser), (counts, _) in zip(reqs, got):
        want = oracle_windows(net, cfg, ser, TOTAL, WIN, NWIN)
        assert want.sum() > 0
        assert np.array_equal(counts, want)
    srv.close()


@pytest.mark.parametrize("ring", [True, False], ids=["int", "float"])
def test_served_lsm_trials_full_size(ring):
    """The LSM trial requests at full size: grid liquids of edge 10 to 37
    (1,000 to 50,653 neurons), 8 bit-stream inputs with one of 4 stream
    pairs sharing bits in each of 10 windows of 50 ms, 1,000 steps, several
    requests in flight together.  Int ring: every request's window counts
    equal its run alone, and the largest and smallest equal the oracle's.
    Float ring: per-request totals within 2 % of the run alone and of the
    oracle, per-neuron window counts within 2 % mean absolute deviation."""
    steps, win, n_win = 1000, 100, 10
    edges = [(41, 10), (42, 37), (43, 20), (44, 28), (45, 15), (46, 33)]
    reqs = []
    for seed, e in edges:
        net, series, _ = W.lsm_request(seed, dt=0.5, edge=e)
        assert series.shape[0] == steps
        cfg = SimConfig(dt_ms=0.5, synapse_tau_ms=5.0, seed=seed, int_weights=ring,
                        int_scale=65536.0, raster_steps=steps)
        reqs.append((net, cfg, series))
    print(f"case: served LSM full size, ring int={ring}", flush=True)
    srv = Server(max_requests=16, max_delay=15, n_windows=n_win, chunk_steps=96,
                 int_weights=int(ring), int_scale=65536.0)
    ids = []
    for k, (net, cfg, ser) in enumerate(reqs):
        ids.append(srv.submit(net, cfg, ser, steps, win))
        time.sleep(0.001 * k)
    got = [srv.wait(r, timeout_ms=120000) for r in ids]
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["peak_active"] >= 2
    for k, ((net, cfg, ser), (counts, info)) in enumerate(zip(reqs, got)):
        cfg1 = SimConfig(dt_ms=0.5, synapse_tau_ms=5.0, seed=cfg.seed, int_weights=ring,
                         int_scale=65536.0, raster_steps=steps)
        s = Sim(net, cfg1)
        for q in range(8):
            s.play_input(q, ser[:, q])
        s.run_steps(steps)
        _, c, i = s.read_spikes()
        s.close()
        alone_w = window_counts(c, i, int(net["n"]), win, n_win)
        assert alone_w.sum() > 0
        if ring:
            assert np.array_equal(counts, alone_w), f"request {k}"
        else:
            g, a = counts.astype(np.float64), alone_w.astype(np.float64)
            assert abs(g.sum() - a.sum()) <= 0.02 * a.sum() + 1, f"request {k}"
            assert np.abs(g - a).mean() <= 0.02 * a.mean() + 1e-3, f"request {k}"
        if k in (0, 1):       # edges 10 and 37 against the oracle
            want = oracle_windows(net, cfg, ser, steps, win, n_win, ring)
            if ring:
                assert np.array_equal(counts, want), f"request {k} vs oracle"
            else:
                g, w = counts.astype(np.float64), want.astype(np.float64)
                print(f"request {k}: spikes served {g.sum():.0f} oracle {w.sum():.0f}, "
                      f"window-count MAD {np.abs(g - w).mean() / max(w.mean(), 1e-9):.4f}")
                assert abs(g.sum() - w.sum()) <= 0.02 * w.sum() + 1
                assert np.abs(g - w).mean() <= 0.02 * w.mean() + 1e-3
    srv.close()
