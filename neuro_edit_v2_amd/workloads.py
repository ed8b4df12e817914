"""Benchmark workloads built from the library's generators.

* :func:`batch_nets` -- 256 independent 3D-grid liquids, sizes log-uniform
  in [1,000, 20,000] neurons, half Izhikevich and half AdEx, each with its
  own seed and 8 seeded Bernoulli(0.05) bit-stream inputs at 1 ms bins.
* :func:`lsm_request` -- one LSM trial: a grid liquid of edge U{10..37}
  driven by 8 bit streams; in each of 10 windows of 50 ms one of the 4
  stream pairs shares its bits (sim/src/script.rs:2849-2894 semantics,
  strength 1).
"""
import ctypes as C
import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _lib as L
from . import net as N

PAIRS = [(0, 1), (2, 3), (4, 5), (6, 7)]


def rng_stream(seed, n_src, p):
    h = C.c_void_p()
    L.check(L.lib().nh_rng_stream_new(seed, n_src, p, C.byref(h)))
    return h


def stream_window(h, n_src, steps, pairs=None, strength=0.0):
    out = np.zeros(steps * n_src, np.float32)
    pr = np.ascontiguousarray([x for p in (pairs or []) for x in p] or [0], np.uint32)
    L.check(L.lib().nh_rng_stream_window(h, steps, int(bool(pairs)), L.ptr(pr),
                                         len(pairs or []), strength, L.ptr(out)))
    return out.reshape(steps, n_src)


def hold(bits, steps_per_bin):
    """1 ms bins held for 1/dt steps (examples/lsm_deep.rhai:98-110)."""
    return np.repeat(bits, steps_per_bin, axis=0)


def grid_dims(n):
    e = max(1, int(round(n ** (1.0 / 3.0))))
    ez = max(1, int(round(n / (e * e))))
    return e, e, ez


def batch_sizes(n_nets=256, lo=1000, hi=20000, seed=2026):
    r = np.random.default_rng(seed)
    return np.rint(np.exp(r.uniform(math.log(lo), math.log(hi), n_nets))).astype(int)


def _one(k, n, dt, steps):
    seed = k + 1
    model = 0 if k % 2 == 0 else 1
    net = N.grid_net(0, model=model, seed=seed, dt_ms=dt, dims=grid_dims(n))
    h = rng_stream(seed * 7919, 8, 0.05)
    try:
        bins = stream_window(h, 8, int(round(steps * dt)))
    finally:
        L.lib().nh_rng_stream_free(h)
    series = hold(bins, int(round(1.0 / dt)))[:steps]
    return net, series


def batch_nets(n_nets=256, dt=0.5, steps=1000, workers=8):
    """(nets, series) -- series[k] is a [steps, 8] input matrix of net k."""
    sizes = batch_sizes(n_nets)
    with ThreadPoolExecutor(workers) as ex:
        out = list(ex.map(lambda a: _one(a[0], int(a[1]), dt, steps), enumerate(sizes)))
    return [o[0] for o in out], [o[1] for o in out]


def lsm_request(seed, dt=0.5, windows=10, win_ms=50, edge=None):
    """(net, series, labels): one LSM trial request."""
    r = np.random.default_rng(seed)
    e = int(edge) if edge is not None else int(r.integers(10, 38))
    model = 0
    net = N.grid_net(e, model=model, seed=seed, dt_ms=dt)
    h = rng_stream(seed * 104729 + 1, 8, 0.05)
    spb = int(round(1.0 / dt))
    rows, labels = [], []
    try:
        for w in range(windows):
            lab = int(r.integers(0, 4))
            labels.append(lab)
            rows.append(stream_window(h, 8, win_ms, [PAIRS[lab]], 1.0))
    finally:
        L.lib().nh_rng_stream_free(h)
    series = hold(np.concatenate(rows, 0), spb)
    return net, series, labels
