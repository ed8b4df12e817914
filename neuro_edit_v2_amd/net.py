"""Network generators (C++ in libneurohip.so) returning the flat simulator
layout as a dict of numpy arrays -- the keys of ``nh_net_desc``:

n, src_ptr, tgt, w, delay, (tau), model_ids, model, n_params, n_state, p,
x0, n_inputs, stim_ptr, stim_input, stim_w.

* :func:`bench_net` -- sim/src/bench.rs:10-53 (fixed out-degree, constant
  weights) at a given dt;
* :func:`cube_net`  -- the random-cube workload (Poisson out-degree,
  U[0,1)*w_scale weights, distance delays 1-20 steps);
* :func:`grid_net`  -- 3D-grid liquid with local wiring and bit-stream
  inputs (net-format/src/builders.rs semantics, bucketed pair loop);
* :func:`load_ne2`  -- a .ne2 file (net-format/src/io.rs), any format
  version 1-7.
"""
import ctypes as C

import numpy as np

from . import _lib as L


def _take(h, extra=None):
    """Copy an nh_net into numpy arrays and free it (``extra(h, out)`` reads
    more from it first)."""
    lib = L.lib()
    d = L.NetDesc()
    try:
        L.check(lib.nh_net_view(h, C.byref(d)))
        n, E = d.n, d.n_syn

        def arr(p, dt, k):
            if not p or k == 0:
                return np.zeros(0, dt)
            buf = (C.c_char * (k * np.dtype(dt).itemsize)).from_address(p)
            return np.frombuffer(buf, dt, k).copy()

        out = dict(
            n=n, src_ptr=arr(d.src_ptr, np.uint32, n + 1), tgt=arr(d.tgt, np.uint32, E),
            w=arr(d.w, np.float32, E), delay=arr(d.delay, np.uint16, E),
            tau=arr(d.tau, np.float32, E) if d.tau else None,
            model_ids=arr(d.model_ids, np.uint8, d.n_models), model=arr(d.model, np.uint8, n),
            n_params=d.n_params, n_state=d.n_state,
            p=arr(d.p, np.float32, n * d.n_params), x0=arr(d.x0, np.float32, n * d.n_state),
            n_inputs=d.n_inputs,
        )
        if d.n_inputs and d.stim_ptr:
            sp = arr(d.stim_ptr, np.uint32, n + 1)
            ns = int(sp[-1])
            out.update(stim_ptr=sp, stim_input=arr(d.stim_input, np.uint32, ns),
                       stim_w=arr(d.stim_w, np.float32, ns))
        else:
            out.update(stim_ptr=np.zeros(n + 1, np.uint32), stim_input=np.zeros(0, np.uint32),
                       stim_w=np.zeros(0, np.float32))
        if extra is not None:
            extra(h, out)
        return out
    finally:
        lib.nh_net_free(h)


def bench_net(n, k=100, seed=1, w_scale=1.0, w_inh=5.0, dt_ms=0.1):
    h = C.c_void_p()
    L.check(L.lib().nh_net_bench(n, k, seed, w_scale, w_inh, dt_ms, C.byref(h)))
    return _take(h)


def cube_net(n, mean_deg=100.0, seed=1, w_scale=0.7):
    h = C.c_void_p()
    L.check(L.lib().nh_net_cube(n, mean_deg, seed, w_scale, C.byref(h)))
    return _take(h)


def grid_net(edge, scale=0.3, lam=2.0, inh_frac=0.2, w_exc=3.0, w_inh=3.0, model=0,
             n_inputs=8, per_input=24, w_in=12.0, seed=1, dt_ms=0.5, dims=None):
    ex, ey, ez = dims if dims is not None else (edge, edge, edge)
    h = C.c_void_p()
    L.check(L.lib().nh_net_grid(ex, ey, ez, scale, lam, inh_frac, w_exc, w_inh, model, n_inputs,
                                per_input, w_in, seed, dt_ms, C.byref(h)))
    return _take(h)


def out_degree(net):
    return np.diff(net["src_ptr"].astype(np.int64))


def load_ne2(path, dt_ms=0.0):
    """Read a .ne2 file: (net dict, info dict).  ``dt_ms`` <= 0 takes the
    file's meta.dt_ms for the delay steps.  ``info`` holds the file's meta,
    table sizes, type names, each neuron's type, input names and groups."""
    data = open(path, "rb").read()
    lib = L.lib()
    h = C.c_void_p()
    inf = L.Ne2Info()
    L.check(lib.nh_net_load_ne2(data, len(data), float(dt_ms), C.byref(h), C.byref(inf)))
    info = {k: getattr(inf, k) for k, _ in L.Ne2Info._fields_}
    buf = C.create_string_buffer(4096)

    def extra(hh, out):
        info["type_names"] = []
        for k in range(inf.n_types):
            L.check(lib.nh_net_type_name(hh, k, buf, 4096))
            info["type_names"].append(buf.value.decode())
        ty = np.zeros(max(out["n"], 1), np.uint16)
        L.check(lib.nh_net_neuron_types(hh, L.ptr(ty), out["n"]))
        info["neuron_types"] = ty[: out["n"]]
        info["input_names"] = []
        for k in range(inf.n_inputs):
            L.check(lib.nh_net_input_name(hh, k, buf, 4096))
            info["input_names"].append(buf.value.decode())
        info["groups"] = []
        for k in range(inf.n_groups):
            m = C.c_uint32()
            L.check(lib.nh_net_group(hh, k, buf, 4096, None, 0, C.byref(m)))
            mem = np.zeros(max(m.value, 1), np.uint32)
            L.check(lib.nh_net_group(hh, k, buf, 4096, L.ptr(mem), m.value, C.byref(m)))
            info["groups"].append((buf.value.decode(), mem[: m.value]))

    return _take(h, extra), info
