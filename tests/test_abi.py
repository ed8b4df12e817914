"""The C-ABI library loads, exports every entry point include/neuro_hip.h
declares, and fails loudly (no CPU fallback) when no GPU is present."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from neuro_edit_v2_amd import _lib as L
from neuro_edit_v2_amd import net as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            txt = open(os.path.join(inc, f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(nh_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(L.LIB_PATH)
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_config_default_matches_reference_simconfig():
    # SimConfig::default -- sim/src/state.rs:35-49
    c = L.Config()
    L.lib().nh_config_default(C.byref(c))
    assert (c.dt_ms, c.synapse_tau_ms, c.readout_tau_ms) == (np.float32(0.1), 2.0, 2.0)
    assert c.noise_rate_hz == 0.0 and c.noise_kick_mv == 20.0 and c.seed == 1
    assert c.record_spikes == 0 and c.spike_record_steps == 512 and c.int_weights == 0
    assert c.int_scale == 4096.0


def test_no_gpu_means_loud_failure():
    from conftest import gpu_present
    if gpu_present():
        pytest.skip("a GPU is present")
    from neuro_edit_v2_amd import Sim
    net = N.cube_net(100, mean_deg=5.0, seed=1)
    with pytest.raises(L.NHError) as e:
        Sim(net)
    assert e.value.code == L.NH_ERR_NO_DEVICE
