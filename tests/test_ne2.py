"""The .ne2 reader (net-format/src/io.rs) against the reference's own
format fixtures (net-format/tests/fixtures/v1-v3.ne2, copied under
tests/golden/ne2/) and their assertions (net-format/tests/format_versions.rs:
20-78), and against examples/lsm.ne2 (format v6; the repository notes
record it as the 216-neuron / 855-synapse liquid, NOTES-for-data.md:3552)."""
import os

import numpy as np
import pytest

from neuro_edit_v2_amd import _lib as L
from neuro_edit_v2_amd import net as N

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ne2")


def load(name, dt_ms=0.1):
    return N.load_ne2(os.path.join(GOLD, name), dt_ms)


def effective_params(net, i):
    return net["p"][i * net["n_params"]: i * net["n_params"] + 4]


@pytest.mark.parametrize("name,version,tau0", [("v1.ne2", 1, 0.0), ("v2.ne2", 2, 0.0),
                                               ("v3.ne2", 3, 7.5)])
def test_format_fixture_lifts_into_typed_net(name, version, tau0):
    """format_versions.rs check_lifted: 4 neurons (2 stock RS exc, 1 stock
    FS inh, 1 custom exc), 2 synapses (0->1 w 1.5, 2->0 w 2.0), types
    exc / inh / izh#.... in first-use order, Dale-signed CSR [1.5, -2.0],
    group [0, 1, 2], input "i", one script."""
    net, info = load(name)
    assert info["version"] == version
    assert net["n"] == 4 and len(net["tgt"]) == 2
    assert info["type_names"][:2] == ["exc", "inh"] and info["type_names"][2].startswith("izh#")
    assert len(info["type_names"]) == 3
    assert list(info["neuron_types"]) == [0, 0, 1, 2]
    assert np.array_equal(effective_params(net, 0), np.float32([0.02, 0.2, -65.0, 8.0]))
    assert np.array_equal(effective_params(net, 2), np.float32([0.1, 0.2, -65.0, 2.0]))
    assert np.array_equal(effective_params(net, 3), np.float32([0.03, 0.25, -60.0, 4.0]))
    # CSR by source: 0->1 (+1.5), 2->0 (inhibitory source: -2.0)
    assert list(net["src_ptr"]) == [0, 1, 1, 2, 2]
    assert list(net["tgt"]) == [1, 0]
    assert np.array_equal(net["w"], np.float32([1.5, -2.0]))
    tau = net["tau"] if net["tau"] is not None else np.zeros(2, np.float32)
    assert tau[0] == np.float32(tau0)
    assert info["groups"][0][1].tolist() == [0, 1, 2]
    assert info["input_names"][0] == "i"
    assert info["n_scripts"] == 1
    # izh init rows (izh.rs:175-178)
    assert np.array_equal(net["x0"][:2], np.float32([-65.0, 0.2 * -65.0]))


def test_lift_hash_name_is_fnv1a_of_params_and_kind():
    """io.rs:118-128: the custom type's name is izh# + the low 16 bits of an
    FNV-1a 64 over the little-endian bytes of its four params, then its
    kind (0 = excitatory), in lower-case hex."""
    _, info = load("v1.ne2")
    h = 0xCBF29CE484222325
    for x in np.float32([0.03, 0.25, -60.0, 4.0]):
        for b in x.tobytes():
            h = ((h ^ b) * 0x100000001B3) & ((1 << 64) - 1)
    h = ((h ^ 0) * 0x100000001B3) & ((1 << 64) - 1)
    assert info["type_names"][2] == "izh#%04x" % (h & 0xFFFF)


def test_lsm_example_loads():
    """examples/lsm.ne2: format v6, 216 neurons, 855 synapses; every delay
    at least one step, Dale signs consistent per source, stimulus targets
    inside the network."""
    net, info = load("lsm.ne2", 0.0)
    assert info["version"] == 6
    assert net["n"] == 216 and len(net["tgt"]) == 855
    assert info["dt_ms"] > 0
    assert net["delay"].min() >= 1
    sp = net["src_ptr"].astype(np.int64)
    for i in range(net["n"]):
        w = net["w"][sp[i]:sp[i + 1]]
        assert np.all(w >= 0) or np.all(w <= 0)
    assert info["n_inputs"] == net["n_inputs"]
    if net["n_inputs"]:
        assert net["stim_input"].max() < net["n_inputs"]


def test_bad_files_are_refused():
    data = open(os.path.join(GOLD, "v3.ne2"), "rb").read()
    lib = L.lib()
    for bad, what in [(b"NE3\0" + data[4:], "magic"), (data[:4] + b"\x09\0\0\0" + data[8:], "version"),
                      (data[:60], "truncated")]:
        h = L.C.c_void_p()
        rc = lib.nh_net_load_ne2(bad, len(bad), 0.1, L.C.byref(h), None)
        assert rc == L.NH_ERR_NET, what
        assert not h.value
