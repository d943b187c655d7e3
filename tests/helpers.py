"""Shared fixtures-loading helpers for the tests."""
import os

import numpy as np

from bcdor_amd import formats, synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_golden(name):
    """Return (ModelSpace, Interaction, Propagator, npz) of a golden case; the
    interaction is regenerated from its seed and checked against the digest."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    ms, vpp = synth.make_system(int(z["nmax"]), int(z["seed"]))
    if "v_v" in z.files:
        assert np.array_equal(z["v_v"], vpp.v) and np.array_equal(z["v_a"], vpp.a), \
            "regenerated interaction differs from the stored one"
    import hashlib
    h = hashlib.sha256()
    for a in (vpp.a, vpp.b, vpp.c, vpp.d, vpp.J):
        h.update(np.ascontiguousarray(a, np.int32).tobytes())
    h.update(np.ascontiguousarray(vpp.v, np.float64).tobytes())
    assert h.hexdigest() == str(z["v_digest"]), "interaction generator drifted from the golden digest"
    g = formats.Propagator(z["prop_np"], z["prop_nh"], z["prop_ep"], (z["prop_fqp"] ** 2).sum(1), z["prop_fqp"],
                           z["prop_eh"], (z["prop_fqh"] ** 2).sum(1), z["prop_fqh"])
    return ms, vpp, g, z


def rel_err(a, b, floor=1.0):
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    if a.size == 0 and b.size == 0:
        return 0.0
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))
