"""C-ABI checks that need no GPU: the library loads, exports every entry point
include/bcdor_amd.h declares, and refuses work without a device (no CPU
fallback)."""
import os
import re

import pytest

from bcdor_amd import capi, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    with open(os.path.join(ROOT, "include", "bcdor_amd.h")) as f:
        txt = f.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bcd_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(capi.EXPORTS)


def test_version_and_errors():
    L = capi.lib()
    assert L.bcd_version() == 1
    assert L.bcd_error_string(-1002)


@pytest.mark.skipif(capi.lib().bcd_device_count() > 0, reason="a GPU is visible")
def test_no_device_is_an_error_not_a_fallback():
    ms, vpp = synth.make_system(1, 0)
    with pytest.raises(capi.BcdError, match="-1002"):
        capi.Context(ms, vpp)
    with pytest.raises(capi.BcdError, match="-1002"):
        capi.Batch([(ms, vpp)], capi.solve_params())
