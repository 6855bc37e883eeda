"""The C-ABI library on the host: it loads, exports every entry point the header declares, and its
host-only calls (version, tuning table, argument checks) behave without a GPU."""

import ctypes
import os
import re

import pytest

from pynear_amd import _native as N

HEADER = os.path.join(os.path.dirname(__file__), "..", "include", "pynear_amd.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pn_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = _declared()
    for must in ("pn_exact_search", "pn_exact_search_device", "pn_ivfbin_search", "pn_ivfl2_search",
                 "pn_kmeans_l2", "pn_pairwise_distance", "pn_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(N.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_cover_every_declared_symbol():
    assert sorted(N.SIGNATURES) == _declared()


def test_version_and_tuning_table():
    assert N.lib().pn_version().decode().count(".") == 2
    old = N.tune_get("l2.wg_per_cu")
    N.tune_set("l2.wg_per_cu", old + 1)
    assert N.tune_get("l2.wg_per_cu") == old + 1
    N.tune_set("l2.wg_per_cu", old)
    with pytest.raises(ValueError):
        N.tune_get("no.such.knob")


def test_null_arguments_are_rejected_with_value_error():
    with pytest.raises(ValueError):
        N.check(N.lib().pn_exact_create(N.METRIC_L2, None))
    with pytest.raises(ValueError):
        N.check(N.lib().pn_exact_create(99, ctypes.byref(ctypes.c_void_p())))
    assert "metric" in N.lib().pn_last_error().decode() or N.lib().pn_last_error()


def test_status_mapping():
    with pytest.raises(RuntimeError):
        N.check(N.PN_ERUNTIME)
    with pytest.raises(RuntimeError):
        N.check(N.PN_EDEVICE)
    N.check(N.PN_OK)
