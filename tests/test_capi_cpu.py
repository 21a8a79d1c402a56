"""The C ABI library loads and exports every declared entry (CPU only)."""
import ctypes as C
import os
import subprocess

import pytest

import hora_amd
from hora_amd._lib import HmParams


def test_library_exports_every_declared_symbol():
    L = hora_amd.lib()
    declared = hora_amd.declared_symbols()
    assert len(declared) >= 15
    out = subprocess.run(["nm", "-D", "--defined-only", hora_amd._lib.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for name in declared:
        assert name in exported, name
        assert hasattr(L, name)


def test_params_default_matches_reference_defaults():
    p = HmParams()
    assert hora_amd.lib().hm_params_default(2, C.byref(p)) == 0
    # IVFPQParams::default (pq_params.rs:80-91)
    assert (p.n_sub, p.sub_bits, p.n_kmeans_center, p.search_n_center, p.train_epoch) == \
        (25, 4, 256, 8, 100)
    # PQParams::default (pq_params.rs:31-40)
    assert hora_amd.lib().hm_params_default(3, C.byref(p)) == 0
    assert (p.n_sub, p.sub_bits, p.train_epoch) == (4, 4, 100)
    assert hora_amd.lib().hm_params_default(99, C.byref(p)) == 1
    assert b"unknown index kind" in hora_amd.lib().hm_last_error()


def test_null_arguments_are_statuses_not_crashes():
    L = hora_amd.lib()
    assert L.hm_index_create(1, 8, None, None) == 1
    assert L.hm_index_add(None, None, 0) == 1
    assert L.hm_index_destroy(None) == 0
    assert L.hm_index_name(None) == b""


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_gpu_fails_loudly():
    # No CPU fallback: without a device the product path reports an error.
    h = C.c_void_p()
    st = hora_amd.lib().hm_index_create(1, 8, None, C.byref(h))
    assert st == 7  # HM_ERR_HIP
    assert hora_amd.lib().hm_last_error()
