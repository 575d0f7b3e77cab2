"""The C ABI's failure-path hooks without a device: an armed entry point
fails once with the injected status and its text, after the requested
number of ordinary calls, and is disarmed afterwards; the allocation cap
refuses a device allocation above it as out of memory before HIP is
asked."""
import ctypes as C

import pytest


@pytest.fixture
def L():
    from wct_amd.capi import lib
    L = lib()
    yield L
    L.wct_debug_inject(None, 0, 0)
    L.wct_debug_alloc_cap(0)


def test_injection_fires_once_after_n_calls(L):
    assert L.wct_debug_inject(b"wct_decon_plan_run", 1, -4) == 0
    # the first call runs normally (a null plan is an argument error)
    assert L.wct_decon_plan_run(None, None) == -1
    assert L.wct_decon_plan_run(None, None) == -4
    assert "wct_decon_plan_run: injected failure (status -4)" in L.wct_last_error().decode()
    assert L.wct_decon_plan_run(None, None) == -1        # one-shot
    # other entry points are not affected
    assert L.wct_debug_inject(b"wct_decon_run", 0, -2) == 0
    assert L.wct_decon_plan_run(None, None) == -1
    assert L.wct_decon_run(None, 0, None, 0, None) == -2
    # disarming
    assert L.wct_debug_inject(b"wct_decon_run", 0, -3) == 0
    assert L.wct_debug_inject(None, 0, 0) == 0
    assert L.wct_decon_run(None, 0, None, 0, None) == -1


def test_injection_argument_checks(L):
    assert L.wct_debug_inject(b"wct_decon_run", -1, -3) == -1
    assert L.wct_debug_inject(b"wct_decon_run", 0, 5) == -1
    assert L.wct_debug_inject(b"wct_decon_run", 0, 0) == 0     # status 0 disarms


def test_allocation_cap(L):
    p = C.c_void_p()
    assert L.wct_debug_alloc_cap(4096) == 0
    assert L.wct_malloc(C.byref(p), 1 << 20) == -4
    assert "above the 4096-byte cap" in L.wct_last_error().decode()
    assert not p.value
    assert L.wct_debug_inject(b"wct_malloc", 0, -2) == 0
    assert L.wct_malloc(C.byref(p), 16) == -2
