"""sv_diagnostics_noalloc (the allocation-free variant of sv_diagnostics)
through the C ABI: its device output equals sv_diagnostics' on two calls in a
row, the second on a smaller context, for every kind, on the context's own
stream and on a caller's stream; on an ensemble context per member."""
import ctypes as C

import numpy as np
import pytest

from sunset_amd.nodes import jittered_lattice, taylor_green_ic

pytestmark = pytest.mark.gpu

NOUT = {0: 3, 1: 6, 2: 5, 3: 1}
PH = dict(Re=1.0, Wi=1.0, beta=0.5, eps=0.1, Ma=0.1, kol_amp=1.0, kol_k=2.0)


def _ctx(n, order, seed):
    import sunset_amd
    ns = jittered_lattice(n, seed=seed, h_over_s={4: 1.8, 6: 2.4}[order])
    s = sunset_amd.Solver(morder=order, form="lc").setup(ns)
    s.set_phys(**PH)
    s.set_initial(taylor_green_ic(ns))
    s.run(3)
    return s


def _hip():
    """The HIP runtime the library itself links (torch ships its own copy, and
    a second runtime in the process does not see the GPU once this one runs)."""
    import sunset_amd
    sunset_amd.lib()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _noalloc(h, kind, member, stream=None):
    """Call sv_diagnostics_noalloc into a caller-owned device buffer of 8
    doubles pre-filled with NaN bytes (0xFF), on the context's stream (None)
    or on a caller's stream; return the 8 values."""
    import sunset_amd
    hip = _hip()
    buf = C.c_void_p()
    assert hip.hipMalloc(C.byref(buf), 64) == 0
    try:
        assert hip.hipMemset(buf, 0xFF, 64) == 0
        assert hip.hipDeviceSynchronize() == 0
        rc = sunset_amd.lib().sv_diagnostics_noalloc(h, kind, member, buf, stream)
        assert rc == 0, sunset_amd.lib().sv_last_error(h).decode()
        if stream is None:
            assert sunset_amd.lib().sv_sync(h) == 0
        else:
            assert hip.hipStreamSynchronize(stream) == 0
        out = np.zeros(8)
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), buf, 64, 2) == 0   # D2H
    finally:
        hip.hipFree(buf)
    return out


def _check(s, kind, member=0, stream=None, ref=None):
    import sunset_amd
    got = _noalloc(s.h, kind, member, stream)
    n = NOUT[kind]
    if ref is None:
        ref = np.zeros(n)
        assert sunset_amd.lib().sv_diagnostics(s.h, kind, member,
                                               ref.ctypes.data_as(sunset_amd.core.PD)) == 0
    assert np.isfinite(got[:n]).all(), (kind, got)
    if kind == 3:
        # the Taylor-Green decay factor is computed on the device here
        assert np.allclose(got[:n], ref, rtol=1e-13, atol=0.0), (kind, got[:n], ref)
    else:
        assert np.array_equal(got[:n], ref), (kind, got[:n], ref)
    assert np.isnan(got[n:]).all()   # nothing written past the kind's values


def test_noalloc_diagnostics_equal_allocating_ones_then_smaller_context():
    big = _ctx(40, 6, 1)
    small = _ctx(13, 4, 2)     # fewer blocks than the first call left in the scratch
    print("[no-alloc] 1600-node context, then 169-node context", flush=True)
    for kind in (0, 1, 2, 3):
        _check(big, kind)
        _check(small, kind)
    hip = _hip()
    side = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(side)) == 0
    for kind in (0, 1, 2, 3):
        _check(small, kind, stream=side)
        _check(big, kind, stream=side)
    assert hip.hipStreamDestroy(side) == 0
    big.close()
    small.close()


def test_noalloc_diagnostics_on_ensemble_members():
    import sunset_amd
    mem = [_ctx(20, 6, 3), _ctx(16, 4, 4), _ctx(24, 6, 5)]
    E = sunset_amd.Ensemble(mem)
    E.run(2)
    print("[no-alloc] ensemble of 3 members", flush=True)
    for m in range(3):
        for kind in (0, 1):
            _check(E, kind, member=m, ref=E.diagnostics(m, {0: "mass_energy", 1: "conf"}[kind]))
    with pytest.raises(AssertionError):
        _check(E, 2, member=0)          # refused on an ensemble: position-based
    E.close()
    for s in mem:
        s.close()
