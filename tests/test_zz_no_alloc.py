"""tsphere_synth_batch_noalloc -- the synthesis whose Fourier-space
intermediate lives in the library's static device scratch, for callers that
cannot allocate -- against tsphere_synth_batch through the C ABI, bit for bit
(the same kernels on the same inputs): two calls in a row, the second on a
smaller input (fewer fields, and then a smaller plan on the same scratch),
and a request beyond the scratch's 4 Mi complex values refused with an error
before anything is launched."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from conftest import emit

pytestmark = pytest.mark.gpu


def _pair(plan, nfield, rng):
    """(no-alloc result, plan-scratch result) of one batched synthesis."""
    from tsphere import _lib
    from tsphere.device import DeviceBuffer

    L = _lib.load()
    cfg = plan.cfg
    s = plan.stream
    nq, ns = 2 * plan.nlm, cfg.nlat * cfg.nphi
    q = np.concatenate([O.random_spectrum(O.Cfg(cfg.lmax, cfg.mmax, cfg.nlat, cfg.nphi),
                                          int(rng.integers(1 << 30))) for _ in range(nfield)])
    qb = DeviceBuffer(nfield * nq, "q")
    qb.upload(q, s)
    a, b = DeviceBuffer(nfield * ns, "a"), DeviceBuffer(nfield * ns, "b")
    _lib.check(L.tsphere_synth_batch_noalloc(plan.handle, nfield, qb.ptr, nq, a.ptr, ns,
                                             s.handle), "synth_batch_noalloc")
    _lib.check(L.tsphere_synth_batch(plan.handle, nfield, qb.ptr, nq, b.ptr, ns, s.handle),
               "synth_batch")
    out = a.download(s), b.download(s)
    for buf in (qb, a, b):
        buf.free()
    return out


def test_noalloc_synthesis_matches_plan_scratch(gpu):
    import tsphere
    from tsphere import _lib

    rng = np.random.default_rng(5)
    plan = tsphere.ShtPlan.create(tsphere.ShtConfig(255, 255, 256, 512), max_fields=4)
    for nfield in (4, 1):  # the second call on a smaller input
        emit(f"no-alloc synthesis lmax 255 256x512, {nfield} field(s)")
        got, want = _pair(plan, nfield, rng)
        assert np.all(np.isfinite(got)) and np.array_equal(got, want), nfield
    plan.destroy()
    small = tsphere.ShtPlan.create(tsphere.ShtConfig(31, 31, 32, 64), max_fields=2)
    emit("no-alloc synthesis lmax 31 32x64, 2 fields (a smaller plan on the same scratch)")
    got, want = _pair(small, 2, rng)
    assert np.all(np.isfinite(got)) and np.array_equal(got, want)
    # beyond the scratch: 5 fields x 1024 m x 1024 latitudes > 4 Mi complex values
    big = tsphere.ShtPlan.create(tsphere.ShtConfig(1023, 1023, 1024, 2048), max_fields=1)
    L = _lib.load()
    rc = L.tsphere_synth_batch_noalloc(big.handle, 5, C.c_void_p(16), 2 * big.nlm,
                                       C.c_void_p(16), 1024 * 2048, small.stream.handle)
    emit(f"  oversize request: rc {rc}, {_lib.last_error()}")
    assert rc == _lib.ERR_INVALID and "more than its" in _lib.last_error()
    big.destroy()
    small.destroy()
