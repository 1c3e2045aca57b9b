"""CPU checks of the C-ABI library: it loads, exports every symbol the header
declares, validates like the reference, and builds the same tables as the
oracle (no compute on a device)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tsphere.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tsphere_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import tsphere._lib as L

    lib = L.load()
    names = _declared()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers the whole header
    assert set(names) <= set(L.SIGNATURES), set(names) - set(L.SIGNATURES)


def test_version_and_errors():
    import tsphere._lib as L

    lib = L.load()
    assert lib.tsphere_version() >= 100
    assert lib.tsphere_plan_query(None, None) == L.ERR_INVALID
    assert "NULL" in L.last_error()


@pytest.mark.parametrize("cfg,msg", [
    ((0, 0, 4, 8), "lmax must be an integer >= 1"),
    ((8, 9, 16, 32), "mmax must be an integer in [0, lmax]"),
    ((8, 8, 8, 32), "nlat must be an integer > lmax"),
    ((8, 8, 16, 16), "nphi must be an integer >= 2*mmax+1"),
])
def test_plan_create_validation(cfg, msg):
    """layout.ts:34-43 messages, before any device work."""
    import tsphere._lib as L

    c = L.ShtConfigC(*cfg, 0, 0, 1)
    h = C.c_void_p()
    rc = L.load().tsphere_plan_create(C.byref(c), C.byref(h))
    assert rc == L.ERR_INVALID
    assert msg in L.last_error()
    assert h.value is None


def test_python_validation_matches():
    import tsphere

    with pytest.raises(ValueError, match=r"nphi must be an integer >= 2\*mmax\+1"):
        tsphere.validate_config(tsphere.ShtConfig(8, 8, 16, 16))
    for lmax, pdeg in [(31, 1), (31, 3), (63, 3), (255, 2), (1023, 1)]:
        assert tsphere.grid_for_lmax(lmax, pdeg) == O.grid_for_lmax(lmax, pdeg)
    for lmax, mmax in [(1, 0), (63, 63), (1023, 1023), (100, 37)]:
        assert tsphere.nlm_calc(lmax, mmax) == O.nlm_calc(lmax, mmax)
        for (l, m) in [(mmax, mmax), (lmax, 0), (lmax, mmax)]:
            assert tsphere.lm_index(lmax, l, m) == O.lm_index(lmax, l, m)


@pytest.mark.parametrize("lmax,nlat,nphi", [(31, 32, 64), (63, 64, 128), (127, 130, 256)])
def test_tables_match_oracle(lmax, nlat, nphi):
    """The library's f32 tables are the oracle's f64 values rounded once."""
    import tsphere

    cfg = tsphere.ShtConfig(lmax, lmax, nlat, nphi)
    t = tsphere.plan_tables(cfg)
    amm, ab = O.legendre_coeffs(lmax, lmax)
    assert np.array_equal(t["ab"], ab.astype(np.float32))
    assert np.array_equal(t["amm"], amm.astype(np.float32))
    x, w = O.gauss_nodes_weights(nlat)
    ct = t["ctstw"][:nlat]
    st = t["ctstw"][nlat:2 * nlat]
    ww = t["ctstw"][2 * nlat:]
    assert np.array_equal(ct, x.astype(np.float32))
    assert np.array_equal(st, np.sqrt(1 - x * x).astype(np.float32))
    assert np.array_equal(ww, (w * (2 * np.pi / nphi)).astype(np.float32))
    k = np.arange(nphi)
    trig = np.stack([np.cos(2 * np.pi * k / nphi), np.sin(2 * np.pi * k / nphi)], 1)
    assert np.array_equal(t["trig"], trig.astype(np.float32).reshape(-1))


def test_seed_table_is_scaled_sin_power():
    """seed_y * 2^(56 seed_ny) == amm[m] sin(theta_j)^m (f64 log domain), with
    ny < 0 exactly when the value is below the significance threshold."""
    import tsphere

    lmax, nlat = 1023, 1024
    cfg = tsphere.ShtConfig(lmax, lmax, nlat, 2048)
    t = tsphere.plan_tables(cfg)
    npair = nlat // 2
    x, _ = O.gauss_nodes_weights(nlat)
    st = np.sqrt(1 - x[:npair] ** 2)
    amm, _ = O.legendre_coeffs(lmax, lmax)
    y = t["seed_y"].reshape(lmax + 1, npair).astype(np.float64)
    ny = t["seed_ny"].reshape(lmax + 1, npair)
    m = np.arange(lmax + 1)[:, None]
    log2_true = np.log2(np.abs(amm))[:, None] + m * np.log2(st)[None, :]
    log2_got = np.log2(np.abs(y)) + 56.0 * ny
    assert np.max(np.abs(log2_got - log2_true)) < 1e-6
    assert np.all(np.sign(y) == np.sign(amm)[:, None])
    thr = 1e-15 * 2.0 ** 56 + 1
    assert np.all(ny <= 0)
    assert np.all(np.abs(y[ny < 0]) <= thr * (1 + 1e-6))
    # unscaled seeds agree with the direct power where f64 can hold it
    ok = ny == 0
    direct = (amm[:, None] * st[None, :] ** m)
    assert np.allclose(y[ok], direct[ok], rtol=1e-6, atol=0)


def test_rtc_compiles_a_kernel_without_a_device():
    """hipRTC targets gfx950 from the CPU container (model-compile time)."""
    import tsphere._lib as L

    src = b"""
extern "C" __global__ void k(long long n, float* __restrict__ out, const float* __restrict__ a,
                             float s) {
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = a[i] * s;
}
"""
    assert L.load().tsphere_rtc_check(src, None) == L.OK, L.last_error()
    rc = L.load().tsphere_rtc_check(b"this is not C++", None)
    assert rc == L.ERR_COMPILE
    assert "error" in L.last_error().lower()
