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


def test_significance_seeds_match_oracle_rows():
    """For every (m, latitude pair): the seed sits at the first l where the
    oracle's f64 ytilde_l^m exceeds 1e-15 (coeffs.ts:64-85, the reference's
    significance rule), rounded down to m + 4k, and holds the oracle's values
    there; pairs never significant up to lmax have no seed."""
    import tsphere

    for lmax, nlat in [(63, 64), (255, 256)]:
        cfg = tsphere.ShtConfig(lmax, lmax, nlat, 2 * nlat)
        t = tsphere.plan_tables(cfg)
        npair = nlat // 2
        x, _ = O.gauss_nodes_weights(nlat)
        amm, ab = O.legendre_coeffs(lmax, lmax)
        y0 = t["seed_y0"].reshape(lmax + 1, npair)
        y1 = t["seed_y1"].reshape(lmax + 1, npair)
        sl = t["seed_l"].reshape(lmax + 1, npair)
        step = 1 if lmax < 100 else 7
        for m in range(0, lmax + 1, step):
            for j in range(0, npair, 1 if lmax < 100 else 5):
                row = O.legendre_row(amm, ab, lmax, m, x[j], np.sqrt(1 - x[j] ** 2))
                sig = np.flatnonzero(np.abs(row) > 1e-15)
                if sig.size == 0:
                    assert sl[m, j] == 2 ** 31 - 1, (m, j)
                    continue
                first = m + sig[0]
                assert sl[m, j] == m + 4 * ((first - m) // 4), (m, j, first, sl[m, j])
                k = sl[m, j] - m
                assert np.isclose(y0[m, j], row[k], rtol=2e-7, atol=0), (m, j)
                if k + 1 <= lmax - m:
                    assert np.isclose(y1[m, j], row[k + 1], rtol=2e-7, atol=1e-30), (m, j)
        # the skipped volume: the part of the (l, m, latitude) triangle the
        # recurrence never has to walk (7% at lmax 63, 20% at 255, 28% at 1023)
        n_tot = sum((lmax - m + 1) * npair for m in range(lmax + 1))
        start = np.minimum(sl, lmax + 1) - np.arange(lmax + 1)[:, None]
        skipped = start.sum() / n_tot
        assert {63: 0.05, 255: 0.15}[lmax] < skipped < 0.4, skipped


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
