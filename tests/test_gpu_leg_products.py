"""Legendre contraction as per-parity matrix products (tile impl 5:
v_mfma_f32_16x16x4_f32 with 16-column blocks of (field, re/im) of 8 fields,
8 or 16 fields per workgroup, the recurrence on the VALU) in the FP32 and the
FP64-recurrence forms, against the VALU contraction (impl 1) of the same
form and the f64 oracle.  Same arithmetic class, different summation order:
the two agree to 2e-6 relative, and impl 5 meets each form's tolerance
against f64 (FP32: the stage tolerance of tests/test_gpu_sht.py; FP64
recurrence: 1e-5 per transform)."""
import numpy as np
import pytest

import oracle as O
from conftest import emit
from test_gpu_m4 import _plan_and_fields, leg_tol

pytestmark = pytest.mark.gpu

# grids: latitude pairs 8 (one wave, idle lanes), an odd grid (the equator
# pairs with itself), lmax + 1 = 1 and 2 mod 4, one to eight waves of pairs
_GRIDS = [(15, 16, 32), (15, 17, 32), (12, 13, 26), (63, 64, 128), (101, 102, 204),
          (127, 128, 256), (255, 256, 512)]
# field counts: a partial group of 8 and of 16, whole groups, one field
_CASES = ([g + (nf, tnf, "fp32") for g in _GRIDS for nf, tnf in ((1, 8), (11, 8), (19, 16))]
          + [g + (16, 16, "fp32") for g in _GRIDS[3:5]]
          + [g + (nf, tnf, "f64rec") for g in _GRIDS[1:4] for nf, tnf in ((11, 8), (19, 16))]
          + [(511, 512, 1024, 16, 16, "fp32"), (1023, 1024, 2048, 5, 8, "fp32"),
             (1023, 1024, 2048, 5, 8, "f64rec")])


@pytest.mark.parametrize("lmax,nlat,nphi,nf,tnf,legendre", _CASES)
def test_parity_product_stages(gpu, lmax, nlat, nphi, nf, tnf, legendre):
    emit(f"products {legendre} lmax{lmax} {nlat}x{nphi} fields {nf} tile nf {tnf}")
    plan, qs, run = _plan_and_fields(lmax, nlat, nphi, nf, legendre, 50)
    ref = O.ShtReference(O.Cfg(lmax, lmax, nlat, nphi))
    s1, b1 = run(impl=1)
    s5, b5 = run(impl=5, nf=tnf)
    for f in range(nf):
        es = O.rel_l2(s5[f], s1[f])
        eb = O.rel_l2(b5[f], b1[f])
        qf = qs[f * 2 * plan.nlm:(f + 1) * 2 * plan.nlm]
        e64 = O.rel_l2(s5[f], ref.synth_fast(qf.astype(np.float64)))
        a64 = O.rel_l2(b5[f], ref.analys_fast(s5[f].astype(np.float64)))
        if f in (0, nf - 1):
            emit(f"  field {f}: vs VALU synth {es:.1e} analys {eb:.1e}; vs f64 synth "
                 f"{e64:.1e} analys {a64:.1e}")
        assert es < 2e-6 and eb < 2e-6
        assert e64 < leg_tol(lmax, legendre) and a64 < leg_tol(lmax, legendre)
    plan.destroy()


def test_parity_product_tile_shapes(gpu):
    """The built-in shape of impl 5 (16 fields from 16 fields up, else 8; a
    wave per 64 latitude pairs) and the shapes it refuses when set."""
    plan, _, _ = _plan_and_fields(127, 128, 256, 16, "fp32", 1)
    plan.set_tile("synth", impl=5)
    assert plan.get_tile("synth", 16) == {"nf": 16, "nwaves": 1, "k": 1, "ctc": 0, "impl": 5,
                                          "rows": 1}
    assert plan.get_tile("synth", 9)["nf"] == 8
    plan.set_tile("analys", impl=5, nwaves=3)  # waves past the pairs are idle, not refused
    for bad in (dict(impl=5, nf=4), dict(impl=5, nf=16, k=2), dict(impl=5, rows=2),
                dict(impl=4)):
        with pytest.raises(Exception, match="tile"):
            plan.set_tile("analys", **bad)
    plan.destroy()


def test_parity_product_waves_past_the_pairs(gpu):
    """More waves than latitude-pair chunks: the extra waves carry no pair
    and every output equals the one-wave launch's bit for bit (each output's
    terms are summed by one wave in the same order; the idle waves add
    exact zeros in the analysis)."""
    plan, _, run = _plan_and_fields(63, 64, 128, 9, "fp32", 5)
    s_a, b_a = run(impl=5, nf=8, nwaves=1)
    s_b, b_b = run(impl=5, nf=8, nwaves=3)
    assert np.array_equal(s_a, s_b) and np.array_equal(b_a, b_b)
    plan.destroy()


def test_parity_products_above_512_pairs_take_the_block_kernels(gpu):
    """impl 5 covers up to 512 latitude pairs (8 waves of 64); beyond, the
    tile resolves to the 4x4 block kernels."""
    import tsphere

    cfg = tsphere.ShtConfig(31, 31, 1026, 64)
    plan = tsphere.ShtPlan.create(cfg, max_fields=4)
    plan.set_tile("synth", impl=5)
    assert plan.get_tile("synth", 4)["impl"] == 3
    plan.destroy()


@pytest.mark.parametrize("lmax,nlat,nphi,nf", [(15, 17, 32, 11), (12, 13, 26, 9),
                                                (63, 64, 128, 8), (101, 102, 204, 19)])
def test_two_row_synthesis(gpu, lmax, nlat, nphi, nf):
    """The 8-field synthesis with two consecutive m-rows per workgroup (lane
    pairs write (re, im) of m and m + 1 as 16 contiguous bytes) equals the
    one-row form and the VALU form bit for bit: every row's sums run over l
    in the same order.  An odd row count (lmax 12) leaves the last workgroup
    one row; partial groups of 8 fields leave idle columns."""
    import tsphere
    from tsphere import _lib

    emit(f"two-row synthesis lmax{lmax} {nlat}x{nphi} fields {nf}")
    L = _lib.load()
    plan, qs, _ = _plan_and_fields(lmax, nlat, nphi, nf, "fp32", 70)
    s = plan.stream
    q = tsphere.DeviceBuffer(nf * 2 * plan.nlm)
    q.upload(qs, s)
    fm = tsphere.DeviceBuffer(nf * 2 * (lmax + 1) * nlat)
    outs = []
    for tile in (dict(impl=5, nf=8, rows=2), dict(impl=5, nf=8, rows=1), dict(impl=1)):
        plan.set_tile("synth", **tile)
        fm.fill(np.nan, s)
        _lib.check(L.tsphere_leg_synth(plan.handle, nf, q.ptr, 2 * plan.nlm, fm.ptr, s.handle))
        outs.append(fm.download(s))
    assert np.all(np.isfinite(outs[0]))
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    with pytest.raises(Exception, match="tile"):
        plan.set_tile("analys", impl=5, nf=8, rows=2)
    with pytest.raises(Exception, match="tile"):
        plan.set_tile("synth", impl=5, nf=16, rows=2)
    plan.destroy()
