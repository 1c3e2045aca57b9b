"""FP32 Legendre contraction on 4x4 matrix blocks (impl 3:
v_mfma_f32_4x4x1_16b_f32, the recurrence on the VALU) in the FP32 and the
FP64-recurrence forms, against the VALU contraction (impl 1) of the same form
and the f64 oracle.  Same arithmetic class, different summation order: the
two agree to 2e-6 relative, and impl 3 meets each form's tolerance against
f64 (FP32: the stage tolerance of tests/test_gpu_sht.py; FP64 recurrence:
1e-5 per transform, 2e-5 per round trip)."""
import numpy as np
import pytest

import oracle as O
from conftest import emit
from test_gpu_sht import _random_field_spectrum

pytestmark = pytest.mark.gpu


def leg_tol(lmax, legendre):
    return 1e-5 if legendre == "f64rec" else max(2.5e-6, 2e-10 * (lmax + 1) ** 2)


# lmax 12 and 101: lmax + 1 = 1 and 2 mod 4, rows whose staged alpha length
# is rounded up before the 16-byte aligned tiles that follow it in LDS
_SMALL = [(15, 16, 32), (15, 17, 32), (12, 13, 26), (63, 64, 128), (127, 128, 256),
          (255, 256, 512), (101, 102, 204)]
_CASES = ([g + (nf, "fp32") for g in _SMALL for nf in (1, 2, 3, 5)]
          + [g + (nf, "f64rec") for g in _SMALL[:4] for nf in (1, 2, 5)]
          + [(511, 512, 1024, 2, "fp32"), (1023, 1024, 2048, 2, "fp32"),
             (1023, 1024, 2048, 2, "f64rec")])


def _plan_and_fields(lmax, nlat, nphi, nf, legendre, seed):
    import tsphere

    cfg = tsphere.ShtConfig(lmax, lmax, nlat, nphi)
    plan = tsphere.ShtPlan.create(cfg, legendre=legendre, max_fields=nf)
    qs = np.concatenate([_random_field_spectrum(O.Cfg(lmax, lmax, nlat, nphi), seed + f)
                         for f in range(nf)])
    q = tsphere.DeviceBuffer(nf * 2 * plan.nlm)
    sp = tsphere.DeviceBuffer(nf * cfg.npts)
    back = tsphere.DeviceBuffer(nf * 2 * plan.nlm)
    q.upload(qs, plan.stream)

    def run(**tile):
        for st in ("synth", "analys"):
            plan.set_tile(st, **tile)
        plan.encode_synth_into(plan.stream, plan.create_synth_binding(q, sp, nf, 2 * plan.nlm,
                                                                      cfg.npts))
        plan.encode_analys_into(plan.stream, plan.create_analys_binding(sp, back, nf, cfg.npts,
                                                                        2 * plan.nlm))
        return (sp.download(plan.stream).reshape(nf, -1),
                back.download(plan.stream).reshape(nf, -1))

    return plan, qs, run


@pytest.mark.parametrize("lmax,nlat,nphi,nf,legendre", _CASES)
def test_block_matrix_stages(gpu, lmax, nlat, nphi, nf, legendre):
    emit(f"m4 {legendre} lmax{lmax} {nlat}x{nphi} nf {nf}")
    plan, qs, run = _plan_and_fields(lmax, nlat, nphi, nf, legendre, 40)
    ref = O.ShtReference(O.Cfg(lmax, lmax, nlat, nphi))
    s1, b1 = run(impl=1)
    s3, b3 = run(impl=3)
    for f in range(nf):
        es = O.rel_l2(s3[f], s1[f])
        eb = O.rel_l2(b3[f], b1[f])
        qf = qs[f * 2 * plan.nlm:(f + 1) * 2 * plan.nlm]
        e64 = O.rel_l2(s3[f], ref.synth_fast(qf.astype(np.float64)))
        a64 = O.rel_l2(b3[f], ref.analys_fast(s3[f].astype(np.float64)))
        if f in (0, nf - 1):
            emit(f"  field {f}: vs VALU synth {es:.1e} analys {eb:.1e}; vs f64 synth "
                 f"{e64:.1e} analys {a64:.1e}")
        assert es < 2e-6 and eb < 2e-6
        assert e64 < leg_tol(lmax, legendre) and a64 < leg_tol(lmax, legendre)
    plan.destroy()


@pytest.mark.parametrize("nwaves,k,rows", [(1, 4, 2), (2, 2, 1), (4, 1, 2), (2, 4, 1),
                                           (4, 4, 2), (4, 2, 1), (8, 1, 1), (8, 2, 2)])
@pytest.mark.parametrize("nfb", [1, 2, 4])
def test_block_matrix_tiles(gpu, nwaves, k, rows, nfb):
    """Every block-matrix workgroup shape gives the VALU result up to
    summation order (three fields, so a four-field workgroup has an idle
    field and a one-field workgroup idle columns)."""
    emit(f"m4 tile nwaves {nwaves} k {k} rows {rows} nf {nfb}")
    plan, qs, run = _plan_and_fields(255, 256, 512, 3, "fp32", 7)
    s1, b1 = run(impl=1)
    s3, b3 = run(impl=3, nwaves=nwaves, k=k, rows=rows, nf=nfb)
    assert O.rel_l2(s3, s1) < 2e-6 and O.rel_l2(b3, b1) < 2e-6
    plan.destroy()


@pytest.mark.parametrize("legendre", ["fp32", "f64rec"])
@pytest.mark.parametrize("nwaves,k,rows", [(2, 1, 1), (1, 2, 2), (4, 1, 2)])
def test_block_matrix_eight_fields(gpu, legendre, nwaves, k, rows):
    """Eight fields per workgroup (a batched-transform tile: the recurrence
    shared by four field pairs) gives the four-field result bit for bit: each
    field's sums run over the same l and pairs in the same order.  Eleven
    fields, so the second workgroup of eight has five idle fields; the
    weights of four field pairs go through the analysis tile in two passes."""
    emit(f"m4 eight fields {legendre} nwaves {nwaves} k {k} rows {rows}")
    plan, qs, run = _plan_and_fields(127, 128, 256, 11, legendre, 3)
    s4, b4 = run(impl=3, nwaves=nwaves, k=k, rows=rows, nf=4)
    s8, b8 = run(impl=3, nwaves=nwaves, k=k, rows=rows, nf=8)
    assert np.array_equal(s4, s8) and np.array_equal(b4, b8)
    s1, b1 = run(impl=1)
    assert O.rel_l2(s8, s1) < 2e-6 and O.rel_l2(b8, b1) < 2e-6
    plan.destroy()


def test_eight_field_tiles_refused_outside_block_kernels(gpu):
    """Eight fields per workgroup exist for the block-matrix kernels with one
    or two chains per lane only; other shapes are refused when set."""
    plan, _, _ = _plan_and_fields(127, 128, 256, 8, "fp32", 1)
    for bad in (dict(impl=1, nf=8, nwaves=2, k=1), dict(impl=3, nf=8, nwaves=1, k=4),
                dict(impl=2, nf=8, nwaves=2, k=1)):
        with pytest.raises(Exception, match="nf in"):
            plan.set_tile("analys", **bad)
    plan.destroy()
