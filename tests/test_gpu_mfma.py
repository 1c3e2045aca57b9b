"""The split-BF16 x3 MFMA Legendre stage vs the f64 oracle: the 4x4x4 block
kernels (the built-in choice up to 512 latitude pairs) and the 16x16x32
column-tile kernels.

Stated tolerance: 1e-5 relative spectral (or grid) L2 against FP64 for a
single transform, at every lmax up to 1023 -- the precision target of the
MFMA form -- and 2e-5 for a round trip (two transforms' errors add; the
split's residual is ~2^-18 per operand, measured 4-6e-6 per transform)."""
import numpy as np
import pytest

import oracle as O
from conftest import emit
from test_gpu_sht import _random_field_spectrum

pytestmark = pytest.mark.gpu
TOL = 1e-5
TOL_RT = 2e-5


def _plan(cfg, **kw):
    import tsphere

    return tsphere.ShtPlan.create(tsphere.ShtConfig(cfg.lmax, cfg.mmax, cfg.nlat, cfg.nphi),
                                  legendre="bf16x3", **kw)


# tile None: the built-in 4x4x4 block kernels; ctc: the 16x16x32 column-tile
# kernels (the form's choice above 512 latitude pairs, and a tuner candidate)
@pytest.mark.parametrize("tile", [None, 1])
@pytest.mark.parametrize("lmax,nlat,nphi", [(15, 16, 32), (15, 17, 32), (63, 64, 128),
                                            (127, 192, 256), (255, 256, 512)])
def test_stages_vs_oracle_bf16x3(gpu, lmax, nlat, nphi, tile):
    cfg = O.Cfg(lmax, lmax, nlat, nphi)
    emit(f"bf16x3 stages lmax{lmax} {nlat}x{nphi} tile {'blocks' if tile is None else 'ctc 1'}")
    plan = _plan(cfg)
    if tile is not None:
        plan.set_tile("synth", ctc=tile)
        plan.set_tile("analys", ctc=tile)
    want = {"impl": 4} if tile is None else {"ctc": tile}
    for stage in ("synth", "analys"):
        got = plan.get_tile(stage, 1)
        assert all(got[k] == v for k, v in want.items()), (stage, got)
    ref = O.ShtReference(cfg)
    q = _random_field_spectrum(cfg, 11)
    s = plan.stream
    plan.qlm_in.upload(q, s)
    plan.encode_stage(s, "legSynth")
    fm = plan.read_fm()
    e1 = O.rel_l2(fm, ref.legendre_synth(q.astype(np.float64)))
    plan.encode_stage(s, "fourSynth")
    plan.encode_stage(s, "fourAnalys")
    gm = plan.read_fm()
    plan.encode_stage(s, "legAnalys")
    qo = plan.qlm_out.download(s)
    e2 = O.rel_l2(qo, ref.legendre_analys(gm.astype(np.float64) * (2 * np.pi / nphi)))
    rt = O.rel_l2(plan.analys(plan.synth(q)), q)
    emit(f"  legSynth {e1:.2e} legAnalys {e2:.2e} round trip {rt:.2e}")
    assert e1 < TOL and e2 < TOL and rt < TOL_RT
    plan.destroy()


def test_batched_many_fields_bf16x3(gpu):
    """40 fields in one launch (3 column tiles): each equals its own transform
    to within the split's rounding, and both meet the f64 tolerance."""
    import tsphere

    cfg = O.Cfg(63, 63, 64, 128)
    nf = 40
    plan = _plan(cfg, max_fields=nf)
    ref = O.ShtReference(cfg)
    qs = np.stack([_random_field_spectrum(cfg, 300 + f) for f in range(nf)])
    s = plan.stream
    qb = tsphere.DeviceBuffer(nf * 2 * cfg.nlm)
    sb = tsphere.DeviceBuffer(nf * cfg.npts)
    qb.upload(qs, s)
    plan.encode_synth_into(s, plan.create_synth_binding(qb, sb, nf, 2 * cfg.nlm, cfg.npts))
    spat = sb.download(s).reshape(nf, -1)
    for f in (0, 17, 39):
        e = O.rel_l2(spat[f], ref.synth_fast(qs[f].astype(np.float64)))
        emit(f"  field {f}: synth vs f64 {e:.2e}")
        assert e < TOL
    plan.encode_analys_into(s, plan.create_analys_binding(sb, qb, nf, cfg.npts, 2 * cfg.nlm))
    back = qb.download(s).reshape(nf, -1)
    errs = [O.rel_l2(back[f], qs[f]) for f in range(nf)]
    ana = [O.rel_l2(back[f], ref.analys_fast(spat[f].astype(np.float64))) for f in range(nf)]
    emit("  round trip per field " + " ".join(f"{e:.1e}" for e in errs))
    emit("  analysis vs f64 per field " + " ".join(f"{e:.1e}" for e in ana))
    assert max(ana) < TOL and max(errs) < TOL_RT
    plan.destroy()


@pytest.mark.parametrize("lmax", [511, 767, 1023])
def test_precision_form_large(gpu, lmax):
    """lmax up to 1023 (nlat = lmax+1, nphi = 2(lmax+1)): synthesis vs the
    f64 oracle and round trip within 1e-5 -- where the plain FP32 recurrence
    is ~1e-4 off (lmax 767: 384 pairs in 6-wave block workgroups)."""
    cfg = O.Cfg(lmax, lmax, lmax + 1, 2 * (lmax + 1))
    emit(f"bf16x3 precision lmax{lmax}")
    plan = _plan(cfg, max_fields=2)
    q = _random_field_spectrum(cfg, 5)
    spat = plan.synth(q)
    ref = O.ShtReference(cfg)
    e_s = O.rel_l2(spat, ref.synth_fast(q.astype(np.float64)))
    e_rt = O.rel_l2(plan.analys(spat), q)
    emit(f"  synth vs f64 {e_s:.2e} round trip {e_rt:.2e}")
    assert e_s < TOL and e_rt < TOL_RT
    plan.destroy()
