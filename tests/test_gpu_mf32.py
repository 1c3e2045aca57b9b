"""FP32 Legendre contraction on the matrix cores (impl 2: v_mfma_f32_16x16x4_f32,
recurrence on the VALU) against the VALU contraction (impl 1) and the f64
oracle.  Same FP32 arithmetic class, different summation order: the two
implementations agree to 2e-6 relative, and both meet the FP32 stage
tolerance of tests/test_gpu_sht.py against f64."""
import numpy as np
import pytest

import oracle as O
from conftest import emit
from test_gpu_sht import _random_field_spectrum

pytestmark = pytest.mark.gpu


def leg_tol(lmax):
    return max(2.5e-6, 2e-10 * (lmax + 1) ** 2)


_SMALL = [(15, 16, 32), (15, 17, 32), (12, 13, 26), (63, 64, 128), (127, 128, 256),
          (255, 256, 512)]
# every field count on the small grids; the two-field case (a step's shape) on the large ones
_CASES = [g + (nf,) for g in _SMALL for nf in (1, 2, 3, 5)] + [(511, 512, 1024, 2),
                                                               (1023, 1024, 2048, 2)]


@pytest.mark.parametrize("lmax,nlat,nphi,nf", _CASES)
def test_matrix_core_stages(gpu, lmax, nlat, nphi, nf):
    import tsphere

    cfg = tsphere.ShtConfig(lmax, lmax, nlat, nphi)
    emit(f"mf32 lmax{lmax} {nlat}x{nphi} nf {nf}")
    plan = tsphere.ShtPlan.create(cfg, max_fields=nf)
    ref = O.ShtReference(O.Cfg(lmax, lmax, nlat, nphi))
    qs = np.concatenate([_random_field_spectrum(O.Cfg(lmax, lmax, nlat, nphi), 40 + f)
                         for f in range(nf)])
    q = tsphere.DeviceBuffer(nf * 2 * plan.nlm)
    sp = tsphere.DeviceBuffer(nf * cfg.npts)
    back = tsphere.DeviceBuffer(nf * 2 * plan.nlm)
    q.upload(qs, plan.stream)

    def run(impl):
        for st in ("synth", "analys"):
            plan.set_tile(st, impl=impl)
        plan.encode_synth_into(plan.stream, plan.create_synth_binding(q, sp, nf, 2 * plan.nlm,
                                                                      cfg.npts))
        plan.encode_analys_into(plan.stream, plan.create_analys_binding(sp, back, nf, cfg.npts,
                                                                        2 * plan.nlm))
        return (sp.download(plan.stream).reshape(nf, -1),
                back.download(plan.stream).reshape(nf, -1))

    s1, b1 = run(1)
    s2, b2 = run(2)
    for f in range(nf):
        es = O.rel_l2(s2[f], s1[f])
        eb = O.rel_l2(b2[f], b1[f])
        qf = qs[f * 2 * plan.nlm:(f + 1) * 2 * plan.nlm]
        e64 = O.rel_l2(s2[f], ref.synth_fast(qf.astype(np.float64)))
        a64 = O.rel_l2(b2[f], ref.analys_fast(s2[f].astype(np.float64)))
        if f in (0, nf - 1):
            emit(f"  field {f}: vs VALU synth {es:.1e} analys {eb:.1e}; vs f64 synth "
                 f"{e64:.1e} analys {a64:.1e}")
        assert es < 2e-6 and eb < 2e-6
        assert e64 < leg_tol(lmax) and a64 < leg_tol(lmax)
    plan.destroy()


@pytest.mark.parametrize("nwaves,k", [(1, 4), (2, 2), (4, 1), (2, 4), (4, 4)])
def test_matrix_core_tiles(gpu, nwaves, k):
    """Every matrix-core workgroup shape gives the built-in shape's result
    up to summation order."""
    import tsphere

    cfg = tsphere.ShtConfig(255, 255, 256, 512)
    plan = tsphere.ShtPlan.create(cfg, max_fields=2)
    qs = np.concatenate([_random_field_spectrum(O.Cfg(255, 255, 256, 512), 7 + f)
                         for f in range(2)])
    emit(f"mf32 tile nwaves {nwaves} k {k}")
    q = tsphere.DeviceBuffer(2 * 2 * plan.nlm)
    sp = tsphere.DeviceBuffer(2 * cfg.npts)
    back = tsphere.DeviceBuffer(2 * 2 * plan.nlm)
    q.upload(qs, plan.stream)
    outs = []
    for tile in ({"impl": 2}, {"impl": 2, "nwaves": nwaves, "k": k}):
        for st in ("synth", "analys"):
            plan.set_tile(st, **tile)
        plan.encode_synth_into(plan.stream, plan.create_synth_binding(q, sp, 2, 2 * plan.nlm,
                                                                      cfg.npts))
        plan.encode_analys_into(plan.stream, plan.create_analys_binding(sp, back, 2, cfg.npts,
                                                                        2 * plan.nlm))
        outs.append((sp.download(plan.stream), back.download(plan.stream)))
    assert O.rel_l2(outs[1][0], outs[0][0]) < 2e-6 and O.rel_l2(outs[1][1], outs[0][1]) < 2e-6
    plan.destroy()
