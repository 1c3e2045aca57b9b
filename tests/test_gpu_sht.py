"""GPU transforms vs the f64 oracle (the role of test/transformChecks.ts and
the stage-wise localisation of scripts/diagnose-sht.ts:163-255)."""
import numpy as np
import pytest

import oracle as O
from conftest import emit

pytestmark = pytest.mark.gpu


def _plan(cfg, **kw):
    import tsphere

    return tsphere.ShtPlan.create(tsphere.ShtConfig(cfg.lmax, cfg.mmax, cfg.nlat, cfg.nphi), **kw)


def leg_tol(lmax):
    """Stated tolerance of one plain-FP32 Legendre stage (and of a round trip)
    vs f64: 2.5e-6 up to lmax ~110, then growing like lmax^2, as the
    rounding of the f32 recurrence state does (measured 1.6e-5 for the
    analysis at lmax 1023).  At lmax 63 the FP32 form measures 0.6-2.3e-6
    per round trip depending on the spectrum, above the 2e-6 target in the
    worst cases: a CPU emulation of the kernel's recurrence attributes
    ~1.0e-6 to the f32 rounding of the coefficients alpha_l and ~2.3e-6 to
    the rounding of the f32 state, each of which would cost recurrence
    instructions to remove.  The FP64-recurrence form meets the target
    (2e-7, tests/test_gpu_f64rec.py)."""
    return max(2.5e-6, 2e-10 * (lmax + 1) ** 2)


def _random_field_spectrum(cfg, seed):
    """sigma_l = 1/(l+1) spectrum (the SHT config's law), m=0 real."""
    rng = np.random.default_rng(seed)
    q = np.zeros((cfg.nlm, 2))
    for m in range(cfg.mmax + 1):
        for l in range(m, cfg.lmax + 1):
            i = O.lm_index(cfg.lmax, l, m)
            s = 1.0 / (l + 1)
            if m == 0:
                q[i, 0] = rng.normal(0, s)
            else:
                q[i] = rng.normal(0, s / np.sqrt(2), 2)
    return q.reshape(-1).astype(np.float32)


# transformChecks.ts:30-52: lmax 31 on gridForLmax(31,1) = 32x64, seed 42, < 1e-4
def test_transform_check_lmax31(gpu):
    nlat, nphi = O.grid_for_lmax(31, 1)
    cfg = O.Cfg(31, 31, nlat, nphi)
    emit(f"transformChecks lmax31 {nlat}x{nphi}")
    plan = _plan(cfg)
    ref = O.ShtReference(cfg)
    q = O.random_spectrum(cfg, 42)
    spat_gpu = plan.synth(q)
    spat_cpu = ref.synth(q.astype(np.float64))
    e_s = O.rel_l2(spat_gpu, spat_cpu)
    q_gpu = plan.analys(spat_cpu.astype(np.float32))
    q_cpu = ref.analys(spat_cpu)
    e_a = O.rel_l2(q_gpu, q_cpu)
    emit(f"  synth {e_s:.2e} analys {e_a:.2e}")
    assert e_s < 1e-4 and e_a < 1e-4
    # tighter than the reference's gate: what fp32 should actually give
    assert e_s < 2e-6 and e_a < 2e-6
    plan.destroy()


@pytest.mark.parametrize("lmax,nlat,nphi,fourier", [
    (15, 16, 32, "auto"),
    (15, 17, 31, "auto"),      # odd nlat (equator pairs with itself), non-pow2 nphi -> DFT
    (31, 48, 64, "dft"),       # explicit DFT
    (63, 64, 128, "auto"),     # the plain-method grid
    (63, 128, 256, "auto"),    # gridForLmax(63, 3)
    (127, 128, 256, "auto"),
])
def test_stages_vs_oracle(gpu, lmax, nlat, nphi, fourier):
    """Each half of each transform against the oracle's half (diagnose-sht.ts)."""
    import tsphere

    cfg = O.Cfg(lmax, lmax, nlat, nphi)
    emit(f"stages lmax{lmax} {nlat}x{nphi} {fourier}")
    plan = _plan(cfg, fourier=fourier)
    ref = O.ShtReference(cfg)
    q = _random_field_spectrum(cfg, lmax)
    s = plan.stream
    plan.qlm_in.upload(q, s)
    plan.encode_stage(s, "legSynth")
    fm = plan.read_fm()
    fm_ref = ref.legendre_synth(q.astype(np.float64))
    e1 = O.rel_l2(fm, fm_ref)
    plan.encode_stage(s, "fourSynth")
    spat = plan.spat_buf.download(s)
    spat_ref = ref.fourier_synth_fast(fm.astype(np.float64))
    e2 = O.rel_l2(spat, spat_ref)
    plan.encode_stage(s, "fourAnalys")
    gm = plan.read_fm()
    gm_ref = ref.fourier_analys_fast(spat.astype(np.float64))
    # the GPU stage leaves the 2pi/nphi factor in the Gauss weight (sht.ts:198)
    e3 = O.rel_l2(gm * (2 * np.pi / nphi), gm_ref)
    plan.encode_stage(s, "legAnalys")
    qo = plan.qlm_out.download(s)
    qo_ref = ref.legendre_analys(gm.astype(np.float64) * (2 * np.pi / nphi))
    e4 = O.rel_l2(qo, qo_ref)
    # end to end, and the round trip is the identity for band-limited input
    rt = plan.analys(plan.synth(q))
    e5 = O.rel_l2(rt, q)
    emit(f"  legSynth {e1:.2e} fourSynth {e2:.2e} fourAnalys {e3:.2e} legAnalys {e4:.2e}"
         f" round trip {e5:.2e}")
    assert max(e2, e3) < 2e-7          # the FFT stages: a few ulps
    assert max(e1, e4) < leg_tol(lmax)
    assert e5 < leg_tol(lmax)
    plan.destroy()


def test_single_mode_packing(gpu):
    """One (l, m) coefficient synthesises Y_lm at its lm_index slot and nowhere
    else: bit-exact packing (layout.ts:30-32)."""
    from scipy.special import sph_harm_y

    cfg = O.Cfg(20, 20, 32, 64)
    plan = _plan(cfg)
    ref = O.ShtReference(cfg)
    theta = np.arccos(ref.ct)
    phi = 2 * np.pi * np.arange(cfg.nphi) / cfg.nphi
    for (l, m) in [(0, 0), (5, 2), (20, 20), (13, 0), (7, 6)]:
        emit(f"single mode l={l} m={m}")
        q = np.zeros(2 * cfg.nlm, np.float32)
        i = O.lm_index(cfg.lmax, l, m)
        q[2 * i] = 1.0
        spat = plan.synth(q).reshape(cfg.nlat, cfg.nphi)
        y = sph_harm_y(l, m, theta[:, None], phi[None, :])
        want = y.real * (1 if m == 0 else 2)
        assert np.max(np.abs(spat - want)) < 5e-6
        back = plan.analys(spat)
        nz = np.flatnonzero(np.abs(back) > 1e-5)
        assert list(nz) == [2 * i], (l, m, nz[:10])
    plan.destroy()


def test_plan_validation_messages(gpu):
    import tsphere

    with pytest.raises(ValueError, match="nlat must be an integer > lmax"):
        tsphere.ShtPlan.create(tsphere.ShtConfig(31, 31, 31, 64))
    with pytest.raises(ValueError, match="fourier:'fft' requires"):
        tsphere.ShtPlan.create(tsphere.ShtConfig(15, 15, 16, 33), fourier="fft")


def test_batched_fields(gpu):
    """A batched transform equals the per-field transforms."""
    import tsphere

    cfg = O.Cfg(63, 63, 64, 128)
    nf = 5
    plan = _plan(cfg, max_fields=nf)
    qs = np.stack([_random_field_spectrum(cfg, 100 + f) for f in range(nf)])
    s = plan.stream
    qb = tsphere.DeviceBuffer(nf * 2 * cfg.nlm)
    sb = tsphere.DeviceBuffer(nf * cfg.npts)
    qb.upload(qs, s)
    plan.encode_synth_into(s, plan.create_synth_binding(qb, sb, nf, 2 * cfg.nlm, cfg.npts))
    spat = sb.download(s).reshape(nf, -1)
    for f in range(nf):
        one = plan.synth(qs[f])
        assert np.array_equal(one, spat[f]), f
    plan.encode_analys_into(s, plan.create_analys_binding(sb, qb, nf, cfg.npts, 2 * cfg.nlm))
    back = qb.download(s).reshape(nf, -1)
    for f in range(nf):
        assert O.rel_l2(back[f], qs[f]) < leg_tol(cfg.lmax)
    plan.destroy()


@pytest.mark.parametrize("lmax,nlat,nphi", [
    (255, 256, 512), (127, 129, 512),      # 256-point rows; band limit below the row
    (511, 512, 1024), (300, 301, 1024),    # 512-point rows
    (1023, 1024, 2048), (600, 601, 2048),  # 1024-point rows (the benchmarked grid)
    (1365, 1366, 4096),                    # 2048-point rows (radix-4 kernels)
])
def test_fourier_stage_rows(gpu, lmax, nlat, nphi):
    """Both Fourier stages alone on random rows against numpy's real FFT (the
    oracle's fast stages), at the row lengths of the one-wave-per-row kernels
    (nphi 512 .. 2048) and beyond, with band limits below the row length and
    odd latitude counts: a few ulps."""
    cfg = O.Cfg(lmax, lmax, nlat, nphi)
    emit(f"fourier stages lmax{lmax} {nlat}x{nphi}")
    plan = _plan(cfg)
    ref = O.ShtReference(cfg)
    rng = np.random.default_rng(lmax + nphi)
    fm_dev = rng.normal(size=(nlat, lmax + 1, 2)).astype(np.float32)
    fm_dev[:, 0, 1] = 0.0  # X[0] = Re F_0
    s = plan.stream
    plan.fm_buf.upload(fm_dev.reshape(-1), s)
    plan.encode_stage(s, "fourSynth")
    spat = plan.spat_buf.download(s)
    fm_ref = np.ascontiguousarray(fm_dev.transpose(1, 0, 2)).reshape(-1)  # [m][ilat]
    e_s = O.rel_l2(spat, ref.fourier_synth_fast(fm_ref.astype(np.float64)))
    g = rng.normal(size=nlat * nphi).astype(np.float32)
    plan.spat_buf.upload(g, s)
    plan.encode_stage(s, "fourAnalys")
    gm = plan.read_fm()
    e_a = O.rel_l2(gm * (2 * np.pi / nphi), ref.fourier_analys_fast(g.astype(np.float64)))
    emit(f"  fourSynth {e_s:.2e} fourAnalys {e_a:.2e}")
    assert e_s < 2e-7 and e_a < 2e-7  # measured 1.0-1.2e-7
    plan.destroy()


@pytest.mark.parametrize("lmax", [255, 511, 1023])
def test_round_trip_large(gpu, lmax):
    """Full-size grids (nlat = lmax+1, nphi = 2(lmax+1)): round trip identity,
    and synthesis vs the f64 oracle's fast transform."""
    cfg = O.Cfg(lmax, lmax, lmax + 1, 2 * (lmax + 1))
    emit(f"round trip lmax{lmax}")
    plan = _plan(cfg)
    q = _random_field_spectrum(cfg, 7)
    spat = plan.synth(q)
    back = plan.analys(spat)
    e_rt = O.rel_l2(back, q)
    ref = O.ShtReference(cfg)
    e_s = O.rel_l2(spat, ref.synth_fast(q.astype(np.float64)))
    emit(f"  round trip {e_rt:.2e} synth vs f64 {e_s:.2e}")
    assert np.all(np.isfinite(back))
    assert e_rt < leg_tol(lmax) and e_s < leg_tol(lmax)
    plan.destroy()


@pytest.mark.parametrize("legendre", ["fp32", "bf16x3"])
def test_sixty_four_fields_lmax511(gpu, legendre):
    """The batched SHT configuration: 64 fields at lmax 511 on 512 x 1024,
    a_lm with sigma_l = 1/(l+1), seed = field index, synthesised and analysed
    in one launch per stage.  Every field's round trip, and the synthesis /
    analysis of three fields against the f64 oracle, within the form's
    tolerance (FP32: the plain-FP32 bound of this file; split-BF16: 1e-5 per
    transform, 2e-5 per round trip)."""
    import tsphere

    nf, lmax = 64, 511
    cfg = O.Cfg(lmax, lmax, 512, 1024)
    emit(f"64 fields lmax 511 {legendre}")
    plan = tsphere.ShtPlan.create(tsphere.ShtConfig(lmax, lmax, 512, 1024), legendre=legendre,
                                  max_fields=nf)
    qs = np.stack([_random_field_spectrum(cfg, f) for f in range(nf)])
    qb = tsphere.DeviceBuffer(nf * 2 * cfg.nlm)
    sb = tsphere.DeviceBuffer(nf * cfg.npts)
    ob = tsphere.DeviceBuffer(nf * 2 * cfg.nlm)
    s = plan.stream
    qb.upload(qs, s)
    plan.encode_synth_into(s, plan.create_synth_binding(qb, sb, nf, 2 * cfg.nlm, cfg.npts))
    plan.encode_analys_into(s, plan.create_analys_binding(sb, ob, nf, cfg.npts, 2 * cfg.nlm))
    spat = sb.download(s).reshape(nf, -1)
    back = ob.download(s).reshape(nf, -1)
    tol, tol_rt = (leg_tol(lmax), leg_tol(lmax)) if legendre == "fp32" else (1e-5, 2e-5)
    rts = [O.rel_l2(back[f], qs[f]) for f in range(nf)]
    emit(f"  round trips: max {max(rts):.2e} mean {np.mean(rts):.2e}")
    assert max(rts) < tol_rt
    ref = O.ShtReference(cfg)
    for f in (0, 31, 63):
        es = O.rel_l2(spat[f], ref.synth_fast(qs[f].astype(np.float64)))
        ea = O.rel_l2(back[f], ref.analys_fast(spat[f].astype(np.float64)))
        emit(f"  field {f}: synth {es:.2e} analys {ea:.2e} vs f64")
        assert es < tol and ea < tol
    plan.destroy()


def test_napi_binding_transform(gpu, tmp_path):
    """A node process drives the transforms through the N-API addon exactly as
    the reference's ShtPlan would (create, device buffers, encode, read
    back); its synthesis matches the f64 oracle and its round trip the input."""
    import os
    import shutil
    import subprocess

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = os.path.join(root, "turing-sphere_amd", "napi", "synth_file.js")
    node = shutil.which("node")
    if node is None or not os.path.exists(os.path.join(os.path.dirname(script),
                                                       "tsphere_napi.node")):
        pytest.skip("node or the built addon is not available")
    cfg = O.Cfg(63, 63, 64, 128)
    emit("N-API synth/analys lmax63 64x128 (node child process)")
    q = _random_field_spectrum(cfg, 17)
    fin, fs, fq = (str(tmp_path / n) for n in ("q.f32", "spat.f32", "back.f32"))
    q.astype(np.float32).tofile(fin)
    r = subprocess.run([node, script, "63", "64", "128", fin, fs, fq], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    spat = np.fromfile(fs, np.float32)
    back = np.fromfile(fq, np.float32)
    ref = O.ShtReference(cfg)
    e_s = O.rel_l2(spat, ref.synth_fast(q.astype(np.float64)))
    e_rt = O.rel_l2(back, q)
    emit(f"  {r.stdout.strip()} synth vs f64 {e_s:.2e} round trip {e_rt:.2e}")
    assert e_s < leg_tol(63) and e_rt < leg_tol(63)
