"""The spectral op's C entry point (tsphere_leg_spectral_imex: Legendre
analysis, the diagonal implicit step and Legendre synthesis of the new state
in one pass per m-row) against the three launches it replaces, called
through the same C ABI: tsphere_leg_analys, the implicit step
X_new = (X + s1 A) / (1 + s2 lam) evaluated op by op in float32 on the host
(IEEE single rounding of each product, sum and quotient, as the kernels
round it), and tsphere_leg_synth of X_new.  The two analyses may sum the
latitude pairs in different orders (the fused kernel picks its own
workgroup shape), so the comparison is to 2e-6 relative, not bit for bit
(the planner-level tests compare the fused and separate step plans
exactly).  One to four fields: three and four fields at lmax 15 have fewer
latitude pairs than one wave, and run in the two-wave workgroup the fused
kernel's combine needs."""
import ctypes as C
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import emit

pytestmark = pytest.mark.gpu


CASES = ([(lmax, nlat, nphi, nf, None) for lmax, nlat, nphi in
          [(15, 16, 32), (63, 64, 128), (255, 256, 512)] for nf in (1, 2, 3, 4)] +
         # the benchmarked shape (512 pairs, 8-wave workgroups) and 768 pairs
         # (two 64-pair chains per lane), two fields; species maps that are
         # not the identity (analysed field f updates state species[f])
         [(1023, 1024, 2048, 2, None), (767, 1536, 2048, 2, None),
          (63, 64, 128, 2, (1, 0)), (255, 256, 512, 3, (2, 0, 1)),
          (1023, 1024, 2048, 2, (1, 0))])


# every case in the FP32 and FP64-recurrence forms; the split-BF16 form's
# spectral op covers up to 512 latitude pairs (its refusal above that is
# tested below)
PARAMS = [(leg,) + c for leg in ("fp32", "f64rec", "bf16x3") for c in CASES
          if leg != "bf16x3" or (c[1] + 1) // 2 <= 512]


@pytest.mark.parametrize("legendre,lmax,nlat,nphi,nfield,species", PARAMS)
def test_spectral_entry_matches_separate_launches(gpu, legendre, lmax, nlat, nphi, nfield,
                                                  species):
    import tsphere
    from tsphere import _lib
    from tsphere.device import DeviceBuffer

    spc_map = list(species) if species is not None else list(range(nfield))
    emit(f"spectral abi {legendre} lmax {lmax} {nlat}x{nphi} nfield {nfield} "
         f"species {spc_map}")
    L = _lib.load()
    cfg = tsphere.ShtConfig(lmax, lmax, nlat, nphi)
    plan = tsphere.ShtPlan.create(cfg, legendre=legendre, max_fields=4)
    s = plan.stream
    nlm, M = plan.nlm, lmax
    rng = np.random.default_rng(1000 + 10 * lmax + nfield)
    nfm = nfield * 2 * (M + 1) * nlat
    fm0 = rng.normal(size=nfm).astype(np.float32)
    x0 = (0.1 * rng.normal(size=(nfield, 2 * nlm))).astype(np.float32)
    ell = np.repeat(np.concatenate([np.arange(m, lmax + 1) for m in range(M + 1)]), 2)
    lam = (ell * (ell + 1)).astype(np.float32)
    s1 = np.float32([0.002, 0.003, 0.001, 0.004])[:nfield]
    s2 = np.float32([4e-7, 3e-6, 1e-5, 2e-6])[:nfield]

    # fused: one launch over the scratch and the state
    fmA = DeviceBuffer(nfm, "fmA")
    fmA.upload(fm0, s)
    xA = [DeviceBuffer(2 * nlm, f"xA{f}") for f in range(nfield)]
    for f in range(nfield):
        xA[f].upload(x0[f], s)
    lamb = DeviceBuffer(2 * nlm, "lam")
    lamb.upload(lam, s)
    ptrs = (C.c_void_p * nfield)(*[b.ptr for b in xA])
    cs1 = (C.c_float * nfield)(*s1.tolist())
    cs2 = (C.c_float * nfield)(*s2.tolist())
    spc = (C.c_int * nfield)(*spc_map)
    _lib.check(L.tsphere_leg_spectral_imex(plan.handle, nfield, fmA.ptr, ptrs, lamb.ptr, cs1, cs2,
                                           spc, s.handle), "spectral")
    got_x = np.stack([b.download(s) for b in xA])
    got_fm = fmA.download(s)

    # separate: analysis, host implicit step, synthesis.  Analysed field f
    # updates state species[f] with that field's scalars; the synthesis
    # writes scratch slot = state index.
    fmB = DeviceBuffer(nfm, "fmB")
    fmB.upload(fm0, s)
    ab = DeviceBuffer(nfield * 2 * nlm, "A")
    _lib.check(L.tsphere_leg_analys(plan.handle, nfield, fmB.ptr, ab.ptr, 2 * nlm, s.handle),
               "analys")
    A = ab.download(s).reshape(nfield, 2 * nlm)
    xn = np.empty_like(x0)
    for f in range(nfield):
        sp = spc_map[f]
        num = x0[sp] + s1[f] * A[f]          # float32 ops, each rounded once
        den = np.float32(1.0) + s2[f] * lam
        xn[sp] = num / den
    xb = DeviceBuffer(nfield * 2 * nlm, "xn")
    xb.upload(xn.reshape(-1), s)
    _lib.check(L.tsphere_leg_synth(plan.handle, nfield, xb.ptr, 2 * nlm, fmB.ptr, s.handle),
               "synth")
    want_fm = fmB.download(s)

    ex = O.rel_l2(got_x, xn)
    ef = O.rel_l2(got_fm, want_fm)
    emit(f"  state {ex:.2e}  scratch {ef:.2e}")
    assert np.all(np.isfinite(got_fm)) and ex < 2e-6 and ef < 2e-6
    plan.destroy()


def test_split_bf16_spectral_entry_refuses_beyond_block_limit(gpu):
    """A split-BF16 plan with 768 latitude pairs: the spectral entry point
    and the fused implicit step return TSPHERE_ERR_INVALID with a message
    (no launch), as the planner expects."""
    import tsphere
    from tsphere import _lib
    from tsphere.device import DeviceBuffer

    lmax, nlat, nphi = 767, 1536, 2048
    emit(f"split-BF16 spectral refusal lmax {lmax} {nlat}x{nphi}")
    L = _lib.load()
    plan = tsphere.ShtPlan.create(tsphere.ShtConfig(lmax, lmax, nlat, nphi), legendre="bf16x3",
                                  max_fields=2)
    nlm = plan.nlm
    fm = DeviceBuffer(2 * 2 * (lmax + 1) * nlat, "fm")
    x = [DeviceBuffer(2 * nlm, f"x{f}") for f in range(2)]
    lam = DeviceBuffer(2 * nlm, "lam")
    ptrs = (C.c_void_p * 2)(*[b.ptr for b in x])
    s1 = (C.c_float * 2)(0.002, 0.002)
    s2 = (C.c_float * 2)(1e-6, 1e-6)
    spc = (C.c_int * 2)(0, 1)
    rc = L.tsphere_leg_spectral_imex(plan.handle, 2, fm.ptr, ptrs, lam.ptr, s1, s2, spc,
                                     plan.stream.handle)
    assert rc == _lib.ERR_INVALID and "512 latitude pairs" in _lib.last_error()
    rc = L.tsphere_leg_imex_synth(plan.handle, 2, ptrs, ptrs, lam.ptr, s1, s2, fm.ptr,
                                  plan.stream.handle)
    assert rc == _lib.ERR_INVALID and "512 latitude pairs" in _lib.last_error()
    plan.destroy()


def test_napi_spectral_entry_matches_ctypes(gpu, tmp_path):
    """legSpectralImex through the N-API addon, in a node child process, gives
    exactly what the ctypes call gives on the same inputs (an FP64-recurrence
    plan, so the addon's legendre option is exercised too)."""
    import tsphere
    from tsphere import _lib
    from tsphere.device import DeviceBuffer

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    addon = os.path.join(root, "turing-sphere_amd", "napi", "tsphere_napi.node")
    node = shutil.which("node")
    if node is None or not os.path.exists(addon):
        pytest.skip("node or the built addon is not available")
    lmax, nlat, nphi, nfield = 63, 64, 128, 2
    emit(f"N-API legSpectralImex f64rec lmax {lmax} (node child process)")
    cfg = tsphere.ShtConfig(lmax, lmax, nlat, nphi)
    plan = tsphere.ShtPlan.create(cfg, legendre="f64rec", max_fields=nfield)
    s = plan.stream
    nlm = plan.nlm
    rng = np.random.default_rng(77)
    fm0 = rng.normal(size=nfield * 2 * (lmax + 1) * nlat).astype(np.float32)
    x0 = (0.1 * rng.normal(size=(nfield, 2 * nlm))).astype(np.float32)
    ell = np.repeat(np.concatenate([np.arange(m, lmax + 1) for m in range(lmax + 1)]), 2)
    lam = (ell * (ell + 1)).astype(np.float32)
    s1 = np.float32([0.002, 0.003])
    s2 = np.float32([4e-7, 3e-6])
    files = {k: str(tmp_path / f"{k}.f32") for k in ("fm", "x0", "x1", "lam", "ofm", "ox0", "ox1")}
    fm0.tofile(files["fm"])
    x0[0].tofile(files["x0"])
    x0[1].tofile(files["x1"])
    lam.tofile(files["lam"])
    js = f"""
const fs = require('fs');
const ts = require({json.dumps(addon)});
const f = {json.dumps(files)};
const rd = p => new Float32Array(fs.readFileSync(p).buffer.slice(0));
const plan = ts.planCreate({{lmax: {lmax}, mmax: {lmax}, nlat: {nlat}, nphi: {nphi},
                             legendre: 'f64rec', maxFields: {nfield}}});
const h = {{}};
for (const k of ['fm', 'x0', 'x1', 'lam']) {{
  const a = rd(f[k]); h[k] = [ts.malloc(4 * a.length), a.length]; ts.upload(h[k][0], a);
}}
ts.legSpectralImex(plan, h.fm[0], [h.x0[0], h.x1[0]], h.lam[0],
                   [{float(s1[0])!r}, {float(s1[1])!r}], [{float(s2[0])!r}, {float(s2[1])!r}], [0, 1]);
for (const [k, o] of [['fm', 'ofm'], ['x0', 'ox0'], ['x1', 'ox1']])
  fs.writeFileSync(f[o], Buffer.from(ts.download(h[k][0], h[k][1]).buffer));
for (const k of Object.keys(h)) ts.free(h[k][0]);
ts.planDestroy(plan);
console.log('ok');
"""
    r = subprocess.run([node, "-e", js], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr

    fmA = DeviceBuffer(fm0.size, "fmA")
    fmA.upload(fm0, s)
    xA = [DeviceBuffer(2 * nlm, f"xA{f}") for f in range(nfield)]
    for f in range(nfield):
        xA[f].upload(x0[f], s)
    lamb = DeviceBuffer(2 * nlm, "lam")
    lamb.upload(lam, s)
    L = _lib.load()
    _lib.check(L.tsphere_leg_spectral_imex(
        plan.handle, nfield, fmA.ptr, (C.c_void_p * nfield)(*[b.ptr for b in xA]), lamb.ptr,
        (C.c_float * nfield)(*s1.tolist()), (C.c_float * nfield)(*s2.tolist()),
        (C.c_int * nfield)(0, 1), s.handle), "spectral")
    assert np.array_equal(np.fromfile(files["ofm"], np.float32), fmA.download(s))
    for f in range(nfield):
        assert np.array_equal(np.fromfile(files[f"ox{f}"], np.float32), xA[f].download(s))
    plan.destroy()


@pytest.mark.parametrize("legendre", ["fp32", "f64rec", "bf16x3"])
def test_batched_fold_entry_matches_per_run_calls(gpu, legendre):
    """tsphere_leg_imex_synth_fields over 2 species x 5 runs (field f =
    species * 5 + run, the scalars of field f at coef[2 f], coef[2 f + 1] in
    device memory) against tsphere_leg_imex_synth called once per run with
    that run's scalars: the implicit step is element-wise and the synthesis
    sums each field over l in one order, so the state and the scratch agree
    bit for bit."""
    import tsphere
    from tsphere import _lib
    from tsphere.device import DeviceBuffer

    lmax, nlat, nphi, nsp, R = 31, 32, 64, 2, 5
    emit(f"batched fold entry {legendre} lmax {lmax} species {nsp} runs {R}")
    L = _lib.load()
    plan = tsphere.ShtPlan.create(tsphere.ShtConfig(lmax, lmax, nlat, nphi), legendre=legendre,
                                  max_fields=nsp * R)
    s, nlm, M = plan.stream, plan.nlm, lmax
    rng = np.random.default_rng(7)
    x0 = (0.1 * rng.normal(size=(nsp, R, 2 * nlm))).astype(np.float32)
    a0 = (0.1 * rng.normal(size=(nsp, R, 2 * nlm))).astype(np.float32)
    ell = np.repeat(np.concatenate([np.arange(m, lmax + 1) for m in range(M + 1)]), 2)
    lamb = DeviceBuffer(2 * nlm, "lam")
    lamb.upload((ell * (ell + 1)).astype(np.float32), s)
    s1 = np.float32(rng.uniform(1e-3, 5e-3, size=(nsp, R)))
    s2 = np.float32(rng.uniform(1e-6, 1e-4, size=(nsp, R)))
    fsz = 2 * (M + 1) * nlat  # floats per field of the scratch

    def buffers(tag):
        xs = [DeviceBuffer(R * 2 * nlm, f"x{tag}{k}") for k in range(nsp)]
        an = [DeviceBuffer(R * 2 * nlm, f"a{tag}{k}") for k in range(nsp)]
        for k in range(nsp):
            xs[k].upload(x0[k].reshape(-1), s)
            an[k].upload(a0[k].reshape(-1), s)
        return xs, an

    xs, an = buffers("B")
    coef = DeviceBuffer(2 * nsp * R, "coef")
    coef.upload(np.stack([s1, s2], axis=-1).reshape(-1), s)
    fmB = DeviceBuffer(nsp * R * fsz, "fmB")
    _lib.check(L.tsphere_leg_imex_synth_fields(
        plan.handle, nsp, R, (C.c_void_p * nsp)(*[b.ptr for b in an]),
        (C.c_void_p * nsp)(*[b.ptr for b in xs]), 2 * nlm, lamb.ptr, coef.ptr, fmB.ptr,
        s.handle), "imex_synth_fields")
    got_x = np.stack([b.download(s).reshape(R, -1) for b in xs])
    got_fm = fmB.download(s).reshape(nsp, R, fsz)

    xr, ar = buffers("R")
    fmR = DeviceBuffer(nsp * fsz, "fmR")
    for r in range(R):
        off = 4 * r * 2 * nlm
        _lib.check(L.tsphere_leg_imex_synth(
            plan.handle, nsp, (C.c_void_p * nsp)(*[b.ptr + off for b in ar]),
            (C.c_void_p * nsp)(*[b.ptr + off for b in xr]), lamb.ptr,
            (C.c_float * nsp)(*s1[:, r].tolist()), (C.c_float * nsp)(*s2[:, r].tolist()),
            fmR.ptr, s.handle), "imex_synth")
        want_fm = fmR.download(s).reshape(nsp, fsz)
        assert np.array_equal(got_fm[:, r], want_fm)
    want_x = np.stack([b.download(s).reshape(R, -1) for b in xr])
    assert np.all(np.isfinite(got_fm)) and np.array_equal(got_x, want_x)
    assert not np.array_equal(got_x, x0)
    plan.destroy()


@pytest.mark.parametrize("legendre,impl", [("fp32", 1), ("fp32", 3), ("fp32", 5), ("f64rec", 1),
                                           ("f64rec", 3), ("f64rec", 5), ("bf16x3", None)])
@pytest.mark.parametrize("lmax,nlat,nphi,nsp,R", [(31, 32, 64, 2, 5), (63, 65, 128, 1, 19)])
def test_batched_analysis_fold_entry(gpu, legendre, impl, lmax, nlat, nphi, nsp, R):
    """tsphere_leg_analys_imex_fields (the analysis with the implicit step of
    each field behind it, the scalars of field f at coef[2 f], coef[2 f + 1])
    against tsphere_leg_analys_fields followed by the implicit step of
    tsphere_leg_imex_synth_fields: the same analysed values, rounded as
    written, enter the same step, so the state agrees bit for bit.  With
    the 4x4-block (impl 3) and per-parity product (impl 5) tiles the step
    runs in the analysis kernel as it forms each analysed value and the
    analysed buffers stay untouched; the VALU tile (impl 1) and the
    split-BF16 form run it as a second kernel.  19 runs leave a partial
    group of fields, an odd grid pairs the equator with itself."""
    import tsphere
    from tsphere import _lib
    from tsphere.device import DeviceBuffer

    emit(f"analysis fold entry {legendre} tile {impl} lmax {lmax} {nlat}x{nphi} "
         f"species {nsp} runs {R}")
    L = _lib.load()
    plan = tsphere.ShtPlan.create(tsphere.ShtConfig(lmax, lmax, nlat, nphi), legendre=legendre,
                                  max_fields=nsp * R)
    if impl is not None:
        plan.set_tile("analys", impl=impl)
    s, nlm, M = plan.stream, plan.nlm, lmax
    rng = np.random.default_rng(17)
    fsz = 2 * (M + 1) * nlat
    fm = DeviceBuffer(nsp * R * fsz, "fm")
    fm.upload(rng.normal(size=nsp * R * fsz).astype(np.float32), s)
    x0 = (0.1 * rng.normal(size=(nsp, R, 2 * nlm))).astype(np.float32)
    ell = np.repeat(np.concatenate([np.arange(m, lmax + 1) for m in range(M + 1)]), 2)
    lamb = DeviceBuffer(2 * nlm, "lam")
    lamb.upload((ell * (ell + 1)).astype(np.float32), s)
    coef = DeviceBuffer(2 * nsp * R, "coef")
    s1 = np.float32(rng.uniform(1e-3, 5e-3, size=(nsp, R)))
    s2 = np.float32(rng.uniform(1e-6, 1e-4, size=(nsp, R)))
    coef.upload(np.stack([s1, s2], axis=-1).reshape(-1), s)

    def buffers(tag):
        xs = [DeviceBuffer(R * 2 * nlm, f"x{tag}{k}") for k in range(nsp)]
        an = [DeviceBuffer(R * 2 * nlm, f"a{tag}{k}") for k in range(nsp)]
        for k in range(nsp):
            xs[k].upload(x0[k].reshape(-1), s)
            an[k].fill(np.nan, s)
        return xs, an

    def ptrs(bs):
        return (C.c_void_p * nsp)(*[b.ptr for b in bs])

    xb, ab = buffers("B")
    _lib.check(L.tsphere_leg_analys_imex_fields(plan.handle, nsp, R, fm.ptr, ptrs(ab), ptrs(xb),
                                                2 * nlm, lamb.ptr, coef.ptr, s.handle),
               "analys_imex_fields")
    got_x = np.stack([b.download(s) for b in xb])
    got_a = np.stack([b.download(s) for b in ab])

    xr, ar = buffers("R")
    _lib.check(L.tsphere_leg_analys_fields(plan.handle, nsp, R, fm.ptr, ptrs(ar), 2 * nlm,
                                           s.handle), "analys_fields")
    want_a = np.stack([b.download(s) for b in ar])
    junk = DeviceBuffer(nsp * R * fsz, "junk")
    _lib.check(L.tsphere_leg_imex_synth_fields(plan.handle, nsp, R, ptrs(ar), ptrs(xr), 2 * nlm,
                                               lamb.ptr, coef.ptr, junk.ptr, s.handle),
               "imex_synth_fields")
    want_x = np.stack([b.download(s) for b in xr])
    assert np.all(np.isfinite(got_x)) and np.array_equal(got_x, want_x)
    assert not np.array_equal(got_x, x0.reshape(nsp, -1))
    if plan.get_tile("analys", nsp * R)["impl"] in (3, 5) and legendre != "bf16x3":
        assert np.all(np.isnan(got_a))
    else:
        assert np.array_equal(got_a, want_a)
    plan.destroy()
