"""Spherical-harmonic transform plan on the gfx950 library.

Mirrors ``ShtPlan`` of src/sht/sht.ts:116-469: a plan is created for one
``ShtConfig`` (validated with the reference's messages), owns its tables and
the Fourier scratch, and transforms caller-owned device buffers; host-array
conveniences ``synth``/``analys``/``synth_from`` upload, run and read back.
Where the reference records into a compute pass, this records onto a HIP
stream (``encode_synth_into`` / ``encode_analys_into``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .device import DeviceBuffer, Stream
from .layout import ShtConfig, nlm_calc, validate_config

FOURIER_MODES = {"auto": _lib.FOURIER_AUTO, "fft": _lib.FOURIER_FFT, "dft": _lib.FOURIER_DFT}
LEGENDRE_MODES = {"fp32": _lib.LEG_FP32, "bf16x3": _lib.LEG_BF16X3, "f64rec": _lib.LEG_F64REC}


@dataclass(frozen=True)
class ShtBinding:
    """createSynthBinding / createAnalysBinding (sht.ts:299-323): the caller's
    buffers a transform reads and writes, plus how many fields it covers."""
    kind: str  # 'synth' | 'analys'
    src: DeviceBuffer
    dst: DeviceBuffer
    nfield: int = 1
    src_stride: int = 0
    dst_stride: int = 0


def _cfg_struct(cfg: ShtConfig, fourier: str, legendre: str, max_fields: int) -> _lib.ShtConfigC:
    if fourier not in FOURIER_MODES:
        raise ValueError(f"fourier must be one of {sorted(FOURIER_MODES)} (got {fourier!r})")
    if legendre not in LEGENDRE_MODES:
        raise ValueError(f"legendre must be one of {sorted(LEGENDRE_MODES)} (got {legendre!r})")
    return _lib.ShtConfigC(cfg.lmax, cfg.mmax, cfg.nlat, cfg.nphi, FOURIER_MODES[fourier],
                           LEGENDRE_MODES[legendre], int(max_fields))


def plan_tables(cfg: ShtConfig) -> dict:
    """The plan's f32 device tables, built on the host (no GPU needed)."""
    validate_config(cfg)
    nlm = nlm_calc(cfg.lmax, cfg.mmax)
    npair = (cfg.nlat + 1) // 2
    out = {
        "ab": np.zeros(2 * nlm, np.float32),
        "amm": np.zeros(cfg.mmax + 1, np.float32),
        "ctstw": np.zeros(3 * cfg.nlat, np.float32),
        "seed_y0": np.zeros((cfg.mmax + 1) * npair, np.float32),
        "seed_y1": np.zeros((cfg.mmax + 1) * npair, np.float32),
        "seed_l": np.zeros((cfg.mmax + 1) * npair, np.int32),
        "trig": np.zeros(2 * cfg.nphi, np.float32),
    }
    c = _cfg_struct(cfg, "auto", "fp32", 1)
    _lib.call("tsphere_plan_tables", C.byref(c), *(out[k].ctypes.data for k in
                                                   ("ab", "amm", "ctstw", "seed_y0", "seed_y1",
                                                    "seed_l", "trig")))
    return out


class ShtPlan:
    """A transform plan on the GPU (sht.ts:116)."""

    def __init__(self, handle, cfg: ShtConfig, info: _lib.PlanInfoC):
        self._h = handle
        self.cfg = cfg
        self.nlm = int(info.nlm)
        self.fourier_mode = "fft" if info.fourier == _lib.FOURIER_FFT else "dft"
        self.legendre = {v: k for k, v in LEGENDRE_MODES.items()}[info.legendre]
        self.leg_lat = int(info.leg_lat)
        self.max_fields = int(info.max_fields)
        self.scratch_bytes = int(info.scratch_bytes)
        n = cfg.nlat
        self.theta = np.zeros(n)
        self.cos_theta = np.zeros(n)
        self.gauss_weights = np.zeros(n)
        _lib.call("tsphere_plan_grid", handle, self.theta.ctypes.data,
                  self.cos_theta.ctypes.data, self.gauss_weights.ctypes.data)
        # Plan-owned staging buffers of the host-array conveniences (sht.ts:213-218).
        self.qlm_in = DeviceBuffer(2 * self.nlm, "sht-qlm-in")
        self.qlm_out = DeviceBuffer(2 * self.nlm, "sht-qlm-out")
        self.spat_buf = DeviceBuffer(cfg.npts, "sht-spat")
        self.fm_buf = DeviceBuffer(2 * (cfg.mmax + 1) * cfg.nlat, "sht-fm")
        self.stream = Stream()

    @staticmethod
    def create(cfg: ShtConfig, fourier: str = "auto", legendre: str = "fp32",
               max_fields: int = 1) -> "ShtPlan":
        """ShtPlan.create (sht.ts:166-183)."""
        validate_config(cfg)
        c = _cfg_struct(cfg, fourier, legendre, max_fields)
        h = C.c_void_p()
        rc = _lib.load().tsphere_plan_create(C.byref(c), C.byref(h))
        if rc == _lib.ERR_INVALID:
            raise ValueError(_lib.last_error())
        _lib.check(rc, "ShtPlan.create")
        info = _lib.PlanInfoC()
        _lib.call("tsphere_plan_query", h.value, C.byref(info))
        return ShtPlan(h.value, cfg, info)

    @property
    def handle(self):
        return self._h

    # ------------------------------------------------------ tile shapes
    STAGES = {"synth": 0, "analys": 1}

    def set_tile(self, stage: str, nf: int = 0, nwaves: int = 0, k: int = 0,
                 ctc: int = 0, impl: int = 0, rows: int = 0) -> None:
        """Fix the Legendre workgroup shape of one stage (0 = built-in); for
        the FP32 and FP64-recurrence forms impl 1 = VALU contraction,
        2 = FP32 matrix cores as 16x16x4 tiles, 3 = FP32 matrix cores as
        16 4x4 blocks per wave, 5 = FP32 matrix cores as per-parity
        products whose 16 columns are (field, re/im) of 8 fields (batched
        transforms: nf 8 or 16, k 1, rows 1 -- 2 for the 8-field synthesis,
        two consecutive m-rows per workgroup -- up to 512 latitude pairs);
        nf = fields per workgroup (1, 2, 4, and 8 for impl 3 with k 1 or 2).
        Shapes the library does not have raise."""
        t = _lib.LegTileC(nf, nwaves, k, ctc, impl, rows)
        _lib.call("tsphere_plan_set_tile", self._h, self.STAGES[stage], C.byref(t))

    def get_tile(self, stage: str, nfield: int) -> dict:
        t = _lib.LegTileC()
        _lib.call("tsphere_plan_get_tile", self._h, self.STAGES[stage], nfield, C.byref(t))
        return {"nf": t.nf, "nwaves": t.nwaves, "k": t.k, "ctc": t.ctc, "impl": t.impl,
                "rows": t.rows}

    def tune(self, nfield: int, reps: int = 5, cache: bool = True) -> dict:
        """Autotune the Legendre tiles for `nfield` fields (timed on the
        device, blocking); results are cached per device and shape (see
        tsphere.tune) so later plans of the same shape skip the timing."""
        from . import tune as _tune

        return _tune.tune_plan(self, nfield, reps, cache)

    # ------------------------------------------------------ bindings
    def create_synth_binding(self, qlm_in: DeviceBuffer, spat_out: DeviceBuffer,
                             nfield: int = 1, qlm_stride: int = 0,
                             spat_stride: int = 0) -> ShtBinding:
        self._check_binding(qlm_in, 2 * self.nlm, nfield, qlm_stride)
        self._check_binding(spat_out, self.cfg.npts, nfield, spat_stride)
        return ShtBinding("synth", qlm_in, spat_out, nfield, qlm_stride, spat_stride)

    def create_analys_binding(self, spat_in: DeviceBuffer, qlm_out: DeviceBuffer,
                              nfield: int = 1, spat_stride: int = 0,
                              qlm_stride: int = 0) -> ShtBinding:
        self._check_binding(spat_in, self.cfg.npts, nfield, spat_stride)
        self._check_binding(qlm_out, 2 * self.nlm, nfield, qlm_stride)
        return ShtBinding("analys", spat_in, qlm_out, nfield, spat_stride, qlm_stride)

    def _check_binding(self, buf: DeviceBuffer, per_field: int, nfield: int, stride: int) -> None:
        if nfield < 1 or nfield > self.max_fields:
            raise ValueError(f"nfield={nfield} outside [1, {self.max_fields}]")
        if nfield > 1 and stride < per_field:
            raise ValueError(f"stride {stride} < field size {per_field}")
        need = (nfield - 1) * stride + per_field
        if buf.count < need:
            raise ValueError(f"buffer '{buf.label}' holds {buf.count} floats, needs {need}")

    # ------------------------------------------------------ recording
    def encode_synth_into(self, stream: Stream, b: ShtBinding) -> None:
        """encodeSynthInto (sht.ts:326-338): Legendre then Fourier, on `stream`."""
        _lib.call("tsphere_synth_batch", self._h, b.nfield, b.src.ptr, b.src_stride, b.dst.ptr,
                  b.dst_stride, stream.handle)

    def encode_analys_into(self, stream: Stream, b: ShtBinding) -> None:
        """encodeAnalysInto (sht.ts:341-353): Fourier then Legendre."""
        _lib.call("tsphere_analys_batch", self._h, b.nfield, b.src.ptr, b.src_stride, b.dst.ptr,
                  b.dst_stride, stream.handle)

    def encode_stage(self, stream: Stream, stage: str) -> None:
        """encodeStage (sht.ts:368-401): one stage alone, on the plan's own
        buffers (qlm_in -> fm_buf -> spat_buf -> fm_buf -> qlm_out)."""
        h, s = self._h, stream.handle
        if stage == "legSynth":
            _lib.call("tsphere_leg_synth", h, 1, self.qlm_in.ptr, 0, self.fm_buf.ptr, s)
        elif stage == "fourSynth":
            _lib.call("tsphere_fourier_synth", h, 1, self.fm_buf.ptr, self.spat_buf.ptr, 0, s)
        elif stage == "fourAnalys":
            _lib.call("tsphere_fourier_analys", h, 1, self.spat_buf.ptr, 0, self.fm_buf.ptr, s)
        elif stage == "legAnalys":
            _lib.call("tsphere_leg_analys", h, 1, self.fm_buf.ptr, self.qlm_out.ptr, 0, s)
        else:
            raise ValueError(f"unknown stage {stage!r}")

    # ------------------------------------------------ host conveniences
    def synth(self, qlm) -> np.ndarray:
        """sht.ts:414-426: spectral (2*nlm) -> spatial (nlat*nphi), host arrays."""
        q = np.asarray(qlm, np.float32).reshape(-1)
        if q.size != 2 * self.nlm:
            raise ValueError(f"qlm must have length {2 * self.nlm}")
        self.qlm_in.upload(q, self.stream)
        self.encode_synth_into(self.stream, self.create_synth_binding(self.qlm_in, self.spat_buf))
        return self.spat_buf.download(self.stream)

    def synth_from(self, qlm_src: DeviceBuffer) -> np.ndarray:
        """sht.ts:435-446: synthesis of coefficients already on the device."""
        self.qlm_in.copy_from(qlm_src, self.stream)
        self.encode_synth_into(self.stream, self.create_synth_binding(self.qlm_in, self.spat_buf))
        return self.spat_buf.download(self.stream)

    def analys(self, spat) -> np.ndarray:
        """sht.ts:449-461: spatial -> spectral, host arrays."""
        a = np.asarray(spat, np.float32).reshape(-1)
        if a.size != self.cfg.npts:
            raise ValueError(f"spat must have length {self.cfg.npts}")
        self.spat_buf.upload(a, self.stream)
        self.encode_analys_into(self.stream, self.create_analys_binding(self.spat_buf, self.qlm_out))
        return self.qlm_out.download(self.stream)

    def read_fm(self) -> np.ndarray:
        """The Fourier-space intermediate of the last staged transform, in the
        reference's fmBuf order (complex [m][ilat], interleaved re/im).  On
        the device the scratch is row-major per latitude ([ilat][m]), so the
        FFT reads and writes contiguous rows; this transposes it back."""
        raw = self.fm_buf.download(self.stream)
        m1, nlat = self.cfg.mmax + 1, self.cfg.nlat
        return np.ascontiguousarray(raw.reshape(nlat, m1, 2).transpose(1, 0, 2)).reshape(-1)

    def destroy(self) -> None:
        """sht.ts:463-468"""
        if self._h:
            self.stream.sync()
            _lib.call("tsphere_plan_destroy", self._h)
            self._h = None
            for b in (self.qlm_in, self.qlm_out, self.spat_buf, self.fm_buf):
                b.free()
            self.stream.destroy()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
