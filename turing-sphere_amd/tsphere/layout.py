"""Grid and spectral layout (src/sht/layout.ts), host side.

Q_lm complex for m >= 0, interleaved [re, im], m-major (for m: for l >= m);
spatial fields float32 [ilat * nphi + iphi], ilat north -> south.
"""
from __future__ import annotations

import math
from dataclasses import dataclass


@dataclass(frozen=True)
class ShtConfig:
    """ShtConfig (layout.ts:17-22)."""
    lmax: int
    mmax: int
    nlat: int
    nphi: int

    @property
    def nlm(self) -> int:
        return nlm_calc(self.lmax, self.mmax)

    @property
    def npts(self) -> int:
        return self.nlat * self.nphi


def nlm_calc(lmax: int, mmax: int) -> int:
    """layout.ts:24-27"""
    return (mmax + 1) * (lmax + 1) - (mmax * (mmax + 1)) // 2


def lm_index(lmax: int, l: int, m: int) -> int:
    """layout.ts:30-32"""
    return m * (lmax + 1) - (m * (m - 1)) // 2 + (l - m)


def validate_config(cfg: ShtConfig) -> None:
    """layout.ts:34-43 (same messages)."""
    lmax, mmax, nlat, nphi = cfg.lmax, cfg.mmax, cfg.nlat, cfg.nphi
    if not isinstance(lmax, int) or lmax < 1:
        raise ValueError(f"lmax must be an integer >= 1 (got {lmax})")
    if not isinstance(mmax, int) or mmax < 0 or mmax > lmax:
        raise ValueError(f"mmax must be an integer in [0, lmax] (got {mmax})")
    if not isinstance(nlat, int) or nlat <= lmax:
        raise ValueError(
            f"nlat must be an integer > lmax for exact Gauss quadrature (got nlat={nlat}, lmax={lmax})")
    if not isinstance(nphi, int) or nphi < 2 * mmax + 1:
        raise ValueError(
            f"nphi must be an integer >= 2*mmax+1 to avoid aliasing (got nphi={nphi}, mmax={mmax})")


def is_power_of_two(n: int) -> bool:
    """layout.ts:45-47"""
    return n > 0 and (n & (n - 1)) == 0


def grid_for_lmax(lmax: int, pdeg: int) -> tuple[int, int]:
    """layout.ts:53-59: dealiased grid for a reaction of polynomial degree pdeg."""
    min_lat = max(lmax + 1, ((pdeg + 1) * lmax + 1) / 2)
    nlat = 2 * math.ceil(min_lat / 2)
    nphi = 1
    while nphi < (pdeg + 1) * lmax + 1:
        nphi *= 2
    return nlat, nphi
