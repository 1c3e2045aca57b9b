// Fourier stage of the transforms: real FFT along longitude, one latitude
// row at a time (src/sht/wgsl/fourier.ts:291-386).
//
// A real row of N = nphi points is an H = N/2-point complex FFT wrapped in
// the real-FFT recombination (fourier.ts:291-308):
//   synthesis  Z[k] = (X[k] + conj X[H-k]) + i e^{+2pi i k/N} (X[k] - conj X[H-k]),
//              z = FFT_H^{+}(Z), x[2m] = Re z[m], x[2m+1] = Im z[m]
//   analysis   z[m] = x[2m] + i x[2m+1], Z = FFT_H^{-}(z),
//              X[k] = (Z[k] + conj Z[H-k])/2 + e^{-2pi i k/N} (-i)(Z[k] - conj Z[H-k])/2
// with X[0] = Re F_0, X[m] = F_m (m <= mmax), 0 above, and no 2pi/nphi
// factor (it lives in the Gauss weight, sht.ts:198).  Twiddles come from the
// host's f64-built table trig[k] = (cos, sin)(2 pi k / N).
//
// MI355X mapping: a natural-order radix-4 Stockham FFT (plus one radix-2
// pass when log2 H is odd) per latitude row, ping-ponging between two LDS
// rows (the building blocks are in fft_row.h, which the planner's generated
// grid kernels share); every thread owns one butterfly per pass, so long rows are spread over
// up to 256 threads and short rows are packed several to a workgroup.  The
// first pass reads global memory directly (with the real-FFT pre-processing
// folded in), the last one is followed by the post-processing that writes
// global memory.  The Fourier scratch is row-major per latitude,
// fm[f][ilat][m], so both sides of the stage move whole contiguous rows; the
// Legendre kernels, which are compute-bound, take the strided side.
//
// Both directions take a FourierScale: a factor on every m >= 1 coefficient
// (the spectrum read by synthesis, the coefficients written by analysis) and
// an optional per-latitude factor on the synthesised rows.  Identity for the
// transforms themselves; the transposes used by reverse-mode gradients are
// synthesis with m >= 1 halved and rows times the quadrature weight, and
// analysis with m >= 1 doubled (then a unit-weight Legendre analysis).
//
// Non-power-of-two nphi uses the direct band-limited DFT (fourier.ts:241-289),
// one thread per output, as the reference's 'dft' fallback does.
#include "tsphere_internal.h"
#include "fft_row.h"

namespace tsphere {
namespace {

using tsfft::FftShape;

// fourier.ts:309-350 (synthesis).  Z[k] is built from the Hermitian spectrum
// X[0] = Re F_0, X[k] = F_k (k <= mmax), 0 above (mmax <= H - 1 since
// nphi >= 2 mmax + 1), read straight from the row fm[f][ilat][0..mmax].
template <int LOG2H>
__global__ __launch_bounds__(256) void k_fft_synth(int mmax, int nlat, int tiles,
                                                   const float2* __restrict__ trig,
                                                   const float2* __restrict__ fm,
                                                   Fields Fs, FourierScale sc) {
  using S = FftShape<LOG2H>;
  __shared__ float2 lds[2][S::RB * S::H];
  const int L = blockIdx.x;
  const int f = L / tiles, tile = L - f * tiles;
  const int r = threadIdx.x / S::TPR, t = threadIdx.x % S::TPR;
  // the pass twiddle tables staged once per workgroup in LDS: these kernels
  // stream many rows, and per-thread copies would triple the bytes each
  // wave reads through its memory pipeline (64 fields of 512 x 1024:
  // synthesis 47 us with the LDS copy, 61 us with per-thread registers)
  using TW = tsfft::PassTwiddlesLds<LOG2H, 256>;
  __shared__ float2 Wl[TW::NTW > 0 ? TW::NTW : 1];
  TW tw;
  tw.W = Wl;
  tw.load(trig);
  const int ilat = tile * S::RB + r;
  const bool row = ilat < nlat;
  const float2* F = fm + ((long long)f * nlat + (row ? ilat : nlat - 1)) * (mmax + 1);
  float2* a = lds[0] + r * S::H;
  float2* b = lds[1] + r * S::H;
  tsfft::first_pass<+1, LOG2H>(tsfft::SynthInput<LOG2H>{F, trig, mmax, sc.mscale}, a, t);
  tw.publish();
  const float2* res = tsfft::rest_passes<+1, LOG2H>(a, b, tw, t);
  if (row) {
    float2* out = reinterpret_cast<float2*>(const_cast<char*>(Fs.at(f))) + (long long)ilat * S::H;
    const float w = sc.rowscale ? sc.rowscale[ilat] : 1.f;
#pragma unroll
    for (int b2 = 0; b2 < 4 * S::BPT; ++b2) {
      const float2 v = res[t + S::TPR * b2];
      out[t + S::TPR * b2] = sc.rowscale ? make_float2(w * v.x, w * v.y) : v;
    }
  }
}

// fourier.ts:352-386 (analysis).
template <int LOG2H>
__global__ __launch_bounds__(256) void k_fft_analys(int mmax, int nlat, int tiles,
                                                    const float2* __restrict__ trig,
                                                    Fields Fs, float2* __restrict__ fm,
                                                    FourierScale sc) {
  using S = FftShape<LOG2H>;
  __shared__ float2 lds[2][S::RB * S::H];
  const int L = blockIdx.x;
  const int f = L / tiles, tile = L - f * tiles;
  const int r = threadIdx.x / S::TPR, t = threadIdx.x % S::TPR;
  using TW = tsfft::PassTwiddlesLds<LOG2H, 256>;  // as in k_fft_synth
  __shared__ float2 Wl[TW::NTW > 0 ? TW::NTW : 1];
  TW tw;
  tw.W = Wl;
  tw.load(trig);
  const int ilat = tile * S::RB + r;
  const bool row = ilat < nlat;
  const float2* in =
      reinterpret_cast<const float2*>(Fs.at(f)) + (long long)(row ? ilat : nlat - 1) * S::H;
  float2* a = lds[0] + r * S::H;
  float2* b = lds[1] + r * S::H;
  tsfft::first_pass<-1, LOG2H>([&](int j) { return in[j]; }, a, t);
  tw.publish();
  const float2* z = tsfft::rest_passes<-1, LOG2H>(a, b, tw, t);
  if (row) {
    float2* G = fm + ((long long)f * nlat + ilat) * (mmax + 1);
    for (int k = t; k <= mmax; k += S::TPR) {
      const float2 g = tsfft::analys_coef<LOG2H>(z, k, trig);
      G[k] = k == 0 ? g : make_float2(sc.mscale * g.x, sc.mscale * g.y);
    }
  }
}

// fourier.ts:241-264: v = Re F_0 + 2 sum_m Re(F_m e^{i m phi}).
__global__ __launch_bounds__(256) void k_dft_synth(int mmax, int nlat, int nphi,
                                                   const float2* __restrict__ trig,
                                                   const float2* __restrict__ fm,
                                                   Fields Fs, FourierScale sc) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int ilat = blockIdx.y;
  const int f = blockIdx.z;
  if (j >= nphi) return;
  const float2* F = fm + ((long long)f * nlat + ilat) * (mmax + 1);
  float v = F[0].x;
  int idx = 0;
  for (int m = 1; m <= mmax; ++m) {
    idx += j;
    if (idx >= nphi) idx -= nphi;
    const float2 t = trig[idx];
    const float2 c = F[m];
    v += (2.0f * sc.mscale) * (c.x * t.x - c.y * t.y);
  }
  if (sc.rowscale) v *= sc.rowscale[ilat];
  reinterpret_cast<float*>(const_cast<char*>(Fs.at(f)))[(long long)ilat * nphi + j] = v;
}

// fourier.ts:266-289: G_m = sum_j f_j e^{-i m phi_j}.
__global__ __launch_bounds__(256) void k_dft_analys(int mmax, int nlat, int nphi,
                                                    const float2* __restrict__ trig,
                                                    Fields Fs, float2* __restrict__ fm,
                                                    FourierScale sc) {
  const int m = blockIdx.x * 256 + threadIdx.x;
  const int ilat = blockIdx.y;
  const int f = blockIdx.z;
  if (m > mmax) return;
  const float* row = reinterpret_cast<const float*>(Fs.at(f)) + (long long)ilat * nphi;
  float re = 0.f, im = 0.f;
  int idx = 0;
  for (int j = 0; j < nphi; ++j) {
    const float2 t = trig[idx];
    re += row[j] * t.x;
    im -= row[j] * t.y;
    idx += m;
    if (idx >= nphi) idx -= nphi;
  }
  const float g = m == 0 ? 1.f : sc.mscale;
  fm[((long long)f * nlat + ilat) * (mmax + 1) + m] = make_float2(g * re, g * im);
}

__global__ void k_fill(float* __restrict__ dst, float v, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = v;
}

__global__ void k_accumulate(float* __restrict__ dst, const float* __restrict__ src, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] += src[i];
}

int ilog2(int n) {
  int r = 0;
  while ((1 << r) < n) ++r;
  return r;
}

}  // namespace

hipError_t launch_fourier_synth(int fourier, int mmax, int nlat, int nphi, const float2* trig,
                                int nfield, const float2* fm, const Fields& spat, hipStream_t s,
                                FourierScale sc) {
  if (nfield <= 0) return hipSuccess;
  if (fourier == TSPHERE_FOURIER_FFT) {
    switch (ilog2(nphi / 2)) {
#define TS_CASE(L)                                                                     \
  case L: {                                                                            \
    const int tiles = (nlat + FftShape<L>::RB - 1) / FftShape<L>::RB;                  \
    hipLaunchKernelGGL(k_fft_synth<L>, dim3(tiles * nfield), dim3(256), 0, s, mmax, nlat, \
                       tiles, trig, fm, spat, sc);                                         \
    break;                                                                             \
  }
      TS_CASE(2) TS_CASE(3) TS_CASE(4) TS_CASE(5) TS_CASE(6) TS_CASE(7) TS_CASE(8) TS_CASE(9)
      TS_CASE(10) TS_CASE(11)
#undef TS_CASE
      default: return hipErrorInvalidValue;
    }
  } else {
    dim3 grid((nphi + 255) / 256, nlat, nfield);
    hipLaunchKernelGGL(k_dft_synth, grid, dim3(256), 0, s, mmax, nlat, nphi, trig, fm, spat, sc);
  }
  return hipGetLastError();
}

hipError_t launch_fourier_analys(int fourier, int mmax, int nlat, int nphi, const float2* trig,
                                 int nfield, const Fields& spat, float2* fm, hipStream_t s,
                                 FourierScale sc) {
  if (nfield <= 0) return hipSuccess;
  if (fourier == TSPHERE_FOURIER_FFT) {
    switch (ilog2(nphi / 2)) {
#define TS_CASE(L)                                                                      \
  case L: {                                                                             \
    const int tiles = (nlat + FftShape<L>::RB - 1) / FftShape<L>::RB;                   \
    hipLaunchKernelGGL(k_fft_analys<L>, dim3(tiles * nfield), dim3(256), 0, s, mmax, nlat, \
                       tiles, trig, spat, fm, sc);                                          \
    break;                                                                              \
  }
      TS_CASE(2) TS_CASE(3) TS_CASE(4) TS_CASE(5) TS_CASE(6) TS_CASE(7) TS_CASE(8) TS_CASE(9)
      TS_CASE(10) TS_CASE(11)
#undef TS_CASE
      default: return hipErrorInvalidValue;
    }
  } else {
    dim3 grid((mmax + 1 + 255) / 256, nlat, nfield);
    hipLaunchKernelGGL(k_dft_analys, grid, dim3(256), 0, s, mmax, nlat, nphi, trig, spat, fm, sc);
  }
  return hipGetLastError();
}

hipError_t launch_accumulate(float* dst, const float* src, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_accumulate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, src,
                     n);
  return hipGetLastError();
}

hipError_t launch_fill(float* dst, float v, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, v, n);
  return hipGetLastError();
}

}  // namespace tsphere
