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
// rows; every thread owns one butterfly per pass, so long rows are spread over
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

namespace tsphere {
namespace {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Twiddle e^{SIGN 2 pi i idx / N} from the (cos, sin) table.
template <int SIGN>
__device__ __forceinline__ float2 tw(const float2* __restrict__ trig, int idx) {
  const float2 t = trig[idx];
  return SIGN > 0 ? t : make_float2(t.x, -t.y);
}

// Launch shape of the H = 2^LOG2H point row FFT.  Every thread owns one
// radix-4 butterfly per pass (two for H = 2048), so a row gets H/4 threads
// (up to 256) and a 256-thread workgroup carries RB = 256/TPR rows.  C2's
// 2048-point rows therefore run 256 threads each: with two fields that is
// 2048 workgroups, 8 waves per SIMD, instead of one wave per row.
template <int LOG2H>
struct FftShape {
  static constexpr int H = 1 << LOG2H;
  static constexpr int N = 2 * H;  // entries of the trig table (real length)
  static constexpr int Q = H / 4;
  static constexpr int TPR = Q < 256 ? Q : 256;
  static constexpr int BPT = Q / TPR;
  static constexpr int RB = 256 / TPR;
  static constexpr bool ODD = LOG2H & 1;
};

// One radix-4 Stockham pass, natural order, from src to dst (both the
// thread's row in LDS).  Butterfly j reads src[j + k Q], k = 0..3, and writes
// dst[4(j - jm) + jm + k Ns], jm = j mod Ns, with twiddles
// e^{SIGN 2 pi i jm k / (4 Ns)} = trig[jm k N / (4 Ns)].
template <int SIGN, int LOG2H>
__device__ __forceinline__ void r4_pass(const float2* src, float2* dst,
                                        const float2* __restrict__ trig, int t, int Ns) {
  using S = FftShape<LOG2H>;
  const int step = S::N / (4 * Ns);
#pragma unroll
  for (int b = 0; b < S::BPT; ++b) {
    const int j = t + S::TPR * b;
    const int jm = j & (Ns - 1);
    const float2 x0 = src[j];
    const float2 x1 = cmul(src[j + S::Q], tw<SIGN>(trig, jm * step));
    const float2 x2 = cmul(src[j + 2 * S::Q], tw<SIGN>(trig, 2 * jm * step));
    const float2 x3 = cmul(src[j + 3 * S::Q], tw<SIGN>(trig, 3 * jm * step));
    const float2 a = cadd(x0, x2), bb = csub(x0, x2);
    const float2 c = cadd(x1, x3), e = csub(x1, x3);
    const float2 d = SIGN > 0 ? make_float2(-e.y, e.x) : make_float2(e.y, -e.x);
    const int o = 4 * (j - jm) + jm;
    dst[o] = cadd(a, c);
    dst[o + Ns] = cadd(bb, d);
    dst[o + 2 * Ns] = csub(a, c);
    dst[o + 3 * Ns] = csub(bb, d);
  }
}

// The first pass reads its inputs through `in(idx)` (global memory, with the
// real-FFT pre-processing folded in) instead of LDS: radix-2 when log2 H is
// odd (Ns = 1: out[2j] = u + v, out[2j+1] = u - v), else radix-4 without
// twiddles.
template <int SIGN, int LOG2H, class In>
__device__ __forceinline__ void first_pass(In in, float2* dst, int t) {
  using S = FftShape<LOG2H>;
  if constexpr (S::ODD) {
#pragma unroll
    for (int b = 0; b < 2 * S::BPT; ++b) {
      const int j = t + S::TPR * b;
      const float2 u = in(j), v = in(j + S::H / 2);
      dst[2 * j] = cadd(u, v);
      dst[2 * j + 1] = csub(u, v);
    }
  } else {
#pragma unroll
    for (int b = 0; b < S::BPT; ++b) {
      const int j = t + S::TPR * b;
      const float2 x0 = in(j), x1 = in(j + S::Q), x2 = in(j + 2 * S::Q), x3 = in(j + 3 * S::Q);
      const float2 a = cadd(x0, x2), bb = csub(x0, x2);
      const float2 c = cadd(x1, x3), e = csub(x1, x3);
      const float2 d = SIGN > 0 ? make_float2(-e.y, e.x) : make_float2(e.y, -e.x);
      dst[4 * j] = cadd(a, c);
      dst[4 * j + 1] = cadd(bb, d);
      dst[4 * j + 2] = csub(a, c);
      dst[4 * j + 3] = csub(bb, d);
    }
  }
}

// Remaining passes, ping-ponging between the two LDS rows (one barrier per
// pass: a pass only writes the buffer the previous pass read).  Returns the
// buffer that holds the result.
template <int SIGN, int LOG2H>
__device__ __forceinline__ float2* rest_passes(float2* a, float2* b,
                                               const float2* __restrict__ trig, int t) {
  using S = FftShape<LOG2H>;
#pragma unroll
  for (int Ns = S::ODD ? 2 : 4; Ns < S::H; Ns <<= 2) {
    __syncthreads();
    r4_pass<SIGN, LOG2H>(a, b, trig, t, Ns);
    float2* tmp = a;
    a = b;
    b = tmp;
  }
  __syncthreads();
  return a;
}

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
  const int ilat = tile * S::RB + r;
  const bool row = ilat < nlat;
  const float2* F = fm + ((long long)f * nlat + (row ? ilat : nlat - 1)) * (mmax + 1);
  auto X = [&](int k) -> float2 {
    if (k > mmax) return make_float2(0.f, 0.f);
    const float2 v = F[k];
    return k == 0 ? v : make_float2(sc.mscale * v.x, sc.mscale * v.y);
  };
  auto Z = [&](int k) -> float2 {
    float2 xk = X(k);
    if (k == 0) xk.y = 0.f;
    const float2 xh = k == 0 ? make_float2(0.f, 0.f) : X(S::H - k);
    const float2 cj = make_float2(xh.x, -xh.y);
    const float2 bq = cmul(csub(xk, cj), trig[k]);
    const float2 s = cadd(xk, cj);
    return make_float2(s.x - bq.y, s.y + bq.x);
  };
  float2* a = lds[0] + r * S::H;
  float2* b = lds[1] + r * S::H;
  first_pass<+1, LOG2H>(Z, a, t);
  float2* res = rest_passes<+1, LOG2H>(a, b, trig, t);
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
  const int ilat = tile * S::RB + r;
  const bool row = ilat < nlat;
  const float2* in =
      reinterpret_cast<const float2*>(Fs.at(f)) + (long long)(row ? ilat : nlat - 1) * S::H;
  float2* a = lds[0] + r * S::H;
  float2* b = lds[1] + r * S::H;
  first_pass<-1, LOG2H>([&](int j) { return in[j]; }, a, t);
  const float2* z = rest_passes<-1, LOG2H>(a, b, trig, t);
  if (row) {
    float2* G = fm + ((long long)f * nlat + ilat) * (mmax + 1);
    for (int k = t; k <= mmax; k += S::TPR) {
      const float2 zk = z[k];
      const float2 zh = z[(S::H - k) & (S::H - 1)];
      const float2 cj = make_float2(zh.x, -zh.y);
      const float2 xe = make_float2(0.5f * (zk.x + cj.x), 0.5f * (zk.y + cj.y));
      const float2 d = make_float2(0.5f * (zk.x - cj.x), 0.5f * (zk.y - cj.y));
      const float2 xo = make_float2(d.y, -d.x);
      const float2 w = make_float2(trig[k].x, -trig[k].y);
      const float2 g = cadd(xe, cmul(xo, w));
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
