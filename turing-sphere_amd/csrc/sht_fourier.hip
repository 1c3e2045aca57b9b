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
// MI355X mapping: one 64-lane wave owns one latitude row and runs a
// natural-order radix-4 Stockham FFT (plus one radix-2 stage when log2 H is
// odd) in its own LDS slice; a workgroup of 4 waves covers 4 consecutive
// latitudes, so the strided Fourier-coefficient accesses fm[m][ilat] touch
// 32 contiguous bytes per m, and the spatial row is read/written as float2.
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

// In-LDS Stockham FFT of H = 2^LOG2H points owned by one wave (natural order
// in and out, unnormalised: X_k = sum_j x_j e^{SIGN 2 pi i j k / H}).  The
// trig table has N = 2H entries, so an angle 2 pi a / (R Ns) is table index
// a * N / (R Ns).  All 4 waves of a block run the same stage sequence, so the
// block barriers line up.  Loops are unrolled at compile time so every
// butterfly operand stays in registers.
template <int SIGN, int LOG2H>
__device__ __forceinline__ void fft_wave(float2* buf, const float2* __restrict__ trig, int lane) {
  constexpr int H = 1 << LOG2H;
  constexpr int N = 2 * H;
  int Ns = 1;
  if constexpr (LOG2H & 1) {  // one radix-2 stage first (Ns = 1: no twiddle)
    constexpr int HALF = H / 2;
    constexpr int NB = (HALF + 63) / 64;
    float2 u[NB], v[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = lane + 64 * b;
      if (j < HALF) {
        u[b] = buf[j];
        v[b] = buf[j + HALF];
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = lane + 64 * b;
      if (j < HALF) {
        buf[2 * j] = cadd(u[b], v[b]);
        buf[2 * j + 1] = csub(u[b], v[b]);
      }
    }
    __syncthreads();
    Ns = 2;
  }
  constexpr int Q = H / 4;
  constexpr int NB = (Q + 63) / 64;
#pragma unroll 1
  for (; Ns < H; Ns <<= 2) {
    const int step = N / (4 * Ns);  // table stride for angle 2 pi jm r / (4 Ns)
    float2 x0[NB], x1[NB], x2[NB], x3[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = lane + 64 * b;
      if (j < Q) {
        const int jm = j & (Ns - 1);
        x0[b] = buf[j];
        x1[b] = cmul(buf[j + Q], tw<SIGN>(trig, jm * step));
        x2[b] = cmul(buf[j + 2 * Q], tw<SIGN>(trig, 2 * jm * step));
        x3[b] = cmul(buf[j + 3 * Q], tw<SIGN>(trig, 3 * jm * step));
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = lane + 64 * b;
      if (j < Q) {
        const int jm = j & (Ns - 1);
        const float2 a = cadd(x0[b], x2[b]);
        const float2 bb = csub(x0[b], x2[b]);
        const float2 c = cadd(x1[b], x3[b]);
        const float2 e = csub(x1[b], x3[b]);
        // d = SIGN * i * e
        const float2 d = SIGN > 0 ? make_float2(-e.y, e.x) : make_float2(e.y, -e.x);
        const int o = (j - jm) * 4 + jm;
        buf[o] = cadd(a, c);
        buf[o + Ns] = cadd(bb, d);
        buf[o + 2 * Ns] = csub(a, c);
        buf[o + 3 * Ns] = csub(bb, d);
      }
    }
    __syncthreads();
  }
}

// fourier.ts:309-350 (synthesis).  Block: 4 waves = 4 latitude rows.
template <int LOG2H>
__global__ __launch_bounds__(256) void k_fft_synth(int mmax, int nlat,
                                                   const float2* __restrict__ trig,
                                                   const float2* __restrict__ fm,
                                                   Fields Fs) {
  constexpr int H = 1 << LOG2H;
  constexpr int nphi = 2 * H;
  __shared__ float2 lds[4 * H];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ilat = blockIdx.x * 4 + wv;
  const int f = blockIdx.y;
  const bool row = ilat < nlat;
  const int il = row ? ilat : nlat - 1;
  float2* buf = lds + wv * H;
  const float2* F = fm + (long long)f * (mmax + 1) * nlat;
  for (int k = lane; k < H; k += 64) {
    float2 xk, xh;
    if (k == 0) xk = make_float2(F[il].x, 0.f);
    else if (k <= mmax) xk = F[(long long)k * nlat + il];
    else xk = make_float2(0.f, 0.f);
    const int kh = H - k;  // k = 0 -> H, above mmax
    if (kh <= mmax) xh = F[(long long)kh * nlat + il];  // kh >= 1 here
    else xh = make_float2(0.f, 0.f);
    const float2 cj = make_float2(xh.x, -xh.y);
    const float2 b = cmul(csub(xk, cj), trig[k]);
    const float2 s = cadd(xk, cj);
    buf[k] = make_float2(s.x - b.y, s.y + b.x);
  }
  __syncthreads();
  fft_wave<+1, LOG2H>(buf, trig, lane);
  if (row) {
    float2* out = reinterpret_cast<float2*>(const_cast<char*>(Fs.at(f))) + (long long)ilat * H;
    for (int m = lane; m < H; m += 64) out[m] = buf[m];
  }
}

// fourier.ts:352-386 (analysis).
template <int LOG2H>
__global__ __launch_bounds__(256) void k_fft_analys(int mmax, int nlat,
                                                    const float2* __restrict__ trig,
                                                    Fields Fs, float2* __restrict__ fm) {
  constexpr int H = 1 << LOG2H;
  constexpr int nphi = 2 * H;
  __shared__ float2 lds[4 * H];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ilat = blockIdx.x * 4 + wv;
  const int f = blockIdx.y;
  const bool row = ilat < nlat;
  const int il = row ? ilat : nlat - 1;
  float2* buf = lds + wv * H;
  const float2* in = reinterpret_cast<const float2*>(Fs.at(f)) + (long long)il * H;
  for (int m = lane; m < H; m += 64) buf[m] = in[m];
  __syncthreads();
  fft_wave<-1, LOG2H>(buf, trig, lane);
  if (row) {
    float2* G = fm + (long long)f * (mmax + 1) * nlat;
    for (int k = lane; k <= mmax; k += 64) {
      const float2 zk = buf[k];
      const float2 zh = buf[(H - k) & (H - 1)];
      const float2 cj = make_float2(zh.x, -zh.y);
      const float2 xe = make_float2(0.5f * (zk.x + cj.x), 0.5f * (zk.y + cj.y));
      const float2 d = make_float2(0.5f * (zk.x - cj.x), 0.5f * (zk.y - cj.y));
      const float2 xo = make_float2(d.y, -d.x);
      const float2 w = make_float2(trig[k].x, -trig[k].y);
      G[(long long)k * nlat + ilat] = cadd(xe, cmul(xo, w));
    }
  }
}

// fourier.ts:241-264: v = Re F_0 + 2 sum_m Re(F_m e^{i m phi}).
__global__ __launch_bounds__(256) void k_dft_synth(int mmax, int nlat, int nphi,
                                                   const float2* __restrict__ trig,
                                                   const float2* __restrict__ fm,
                                                   Fields Fs) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  const int ilat = blockIdx.y;
  const int f = blockIdx.z;
  if (j >= nphi) return;
  const float2* F = fm + (long long)f * (mmax + 1) * nlat;
  float v = F[ilat].x;
  int idx = 0;
  for (int m = 1; m <= mmax; ++m) {
    idx += j;
    if (idx >= nphi) idx -= nphi;
    const float2 t = trig[idx];
    const float2 c = F[(long long)m * nlat + ilat];
    v += 2.0f * (c.x * t.x - c.y * t.y);
  }
  reinterpret_cast<float*>(const_cast<char*>(Fs.at(f)))[(long long)ilat * nphi + j] = v;
}

// fourier.ts:266-289: G_m = sum_j f_j e^{-i m phi_j}.
__global__ __launch_bounds__(256) void k_dft_analys(int mmax, int nlat, int nphi,
                                                    const float2* __restrict__ trig,
                                                    Fields Fs, float2* __restrict__ fm) {
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
  fm[(long long)f * (mmax + 1) * nlat + (long long)m * nlat + ilat] = make_float2(re, im);
}

__global__ void k_fill(float* __restrict__ dst, float v, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = v;
}

int ilog2(int n) {
  int r = 0;
  while ((1 << r) < n) ++r;
  return r;
}

}  // namespace

hipError_t launch_fourier_synth(int fourier, int mmax, int nlat, int nphi, const float2* trig,
                                int nfield, const float2* fm, const Fields& spat, hipStream_t s) {
  if (nfield <= 0) return hipSuccess;
  if (fourier == TSPHERE_FOURIER_FFT) {
    dim3 grid((nlat + 3) / 4, nfield);
    switch (ilog2(nphi / 2)) {
#define TS_CASE(L)                                                                          \
  case L:                                                                                   \
    hipLaunchKernelGGL(k_fft_synth<L>, grid, dim3(256), 0, s, mmax, nlat, trig, fm, spat);                \
                                                                                            \
    break;
      TS_CASE(2) TS_CASE(3) TS_CASE(4) TS_CASE(5) TS_CASE(6) TS_CASE(7) TS_CASE(8) TS_CASE(9)
      TS_CASE(10) TS_CASE(11)
#undef TS_CASE
      default: return hipErrorInvalidValue;
    }
  } else {
    dim3 grid((nphi + 255) / 256, nlat, nfield);
    hipLaunchKernelGGL(k_dft_synth, grid, dim3(256), 0, s, mmax, nlat, nphi, trig, fm, spat);
  }
  return hipGetLastError();
}

hipError_t launch_fourier_analys(int fourier, int mmax, int nlat, int nphi, const float2* trig,
                                 int nfield, const Fields& spat, float2* fm, hipStream_t s) {
  if (nfield <= 0) return hipSuccess;
  if (fourier == TSPHERE_FOURIER_FFT) {
    dim3 grid((nlat + 3) / 4, nfield);
    switch (ilog2(nphi / 2)) {
#define TS_CASE(L)                                                                           \
  case L:                                                                                    \
    hipLaunchKernelGGL(k_fft_analys<L>, grid, dim3(256), 0, s, mmax, nlat, trig, spat, fm);\
                                                                                             \
    break;
      TS_CASE(2) TS_CASE(3) TS_CASE(4) TS_CASE(5) TS_CASE(6) TS_CASE(7) TS_CASE(8) TS_CASE(9)
      TS_CASE(10) TS_CASE(11)
#undef TS_CASE
      default: return hipErrorInvalidValue;
    }
  } else {
    dim3 grid((mmax + 1 + 255) / 256, nlat, nfield);
    hipLaunchKernelGGL(k_dft_analys, grid, dim3(256), 0, s, mmax, nlat, nphi, trig, spat, fm);
  }
  return hipGetLastError();
}

hipError_t launch_fill(float* dst, float v, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, v, n);
  return hipGetLastError();
}

}  // namespace tsphere
