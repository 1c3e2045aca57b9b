// Legendre stage of the spherical-harmonic transforms, plain FP32 on the
// gfx950 VALU.
//
//   synthesis  F_m(theta_i) = sum_{l=m..lmax} Q_lm ytilde_l^m(theta_i)
//              (src/sht/wgsl/leg.ts:43-137, reference.ts:37-57)
//   analysis   Q_lm = sum_i w_i G_m(theta_i) ytilde_l^m(theta_i)
//              (src/sht/wgsl/leg.ts:139-352, reference.ts:103-116)
//
// The associated Legendre functions are generated on the fly by the 3-term
// recurrence over l with f32 coefficients (a_l^m, b_l^m) (coeffs.ts:28-56).
// Both kernels fold north/south latitude pairs onto one recurrence
// (ytilde_l^m(-x) = (-1)^(l-m) ytilde_l^m(x)); even (l-m) terms accumulate in
// E and odd ones in O, north = E + O and south = E - O (leg.ts:129-134,
// 207-219).  Values whose magnitude is still below ACCURACY carry a negative
// exponent ny and are not accumulated (leg.ts:140-157, common.ts); the seed
// ytilde_m^m comes from the host f64 table, already scaled.
//
// MI355X mapping:
//  * synthesis: one 64-lane wave per (m, 64 latitude pairs).  m, l and the
//    coefficient / spectral addresses are wave-uniform, so (a,b) and Q_lm are
//    scalar loads shared by the 64 lanes; a lane does 3 VALU ops of
//    recurrence + 2 FMAs per field per l.  Waves are ordered by m ascending,
//    i.e. longest l-loop first, which balances the triangle.
//  * analysis: one workgroup per m, up to 4 waves, K latitude pairs per lane
//    in registers.  Each lane accumulates a span of l-values (64 partial sums
//    per lane: span_l * fields * re/im) and the span is reduced across the
//    wave by a transposing butterfly -- v_permlane32_swap, v_permlane16_swap,
//    then DPP row_mirror / row_half_mirror / quad_perm -- which leaves lane i
//    holding the complete sum of partial i after 141 VALU ops for all 64
//    values (instead of 6 shuffles per value).  Waves combine through LDS,
//    one barrier per span.
#include "tsphere_internal.h"

namespace tsphere {
namespace {

__device__ __forceinline__ long long lm_base(int lmax, int m) {
  return (long long)m * (lmax + 1) - ((long long)m * (m - 1)) / 2;
}

// a * cos(theta) * y with cos(theta) = ct.x + ct.y (f32 hi + lo of the f64
// node), plus c: rounding the node to f32 perturbs theta and, through it,
// every ytilde_l by ~ l * 2^-24 / sin(theta) -- the dominant fp32 error of the
// plain recurrence.  Carrying the low half costs one FMA.
__device__ __forceinline__ float act_fma(float a, float y, float2 ct, float c) {
  const float u = a * y;
  return fmaf(u, ct.x, fmaf(u, ct.y, c));
}

// One step of the recurrence: (y_l, y_{l+1}) -> (y_{l+2}, y_{l+3}), written
// as ytilde_l = a_l (cos theta) ytilde_{l-1} + b_l ytilde_{l-2}.
__device__ __forceinline__ void advance(float2 a0, float2 a1, float2 ct, float& y0, float& y1) {
  const float t0 = act_fma(a0.x, y1, ct, a0.y * y0);
  y1 = act_fma(a1.x, t0, ct, a1.y * y1);
  y0 = t0;
}

__device__ __forceinline__ void cfma(float2& acc, float y, float2 q) {
  acc.x = fmaf(y, q.x, acc.x);
  acc.y = fmaf(y, q.y, acc.y);
}

// ------------------------------------------------------------- synthesis
template <int NF>
__global__ __launch_bounds__(256) void k_leg_synth(LegTables T, int nfield, int nchunk,
                                                   int nwork, Fields Fq,
                                                   float2* __restrict__ fm) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wave >= nwork) return;  // whole wave
  const int m = __builtin_amdgcn_readfirstlane(wave / nchunk);
  const int chunk = __builtin_amdgcn_readfirstlane(wave - m * nchunk);
  const int f0 = blockIdx.y * NF;
  const int lmax = T.lmax;
  const int j = chunk * 64 + lane;
  const bool active = j < T.npair;
  const int jc = active ? j : T.npair - 1;

  const float2 ct = T.ct[jc];
  const Seed sd = T.seed[(long long)m * T.npair + jc];
  const long long base = lm_base(lmax, m);
  const float2* __restrict__ ab = T.ab + base;  // ab[l - m]
  float y0 = sd.y;
  int ny = sd.ny;
  float y1 = (m < lmax) ? act_fma(ab[1].x, y0, ct, 0.0f) : 0.0f;

  const float2* qf[NF];
  float2 accE[NF], accO[NF];
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const int fld = min(f0 + f, nfield - 1);
    qf[f] = reinterpret_cast<const float2*>(Fq.at(fld)) + base;
    accE[f] = make_float2(0.f, 0.f);
    accO[f] = make_float2(0.f, 0.f);
  }

  int l = m;
  bool done = false;
  // Phase 1: some lane still carries a scaled (negligible) value: the
  // reference's per-lane rule, once per l-pair.
  while (__builtin_amdgcn_ballot_w64(ny != 0) != 0) {
    if (ny == 0) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        cfma(accE[f], y0, qf[f][l - m]);
        if (l + 1 <= lmax) cfma(accO[f], y1, qf[f][l + 1 - m]);
      }
    } else if (fabsf(y0) > kRescaleThr) {
      ny += 1;
      y0 *= kInvScale;
      y1 *= kInvScale;
    }
    if (l + 2 > lmax) {
      done = true;
      break;
    }
    const float2 a0 = ab[l + 2 - m];
    const float2 a1 = (l + 3 <= lmax) ? ab[l + 3 - m] : make_float2(0.f, 0.f);
    advance(a0, a1, ct, y0, y1);
    l += 2;
  }
  if (!done) {
    // Phase 2: every lane significant -- no per-lane tests.
#pragma unroll 2
    for (; l + 3 <= lmax; l += 2) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        cfma(accE[f], y0, qf[f][l - m]);
        cfma(accO[f], y1, qf[f][l + 1 - m]);
      }
      advance(ab[l + 2 - m], ab[l + 3 - m], ct, y0, y1);
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      if (l <= lmax) cfma(accE[f], y0, qf[f][l - m]);
      if (l + 1 <= lmax) cfma(accO[f], y1, qf[f][l + 1 - m]);
    }
    if (l + 2 <= lmax) {
      const float2 a0 = ab[l + 2 - m];
      const float t0 = act_fma(a0.x, y1, ct, a0.y * y0);
#pragma unroll
      for (int f = 0; f < NF; ++f) cfma(accE[f], t0, qf[f][l + 2 - m]);
    }
  }

  if (active) {
    const int nlat = T.nlat;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      if (f0 + f < nfield) {
        float2* out = fm + ((long long)(f0 + f) * (T.mmax + 1) + m) * nlat;
        out[j] = make_float2(accE[f].x + accO[f].x, accE[f].y + accO[f].y);
        out[nlat - 1 - j] = make_float2(accE[f].x - accO[f].x, accE[f].y - accO[f].y);
      }
    }
  }
}

// ------------------------------------------------ transposing wave reduction
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, true));
}

template <int CTRL, int N>
__device__ __forceinline__ void dpp_stage(float (&v)[64], bool hi) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const float keep = hi ? v[i + N] : v[i];
    const float send = hi ? v[i] : v[i + N];
    v[i] = keep + dpp<CTRL>(send);
  }
}

// 64 partial values per lane -> lane i holds the sum over the wave of v[i].
__device__ __forceinline__ float transpose_reduce64(float (&v)[64]) {
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[i]),
                                                    __float_as_uint(v[i + 32]), false, false);
    v[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[i]),
                                                    __float_as_uint(v[i + 16]), false, false);
    v[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
  }
  const int lane = threadIdx.x & 63;
  dpp_stage<0x140, 8>(v, lane & 8);  // row_mirror      (lane ^ 15)
  dpp_stage<0x141, 4>(v, lane & 4);  // row_half_mirror (lane ^ 7)
  dpp_stage<0x1B, 2>(v, lane & 2);   // quad_perm 3210  (lane ^ 3)
  dpp_stage<0xB1, 1>(v, lane & 1);   // quad_perm 1032  (lane ^ 1)
  return v[0];
}

// -------------------------------------------------------------- analysis
template <int NF, int K, bool CHECKED>
__device__ __forceinline__ void analys_span(float (&acc)[64], const float2* __restrict__ ab,
                                            int ls, int m, int lmax, const float2 (&ct)[K],
                                            float (&y0)[K], float (&y1)[K], int (&ny)[K],
                                            const float2 (&wp)[K][NF],
                                            const float2 (&wm)[K][NF]) {
  constexpr int SPAN = 32 / NF;
#pragma unroll
  for (int p = 0; p < SPAN; p += 2) {
    const int l = ls + p;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (!CHECKED || ny[k] == 0) {
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          float* e = &acc[(p * NF + f) * 2];
          float* o = &acc[((p + 1) * NF + f) * 2];
          e[0] = fmaf(y0[k], wp[k][f].x, e[0]);
          e[1] = fmaf(y0[k], wp[k][f].y, e[1]);
          o[0] = fmaf(y1[k], wm[k][f].x, o[0]);
          o[1] = fmaf(y1[k], wm[k][f].y, o[1]);
        }
      } else if (fabsf(y0[k]) > kRescaleThr) {
        ny[k] += 1;
        y0[k] *= kInvScale;
        y1[k] *= kInvScale;
      }
    }
    const float2 a0 = (l + 2 <= lmax) ? ab[l + 2 - m] : make_float2(0.f, 0.f);
    const float2 a1 = (l + 3 <= lmax) ? ab[l + 3 - m] : make_float2(0.f, 0.f);
#pragma unroll
    for (int k = 0; k < K; ++k) advance(a0, a1, ct[k], y0[k], y1[k]);
  }
}

template <int NF, int K>
__global__ __launch_bounds__(256) void k_leg_analys(LegTables T, int nfield, int nwaves,
                                                    const float2* __restrict__ fm,
                                                    Fields Fq) {
  constexpr int SPAN = 32 / NF;
  static_assert(SPAN % 2 == 0 && SPAN * NF * 2 == 64, "64 partials per lane");
  __shared__ float red[2][4][64];
  const int m = blockIdx.x;
  const int f0 = blockIdx.y * NF;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lmax = T.lmax, nlat = T.nlat, npair = T.npair;
  const long long base = lm_base(lmax, m);
  const float2* __restrict__ ab = T.ab + base;
  const int nthr = nwaves * 64;

  float2 ct[K];
  float y0[K], y1[K];
  int ny[K];
  float2 wp[K][NF], wm[K][NF];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int j = tid + k * nthr;
    if (j < npair) {
      ct[k] = T.ct[j];
      const Seed sd = T.seed[(long long)m * npair + j];
      y0[k] = sd.y;
      ny[k] = sd.ny;
      y1[k] = (m < lmax) ? act_fma(ab[1].x, y0[k], ct[k], 0.0f) : 0.0f;
      const float w = T.w[j];
      const int js = nlat - 1 - j;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const int fld = min(f0 + f, nfield - 1);
        const float2* g = fm + ((long long)fld * (T.mmax + 1) + m) * nlat;
        const float2 gN = g[j];
        const float2 gS = g[js];
        if (js == j) {  // equator of an odd grid pairs with itself
          wp[k][f] = make_float2(gN.x * w, gN.y * w);
          wm[k][f] = make_float2(0.f, 0.f);
        } else {
          wp[k][f] = make_float2((gN.x + gS.x) * w, (gN.y + gS.y) * w);
          wm[k][f] = make_float2((gN.x - gS.x) * w, (gN.y - gS.y) * w);
        }
      }
    } else {
      ct[k] = make_float2(0.f, 0.f);
      y0[k] = 0.f;
      y1[k] = 0.f;
      ny[k] = 0;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        wp[k][f] = make_float2(0.f, 0.f);
        wm[k][f] = make_float2(0.f, 0.f);
      }
    }
  }

  float acc[64];
  int buf = 0;
  for (int ls = m; ls <= lmax; ls += SPAN) {
#pragma unroll
    for (int i = 0; i < 64; ++i) acc[i] = 0.f;
    bool sc = false;
#pragma unroll
    for (int k = 0; k < K; ++k) sc |= (ny[k] != 0);
    if (__builtin_amdgcn_ballot_w64(sc) != 0)
      analys_span<NF, K, true>(acc, ab, ls, m, lmax, ct, y0, y1, ny, wp, wm);
    else
      analys_span<NF, K, false>(acc, ab, ls, m, lmax, ct, y0, y1, ny, wp, wm);
    const float v = transpose_reduce64(acc);
    red[buf][wv][lane] = v;
    __syncthreads();
    if (wv == 0) {
      float s = red[buf][0][lane];
      for (int w = 1; w < nwaves; ++w) s += red[buf][w][lane];
      const int p = lane / (2 * NF);
      const int rem = lane - p * 2 * NF;
      const int f = rem >> 1;
      const int l = ls + p;
      if (l <= lmax && f0 + f < nfield) {
        float* dst = reinterpret_cast<float*>(const_cast<char*>(Fq.at(f0 + f))) +
                     2 * (base + (l - m));
        dst[rem & 1] = s;
      }
    }
    buf ^= 1;
  }
}

template <int NF>
hipError_t leg_synth_nf(const LegTables& T, int nfield, const Fields& q, float2* fm,
                        hipStream_t s) {
  const int nchunk = (T.npair + 63) / 64;
  const int nwork = (T.mmax + 1) * nchunk;
  dim3 grid((nwork + 3) / 4, (nfield + NF - 1) / NF);
  hipLaunchKernelGGL(k_leg_synth<NF>, grid, dim3(256), 0, s, T, nfield, nchunk, nwork, q, fm);
  return hipGetLastError();
}

template <int NF, int K>
hipError_t leg_analys_nfk(const LegTables& T, int nfield, int nwaves, const float2* fm,
                          const Fields& q, hipStream_t s) {
  dim3 grid(T.mmax + 1, (nfield + NF - 1) / NF);
  hipLaunchKernelGGL((k_leg_analys<NF, K>), grid, dim3(64 * nwaves), 0, s, T, nfield, nwaves,
                     fm, q);
  return hipGetLastError();
}

template <int NF>
hipError_t leg_analys_nf(const LegTables& T, int nfield, const float2* fm, const Fields& q,
                         hipStream_t s) {
  const int nwaves = std::min(4, (T.npair + 63) / 64);
  const int K = (T.npair + 64 * nwaves - 1) / (64 * nwaves);
  switch (K) {
    case 1: return leg_analys_nfk<NF, 1>(T, nfield, nwaves, fm, q, s);
    case 2: return leg_analys_nfk<NF, 2>(T, nfield, nwaves, fm, q, s);
    case 3: return leg_analys_nfk<NF, 3>(T, nfield, nwaves, fm, q, s);
    case 4: return leg_analys_nfk<NF, 4>(T, nfield, nwaves, fm, q, s);
    case 5:
    case 6: return leg_analys_nfk<NF, 6>(T, nfield, nwaves, fm, q, s);
    case 7:
    case 8: return leg_analys_nfk<NF, 8>(T, nfield, nwaves, fm, q, s);
    default: return hipErrorInvalidValue;  // nlat > 4096: rejected at plan creation
  }
}

}  // namespace

hipError_t launch_leg_synth(const LegTables& T, int nfield, const Fields& q, float2* fm,
                            hipStream_t s) {
  if (nfield <= 0) return hipSuccess;
  switch (nfield) {
    case 1: return leg_synth_nf<1>(T, nfield, q, fm, s);
    case 2: return leg_synth_nf<2>(T, nfield, q, fm, s);
    case 3: return leg_synth_nf<3>(T, nfield, q, fm, s);
    default: return leg_synth_nf<4>(T, nfield, q, fm, s);
  }
}

hipError_t launch_leg_analys(const LegTables& T, int nfield, const float2* fm, const Fields& q,
                             hipStream_t s) {
  if (nfield <= 0) return hipSuccess;
  if (nfield == 1) return leg_analys_nf<1>(T, nfield, fm, q, s);
  if (nfield == 2) return leg_analys_nf<2>(T, nfield, fm, q, s);
  return leg_analys_nf<4>(T, nfield, fm, q, s);
}

}  // namespace tsphere
