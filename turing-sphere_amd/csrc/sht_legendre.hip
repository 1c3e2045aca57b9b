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
// North/south latitude pairs share one recurrence (ytilde_l^m(-x) =
// (-1)^(l-m) ytilde_l^m(x)): even (l-m) terms accumulate in E and odd ones in
// O, north = E + O, south = E - O (leg.ts:129-134, 207-219).
//
// Significance seeding.  The reference carries a scaled exponent through the
// recurrence and skips values below ACCURACY = 1e-15 (common.ts, leg.ts:
// 140-157).  Here the host walks that part of the recurrence once, in f64,
// and stores for every (m, latitude pair) the first l (rounded down to a
// multiple of 4 above m) where the value is significant, with the exact f64
// values ytilde_l, ytilde_{l+1} there.  A lane starts its f32 recurrence at
// that point; before it, it contributes exactly 0, as the reference's lane
// does.  No scaling, no per-lane tests in the loop, and the negligible part
// of the (l, m, latitude) volume (about 1 - pi/4 of it) is skipped.
//
// MI355X mapping.  One workgroup per m-pair (m, mmax - m), whose l-loops
// together have a constant length lmax + 2, so the triangle is balanced
// without a tail, or -- for launches of one or two fields -- one workgroup per
// m-row, for twice the resident waves.  The row's Q_lm (all batched fields)
// and (a, b) are staged once into LDS with coalesced loads; the inner loop
// reads them as LDS broadcasts and has no global or scalar loads.  Each lane
// carries K chains (K independent dependency chains), each a chunk of 64
// consecutive latitude pairs that starts at its own significance point.
//  * synthesis: a lane does 4 VALU ops of recurrence (a * cos(theta) with the
//    f32 hi+lo node) + 2 FMAs per field per l;
//  * analysis: each lane accumulates a span of l-values (64 partial sums:
//    span_l * fields * re/im), reduced across the wave by a transposing
//    butterfly -- v_permlane32_swap, v_permlane16_swap, DPP row_mirror /
//    row_half_mirror / quad_perm -- after which lane i holds the complete sum
//    of partial i (141 VALU ops for all 64); waves combine through LDS.
#include "tsphere_internal.h"

namespace tsphere {
namespace {

constexpr int kTile = 4;   // l-values between seed injections (seeds are 4-aligned)
constexpr int kPadQ = 24;  // zero entries after a staged row (reads run past lmax)
constexpr int kPadAnalys = 40;  // analysis spans read up to 32 + 3 entries past lmax
// entries of the staged alpha row in an analysis workgroup's LDS, rounded to
// 4 so that the arrays after it start 16-byte aligned (their 16-byte loads)
__host__ __device__ constexpr int analys_ab_len(int lmax) { return (lmax + 1 + kPadAnalys + 3) & ~3; }

__device__ __forceinline__ long long lm_base(int lmax, int m) {
  return (long long)m * (lmax + 1) - ((long long)m * (m - 1)) / 2;
}

__device__ __forceinline__ void cfma(float2& acc, float y, float2 q) {
  acc.x = fmaf(y, q.x, acc.x);
  acc.y = fmaf(y, q.y, acc.y);
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return __builtin_amdgcn_readfirstlane(v);
}

// Last seed position among the lanes of a chain (-1 if no lane is ever
// significant): past it no lane of the chain takes a seed, and the loops run
// without the per-lane seed test.
__device__ __forceinline__ int last_seed(int ls, int L) { return wave_max(ls <= L ? ls : -1); }

// Latitude pair of this lane's chain k (of K).  A chain is a chunk of 64
// consecutive pairs, so the lanes of one chain span a narrow band of
// latitudes whose significance points are close and the chain can start at
// the smallest of them (near the poles at high m that is far above m).  Chunks
// are dealt to the waves boustrophedon -- wave w takes chunks w, 2 nw - 1 - w,
// 2 nw + w, ... -- so every wave gets polar and equatorial chunks alike, and
// chunk numbers (and so latitudes, toward the equator) increase with k.
__device__ __forceinline__ int pair_of(int k) {
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int chunk = k * nw + ((k & 1) ? nw - 1 - w : w);
  return chunk * 64 + (threadIdx.x & 63);
}

// First chain (of K) that is live at l: chains with no lane significant
// before l are still exactly 0, so chains [kf, K) carry every contribution.
// (Significance moves toward the poles as l grows, so the live chains are
// the last ones; if they are not, kf only keeps some zero chains running.)
template <int K>
__device__ __forceinline__ int first_live(const int (&lk)[K], int l) {
  int kf = K - 1;
#pragma unroll
  for (int k = K - 1; k >= 0; --k)
    if (lk[k] <= l) kf = k;
  return kf;
}

// The m-row this workgroup handles in its pass r (0 or 1), or -1.  Paired
// (rows = 2): workgroup b takes m = b and M - b, whose l-loops sum to
// lmax + 2, laid out XCD by XCD.  Single (rows = 1): each workgroup takes one
// row -- twice the workgroups, each with half the work, so more waves are
// resident when the field count is small.  Workgroups go to XCD b % 8 in
// turn and, in the first pass, to the XCD's CUs in turn, so the t-th
// workgroup of an XCD shares a CU with the (t + C)-th, (t + 2C)-th and
// (t + 3C)-th (C = a quarter of the XCD's workgroups).  The quarters take the
// rows u, n - 1 - u, n/4 + u and 3n/4 - 1 - u (u = x C + t mod C on XCD x):
// two complementary pairs, so every CU gets the same l-length in all, and
// consecutive u on one XCD are consecutive m (a 128-byte line of the Fourier
// scratch holds 16 m of a latitude, shared in one L2).  The longest rows are
// dispatched first.
// (Kernels with fewer resident workgroups per CU than four -- the analysis,
// whose LDS tiles allow two -- keep the older dealing, quarters = false: rows
// in groups of 16 consecutive m dealt boustrophedon over the XCDs, which
// measured slightly ahead there.)
__device__ __forceinline__ int row_m(int r, int rows, int M, bool quarters = true) {
  if (rows == 1) {
    if (r != 0) return -1;
    const int b = blockIdx.x, n = gridDim.x;
    const int x = b & 7, t = b >> 3;  // XCD, position on it
    if (!quarters) {
      if (n % 256 != 0) return b;  // fewer than two groups per XCD: no balance
      const int gi = t >> 4, g = gi * 8 + ((gi & 1) ? 7 - x : x);
      return g * 16 + (t & 15);
    }
    if (n % 32 != 0) return b;  // too few workgroups to deal in quarters
    const int C = n >> 5, Q = n >> 2;
    const int q = t / C, u = x * C + (t - q * C);
    return q == 0 ? u : q == 1 ? n - 1 - u : q == 2 ? Q + u : 3 * Q - 1 - u;
  }
  const int mp = xcd_chunk(blockIdx.x, gridDim.x);
  const int m = r == 0 ? mp : M - mp;
  return (r == 1 && m <= mp) ? -1 : m;
}

// Recurrence precision of the Legendre kernels, both on the rescaled
// functions z_l = ytilde_l / c_l of LegTables:
//   z_l = alpha_l cos(theta) z_{l-1} + z_{l-2}.
// RecF: f32 state, cos(theta) as an f32 hi + lo pair (3 VALU ops per l:
// one product, two FMAs; rounding the node to f32 would move theta, and every
// value with it, by ~ l 2^-24 / sin(theta)).  RecD: f64 state from the f64 alpha, node and seeds
// (2 f64 ops per l), rounded to f32 once per value before the FP32 sums: the
// FP32 form's error at high degree comes from rounding the recurrence state,
// which this removes, at about the FP32 form's cost.
// (z_l, z_{l+1}) -> (z_{l+2}, z_{l+3}) with a0 = alpha_{l+2}, a1 = alpha_{l+3}
struct RecF {
  using Tables = LegTables;
  using ab_t = float;
  using ct_t = float2;
  using y_t = float;
  static __device__ __forceinline__ ab_t zero_ab() { return 0.f; }
  static __device__ __forceinline__ void advance(ab_t a0, ab_t a1, ct_t ct, y_t& y0, y_t& y1) {
    const float u0 = a0 * y1;
    const float t0 = fmaf(u0, ct.x, fmaf(u0, ct.y, y0));
    const float u1 = a1 * t0;
    y1 = fmaf(u1, ct.x, fmaf(u1, ct.y, y1));
    y0 = t0;
  }
};
struct RecD {
  using Tables = LegTablesZ64;
  using ab_t = double;
  using ct_t = double;
  using y_t = double;
  static __device__ __forceinline__ ab_t zero_ab() { return 0.0; }
  static __device__ __forceinline__ void advance(ab_t a0, ab_t a1, ct_t ct, y_t& y0, y_t& y1) {
    const double t0 = fma(a0 * y1, ct, y0);
    y1 = fma(a1 * t0, ct, y1);
    y0 = t0;
  }
};

template <class P>
struct ChainT {
  typename P::ct_t ct;
  typename P::y_t y0, y1, s0, s1;
  int ls;  // seed position; INT_MAX = never significant up to lmax
};

template <class P>
__device__ __forceinline__ void load_chain_t(const typename P::Tables& T, int m, int j,
                                             ChainT<P>& c) {
  if (j < T.npair) {
    c.ct = T.ct[j];
    const auto sd = T.seed[(long long)m * T.npair + j];
    c.s0 = sd.y0;
    c.s1 = sd.y1;
    c.ls = sd.l;
  } else {
    c.ct = typename P::ct_t{};
    c.s0 = c.s1 = 0;
    c.ls = 0x7fffffff;
  }
  c.y0 = c.y1 = 0;
}

// Stage a row's alpha (in the recurrence's precision) and the fields' Q
// scaled by c_l (the synthesis sums z_l (c_l Q_l)).
template <int NF, class P>
__device__ __forceinline__ void stage_row_t(const typename P::Tables& T, int m, int f0,
                                            int nfield, const Fields& Fq, bool with_q,
                                            typename P::ab_t* sab, float2* sq, int pad) {
  const int n = T.lmax - m + 1;
  const long long base = lm_base(T.lmax, m);
  for (int i = threadIdx.x; i < n + pad; i += blockDim.x) {
    sab[i] = i < n ? T.ab[base + i] : P::zero_ab();
    if (with_q) {
      const float cz = i < n ? T.cz[base + i] : 0.f;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const int fld = min(f0 + f, nfield - 1);
        const float2 q =
            i < n ? reinterpret_cast<const float2*>(Fq.at(fld))[base + i] : make_float2(0.f, 0.f);
        sq[i * NF + f] = make_float2(q.x * cz, q.y * cz);
      }
    }
  }
}

// The implicit step of one real coefficient, each product and sum rounded
// as written (the generated kernel's (X + (s1 * A)) / (1 + (s2 * lam))).
__device__ __forceinline__ float imex_step(float x, float s1, float a, float s2, float lam) {
#pragma clang fp contract(off)
  return (x + (s1 * a)) / (1.0f + (s2 * lam));
}

// The implicit step's scalars of field fld: from the device array of a
// batched step, else from the launch arguments (fld < 4 there).
__device__ __forceinline__ void imex_coefs(const ImexCoef& ic, int fld, float& s1, float& s2) {
  if (ic.dev) {
    s1 = ic.dev[2 * fld];
    s2 = ic.dev[2 * fld + 1];
  } else {
    s1 = ic.s1[fld & 3];
    s2 = ic.s2[fld & 3];
  }
}

// X = (X + s1 a) / (1 + s2 lam) of one complex coefficient, (s1, s2) =
// sc[0], sc[1]
__device__ __forceinline__ void imex_apply(float2* X, float2 a, float2 lm, const float* sc) {
  const float s1 = sc[0], s2 = sc[1];
  const float2 x = *X;
  *X = make_float2(imex_step(x.x, s1, a.x, s2, lm.x), imex_step(x.y, s1, a.y, s2, lm.y));
}

// stage_row_t with the implicit step in front: field f's state X is replaced
// by X_new = (X + s1 A) / (1 + s2 lam) (A: its analysed values), written back
// to the state, and c_l X_new staged as the synthesis coefficients.
template <int NF, class P>
__device__ __forceinline__ void stage_row_imex(const typename P::Tables& T, int m, int f0,
                                               int nfield, const Fields& Fq, const Fields& Fa,
                                               const float* lam, const ImexCoef& ic,
                                               typename P::ab_t* sab, float2* sq, int pad) {
  const int n = T.lmax - m + 1;
  const long long base = lm_base(T.lmax, m);
  for (int i = threadIdx.x; i < n + pad; i += blockDim.x) {
    sab[i] = i < n ? T.ab[base + i] : P::zero_ab();
    const float cz = i < n ? T.cz[base + i] : 0.f;
    const float2 lm2 =
        i < n ? reinterpret_cast<const float2*>(lam)[base + i] : make_float2(0.f, 0.f);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const int fld = f0 + f;
      float2 xn = make_float2(0.f, 0.f);
      if (i < n && fld < nfield) {
        float2* X = reinterpret_cast<float2*>(const_cast<char*>(Fq.at(fld))) + base + i;
        const float2 x = *X;
        const float2 a = reinterpret_cast<const float2*>(Fa.at(fld))[base + i];
        float c1, c2;
        imex_coefs(ic, fld, c1, c2);
        xn.x = imex_step(x.x, c1, a.x, c2, lm2.x);
        xn.y = imex_step(x.y, c1, a.y, c2, lm2.y);
        *X = xn;
      }
      sq[i * NF + f] = make_float2(xn.x * cz, xn.y * cz);
    }
  }
}

// ------------------------------------------------------------- synthesis
// kTile l-values of chains [K0, K) from l = m + i (INJ: lanes may take
// their seed at l).
template <int K0, bool INJ, int NF, int K, class P>
__device__ __forceinline__ void synth_tile(const typename P::ab_t* sab, const float2* sq, int i,
                                           int l, ChainT<P> (&c)[K], float2 (&accE)[K][NF],
                                           float2 (&accO)[K][NF]) {
  if constexpr (INJ) {
#pragma unroll
    for (int k = K0; k < K; ++k)
      if (c[k].ls == l) {
        c[k].y0 = c[k].s0;
        c[k].y1 = c[k].s1;
      }
  }
#pragma unroll
  for (int p = 0; p < kTile; p += 2) {
    float2 q0[NF], q1[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      q0[f] = sq[(i + p) * NF + f];
      q1[f] = sq[(i + p + 1) * NF + f];
    }
    const auto a0 = sab[i + p + 2];
    const auto a1 = sab[i + p + 3];
#pragma unroll
    for (int k = K0; k < K; ++k) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        cfma(accE[k][f], (float)c[k].y0, q0[f]);
        cfma(accO[k][f], (float)c[k].y1, q1[f]);
      }
      P::advance(a0, a1, c[k].ct, c[k].y0, c[k].y1);
    }
  }
}

// Tiles from l on, in phases: phase K0 runs chains [K0, K) until a chain
// below K0 becomes live (lk[k] <= l), then hands over to phase K0 - 1.  Each
// phase is a branch-free loop; a chain run before it is live only adds zeros.
// Within a phase, tiles past the last seed of its chains (lz) skip the seed
// test.
template <int K0, int NF, int K, class P>
__device__ __forceinline__ void synth_phases(int l, int L, int m, const int (&lk)[K],
                                             const int (&lz)[K], const typename P::ab_t* sab,
                                             const float2* sq, ChainT<P> (&c)[K],
                                             float2 (&accE)[K][NF], float2 (&accO)[K][NF]) {
  int lend = L + 1, injend = 0;
#pragma unroll
  for (int k = 0; k < K0; ++k) lend = min(lend, lk[k]);
#pragma unroll
  for (int k = K0; k < K; ++k) injend = max(injend, lz[k] + 1);
  const int l1 = min(lend, injend);
  for (; l < l1; l += kTile) synth_tile<K0, true, NF, K, P>(sab, sq, l - m, l, c, accE, accO);
  for (; l < lend; l += kTile) synth_tile<K0, false, NF, K, P>(sab, sq, l - m, l, c, accE, accO);
  if constexpr (K0 > 0)
    synth_phases<K0 - 1, NF, K, P>(l, L, m, lk, lz, sab, sq, c, accE, accO);
}

template <int NF, int K, class P>
__global__ __launch_bounds__(256) void k_leg_synth(typename P::Tables T, int nfield, Fields Fq,
                                                   float2* __restrict__ fm, int rows) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  const int L = T.lmax, M = T.mmax;
  auto* sab = reinterpret_cast<typename P::ab_t*>(lds_raw);
  float2* sq = reinterpret_cast<float2*>(sab + (L + 1 + kPadQ));
  const int f0 = blockIdx.y * NF;

  for (int r = 0; r < 2; ++r) {
    const int m = row_m(r, rows, M);
    if (m < 0) break;
    __syncthreads();  // the previous row's readers are done
    stage_row_t<NF, P>(T, m, f0, nfield, Fq, true, sab, sq, kPadQ);
    __syncthreads();

    ChainT<P> c[K];
    int lk[K], lz[K];  // first and last l where a lane of chain k takes its seed
    int lmin = 0x7fffffff;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      load_chain_t<P>(T, m, pair_of(k), c[k]);
      lk[k] = wave_min(c[k].ls);
      lz[k] = last_seed(c[k].ls, L);
      lmin = min(lmin, lk[k]);
    }

    float2 accE[K][NF], accO[K][NF];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int f = 0; f < NF; ++f) accE[k][f] = accO[k][f] = make_float2(0.f, 0.f);

    if (lmin <= L) synth_phases<K - 1, NF, K, P>(lmin, L, m, lk, lz, sab, sq, c, accE, accO);

    const int nlat = T.nlat;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int j = pair_of(k);
      if (j >= T.npair) continue;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        if (f0 + f >= nfield) continue;
        // fm row-major per latitude: (f, ilat, m) at (f nlat + ilat)(M+1) + m
        float2* out = fm + (long long)(f0 + f) * nlat * (M + 1) + m;
        out[(long long)j * (M + 1)] =
            make_float2(accE[k][f].x + accO[k][f].x, accE[k][f].y + accO[k][f].y);
        out[(long long)(nlat - 1 - j) * (M + 1)] =
            make_float2(accE[k][f].x - accO[k][f].x, accE[k][f].y - accO[k][f].y);
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
// One span of SPAN l-values from ls for chains [K0, K): partial sums acc
// (zeroed here), laid out (l, field, re/im).
template <int K0, bool INJ, int NF, int K, class P>
__device__ __forceinline__ void analys_span(const typename P::ab_t* sab, int ls, int m,
                                            ChainT<P> (&c)[K], const float2 (&wp)[K][NF],
                                            const float2 (&wm)[K][NF], float (&acc)[64]) {
  constexpr int SPAN = 32 / NF;
#pragma unroll
  for (int i = 0; i < 64; ++i) acc[i] = 0.f;
#pragma unroll
  for (int p = 0; p < SPAN; p += 2) {
    const int l = ls + p;
    if (INJ && p % kTile == 0) {
#pragma unroll
      for (int k = K0; k < K; ++k)
        if (c[k].ls == l) {
          c[k].y0 = c[k].s0;
          c[k].y1 = c[k].s1;
        }
    }
#pragma unroll
    for (int k = K0; k < K; ++k) {
      const float y0 = (float)c[k].y0, y1 = (float)c[k].y1;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        float* e = &acc[(p * NF + f) * 2];
        float* o = &acc[((p + 1) * NF + f) * 2];
        e[0] = fmaf(y0, wp[k][f].x, e[0]);
        e[1] = fmaf(y0, wp[k][f].y, e[1]);
        o[0] = fmaf(y1, wm[k][f].x, o[0]);
        o[1] = fmaf(y1, wm[k][f].y, o[1]);
      }
    }
    const int i = l - m;
    const auto a0 = sab[i + 2];
    const auto a1 = sab[i + 3];
#pragma unroll
    for (int k = K0; k < K; ++k) P::advance(a0, a1, c[k].ct, c[k].y0, c[k].y1);
  }
}

template <int K0, bool INJ, int NF, int K, class P>
__device__ __forceinline__ void analys_span_from(int kf, const typename P::ab_t* sab, int ls,
                                                 int m, ChainT<P> (&c)[K],
                                                 const float2 (&wp)[K][NF],
                                                 const float2 (&wm)[K][NF], float (&acc)[64]) {
  if constexpr (K0 + 1 >= K) {
    analys_span<K0, INJ, NF, K, P>(sab, ls, m, c, wp, wm, acc);
  } else {
    if (kf <= K0)
      analys_span<K0, INJ, NF, K, P>(sab, ls, m, c, wp, wm, acc);
    else
      analys_span_from<K0 + 1, INJ, NF, K, P>(kf, sab, ls, m, c, wp, wm, acc);
  }
}

template <int NF, int K, class P>
__global__ __launch_bounds__(256) void k_leg_analys(typename P::Tables T, int nfield,
                                                    const float2* __restrict__ fm, Fields Fq,
                                                    int rows) {
  constexpr int SPAN = 32 / NF;  // l-values per reduced span
  static_assert(SPAN % kTile == 0 && SPAN * NF * 2 == 64, "64 partials per lane");
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  __shared__ float red[2][4][64];
  __shared__ int smin[4];
  const int L = T.lmax, M = T.mmax, nlat = T.nlat;
  auto* sab = reinterpret_cast<typename P::ab_t*>(lds_raw);
  const int f0 = blockIdx.y * NF;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nthr = blockDim.x, nwaves = nthr >> 6;

  for (int r = 0; r < 2; ++r) {
    const int m = row_m(r, rows, M, false);
    if (m < 0) break;
    const long long base = lm_base(L, m);
    __syncthreads();
    stage_row_t<NF, P>(T, m, f0, nfield, Fq, false, sab, nullptr, kPadAnalys);

    ChainT<P> c[K];
    float2 wp[K][NF], wm[K][NF];
    int lk[K];  // first l where some lane of chain k is significant
    int lzmax = -1;  // last l where a lane of this wave takes its seed
    int lmin = 0x7fffffff;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int j = pair_of(k);
      load_chain_t<P>(T, m, j, c[k]);
      lk[k] = wave_min(c[k].ls);
      lzmax = max(lzmax, last_seed(c[k].ls, L));
      lmin = min(lmin, lk[k]);
      const int jc = min(j, T.npair - 1);
      const int js = nlat - 1 - jc;
      const float w = j < T.npair ? T.w[jc] : 0.f;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const int fld = min(f0 + f, nfield - 1);
        const float2* g = fm + (long long)fld * nlat * (M + 1) + m;
        const float2 gN = g[(long long)jc * (M + 1)];
        const float2 gS = g[(long long)js * (M + 1)];
        if (js == jc) {  // equator of an odd grid pairs with itself
          wp[k][f] = make_float2(gN.x * w, gN.y * w);
          wm[k][f] = make_float2(0.f, 0.f);
        } else {
          wp[k][f] = make_float2((gN.x + gS.x) * w, (gN.y + gS.y) * w);
          wm[k][f] = make_float2((gN.x - gS.x) * w, (gN.y - gS.y) * w);
        }
      }
    }
    if (lane == 0) smin[wv] = lmin;
    __syncthreads();  // also publishes the staged row
    int lblk = smin[0];
    for (int w = 1; w < nwaves; ++w) lblk = min(lblk, smin[w]);
    lblk = min(lblk, L + 1);

    // coefficients below every lane's significance point are exactly 0
    for (int i = m + tid; i < lblk; i += nthr)
#pragma unroll
      for (int f = 0; f < NF; ++f)
        if (f0 + f < nfield)
          reinterpret_cast<float2*>(const_cast<char*>(Fq.at(f0 + f)))[base + (i - m)] =
              make_float2(0.f, 0.f);

    int buf = 0;
    for (int ls = lblk; ls <= L; ls += SPAN) {
      if (lmin >= ls + SPAN) {
        // No lane of this wave is significant before the next span: its
        // chains are all still 0 and no seed falls in this span, so its
        // partial sums are exactly 0.
        red[buf][wv][lane] = 0.f;
      } else {
        float acc[64];
        const int kf = first_live<K>(lk, ls + SPAN - 1);
        if (ls <= lzmax)  // a lane may still take its seed in this span
          analys_span_from<0, true, NF, K, P>(kf, sab, ls, m, c, wp, wm, acc);
        else
          analys_span_from<0, false, NF, K, P>(kf, sab, ls, m, c, wp, wm, acc);
        red[buf][wv][lane] = transpose_reduce64(acc);
      }
      __syncthreads();
      if (wv == 0) {
        float s = red[buf][0][lane];
        for (int w = 1; w < nwaves; ++w) s += red[buf][w][lane];
        const int p = lane / (2 * NF);
        const int rem = lane - p * 2 * NF;
        const int f = rem >> 1;
        const int l = ls + p;
        if (l <= L && f0 + f < nfield) {
          float* dst = reinterpret_cast<float*>(const_cast<char*>(Fq.at(f0 + f))) +
                       2 * (base + (l - m));
          dst[rem & 1] = s * T.cz[base + (l - m)];
        }
      }
      buf ^= 1;
    }
  }
}

template <class P>
size_t synth_lds(int lmax, int nf) {
  return (sizeof(typename P::ab_t) + sizeof(float2) * nf) * (lmax + 1 + kPadQ);
}

template <int NF, int K, class P>
hipError_t synth_nfk(const typename P::Tables& T, int nfield, int nwaves, int rows,
                     const Fields& q, float2* fm, hipStream_t s) {
  dim3 grid(rows == 1 ? T.mmax + 1 : T.mmax / 2 + 1, (nfield + NF - 1) / NF);
  hipLaunchKernelGGL((k_leg_synth<NF, K, P>), grid, dim3(64 * nwaves), synth_lds<P>(T.lmax, NF),
                     s, T, nfield, q, fm, rows);
  return hipGetLastError();
}

template <int NF, int K, class P>
hipError_t analys_nfk(const typename P::Tables& T, int nfield, int nwaves, int rows,
                      const float2* fm, const Fields& q, hipStream_t s) {
  dim3 grid(rows == 1 ? T.mmax + 1 : T.mmax / 2 + 1, (nfield + NF - 1) / NF);
  hipLaunchKernelGGL((k_leg_analys<NF, K, P>), grid, dim3(64 * nwaves),
                     sizeof(typename P::ab_t) * (T.lmax + 1 + kPadAnalys), s, T, nfield, fm, q,
                     rows);
  return hipGetLastError();
}

template <int NF, class P>
hipError_t synth_nf(const typename P::Tables& T, int nfield, const LegTile& t, const Fields& q,
                    float2* fm, hipStream_t s) {
  switch (t.k) {
    case 1: return synth_nfk<NF, 1, P>(T, nfield, t.nwaves, t.rows, q, fm, s);
    case 2: return synth_nfk<NF, 2, P>(T, nfield, t.nwaves, t.rows, q, fm, s);
    case 4: return synth_nfk<NF, 4, P>(T, nfield, t.nwaves, t.rows, q, fm, s);
    case 8: return synth_nfk<NF, 8, P>(T, nfield, t.nwaves, t.rows, q, fm, s);
    default: return hipErrorInvalidValue;
  }
}

template <int NF, class P>
hipError_t analys_nf(const typename P::Tables& T, int nfield, const LegTile& t, const float2* fm,
                     const Fields& q, hipStream_t s) {
  switch (t.k) {
    case 1: return analys_nfk<NF, 1, P>(T, nfield, t.nwaves, t.rows, fm, q, s);
    case 2: return analys_nfk<NF, 2, P>(T, nfield, t.nwaves, t.rows, fm, q, s);
    case 4: return analys_nfk<NF, 4, P>(T, nfield, t.nwaves, t.rows, fm, q, s);
    case 8: return analys_nfk<NF, 8, P>(T, nfield, t.nwaves, t.rows, fm, q, s);
    default: return hipErrorInvalidValue;
  }
}

// ------------------------------------------ contraction on FP32 matrix cores
// The same Legendre stage with the sum over l (synthesis) or over latitude
// pairs (analysis) on v_mfma_f32_16x16x4_f32 -- FP32 operands, FP32
// accumulation -- while the VALU only runs the recurrence (4 ops per l per
// pair instead of 8-13).  Operand layout (checked on gfx950): A lane L holds
// A[L % 16][L / 16], B lane L holds B[L / 16][L % 16], D lane L holds
// D[4 (L / 16) + r][L % 16].  Columns are (parity, field, re/im): even and
// odd l - m terms go to separate columns, so one product serves both (the
// north / south recombination happens on the way out).  Each lane carries KC
// latitude pairs; a wave's pairs pass through a wave-private LDS tile of 16 l
// that turns "lane = pair" into the operand layout.
typedef __attribute__((ext_vector_type(4))) float f4v;
constexpr int MF_TL = 16;   // l per tile
constexpr int MF_PAD = 24;  // zero (a, b) / Q entries after a staged row

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ f4v mf4(float a, float b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// The waves' partial sums of one output, from a buffer with a slot for each
// of 8 waves (at(w): slot w), added in wave order -- the order of a loop over
// the first nwaves slots -- with the reads issued four at a time (slots past
// nwaves exist and are ignored): two LDS round trips instead of one per wave,
// for four registers.
template <class At>
__device__ __forceinline__ float sum_waves(At at, int nwaves) {
  float v[4];
#pragma unroll
  for (int w = 0; w < 4; ++w) v[w] = at(w);
  float s = v[0];
#pragma unroll
  for (int w = 1; w < 4; ++w)
    if (w < nwaves) s += v[w];
  if (nwaves > 4) {
#pragma unroll
    for (int w = 0; w < 4; ++w) v[w] = at(w + 4);
#pragma unroll
    for (int w = 0; w < 4; ++w)
      if (w + 4 < nwaves) s += v[w];
  }
  return s;
}

// Synthesis: D[pair][col] = sum_l P[pair][l] B[l][col], B = Q masked by parity.
// A row tile is 16 pairs; K-block kb of a 16-l tile holds l = lt + 4 (L/16) + kb
// for lane L, so each lane reads its 4 K values as one 16-byte LDS load.
template <int NF, int KC, class P>
__global__ __launch_bounds__(256) void k_leg_synth_mf(typename P::Tables T, int nfield, Fields Fq,
                                                      float2* __restrict__ fm) {
  constexpr int NCOL = 4 * NF;
  constexpr int QW = 64 * KC;        // pairs per wave
  constexpr int NRT = QW / 16;       // row tiles per wave
  constexpr int PROW = MF_TL + 4;    // floats per pair row of the tile (padded)
  extern __shared__ float smem[];
  const int L = T.lmax, M = T.mmax, nlat = T.nlat;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int r16 = lane & 15, kg = lane >> 4;
  const int nrow = (L + 1 + MF_PAD + 1) & ~1;  // even: keeps the tile 16-byte aligned
  auto* sab = reinterpret_cast<typename P::ab_t*>(smem);
  float2* sq = reinterpret_cast<float2*>(sab + nrow);
  float* tl = reinterpret_cast<float*>(sq + nrow * NF) + wv * (QW * PROW);
  const int f0 = blockIdx.y * NF;
  const int wp0 = wv * QW;  // this wave's first pair
  // this lane's output column
  const int par = r16 / (2 * NF), cf = (r16 % (2 * NF)) >> 1, ri = r16 & 1;

  for (int r = 0; r < 2; ++r) {
    const int mp = xcd_chunk(blockIdx.x, gridDim.x);
    const int m = r == 0 ? mp : M - mp;
    if (r == 1 && m <= mp) break;
    __syncthreads();
    stage_row_t<NF, P>(T, m, f0, nfield, Fq, true, sab, sq, MF_PAD);
    __syncthreads();

    ChainT<P> c[KC];
    int lmin = 0x7fffffff;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      load_chain_t<P>(T, m, wp0 + lane + 64 * k, c[k]);
      lmin = min(lmin, c[k].ls);
    }
    lmin = wave_min(lmin);

    f4v acc[NRT];
#pragma unroll
    for (int t = 0; t < NRT; ++t) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};

    for (int lt = lmin; lt <= L; lt += MF_TL) {
      // 1) 16 values of each chain -> its pair row, four 16-byte stores
      float v[KC][MF_TL];
#pragma unroll
      for (int t = 0; t < MF_TL; t += 2) {
        if (t % kTile == 0) {
#pragma unroll
          for (int k = 0; k < KC; ++k)
            if (c[k].ls == lt + t) {
              c[k].y0 = c[k].s0;
              c[k].y1 = c[k].s1;
            }
        }
        const int i = lt + t - m;
        const auto a0 = sab[i + 2], a1 = sab[i + 3];
#pragma unroll
        for (int k = 0; k < KC; ++k) {
          v[k][t] = (float)c[k].y0;
          v[k][t + 1] = (float)c[k].y1;
          P::advance(a0, a1, c[k].ct, c[k].y0, c[k].y1);
        }
      }
#pragma unroll
      for (int k = 0; k < KC; ++k)
#pragma unroll
        for (int t4 = 0; t4 < MF_TL / 4; ++t4)
          *reinterpret_cast<float4*>(tl + (lane + 64 * k) * PROW + 4 * t4) =
              make_float4(v[k][4 * t4], v[k][4 * t4 + 1], v[k][4 * t4 + 2], v[k][4 * t4 + 3]);
      wave_sync_lds();
      // 2) B: Q of l = lt + 4 kg + kb for this lane's column, zero off its parity
      float b[4];
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        const int l = lt + 4 * kg + kb;
        const float2 q = sq[(l - m) * NF + cf];
        b[kb] = (r16 < NCOL && ((l - m) & 1) == par && l <= L) ? (ri ? q.y : q.x) : 0.f;
      }
      // 3) 4 products per row tile
#pragma unroll
      for (int rt = 0; rt < NRT; ++rt) {
        const float4 a4 = *reinterpret_cast<const float4*>(tl + (16 * rt + r16) * PROW + 4 * kg);
        acc[rt] = mf4(a4.x, b[0], acc[rt]);
        acc[rt] = mf4(a4.y, b[1], acc[rt]);
        acc[rt] = mf4(a4.z, b[2], acc[rt]);
        acc[rt] = mf4(a4.w, b[3], acc[rt]);
      }
      wave_sync_lds();  // the tile is rewritten next iteration
    }

    // north = E + O, south = E - O: the odd column of (field, re/im) is 2 NF lanes up
#pragma unroll
    for (int rt = 0; rt < NRT; ++rt)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float e = acc[rt][q];
        const float o = __shfl(e, lane + 2 * NF);
        const int j = wp0 + 16 * rt + 4 * kg + q;
        if (r16 < 2 * NF && f0 + cf < nfield && j < T.npair) {
          float* out = reinterpret_cast<float*>(fm + (long long)(f0 + cf) * nlat * (M + 1) + m) + ri;
          out[2LL * j * (M + 1)] = e + o;
          out[2LL * (nlat - 1 - j) * (M + 1)] = e - o;
        }
      }
  }
}

// Analysis: D[l][col] = sum_pairs P[pair][l] W[pair][col], W = (G_N +/- G_S) w
// by parity (leg.ts:207-219).  K-block kb of the wave holds pairs
// (L / 16) * NKB + kb for lane L, so a lane's A values for four K-blocks are one
// 16-byte load; W stays in registers for the whole row.  The four waves' partial
// sums of a tile meet in LDS.
template <int NF, int KC, class P>
__global__ __launch_bounds__(256) void k_leg_analys_mf(typename P::Tables T, int nfield,
                                                       const float2* __restrict__ fm, Fields Fq) {
  constexpr int NCOL = 4 * NF;
  constexpr int QW = 64 * KC;   // pairs per wave
  constexpr int NKB = QW / 4;   // K-blocks per wave
  constexpr int TROW = QW + 4;  // floats per l row of the tile (padded)
  extern __shared__ float smem[];
  __shared__ int smin[4];
  const int L = T.lmax, M = T.mmax, nlat = T.nlat;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nthr = blockDim.x, nwaves = nthr >> 6;
  const int r16 = lane & 15, kg = lane >> 4;
  auto* sab = reinterpret_cast<typename P::ab_t*>(smem);
  float* tiles = reinterpret_cast<float*>(sab + analys_ab_len(L));
  float* tl = tiles + wv * (2 * MF_TL * TROW);  // two tiles per wave
  float* red = tiles + 4 * 2 * MF_TL * TROW;     // [4][256]
  const int f0 = blockIdx.y * NF;
  const int wp0 = wv * QW;
  const int par = r16 / (2 * NF), cf = (r16 % (2 * NF)) >> 1, ri = r16 & 1;

  for (int r = 0; r < 2; ++r) {
    const int mp = xcd_chunk(blockIdx.x, gridDim.x);
    const int m = r == 0 ? mp : M - mp;
    if (r == 1 && m <= mp) break;
    const long long base = lm_base(L, m);
    __syncthreads();
    stage_row_t<NF, P>(T, m, f0, nfield, Fq, false, sab, nullptr, kPadAnalys);

    ChainT<P> c[KC];
    int lmin = 0x7fffffff;
#pragma unroll
    for (int k = 0; k < KC; ++k) {
      load_chain_t<P>(T, m, wp0 + lane + 64 * k, c[k]);
      lmin = min(lmin, c[k].ls);
    }
    // B operands for the whole row: pair wp0 + kg NKB + kb, this lane's column
    float wB[NKB];
    {
      const bool live = r16 < NCOL && f0 + cf < nfield;
      const int fld = min(f0 + cf, nfield - 1);
      const float* g =
          reinterpret_cast<const float*>(fm + (long long)fld * nlat * (M + 1) + m) + ri;
#pragma unroll
      for (int kb = 0; kb < NKB; ++kb) {
        const int j = wp0 + kg * NKB + kb;
        float v = 0.f;
        if (live && j < T.npair) {
          const int js = nlat - 1 - j;
          const float w = T.w[j];
          const float gN = g[2LL * j * (M + 1)], gS = g[2LL * js * (M + 1)];
          if (js == j)  // equator of an odd grid pairs with itself
            v = par == 0 ? gN * w : 0.f;
          else
            v = par == 0 ? (gN + gS) * w : (gN - gS) * w;
        }
        wB[kb] = v;
      }
    }
    lmin = wave_min(lmin);
    if (lane == 0) smin[wv] = lmin;
    __syncthreads();  // also publishes the staged row
    int lblk = smin[0];
    for (int w = 1; w < nwaves; ++w) lblk = min(lblk, smin[w]);
    lblk = min(lblk, L + 1);

    // coefficients below every lane's significance point are exactly 0
    for (int i = m + tid; i < lblk; i += nthr)
#pragma unroll
      for (int f = 0; f < NF; ++f)
        if (f0 + f < nfield)
          reinterpret_cast<float2*>(const_cast<char*>(Fq.at(f0 + f)))[base + (i - m)] =
              make_float2(0.f, 0.f);

    // Software pipeline: while the matrix core contracts tile i, the VALU runs
    // the recurrence of tile i + 1 into the other half of the wave's double
    // tile (one step of two l per four products).
    auto rec_step = [&](float* dst, int lt, int s) {
      const int t = 2 * s;
      if (t % kTile == 0) {
#pragma unroll
        for (int k = 0; k < KC; ++k)
          if (c[k].ls == lt + t) {
            c[k].y0 = c[k].s0;
            c[k].y1 = c[k].s1;
          }
      }
      const int i = lt + t - m;
      const auto a0 = sab[i + 2], a1 = sab[i + 3];
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        dst[t * TROW + lane + 64 * k] = (float)c[k].y0;
        dst[(t + 1) * TROW + lane + 64 * k] = (float)c[k].y1;
        P::advance(a0, a1, c[k].ct, c[k].y0, c[k].y1);
      }
    };
    constexpr int G = NKB / 4;  // groups of four products per tile
    if (lblk <= L) {
#pragma unroll
      for (int st = 0; st < MF_TL / 2; ++st) rec_step(tl, lblk, st);
    }
    int cur = 0;
    for (int lt = lblk; lt <= L; lt += MF_TL) {
      wave_sync_lds();  // tile `cur` complete; the other half is free
      const float* tc = tl + cur * (MF_TL * TROW);
      float* tn = tl + (cur ^ 1) * (MF_TL * TROW);
      f4v acc = f4v{0.f, 0.f, 0.f, 0.f};
      const float* arow = tc + r16 * TROW + kg * NKB;
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const float4 a4 = *reinterpret_cast<const float4*>(arow + 4 * g);
        acc = mf4(a4.x, wB[4 * g], acc);
        acc = mf4(a4.y, wB[4 * g + 1], acc);
        acc = mf4(a4.z, wB[4 * g + 2], acc);
        acc = mf4(a4.w, wB[4 * g + 3], acc);
#pragma unroll
        for (int st = (MF_TL / 2) * g / G; st < (MF_TL / 2) * (g + 1) / G; ++st)
          rec_step(tn, lt + MF_TL, st);
      }
      cur ^= 1;
      // the waves' partial sums meet in LDS; one thread per (l, column)
#pragma unroll
      for (int q = 0; q < 4; ++q) red[wv * 256 + (4 * kg + q) * 16 + r16] = acc[q];
      __syncthreads();
      for (int idx = tid; idx < 256; idx += nthr) {
        const int row = idx >> 4, col = idx & 15;
        const int l = lt + row;
        const int cpar = col / (2 * NF);
        if (col < NCOL && l <= L && ((l - m) & 1) == cpar) {
          float s = red[idx];
          for (int w = 1; w < nwaves; ++w) s += red[w * 256 + idx];
          const int f = (col % (2 * NF)) >> 1;
          if (f0 + f < nfield)
            reinterpret_cast<float*>(const_cast<char*>(Fq.at(f0 + f)))[2 * (base + (l - m)) +
                                                                        (col & 1)] =
                s * T.cz[base + (l - m)];
        }
      }
      __syncthreads();  // red is rewritten next iteration
    }
  }
}

template <int NF, int KC, class P>
hipError_t synth_mf(const typename P::Tables& T, int nfield, int nwaves, const Fields& q, float2* fm,
                    hipStream_t s) {
  dim3 grid(T.mmax / 2 + 1, (nfield + NF - 1) / NF);
  const size_t lds = (sizeof(typename P::ab_t) + sizeof(float2) * NF) * ((T.lmax + 1 + MF_PAD + 1) & ~1) +
                     sizeof(float) * (size_t)nwaves * 64 * KC * (MF_TL + 4);
  hipLaunchKernelGGL((k_leg_synth_mf<NF, KC, P>), grid, dim3(64 * nwaves), lds, s, T, nfield, q, fm);
  return hipGetLastError();
}

template <int NF, int KC, class P>
hipError_t analys_mf(const typename P::Tables& T, int nfield, int nwaves, const float2* fm,
                     const Fields& q, hipStream_t s) {
  dim3 grid(T.mmax / 2 + 1, (nfield + NF - 1) / NF);
  const size_t lds = sizeof(typename P::ab_t) * analys_ab_len(T.lmax) +
                     sizeof(float) * (4 * 2 * MF_TL * (size_t)(64 * KC + 4) + 4 * 256);
  hipLaunchKernelGGL((k_leg_analys_mf<NF, KC, P>), grid, dim3(64 * nwaves), lds, s, T, nfield, fm,
                     q);
  return hipGetLastError();
}

template <int NF, class P>
hipError_t synth_mf_nf(const typename P::Tables& T, int nfield, const LegTile& t, const Fields& q,
                       float2* fm, hipStream_t s) {
  switch (t.k) {
    case 1: return synth_mf<NF, 1, P>(T, nfield, t.nwaves, q, fm, s);
    case 2: return synth_mf<NF, 2, P>(T, nfield, t.nwaves, q, fm, s);
    case 4: return synth_mf<NF, 4, P>(T, nfield, t.nwaves, q, fm, s);
    default: return hipErrorInvalidValue;
  }
}

template <int NF, class P>
hipError_t analys_mf_nf(const typename P::Tables& T, int nfield, const LegTile& t, const float2* fm,
                        const Fields& q, hipStream_t s) {
  switch (t.k) {
    case 1: return analys_mf<NF, 1, P>(T, nfield, t.nwaves, fm, q, s);
    case 2: return analys_mf<NF, 2, P>(T, nfield, t.nwaves, fm, q, s);
    case 4: return analys_mf<NF, 4, P>(T, nfield, t.nwaves, fm, q, s);
    default: return hipErrorInvalidValue;
  }
}

// ---------------------------------- contraction on 4x4 FP32 matrix blocks
// v_mfma_f32_4x4x1_16b_f32 runs 16 independent 4 x 4 x 1 products per wave:
// block b is lanes 4b..4b+3, A_b[r] comes from lane 4b + r, B_b[j] from lane
// 4b + j, and D_b[r][j] lands in register r of lane 4b + j (checked on
// gfx950).  Its four columns are exactly (field pair) x (re, im) of two
// fields, so with "lane = latitude pair" no column is idle and no operand
// needs a transpose for the synthesis: the matrix cores take every sum
// (2 x 64 FP32 FMAs per wave per l and field pair, at twice the VALU's FMA
// rate and on their own pipe) and the VALU runs the recurrence alone.
__device__ __forceinline__ f4v mf441(float a, float b, f4v c) {
  return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
}

// Q of l-offset i, field pair fp, for this lane's column j (0 for the idle
// half of a one-field launch).  sqf: the staged row as floats, (i, f, re/im).
template <int NF>
__device__ __forceinline__ float qcol(const float* sqf, int i, int fp, int j) {
  if constexpr (NF == 1) {
    const float v = sqf[i * 2 + (j & 1)];
    return j < 2 ? v : 0.f;
  } else {
    return sqf[i * 2 * NF + 4 * fp + j];
  }
}

// Operands of NT l-values (4 or 8) from l = m + i: the B columns of Q and
// the recurrence's alpha.
template <int NT, int NF, class P>
struct TileOps {
  float bq[NT][(NF + 1) / 2];
  typename P::ab_t al[NT];
  __device__ __forceinline__ void load(const typename P::ab_t* gal, const float* sqf, int i, int j) {
#pragma unroll
    for (int p = 0; p < NT; ++p) {
#pragma unroll
      for (int fp = 0; fp < (NF + 1) / 2; ++fp) bq[p][fp] = qcol<NF>(sqf, i + p, fp, j);
      al[p] = gal[i + p + 2];
    }
  }
};

// NT l-values of chains [K0, K) from l on, with operands already loaded.
// The recurrence of the whole tile runs first and its products follow as
// one group: on gfx950 the block products and the VALU do not overlap, and
// every switch between the two costs extra issue time (measured: 8 products
// and 24 FMAs take 94 ns interleaved, 71 ns grouped, per SIMD).
template <int NT, int K0, bool INJ, int NF, int K, class P>
__device__ __forceinline__ void synth_tile_m4(const TileOps<NT, NF, P>& o, int l, ChainT<P> (&c)[K],
                                              f4v (&accE)[K][(NF + 1) / 2],
                                              f4v (&accO)[K][(NF + 1) / 2]) {
  constexpr int FP = (NF + 1) / 2;
  float zv[NT][K];
#pragma unroll
  for (int p = 0; p < NT; p += 2) {
    if constexpr (INJ) {
      if (p % kTile == 0) {
#pragma unroll
        for (int k = K0; k < K; ++k)
          if (c[k].ls == l + p) {
            c[k].y0 = c[k].s0;
            c[k].y1 = c[k].s1;
          }
      }
    }
#pragma unroll
    for (int k = K0; k < K; ++k) {
      zv[p][k] = (float)c[k].y0;
      zv[p + 1][k] = (float)c[k].y1;
      P::advance(o.al[p], o.al[p + 1], c[k].ct, c[k].y0, c[k].y1);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int p = 0; p < NT; p += 2)
#pragma unroll
    for (int k = K0; k < K; ++k)
#pragma unroll
      for (int fp = 0; fp < FP; ++fp) {
        accE[k][fp] = mf441(zv[p][k], o.bq[p][fp], accE[k][fp]);
        accO[k][fp] = mf441(zv[p + 1][k], o.bq[p + 1][fp], accO[k][fp]);
      }
  __builtin_amdgcn_sched_barrier(0);
}

template <int NT, int K0, bool INJ, int NF, int K, class P>
__device__ __forceinline__ void synth_tile_m4_at(const typename P::ab_t* gal, const float* sqf,
                                                 int m, int l, int j, ChainT<P> (&c)[K],
                                                 f4v (&accE)[K][(NF + 1) / 2],
                                                 f4v (&accO)[K][(NF + 1) / 2]) {
  TileOps<NT, NF, P> o;
  o.load(gal, sqf, l - m, j);
  synth_tile_m4<NT, K0, INJ, NF, K, P>(o, l, c, accE, accO);
}

// The phase structure of synth_phases (chains [K0, K) until a lower chain
// becomes live) around synth_tile_m4: 8-l tiles where they fit before the
// phase's end, a 4-l tile for the rest (phase ends are 4-aligned in l - m).
template <int K0, int NF, int K, class P>
__device__ __forceinline__ void synth_phases_m4(int l, int L, int m, int j, const int (&lk)[K],
                                                const int (&lz)[K], const typename P::ab_t* gal,
                                                const float* sqf, ChainT<P> (&c)[K],
                                                f4v (&accE)[K][(NF + 1) / 2],
                                                f4v (&accO)[K][(NF + 1) / 2]) {
  int lend = L + 1, injend = 0;
#pragma unroll
  for (int k = 0; k < K0; ++k) lend = min(lend, lk[k]);
#pragma unroll
  for (int k = K0; k < K; ++k) injend = max(injend, lz[k] + 1);
  const int l1 = min(lend, injend);
  while (l < l1) {
    if (l + 8 <= lend) {
      synth_tile_m4_at<8, K0, true, NF, K, P>(gal, sqf, m, l, j, c, accE, accO);
      l += 8;
    } else {
      synth_tile_m4_at<4, K0, true, NF, K, P>(gal, sqf, m, l, j, c, accE, accO);
      l += 4;
    }
  }
  for (; l + 8 <= lend; l += 8)
    synth_tile_m4_at<8, K0, false, NF, K, P>(gal, sqf, m, l, j, c, accE, accO);
  for (; l < lend; l += 4)
    synth_tile_m4_at<4, K0, false, NF, K, P>(gal, sqf, m, l, j, c, accE, accO);
  if constexpr (K0 > 0)
    synth_phases_m4<K0 - 1, NF, K, P>(l, L, m, j, lk, lz, gal, sqf, c, accE, accO);
}

// Synthesis: per chain and field pair, D_E / D_O (E: even l - m, O: odd)
// accumulate P[pair][l] * Q[l][col] over l with the row's 4 pairs of a block
// as rows; the lane of column j ends with the sums of the 4 pairs of its block.
// The 4-field, one-chain FP32 form compiles to 100 VGPRs (4 waves per SIMD);
// held to 96 it takes 5, which its LDS (alpha and 4 fields' Q per row) allows.
template <int NF, int K, class P>
constexpr int m4_synth_waves_per_eu() {
  return std::is_same<typename P::ab_t, float>::value && NF == 4 && K == 1 ? 5 : 1;
}

template <int NF, int K, class P, bool IMEX = false>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(m4_synth_waves_per_eu<NF, K, P>())))
void k_leg_synth_m4(typename P::Tables T, int nfield, Fields Fq, float2* __restrict__ fm, int rows,
                    Fields Fa, const float* lam, ImexCoef ic) {
  constexpr int FP = (NF + 1) / 2;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  const int L = T.lmax, M = T.mmax, nlat = T.nlat;
  auto* sab = reinterpret_cast<typename P::ab_t*>(lds_raw);
  float2* sq = reinterpret_cast<float2*>(sab + (L + 1 + kPadQ));
  const float* sqf = reinterpret_cast<const float*>(sq);
  const int f0 = blockIdx.y * NF;
  const int lane = threadIdx.x & 63, j = lane & 3, blk = lane >> 2;

  for (int r = 0; r < 2; ++r) {
    const int m = row_m(r, rows, M);
    if (m < 0) break;
    __syncthreads();  // the previous row's readers are done
    if constexpr (IMEX)
      stage_row_imex<NF, P>(T, m, f0, nfield, Fq, Fa, lam, ic, sab, sq, kPadQ);
    else
      stage_row_t<NF, P>(T, m, f0, nfield, Fq, true, sab, sq, kPadQ);
    __syncthreads();

    ChainT<P> c[K];
    int lk[K], lz[K];
    int lmin = 0x7fffffff;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      load_chain_t<P>(T, m, pair_of(k), c[k]);
      lk[k] = wave_min(c[k].ls);
      lz[k] = last_seed(c[k].ls, L);
      lmin = min(lmin, lk[k]);
    }

    f4v accE[K][FP], accO[K][FP];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int fp = 0; fp < FP; ++fp) accE[k][fp] = accO[k][fp] = f4v{0.f, 0.f, 0.f, 0.f};

    if (lmin <= L) synth_phases_m4<K - 1, NF, K, P>(lmin, L, m, j, lk, lz, sab, sqf, c, accE, accO);

#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int pb = pair_of(k) - lane + 4 * blk;  // pair of block row 0
#pragma unroll
      for (int fp = 0; fp < FP; ++fp) {
        const int f = 2 * fp + (j >> 1);
        if (f >= NF || f0 + f >= nfield) continue;
        // fm row-major per latitude: (f, ilat, m) at (f nlat + ilat)(M+1) + m
        float* out = reinterpret_cast<float*>(fm + (long long)(f0 + f) * nlat * (M + 1) + m) +
                     (j & 1);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int jp = pb + q;
          if (jp >= T.npair) continue;
          const float e = accE[k][fp][q], o = accO[k][fp][q];
          out[2LL * jp * (M + 1)] = e + o;
          out[2LL * (nlat - 1 - jp) * (M + 1)] = e - o;
        }
      }
    }
  }
}

// Analysis: Q[l][col] = sum_pairs P[pair][l] W[pair][col].  The sum runs
// over latitude pairs, which are the lanes, so each chain's 16 values of a
// span go through a wave-private LDS tile (row = l, column = pair) and come
// back as 16-byte loads of 4 consecutive pairs at one l.  Block b is an
// l-group lg (4 values of l - m of one parity: offsets {0,2,4,6}, {8,..,14},
// {1,..,7}, {9,..,15} of the span) times a pair-group pg (16 pairs of the
// chain): its lane r supplies P[pair][l_r] as A, its lane j the weight
// W[pair][col j] of the block's parity as B ((G_N + G_S) w even, (G_N - G_S) w
// odd), the pair running over the 16 of pg in turn.  Register r of lane
// 4b + j then sums P W over the pair-group for l_r, so a span ends with a sum
// over the 4 pair-groups (lane bits 4, 5: two swaps) and the waves' sums
// meeting in LDS.  Seeds sit at l - m = 0 mod 4, so a span's offsets keep the
// parity of l - m.  Lane bits 2, 3 map to l-groups (0, 2, 1, 3), and the
// tile stores offset p in physical row m4_row(p), so that block q of a
// pair-group reads physical rows 4q .. 4q + 3: with rows 68 floats apart, the
// 16-byte load of lane 16 pg + 4 q + j starts in 4-bank group
// (4 q + j + 4 pg + t) mod 16, distinct over any 16 lanes that differ in
// (q, j) or in (j, pg), whether an LDS pass takes 16 consecutive lanes or
// 4 lanes from each 16.  (The bank-conflict counter reads 4 cycles per
// 16-byte operand load under this placement and the earlier one alike: the
// passes of a full-wave 16-byte load, not avoidable conflicts.)
constexpr int M4_SPAN = 16;
// spans of the block analysis per workgroup barrier: two keep the
// workgroup's LDS (tiles, alpha, the waves' sums) within a third of the CU's
// 160 KB at lmax 1023 with FP64 alpha
constexpr int M4_NH = 2;
constexpr int M4_TROW = 64 + 4;  // floats per l row of the tile
constexpr int M4_WROW = 64 + 2;  // floats per column row of the row's staged weights
static_assert(8 * 2 * M4_WROW <= M4_SPAN * M4_TROW, "two field pairs' weights fit a tile");

__device__ __forceinline__ int m4_lgroup(int blk) { return ((blk & 1) << 1) | ((blk >> 1) & 1); }
// l offset in the span of row r of l-group g
__device__ __forceinline__ int m4_off(int g, int r) { return (g >> 1) + 2 * r + 8 * (g & 1); }
// physical tile row of span offset p: m4_row(m4_off(m4_lgroup(q), r)) = 4 q + r
__host__ __device__ constexpr int m4_row(int p) { return (p & 8) | ((p & 1) << 2) | ((p & 7) >> 1); }
static_assert(m4_row(0) == 0 && m4_row(2) == 1 && m4_row(1) == 4 && m4_row(7) == 7 &&
                  m4_row(8) == 8 && m4_row(14) == 11 && m4_row(9) == 12 && m4_row(15) == 15,
              "block q's rows are physical rows 4q .. 4q + 3");

// Three 8-wave workgroups per CU (6 waves per SIMD) need at most 80 VGPRs and
// a third of the CU's 160 KB of LDS per workgroup: the FP32-recurrence,
// one-chain, one-field-pair form is held to that register count (measured ~7%
// faster than at 84 VGPRs and two workgroups); the FP64 recurrence and wider
// forms would spill under it.
template <int NF, int K, class P>
constexpr int m4_analys_waves_per_eu() {
  return std::is_same<typename P::ab_t, float>::value && K == 1
             ? ((NF + 1) / 2 == 1 ? 6 : (NF + 1) / 2 == 2 ? 5 : 1)
             : 1;
}

// IMEX: Fq is the state, and each analysed real coefficient a of field f
// goes straight into its implicit step (s1, s2 = coef[2 f], coef[2 f + 1];
// lam per real coefficient) instead of to memory
template <int NF, int K, class P, bool IMEX = false>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(m4_analys_waves_per_eu<NF, K, P>())))
void k_leg_analys_m4(typename P::Tables T, int nfield, const float2* __restrict__ fm, Fields Fq,
                     int rows, const float* __restrict__ lam, const float* __restrict__ coef) {
  constexpr int FP = (NF + 1) / 2;
  constexpr int NRED = M4_NH * 64 * FP;  // a wave's sums between two barriers
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  __shared__ int smin[8];
  const int L = T.lmax, M = T.mmax, nlat = T.nlat;
  auto* sab = reinterpret_cast<typename P::ab_t*>(lds_raw);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nthr = blockDim.x, nwaves = nthr >> 6;
  const int j = lane & 3, blk = lane >> 2;
  const int lg = m4_lgroup(blk & 3), pg = blk >> 2;
  const bool odd = lg >> 1;  // this block's parity of l - m
  float* const tiles = reinterpret_cast<float*>(sab + analys_ab_len(L));
  float* tile = tiles + wv * (K * M4_SPAN * M4_TROW);  // one tile per chain
  // the waves' sums, [buffer][wave][NRED], after the tiles: sized by the
  // launch's wave count, so a 4-wave workgroup takes half of 8 waves' space
  float* const red = tiles + nwaves * (K * M4_SPAN * M4_TROW);
  const int f0 = blockIdx.y * NF;

  for (int r = 0; r < 2; ++r) {
    const int m = row_m(r, rows, M, false);
    if (m < 0) break;
    const long long base = lm_base(L, m);
    __syncthreads();
    stage_row_t<NF, P>(T, m, f0, nfield, Fq, false, sab, nullptr, kPadAnalys);

    ChainT<P> c[K];
    int lk[K];
    int lzmax = -1, lmin = 0x7fffffff;
    // B operands for the whole row: each lane computes W of its own pair
    // (every column, both parities) into the chain's tile, then gathers the
    // 16 of its block: pair 16 pg + t, column j, the block's parity.  The
    // tile holds them as rows q = (parity FPS + fp) 4 + column of M4_WROW
    // floats, one per pair: the writes are consecutive, and the gather's
    // 32 distinct addresses (q of the block, 16 pg + t) fall in 32 banks.
    // A tile holds two field pairs' weights: more go through it in passes.
    constexpr int FPS = FP < 2 ? FP : 2;
    float W[K][16][FP];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      load_chain_t<P>(T, m, pair_of(k), c[k]);
      lk[k] = wave_min(c[k].ls);
      lzmax = max(lzmax, last_seed(c[k].ls, L));
      lmin = min(lmin, lk[k]);
    }
#pragma unroll
    for (int fh = 0; fh < FP; fh += FPS) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int jp = pair_of(k);
        const int jc = min(jp, T.npair - 1);
        const int js = nlat - 1 - jc;
        const float w = jp < T.npair ? T.w[jc] : 0.f;
        float* tw = tile + k * (M4_SPAN * M4_TROW) + lane;  // [parity][fp][col][pair]
#pragma unroll
        for (int f = 2 * fh; f < 2 * (fh + FPS); ++f) {
          const bool live = f < NF && f0 + f < nfield;
          const int fld = min(f0 + min(f, NF - 1), nfield - 1);
          const float2* g = fm + (long long)fld * nlat * (M + 1) + m;
          const float2 gN = g[(long long)jc * (M + 1)], gS = g[(long long)js * (M + 1)];
          float4 we{0.f, 0.f, 0.f, 0.f};
          if (live) {
            if (js == jc)  // equator of an odd grid pairs with itself
              we = make_float4(gN.x * w, gN.y * w, 0.f, 0.f);
            else
              we = make_float4((gN.x + gS.x) * w, (gN.y + gS.y) * w, (gN.x - gS.x) * w,
                               (gN.y - gS.y) * w);
          }
          const int fq = (f >> 1) - fh, c0 = 2 * (f & 1);
          tw[((0 * FPS + fq) * 4 + c0) * M4_WROW] = we.x;
          tw[((0 * FPS + fq) * 4 + c0 + 1) * M4_WROW] = we.y;
          tw[((1 * FPS + fq) * 4 + c0) * M4_WROW] = we.z;
          tw[((1 * FPS + fq) * 4 + c0 + 1) * M4_WROW] = we.w;
        }
      }
      wave_sync_lds();
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int t = 0; t < 16; ++t)
#pragma unroll
          for (int fq = 0; fq < FPS; ++fq)
            W[k][t][fh + fq] = tile[k * (M4_SPAN * M4_TROW) +
                                    (((odd ? FPS : 0) + fq) * 4 + j) * M4_WROW + 16 * pg + t];
      wave_sync_lds();  // the tiles are rewritten by the next pass or the spans
    }
    if (lane == 0) smin[wv] = lmin;
    __syncthreads();  // also publishes the staged row
    int lblk = smin[0];
    for (int w = 1; w < nwaves; ++w) lblk = min(lblk, smin[w]);
    lblk = min(lblk, L + 1);

    // coefficients below every lane's significance point are exactly 0 (the
    // implicit step with a = 0 there)
    for (int i = m + tid; i < lblk; i += nthr)
#pragma unroll
      for (int f = 0; f < NF; ++f)
        if (f0 + f < nfield) {
          float2* const X =
              reinterpret_cast<float2*>(const_cast<char*>(Fq.at(f0 + f))) + base + (i - m);
          if constexpr (IMEX)
            imex_apply(X, make_float2(0.f, 0.f),
                       reinterpret_cast<const float2*>(lam)[base + (i - m)], coef + 2 * (f0 + f));
          else
            *X = make_float2(0.f, 0.f);
        }

    // Spans go M4_NH at a time between workgroup barriers: each wave reduces
    // its half-spans into red, then all threads combine the waves' sums.
    int buf = 0;
    for (int ls0 = lblk; ls0 <= L; ls0 += M4_NH * M4_SPAN) {
#pragma unroll
      for (int h = 0; h < M4_NH; ++h) {
        const int ls = ls0 + h * M4_SPAN;
        float* rw = red + (buf * nwaves + wv) * NRED + h * 64 * FP;
        if (lmin >= ls + M4_SPAN || ls > L) {
          // every chain of this wave is still exactly 0 in this span (or the
          // span is past lmax)
#pragma unroll
          for (int fp = 0; fp < FP; ++fp) rw[lane * FP + fp] = 0.f;
          continue;
        }
        // 1) each live chain's 16 values of the span -> column `lane` of its
        // tile.  The span's alpha are read first, in one batch: read between
        // the tile stores, each would wait for the stores before it.
        typename P::ab_t al[M4_SPAN];
#pragma unroll
        for (int p = 0; p < M4_SPAN; ++p) al[p] = sab[ls - m + p + 2];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (lk[k] >= ls + M4_SPAN) continue;  // chain still 0 (wave-uniform)
          float* tk = tile + k * (M4_SPAN * M4_TROW) + lane;
#pragma unroll
          for (int p = 0; p < M4_SPAN; p += 2) {
            const int l = ls + p;
            if (p % kTile == 0 && ls <= lzmax && c[k].ls == l) {
              c[k].y0 = c[k].s0;
              c[k].y1 = c[k].s1;
            }
            tk[m4_row(p) * M4_TROW] = (float)c[k].y0;
            tk[m4_row(p + 1) * M4_TROW] = (float)c[k].y1;
            P::advance(al[p], al[p + 1], c[k].ct, c[k].y0, c[k].y1);
          }
        }
        wave_sync_lds();
        // 2) block products over the pair-group's 16 pairs of each live chain
        f4v acc[FP];
#pragma unroll
        for (int fp = 0; fp < FP; ++fp) acc[fp] = f4v{0.f, 0.f, 0.f, 0.f};
        const int arow = (4 * (blk & 3) + j) * M4_TROW + 16 * pg;  // lane j: row j of lg
        // the products as one group (no VALU work scheduled between them: on
        // gfx950 they do not overlap it, and every switch costs issue time)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (lk[k] >= ls + M4_SPAN) continue;
          const float* tk = tile + k * (M4_SPAN * M4_TROW) + arow;
          float4 a4[4];
#pragma unroll
          for (int t = 0; t < 4; ++t) a4[t] = *reinterpret_cast<const float4*>(tk + 4 * t);
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const float a[4] = {a4[t].x, a4[t].y, a4[t].z, a4[t].w};
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
              for (int fp = 0; fp < FP; ++fp)
                acc[fp] = mf441(a[u], W[k][4 * t + u][fp], acc[fp]);
          }
        }
        __builtin_amdgcn_sched_barrier(0);
        wave_sync_lds();  // the tiles are rewritten in the next span
        // 3) sum over the 4 pair-groups: lane bit 5 keeps rows {0,1} / {2,3},
        // lane bit 4 row {0} / {1} of those: lane 4b + j ends with row pg
#pragma unroll
        for (int fp = 0; fp < FP; ++fp) {
          float v[4] = {acc[fp][0], acc[fp][1], acc[fp][2], acc[fp][3]};
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            const auto s2 = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[i]),
                                                             __float_as_uint(v[i + 2]), false, false);
            v[i] = __uint_as_float(s2[0]) + __uint_as_float(s2[1]);
          }
          const auto s1 = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[0]),
                                                           __float_as_uint(v[1]), false, false);
          rw[lane * FP + fp] = __uint_as_float(s1[0]) + __uint_as_float(s1[1]);
        }
      }
      __syncthreads();
      // 4) the waves' sums, in wave order, one output per thread
      for (int o = tid; o < NRED; o += nthr) {
        const int fp = o % FP, ln = (o / FP) & 63, h = o / (64 * FP);
        const int jj = ln & 3, bb = ln >> 2;
        const int l = ls0 + h * M4_SPAN + m4_off(m4_lgroup(bb & 3), bb >> 2);
        const int f = 2 * fp + (jj >> 1);
        if (l <= L && f < NF && f0 + f < nfield) {
          const float s = sum_waves([&](int w) { return red[(buf * nwaves + w) * NRED + o]; },
                                    nwaves);
          float* dst = reinterpret_cast<float*>(const_cast<char*>(Fq.at(f0 + f))) +
                       2 * (base + (l - m));
          const float a = s * T.cz[base + (l - m)];
          if constexpr (IMEX) {
            const float* sc = coef + 2 * (f0 + f);
            dst[jj & 1] =
                imex_step(dst[jj & 1], sc[0], a, sc[1], lam[2 * (base + (l - m)) + (jj & 1)]);
          } else {
            dst[jj & 1] = a;
          }
        }
      }
      buf ^= 1;
    }
  }
}

// ------------------------------------------------------- the spectral op
// Legendre analysis of step n, the diagonal implicit solve and the Legendre
// synthesis of step n + 1 in one pass over each m-row: the recurrence value
// of every (l, m, latitude pair) is generated once and used by both
// contractions.  Per row, as k_leg_analys_m4 up to the waves' combined sums
// A(l) of a group of spans; then, per coefficient (bench/shtns/spec.h:
// 166-169's implicit step, the reference models' last lines)
//   X_new = (X + s1 A) / (1 + s2 lam)     (each product and sum rounded as
//                                          written: the generated kernel's
//                                          arithmetic)
// written to the state and, scaled by c_l, to LDS; then the synthesis block
// products of those l with the chain values the wave kept in registers (the
// synthesis of k_leg_synth_m4, in the same order over l).  The spans run as a
// pipeline with one workgroup barrier each: iteration i takes span i's
// analysis up to the waves' sums, the barrier, then the combine and implicit
// step of span i (the first 64 FP threads, whose state, lam and c_l loads
// were issued before the analysis) beside every wave's synthesis products of
// span i - 1 (solved in the previous iteration, published by this barrier).
// The row's Fourier scratch entries are read at the row's start and its
// synthesis written at its end, so the scratch is updated in place.
// Two 8-wave workgroups per CU need at most 128 VGPRs: the one-chain,
// two-field form is held to that, without spills in either recurrence form
// (the FP64 form reads its alpha four at a time and its tile operands where
// used).
// Measured at lmax 1023 (two fields, ~110 us): the recurrence arithmetic is
// not on the critical path (removing it leaves the time unchanged), nor are
// the combine's loads; the matrix products, the tile stores and the span
// barrier are (one more barrier per span: +15 us).  Spans grouped four to a
// barrier, with the synthesis regenerating its recurrence values instead of
// keeping them in registers, measured 130 us: the waves lose the overlap of
// one span's synthesis with the next span's analysis.
template <int NF, int K, class P>
constexpr int m4_spectral_waves_per_eu() { return NF == 2 && K == 1 ? 4 : 1; }

// a[f] for a lane-varying f < NF, as a select over the entries (which stay
// in scalar registers, where a dynamic index would load from memory)
template <int NF, class T>
__device__ __forceinline__ T pick(const T (&a)[4], int f) {
  T v = a[0];
#pragma unroll
  for (int i = 1; i < NF; ++i) v = f == i ? a[i] : v;
  return v;
}

template <int NF, int K, class P>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(m4_spectral_waves_per_eu<NF, K, P>())))
void k_leg_spectral_m4(typename P::Tables T, int nfield, float2* fm, Fields St,
                       const float* __restrict__ lam, ImexCoef ic, int rows) {
  constexpr int FP = (NF + 1) / 2;
  constexpr int NO = 64 * FP;  // combined sums per span
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  __shared__ float red[2][8][NO];
  // c_l X_new of a span, (species, re/im) column-major over the span's l:
  // a lane's B operands of its column are 16 consecutive floats
  __shared__ __attribute__((aligned(16))) float qs[2][4 * FP][M4_SPAN];
  __shared__ int smin[8];
  const int L = T.lmax, M = T.mmax, nlat = T.nlat;
  auto* sab = reinterpret_cast<typename P::ab_t*>(lds_raw);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nthr = blockDim.x, nwaves = nthr >> 6;
  const int j = lane & 3, blk = lane >> 2;
  const int lg = m4_lgroup(blk & 3), pg = blk >> 2;
  const bool odd = lg >> 1;
  float* tile = reinterpret_cast<float*>(sab + analys_ab_len(L)) +
                wv * (K * M4_SPAN * M4_TROW);
  // this thread's coefficient of a span's combine (threads below NO); its
  // field's scalars are selected from the argument registers with pick (a
  // dynamic index would re-read them from argument memory after every
  // barrier: the state stores between may alias it as far as the compiler
  // knows)
  const int cfp = tid % FP, cln = (tid / FP) & 63;
  const int cjj = cln & 3, cbb = cln >> 2;
  const int clo = m4_off(m4_lgroup(cbb & 3), cbb >> 2);
  const int cf = 2 * cfp + (cjj >> 1);
  const bool cthr = tid < NO && cf < NF;
  const bool cfield = cthr && cf < nfield;
  float* const cX = cfield ? reinterpret_cast<float*>(const_cast<char*>(St.at(ic.spc[cf])))
                           : nullptr;

  for (int r = 0; r < 2; ++r) {
    const int m = row_m(r, rows, M, false);
    if (m < 0) break;
    const long long base = lm_base(L, m);
    __syncthreads();
    stage_row_t<NF, P>(T, m, 0, nfield, Fields{}, false, sab, nullptr, kPadAnalys);

    ChainT<P> c[K];
    int lk[K];
    int lzmax = -1, lmin = 0x7fffffff;
    float W[K][16][FP];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      load_chain_t<P>(T, m, pair_of(k), c[k]);
      lk[k] = wave_min(c[k].ls);
      lzmax = max(lzmax, last_seed(c[k].ls, L));
      lmin = min(lmin, lk[k]);
      const int jp = pair_of(k);
      const int jc = min(jp, T.npair - 1);
      const int js = nlat - 1 - jc;
      const float w = jp < T.npair ? T.w[jc] : 0.f;
      float* tw = tile + k * (M4_SPAN * M4_TROW) + lane;  // [parity][fp][col][pair]
#pragma unroll
      for (int f = 0; f < 2 * FP; ++f) {
        const bool live = f < NF && f < nfield;
        const int fld = min(min(f, NF - 1), nfield - 1);
        const float2* g = fm + (long long)fld * nlat * (M + 1) + m;
        const float2 gN = g[(long long)jc * (M + 1)], gS = g[(long long)js * (M + 1)];
        float4 we{0.f, 0.f, 0.f, 0.f};
        if (live) {
          if (js == jc)
            we = make_float4(gN.x * w, gN.y * w, 0.f, 0.f);
          else
            we = make_float4((gN.x + gS.x) * w, (gN.y + gS.y) * w, (gN.x - gS.x) * w,
                             (gN.y - gS.y) * w);
        }
        const int fp = f >> 1, c0 = 2 * (f & 1);
        tw[((0 * FP + fp) * 4 + c0) * M4_WROW] = we.x;
        tw[((0 * FP + fp) * 4 + c0 + 1) * M4_WROW] = we.y;
        tw[((1 * FP + fp) * 4 + c0) * M4_WROW] = we.z;
        tw[((1 * FP + fp) * 4 + c0 + 1) * M4_WROW] = we.w;
      }
    }
    wave_sync_lds();
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int t = 0; t < 16; ++t)
#pragma unroll
        for (int fp = 0; fp < FP; ++fp)
          W[k][t][fp] = tile[k * (M4_SPAN * M4_TROW) +
                             (((odd ? FP : 0) + fp) * 4 + j) * M4_WROW + 16 * pg + t];
    wave_sync_lds();
    if (lane == 0) smin[wv] = lmin;
    __syncthreads();  // also publishes the staged row; every W is read
    int lblk = smin[0];
    for (int w = 1; w < nwaves; ++w) lblk = min(lblk, smin[w]);
    lblk = min(lblk, L + 1);

    // below every lane's significance point the analysed values are exactly
    // 0: the implicit step alone (and no synthesis contribution)
    for (int o = tid; o < (lblk - m) * 2 * NF; o += nthr) {
      const int f = o % NF, q = o / NF;
      if (f >= nfield) continue;
      const long long i = 2 * base + q;
      float* X = reinterpret_cast<float*>(const_cast<char*>(St.at(ic.spc[f])));
      X[i] = imex_step(X[i], ic.s1[f], 0.0f, ic.s2[f], lam[i]);
    }

    f4v accE[K][FP], accO[K][FP];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int fp = 0; fp < FP; ++fp) accE[k][fp] = accO[k][fp] = f4v{0.f, 0.f, 0.f, 0.f};

    // span ls's analysis into this wave's sums red[b][wv] and the chain
    // values zc; false when every chain of the wave is still 0 there.  INJ:
    // lanes may take their seed in this span (spans past the chains' last
    // seed run without the test)
    auto analyse = [&](int ls, int b, float (&zc)[K][M4_SPAN], auto inj) -> bool {
      constexpr bool INJ = decltype(inj)::value;
      float* rw = &red[b][wv][0];
      if (lmin >= ls + M4_SPAN) {
#pragma unroll
        for (int fp = 0; fp < FP; ++fp) rw[lane * FP + fp] = 0.f;
        return false;
      }
      // the span's alpha in batches read ahead of the tile stores: all 16 of
      // the f32 form at once, the f64 form's four at a time (16 doubles would
      // take 32 VGPRs of a kernel held to 128; with 4 it spills 28 bytes per
      // lane instead of 120, and its step measured 152 -> 140 us at lmax 1023 with two fields)
      constexpr int AB = sizeof(typename P::ab_t) == 8 ? M4_SPAN / 4 : M4_SPAN;
#pragma unroll
      for (int pb = 0; pb < M4_SPAN; pb += AB) {
        typename P::ab_t al[AB];
#pragma unroll
        for (int p = 0; p < AB; ++p) al[p] = sab[ls - m + pb + p + 2];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          float* tk = tile + k * (M4_SPAN * M4_TROW) + lane;
#pragma unroll
          for (int p = pb; p < pb + AB; p += 2) {
            const int l = ls + p;
            if (INJ && p % kTile == 0 && ls <= lzmax && c[k].ls == l) {
              c[k].y0 = c[k].s0;
              c[k].y1 = c[k].s1;
            }
            zc[k][p] = (float)c[k].y0;
            zc[k][p + 1] = (float)c[k].y1;
            tk[m4_row(p) * M4_TROW] = zc[k][p];
            tk[m4_row(p + 1) * M4_TROW] = zc[k][p + 1];
            P::advance(al[p - pb], al[p - pb + 1], c[k].ct, c[k].y0, c[k].y1);
          }
        }
      }
      wave_sync_lds();
      f4v acc[FP];
#pragma unroll
      for (int fp = 0; fp < FP; ++fp) acc[fp] = f4v{0.f, 0.f, 0.f, 0.f};
      const int arow = (4 * (blk & 3) + j) * M4_TROW + 16 * pg;
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const float* tk = tile + k * (M4_SPAN * M4_TROW) + arow;
        // each 16-byte operand load issued one group ahead of the products
        // that use it (the fences keep that order), so a wave waits for one
        // LDS round trip per span here, not four; the FP64-recurrence form,
        // whose registers are full, loads each group where it is used
        constexpr bool AHEAD = sizeof(typename P::ab_t) == 4;
        float4 cur = *reinterpret_cast<const float4*>(tk);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          float4 nxt = cur;
          if (AHEAD && t < 3) nxt = *reinterpret_cast<const float4*>(tk + 4 * (t + 1));
          if (!AHEAD && t > 0) cur = *reinterpret_cast<const float4*>(tk + 4 * t);
          __builtin_amdgcn_sched_barrier(0);
          const float a[4] = {cur.x, cur.y, cur.z, cur.w};
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int fp = 0; fp < FP; ++fp) acc[fp] = mf441(a[u], W[k][4 * t + u][fp], acc[fp]);
          __builtin_amdgcn_sched_barrier(0);
          cur = nxt;
        }
      }
      __builtin_amdgcn_sched_barrier(0);
      wave_sync_lds();  // the tile is rewritten by the next span
#pragma unroll
      for (int fp = 0; fp < FP; ++fp) {
        float v[4] = {acc[fp][0], acc[fp][1], acc[fp][2], acc[fp][3]};
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const auto s2 = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[i]),
                                                           __float_as_uint(v[i + 2]), false, false);
          v[i] = __uint_as_float(s2[0]) + __uint_as_float(s2[1]);
        }
        const auto s1 = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[0]),
                                                         __float_as_uint(v[1]), false, false);
        rw[lane * FP + fp] = __uint_as_float(s1[0]) + __uint_as_float(s1[1]);
      }
      return true;
    };
    // synthesis block products of one span's l, in order of l
    auto synthesise = [&](const float (&zp)[K][M4_SPAN], int b) {
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int fp = 0; fp < FP; ++fp) {
        // the lane's column of the span as four 16-byte loads, each issued
        // one group of products ahead
        const float4* q4 = reinterpret_cast<const float4*>(&qs[b][4 * fp + j][0]);
        float4 cur = q4[0];
#pragma unroll
        for (int t = 0; t < M4_SPAN / 4; ++t) {
          float4 nxt = cur;
          if (t < M4_SPAN / 4 - 1) nxt = q4[t + 1];
          __builtin_amdgcn_sched_barrier(0);
          float bq[4] = {cur.x, cur.y, cur.z, cur.w};
          if (NF == 1 && j >= 2) bq[0] = bq[1] = bq[2] = bq[3] = 0.f;
#pragma unroll
          for (int u = 0; u < 4; u += 2)
#pragma unroll
            for (int k = 0; k < K; ++k) {
              accE[k][fp] = mf441(zp[k][4 * t + u], bq[u], accE[k][fp]);
              accO[k][fp] = mf441(zp[k][4 * t + u + 1], bq[u + 1], accO[k][fp]);
            }
          __builtin_amdgcn_sched_barrier(0);
          cur = nxt;
        }
      }
    };
    // one pipeline iteration: span ls's analysis, the barrier, its combine and
    // implicit step beside the synthesis of the span before it (if any)
    auto iteration = [&](int ls, int b, float (&zc)[K][M4_SPAN], const float (&zp)[K][M4_SPAN],
                         bool livep) -> bool {
      const int l = ls + clo;
      const bool has = cfield && l <= L;
      long long ci = 0;
      float xo = 0.f, lv = 0.f, cz = 0.f;
      if (has) {  // issued before the analysis, used after the barrier
        ci = 2 * (base + (l - m)) + (cjj & 1);
        xo = cX[ci];
        lv = lam[ci];
        cz = T.cz[base + (l - m)];
      }
      const bool live = analyse(ls, b, zc, std::true_type{});
      __syncthreads();  // the waves' sums of span ls; qs[b ^ 1] of the span before
      if (cthr) {
        float qv = 0.f;
        if (has) {
          // (the analysis kernel's batched sum_waves measured slower here:
          // this combine is not on the span's critical path)
          float s = red[b][0][tid];
          for (int w = 1; w < nwaves; ++w) s += red[b][w][tid];
          const float xn = imex_step(xo, pick<NF>(ic.s1, cf), s * cz, pick<NF>(ic.s2, cf), lv);
          cX[ci] = xn;
          qv = xn * cz;
        }
        qs[b][2 * (cfield ? pick<NF>(ic.spc, cf) : cf) + (cjj & 1)][clo] = qv;
      }
      if (livep) synthesise(zp, b ^ 1);
      return live;
    };
    float zA[K][M4_SPAN], zB[K][M4_SPAN];
    bool liveA = false, liveB = false, lastA = false;
    // spans two at a time, the chain values alternating between zA and zB
    for (int ls = lblk; ls <= L; ls += 2 * M4_SPAN) {
      liveA = iteration(ls, 0, zA, zB, liveB);
      lastA = true;
      if (ls + M4_SPAN > L) break;
      liveB = iteration(ls + M4_SPAN, 1, zB, zA, liveA);
      lastA = false;
    }
    if (lblk <= L) {
      __syncthreads();  // qs of the last span
      if (lastA) {
        if (liveA) synthesise(zA, 0);
      } else if (liveB) {
        synthesise(zB, 1);
      }
    }

    // the synthesis of the row into the Fourier scratch (north = E + O,
    // south = E - O), species s in field slot s.  The offsets are formed
    // here from an opaque copy of the row stride: hoisted to the kernel's
    // start (as the compiler otherwise does) they would hold 16 VGPRs
    // through the span loop
    int M1 = M + 1;
    asm volatile("" : "+s"(M1));
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int pb = pair_of(k) - lane + 4 * blk;
#pragma unroll
      for (int fp = 0; fp < FP; ++fp) {
        const int s = 2 * fp + (j >> 1);
        if (s >= NF || s >= nfield) continue;
        float* out = reinterpret_cast<float*>(fm + (long long)s * nlat * M1 + m) + (j & 1);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int jp = pb + q;
          if (jp >= T.npair) continue;
          const float e = accE[k][fp][q], o = accO[k][fp][q];
          out[2LL * jp * M1] = e + o;
          out[2LL * (nlat - 1 - jp) * M1] = e - o;
        }
      }
    }
  }
}

template <int NF, int K, class P>
hipError_t spectral_m4(const typename P::Tables& T, int nfield, int nwaves, float2* fm,
                       const Fields& st, const float* lam, const ImexCoef& ic, hipStream_t s) {
  if (nwaves < (NF + 1) / 2 || nwaves > 8) return hipErrorInvalidValue;
  dim3 grid(T.mmax + 1, 1);
  const size_t lds = sizeof(typename P::ab_t) * analys_ab_len(T.lmax) +
                     sizeof(float) * (size_t)nwaves * K * M4_SPAN * M4_TROW;
  hipLaunchKernelGGL((k_leg_spectral_m4<NF, K, P>), grid, dim3(64 * nwaves), lds, s, T, nfield, fm,
                     st, lam, ic, 1);
  return hipGetLastError();
}

template <class P>
hipError_t spectral_launch(const typename P::Tables& T, int nfield, float2* fm, const Fields& st,
                           const float* lam, const ImexCoef& ic, hipStream_t s) {
  if (nfield < 1 || nfield > 4) return hipErrorInvalidValue;
  // The shape the block-matrix analysis takes by default for one or two
  // fields (resolve_leg_tile: up to 8 waves, then the fewest chains per lane
  // that cover the pairs), so the chunks are dealt to the same waves and the
  // waves' sums meet in the same order: a step through this kernel equals
  // its separate launches bit for bit at every pair count it takes.
  const int npair = T.npair;
  const int nw0 = std::min(8, (npair + 63) / 64);
  const int kk = (npair + 64 * nw0 - 1) / (64 * nw0);
  if (kk > 2) return hipErrorInvalidValue;
  const int k = kk <= 1 ? 1 : 2;
  // the combine takes one thread per sum of a span: 64 per field pair
  const int nwaves = std::max(nw0, nfield > 2 ? 2 : 1);
  if (nfield <= 2) {
    if (nfield == 1)
      return k == 1 ? spectral_m4<1, 1, P>(T, nfield, nwaves, fm, st, lam, ic, s)
                    : spectral_m4<1, 2, P>(T, nfield, nwaves, fm, st, lam, ic, s);
    return k == 1 ? spectral_m4<2, 1, P>(T, nfield, nwaves, fm, st, lam, ic, s)
                  : spectral_m4<2, 2, P>(T, nfield, nwaves, fm, st, lam, ic, s);
  }
  return k == 1 ? spectral_m4<4, 1, P>(T, nfield, nwaves, fm, st, lam, ic, s)
                : spectral_m4<4, 2, P>(T, nfield, nwaves, fm, st, lam, ic, s);
}

template <int NF, int K, class P, bool IMEX = false>
hipError_t synth_m4(const typename P::Tables& T, int nfield, int nwaves, int rows, const Fields& q,
                    float2* fm, hipStream_t s, const Fields& a = Fields{},
                    const float* lam = nullptr, const ImexCoef& ic = ImexCoef{}) {
  dim3 grid(rows == 1 ? T.mmax + 1 : T.mmax / 2 + 1, (nfield + NF - 1) / NF);
  hipLaunchKernelGGL((k_leg_synth_m4<NF, K, P, IMEX>), grid, dim3(64 * nwaves),
                     synth_lds<P>(T.lmax, NF), s, T, nfield, q, fm, rows, a, lam, ic);
  return hipGetLastError();
}

// The implicit step of every field in front of the block-matrix synthesis
// (the tile's waves, chains and rows; impl 3 only).
template <class P>
hipError_t imex_synth_launch(const typename P::Tables& T, int nfield, const LegTile& t,
                             const Fields& q, const Fields& a, const float* lam,
                             const ImexCoef& ic, float2* fm, hipStream_t s) {
#define TS_IMEX_K(NF)                                                                        \
  switch (t.k) {                                                                             \
    case 1: return synth_m4<NF, 1, P, true>(T, nfield, t.nwaves, t.rows, q, fm, s, a, lam, ic); \
    case 2: return synth_m4<NF, 2, P, true>(T, nfield, t.nwaves, t.rows, q, fm, s, a, lam, ic); \
    case 4: return synth_m4<NF, 4, P, true>(T, nfield, t.nwaves, t.rows, q, fm, s, a, lam, ic); \
    default: return hipErrorInvalidValue;                                                    \
  }
  switch (t.nf) {
    case 1: TS_IMEX_K(1)
    case 2: TS_IMEX_K(2)
    default: TS_IMEX_K(4)
  }
#undef TS_IMEX_K
}

template <int NF, int K, class P, bool IMEX = false>
hipError_t analys_m4(const typename P::Tables& T, int nfield, int nwaves, int rows,
                     const float2* fm, const Fields& q, hipStream_t s,
                     const float* lam = nullptr, const float* coef = nullptr) {
  dim3 grid(rows == 1 ? T.mmax + 1 : T.mmax / 2 + 1, (nfield + NF - 1) / NF);
  // alpha, one tile per wave and chain, the waves' sums (two buffers of
  // 2 spans x 64 lanes x field pairs per wave)
  const size_t lds = sizeof(typename P::ab_t) * analys_ab_len(T.lmax) +
                     sizeof(float) * (size_t)nwaves * K * M4_SPAN * M4_TROW +
                     sizeof(float) * (size_t)2 * nwaves * M4_NH * 64 * ((NF + 1) / 2);
  hipLaunchKernelGGL((k_leg_analys_m4<NF, K, P, IMEX>), grid, dim3(64 * nwaves), lds, s, T, nfield,
                     fm, q, rows, lam, coef);
  return hipGetLastError();
}

template <int NF, class P>
hipError_t synth_m4_nf(const typename P::Tables& T, int nfield, const LegTile& t, const Fields& q,
                       float2* fm, hipStream_t s) {
  switch (t.k) {
    case 1: return synth_m4<NF, 1, P>(T, nfield, t.nwaves, t.rows, q, fm, s);
    case 2: return synth_m4<NF, 2, P>(T, nfield, t.nwaves, t.rows, q, fm, s);
    case 4: return synth_m4<NF, 4, P>(T, nfield, t.nwaves, t.rows, q, fm, s);
    default: return hipErrorInvalidValue;
  }
}

template <int NF, class P, bool IMEX = false>
hipError_t analys_m4_nf(const typename P::Tables& T, int nfield, const LegTile& t,
                        const float2* fm, const Fields& q, hipStream_t s,
                        const float* lam = nullptr, const float* coef = nullptr) {
  switch (t.k) {
    case 1: return analys_m4<NF, 1, P, IMEX>(T, nfield, t.nwaves, t.rows, fm, q, s, lam, coef);
    case 2: return analys_m4<NF, 2, P, IMEX>(T, nfield, t.nwaves, t.rows, fm, q, s, lam, coef);
    case 4: return analys_m4<NF, 4, P, IMEX>(T, nfield, t.nwaves, t.rows, fm, q, s, lam, coef);
    default: return hipErrorInvalidValue;
  }
}

// Eight fields per workgroup (batched transforms: the recurrence shared by
// four field pairs), with one or two chains per lane.
template <class P>
hipError_t synth_m4_nf8(const typename P::Tables& T, int nfield, const LegTile& t,
                        const Fields& q, float2* fm, hipStream_t s) {
  switch (t.k) {
    case 1: return synth_m4<8, 1, P>(T, nfield, t.nwaves, t.rows, q, fm, s);
    case 2: return synth_m4<8, 2, P>(T, nfield, t.nwaves, t.rows, q, fm, s);
    default: return hipErrorInvalidValue;
  }
}

template <class P, bool IMEX = false>
hipError_t analys_m4_nf8(const typename P::Tables& T, int nfield, const LegTile& t,
                         const float2* fm, const Fields& q, hipStream_t s,
                         const float* lam = nullptr, const float* coef = nullptr) {
  switch (t.k) {
    case 1: return analys_m4<8, 1, P, IMEX>(T, nfield, t.nwaves, t.rows, fm, q, s, lam, coef);
    case 2: return analys_m4<8, 2, P, IMEX>(T, nfield, t.nwaves, t.rows, fm, q, s, lam, coef);
    default: return hipErrorInvalidValue;
  }
}

// ---------------------------- per-parity matrix products (tile impl 5)
// For launches of many fields.  Per m-row and parity of l - m both
// contractions are plain matrix products on v_mfma_f32_16x16x4_f32 (operand
// layout as for the 16x16x4 kernels above), with 16-column blocks that are
// (field, re/im) of 8 fields:
//   synthesis  D_par[pair][col] = sum_{l of par} z_l[pair] (c_l Q_l)[col]
//   analysis   D_par[l][col]    = sum_pairs z_l[pair] W_par[pair][col]
// Every column is live and every product adds a nonzero term (the 16x16x4
// kernels above carry both parities in their columns, so half their products
// are zero), and one recurrence value serves 8 CB fields (the 4x4 block
// kernels: 2 or 4).  One workgroup per m-row and group of NF = 8 CB fields;
// wave w runs the recurrence of pairs 64 w .. 64 w + 63 (lane = pair) and
// passes each span's values through a wave-private LDS tile that holds
// (z_l, z_{l+1}) of a pair side by side, so one 8-byte load gives a lane its
// operands of both parities.  Spans start at the row's first significant l,
// which is m + 4i, so span offset p has the parity of l - m.
constexpr int G16_SSPAN = 16;        // synthesis: l per span (two 4-l steps per parity)
constexpr int G16_ASPAN = 32;        // analysis: l per span (16 rows per parity)
constexpr int G16_STROW = 128 + 32;  // synthesis tile: floats per row (l-pair) of 64 pairs x 2
constexpr int G16_ATROW = 128 + 4;   // analysis tile: floats per row (l-pair)
// staged c_l Q_l of the synthesis: floats per row (l-pair) of 16 CB columns x
// 2 parities, padded so that the two rows a half-wave reads fall in distinct banks
__host__ __device__ constexpr int g16_qrow(int cb) { return cb == 1 ? 32 : 96; }

// The row's alpha shifted by two (sal[i] = alpha at l - m = i + 2), zero past
// the row: a span's alphas then start 16-byte aligned.
template <class P>
__device__ __forceinline__ void g16_stage_alpha(const typename P::Tables& T, int m,
                                                typename P::ab_t* sal) {
  const int n = T.lmax - m + 1;
  const long long base = lm_base(T.lmax, m);
  for (int i = threadIdx.x; i < analys_ab_len(T.lmax); i += blockDim.x)
    sal[i] = i + 2 < n ? T.ab[base + i + 2] : P::zero_ab();
}

// One span of SP values of this lane's chain from ls into its tile row
// (p / 2: (z_{ls+p}, z_{ls+p+1})), the seed taken at its position.
template <int SP, int TROW, class P>
__device__ __forceinline__ void g16_recur(ChainT<P>& c, int ls, int lz, int m,
                                          const typename P::ab_t* sal, float* tile) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int pb = 0; pb < SP; pb += 8) {
    typename P::ab_t al[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) al[q] = sal[ls - m + pb + q];
#pragma unroll
    for (int p = pb; p < pb + 8; p += 2) {
      if (p % kTile == 0 && ls <= lz && c.ls == ls + p) {
        c.y0 = c.s0;
        c.y1 = c.s1;
      }
      *reinterpret_cast<float2*>(tile + (p / 2) * TROW + 2 * lane) =
          make_float2((float)c.y0, (float)c.y1);
      P::advance(al[p - pb], al[p - pb + 1], c.ct, c.y0, c.y1);
    }
  }
}

// ROWS = 2 (8 fields): the workgroup takes the consecutive rows m0 = 2 q and
// m0 + 1 one after the other, keeps both rows' sums, and neighbouring lanes
// (re, im of one field) swap values so that each writes one (re, im) pair:
// a lane pair writes 16 contiguous bytes of (field, latitude), m0 and m0 + 1.
template <int CB, int ROWS, class P>
__global__ __launch_bounds__(512) void k_leg_synth_g16(typename P::Tables T, int nfield, Fields Fq,
                                                       float2* __restrict__ fm) {
  static_assert(ROWS == 1 || (ROWS == 2 && CB == 1), "two rows with 8 fields per workgroup");
  constexpr int NF = 8 * CB, SP = G16_SSPAN, QROW = g16_qrow(CB), TROW = G16_STROW;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  __shared__ int smin[8];
  const int L = T.lmax, M = T.mmax, nlat = T.nlat;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nthr = blockDim.x;
  const int nwaves = nthr >> 6;
  const int r16 = lane & 15, g = lane >> 4;
  auto* sal = reinterpret_cast<typename P::ab_t*>(lds_raw);
  float* const sq = reinterpret_cast<float*>(sal + analys_ab_len(L));  // [2][SP / 2][QROW]
  float* const tile = sq + 2 * (SP / 2) * QROW + wv * ((SP / 2) * TROW);
  const int f0 = blockIdx.y * NF;
  const int m0 = ROWS * row_m(0, 1, M, false);

  f4v acc[ROWS][2][4][CB];  // [row][parity][16-pair block][column block]
#pragma unroll
  for (int rr = 0; rr < ROWS; ++rr)
#pragma unroll
    for (int pr = 0; pr < 2; ++pr)
#pragma unroll
      for (int pb = 0; pb < 4; ++pb)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) acc[rr][pr][pb][cb] = f4v{0.f, 0.f, 0.f, 0.f};

#pragma unroll
  for (int rr = 0; rr < ROWS; ++rr) {
  const int m = m0 + rr;
  if (m > M) break;
  if (rr > 0) __syncthreads();  // the previous row's readers of alpha, c_l Q_l and tiles are done
  const long long base = lm_base(L, m);
  g16_stage_alpha<P>(T, m, sal);

  ChainT<P> c;
  load_chain_t<P>(T, m, 64 * wv + lane, c);
  const int lk = wave_min(c.ls), lz = last_seed(c.ls, L);
  if (lane == 0) smin[wv] = lk;
  __syncthreads();
  int lblk = smin[0];
  for (int w = 1; w < nwaves; ++w) lblk = min(lblk, smin[w]);

  // c_l Q_l of span [ls, ls + SP): row i holds l = ls + 2 i (even l - m) and
  // l + 1 (odd), column c = (field, re/im) at floats 2 c + parity, so a
  // field's four values of a row are one 16-byte store; zero past lmax
  auto stage = [&](int ls, float* d) {
    for (int it = tid; it < NF * (SP / 2); it += nthr) {
      const int f = it / (SP / 2), i = it % (SP / 2);
      const int l = ls + 2 * i;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (f0 + f < nfield && l <= L) {
        const float2* src = reinterpret_cast<const float2*>(Fq.at(f0 + f)) + base + (l - m);
        const float2 a = src[0];
        const float ca = T.cz[base + (l - m)];
        float2 b = make_float2(0.f, 0.f);
        float cb = 0.f;
        if (l + 1 <= L) {
          b = src[1];
          cb = T.cz[base + (l - m) + 1];
        }
        v = make_float4(a.x * ca, b.x * cb, a.y * ca, b.y * cb);
      }
      *reinterpret_cast<float4*>(d + i * QROW + 4 * f) = v;
    }
  };

  if (lblk <= L) stage(lblk, sq);
  int buf = 0;
  for (int ls = lblk; ls <= L; ls += SP) {
    __syncthreads();  // span ls staged; the other buffer's readers are done
    if (ls + SP <= L) stage(ls + SP, sq + (buf ^ 1) * ((SP / 2) * QROW));
    if (lk < ls + SP) {  // a lane of this wave is significant by the span's end
      g16_recur<SP, TROW, P>(c, ls, lz, m, sal, tile);
      wave_sync_lds();
      const float* d = sq + buf * ((SP / 2) * QROW);
#pragma unroll
      for (int ks = 0; ks < SP / 8; ++ks) {
        // k = g: l = ls + 8 ks + 2 g (+1 for the odd parity)
        float2 b[CB];
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
          b[cb] = *reinterpret_cast<const float2*>(d + (4 * ks + g) * QROW + 32 * cb + 2 * r16);
#pragma unroll
        for (int pb = 0; pb < 4; ++pb) {
          const float2 a =
              *reinterpret_cast<const float2*>(tile + (4 * ks + g) * TROW + 32 * pb + 2 * r16);
#pragma unroll
          for (int cb = 0; cb < CB; ++cb) {
            acc[rr][0][pb][cb] = mf4(a.x, b[cb].x, acc[rr][0][pb][cb]);
            acc[rr][1][pb][cb] = mf4(a.y, b[cb].y, acc[rr][1][pb][cb]);
          }
        }
      }
      wave_sync_lds();  // the tile is rewritten in the next span
    }
    buf ^= 1;
  }
  }

  // north = E + O, south = E - O; D row 4 g + r is pair 64 w + 16 pb + 4 g + r
  if constexpr (ROWS == 1) {
#pragma unroll
    for (int pb = 0; pb < 4; ++pb)
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) {
        const int f = f0 + 8 * cb + (r16 >> 1);
        if (f >= nfield) continue;
        float* out = reinterpret_cast<float*>(fm + (long long)f * nlat * (M + 1) + m0) + (r16 & 1);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int jp = 64 * wv + 16 * pb + 4 * g + r;
          if (jp >= T.npair) continue;
          const float e = acc[0][0][pb][cb][r], o = acc[0][1][pb][cb][r];
          out[2LL * jp * (M + 1)] = e + o;
          out[2LL * (nlat - 1 - jp) * (M + 1)] = e - o;
        }
      }
  } else {
    // lane 2 i holds re, lane 2 i + 1 im of one field for both rows: the
    // even lane takes (re, im) of row m0, the odd lane of row m0 + 1
    const int comp = r16 & 1;
    const int f = f0 + (r16 >> 1);
    float2* out = fm + (long long)min(f, nfield - 1) * nlat * (M + 1) + m0 + comp;
    const bool live = f < nfield && (comp == 0 || m0 + 1 <= M);
#pragma unroll
    for (int pb = 0; pb < 4; ++pb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int jp = 64 * wv + 16 * pb + 4 * g + r;
        const float e0 = acc[0][0][pb][0][r], o0 = acc[0][1][pb][0][r];
        const float e1 = acc[1][0][pb][0][r], o1 = acc[1][1][pb][0][r];
        const float nA = e0 + o0, sA = e0 - o0, nB = e1 + o1, sB = e1 - o1;
        const float gn = __shfl_xor(comp ? nA : nB, 1), gs = __shfl_xor(comp ? sA : sB, 1);
        if (!live || jp >= T.npair) continue;
        out[(long long)jp * (M + 1)] = comp ? make_float2(gn, nB) : make_float2(nA, gn);
        out[(long long)(nlat - 1 - jp) * (M + 1)] = comp ? make_float2(gs, sB) : make_float2(sA, gs);
      }
  }
}

// IMEX: Fq is the state, and each analysed coefficient a of field f goes
// straight into its implicit step (s1, s2 = coef[2 f], coef[2 f + 1]; lam
// per coefficient) instead of to memory
template <int CB, bool IMEX, class P>
__global__ __launch_bounds__(512) void k_leg_analys_g16(typename P::Tables T, int nfield,
                                                        const float2* __restrict__ fm, Fields Fq,
                                                        const float* __restrict__ lam,
                                                        const float* __restrict__ coef) {
  constexpr int NF = 8 * CB, SP = G16_ASPAN, TROW = G16_ATROW;
  constexpr int RS = (SP / 2) * TROW;  // floats of a wave's tile
  static_assert(2 * CB * 256 <= RS, "a wave's sums of a span fit its tile");
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  __shared__ int smin[8];
  const int L = T.lmax, M = T.mmax, nlat = T.nlat;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nthr = blockDim.x;
  const int nwaves = nthr >> 6;
  const int r16 = lane & 15, g = lane >> 4;
  auto* sal = reinterpret_cast<typename P::ab_t*>(lds_raw);
  float* const tiles = reinterpret_cast<float*>(sal + analys_ab_len(L));
  float* const tile = tiles + wv * RS;
  // a wave's sums of a span go to its own tile once its products have read
  // it, and the workgroup combines them from the tiles: no LDS of their own,
  // so that four 4-wave workgroups of 16 fields fit a CU (a separate buffer:
  // three per CU, 185 us for 64 fields at lmax 511; two buffers, one barrier
  // per span: two per CU, 232 us)
  const int f0 = blockIdx.y * NF;
  const int m = row_m(0, 1, M, false);
  const long long base = lm_base(L, m);
  g16_stage_alpha<P>(T, m, sal);

  ChainT<P> c;
  load_chain_t<P>(T, m, 64 * wv + lane, c);
  const int lk = wave_min(c.ls), lz = last_seed(c.ls, L);
  if (lane == 0) smin[wv] = lk;

  __syncthreads();  // smin and the staged alpha
  int lblk = smin[0];
  for (int w = 1; w < nwaves; ++w) lblk = min(lblk, smin[w]);
  lblk = min(lblk, L + 1);

  // coefficients below every lane's significance point are exactly 0 (the
  // implicit step with a = 0 there)
  const int nz = lblk - m;
  for (int o = tid; o < NF * nz; o += nthr) {
    const int f = o / nz, i = o - f * nz;
    if (f0 + f >= nfield) continue;
    float2* const X = reinterpret_cast<float2*>(const_cast<char*>(Fq.at(f0 + f))) + base + i;
    if constexpr (IMEX)
      imex_apply(X, make_float2(0.f, 0.f), reinterpret_cast<const float2*>(lam)[base + i],
                 coef + 2 * (f0 + f));
    else
      *X = make_float2(0.f, 0.f);
  }

  // B operands of the whole row: W_par of pair 64 w + 4 kp + g, column
  // 16 cb + r16 (W_E = (G_N + G_S) w, W_O = (G_N - G_S) w, leg.ts:207-219);
  // loaded after the row's barrier, so that their latency overlaps the first
  // span's recurrence instead of holding the barrier
  float W[2][16][CB];
#pragma unroll
  for (int kp = 0; kp < 16; ++kp)
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
      const int jp = 64 * wv + 4 * kp + g;
      const int f = f0 + 8 * cb + (r16 >> 1);
      float e = 0.f, o = 0.f;
      if (jp < T.npair && f < nfield) {
        const float* gp =
            reinterpret_cast<const float*>(fm + (long long)f * nlat * (M + 1) + m) + (r16 & 1);
        const int js = nlat - 1 - jp;
        const float w = T.w[jp];
        const float gN = gp[2LL * jp * (M + 1)], gS = gp[2LL * js * (M + 1)];
        if (js == jp) {  // equator of an odd grid pairs with itself
          e = gN * w;
        } else {
          e = (gN + gS) * w;
          o = (gN - gS) * w;
        }
      }
      W[0][kp][cb] = e;
      W[1][kp][cb] = o;
    }

  const int cf = tid >> 4, crow = tid & 15;  // this thread's first combine item
  const int cidx = (((cf >> 3) * 16 + crow) * 8 + ((cf & 7) ^ (crow & 7))) * 4;
  float2* const cq =
      reinterpret_cast<float2*>(const_cast<char*>(Fq.at(min(f0 + cf, nfield - 1)))) + base;
  for (int ls = lblk; ls <= L; ls += SP) {
    f4v acc[2][CB];
#pragma unroll
    for (int pr = 0; pr < 2; ++pr)
#pragma unroll
      for (int cb = 0; cb < CB; ++cb) acc[pr][cb] = f4v{0.f, 0.f, 0.f, 0.f};
    if (lk < ls + SP) {  // a lane of this wave is significant by the span's end
      g16_recur<SP, TROW, P>(c, ls, lz, m, sal, tile);
      wave_sync_lds();
#pragma unroll
      for (int kp = 0; kp < 16; ++kp) {
        // A[row r16][k g]: z at l = ls + 2 r16 (+1) of pair 4 kp + g
        const float2 a = *reinterpret_cast<const float2*>(tile + r16 * TROW + 2 * (4 * kp + g));
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
          acc[0][cb] = mf4(a.x, W[0][kp][cb], acc[0][cb]);
          acc[1][cb] = mf4(a.y, W[1][kp][cb], acc[1][cb]);
        }
      }
      wave_sync_lds();  // the tile is rewritten in the next span
    }
    // this wave's sums of the span: D row 4 g + r is l = ls + 2 (4 g + r) +
    // parity.  The wave's tile then holds [cb][row][field slot][re/im][parity]:
    // one field's four values of a row (l, l + 1; re, im) are 16 contiguous
    // bytes, its slot the field's index XOR the row's low bits, so that the
    // combine's 16-byte reads of consecutive rows fall in distinct banks
    float* rw = tile;
#pragma unroll
    for (int cb = 0; cb < CB; ++cb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 4 * g + r;
        *reinterpret_cast<float2*>(
            rw + ((cb * 16 + row) * 8 + ((r16 >> 1) ^ (row & 7))) * 4 + 2 * (r16 & 1)) =
            make_float2(acc[0][cb][r], acc[1][cb][r]);
      }
    __syncthreads();
    // the waves' sums in wave order: thread (field, row) takes l = ls + 2 row
    // and l + 1, re and im; consecutive threads on consecutive l of a field
    // (a thread's first item is the same in every span: its field's row
    // pointer and sum index are formed before the span loop)
    for (int o = tid; o < NF * 16; o += nthr) {
      const bool first = o == tid;
      const int f = first ? cf : o >> 4, row = first ? crow : o & 15;
      const int l = ls + 2 * row;
      if (l > L || f0 + f >= nfield) continue;
      const int idx = first ? cidx : (((f >> 3) * 16 + row) * 8 + ((f & 7) ^ (row & 7))) * 4;
      float4 v = *reinterpret_cast<const float4*>(tiles + idx);
      for (int w = 1; w < nwaves; ++w) {
        const float4 u = *reinterpret_cast<const float4*>(tiles + w * RS + idx);
        v.x += u.x;
        v.y += u.y;
        v.z += u.z;
        v.w += u.w;
      }
      float2* const q =
          first ? cq : reinterpret_cast<float2*>(const_cast<char*>(Fq.at(f0 + f))) + base;
      float2* dst = q + (l - m);
      const float* cz = T.cz + base + (l - m);
      const float c0 = cz[0];
      const float2 a0 = make_float2(v.x * c0, v.z * c0);  // (re, im) of l: parity 0
      if constexpr (IMEX)
        imex_apply(dst, a0, reinterpret_cast<const float2*>(lam)[base + (l - m)],
                   coef + 2 * (f0 + f));
      else
        dst[0] = a0;
      if (l + 1 <= L) {
        const float c1 = cz[1];
        const float2 a1 = make_float2(v.y * c1, v.w * c1);
        if constexpr (IMEX)
          imex_apply(dst + 1, a1, reinterpret_cast<const float2*>(lam)[base + (l - m) + 1],
                     coef + 2 * (f0 + f));
        else
          dst[1] = a1;
      }
    }
    __syncthreads();  // the tiles are rewritten in the next span
  }
}

template <int CB, int ROWS, class P>
hipError_t synth_g16(const typename P::Tables& T, int nfield, int nwaves, const Fields& q,
                     float2* fm, hipStream_t s) {
  if (nwaves < 1 || nwaves > 8 || 64 * nwaves < T.npair) return hipErrorInvalidValue;
  dim3 grid((T.mmax + ROWS) / ROWS, (nfield + 8 * CB - 1) / (8 * CB));
  const size_t lds = sizeof(typename P::ab_t) * analys_ab_len(T.lmax) +
                     sizeof(float) * (2 * (G16_SSPAN / 2) * g16_qrow(CB) +
                                      (size_t)nwaves * (G16_SSPAN / 2) * G16_STROW);
  hipLaunchKernelGGL((k_leg_synth_g16<CB, ROWS, P>), grid, dim3(64 * nwaves), lds, s, T, nfield, q,
                     fm);
  return hipGetLastError();
}

template <int CB, bool IMEX, class P>
hipError_t analys_g16(const typename P::Tables& T, int nfield, int nwaves, const float2* fm,
                      const Fields& q, const float* lam, const float* coef, hipStream_t s) {
  if (nwaves < 1 || nwaves > 8 || 64 * nwaves < T.npair) return hipErrorInvalidValue;
  dim3 grid(T.mmax + 1, (nfield + 8 * CB - 1) / (8 * CB));
  const size_t lds = sizeof(typename P::ab_t) * analys_ab_len(T.lmax) +
                     sizeof(float) * (size_t)nwaves * (G16_ASPAN / 2) * G16_ATROW;
  hipLaunchKernelGGL((k_leg_analys_g16<CB, IMEX, P>), grid, dim3(64 * nwaves), lds, s, T, nfield,
                     fm, q, lam, coef);
  return hipGetLastError();
}

template <class P>
hipError_t synth_g16_nf(const typename P::Tables& T, int nfield, const LegTile& t,
                        const Fields& q, float2* fm, hipStream_t s) {
  if (t.nf == 8)
    return t.rows == 2 ? synth_g16<1, 2, P>(T, nfield, t.nwaves, q, fm, s)
                       : synth_g16<1, 1, P>(T, nfield, t.nwaves, q, fm, s);
  return synth_g16<2, 1, P>(T, nfield, t.nwaves, q, fm, s);
}

template <class P, bool IMEX = false>
hipError_t analys_g16_nf(const typename P::Tables& T, int nfield, const LegTile& t,
                         const float2* fm, const Fields& q, hipStream_t s,
                         const float* lam = nullptr, const float* coef = nullptr) {
  if (t.rows != 1) return hipErrorInvalidValue;  // one row per analysis workgroup
  return t.nf == 8 ? analys_g16<1, IMEX, P>(T, nfield, t.nwaves, fm, q, lam, coef, s)
                   : analys_g16<2, IMEX, P>(T, nfield, t.nwaves, fm, q, lam, coef, s);
}

// ------------------------------ split-BF16 x3 on 4x4x4 matrix blocks (b4)
// v_mfma_f32_4x4x4_16b_bf16 keeps the block and output layout of the 4x4x1
// FP32 product above (A_b[r][k] = the k-th bf16 of lane 4b + r, B_b[k][j] =
// the k-th bf16 of lane 4b + j, D_b[r][j] in register r of lane 4b + j;
// checked on gfx950) and its issue cost (~4 ns per instruction per SIMD,
// measured) for four times the multiply-adds.  K runs over four l of one
// parity of l - m in the synthesis (a lane's own pair, so no operand moves
// between lanes) and over four latitude pairs in the analysis.  Every operand
// is split x = hi + lo into bf16 values rounded to nearest even, and a
// product is lo hi + hi lo + hi hi accumulated in FP32 (lo lo, below 2^-16
// relative, is dropped).  The recurrence is the FP64 one (RecD on the
// LegTablesZ64 tables); each value is rounded to f32 once, then split.
// Groups of four l of one parity are taken from 8-l blocks aligned to
// l - m = 0 mod 8 in every kernel (chains are exactly 0 before their seeds,
// so starting earlier adds exact zeros): the synthesis, the analysis and
// the spectral op then sum the same groups in the same order, and a step
// through the spectral op equals its separate launches bit for bit.
// One chain per lane (up to 512 latitude pairs with 8 waves).
typedef __attribute__((ext_vector_type(4))) short s4v;

__device__ __forceinline__ short bf16_rn(float v) { return __builtin_bit_cast(short, (__bf16)v); }
__device__ __forceinline__ float bf16_f32(short h) {
  return __uint_as_float((unsigned)(unsigned short)h << 16);
}
struct HL {
  short hi, lo;
};
// The split of the f32 value v itself, in every kernel alike: v is made
// opaque first, since for v = (float)d of a double the compiler may fold the
// two roundings into one correctly rounded double -> bf16 conversion (a
// round-to-odd sequence of ~10 VALU ops per value, whose hi differs from
// bf16(float(d)) near ties), and lo = v - hi is rounded as written (a
// product v must not be fused into the subtraction).  Either would make the
// kernels disagree bit for bit.
__device__ __forceinline__ HL split_bf16(float v) {
#pragma clang fp contract(off)
  asm("" : "+v"(v));
  const short h = bf16_rn(v);
  return HL{h, bf16_rn(v - bf16_f32(h))};
}
// four values as the hi and lo operands of one K = 4 product
struct __attribute__((aligned(16))) Split4 {
  s4v hi, lo;
};
__device__ __forceinline__ Split4 split4(float a, float b, float c, float d) {
  const HL x = split_bf16(a), y = split_bf16(b), z = split_bf16(c), w = split_bf16(d);
  return Split4{s4v{x.hi, y.hi, z.hi, w.hi}, s4v{x.lo, y.lo, z.lo, w.lo}};
}
__device__ __forceinline__ f4v mb3(const Split4& a, const Split4& b, f4v c) {
  c = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(a.lo, b.hi, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(a.hi, b.lo, c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(a.hi, b.hi, c, 0, 0, 0);
}
// group (E0, O0, E1, O1) and position k of offset p of a 16-l span whose
// start has l - m = 0 mod 8
__device__ __forceinline__ constexpr int b4_group(int p) { return ((p >> 3) << 1) | (p & 1); }
__device__ __forceinline__ constexpr int b4_pos(int p) { return (p & 7) >> 1; }

// Stage a row's alpha (f64) and its split synthesis coefficients c_l Q_l
// (IMEX: of X_new = (X + s1 A) / (1 + s2 lam), written back to the state as
// stage_row_imex does): 8-l block t, parity par, column col (4 fp + j) at
// shorts ((t 2 + par) NC + col) 8 + {k: hi, 4 + k: lo}, l - m = 8 t + 2 k + par.
template <int NF, bool IMEX>
__device__ __forceinline__ void stage_row_b4(const LegTablesZ64& T, int m, int f0, int nfield,
                                             const Fields& Fq, const Fields& Fa, const float* lam,
                                             const ImexCoef& ic, double* sab, short* sqb, int nt) {
  constexpr int NC = 4 * ((NF + 1) / 2);
  const int n = T.lmax - m + 1;
  const long long base = lm_base(T.lmax, m);
  for (int i = threadIdx.x; i < 8 * nt + kPadQ; i += blockDim.x)
    sab[i] = i < n ? T.ab[base + i] : 0.0;
  for (int i = threadIdx.x; i < 8 * nt; i += blockDim.x) {
    const float cz = i < n ? T.cz[base + i] : 0.f;
    short* dst = sqb + ((i >> 3) * 2 + (i & 1)) * (NC * 8) + ((i & 7) >> 1);
#pragma unroll
    for (int f = 0; f < NC / 2; ++f) {
      const int fld = f0 + f;
      float2 v = make_float2(0.f, 0.f);
      if (f < NF && i < n && fld < nfield) {
        if constexpr (IMEX) {
          float2* X = reinterpret_cast<float2*>(const_cast<char*>(Fq.at(fld))) + base + i;
          const float2 x = *X;
          const float2 a = reinterpret_cast<const float2*>(Fa.at(fld))[base + i];
          const float2 lm2 = reinterpret_cast<const float2*>(lam)[base + i];
          float2 xn;
          float c1, c2;
          imex_coefs(ic, fld, c1, c2);
          xn.x = imex_step(x.x, c1, a.x, c2, lm2.x);
          xn.y = imex_step(x.y, c1, a.y, c2, lm2.y);
          *X = xn;
          v = make_float2(xn.x * cz, xn.y * cz);
        } else {
          const float2 q = reinterpret_cast<const float2*>(Fq.at(fld))[base + i];
          v = make_float2(q.x * cz, q.y * cz);
        }
      }
      const HL re = split_bf16(v.x), im = split_bf16(v.y);
      dst[(2 * f) * 8] = re.hi;
      dst[(2 * f) * 8 + 4] = re.lo;
      dst[(2 * f + 1) * 8] = im.hi;
      dst[(2 * f + 1) * 8 + 4] = im.lo;
    }
  }
}

// One 8-l block of the synthesis from l - m = i (8-aligned): the recurrence
// of the 8 values, their split, then the products as one group (the matrix
// products and the VALU do not overlap on gfx950).
template <bool INJ, int NF>
__device__ __forceinline__ void synth_tile_b4(const double* sab, const short* sqb, int m, int i,
                                              int j, ChainT<RecD>& c, f4v (&accE)[(NF + 1) / 2],
                                              f4v (&accO)[(NF + 1) / 2]) {
  constexpr int FP = (NF + 1) / 2, NC = 4 * FP;
  const short* tq = sqb + (i >> 3) * (2 * NC * 8);
  Split4 bE[FP], bO[FP];
#pragma unroll
  for (int fp = 0; fp < FP; ++fp) {
    bE[fp] = *reinterpret_cast<const Split4*>(tq + (4 * fp + j) * 8);
    bO[fp] = *reinterpret_cast<const Split4*>(tq + (NC + 4 * fp + j) * 8);
  }
  double al[8];
#pragma unroll
  for (int p = 0; p < 8; ++p) al[p] = sab[i + p + 2];
  float z[8];
  const int l = m + i;
#pragma unroll
  for (int p = 0; p < 8; p += 2) {
    if (INJ && (p & 3) == 0 && c.ls == l + p) {
      c.y0 = c.s0;
      c.y1 = c.s1;
    }
    z[p] = (float)c.y0;
    z[p + 1] = (float)c.y1;
    RecD::advance(al[p], al[p + 1], c.ct, c.y0, c.y1);
  }
  const Split4 aE = split4(z[0], z[2], z[4], z[6]);
  const Split4 aO = split4(z[1], z[3], z[5], z[7]);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int fp = 0; fp < FP; ++fp) {
    accE[fp] = mb3(aE, bE[fp], accE[fp]);
    accO[fp] = mb3(aO, bO[fp], accO[fp]);
  }
  __builtin_amdgcn_sched_barrier(0);
}

__host__ __device__ constexpr int b4_tiles(int lmax) { return (lmax + 8) / 8; }

// Synthesis (IMEX: the implicit step of every field in front, as
// k_leg_synth_m4<..., true>): the 8-l blocks of the row from the wave's first
// seed (rounded down to 8 in l - m), with the seed test up to its last seed.
template <int NF, bool IMEX = false>
__global__ __launch_bounds__(512) void k_leg_synth_b4(LegTablesZ64 T, int nfield, Fields Fq,
                                                      float2* __restrict__ fm, int rows,
                                                      Fields Fa, const float* lam, ImexCoef ic) {
  using P = RecD;
  constexpr int FP = (NF + 1) / 2;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  const int L = T.lmax, M = T.mmax, nlat = T.nlat;
  double* sab = reinterpret_cast<double*>(lds_raw);
  short* sqb = reinterpret_cast<short*>(sab + 8 * b4_tiles(L) + kPadQ);
  const int f0 = blockIdx.y * NF;
  const int lane = threadIdx.x & 63, j = lane & 3, blk = lane >> 2;

  for (int r = 0; r < 2; ++r) {
    const int m = row_m(r, rows, M);
    if (m < 0) break;
    const int n = L - m + 1;
    __syncthreads();  // the previous row's readers are done
    stage_row_b4<NF, IMEX>(T, m, f0, nfield, Fq, Fa, lam, ic, sab, sqb, (n + 7) / 8);
    __syncthreads();

    ChainT<P> c;
    load_chain_t<P>(T, m, pair_of(0), c);
    const int lk = wave_min(c.ls), lz = last_seed(c.ls, L);
    f4v accE[FP], accO[FP];
#pragma unroll
    for (int fp = 0; fp < FP; ++fp) accE[fp] = accO[fp] = f4v{0.f, 0.f, 0.f, 0.f};
    if (lk <= L) {
      int i = (lk - m) & ~7;
      for (; i < n && m + i <= lz; i += 8) synth_tile_b4<true, NF>(sab, sqb, m, i, j, c, accE, accO);
      for (; i < n; i += 8) synth_tile_b4<false, NF>(sab, sqb, m, i, j, c, accE, accO);
    }

    const int pb = pair_of(0) - lane + 4 * blk;  // pair of block row 0
#pragma unroll
    for (int fp = 0; fp < FP; ++fp) {
      const int f = 2 * fp + (j >> 1);
      if (f >= NF || f0 + f >= nfield) continue;
      float* out = reinterpret_cast<float*>(fm + (long long)(f0 + f) * nlat * (M + 1) + m) + (j & 1);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int jp = pb + q;
        if (jp >= T.npair) continue;
        const float e = accE[fp][q], o = accO[fp][q];
        out[2LL * jp * (M + 1)] = e + o;
        out[2LL * (nlat - 1 - jp) * (M + 1)] = e - o;
      }
    }
  }
}

// Analysis weights of a row for the b4 kernels: each lane computes W =
// (G_N +/- G_S) w of its own pair (every column, both parities) into its
// wave's tile, then gathers and splits the 16 pairs of its block's
// pair-group for its column and parity, four pairs per K group.
template <int NF>
__device__ __forceinline__ void b4_weights(const LegTablesZ64& T, const float2* fm, int m, int f0,
                                           int nfield, float* tile, int lane, int j, int pg,
                                           bool odd, Split4 (&W)[4][(NF + 1) / 2]) {
  constexpr int FP = (NF + 1) / 2;
  const int M = T.mmax, nlat = T.nlat;
  const int jp = pair_of(0);
  const int jc = min(jp, T.npair - 1);
  const int js = nlat - 1 - jc;
  const float w = jp < T.npair ? T.w[jc] : 0.f;
  float* tw = tile + lane;  // [parity][fp][col][pair]
#pragma unroll
  for (int f = 0; f < 2 * FP; ++f) {
    const bool live = f < NF && f0 + f < nfield;
    const int fld = min(f0 + min(f, NF - 1), nfield - 1);
    const float2* g = fm + (long long)fld * nlat * (M + 1) + m;
    const float2 gN = g[(long long)jc * (M + 1)], gS = g[(long long)js * (M + 1)];
    float4 we{0.f, 0.f, 0.f, 0.f};
    if (live) {
      if (js == jc)  // equator of an odd grid pairs with itself
        we = make_float4(gN.x * w, gN.y * w, 0.f, 0.f);
      else
        we = make_float4((gN.x + gS.x) * w, (gN.y + gS.y) * w, (gN.x - gS.x) * w,
                         (gN.y - gS.y) * w);
    }
    const int fp = f >> 1, c0 = 2 * (f & 1);
    tw[((0 * FP + fp) * 4 + c0) * M4_WROW] = we.x;
    tw[((0 * FP + fp) * 4 + c0 + 1) * M4_WROW] = we.y;
    tw[((1 * FP + fp) * 4 + c0) * M4_WROW] = we.z;
    tw[((1 * FP + fp) * 4 + c0 + 1) * M4_WROW] = we.w;
  }
  wave_sync_lds();
#pragma unroll
  for (int fp = 0; fp < FP; ++fp) {
    const float* src = tile + (((odd ? FP : 0) + fp) * 4 + j) * M4_WROW + 16 * pg;
#pragma unroll
    for (int t = 0; t < 4; ++t)
      W[t][fp] = split4(src[4 * t], src[4 * t + 1], src[4 * t + 2], src[4 * t + 3]);
  }
  wave_sync_lds();  // the tile is rewritten by the spans
}

// A span's tile as bf16: physical row m4_row(p) of span offset p, the pairs
// in quads of [hi x 4, lo x 4] (16 bytes), so that the 16-byte operand load
// of lane 16 pg + 4 q + r (row 4 q + r, pairs 16 pg + 4 t ..) is the A
// operand of K group t.  A row is 64 pairs x 4 bytes, as the FP32 tile's.
__device__ __forceinline__ void b4_tile_store(short* tks, int p, int lane, short hi, short lo) {
  short* d = tks + m4_row(p) * (2 * M4_TROW) + (lane >> 2) * 8 + (lane & 3);
  d[0] = hi;
  d[4] = lo;
}

// The analysis products of one span: block (l-group, pair-group), K groups
// t = 0..3 of four pairs each.
template <int FP>
__device__ __forceinline__ void b4_span_products(const short* tks, int blk, int j, int pg,
                                                 const Split4 (&W)[4][FP], f4v (&acc)[FP]) {
  const short* ta = tks + (4 * (blk & 3) + j) * (2 * M4_TROW) + 32 * pg;
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const Split4 a = *reinterpret_cast<const Split4*>(ta + 8 * t);
#pragma unroll
    for (int fp = 0; fp < FP; ++fp) acc[fp] = mb3(a, W[t][fp], acc[fp]);
  }
  __builtin_amdgcn_sched_barrier(0);
}

// Sum of the block rows over the 4 pair-groups (as in the FP32 kernels):
// lane 4 b + j ends with the sum for row pg of its block.
__device__ __forceinline__ float b4_pair_group_sum(const f4v& a) {
  float v[4] = {a[0], a[1], a[2], a[3]};
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const auto s2 = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[i]), __float_as_uint(v[i + 2]),
                                                     false, false);
    v[i] = __uint_as_float(s2[0]) + __uint_as_float(s2[1]);
  }
  const auto s1 = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[0]), __float_as_uint(v[1]),
                                                   false, false);
  return __uint_as_float(s1[0]) + __uint_as_float(s1[1]);
}

// The first span of a row for the b4 kernels: the workgroup's first seed,
// rounded down to l - m = 0 mod 8 (L + 1: the row is 0).
__device__ __forceinline__ int b4_row_start(const int* smin, int nwaves, int m, int L) {
  int lblk = smin[0];
  for (int w = 1; w < nwaves; ++w) lblk = min(lblk, smin[w]);
  return lblk <= L ? m + ((lblk - m) & ~7) : L + 1;
}

// Analysis: k_leg_analys_m4's spans (one chain per lane) with the tile in
// bf16 and the products as split 4x4x4 blocks.
template <int NF>
__global__ __launch_bounds__(512) void k_leg_analys_b4(LegTablesZ64 T, int nfield,
                                                       const float2* __restrict__ fm, Fields Fq,
                                                       int rows) {
  using P = RecD;
  constexpr int FP = (NF + 1) / 2;
  constexpr int NH = M4_NH;  // spans per workgroup barrier
  constexpr int NRED = NH * 64 * FP;  // a wave's sums between two barriers
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  __shared__ int smin[8];
  const int L = T.lmax, M = T.mmax;
  double* sab = reinterpret_cast<double*>(lds_raw);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nthr = blockDim.x, nwaves = nthr >> 6;
  const int j = lane & 3, blk = lane >> 2;
  const int pg = blk >> 2;
  const bool odd = m4_lgroup(blk & 3) >> 1;
  float* const tiles = reinterpret_cast<float*>(sab + analys_ab_len(L));
  float* tile = tiles + wv * (M4_SPAN * M4_TROW);
  // the waves' sums, [buffer][wave][NRED], after the tiles (as k_leg_analys_m4)
  float* const red = tiles + nwaves * (M4_SPAN * M4_TROW);
  short* tks = reinterpret_cast<short*>(tile);
  const int f0 = blockIdx.y * NF;

  for (int r = 0; r < 2; ++r) {
    const int m = row_m(r, rows, M, false);
    if (m < 0) break;
    const long long base = lm_base(L, m);
    __syncthreads();
    stage_row_t<NF, P>(T, m, f0, nfield, Fq, false, sab, nullptr, kPadAnalys);
    ChainT<P> c;
    load_chain_t<P>(T, m, pair_of(0), c);
    const int lmin = wave_min(c.ls), lzmax = last_seed(c.ls, L);
    Split4 W[4][FP];
    b4_weights<NF>(T, fm, m, f0, nfield, tile, lane, j, pg, odd, W);
    if (lane == 0) smin[wv] = lmin;
    __syncthreads();  // also publishes the staged row
    const int lblk = b4_row_start(smin, nwaves, m, L);

    for (int i = m + tid; i < lblk; i += nthr)
#pragma unroll
      for (int f = 0; f < NF; ++f)
        if (f0 + f < nfield)
          reinterpret_cast<float2*>(const_cast<char*>(Fq.at(f0 + f)))[base + (i - m)] =
              make_float2(0.f, 0.f);

    int buf = 0;
    for (int ls0 = lblk; ls0 <= L; ls0 += NH * M4_SPAN) {
#pragma unroll
      for (int h = 0; h < NH; ++h) {
        const int ls = ls0 + h * M4_SPAN;
        float* rw = red + (buf * nwaves + wv) * NRED + h * 64 * FP;
        if (lmin >= ls + M4_SPAN || ls > L) {
#pragma unroll
          for (int fp = 0; fp < FP; ++fp) rw[lane * FP + fp] = 0.f;
          continue;
        }
        // the span's values: recurrence (alpha four at a time), split, tile
#pragma unroll
        for (int pb = 0; pb < M4_SPAN; pb += 4) {
          double al[4];
#pragma unroll
          for (int p = 0; p < 4; ++p) al[p] = sab[ls - m + pb + p + 2];
#pragma unroll
          for (int p = pb; p < pb + 4; p += 2) {
            const int l = ls + p;
            if (p % kTile == 0 && ls <= lzmax && c.ls == l) {
              c.y0 = c.s0;
              c.y1 = c.s1;
            }
            const HL v0 = split_bf16((float)c.y0), v1 = split_bf16((float)c.y1);
            b4_tile_store(tks, p, lane, v0.hi, v0.lo);
            b4_tile_store(tks, p + 1, lane, v1.hi, v1.lo);
            P::advance(al[p - pb], al[p - pb + 1], c.ct, c.y0, c.y1);
          }
        }
        wave_sync_lds();
        f4v acc[FP];
#pragma unroll
        for (int fp = 0; fp < FP; ++fp) acc[fp] = f4v{0.f, 0.f, 0.f, 0.f};
        b4_span_products<FP>(tks, blk, j, pg, W, acc);
        wave_sync_lds();  // the tile is rewritten in the next span
#pragma unroll
        for (int fp = 0; fp < FP; ++fp) rw[lane * FP + fp] = b4_pair_group_sum(acc[fp]);
      }
      __syncthreads();
      for (int o = tid; o < NRED; o += nthr) {
        const int fp = o % FP, ln = (o / FP) & 63, h = o / (64 * FP);
        const int jj = ln & 3, bb = ln >> 2;
        const int l = ls0 + h * M4_SPAN + m4_off(m4_lgroup(bb & 3), bb >> 2);
        const int f = 2 * fp + (jj >> 1);
        if (l <= L && f < NF && f0 + f < nfield) {
          const float s = sum_waves([&](int w) { return red[(buf * nwaves + w) * NRED + o]; },
                                    nwaves);
          float* dst = reinterpret_cast<float*>(const_cast<char*>(Fq.at(f0 + f))) +
                       2 * (base + (l - m));
          dst[jj & 1] = s * T.cz[base + (l - m)];
        }
      }
      buf ^= 1;
    }
  }
}

// The spectral op in the split-BF16 form: k_leg_spectral_m4's pipeline (one
// chain per lane) with the b4 analysis of each span, the solved span's
// coefficients staged split in LDS, and the synthesis products of the span's
// four parity groups from the split values the wave kept in registers.
template <int NF>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(NF == 2 ? 4 : 1)))
void k_leg_spectral_b4(LegTablesZ64 T, int nfield, float2* fm, Fields St,
                       const float* __restrict__ lam, ImexCoef ic) {
  using P = RecD;
  constexpr int FP = (NF + 1) / 2;
  constexpr int NO = 64 * FP;  // combined sums per span
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  __shared__ float red[2][8][NO];
  // c_l X_new of a span, split: [column][group (E0, O0, E1, O1)][hi x 4, lo x 4]
  __shared__ __attribute__((aligned(16))) short qsb[2][4 * FP][4][8];
  __shared__ int smin[8];
  const int L = T.lmax, M = T.mmax, nlat = T.nlat;
  double* sab = reinterpret_cast<double*>(lds_raw);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int nthr = blockDim.x, nwaves = nthr >> 6;
  const int j = lane & 3, blk = lane >> 2;
  const int pg = blk >> 2;
  const bool odd = m4_lgroup(blk & 3) >> 1;
  float* tile = reinterpret_cast<float*>(sab + analys_ab_len(L)) + wv * (M4_SPAN * M4_TROW);
  short* tks = reinterpret_cast<short*>(tile);
  // this thread's coefficient of a span's combine (as k_leg_spectral_m4)
  const int cfp = tid % FP, cln = (tid / FP) & 63;
  const int cjj = cln & 3, cbb = cln >> 2;
  const int clo = m4_off(m4_lgroup(cbb & 3), cbb >> 2);
  const int cf = 2 * cfp + (cjj >> 1);
  const bool cthr = tid < NO && cf < NF;
  const bool cfield = cthr && cf < nfield;
  float* const cX = cfield ? reinterpret_cast<float*>(const_cast<char*>(St.at(ic.spc[cf])))
                           : nullptr;
  const int m = row_m(0, 1, M, false);  // rows dealt as in k_leg_spectral_m4
  const long long base = lm_base(L, m);
  stage_row_t<NF, P>(T, m, 0, nfield, Fields{}, false, sab, nullptr, kPadAnalys);

  ChainT<P> c;
  load_chain_t<P>(T, m, pair_of(0), c);
  const int lmin = wave_min(c.ls), lzmax = last_seed(c.ls, L);
  Split4 W[4][FP];
  b4_weights<NF>(T, fm, m, 0, nfield, tile, lane, j, pg, odd, W);
  if (lane == 0) smin[wv] = lmin;
  __syncthreads();  // also publishes the staged row; every W is read
  const int lblk = b4_row_start(smin, nwaves, m, L);

  // below the row start the analysed values are exactly 0: the implicit
  // step alone (and no synthesis contribution)
  for (int o = tid; o < (lblk - m) * 2 * NF; o += nthr) {
    const int f = o % NF, q = o / NF;
    if (f >= nfield) continue;
    const long long i = 2 * base + q;
    float* X = reinterpret_cast<float*>(const_cast<char*>(St.at(ic.spc[f])));
    X[i] = imex_step(X[i], ic.s1[f], 0.0f, ic.s2[f], lam[i]);
  }

  f4v accE[FP], accO[FP];
#pragma unroll
  for (int fp = 0; fp < FP; ++fp) accE[fp] = accO[fp] = f4v{0.f, 0.f, 0.f, 0.f};

  // span ls's analysis into this wave's sums red[b][wv] and its split values
  // zs (groups E0, O0, E1, O1); false when the wave's chain is still 0 there
  auto analyse = [&](int ls, int b, Split4 (&zs)[4]) -> bool {
    float* rw = &red[b][wv][0];
    if (lmin >= ls + M4_SPAN) {
#pragma unroll
      for (int fp = 0; fp < FP; ++fp) rw[lane * FP + fp] = 0.f;
      return false;
    }
    float z[8];
#pragma unroll
    for (int pb = 0; pb < M4_SPAN; pb += 4) {
      double al[4];
#pragma unroll
      for (int p = 0; p < 4; ++p) al[p] = sab[ls - m + pb + p + 2];
#pragma unroll
      for (int p = pb; p < pb + 4; p += 2) {
        const int l = ls + p;
        if (p % kTile == 0 && ls <= lzmax && c.ls == l) {
          c.y0 = c.s0;
          c.y1 = c.s1;
        }
        z[p & 7] = (float)c.y0;
        z[(p & 7) + 1] = (float)c.y1;
        P::advance(al[p - pb], al[p - pb + 1], c.ct, c.y0, c.y1);
      }
      if ((pb & 7) == 4) {  // an 8-l block complete: its two parity groups
        const int g = (pb >> 3) << 1;
        zs[g] = split4(z[0], z[2], z[4], z[6]);
        zs[g + 1] = split4(z[1], z[3], z[5], z[7]);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          b4_tile_store(tks, (pb & 8) + 2 * k, lane, zs[g].hi[k], zs[g].lo[k]);
          b4_tile_store(tks, (pb & 8) + 2 * k + 1, lane, zs[g + 1].hi[k], zs[g + 1].lo[k]);
        }
      }
    }
    wave_sync_lds();
    f4v acc[FP];
#pragma unroll
    for (int fp = 0; fp < FP; ++fp) acc[fp] = f4v{0.f, 0.f, 0.f, 0.f};
    b4_span_products<FP>(tks, blk, j, pg, W, acc);
    wave_sync_lds();  // the tile is rewritten by the next span
#pragma unroll
    for (int fp = 0; fp < FP; ++fp) rw[lane * FP + fp] = b4_pair_group_sum(acc[fp]);
    return true;
  };
  // synthesis products of one span's four groups, E0 and E1 into accE, O0
  // and O1 into accO (the order of the synthesis kernel's 8-l blocks)
  auto synthesise = [&](const Split4 (&zp)[4], int b) {
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int fp = 0; fp < FP; ++fp) {
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        Split4 bq = *reinterpret_cast<const Split4*>(&qsb[b][4 * fp + j][g][0]);
        if (NF == 1 && j >= 2) bq = Split4{};  // the idle columns of one field
        if (g & 1)
          accO[fp] = mb3(zp[g], bq, accO[fp]);
        else
          accE[fp] = mb3(zp[g], bq, accE[fp]);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  };
  auto iteration = [&](int ls, int b, Split4 (&zc)[4], const Split4 (&zp)[4], bool livep) -> bool {
    const int l = ls + clo;
    const bool has = cfield && l <= L;
    long long ci = 0;
    float xo = 0.f, lv = 0.f, cz = 0.f;
    if (has) {  // issued before the analysis, used after the barrier
      ci = 2 * (base + (l - m)) + (cjj & 1);
      xo = cX[ci];
      lv = lam[ci];
      cz = T.cz[base + (l - m)];
    }
    const bool live = analyse(ls, b, zc);
    __syncthreads();  // the waves' sums of span ls; qsb[b ^ 1] of the span before
    if (cthr) {
      float qv = 0.f;
      if (has) {
        float s = red[b][0][tid];
        for (int w = 1; w < nwaves; ++w) s += red[b][w][tid];
        const float xn = imex_step(xo, pick<NF>(ic.s1, cf), s * cz, pick<NF>(ic.s2, cf), lv);
        cX[ci] = xn;
        qv = xn * cz;
      }
      const HL h = split_bf16(qv);
      short* d = &qsb[b][2 * (cfield ? pick<NF>(ic.spc, cf) : cf) + (cjj & 1)][b4_group(clo)]
                     [b4_pos(clo)];
      d[0] = h.hi;
      d[4] = h.lo;
    }
    if (livep) synthesise(zp, b ^ 1);
    return live;
  };
  Split4 zA[4], zB[4];
  bool liveA = false, liveB = false, lastA = false;
  for (int ls = lblk; ls <= L; ls += 2 * M4_SPAN) {
    liveA = iteration(ls, 0, zA, zB, liveB);
    lastA = true;
    if (ls + M4_SPAN > L) break;
    liveB = iteration(ls + M4_SPAN, 1, zB, zA, liveA);
    lastA = false;
  }
  if (lblk <= L) {
    __syncthreads();  // qsb of the last span
    if (lastA) {
      if (liveA) synthesise(zA, 0);
    } else if (liveB) {
      synthesise(zB, 1);
    }
  }

  int M1 = M + 1;
  asm volatile("" : "+s"(M1));
  const int pb = pair_of(0) - lane + 4 * blk;
#pragma unroll
  for (int fp = 0; fp < FP; ++fp) {
    const int s = 2 * fp + (j >> 1);
    if (s >= NF || s >= nfield) continue;
    float* out = reinterpret_cast<float*>(fm + (long long)s * nlat * M1 + m) + (j & 1);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int jp = pb + q;
      if (jp >= T.npair) continue;
      const float e = accE[fp][q], o = accO[fp][q];
      out[2LL * jp * M1] = e + o;
      out[2LL * (nlat - 1 - jp) * M1] = e - o;
    }
  }
}

template <int NF, bool IMEX = false>
hipError_t synth_b4(const LegTablesZ64& T, int nfield, int nwaves, int rows, const Fields& q,
                    float2* fm, hipStream_t s, const Fields& a = Fields{},
                    const float* lam = nullptr, const ImexCoef& ic = ImexCoef{}) {
  const int nt = b4_tiles(T.lmax);
  const size_t lds = sizeof(double) * (8 * nt + kPadQ) +
                     sizeof(short) * (size_t)nt * 2 * (4 * ((NF + 1) / 2)) * 8;
  dim3 grid(rows == 1 ? T.mmax + 1 : T.mmax / 2 + 1, (nfield + NF - 1) / NF);
  hipLaunchKernelGGL((k_leg_synth_b4<NF, IMEX>), grid, dim3(64 * nwaves), lds, s, T, nfield, q, fm,
                     rows, a, lam, ic);
  return hipGetLastError();
}

template <int NF>
hipError_t analys_b4(const LegTablesZ64& T, int nfield, int nwaves, int rows, const float2* fm,
                     const Fields& q, hipStream_t s) {
  dim3 grid(rows == 1 ? T.mmax + 1 : T.mmax / 2 + 1, (nfield + NF - 1) / NF);
  const size_t lds = sizeof(double) * analys_ab_len(T.lmax) +
                     sizeof(float) * (size_t)nwaves * M4_SPAN * M4_TROW +
                     sizeof(float) * (size_t)2 * nwaves * M4_NH * 64 * ((NF + 1) / 2);
  hipLaunchKernelGGL((k_leg_analys_b4<NF>), grid, dim3(64 * nwaves), lds, s, T, nfield, fm, q, rows);
  return hipGetLastError();
}

template <int NF>
hipError_t spectral_b4(const LegTablesZ64& T, int nfield, int nwaves, float2* fm,
                       const Fields& st, const float* lam, const ImexCoef& ic, hipStream_t s) {
  const size_t lds = sizeof(double) * analys_ab_len(T.lmax) +
                     sizeof(float) * (size_t)nwaves * M4_SPAN * M4_TROW;
  hipLaunchKernelGGL((k_leg_spectral_b4<NF>), dim3(T.mmax + 1), dim3(64 * nwaves), lds, s, T,
                     nfield, fm, st, lam, ic);
  return hipGetLastError();
}

// Built-in choice of the contraction unit (1 VALU, 2 16x16x4 matrix tiles,
// 3 4x4 matrix blocks).
constexpr int kDefaultImpl = 3;

int round_k(int k) { return k <= 1 ? 1 : k <= 2 ? 2 : k <= 4 ? 4 : k <= 8 ? 8 : 0; }

}  // namespace

int mfma_col_tiles(int nfield);

LegTile resolve_leg_tile(int legendre, int stage, int npair, int nfield, LegTile t) {
  if (legendre == TSPHERE_LEG_BF16X3) {
    // built-in: the 4x4x4 block kernels (impl 4) where one chain per lane
    // covers the pairs; the 16x16x32 column-tile kernels (ctc) beyond that
    // or when a column-tile count is asked for
    if (t.impl == 0 && t.ctc == 0 && npair <= kB4MaxPairs) t.impl = 4;
    if (t.impl == 4 && npair <= kB4MaxPairs) {
      if (t.nf == 0) t.nf = nfield == 1 ? 1 : nfield == 2 ? 2 : 4;
      if (t.rows == 0) t.rows = nfield <= 2 ? 1 : 2;
      if (t.nwaves == 0) t.nwaves = (npair + 63) / 64;
      if (t.k == 0) t.k = 1;
      t.ctc = 0;
      return t;
    }
    t.impl = 0;
    const int ct = mfma_col_tiles(nfield);
    if (t.ctc == 0)
      t.ctc = stage == 0 ? (ct >= 4 ? 4 : ct >= 2 ? 2 : 1) : (ct >= 2 ? 2 : 1);
    return t;
  }
  if (t.impl == 0) t.impl = kDefaultImpl;
  // per-parity matrix products: one m-row per workgroup, 8 or 16 fields, one
  // chain per lane, a wave per 64 pairs (up to 512 pairs)
  if (t.impl == 5 && npair > 64 * 8) t = LegTile{0, 0, 0, 0, 3, 0};
  if (t.impl == 5) {
    if (t.nf == 0) t.nf = nfield >= 16 ? 16 : 8;
    if (t.nwaves == 0) t.nwaves = (npair + 63) / 64;
    if (t.k == 0) t.k = 1;
    if (t.rows == 0) t.rows = 1;
    t.ctc = 0;
    return t;
  }
  // one row per workgroup when a launch has one or two fields (its workgroups
  // are few and long); the 16x16x4 matrix-core kernels always pair rows
  if (t.rows == 0) t.rows = nfield <= 2 ? 1 : 2;
  if (t.impl == 2) t.rows = 2;
  // beyond the matrix-core kernels' reach
  if (t.impl == 2 && npair > 64 * 4 * 4) t.impl = 1;
  if (t.impl == 3 && npair > 64 * 8 * 4) t.impl = 1;
  if (t.nf == 0) t.nf = nfield == 1 ? 1 : nfield == 2 ? 2 : 4;
  // block-matrix kernels of one or two fields per workgroup: up to 8 waves
  // of one chain each (measured ahead of 4 waves of two)
  const int wmax = (t.impl == 3 && t.nf <= 2) ? 8 : 4;
  if (t.nwaves == 0 && t.k == 0) t.nwaves = std::min(wmax, (npair + 63) / 64);
  if (t.nwaves == 0) t.nwaves = std::min(wmax, (npair + 64 * t.k - 1) / (64 * t.k));
  if (t.k == 0) t.k = round_k((npair + 64 * t.nwaves - 1) / (64 * t.nwaves));
  return t;
}

namespace {

bool fp32_tile_ok(const LegTile& t, int npair) {
  if (t.impl == 5)  // two rows per workgroup: the 8-field synthesis
    return (t.nf == 8 || t.nf == 16) && t.k == 1 && (t.rows == 1 || (t.rows == 2 && t.nf == 8)) &&
           t.nwaves >= 1 && t.nwaves <= 8 && 64 * t.nwaves >= npair;
  const bool kok = t.impl >= 2 ? (t.k == 1 || t.k == 2 || t.k == 4) : round_k(t.k) == t.k;
  const int wmax = t.impl == 3 ? 8 : 4;
  // eight fields per workgroup: block-matrix kernels with one or two chains
  const bool nfok = t.nf == 1 || t.nf == 2 || t.nf == 4 || (t.nf == 8 && t.impl == 3 && t.k <= 2);
  return (t.impl >= 1 && t.impl <= 3) && (t.rows == 1 || t.rows == 2) && nfok &&
         t.nwaves >= 1 && t.nwaves <= wmax && kok && 64 * t.nwaves * t.k >= npair;
}

}  // namespace

hipError_t launch_leg_synth(const LegTables& T, int nfield, const Fields& q, float2* fm,
                            LegTile tile, hipStream_t s) {
  if (nfield <= 0) return hipSuccess;
  const LegTile t = resolve_leg_tile(TSPHERE_LEG_FP32, 0, T.npair, nfield, tile);
  if (!fp32_tile_ok(t, T.npair)) return hipErrorInvalidValue;
  if (t.impl == 5) return synth_g16_nf<RecF>(T, nfield, t, q, fm, s);
  if (t.impl == 2) switch (t.nf) {
      case 1: return synth_mf_nf<1, RecF>(T, nfield, t, q, fm, s);
      case 2: return synth_mf_nf<2, RecF>(T, nfield, t, q, fm, s);
      default: return synth_mf_nf<4, RecF>(T, nfield, t, q, fm, s);
    }
  if (t.impl == 3) switch (t.nf) {
      case 1: return synth_m4_nf<1, RecF>(T, nfield, t, q, fm, s);
      case 2: return synth_m4_nf<2, RecF>(T, nfield, t, q, fm, s);
      case 8: return synth_m4_nf8<RecF>(T, nfield, t, q, fm, s);
      default: return synth_m4_nf<4, RecF>(T, nfield, t, q, fm, s);
    }
  switch (t.nf) {
    case 1: return synth_nf<1, RecF>(T, nfield, t, q, fm, s);
    case 2: return synth_nf<2, RecF>(T, nfield, t, q, fm, s);
    default: return synth_nf<4, RecF>(T, nfield, t, q, fm, s);
  }
}

hipError_t launch_leg_analys(const LegTables& T, int nfield, const float2* fm, const Fields& q,
                             LegTile tile, hipStream_t s) {
  if (nfield <= 0) return hipSuccess;
  const LegTile t = resolve_leg_tile(TSPHERE_LEG_FP32, 1, T.npair, nfield, tile);
  if (!fp32_tile_ok(t, T.npair)) return hipErrorInvalidValue;
  if (t.impl == 5) return analys_g16_nf<RecF>(T, nfield, t, fm, q, s);
  if (t.impl == 2) switch (t.nf) {
      case 1: return analys_mf_nf<1, RecF>(T, nfield, t, fm, q, s);
      case 2: return analys_mf_nf<2, RecF>(T, nfield, t, fm, q, s);
      default: return analys_mf_nf<4, RecF>(T, nfield, t, fm, q, s);
    }
  if (t.impl == 3) switch (t.nf) {
      case 1: return analys_m4_nf<1, RecF>(T, nfield, t, fm, q, s);
      case 2: return analys_m4_nf<2, RecF>(T, nfield, t, fm, q, s);
      case 8: return analys_m4_nf8<RecF>(T, nfield, t, fm, q, s);
      default: return analys_m4_nf<4, RecF>(T, nfield, t, fm, q, s);
    }
  switch (t.nf) {
    case 1: return analys_nf<1, RecF>(T, nfield, t, fm, q, s);
    case 2: return analys_nf<2, RecF>(T, nfield, t, fm, q, s);
    default: return analys_nf<4, RecF>(T, nfield, t, fm, q, s);
  }
}

hipError_t launch_leg_synth_f64r(const LegTablesZ64& T, int nfield, const Fields& q, float2* fm,
                                 LegTile tile, hipStream_t s) {
  if (nfield <= 0) return hipSuccess;
  const LegTile t = resolve_leg_tile(TSPHERE_LEG_F64REC, 0, T.npair, nfield, tile);
  if (!fp32_tile_ok(t, T.npair)) return hipErrorInvalidValue;
  if (t.impl == 5) return synth_g16_nf<RecD>(T, nfield, t, q, fm, s);
  if (t.impl == 2) switch (t.nf) {
      case 1: return synth_mf_nf<1, RecD>(T, nfield, t, q, fm, s);
      case 2: return synth_mf_nf<2, RecD>(T, nfield, t, q, fm, s);
      default: return synth_mf_nf<4, RecD>(T, nfield, t, q, fm, s);
    }
  if (t.impl == 3) switch (t.nf) {
      case 1: return synth_m4_nf<1, RecD>(T, nfield, t, q, fm, s);
      case 2: return synth_m4_nf<2, RecD>(T, nfield, t, q, fm, s);
      case 8: return synth_m4_nf8<RecD>(T, nfield, t, q, fm, s);
      default: return synth_m4_nf<4, RecD>(T, nfield, t, q, fm, s);
    }
  switch (t.nf) {
    case 1: return synth_nf<1, RecD>(T, nfield, t, q, fm, s);
    case 2: return synth_nf<2, RecD>(T, nfield, t, q, fm, s);
    default: return synth_nf<4, RecD>(T, nfield, t, q, fm, s);
  }
}

hipError_t launch_leg_analys_f64r(const LegTablesZ64& T, int nfield, const float2* fm,
                                  const Fields& q, LegTile tile, hipStream_t s) {
  if (nfield <= 0) return hipSuccess;
  const LegTile t = resolve_leg_tile(TSPHERE_LEG_F64REC, 1, T.npair, nfield, tile);
  if (!fp32_tile_ok(t, T.npair)) return hipErrorInvalidValue;
  if (t.impl == 5) return analys_g16_nf<RecD>(T, nfield, t, fm, q, s);
  if (t.impl == 2) switch (t.nf) {
      case 1: return analys_mf_nf<1, RecD>(T, nfield, t, fm, q, s);
      case 2: return analys_mf_nf<2, RecD>(T, nfield, t, fm, q, s);
      default: return analys_mf_nf<4, RecD>(T, nfield, t, fm, q, s);
    }
  if (t.impl == 3) switch (t.nf) {
      case 1: return analys_m4_nf<1, RecD>(T, nfield, t, fm, q, s);
      case 2: return analys_m4_nf<2, RecD>(T, nfield, t, fm, q, s);
      case 8: return analys_m4_nf8<RecD>(T, nfield, t, fm, q, s);
      default: return analys_m4_nf<4, RecD>(T, nfield, t, fm, q, s);
    }
  switch (t.nf) {
    case 1: return analys_nf<1, RecD>(T, nfield, t, fm, q, s);
    case 2: return analys_nf<2, RecD>(T, nfield, t, fm, q, s);
    default: return analys_nf<4, RecD>(T, nfield, t, fm, q, s);
  }
}

// ------------------------------- the analysis with the implicit step behind
// X = (X + s1 A) / (1 + s2 lam) over the nlm complex values of each field f
// (s1, s2 = coef[2 f], coef[2 f + 1]; A and X distinct buffers): grid
// (chunks of 1024 values, field), four values a thread, loads first
__global__ __launch_bounds__(256) void k_imex_fields(long long nlm, const float2* __restrict__ lam,
                                                     const float* __restrict__ coef, Fields Fa,
                                                     Fields Fq) {
  const int f = blockIdx.y;
  const float s1 = coef[2 * f], s2 = coef[2 * f + 1];
  const float2* __restrict__ A = reinterpret_cast<const float2*>(Fa.at(f));
  float2* __restrict__ X = reinterpret_cast<float2*>(const_cast<char*>(Fq.at(f)));
  const long long i0 = (long long)blockIdx.x * 1024 + threadIdx.x;
  float2 a[4], x[4], lm[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const long long i = i0 + 256 * k;
    if (i < nlm) {
      a[k] = A[i];
      x[k] = X[i];
      lm[k] = lam[i];
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const long long i = i0 + 256 * k;
    if (i < nlm)
      X[i] = make_float2(imex_step(x[k].x, s1, a[k].x, s2, lm[k].x),
                         imex_step(x[k].y, s1, a[k].y, s2, lm[k].y));
  }
}

hipError_t launch_imex_fields(long long nlm, int nfield, const Fields& a, const Fields& q,
                              const float* lam, const float* coef, hipStream_t s) {
  if (nfield <= 0) return hipSuccess;
  dim3 grid((unsigned)((nlm + 1023) / 1024), nfield);
  hipLaunchKernelGGL(k_imex_fields, grid, dim3(256), 0, s, nlm, reinterpret_cast<const float2*>(lam),
                     coef, a, q);
  return hipGetLastError();
}

hipError_t launch_leg_analys_imex(const LegTables& T, int nfield, const float2* fm,
                                  const Fields& a, const Fields& q, const float* lam,
                                  const float* coef, LegTile tile, hipStream_t s) {
  if (nfield <= 0) return hipSuccess;
  const LegTile t = resolve_leg_tile(TSPHERE_LEG_FP32, 1, T.npair, nfield, tile);
  if (!fp32_tile_ok(t, T.npair)) return hipErrorInvalidValue;
  if (t.impl == 5) return analys_g16_nf<RecF, true>(T, nfield, t, fm, q, s, lam, coef);
  if (t.impl == 3) switch (t.nf) {
      case 1: return analys_m4_nf<1, RecF, true>(T, nfield, t, fm, q, s, lam, coef);
      case 2: return analys_m4_nf<2, RecF, true>(T, nfield, t, fm, q, s, lam, coef);
      case 8: return analys_m4_nf8<RecF, true>(T, nfield, t, fm, q, s, lam, coef);
      default: return analys_m4_nf<4, RecF, true>(T, nfield, t, fm, q, s, lam, coef);
    }
  if (hipError_t e = launch_leg_analys(T, nfield, fm, a, tile, s)) return e;
  return launch_imex_fields(nlm_calc(T.lmax, T.mmax), nfield, a, q, lam, coef, s);
}

hipError_t launch_leg_analys_imex_f64r(const LegTablesZ64& T, int nfield, const float2* fm,
                                       const Fields& a, const Fields& q, const float* lam,
                                       const float* coef, LegTile tile, hipStream_t s) {
  if (nfield <= 0) return hipSuccess;
  const LegTile t = resolve_leg_tile(TSPHERE_LEG_F64REC, 1, T.npair, nfield, tile);
  if (!fp32_tile_ok(t, T.npair)) return hipErrorInvalidValue;
  if (t.impl == 5) return analys_g16_nf<RecD, true>(T, nfield, t, fm, q, s, lam, coef);
  if (t.impl == 3) switch (t.nf) {
      case 1: return analys_m4_nf<1, RecD, true>(T, nfield, t, fm, q, s, lam, coef);
      case 2: return analys_m4_nf<2, RecD, true>(T, nfield, t, fm, q, s, lam, coef);
      case 8: return analys_m4_nf8<RecD, true>(T, nfield, t, fm, q, s, lam, coef);
      default: return analys_m4_nf<4, RecD, true>(T, nfield, t, fm, q, s, lam, coef);
    }
  if (hipError_t e = launch_leg_analys_f64r(T, nfield, fm, a, tile, s)) return e;
  return launch_imex_fields(nlm_calc(T.lmax, T.mmax), nfield, a, q, lam, coef, s);
}

hipError_t launch_leg_imex_synth(const LegTables& T, int nfield, const Fields& q, const Fields& a,
                                const float* lam, const ImexCoef& ic, float2* fm, LegTile tile,
                                hipStream_t s) {
  LegTile t = resolve_leg_tile(TSPHERE_LEG_FP32, 0, T.npair, nfield, tile);
  if (t.impl != 3) t = resolve_leg_tile(TSPHERE_LEG_FP32, 0, T.npair, nfield, LegTile{0, 0, 0, 0, 3, 0});
  if (!fp32_tile_ok(t, T.npair) || t.impl != 3) return hipErrorInvalidValue;
  return imex_synth_launch<RecF>(T, nfield, t, q, a, lam, ic, fm, s);
}

hipError_t launch_leg_imex_synth_f64r(const LegTablesZ64& T, int nfield, const Fields& q,
                                     const Fields& a, const float* lam, const ImexCoef& ic,
                                     float2* fm, LegTile tile, hipStream_t s) {
  LegTile t = resolve_leg_tile(TSPHERE_LEG_F64REC, 0, T.npair, nfield, tile);
  if (t.impl != 3)
    t = resolve_leg_tile(TSPHERE_LEG_F64REC, 0, T.npair, nfield, LegTile{0, 0, 0, 0, 3, 0});
  if (!fp32_tile_ok(t, T.npair) || t.impl != 3) return hipErrorInvalidValue;
  return imex_synth_launch<RecD>(T, nfield, t, q, a, lam, ic, fm, s);
}

hipError_t launch_leg_spectral(const LegTables& T, int nfield, float2* fm, const Fields& st,
                               const float* lam, const ImexCoef& ic, hipStream_t s) {
  return spectral_launch<RecF>(T, nfield, fm, st, lam, ic, s);
}

hipError_t launch_leg_spectral_f64r(const LegTablesZ64& T, int nfield, float2* fm,
                                    const Fields& st, const float* lam, const ImexCoef& ic,
                                    hipStream_t s) {
  return spectral_launch<RecD>(T, nfield, fm, st, lam, ic, s);
}

}  // namespace tsphere

namespace tsphere {

namespace {

bool b4_tile_ok(const LegTile& t, int npair) {
  return t.impl == 4 && (t.nf == 1 || t.nf == 2 || t.nf == 4) && (t.rows == 1 || t.rows == 2) &&
         t.k == 1 && t.nwaves >= 1 && t.nwaves <= 8 && 64 * t.nwaves >= npair;
}

}  // namespace

hipError_t launch_leg_synth_b4(const LegTablesZ64& T, int nfield, const Fields& q, float2* fm,
                               LegTile tile, hipStream_t s) {
  if (nfield <= 0) return hipSuccess;
  const LegTile t = resolve_leg_tile(TSPHERE_LEG_BF16X3, 0, T.npair, nfield, tile);
  if (!b4_tile_ok(t, T.npair)) return hipErrorInvalidValue;
  switch (t.nf) {
    case 1: return synth_b4<1>(T, nfield, t.nwaves, t.rows, q, fm, s);
    case 2: return synth_b4<2>(T, nfield, t.nwaves, t.rows, q, fm, s);
    default: return synth_b4<4>(T, nfield, t.nwaves, t.rows, q, fm, s);
  }
}

hipError_t launch_leg_analys_b4(const LegTablesZ64& T, int nfield, const float2* fm,
                                const Fields& q, LegTile tile, hipStream_t s) {
  if (nfield <= 0) return hipSuccess;
  const LegTile t = resolve_leg_tile(TSPHERE_LEG_BF16X3, 1, T.npair, nfield, tile);
  if (!b4_tile_ok(t, T.npair)) return hipErrorInvalidValue;
  switch (t.nf) {
    case 1: return analys_b4<1>(T, nfield, t.nwaves, t.rows, fm, q, s);
    case 2: return analys_b4<2>(T, nfield, t.nwaves, t.rows, fm, q, s);
    default: return analys_b4<4>(T, nfield, t.nwaves, t.rows, fm, q, s);
  }
}

hipError_t launch_leg_imex_synth_b4(const LegTablesZ64& T, int nfield, const Fields& q,
                                    const Fields& a, const float* lam, const ImexCoef& ic,
                                    float2* fm, LegTile tile, hipStream_t s) {
  LegTile t = resolve_leg_tile(TSPHERE_LEG_BF16X3, 0, T.npair, nfield, tile);
  if (t.impl != 4)
    t = resolve_leg_tile(TSPHERE_LEG_BF16X3, 0, T.npair, nfield, LegTile{0, 0, 0, 0, 4, 0});
  if (!b4_tile_ok(t, T.npair)) return hipErrorInvalidValue;
  switch (t.nf) {
    case 1: return synth_b4<1, true>(T, nfield, t.nwaves, t.rows, q, fm, s, a, lam, ic);
    case 2: return synth_b4<2, true>(T, nfield, t.nwaves, t.rows, q, fm, s, a, lam, ic);
    default: return synth_b4<4, true>(T, nfield, t.nwaves, t.rows, q, fm, s, a, lam, ic);
  }
}

hipError_t launch_leg_spectral_b4(const LegTablesZ64& T, int nfield, float2* fm, const Fields& st,
                                  const float* lam, const ImexCoef& ic, hipStream_t s) {
  if (nfield < 1 || nfield > 4 || T.npair > kB4MaxPairs) return hipErrorInvalidValue;
  // the analysis kernel's default waves (the same chunk-to-wave mapping, so
  // the same combine order), and 64 combine threads per field pair
  const int nwaves = std::max((T.npair + 63) / 64, nfield > 2 ? 2 : 1);
  if (nfield == 1) return spectral_b4<1>(T, nfield, nwaves, fm, st, lam, ic, s);
  if (nfield == 2) return spectral_b4<2>(T, nfield, nwaves, fm, st, lam, ic, s);
  return spectral_b4<4>(T, nfield, nwaves, fm, st, lam, ic, s);
}

}  // namespace tsphere
