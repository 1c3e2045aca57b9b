// Legendre stage on the gfx950 matrix cores: split-BF16 x3 products with
// FP32 accumulation (v_mfma_f32_16x16x32_bf16).
//
// Per m the Legendre stage is a dense contraction over a per-m matrix
// P[pair][l] = ytilde_l^m(theta_pair) (parity-split into even and odd l-m):
//   synthesis  E/O[pair][col] = sum_{l even/odd} P[pair][l] Q[l][col]
//   analysis   Q[l][col]      = sum_pair P[pair][l] W[pair][col],
//              W = (G_N +/- G_S) w for even/odd l-m (leg.ts:207-219)
// with col = (field, re/im).  Every operand is split as x = hi + lo with hi,
// lo in BF16, and the product is hi*hi + hi*lo + lo*hi (the lo*lo term,
// ~2^-18 relative, is dropped); the MFMA accumulates in FP32.
//
// P is generated on the fly by the 3-term recurrence in FP64 -- the FP32
// recurrence's own rounding reaches 1e-4 at lmax 1023, while on gfx950 a
// v_fma_f64 issues at ~90% of the v_fma_f32 rate, so the FP64 state costs
// about what the FP32 one does -- from the significance seeds (the host's
// f64 values at the first significant l, see sht_legendre.hip).  BF16 has
// FP32's exponent range, so seeded values (> 1e-17) need no row scaling.
//
// Layout on the chip.  Each lane carries one latitude pair's recurrence and
// produces, per tile of consecutive l, a row of P, which must reach the MFMA
// operand layout (A/B: lane L holds 8 consecutive K values of row/column
// L % 16, K block L / 16; D: lane L holds rows 4(L/16)..+3 of column L % 16):
//  * synthesis: A = P rows (64 l = 32 even + 32 odd per tile), formed in
//    registers by a permlane transpose (no LDS), B = Q (loaded from HBM per
//    tile, split on the fly); the tile loop is wave-uniform.
//  * analysis: tile [l][pair] (32 l), A = P^T rows, B = W (split once per
//    row and held in registers); per tile the four waves' partial sums over
//    their pairs are added through LDS.
#include "tsphere_internal.h"

namespace tsphere {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f4;

__device__ __forceinline__ long long lm_base(int lmax, int m) {
  return (long long)m * (lmax + 1) - ((long long)m * (m - 1)) / 2;
}

__device__ __forceinline__ void split(float v, __bf16& hi, __bf16& lo) {
  hi = (__bf16)v;
  lo = (__bf16)(v - (float)hi);
}

__device__ __forceinline__ f4 mma3(const bf16x8& ah, const bf16x8& al, const bf16x8& bh,
                                   const bf16x8& bl, f4 c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
  return c;
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return __builtin_amdgcn_readfirstlane(v);
}

struct ChainD {
  double ct, y0, y1, s0, s1;
  int ls;
};

__device__ __forceinline__ void load_chain(const LegTables64& T, int m, int j, ChainD& c) {
  if (j < T.npair) {
    c.ct = T.ct[j];
    const SeedD sd = T.seed[(long long)m * T.npair + j];
    c.s0 = sd.y0;
    c.s1 = sd.y1;
    c.ls = sd.l;
  } else {
    c.ct = 0.0;
    c.s0 = c.s1 = 0.0;
    c.ls = 0x7fffffff;
  }
  c.y0 = c.y1 = 0.0;
}

__device__ __forceinline__ void step2(double2 a0, double2 a1, double ct, double& y0, double& y1) {
  const double t0 = fma(a0.x * y1, ct, a0.y * y0);
  y1 = fma(a1.x * t0, ct, a1.y * y1);
  y0 = t0;
}

// ------------------------------------------------------------- synthesis
// Block: 4 waves, 256 latitude pairs (one per lane); grid (m-pair, pair chunk,
// column chunk of CTC 16-column tiles).  Per tile of 64 l (32 even + 32 odd)
// each lane generates its pair's values as four BF16 planes (E hi/lo, O
// hi/lo), each four K-blocks of 8; a 4 x 4 transpose of (16-lane group x
// K-block) -- two v_permlane32_swap then two v_permlane16_swap per dword --
// leaves in register block rt exactly the MFMA A operand of row tile rt
// (pairs 16 rt .. 16 rt + 15).  No LDS tile: LDS holds only the row's (a, b).
constexpr int SY_TL = 64;                  // l per tile (32 even + 32 odd)
constexpr int SY_PADL = SY_TL + 8;         // zero (a,b) entries after a staged row

typedef __attribute__((ext_vector_type(4))) unsigned int u4;

// X[b][g] (lane group b, register block g) -> X[g][b], per dword.
__device__ __forceinline__ void transpose4x4(bf16x8 (&P)[4]) {
  u4 R[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) R[g] = __builtin_bit_cast(u4, P[g]);
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const auto r02 = __builtin_amdgcn_permlane32_swap(R[0][w], R[2][w], false, false);
    const auto r13 = __builtin_amdgcn_permlane32_swap(R[1][w], R[3][w], false, false);
    const auto r01 = __builtin_amdgcn_permlane16_swap(r02[0], r13[0], false, false);
    const auto r23 = __builtin_amdgcn_permlane16_swap(r02[1], r13[1], false, false);
    R[0][w] = r01[0];
    R[1][w] = r01[1];
    R[2][w] = r23[0];
    R[3][w] = r23[1];
  }
#pragma unroll
  for (int g = 0; g < 4; ++g) P[g] = __builtin_bit_cast(bf16x8, R[g]);
}

template <int CTC>
__global__ __launch_bounds__(256) void k_leg_synth_mfma(LegTables64 T, int nfield, Fields Fq,
                                                        float2* __restrict__ fm) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  const int L = T.lmax, M = T.mmax;
  double2* sab = reinterpret_cast<double2*>(smem);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int pair0 = blockIdx.y * 256 + wv * 64;  // this wave's first pair
  const int ct0 = blockIdx.z * CTC;              // first 16-column tile
  const int g = lane >> 4, r16 = lane & 15;

  for (int r = 0; r < 2; ++r) {
    const int mp = xcd_chunk(blockIdx.x, gridDim.x);
    const int m = r == 0 ? mp : M - mp;
    if (r == 1 && m <= mp) break;
    const int n = L - m + 1;
    const long long base = lm_base(L, m);
    __syncthreads();
    for (int i = tid; i < n + SY_PADL; i += 256)
      sab[i] = i < n ? T.ab[base + i] : make_double2(0.0, 0.0);
    __syncthreads();

    ChainD c;
    load_chain(T, m, pair0 + lane, c);
    const int lmin = wave_min(c.ls);

    f4 accE[4][CTC], accO[4][CTC];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int t = 0; t < CTC; ++t) accE[rt][t] = accO[rt][t] = f4{0.f, 0.f, 0.f, 0.f};

    for (int lt = lmin; lt <= L; lt += SY_TL) {
      // 1) this lane's P row for l = lt .. lt+63: planes [E hi, E lo, O hi, O lo],
      //    K-block kb holds even (odd) values k = 8 kb .. 8 kb + 7, l = lt + 2k (+1)
      bf16x8 P[4][4];
      const int i0 = lt - m;
#pragma unroll
      for (int tt = 0; tt < SY_TL; tt += 2) {
        if ((tt & 3) == 0 && c.ls == lt + tt) {
          c.y0 = c.s0;
          c.y1 = c.s1;
        }
        const int k = tt >> 1, kb = k >> 3, s8 = k & 7;
        __bf16 h, lo;
        split((float)c.y0, h, lo);
        P[0][kb][s8] = h;
        P[1][kb][s8] = lo;
        split((float)c.y1, h, lo);
        P[2][kb][s8] = h;
        P[3][kb][s8] = lo;
        step2(sab[i0 + tt + 2], sab[i0 + tt + 3], c.ct, c.y0, c.y1);
      }
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) transpose4x4(P[pl]);

      // 2) B operands: Q for this tile, lane column r16 of each column tile,
      //    K block g: l = lt + 2(8g + s) (even) and +1 (odd)
      bf16x8 bEh[CTC], bEl[CTC], bOh[CTC], bOl[CTC];
#pragma unroll
      for (int t = 0; t < CTC; ++t) {
        const int col = (ct0 + t) * 16 + r16;
        const int f = col >> 1;
        const bool live = f < nfield;
        const float* q = reinterpret_cast<const float*>(Fq.at(live ? f : 0)) + 2 * base;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          const int le = i0 + 2 * (8 * g + s);  // l - m of the even value
          float ve = 0.f, vo = 0.f;
          if (live && le < n) ve = q[2 * le + (col & 1)];
          if (live && le + 1 < n) vo = q[2 * (le + 1) + (col & 1)];
          __bf16 h, lo;
          split(ve, h, lo);
          bEh[t][s] = h;
          bEl[t][s] = lo;
          split(vo, h, lo);
          bOh[t][s] = h;
          bOl[t][s] = lo;
        }
      }

      // 3) 4 row tiles x CTC column tiles x (E, O) x 3 products
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
        for (int t = 0; t < CTC; ++t) {
          accE[rt][t] = mma3(P[0][rt], P[1][rt], bEh[t], bEl[t], accE[rt][t]);
          accO[rt][t] = mma3(P[2][rt], P[3][rt], bOh[t], bOl[t], accO[rt][t]);
        }
      }
    }

    // 4) north = E + O, south = E - O
    const int nlat = T.nlat;
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int t = 0; t < CTC; ++t) {
        const int col = (ct0 + t) * 16 + r16;
        const int f = col >> 1;
        if (f >= nfield) continue;
        // fm row-major per latitude: complex (f, ilat, m) at (f nlat + ilat)(M+1) + m
        float* out = reinterpret_cast<float*>(fm + (long long)f * nlat * (M + 1) + m) + (col & 1);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int j = pair0 + 16 * rt + 4 * g + q;
          if (j >= T.npair) continue;
          const float e = accE[rt][t][q], o = accO[rt][t][q];
          out[2LL * j * (M + 1)] = e + o;
          out[2LL * (nlat - 1 - j) * (M + 1)] = e - o;
        }
      }
  }
}

// -------------------------------------------------------------- analysis
// Block: one m-pair, 4 waves over all latitude pairs; wave w owns pairs
// [w*PW, (w+1)*PW), PW = 32*KB (KB 32-pair K blocks), each lane KB/2 chains.
constexpr int AN_TL = 32;                  // l per tile (16 even + 16 odd)

template <int KB, int CTC>
__global__ __launch_bounds__(256) void k_leg_analys_mfma(LegTables64 T, int nfield,
                                                         const float2* __restrict__ fm,
                                                         Fields Fq) {
  constexpr int PW = 32 * KB;              // pairs per wave
  constexpr int CH = PW / 64;              // chains per lane (KB even)
  constexpr int ROWB = PW * 2 + 16;        // bytes per LDS row (one l): PW bf16 + pad
  constexpr int PLANE = 16 * ROWB;         // 16 l of one parity, hi or lo
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ float red[4][2 * CTC][64][4];
  __shared__ int smin[4];
  const int L = T.lmax, M = T.mmax, nlat = T.nlat;
  double2* sab = reinterpret_cast<double2*>(smem);
  unsigned char* tile = smem + sizeof(double2) * (L + 1 + AN_TL + 8);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  unsigned char* wt = tile + wv * 4 * PLANE;
  const int pair0 = wv * PW;
  const int ct0 = blockIdx.y * CTC;
  const int g = lane >> 4, r16 = lane & 15;

  for (int r = 0; r < 2; ++r) {
    const int mp = xcd_chunk(blockIdx.x, gridDim.x);
    const int m = r == 0 ? mp : M - mp;
    if (r == 1 && m <= mp) break;
    const int n = L - m + 1;
    const long long base = lm_base(L, m);
    __syncthreads();
    for (int i = tid; i < n + AN_TL + 8; i += 256)
      sab[i] = i < n ? T.ab[base + i] : make_double2(0.0, 0.0);

    ChainD c[CH];
    int lmin = 0x7fffffff;
#pragma unroll
    for (int h = 0; h < CH; ++h) {
      load_chain(T, m, pair0 + h * 64 + lane, c[h]);
      lmin = min(lmin, c[h].ls);
    }
    lmin = wave_min(lmin);
    if (lane == 0) smin[wv] = lmin;

    // B operands W = (G_N +/- G_S) w, split once per row: K block kb, column
    // r16 of tile t, pairs 32 kb + 8 g + s
    bf16x8 wEh[KB][CTC], wEl[KB][CTC], wOh[KB][CTC], wOl[KB][CTC];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
      for (int t = 0; t < CTC; ++t) {
        const int col = (ct0 + t) * 16 + r16;
        const int f = col >> 1;
        const bool live = f < nfield;
        const float* gg = reinterpret_cast<const float*>(
                              fm + (long long)(live ? f : 0) * nlat * (M + 1) + m) + (col & 1);
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          const int j = pair0 + 32 * kb + 8 * g + s;
          float vp = 0.f, vm = 0.f;
          if (live && j < T.npair) {
            const int js = nlat - 1 - j;
            const float w = T.w[j];
            const float gN = gg[2LL * j * (M + 1)], gS = gg[2LL * js * (M + 1)];
            if (js == j) {
              vp = gN * w;
            } else {
              vp = (gN + gS) * w;
              vm = (gN - gS) * w;
            }
          }
          __bf16 h, lo;
          split(vp, h, lo);
          wEh[kb][t][s] = h;
          wEl[kb][t][s] = lo;
          split(vm, h, lo);
          wOh[kb][t][s] = h;
          wOl[kb][t][s] = lo;
        }
      }
    __syncthreads();  // sab and smin published
    int lblk = smin[0];
    for (int w = 1; w < 4; ++w) lblk = min(lblk, smin[w]);
    lblk = min(lblk, L + 1);

    for (int i = m + tid; i < lblk; i += 256)
      for (int f = ct0 * 8; f < min(nfield, (ct0 + CTC) * 8); ++f)
        reinterpret_cast<float2*>(const_cast<char*>(Fq.at(f)))[base + (i - m)] =
            make_float2(0.f, 0.f);

    int buf = 0;
    for (int lt = lblk; lt <= L; lt += AN_TL) {
      const int i0 = lt - m;
      // 1) P^T tile: row = l (16 even / 16 odd), column = this wave's pairs
#pragma unroll
      for (int h = 0; h < CH; ++h) {
        const int col = h * 64 + lane;  // pair within the wave
#pragma unroll
        for (int tt = 0; tt < AN_TL; tt += 2) {
          if ((tt & 3) == 0 && c[h].ls == lt + tt) {
            c[h].y0 = c[h].s0;
            c[h].y1 = c[h].s1;
          }
          const int k = tt >> 1;
          __bf16 hi, lo;
          split((float)c[h].y0, hi, lo);
          *reinterpret_cast<__bf16*>(wt + 0 * PLANE + k * ROWB + 2 * col) = hi;
          *reinterpret_cast<__bf16*>(wt + 1 * PLANE + k * ROWB + 2 * col) = lo;
          split((float)c[h].y1, hi, lo);
          *reinterpret_cast<__bf16*>(wt + 2 * PLANE + k * ROWB + 2 * col) = hi;
          *reinterpret_cast<__bf16*>(wt + 3 * PLANE + k * ROWB + 2 * col) = lo;
          step2(sab[i0 + tt + 2], sab[i0 + tt + 3], c[h].ct, c[h].y0, c[h].y1);
        }
      }
      wave_lds_sync();

      // 2) D[l][col] over this wave's K blocks
      f4 dE[CTC], dO[CTC];
#pragma unroll
      for (int t = 0; t < CTC; ++t) dE[t] = dO[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        const unsigned char* rowp = wt + r16 * ROWB + (32 * kb + 8 * g) * 2;
        const bf16x8 aEh = *reinterpret_cast<const bf16x8*>(rowp + 0 * PLANE);
        const bf16x8 aEl = *reinterpret_cast<const bf16x8*>(rowp + 1 * PLANE);
        const bf16x8 aOh = *reinterpret_cast<const bf16x8*>(rowp + 2 * PLANE);
        const bf16x8 aOl = *reinterpret_cast<const bf16x8*>(rowp + 3 * PLANE);
#pragma unroll
        for (int t = 0; t < CTC; ++t) {
          dE[t] = mma3(aEh, aEl, wEh[kb][t], wEl[kb][t], dE[t]);
          dO[t] = mma3(aOh, aOl, wOh[kb][t], wOl[kb][t], dO[t]);
        }
      }
      wave_lds_sync();  // this wave's tile reads are done before its next writes

      // 3) add the four waves' partials; D lane layout: rows 4g+q, column r16
#pragma unroll
      for (int t = 0; t < CTC; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          red[wv][2 * t][lane][q] = dE[t][q];
          red[wv][2 * t + 1][lane][q] = dO[t][q];
        }
      __syncthreads();
      for (int e = tid; e < 2 * CTC * 64 * 4; e += 256) {
        const int q = e & 3, ln = (e >> 2) & 63, tp = e >> 8;
        const float s = red[0][tp][ln][q] + red[1][tp][ln][q] + red[2][tp][ln][q] +
                        red[3][tp][ln][q];
        const int t = tp >> 1, par = tp & 1;
        const int row = 4 * (ln >> 4) + q;  // 0..15: even/odd l index
        const int col = (ct0 + t) * 16 + (ln & 15);
        const int f = col >> 1;
        const int l = lt + 2 * row + par;
        if (l <= L && f < nfield)
          reinterpret_cast<float*>(const_cast<char*>(Fq.at(f)))[2 * (base + (l - m)) + (col & 1)] = s;
      }
      __syncthreads();  // red is rewritten next tile
      (void)buf;
    }
  }
}

template <int CTC>
hipError_t synth_ctc(const LegTables64& T, int nfield, const Fields& q, float2* fm, hipStream_t s) {
  const int ncol = 2 * nfield;
  const int ct = (ncol + 15) / 16;
  dim3 grid(T.mmax / 2 + 1, (T.npair + 255) / 256, (ct + CTC - 1) / CTC);
  const size_t lds = sizeof(double2) * (T.lmax + 1 + SY_PADL);
  hipLaunchKernelGGL(k_leg_synth_mfma<CTC>, grid, dim3(256), lds, s, T, nfield, q, fm);
  return hipGetLastError();
}

template <int KB, int CTC>
hipError_t analys_kb(const LegTables64& T, int nfield, const float2* fm, const Fields& q,
                     hipStream_t s) {
  const int ncol = 2 * nfield;
  const int ct = (ncol + 15) / 16;
  constexpr int PW = 32 * KB;
  constexpr int ROWB = PW * 2 + 16;
  dim3 grid(T.mmax / 2 + 1, (ct + CTC - 1) / CTC);
  const size_t lds = sizeof(double2) * (T.lmax + 1 + AN_TL + 8) + 4 * 4 * 16 * (size_t)ROWB;
  hipLaunchKernelGGL((k_leg_analys_mfma<KB, CTC>), grid, dim3(256), lds, s, T, nfield, fm, q);
  return hipGetLastError();
}

}  // namespace

// CTC (column tiles per block) is the tile-shape knob of the synthesis.
int mfma_col_tiles(int nfield) { return (2 * nfield + 15) / 16; }

hipError_t launch_leg_synth_mfma(const LegTables64& T, int nfield, const Fields& q, float2* fm,
                                 LegTile tile, hipStream_t s) {
  if (nfield <= 0) return hipSuccess;
  switch (resolve_leg_tile(TSPHERE_LEG_BF16X3, 0, T.npair, nfield, tile).ctc) {
    case 1: return synth_ctc<1>(T, nfield, q, fm, s);
    case 2: return synth_ctc<2>(T, nfield, q, fm, s);
    case 4: return synth_ctc<4>(T, nfield, q, fm, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_leg_analys_mfma(const LegTables64& T, int nfield, const float2* fm,
                                  const Fields& q, LegTile tile, hipStream_t s) {
  if (nfield <= 0) return hipSuccess;
  // pairs per wave: 4 waves cover npair, rounded up to 64-pair multiples
  const int pw = ((T.npair + 3) / 4 + 63) / 64 * 64;
  const int kb = pw / 32;
  const int ctc = resolve_leg_tile(TSPHERE_LEG_BF16X3, 1, T.npair, nfield, tile).ctc;
  if (ctc == 2) {
    switch (kb) {
      case 2: return analys_kb<2, 2>(T, nfield, fm, q, s);
      case 4: return analys_kb<4, 2>(T, nfield, fm, q, s);
      case 6: return analys_kb<6, 2>(T, nfield, fm, q, s);
      case 8: return analys_kb<8, 2>(T, nfield, fm, q, s);
      default: return hipErrorInvalidValue;
    }
  }
  if (ctc != 1) return hipErrorInvalidValue;
  switch (kb) {
    case 2: return analys_kb<2, 1>(T, nfield, fm, q, s);
    case 4: return analys_kb<4, 1>(T, nfield, fm, q, s);
    case 6: return analys_kb<6, 1>(T, nfield, fm, q, s);
    case 8: return analys_kb<8, 1>(T, nfield, fm, q, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace tsphere
