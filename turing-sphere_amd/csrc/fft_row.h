// Row FFT building blocks of the Fourier stage (src/sht/wgsl/fourier.ts:
// 291-386): a natural-order radix-4 Stockham FFT (plus one radix-2 pass when
// log2 H is odd) over one latitude row, ping-ponging between two LDS rows.
//
// This header is self-contained device code: the library's Fourier kernels
// (sht_fourier.hip) include it, and the planner's generated grid kernels
// (FFT synthesis -> element-wise statements -> FFT analysis per latitude row,
// compiled by hipRTC) embed its text, so both run the same arithmetic.  To
// make that hold bit for bit whatever the surrounding code and compile flags,
// no function here leaves rounding to the compiler's contraction pass: every
// body says contract(off) and the complex products are written with explicit
// fmaf (one rounding per component's second product).
//
// A real row of N = nphi points is an H = N/2-point complex FFT wrapped in
// the real-FFT recombination (fourier.ts:291-308):
//   synthesis  Z[k] = (X[k] + conj X[H-k]) + i e^{+2pi i k/N} (X[k] - conj X[H-k]),
//              z = FFT_H^{+}(Z), x[2m] = Re z[m], x[2m+1] = Im z[m]
//   analysis   z[m] = x[2m] + i x[2m+1], Z = FFT_H^{-}(z),
//              X[k] = (Z[k] + conj Z[H-k])/2 + e^{-2pi i k/N} (-i)(Z[k] - conj Z[H-k])/2
// with X[0] = Re F_0, X[m] = F_m (m <= mmax), 0 above.  Twiddles come from the
// plan's f64-built table trig[k] = (cos, sin)(2 pi k / N); the radix-4
// passes take theirs from per-pass copies of trig's entries stored after the
// table on the device, held per thread in registers or per workgroup in LDS.
#pragma once

namespace tsfft {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
#pragma clang fp contract(off)
  return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) {
#pragma clang fp contract(off)
  return make_float2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ float2 csub(float2 a, float2 b) {
#pragma clang fp contract(off)
  return make_float2(a.x - b.x, a.y - b.y);
}

// Twiddle e^{SIGN 2 pi i idx / N} from the plan's (cos, sin) table.
template <int SIGN>
__device__ __forceinline__ float2 tw(const float2* __restrict__ trig, int idx) {
  const float2 t = trig[idx];
  return SIGN > 0 ? t : make_float2(t.x, -t.y);
}

// e^{SIGN i phi} from a staged e^{+i phi}.
template <int SIGN>
__device__ __forceinline__ float2 twc(float2 t) {
  return SIGN > 0 ? t : make_float2(t.x, -t.y);
}

// Launch shape of the H = 2^LOG2H point row FFT.  Every thread owns one
// radix-4 butterfly per pass (two for H = 2048), so a row gets H/4 threads
// (up to 256) and a 256-thread workgroup carries RB = 256/TPR rows.  Thread t
// of a row ends up owning the complex outputs t + TPR b, b < 4 BPT.
template <int LOG2H>
struct FftShape {
  static constexpr int H = 1 << LOG2H;
  static constexpr int N = 2 * H;  // entries of the trig table (real length)
  static constexpr int Q = H / 4;
  static constexpr int TPR = Q < 256 ? Q : 256;
  static constexpr int BPT = Q / TPR;
  static constexpr int RB = 256 / TPR;
  static constexpr bool ODD = LOG2H & 1;
  static constexpr int NS0 = ODD ? 2 : 4;  // Ns of the first pass after the first
  static constexpr int NPASS = (LOG2H - (ODD ? 1 : 2)) / 2;  // radix-4 passes after the first
};

// The pass tables that follow the N entries of a plan's trig table on the
// device: pass Ns (Ns = NS0 4^p) holds W[k - 1][jm] = trig[k jm N / (4 Ns)],
// k = 1..3, jm < Ns, k-major at offset Ns - NS0, so a thread's entries for
// consecutive jm are consecutive.  Entries in all: NS0 (4^NPASS - 1).
__host__ __device__ constexpr int pass_table_size(int log2h) {
  const int ns0 = (log2h & 1) ? 2 : 4;
  const int npass = (log2h - ((log2h & 1) ? 1 : 2)) / 2;
  return npass > 0 ? ns0 * ((1 << (2 * npass)) - 1) : 0;
}

// Twiddles of the radix-4 passes after the first: pass P (Ns = NS0 4^P)
// multiplies input k = 1..3 of butterfly j = t + TPR b by
// e^{+2 pi i k jm / (4 Ns)} = W[k - 1][jm], jm = j mod Ns, an entry of the
// pass tables after the plan's trig table.  Two holders with the same
// values, so the same arithmetic: at<P>(b, k, jm) returns the entry.
//
// PassTwiddles keeps one thread's entries in registers (3 per pass and
// distinct jm), read once when a kernel starts (the reads overlap the first
// pass's row loads; a wave's read of one table row covers at most 512
// contiguous bytes), so the passes read no twiddles from memory; for
// Ns <= TPR all of a thread's butterflies share jm and the entries of b = 0.
// For kernels whose LDS, not registers, limits the waves a CU holds.
template <int LOG2H>
struct PassTwiddles {
  using S = FftShape<LOG2H>;
  static constexpr int NP = S::NPASS > 0 ? S::NPASS : 1;
  float2 w[NP][S::BPT][3];
  __device__ __forceinline__ void load(const float2* __restrict__ trig, int t) {
    const float2* __restrict__ W = trig + S::N;
#pragma unroll
    for (int p = 0; p < S::NPASS; ++p) {
      const int Ns = S::NS0 << (2 * p);
#pragma unroll
      for (int b = 0; b < S::BPT; ++b) {
#pragma unroll
        for (int k = 1; k <= 3; ++k) {
          if (b > 0 && Ns <= S::TPR) {
            w[p][b][k - 1] = w[p][0][k - 1];
          } else {
            const int jm = (t + S::TPR * b) & (Ns - 1);
            w[p][b][k - 1] = W[(Ns - S::NS0) + (k - 1) * Ns + jm];
          }
        }
      }
    }
  }
  template <int P>
  __device__ __forceinline__ float2 at(int b, int k, int) const {
    return w[P][b][k - 1];
  }
};

// PassTwiddlesLds keeps one copy of the pass tables per workgroup of NT
// threads in LDS (`W`, pass_table_size entries): each thread copies a few
// entries, load() into registers when the kernel starts and publish() to W
// after the first pass (which needs none), whose row loads overlap the
// table's; the first barrier of the remaining passes makes W visible.  A
// workgroup then reads the tables from memory once, where per-thread
// registers read each entry once per thread: for kernels that stream many
// rows, whose memory pipeline the per-thread reads would load.
template <int LOG2H, int NT>
struct PassTwiddlesLds {
  using S = FftShape<LOG2H>;
  static constexpr int NTW = pass_table_size(LOG2H);
  static constexpr int PER = NTW > 0 ? (NTW + NT - 1) / NT : 1;
  float2* W;
  float2 v[PER];
  __device__ __forceinline__ void load(const float2* __restrict__ trig) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int i = threadIdx.x + p * NT;
      if (i < NTW) v[p] = trig[S::N + i];
    }
  }
  __device__ __forceinline__ void publish() const {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      const int i = threadIdx.x + p * NT;
      if (i < NTW) W[i] = v[p];
    }
  }
  template <int P>
  __device__ __forceinline__ float2 at(int, int k, int jm) const {
    constexpr int Ns = S::NS0 << (2 * P);
    return W[(Ns - S::NS0) + (k - 1) * Ns + jm];
  }
};

// One radix-4 Stockham pass, natural order, from src to dst (both the
// thread's row in LDS), pass P of those after the first (Ns = NS0 4^P).
// Butterfly j reads src[j + k Q], k = 0..3, and writes
// dst[4(j - jm) + jm + k Ns], jm = j mod Ns, with twiddles
// e^{SIGN 2 pi i jm k / (4 Ns)} from tw.
template <int SIGN, int LOG2H, int P, class TW>
__device__ __forceinline__ void r4_pass(const float2* src, float2* dst, const TW& tw, int t) {
#pragma clang fp contract(off)
  using S = FftShape<LOG2H>;
  constexpr int NS = S::NS0 << (2 * P);
#pragma unroll
  for (int b = 0; b < S::BPT; ++b) {
    const int j = t + S::TPR * b;
    const int jm = j & (NS - 1);
    const float2 x0 = src[j];
    const float2 x1 = cmul(src[j + S::Q], twc<SIGN>(tw.template at<P>(b, 1, jm)));
    const float2 x2 = cmul(src[j + 2 * S::Q], twc<SIGN>(tw.template at<P>(b, 2, jm)));
    const float2 x3 = cmul(src[j + 3 * S::Q], twc<SIGN>(tw.template at<P>(b, 3, jm)));
    const float2 a = cadd(x0, x2), bb = csub(x0, x2);
    const float2 c = cadd(x1, x3), e = csub(x1, x3);
    const float2 d = SIGN > 0 ? make_float2(-e.y, e.x) : make_float2(e.y, -e.x);
    const int o = 4 * (j - jm) + jm;
    dst[o] = cadd(a, c);
    dst[o + NS] = cadd(bb, d);
    dst[o + 2 * NS] = csub(a, c);
    dst[o + 3 * NS] = csub(bb, d);
  }
}

// The first pass reads its inputs through `in(idx)` (global memory or an LDS
// row, with any pre-processing folded in) instead of the ping-pong row:
// radix-2 when log2 H is odd (Ns = 1: out[2j] = u + v, out[2j+1] = u - v),
// else radix-4 without twiddles.
template <int SIGN, int LOG2H, class In>
__device__ __forceinline__ void first_pass(In in, float2* dst, int t) {
#pragma clang fp contract(off)
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

// Passes P.. of the radix-4 passes after the first, ping-ponging between the
// two LDS rows (one barrier per pass: a pass only writes the buffer the
// previous pass read).  `active` (uniform per 256-thread group) masks a
// group's arithmetic but never its barriers, so every thread of the workgroup
// meets the same barriers.  Returns the buffer that holds the result, after a
// final barrier.
template <int SIGN, int LOG2H, int P, class TW>
__device__ __forceinline__ float2* passes_from(bool active, float2* a, float2* b, const TW& tw,
                                               int t) {
  using S = FftShape<LOG2H>;
  __syncthreads();
  if constexpr (P >= S::NPASS) {
    return a;
  } else {
    if (active) r4_pass<SIGN, LOG2H, P>(a, b, tw, t);
    return passes_from<SIGN, LOG2H, P + 1>(active, b, a, tw, t);
  }
}

// Remaining passes of one row per 256-thread group (after first_pass).
template <int SIGN, int LOG2H, class TW>
__device__ __forceinline__ float2* rest_passes(float2* a, float2* b, const TW& tw, int t) {
  return passes_from<SIGN, LOG2H, 0>(true, a, b, tw, t);
}

// The same for workgroups in which several groups of 256 threads each
// transform their own row, some of them idle (`active` false).
template <int SIGN, int LOG2H, class TW>
__device__ __forceinline__ float2* rest_passes_masked(bool active, float2* a, float2* b,
                                                      const TW& tw, int t) {
  return passes_from<SIGN, LOG2H, 0>(active, a, b, tw, t);
}

// Synthesis pre-processing, fourier.ts:309-350: Z[k] from the row's
// Hermitian spectrum F[0..mmax] (mmax <= H - 1 since nphi >= 2 mmax + 1),
// m >= 1 entries scaled by mscale.
template <int LOG2H>
struct SynthInput {
  const float2* __restrict__ F;
  const float2* __restrict__ trig;
  int mmax;
  float mscale;
  // The load is unconditional, at an index clamped into the row, and the
  // entries above mmax are zeroed by a select: a load under a branch waits
  // for its data before the next one is issued.
  __device__ __forceinline__ float2 X(int k) const {
#pragma clang fp contract(off)
    const float2 v = F[k <= mmax ? k : mmax];
    const float2 s = k == 0 ? v : make_float2(mscale * v.x, mscale * v.y);
    return k <= mmax ? s : make_float2(0.f, 0.f);
  }
  __device__ __forceinline__ float2 operator()(int k) const {
#pragma clang fp contract(off)
    float2 xk = X(k);
    if (k == 0) xk.y = 0.f;
    const float2 xhk = X(FftShape<LOG2H>::H - k);  // X(H) is 0 (mmax < H): no branch
    const float2 xh = k == 0 ? make_float2(0.f, 0.f) : xhk;
    const float2 cj = make_float2(xh.x, -xh.y);
    const float2 bq = cmul(csub(xk, cj), trig[k]);
    const float2 s = cadd(xk, cj);
    return make_float2(s.x - bq.y, s.y + bq.x);
  }
};

// One row's synthesis into the ping-pong rows a, b; returns the row holding
// x[2m] = Re z[m], x[2m+1] = Im z[m] (thread t reads back its own entries
// t + TPR b, b < 4 BPT, after the final barrier).
template <int LOG2H, class TW>
__device__ __forceinline__ float2* synth_row(const float2* __restrict__ F, int mmax, float mscale,
                                            const float2* __restrict__ trig, const TW& tw,
                                            float2* a, float2* b, int t) {
  const SynthInput<LOG2H> in{F, trig, mmax, mscale};
  first_pass<+1, LOG2H>(in, a, t);
  return rest_passes<+1, LOG2H>(a, b, tw, t);
}

// One row's analysis FFT from in(j) = x[2j] + i x[2j+1]; returns z.
template <int LOG2H, class In, class TW>
__device__ __forceinline__ const float2* analys_row(In in, const TW& tw, float2* a, float2* b,
                                                   int t) {
  first_pass<-1, LOG2H>(in, a, t);
  return rest_passes<-1, LOG2H>(a, b, tw, t);
}

// Analysis post-processing, fourier.ts:352-386: X[k] from z, 0 <= k < H.
template <int LOG2H>
__device__ __forceinline__ float2 analys_coef(const float2* z, int k,
                                              const float2* __restrict__ trig) {
#pragma clang fp contract(off)
  constexpr int H = FftShape<LOG2H>::H;
  const float2 zk = z[k];
  const float2 zh = z[(H - k) & (H - 1)];
  const float2 cj = make_float2(zh.x, -zh.y);
  const float2 xe = make_float2(0.5f * (zk.x + cj.x), 0.5f * (zk.y + cj.y));
  const float2 d = make_float2(0.5f * (zk.x - cj.x), 0.5f * (zk.y - cj.y));
  const float2 xo = make_float2(d.y, -d.x);
  const float2 w = make_float2(trig[k].x, -trig[k].y);
  return cadd(xe, cmul(xo, w));
}

}  // namespace tsfft
