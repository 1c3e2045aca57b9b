// Internal declarations shared by the C-ABI runtime (tsphere_runtime.hip)
// and the gfx950 kernels (sht_*.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "tsphere.h"

namespace tsphere {

// Where the f32 recurrence of one (m, latitude pair) starts: the first l at
// which |ytilde_l^m| exceeds ACCURACY, rounded down to m + 4k, with the f64
// values ytilde_l and ytilde_{l+1} there (l = INT_MAX: never significant up
// to lmax).  Built on the host by a scaled f64 recurrence.
struct SeedL {
  float y0, y1;
  int l;
  int pad;
};

// fp32 extended-range constants of the recurrence (the SHTNS values the
// reference injects, src/sht/wgsl/common.ts:17-20): a lane whose running value
// carries a negative exponent `ny` does not accumulate until the value is back
// above ACCURACY in true magnitude.
constexpr float kScale = 72057594037927936.0f;         // 2^56
constexpr float kInvScale = 1.0f / 72057594037927936.0f;  // 2^-56, exact
constexpr float kAccuracy = 1e-15f;
constexpr float kRescaleThr = kAccuracy * kScale + 1.0f;  // ~73.06

// Device tables of the FP32 form.  The recurrence runs on the rescaled
// functions z_l = ytilde_l / c_l with c_m = c_{m+1} = 1 and c_l = b_l c_{l-2}
// (coeffs.ts:28-56's b_l^m), for which it reads
//   z_l = alpha_l cos(theta) z_{l-1} + z_{l-2},  alpha_l = a_l c_{l-1} / c_l
// -- one coefficient per step instead of two, so one VALU op less per l.
// |c_l| stays within [0.24, 1.13] up to lmax 1023 (no range issue in f32);
// the synthesis takes c_l Q_lm as its coefficients and the analysis
// multiplies its sums by c_l.
struct LegTables {
  const float* ab;       // [nlm + pad]: alpha_l^m, m-major like Q (0 at l = m)
  const float* cz;       // [nlm + pad]: c_l^m
  const float2* ct;      // [npair]: cos(theta_j) of the northern latitude of pair j,
                         //   as an unevaluated f32 sum (hi, lo) of the f64 node
  const float* w;        // [nlat]: Gauss weight * 2 pi / nphi
  const SeedL* seed;     // [(mmax+1) * npair]: z at the seed position
  int lmax, mmax, nlat, npair;
};

// The same seeds in f64, for the FP64 recurrence of the MFMA path.
struct SeedD {
  double y0, y1;
  int l;
  int pad;
};

// Tables of the split-BF16 MFMA path (its recurrence runs in FP64).
struct LegTables64 {
  const double2* ab;     // [nlm + pad]: (a_l^m, b_l^m) in f64
  const double* ct;      // [npair]: cos(theta_j), f64
  const float* w;        // [nlat]: Gauss weight * 2 pi / nphi
  const SeedD* seed;     // [(mmax+1) * npair]
  int lmax, mmax, nlat, npair;
};

// Tables of the FP64-recurrence form: the rescaled recurrence of LegTables
// in f64.
struct LegTablesZ64 {
  const double* ab;      // [nlm + pad]: alpha_l^m in f64
  const float* cz;       // [nlm + pad]: c_l^m
  const double* ct;      // [npair]: cos(theta_j), f64
  const float* w;        // [nlat]: Gauss weight * 2 pi / nphi
  const SeedD* seed;     // [(mmax+1) * npair]: z at the seed position, f64
  int lmax, mmax, nlat, npair;
};

// Tile shape of a Legendre launch (0 = the built-in choice for that entry).
//  FP32 form: impl 1 = contraction on the VALU, 2 = on the FP32 matrix cores;
//  nf fields per workgroup (1, 2, 4; 8 for the block kernels with k <= 2),
//  nwaves waves, k latitude pairs per lane
//  (1, 2, 4, 8 on the VALU, 1, 2, 4 on the matrix cores; nwaves * 64 * k must
//  cover the pairs).
//  rows (VALU kernels): m-rows per workgroup, 2 = the pair (m, M - m), 1 = one.
//  impl 5 = per-parity matrix products (16x16x4, columns (field, re/im) of 8
//  fields): nf 8 or 16, k 1, rows 1 (the 8-field synthesis also 2: the
//  consecutive rows 2q, 2q + 1), nwaves * 64 >= pairs (<= 512 pairs).
//  Split-BF16 form: impl 4 = 4x4x4 matrix blocks (nf, nwaves, rows as the
//  FP32 block kernels, k = 1), else ctc column tiles (16 columns each) of the
//  16x16x32 kernels per workgroup.
struct LegTile {
  int nf, nwaves, k, ctc, impl, rows;
};

struct Plan {
  tsphere_sht_config cfg{};
  int fourier = TSPHERE_FOURIER_FFT;
  int npair = 0;
  long long nlm = 0;
  float* d_alz = nullptr;     // FP32 form: alpha, z seeds
  float* d_cz = nullptr;      // FP32 and FP64-recurrence forms: c_l
  float2* d_ct = nullptr;
  float* d_w = nullptr;
  float* d_ones = nullptr;  // [nlat] of 1.0: the unit weights of the synthesis transpose
  SeedL* d_seedz = nullptr;
  double* d_alz64 = nullptr;  // FP64-recurrence form: alpha, z seeds in f64
  SeedD* d_seedz64 = nullptr;
  double2* d_ab64 = nullptr;   // split-BF16 form: (a, b), seeds in f64
  double* d_ct64 = nullptr;    // both f64 forms
  SeedD* d_seed64 = nullptr;
  LegTile tile[2] = {};  // [synthesis, analysis]; zeros = built-in heuristic
  float2* d_trig = nullptr;
  float2* d_fm = nullptr;  // Fourier scratch, max_fields * (mmax+1) * nlat
  int device = 0;          // the GPU the plan's tables live on
  size_t bytes = 0;
  std::vector<double> theta, cos_theta, weights;
  LegTables tables(bool unit_w = false) const {
    return LegTables{d_alz, d_cz, d_ct, unit_w ? d_ones : d_w, d_seedz, cfg.lmax, cfg.mmax,
                     cfg.nlat, npair};
  }
  LegTablesZ64 tablesz64(bool unit_w = false) const {
    return LegTablesZ64{d_alz64,  d_cz,     d_ct64,   unit_w ? d_ones : d_w, d_seedz64,
                        cfg.lmax, cfg.mmax, cfg.nlat, npair};
  }
  LegTables64 tables64(bool unit_w = false) const {
    return LegTables64{d_ab64,    d_ct64,   unit_w ? d_ones : d_w, d_seed64, cfg.lmax,
                       cfg.mmax, cfg.nlat, npair};
  }
};

// Where the fields of a batched call live: field f is at
// p[f / per] + (f % per) * stride bytes -- a pointer list (stride 0, per 1),
// one strided base (per = INT_MAX), or several strided bases (an ensemble's
// state variables, each holding `per` runs).  Passed by value in the kernel
// arguments; `f` is wave-uniform, so the lookup is scalar.
constexpr int kMaxList = 8;
struct Fields {
  const char* p[kMaxList];
  long long stride;  // bytes
  int per;           // fields per pointer
  __host__ __device__ const char* at(int f) const {
    return per <= 1 ? p[f] : p[f / per] + (long long)(f % per) * stride;
  }
};
inline Fields strided(const void* base, long long stride_bytes) {
  Fields F{};
  F.p[0] = static_cast<const char*>(base);
  F.stride = stride_bytes;
  F.per = 0x7fffffff;
  return F;
}


// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b runs on
// XCD b % 8, each XCD with its own L2).  This maps b to a contiguous range
// per XCD, so neighbouring m (which share the cache lines of the latitude
// rows of the Fourier scratch) are handled under one L2.  A bijection on
// [0, nb) for any nb.
__device__ __forceinline__ int xcd_chunk(int b, int nb) {
  const int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + (x < r ? x : r) + (b >> 3);
}

// Host-side table construction (no device needed).
struct HostTables {
  std::vector<float> ab;       // 2*(nlm + pad)
  std::vector<float> amm;      // mmax+1
  std::vector<float> ct, st, w;  // nlat each
  std::vector<float> ct_lo;      // nlat: f32(x - f32(x)), the low half of cos(theta)
  std::vector<SeedL> seed;     // (mmax+1)*npair
  std::vector<SeedD> seed64;   // (mmax+1)*npair, same seeds in f64
  std::vector<double> ab64;    // 2*(nlm + pad), f64 coefficients
  std::vector<double> alz64, cz64;  // nlm + pad: alpha_l, c_l of the rescaled recurrence
  std::vector<SeedD> seedz64;       // (mmax+1)*npair: the seeds as z = ytilde / c
  std::vector<float> trig;     // 2*nphi
  std::vector<double> x, wg;   // f64 Gauss nodes / weights
};
constexpr int kAbPad = 40;  // zero entries after the last row (kernels read alpha past a row)
void build_host_tables(const tsphere_sht_config& cfg, HostTables& t);
long long nlm_calc(int lmax, int mmax);

// Kernel launchers (sht_legendre.hip, sht_fourier.hip).  All enqueue on `s`.
// Spectral fields (Q, 2*nlm floats each) and spatial fields (nlat*nphi
// floats each) are addressed through Fields; the Fourier scratch fm is the
// plan's, contiguous per field.
// `tile` fields left 0 take the launcher's own choice; an invalid shape
// returns hipErrorInvalidValue without launching.
hipError_t launch_leg_synth(const LegTables& T, int nfield, const Fields& q, float2* fm,
                            LegTile tile, hipStream_t s);
hipError_t launch_leg_analys(const LegTables& T, int nfield, const float2* fm, const Fields& q,
                             LegTile tile, hipStream_t s);
hipError_t launch_leg_synth_mfma(const LegTables64& T, int nfield, const Fields& q, float2* fm,
                                 LegTile tile, hipStream_t s);
hipError_t launch_leg_analys_mfma(const LegTables64& T, int nfield, const float2* fm,
                                  const Fields& q, LegTile tile, hipStream_t s);
// The spectral op: Legendre analysis of the Fourier scratch (nfield fields),
// the diagonal implicit step X_new = (X + s1 A) / (1 + s2 lam) of state field
// spc[f] for analysed field f, and the Legendre synthesis of the new state
// into the same scratch (field slot = state index).  FP32 and FP64-recurrence
// forms; up to 4 fields.
struct ImexCoef {
  float s1[4], s2[4];  // per analysed field
  int spc[4];          // the state field (species) each analysed field updates
  // batched implicit step (fold only): (s1, s2) of field f at dev[2 f],
  // dev[2 f + 1] in device memory, for any field count; null: s1, s2 above
  const float* dev;
};
// The implicit step (s1[f], s2[f] per field; spc unused) of every state
// field q[f] from its analysed values a[f], in front of the Legendre
// synthesis of the new state into fm: the next step's first stage with the
// implicit solve of the last one folded in.  Block-matrix kernels (impl 3).
hipError_t launch_leg_imex_synth(const LegTables& T, int nfield, const Fields& q, const Fields& a,
                                const float* lam, const ImexCoef& ic, float2* fm, LegTile tile,
                                hipStream_t s);
hipError_t launch_leg_imex_synth_f64r(const LegTablesZ64& T, int nfield, const Fields& q,
                                     const Fields& a, const float* lam, const ImexCoef& ic,
                                     float2* fm, LegTile tile, hipStream_t s);
hipError_t launch_leg_spectral(const LegTables& T, int nfield, float2* fm, const Fields& st,
                               const float* lam, const ImexCoef& ic, hipStream_t s);
hipError_t launch_leg_spectral_f64r(const LegTablesZ64& T, int nfield, float2* fm,
                                    const Fields& st, const float* lam, const ImexCoef& ic,
                                    hipStream_t s);
// The split-BF16 form on 4x4x4 matrix blocks (tile impl 4, the built-in
// choice up to kB4MaxPairs latitude pairs: one chain per lane, 8 waves), on
// the FP64-recurrence tables; the spectral op and the fused implicit step as
// in the FP32 forms.
constexpr int kB4MaxPairs = 512;
hipError_t launch_leg_synth_b4(const LegTablesZ64& T, int nfield, const Fields& q, float2* fm,
                               LegTile tile, hipStream_t s);
hipError_t launch_leg_analys_b4(const LegTablesZ64& T, int nfield, const float2* fm,
                                const Fields& q, LegTile tile, hipStream_t s);
hipError_t launch_leg_imex_synth_b4(const LegTablesZ64& T, int nfield, const Fields& q,
                                    const Fields& a, const float* lam, const ImexCoef& ic,
                                    float2* fm, LegTile tile, hipStream_t s);
hipError_t launch_leg_spectral_b4(const LegTablesZ64& T, int nfield, float2* fm, const Fields& st,
                                  const float* lam, const ImexCoef& ic, hipStream_t s);
// The FP64-recurrence form: the VALU kernels with an f64 recurrence.
hipError_t launch_leg_synth_f64r(const LegTablesZ64& T, int nfield, const Fields& q, float2* fm,
                                 LegTile tile, hipStream_t s);
hipError_t launch_leg_analys_f64r(const LegTablesZ64& T, int nfield, const float2* fm,
                                  const Fields& q, LegTile tile, hipStream_t s);
// The analysis of nfield fields with the implicit step of each behind it:
// state q[f] = (q[f] + s1 A) / (1 + s2 lam), A its analysed values, (s1, s2)
// = coef[2 f], coef[2 f + 1] in device memory.  The 4x4-block and per-parity
// product tiles (impl 3, 5) apply the step where they form each analysed
// value and leave a[] untouched; other tiles write A to a[] and run the step
// (launch_imex_fields) behind.
hipError_t launch_imex_fields(long long nlm, int nfield, const Fields& a, const Fields& q,
                              const float* lam, const float* coef, hipStream_t s);
hipError_t launch_leg_analys_imex(const LegTables& T, int nfield, const float2* fm,
                                  const Fields& a, const Fields& q, const float* lam,
                                  const float* coef, LegTile tile, hipStream_t s);
hipError_t launch_leg_analys_imex_f64r(const LegTablesZ64& T, int nfield, const float2* fm,
                                       const Fields& a, const Fields& q, const float* lam,
                                       const float* coef, LegTile tile, hipStream_t s);
// The shape a launch of `nfield` fields would use (tile fields resolved).
LegTile resolve_leg_tile(int legendre, int stage, int npair, int nfield, LegTile tile);
// Scaling hooks of the Fourier stage (identity by default): mscale multiplies
// every m >= 1 coefficient, rowscale (if set) each synthesised latitude row.
struct FourierScale {
  float mscale = 1.f;
  const float* rowscale = nullptr;
};
hipError_t launch_fourier_synth(int fourier, int mmax, int nlat, int nphi, const float2* trig,
                                int nfield, const float2* fm, const Fields& spat, hipStream_t s,
                                FourierScale sc = {});
hipError_t launch_fourier_analys(int fourier, int mmax, int nlat, int nphi, const float2* trig,
                                 int nfield, const Fields& spat, float2* fm, hipStream_t s,
                                 FourierScale sc = {});
hipError_t launch_accumulate(float* dst, const float* src, long long n, hipStream_t s);
hipError_t launch_fill(float* dst, float v, long long n, hipStream_t s);

}  // namespace tsphere
