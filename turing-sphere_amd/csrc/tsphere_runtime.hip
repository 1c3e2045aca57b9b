// C-ABI runtime of the gfx950 spectral IMEX library (declared in
// include/tsphere.h): error plumbing, device memory, streams/events, the
// transform plan (host f64 precompute -> f32 device tables), hipRTC modules
// for the generated element-wise kernels, and HIP-graph step capture.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "fft_row.h"
#include "tsphere_internal.h"

using namespace tsphere;

// ---------------------------------------------------------------- errors
namespace {
thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[2048];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
    return fail(TSPHERE_ERR_NODEVICE, "%s: no GPU (%s)", what, hipGetErrorString(e));
  return fail(TSPHERE_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define TS_HIP(call, what)                       \
  do {                                           \
    hipError_t e_ = (call);                      \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

int ok() {
  g_err.clear();
  return TSPHERE_OK;
}

hipStream_t S(tsphere_stream s) { return reinterpret_cast<hipStream_t>(s); }

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
}  // namespace

// ------------------------------------------------------------ host tables
namespace tsphere {

long long nlm_calc(int lmax, int mmax) {
  return (long long)(mmax + 1) * (lmax + 1) - ((long long)mmax * (mmax + 1)) / 2;
}

namespace {
long long lm_index(int lmax, int l, int m) {
  return (long long)m * (lmax + 1) - ((long long)m * (m - 1)) / 2 + (l - m);
}

// Gauss-Legendre nodes (decreasing) and weights: Newton on P_n from the same
// initial guess and stopping rule as src/sht/gauss.ts:11-51.
void gauss(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const int half = (n + 1) >> 1;
  for (int i = 0; i < half; i++) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double pp = 0.0;
    for (int iter = 0; iter < 100; iter++) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      const double dz = p1 / pp;
      z -= dz;
      if (std::fabs(dz) < 1e-15 * std::fabs(z) + 1e-300) {
        double q1 = 1.0, q2 = 0.0;
        for (int j = 1; j <= n; j++) {
          const double q3 = q2;
          q2 = q1;
          q1 = ((2.0 * j - 1.0) * z * q2 - (j - 1.0) * q3) / j;
        }
        pp = n * (z * q1 - q2) / (z * z - 1.0);
        z -= q1 / pp;
        break;
      }
    }
    x[i] = z;
    x[n - 1 - i] = -z;
    const double wi = 2.0 / ((1.0 - z * z) * pp * pp);
    w[i] = wi;
    w[n - 1 - i] = wi;
  }
  if (n & 1) x[half - 1] = 0.0;
}
}  // namespace

void build_host_tables(const tsphere_sht_config& c, HostTables& t) {
  const int lmax = c.lmax, mmax = c.mmax, nlat = c.nlat, nphi = c.nphi;
  const long long nlm = nlm_calc(lmax, mmax);
  const int npair = (nlat + 1) / 2;

  // Recurrence coefficients, coeffs.ts:28-56 (f64, then f32).
  std::vector<double> amm(mmax + 1);
  double tt = 1.0 / (4.0 * M_PI);
  amm[0] = std::sqrt(tt);
  for (int m = 1; m <= mmax; m++) {
    tt *= (2.0 * m + 1.0) / (2.0 * m);
    amm[m] = -std::sqrt(tt);
    if (m % 2 == 0) amm[m] = -amm[m];
  }
  t.ab64.assign(2 * (nlm + kAbPad), 0.0);
  for (int m = 0; m <= mmax; m++) {
    if (m + 1 <= lmax) t.ab64[2 * lm_index(lmax, m + 1, m)] = std::sqrt(2.0 * m + 3.0);
    for (int l = m + 2; l <= lmax; l++) {
      const long long lm = lm_index(lmax, l, m);
      const double t1 = (double)(l + m) * (l - m);
      const double t2 = (double)(l - 1 + m) * (l - 1 - m);
      t.ab64[2 * lm] = std::sqrt(((2.0 * l + 1.0) * (2.0 * l - 1.0)) / t1);
      t.ab64[2 * lm + 1] = -std::sqrt(((2.0 * l + 1.0) / (2.0 * l - 3.0)) * (t2 / t1));
    }
  }
  t.ab.assign(2 * (nlm + kAbPad), 0.0f);
  for (size_t i = 0; i < t.ab64.size(); i++) t.ab[i] = (float)t.ab64[i];
  // Rescaled recurrence: c_m = c_{m+1} = 1, c_l = b_l c_{l-2},
  // alpha_l = a_l c_{l-1} / c_l (alpha_{m+1} = a_{m+1}).
  t.alz64.assign(nlm + kAbPad, 0.0);
  t.cz64.assign(nlm + kAbPad, 0.0);
  for (int m = 0; m <= mmax; m++) {
    const long long b0 = lm_index(lmax, m, m);
    t.cz64[b0] = 1.0;
    if (m + 1 <= lmax) {
      t.cz64[b0 + 1] = 1.0;
      t.alz64[b0 + 1] = t.ab64[2 * (b0 + 1)];
    }
    for (int l = m + 2; l <= lmax; l++) {
      const long long i = b0 + (l - m);
      t.cz64[i] = t.ab64[2 * i + 1] * t.cz64[i - 2];
      t.alz64[i] = t.ab64[2 * i] * t.cz64[i - 1] / t.cz64[i];
    }
  }
  t.amm.resize(mmax + 1);
  for (int m = 0; m <= mmax; m++) t.amm[m] = (float)amm[m];

  // Grid tables, sht.ts:194-199.
  gauss(nlat, t.x, t.wg);
  t.ct.resize(nlat);
  t.st.resize(nlat);
  t.w.resize(nlat);
  t.ct_lo.resize(nlat);
  for (int i = 0; i < nlat; i++) {
    t.ct[i] = (float)t.x[i];
    t.ct_lo[i] = (float)(t.x[i] - (double)t.ct[i]);
    t.st[i] = (float)std::sqrt(1.0 - t.x[i] * t.x[i]);
    t.w[i] = (float)(t.wg[i] * ((2.0 * M_PI) / nphi));
  }

  // Significance seeds.  Walk the recurrence of every (m, latitude pair) in
  // f64 with an explicit power-of-two exponent (true value = y * 2^E, so
  // amm[m] sin^m(theta) never underflows) until the first l where
  // |ytilde_l| > ACCURACY (the reference's significance rule, common.ts:19),
  // and keep the values at the last l = m + 4k at or before it.
  t.seed.resize((size_t)(mmax + 1) * npair);
  t.seed64.resize((size_t)(mmax + 1) * npair);
  t.seedz64.resize((size_t)(mmax + 1) * npair);
  const double thr = (double)kAccuracy;
  for (int m = 0; m <= mmax; m++) {
    const double la = std::log2(std::fabs(amm[m]));
    const double sg = amm[m] < 0 ? -1.0 : 1.0;
    for (int j = 0; j < npair; j++) {
      const double x = t.x[j];
      const double st = std::sqrt(1.0 - x * x);
      const double e0 = la + (m == 0 ? 0.0 : m * std::log2(st));
      int E = (int)std::floor(e0);
      double y0 = sg * std::exp2(e0 - E);
      double y1 = (m < lmax) ? std::sqrt(2.0 * m + 3.0) * x * y0 : 0.0;
      SeedL seed{0.f, 0.f, 0x7fffffff, 0};
      SeedD seed64{0.0, 0.0, 0x7fffffff, 0};
      double c0 = y0, c1 = y1;
      int cE = E, cl = m;
      for (int l = m; l <= lmax; l += 2) {
        if (((l - m) & 3) == 0) {  // candidate seed position
          c0 = y0;
          c1 = y1;
          cE = E;
          cl = l;
        }
        const bool sig0 = std::fabs(y0) * std::exp2((double)E) > thr;
        const bool sig1 = l + 1 <= lmax && std::fabs(y1) * std::exp2((double)E) > thr;
        if (sig0 || sig1) {
          seed64.y0 = std::ldexp(c0, cE);
          seed64.y1 = (cl + 1 <= lmax) ? std::ldexp(c1, cE) : 0.0;
          seed64.l = cl;
          seed.y0 = (float)seed64.y0;
          seed.y1 = (float)seed64.y1;
          seed.l = cl;
          break;
        }
        // two steps of coeffs.ts:28-56's recurrence, in f64
        for (int s = 0; s < 2; s++) {
          const int ll = l + 2 + s;
          if (ll > lmax) break;
          const double t1 = (double)(ll + m) * (ll - m);
          const double t2 = (double)(ll - 1 + m) * (ll - 1 - m);
          const double a = std::sqrt(((2.0 * ll + 1.0) * (2.0 * ll - 1.0)) / t1);
          const double b = -std::sqrt(((2.0 * ll + 1.0) / (2.0 * ll - 3.0)) * (t2 / t1));
          const double yn = a * x * (s == 0 ? y1 : y0) + b * (s == 0 ? y0 : y1);
          if (s == 0) y0 = yn; else y1 = yn;
        }
        if (std::fabs(y0) > 0x1p200 || std::fabs(y1) > 0x1p200) {
          y0 *= 0x1p-200;
          y1 *= 0x1p-200;
          E += 200;
        }
      }
      t.seed[(size_t)m * npair + j] = seed;
      t.seed64[(size_t)m * npair + j] = seed64;
      SeedD sz = seed64;
      if (sz.l <= lmax) {
        const long long i = lm_index(lmax, sz.l, m);
        sz.y0 = seed64.y0 / t.cz64[i];
        sz.y1 = sz.l + 1 <= lmax ? seed64.y1 / t.cz64[i + 1] : 0.0;
      }
      t.seedz64[(size_t)m * npair + j] = sz;
    }
  }

  // Twiddles in f64 (sht.ts:200-205).
  t.trig.resize(2 * (size_t)nphi);
  for (int k = 0; k < nphi; k++) {
    t.trig[2 * k] = (float)std::cos((2.0 * M_PI * k) / nphi);
    t.trig[2 * k + 1] = (float)std::sin((2.0 * M_PI * k) / nphi);
  }
}

}  // namespace tsphere

// ================================================================= C ABI
extern "C" {

int tsphere_version(void) { return 100; }

const char* tsphere_last_error(void) { return g_err.c_str(); }

int tsphere_device_count(int* count) {
  if (!count) return fail(TSPHERE_ERR_INVALID, "count is NULL");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  (void)hipGetLastError();
  *count = n;
  return ok();
}

// ---------------------------------------------------------- plumbing
int tsphere_stream_create(tsphere_stream* out) {
  if (!out) return fail(TSPHERE_ERR_INVALID, "out is NULL");
  hipStream_t s = nullptr;
  TS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  *out = s;
  return ok();
}

int tsphere_stream_destroy(tsphere_stream s) {
  if (s) TS_HIP(hipStreamDestroy(S(s)), "hipStreamDestroy");
  return ok();
}

int tsphere_stream_sync(tsphere_stream s) {
  TS_HIP(hipStreamSynchronize(S(s)), "hipStreamSynchronize");
  return ok();
}

int tsphere_device_sync(void) {
  TS_HIP(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return ok();
}

int tsphere_malloc(size_t bytes, void** out) {
  if (!out) return fail(TSPHERE_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (bytes == 0) return ok();
  TS_HIP(hipMalloc(out, bytes), "hipMalloc");
  return ok();
}

int tsphere_free(void* ptr) {
  if (ptr) TS_HIP(hipFree(ptr), "hipFree");
  return ok();
}

int tsphere_memcpy_htod(void* dst, const void* src, size_t bytes, tsphere_stream s) {
  if (bytes == 0) return ok();
  if (!dst || !src) return fail(TSPHERE_ERR_INVALID, "memcpy_htod: NULL pointer");
  TS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(s)), "hipMemcpyAsync(H2D)");
  return ok();
}

int tsphere_memcpy_dtoh(void* dst, const void* src, size_t bytes, tsphere_stream s) {
  if (bytes == 0) return ok();
  if (!dst || !src) return fail(TSPHERE_ERR_INVALID, "memcpy_dtoh: NULL pointer");
  TS_HIP(hipStreamSynchronize(S(s)), "hipStreamSynchronize");
  TS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy(D2H)");
  return ok();
}

int tsphere_memcpy_dtod(void* dst, const void* src, size_t bytes, tsphere_stream s) {
  if (bytes == 0) return ok();
  if (!dst || !src) return fail(TSPHERE_ERR_INVALID, "memcpy_dtod: NULL pointer");
  TS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(s)), "hipMemcpyAsync(D2D)");
  return ok();
}

int tsphere_fill_f32(float* dst, float value, long long count, tsphere_stream s) {
  if (count < 0 || (count > 0 && !dst)) return fail(TSPHERE_ERR_INVALID, "fill: bad arguments");
  TS_HIP(launch_fill(dst, value, count, S(s)), "fill kernel");
  return ok();
}

struct tsphere_event_t {
  hipEvent_t e;
};

int tsphere_event_create(tsphere_event* out) {
  if (!out) return fail(TSPHERE_ERR_INVALID, "out is NULL");
  hipEvent_t e;
  TS_HIP(hipEventCreate(&e), "hipEventCreate");
  *out = new tsphere_event_t{e};
  return ok();
}

int tsphere_event_destroy(tsphere_event e) {
  if (e) {
    hipError_t r = hipEventDestroy(e->e);
    delete e;
    if (r != hipSuccess) return hip_fail(r, "hipEventDestroy");
  }
  return ok();
}

int tsphere_stream_wait(tsphere_stream s, tsphere_event e) {
  if (!e) return fail(TSPHERE_ERR_INVALID, "event is NULL");
  TS_HIP(hipStreamWaitEvent(S(s), e->e, 0), "hipStreamWaitEvent");
  return ok();
}

int tsphere_event_record(tsphere_event e, tsphere_stream s) {
  if (!e) return fail(TSPHERE_ERR_INVALID, "event is NULL");
  TS_HIP(hipEventRecord(e->e, S(s)), "hipEventRecord");
  return ok();
}

int tsphere_event_elapsed(tsphere_event a, tsphere_event b, float* ms) {
  if (!a || !b || !ms) return fail(TSPHERE_ERR_INVALID, "event_elapsed: NULL argument");
  TS_HIP(hipEventSynchronize(b->e), "hipEventSynchronize");
  TS_HIP(hipEventElapsedTime(ms, a->e, b->e), "hipEventElapsedTime");
  return ok();
}

// ------------------------------------------------------------- the plan
struct tsphere_plan_t : Plan {};

static int validate(const tsphere_sht_config* c) {
  if (!c) return fail(TSPHERE_ERR_INVALID, "config is NULL");
  // Messages of layout.ts:34-43.
  if (c->lmax < 1) return fail(TSPHERE_ERR_INVALID, "lmax must be an integer >= 1 (got %d)", c->lmax);
  if (c->mmax < 0 || c->mmax > c->lmax)
    return fail(TSPHERE_ERR_INVALID, "mmax must be an integer in [0, lmax] (got %d)", c->mmax);
  if (c->nlat <= c->lmax)
    return fail(TSPHERE_ERR_INVALID,
                "nlat must be an integer > lmax for exact Gauss quadrature (got nlat=%d, lmax=%d)",
                c->nlat, c->lmax);
  if (c->nphi < 2 * c->mmax + 1)
    return fail(TSPHERE_ERR_INVALID,
                "nphi must be an integer >= 2*mmax+1 to avoid aliasing (got nphi=%d, mmax=%d)",
                c->nphi, c->mmax);
  if (c->nlat > 4096)
    return fail(TSPHERE_ERR_UNSUPPORTED, "nlat=%d exceeds the library's limit of 4096", c->nlat);
  if (c->fourier < TSPHERE_FOURIER_AUTO || c->fourier > TSPHERE_FOURIER_DFT)
    return fail(TSPHERE_ERR_INVALID, "unknown fourier mode %d", c->fourier);
  if (c->legendre != TSPHERE_LEG_FP32 && c->legendre != TSPHERE_LEG_BF16X3 &&
      c->legendre != TSPHERE_LEG_F64REC)
    return fail(TSPHERE_ERR_INVALID, "unknown legendre mode %d", c->legendre);
  if (c->max_fields < 1) return fail(TSPHERE_ERR_INVALID, "max_fields must be >= 1");
  return TSPHERE_OK;
}

// The FFT path needs a power-of-two nphi whose half fits the per-wave LDS
// slice (the reference's fftFits test, sht.ts:169-172, with gfx950 limits).
static bool fft_fits(int nphi) { return is_pow2(nphi) && nphi >= 8 && nphi <= 4096; }

int tsphere_plan_tables(const tsphere_sht_config* cfg, float* ab, float* amm, float* ctstw,
                        float* seed_y0, float* seed_y1, int* seed_l, float* trig) {
  if (int r = validate(cfg)) return r;
  HostTables t;
  build_host_tables(*cfg, t);
  const long long nlm = nlm_calc(cfg->lmax, cfg->mmax);
  const int nlat = cfg->nlat, npair = (nlat + 1) / 2;
  if (ab) std::memcpy(ab, t.ab.data(), sizeof(float) * 2 * nlm);
  if (amm) std::memcpy(amm, t.amm.data(), sizeof(float) * (cfg->mmax + 1));
  if (ctstw) {
    for (int i = 0; i < nlat; i++) {
      ctstw[i] = t.ct[i];
      ctstw[nlat + i] = t.st[i];
      ctstw[2 * nlat + i] = t.w[i];
    }
  }
  const size_t ns = (size_t)(cfg->mmax + 1) * npair;
  for (size_t i = 0; i < ns; i++) {
    if (seed_y0) seed_y0[i] = t.seed[i].y0;
    if (seed_y1) seed_y1[i] = t.seed[i].y1;
    if (seed_l) seed_l[i] = t.seed[i].l;
  }
  if (trig) std::memcpy(trig, t.trig.data(), sizeof(float) * 2 * cfg->nphi);
  return ok();
}

static void plan_free(Plan* p) {
  if (!p) return;
  for (void* ptr : {(void*)p->d_alz, (void*)p->d_cz, (void*)p->d_ct, (void*)p->d_w,
                    (void*)p->d_ones, (void*)p->d_seedz, (void*)p->d_alz64, (void*)p->d_seedz64,
                    (void*)p->d_ab64, (void*)p->d_ct64, (void*)p->d_seed64, (void*)p->d_trig,
                    (void*)p->d_fm})
    if (ptr) (void)hipFree(ptr);
}

}  // extern "C"

template <typename T>
static hipError_t upload(T** dst, const T* src, size_t n) {
  hipError_t e = hipMalloc((void**)dst, sizeof(T) * n);
  if (e != hipSuccess) return e;
  return hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice);
}

extern "C" {

// ------------------------------------- the synthesis that allocates nothing
// tsphere_synth_batch_noalloc keeps its Fourier-space intermediate here
// instead of in the plan's scratch: 4 Mi complex values (32 MiB), which every
// GPU the library runs on holds until the process ends.  Host code never
// uses the array's name as an address: the address on each device comes from
// hipGetSymbolAddress, taken once per device (with that device current) when
// the first plan on it is created.
constexpr size_t kNoAllocScratch = size_t(4) << 20;  // float2 entries
static __device__ __align__(256) float2 scratch[kNoAllocScratch];
constexpr int kMaxDevices = 64;
static float2* g_scratch[kMaxDevices];  // per device: scratch's device address
static std::mutex g_scratch_mu;

// With device dev current: resolve its copy of scratch (once).
static int resolve_scratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices)
    return fail(TSPHERE_ERR_UNSUPPORTED, "device %d beyond the library's %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[dev]) return TSPHERE_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  g_scratch[dev] = static_cast<float2*>(ptr);
  return TSPHERE_OK;
}

int tsphere_plan_create(const tsphere_sht_config* cfg, tsphere_plan* out) {
  if (!out) return fail(TSPHERE_ERR_INVALID, "out is NULL");
  *out = nullptr;
  if (int r = validate(cfg)) return r;
  const bool fits = fft_fits(cfg->nphi);
  if (cfg->fourier == TSPHERE_FOURIER_FFT && !fits)
    return fail(TSPHERE_ERR_INVALID,
                "fourier:'fft' requires power-of-two nphi in [8, 4096] (nphi=%d)", cfg->nphi);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return fail(TSPHERE_ERR_NODEVICE, "tsphere_plan_create: no GPU visible");
  }

  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return fail(TSPHERE_ERR_NODEVICE, "tsphere_plan_create: no current GPU");
  }
  if (int r = resolve_scratch(dev)) return r;

  auto* p = new (std::nothrow) tsphere_plan_t();
  if (!p) return fail(TSPHERE_ERR_INVALID, "out of host memory");
  p->cfg = *cfg;
  p->device = dev;
  p->fourier = (cfg->fourier == TSPHERE_FOURIER_DFT || !fits) ? TSPHERE_FOURIER_DFT
                                                               : TSPHERE_FOURIER_FFT;
  p->npair = (cfg->nlat + 1) / 2;
  p->nlm = nlm_calc(cfg->lmax, cfg->mmax);

  HostTables t;
  build_host_tables(*cfg, t);
  p->theta.resize(cfg->nlat);
  p->cos_theta = t.x;
  p->weights = t.wg;
  for (int i = 0; i < cfg->nlat; i++) p->theta[i] = std::acos(t.x[i]);

  hipError_t e = hipSuccess;
  const size_t nfm = (size_t)cfg->max_fields * (cfg->mmax + 1) * cfg->nlat;
  const size_t ntab = (size_t)(p->nlm + kAbPad);
  std::vector<float2> ct2(p->npair);
  for (int j = 0; j < p->npair; j++) ct2[j] = make_float2(t.ct[j], t.ct_lo[j]);
  if (e == hipSuccess) e = upload(&p->d_w, t.w.data(), (size_t)cfg->nlat);
  const std::vector<float> ones(cfg->nlat, 1.0f);
  if (e == hipSuccess) e = upload(&p->d_ones, ones.data(), (size_t)cfg->nlat);
  size_t bytes_tab = 0;
  {  // c_l of the rescaled recurrence (every form: split-BF16 runs it in f64)
    std::vector<float> cz(ntab);
    for (size_t i = 0; i < ntab; i++) cz[i] = (float)t.cz64[i];
    if (e == hipSuccess) e = upload(&p->d_cz, cz.data(), ntab);
    bytes_tab += sizeof(float) * ntab;
  }
  if (cfg->legendre == TSPHERE_LEG_FP32) {
    std::vector<float> al(ntab);
    for (size_t i = 0; i < ntab; i++) al[i] = (float)t.alz64[i];
    std::vector<SeedL> sz(t.seedz64.size());
    for (size_t i = 0; i < sz.size(); i++)
      sz[i] = SeedL{(float)t.seedz64[i].y0, (float)t.seedz64[i].y1, t.seedz64[i].l, 0};
    if (e == hipSuccess) e = upload(&p->d_alz, al.data(), ntab);
    if (e == hipSuccess) e = upload(&p->d_ct, ct2.data(), (size_t)p->npair);
    if (e == hipSuccess) e = upload(&p->d_seedz, sz.data(), sz.size());
    bytes_tab += sizeof(float) * ntab + sizeof(float2) * p->npair + sizeof(SeedL) * sz.size();
  } else {
    if (e == hipSuccess) e = upload(&p->d_ct64, t.x.data(), (size_t)p->npair);
    bytes_tab += sizeof(double) * p->npair;
  }
  if (cfg->legendre == TSPHERE_LEG_F64REC || cfg->legendre == TSPHERE_LEG_BF16X3) {
    // the rescaled f64 recurrence: the FP64-recurrence form and the split-BF16
    // 4x4x4 block kernels
    if (e == hipSuccess) e = upload(&p->d_alz64, t.alz64.data(), ntab);
    if (e == hipSuccess) e = upload(&p->d_seedz64, t.seedz64.data(), t.seedz64.size());
    bytes_tab += sizeof(double) * ntab + sizeof(SeedD) * t.seedz64.size();
  }
  if (cfg->legendre == TSPHERE_LEG_BF16X3) {
    if (e == hipSuccess)
      e = upload(&p->d_ab64, reinterpret_cast<const double2*>(t.ab64.data()), ntab);
    if (e == hipSuccess) e = upload(&p->d_seed64, t.seed64.data(), t.seed64.size());
    bytes_tab += sizeof(double2) * ntab + sizeof(SeedD) * t.seed64.size();
  }
  // FFT plans: the row FFT's pass tables follow the nphi (cos, sin) entries
  // (entries of the f32 table, so the row FFT rounds as with trig itself)
  std::vector<float2> dtrig(cfg->nphi);
  for (int k = 0; k < cfg->nphi; ++k) dtrig[k] = make_float2(t.trig[2 * k], t.trig[2 * k + 1]);
  if (p->fourier == TSPHERE_FOURIER_FFT) {
    const int N = cfg->nphi;
    int log2h = 0;
    while ((2 << log2h) < N) ++log2h;  // N = 2 H, H = 2^log2h
    const int ns0 = (log2h & 1) ? 2 : 4;
    for (int Ns = ns0; Ns < N / 2; Ns <<= 2)
      for (int k = 1; k <= 3; ++k)
        for (int jm = 0; jm < Ns; ++jm) {
          const float2 w = dtrig[(size_t)k * jm * (N / (4 * Ns))];
          dtrig.push_back(w);
        }
    if (dtrig.size() != (size_t)N + tsfft::pass_table_size(log2h)) e = hipErrorInvalidValue;
  }
  if (e == hipSuccess) e = upload(&p->d_trig, dtrig.data(), dtrig.size());
  if (e == hipSuccess) e = hipMalloc((void**)&p->d_fm, sizeof(float2) * nfm);
  if (e != hipSuccess) {
    plan_free(p);
    delete p;
    return hip_fail(e, "tsphere_plan_create");
  }
  p->bytes = bytes_tab + 2 * sizeof(float) * cfg->nlat + sizeof(float2) * dtrig.size() +
             sizeof(float2) * nfm;
  *out = p;
  return ok();
}

int tsphere_plan_destroy(tsphere_plan p) {
  if (p) {
    plan_free(p);
    delete p;
  }
  return ok();
}

int tsphere_plan_query(tsphere_plan p, tsphere_plan_info* info) {
  if (!p || !info) return fail(TSPHERE_ERR_INVALID, "plan_query: NULL argument");
  info->lmax = p->cfg.lmax;
  info->mmax = p->cfg.mmax;
  info->nlat = p->cfg.nlat;
  info->nphi = p->cfg.nphi;
  info->nlm = p->nlm;
  info->fourier = p->fourier;
  info->legendre = p->cfg.legendre;
  info->leg_lat = p->npair;
  info->max_fields = p->cfg.max_fields;
  info->scratch_bytes = p->bytes;
  return ok();
}

int tsphere_plan_grid(tsphere_plan p, double* theta, double* cos_theta, double* weights) {
  if (!p) return fail(TSPHERE_ERR_INVALID, "plan is NULL");
  const size_t n = sizeof(double) * p->cfg.nlat;
  if (theta) std::memcpy(theta, p->theta.data(), n);
  if (cos_theta) std::memcpy(cos_theta, p->cos_theta.data(), n);
  if (weights) std::memcpy(weights, p->weights.data(), n);
  return ok();
}

// ------------------------------------------------------------ transforms
static int check_batch(tsphere_plan p, int nfield, const void* a, const void* b) {
  if (!p) return fail(TSPHERE_ERR_INVALID, "plan is NULL");
  if (nfield < 1 || nfield > p->cfg.max_fields)
    return fail(TSPHERE_ERR_INVALID, "nfield=%d outside [1, max_fields=%d]", nfield,
                p->cfg.max_fields);
  if (!a || !b) return fail(TSPHERE_ERR_INVALID, "NULL buffer");
  return TSPHERE_OK;
}

static int check_stride(long long stride, long long per_field, int nfield, const char* what) {
  if (nfield > 1 && stride < per_field)
    return fail(TSPHERE_ERR_INVALID, "%s=%lld smaller than one field (%lld floats)", what, stride,
                per_field);
  if (stride < 0) return fail(TSPHERE_ERR_INVALID, "%s must be >= 0", what);
  return TSPHERE_OK;
}

static int check_list(tsphere_plan p, int nfield, const void* const* a, const void* const* b) {
  if (!p) return fail(TSPHERE_ERR_INVALID, "plan is NULL");
  if (nfield < 1 || nfield > p->cfg.max_fields || nfield > kMaxList)
    return fail(TSPHERE_ERR_INVALID, "nfield=%d outside [1, min(max_fields=%d, %d)]", nfield,
                p->cfg.max_fields, kMaxList);
  if (!a || !b) return fail(TSPHERE_ERR_INVALID, "NULL pointer list");
  for (int f = 0; f < nfield; f++)
    if (!a[f] || !b[f]) return fail(TSPHERE_ERR_INVALID, "NULL buffer for field %d", f);
  return TSPHERE_OK;
}

static Fields listed(const void* const* v, int n) {
  Fields F{};
  for (int f = 0; f < n; f++) F.p[f] = static_cast<const char*>(v[f]);
  F.stride = 0;
  F.per = 1;
  return F;
}

// split-BF16: the 4x4x4 block kernels where the tile resolves to them
static bool b4(const Plan* p, int stage, int nfield, const LegTile& t) {
  return p->cfg.legendre == TSPHERE_LEG_BF16X3 &&
         resolve_leg_tile(TSPHERE_LEG_BF16X3, stage, p->npair, nfield, t).impl == 4;
}

static hipError_t leg_synth(Plan* p, int nfield, const Fields& q, float2* fm, hipStream_t s) {
  if (p->cfg.legendre == TSPHERE_LEG_F64REC)
    return launch_leg_synth_f64r(p->tablesz64(), nfield, q, fm, p->tile[0], s);
  if (b4(p, 0, nfield, p->tile[0]))
    return launch_leg_synth_b4(p->tablesz64(), nfield, q, fm, p->tile[0], s);
  if (p->cfg.legendre == TSPHERE_LEG_BF16X3)
    return launch_leg_synth_mfma(p->tables64(), nfield, q, fm, p->tile[0], s);
  return launch_leg_synth(p->tables(), nfield, q, fm, p->tile[0], s);
}

static hipError_t leg_analys(Plan* p, int nfield, const float2* fm, const Fields& q,
                             hipStream_t s, bool unit_w = false) {
  if (p->cfg.legendre == TSPHERE_LEG_F64REC)
    return launch_leg_analys_f64r(p->tablesz64(unit_w), nfield, fm, q, p->tile[1], s);
  if (b4(p, 1, nfield, p->tile[1]))
    return launch_leg_analys_b4(p->tablesz64(unit_w), nfield, fm, q, p->tile[1], s);
  if (p->cfg.legendre == TSPHERE_LEG_BF16X3)
    return launch_leg_analys_mfma(p->tables64(unit_w), nfield, fm, q, p->tile[1], s);
  return launch_leg_analys(p->tables(unit_w), nfield, fm, q, p->tile[1], s);
}

static int run_synth(Plan* p, int nfield, const Fields& q, const Fields& spat, hipStream_t s) {
  TS_HIP(leg_synth(p, nfield, q, p->d_fm, s), "leg_synth");
  TS_HIP(launch_fourier_synth(p->fourier, p->cfg.mmax, p->cfg.nlat, p->cfg.nphi, p->d_trig,
                              nfield, p->d_fm, spat, s),
         "fourier_synth");
  return ok();
}

static int run_analys(Plan* p, int nfield, const Fields& spat, const Fields& q, hipStream_t s) {
  TS_HIP(launch_fourier_analys(p->fourier, p->cfg.mmax, p->cfg.nlat, p->cfg.nphi, p->d_trig,
                               nfield, spat, p->d_fm, s),
         "fourier_analys");
  TS_HIP(leg_analys(p, nfield, p->d_fm, q, s), "leg_analys");
  return ok();
}

int tsphere_leg_synth(tsphere_plan p, int nfield, const float* qlm, long long qlm_stride,
                      float* fm, tsphere_stream s) {
  if (int r = check_batch(p, nfield, qlm, fm)) return r;
  if (int r = check_stride(qlm_stride, 2 * p->nlm, nfield, "qlm_stride")) return r;
  TS_HIP(leg_synth(p, nfield, strided(qlm, 4 * qlm_stride), reinterpret_cast<float2*>(fm),
                   S(s)),
         "leg_synth");
  return ok();
}

int tsphere_leg_analys(tsphere_plan p, int nfield, const float* fm, float* qlm,
                       long long qlm_stride, tsphere_stream s) {
  if (int r = check_batch(p, nfield, fm, qlm)) return r;
  if (int r = check_stride(qlm_stride, 2 * p->nlm, nfield, "qlm_stride")) return r;
  TS_HIP(leg_analys(p, nfield, reinterpret_cast<const float2*>(fm),
                    strided(qlm, 4 * qlm_stride), S(s)),
         "leg_analys");
  return ok();
}

int tsphere_fourier_synth(tsphere_plan p, int nfield, const float* fm, float* spat,
                          long long spat_stride, tsphere_stream s) {
  if (int r = check_batch(p, nfield, fm, spat)) return r;
  if (int r = check_stride(spat_stride, (long long)p->cfg.nlat * p->cfg.nphi, nfield,
                           "spat_stride"))
    return r;
  TS_HIP(launch_fourier_synth(p->fourier, p->cfg.mmax, p->cfg.nlat, p->cfg.nphi, p->d_trig,
                              nfield, reinterpret_cast<const float2*>(fm),
                              strided(spat, 4 * spat_stride), S(s)),
         "fourier_synth");
  return ok();
}

int tsphere_fourier_analys(tsphere_plan p, int nfield, const float* spat, long long spat_stride,
                           float* fm, tsphere_stream s) {
  if (int r = check_batch(p, nfield, spat, fm)) return r;
  if (int r = check_stride(spat_stride, (long long)p->cfg.nlat * p->cfg.nphi, nfield,
                           "spat_stride"))
    return r;
  TS_HIP(launch_fourier_analys(p->fourier, p->cfg.mmax, p->cfg.nlat, p->cfg.nphi, p->d_trig,
                               nfield, strided(spat, 4 * spat_stride),
                               reinterpret_cast<float2*>(fm), S(s)),
         "fourier_analys");
  return ok();
}

int tsphere_synth_batch(tsphere_plan p, int nfield, const float* qlm, long long qlm_stride,
                        float* spat, long long spat_stride, tsphere_stream s) {
  if (int r = check_batch(p, nfield, qlm, spat)) return r;
  if (int r = check_stride(qlm_stride, 2 * p->nlm, nfield, "qlm_stride")) return r;
  if (int r = check_stride(spat_stride, (long long)p->cfg.nlat * p->cfg.nphi, nfield,
                           "spat_stride"))
    return r;
  return run_synth(p, nfield, strided(qlm, 4 * qlm_stride), strided(spat, 4 * spat_stride),
                   S(s));
}

int tsphere_analys_batch(tsphere_plan p, int nfield, const float* spat, long long spat_stride,
                         float* qlm, long long qlm_stride, tsphere_stream s) {
  if (int r = check_batch(p, nfield, spat, qlm)) return r;
  if (int r = check_stride(qlm_stride, 2 * p->nlm, nfield, "qlm_stride")) return r;
  if (int r = check_stride(spat_stride, (long long)p->cfg.nlat * p->cfg.nphi, nfield,
                           "spat_stride"))
    return r;
  return run_analys(p, nfield, strided(spat, 4 * spat_stride), strided(qlm, 4 * qlm_stride),
                    S(s));
}

// The synthesis of tsphere_synth_batch with its Fourier-space intermediate in
// the static scratch above (no allocation, no plan scratch).
int tsphere_synth_batch_noalloc(tsphere_plan p, int nfield, const float* qlm,
                                long long qlm_stride, float* spat, long long spat_stride,
                                tsphere_stream s) {
  if (!p) return fail(TSPHERE_ERR_INVALID, "plan is NULL");
  if (nfield < 1) return fail(TSPHERE_ERR_INVALID, "nfield=%d must be >= 1", nfield);
  if (!qlm || !spat) return fail(TSPHERE_ERR_INVALID, "NULL buffer");
  if (int r = check_stride(qlm_stride, 2 * p->nlm, nfield, "qlm_stride")) return r;
  if (int r = check_stride(spat_stride, (long long)p->cfg.nlat * p->cfg.nphi, nfield,
                           "spat_stride"))
    return r;
  const size_t need = (size_t)nfield * (p->cfg.mmax + 1) * p->cfg.nlat;
  if (need > kNoAllocScratch)
    return fail(TSPHERE_ERR_INVALID,
                "synth_batch_noalloc: %d field(s) at mmax=%d, nlat=%d need %zu complex values "
                "of Fourier scratch, more than its %zu (32 MiB)",
                nfield, p->cfg.mmax, p->cfg.nlat, need, kNoAllocScratch);
  int dev = -1;
  TS_HIP(hipGetDevice(&dev), "hipGetDevice");
  if (dev != p->device)
    return fail(TSPHERE_ERR_INVALID, "synth_batch_noalloc: the plan lives on device %d, "
                "device %d is current", p->device, dev);
  float2* fm = g_scratch[dev];  // resolved when the plan was created
  if (!fm) return fail(TSPHERE_ERR_INVALID, "synth_batch_noalloc: no scratch on device %d", dev);
  const Fields q = strided(qlm, 4 * qlm_stride), out = strided(spat, 4 * spat_stride);
  TS_HIP(leg_synth(p, nfield, q, fm, S(s)), "leg_synth");
  TS_HIP(launch_fourier_synth(p->fourier, p->cfg.mmax, p->cfg.nlat, p->cfg.nphi, p->d_trig,
                              nfield, fm, out, S(s)),
         "fourier_synth");
  return ok();
}

int tsphere_synth_multi(tsphere_plan p, int nfield, const float* const* qlm, float* const* spat,
                        tsphere_stream s) {
  if (int r = check_list(p, nfield, (const void* const*)qlm, (const void* const*)spat)) return r;
  return run_synth(p, nfield, listed((const void* const*)qlm, nfield),
                   listed((const void* const*)spat, nfield), S(s));
}

int tsphere_analys_multi(tsphere_plan p, int nfield, const float* const* spat, float* const* qlm,
                         tsphere_stream s) {
  if (int r = check_list(p, nfield, (const void* const*)spat, (const void* const*)qlm)) return r;
  return run_analys(p, nfield, listed((const void* const*)spat, nfield),
                    listed((const void* const*)qlm, nfield), S(s));
}

int tsphere_leg_synth_multi(tsphere_plan p, int nfield, const float* const* qlm, float* fm,
                            tsphere_stream s) {
  std::vector<const void*> fmv(nfield > 0 && nfield <= kMaxList ? nfield : 0, fm);
  if (int r = check_list(p, nfield, (const void* const*)qlm, fmv.data())) return r;
  TS_HIP(leg_synth(p, nfield, listed((const void* const*)qlm, nfield),
                   reinterpret_cast<float2*>(fm), S(s)),
         "leg_synth");
  return ok();
}

int tsphere_leg_analys_multi(tsphere_plan p, int nfield, const float* fm, float* const* qlm,
                             tsphere_stream s) {
  std::vector<const void*> fmv(nfield > 0 && nfield <= kMaxList ? nfield : 0, fm);
  if (int r = check_list(p, nfield, fmv.data(), (const void* const*)qlm)) return r;
  TS_HIP(leg_analys(p, nfield, reinterpret_cast<const float2*>(fm),
                    listed((const void* const*)qlm, nfield), S(s)),
         "leg_analys");
  return ok();
}

int tsphere_leg_imex_synth(tsphere_plan p, int nfield, const float* const* analysed,
                           float* const* state, const float* lam, const float* s1,
                           const float* s2, float* fm, tsphere_stream s) {
  if (!p) return fail(TSPHERE_ERR_INVALID, "plan is NULL");
  if (nfield < 1 || nfield > 4 || nfield > p->cfg.max_fields)
    return fail(TSPHERE_ERR_INVALID, "nfield=%d: need 1..4 and <= max_fields=%d", nfield,
                p->cfg.max_fields);
  if (p->cfg.legendre == TSPHERE_LEG_BF16X3 && p->npair > kB4MaxPairs)
    return fail(TSPHERE_ERR_INVALID, "the fused implicit step of the split-BF16 form covers up "
                                     "to %d latitude pairs", kB4MaxPairs);
  if (!analysed || !state || !lam || !s1 || !s2 || !fm)
    return fail(TSPHERE_ERR_INVALID, "NULL argument");
  ImexCoef ic{};
  for (int f = 0; f < nfield; f++) {
    if (!analysed[f] || !state[f]) return fail(TSPHERE_ERR_INVALID, "NULL buffer for field %d", f);
    ic.s1[f] = s1[f];
    ic.s2[f] = s2[f];
    ic.spc[f] = f;
  }
  const Fields q = listed((const void* const*)state, nfield);
  const Fields a = listed((const void* const*)analysed, nfield);
  float2* f2 = reinterpret_cast<float2*>(fm);
  if (p->cfg.legendre == TSPHERE_LEG_F64REC)
    TS_HIP(launch_leg_imex_synth_f64r(p->tablesz64(), nfield, q, a, lam, ic, f2, p->tile[0], S(s)),
           "leg_imex_synth");
  else if (p->cfg.legendre == TSPHERE_LEG_BF16X3)
    TS_HIP(launch_leg_imex_synth_b4(p->tablesz64(), nfield, q, a, lam, ic, f2, p->tile[0], S(s)),
           "leg_imex_synth");
  else
    TS_HIP(launch_leg_imex_synth(p->tables(), nfield, q, a, lam, ic, f2, p->tile[0], S(s)),
           "leg_imex_synth");
  return ok();
}

int tsphere_leg_spectral_imex(tsphere_plan p, int nfield, float* fm, float* const* state,
                              const float* lam, const float* s1, const float* s2,
                              const int* species, tsphere_stream s) {
  if (!p) return fail(TSPHERE_ERR_INVALID, "plan is NULL");
  if (nfield < 1 || nfield > 4 || nfield > p->cfg.max_fields)
    return fail(TSPHERE_ERR_INVALID, "nfield=%d: need 1..4 and <= max_fields=%d", nfield,
                p->cfg.max_fields);
  if (p->cfg.legendre == TSPHERE_LEG_BF16X3 && p->npair > kB4MaxPairs)
    return fail(TSPHERE_ERR_INVALID, "the split-BF16 spectral op covers up to %d latitude pairs",
                kB4MaxPairs);
  if (!fm || !state || !lam || !s1 || !s2 || !species)
    return fail(TSPHERE_ERR_INVALID, "NULL argument");
  ImexCoef ic{};
  bool seen[4] = {false, false, false, false};
  for (int f = 0; f < nfield; f++) {
    if (!state[f]) return fail(TSPHERE_ERR_INVALID, "NULL state buffer %d", f);
    if (species[f] < 0 || species[f] >= nfield || seen[species[f]])
      return fail(TSPHERE_ERR_INVALID, "species must be a permutation of 0..%d", nfield - 1);
    seen[species[f]] = true;
    ic.s1[f] = s1[f];
    ic.s2[f] = s2[f];
    ic.spc[f] = species[f];
  }
  if (p->npair > 64 * 8 * 2)
    return fail(TSPHERE_ERR_INVALID, "the spectral op covers up to 1024 latitude pairs");
  const Fields st = listed((const void* const*)state, nfield);
  float2* f2 = reinterpret_cast<float2*>(fm);
  if (p->cfg.legendre == TSPHERE_LEG_F64REC)
    TS_HIP(launch_leg_spectral_f64r(p->tablesz64(), nfield, f2, st, lam, ic, S(s)), "leg_spectral");
  else if (p->cfg.legendre == TSPHERE_LEG_BF16X3)
    TS_HIP(launch_leg_spectral_b4(p->tablesz64(), nfield, f2, st, lam, ic, S(s)), "leg_spectral");
  else
    TS_HIP(launch_leg_spectral(p->tables(), nfield, f2, st, lam, ic, S(s)), "leg_spectral");
  return ok();
}

int tsphere_plan_fourier_table(tsphere_plan p, const void** trig) {
  if (!p || !trig) return fail(TSPHERE_ERR_INVALID, "NULL argument");
  *trig = p->d_trig;
  return ok();
}

static int check_fields(tsphere_plan p, int nptr, int per, const void* const* a,
                        const void* const* b) {
  if (!p) return fail(TSPHERE_ERR_INVALID, "plan is NULL");
  if (nptr < 1 || nptr > kMaxList || per < 1)
    return fail(TSPHERE_ERR_INVALID, "nptr=%d per=%d: need 1 <= nptr <= %d, per >= 1", nptr, per,
                kMaxList);
  if ((long long)nptr * per > p->cfg.max_fields)
    return fail(TSPHERE_ERR_INVALID, "nptr*per=%lld exceeds max_fields=%d",
                (long long)nptr * per, p->cfg.max_fields);
  if (!a || !b) return fail(TSPHERE_ERR_INVALID, "NULL pointer list");
  for (int f = 0; f < nptr; f++)
    if (!a[f] || !b[f]) return fail(TSPHERE_ERR_INVALID, "NULL buffer for pointer %d", f);
  return TSPHERE_OK;
}

static Fields grouped(const void* const* v, int nptr, int per, long long stride_floats) {
  Fields F{};
  for (int f = 0; f < nptr; f++) F.p[f] = static_cast<const char*>(v[f]);
  F.stride = 4 * stride_floats;
  F.per = per;
  return F;
}

int tsphere_synth_fields(tsphere_plan p, int nptr, int per, const float* const* qlm,
                         long long qlm_stride, float* const* spat, long long spat_stride,
                         tsphere_stream s) {
  if (int r = check_fields(p, nptr, per, (const void* const*)qlm, (const void* const*)spat))
    return r;
  if (int r = check_stride(qlm_stride, 2 * p->nlm, per, "qlm_stride")) return r;
  if (int r = check_stride(spat_stride, (long long)p->cfg.nlat * p->cfg.nphi, per, "spat_stride"))
    return r;
  return run_synth(p, nptr * per, grouped((const void* const*)qlm, nptr, per, qlm_stride),
                   grouped((const void* const*)spat, nptr, per, spat_stride), S(s));
}

int tsphere_analys_fields(tsphere_plan p, int nptr, int per, const float* const* spat,
                          long long spat_stride, float* const* qlm, long long qlm_stride,
                          tsphere_stream s) {
  if (int r = check_fields(p, nptr, per, (const void* const*)spat, (const void* const*)qlm))
    return r;
  if (int r = check_stride(qlm_stride, 2 * p->nlm, per, "qlm_stride")) return r;
  if (int r = check_stride(spat_stride, (long long)p->cfg.nlat * p->cfg.nphi, per, "spat_stride"))
    return r;
  return run_analys(p, nptr * per, grouped((const void* const*)spat, nptr, per, spat_stride),
                    grouped((const void* const*)qlm, nptr, per, qlm_stride), S(s));
}

int tsphere_leg_synth_fields(tsphere_plan p, int nptr, int per, const float* const* qlm,
                             long long qlm_stride, float* fm, tsphere_stream s) {
  std::vector<const void*> fmv(nptr > 0 && nptr <= kMaxList ? nptr : 0, fm);
  if (int r = check_fields(p, nptr, per, (const void* const*)qlm, fmv.data())) return r;
  if (int r = check_stride(qlm_stride, 2 * p->nlm, per, "qlm_stride")) return r;
  TS_HIP(leg_synth(p, nptr * per, grouped((const void* const*)qlm, nptr, per, qlm_stride),
                   reinterpret_cast<float2*>(fm), S(s)),
         "leg_synth");
  return ok();
}

int tsphere_leg_analys_fields(tsphere_plan p, int nptr, int per, const float* fm,
                              float* const* qlm, long long qlm_stride, tsphere_stream s) {
  std::vector<const void*> fmv(nptr > 0 && nptr <= kMaxList ? nptr : 0, fm);
  if (int r = check_fields(p, nptr, per, fmv.data(), (const void* const*)qlm)) return r;
  if (int r = check_stride(qlm_stride, 2 * p->nlm, per, "qlm_stride")) return r;
  TS_HIP(leg_analys(p, nptr * per, reinterpret_cast<const float2*>(fm),
                    grouped((const void* const*)qlm, nptr, per, qlm_stride), S(s)),
         "leg_analys");
  return ok();
}

int tsphere_leg_imex_synth_fields(tsphere_plan p, int nptr, int per,
                                  const float* const* analysed, float* const* state,
                                  long long qlm_stride, const float* lam, const float* coef,
                                  float* fm, tsphere_stream s) {
  if (int r = check_fields(p, nptr, per, (const void* const*)analysed, (const void* const*)state))
    return r;
  if (int r = check_stride(qlm_stride, 2 * p->nlm, per, "qlm_stride")) return r;
  if (!lam || !coef || !fm) return fail(TSPHERE_ERR_INVALID, "NULL argument");
  if (p->cfg.legendre == TSPHERE_LEG_BF16X3 && p->npair > kB4MaxPairs)
    return fail(TSPHERE_ERR_INVALID, "the fused implicit step of the split-BF16 form covers up "
                                     "to %d latitude pairs", kB4MaxPairs);
  const int nf = nptr * per;
  const Fields q = grouped((const void* const*)state, nptr, per, qlm_stride);
  const Fields a = grouped((const void* const*)analysed, nptr, per, qlm_stride);
  ImexCoef ic{};
  ic.dev = coef;
  float2* f2 = reinterpret_cast<float2*>(fm);
  if (p->cfg.legendre == TSPHERE_LEG_F64REC)
    TS_HIP(launch_leg_imex_synth_f64r(p->tablesz64(), nf, q, a, lam, ic, f2, p->tile[0], S(s)),
           "leg_imex_synth");
  else if (p->cfg.legendre == TSPHERE_LEG_BF16X3)
    TS_HIP(launch_leg_imex_synth_b4(p->tablesz64(), nf, q, a, lam, ic, f2, p->tile[0], S(s)),
           "leg_imex_synth");
  else
    TS_HIP(launch_leg_imex_synth(p->tables(), nf, q, a, lam, ic, f2, p->tile[0], S(s)),
           "leg_imex_synth");
  return ok();
}

int tsphere_leg_analys_imex_fields(tsphere_plan p, int nptr, int per, const float* fm,
                                   float* const* analysed, float* const* state,
                                   long long qlm_stride, const float* lam, const float* coef,
                                   tsphere_stream s) {
  if (int r = check_fields(p, nptr, per, (const void* const*)analysed, (const void* const*)state))
    return r;
  if (int r = check_stride(qlm_stride, 2 * p->nlm, per, "qlm_stride")) return r;
  if (!fm || !lam || !coef) return fail(TSPHERE_ERR_INVALID, "NULL argument");
  const int nf = nptr * per;
  const Fields q = grouped((const void* const*)state, nptr, per, qlm_stride);
  const Fields a = grouped((const void* const*)analysed, nptr, per, qlm_stride);
  const float2* f2 = reinterpret_cast<const float2*>(fm);
  if (p->cfg.legendre == TSPHERE_LEG_F64REC) {
    TS_HIP(launch_leg_analys_imex_f64r(p->tablesz64(), nf, f2, a, q, lam, coef, p->tile[1], S(s)),
           "leg_analys_imex");
  } else if (p->cfg.legendre == TSPHERE_LEG_FP32) {
    TS_HIP(launch_leg_analys_imex(p->tables(), nf, f2, a, q, lam, coef, p->tile[1], S(s)),
           "leg_analys_imex");
  } else {
    TS_HIP(leg_analys(p, nf, f2, a, S(s)), "leg_analys");
    TS_HIP(launch_imex_fields(p->nlm, nf, a, q, lam, coef, S(s)), "imex_fields");
  }
  return ok();
}

// ------------------------------------------------------- transposes
// synth^T: G = 2^{[m>=1]} * FourierAnalysis(g), q = sum_i P_lm(theta_i) G_m(theta_i)
// analys^T: F = LegendreSynth(q), g_ij = w_i * FourierSynth(F with m>=1 halved)
int tsphere_synth_adjoint(tsphere_plan p, int nptr, int per, const float* const* gbar,
                          long long g_stride, float* const* qbar, long long q_stride,
                          tsphere_stream s) {
  if (int r = check_fields(p, nptr, per, (const void* const*)gbar, (const void* const*)qbar))
    return r;
  if (int r = check_stride(q_stride, 2 * p->nlm, per, "q_stride")) return r;
  if (int r = check_stride(g_stride, (long long)p->cfg.nlat * p->cfg.nphi, per, "g_stride"))
    return r;
  const int nf = nptr * per;
  FourierScale sc;
  sc.mscale = 2.f;
  TS_HIP(launch_fourier_analys(p->fourier, p->cfg.mmax, p->cfg.nlat, p->cfg.nphi, p->d_trig, nf,
                               grouped((const void* const*)gbar, nptr, per, g_stride), p->d_fm,
                               S(s), sc),
         "fourier_analys (transpose)");
  TS_HIP(leg_analys(p, nf, p->d_fm, grouped((const void* const*)qbar, nptr, per, q_stride), S(s),
                    true),
         "leg_analys (transpose)");
  return ok();
}

int tsphere_analys_adjoint(tsphere_plan p, int nptr, int per, const float* const* qbar,
                           long long q_stride, float* const* gbar, long long g_stride,
                           tsphere_stream s) {
  if (int r = check_fields(p, nptr, per, (const void* const*)qbar, (const void* const*)gbar))
    return r;
  if (int r = check_stride(q_stride, 2 * p->nlm, per, "q_stride")) return r;
  if (int r = check_stride(g_stride, (long long)p->cfg.nlat * p->cfg.nphi, per, "g_stride"))
    return r;
  const int nf = nptr * per;
  TS_HIP(leg_synth(p, nf, grouped((const void* const*)qbar, nptr, per, q_stride), p->d_fm, S(s)),
         "leg_synth (transpose)");
  FourierScale sc;
  sc.mscale = 0.5f;
  sc.rowscale = p->d_w;
  TS_HIP(launch_fourier_synth(p->fourier, p->cfg.mmax, p->cfg.nlat, p->cfg.nphi, p->d_trig, nf,
                              p->d_fm, grouped((const void* const*)gbar, nptr, per, g_stride),
                              S(s), sc),
         "fourier_synth (transpose)");
  return ok();
}

int tsphere_accumulate(float* dst, const float* src, long long n, tsphere_stream s) {
  if (n < 0) return fail(TSPHERE_ERR_INVALID, "n must be >= 0");
  if (n > 0 && (!dst || !src)) return fail(TSPHERE_ERR_INVALID, "NULL buffer");
  TS_HIP(launch_accumulate(dst, src, n, S(s)), "accumulate");
  return ok();
}

// ------------------------------------------------------------ tile tuning
static LegTile to_tile(const tsphere_leg_tile& t) {
  return LegTile{t.nf, t.nwaves, t.k, t.ctc, t.impl, t.rows};
}
static tsphere_leg_tile from_tile(const LegTile& t) {
  tsphere_leg_tile o{};
  o.nf = t.nf;
  o.nwaves = t.nwaves;
  o.k = t.k;
  o.ctc = t.ctc;
  o.impl = t.impl;
  o.rows = t.rows;
  return o;
}

int tsphere_plan_set_tile(tsphere_plan p, int stage, const tsphere_leg_tile* tile) {
  if (!p) return fail(TSPHERE_ERR_INVALID, "plan is NULL");
  if (stage != 0 && stage != 1) return fail(TSPHERE_ERR_INVALID, "stage must be 0 or 1");
  const LegTile t = tile ? to_tile(*tile) : LegTile{};
  const bool bf = p->cfg.legendre == TSPHERE_LEG_BF16X3;
  if (t.nf < 0 || t.nwaves < 0 || t.k < 0 || t.ctc < 0 || t.impl < 0 || t.impl > 5 ||
      t.rows < 0 || t.rows > 2 || (bf && t.impl != 0 && t.impl != 4) || (!bf && t.impl == 4))
    return fail(TSPHERE_ERR_INVALID, "tile fields must be >= 0 (impl 0..3 or 5, split-BF16 0 or 4; "
                                     "rows 0, 1 or 2)");
  // validate against a one-field and a max-fields launch
  for (int nf : {1, p->cfg.max_fields}) {
    const LegTile r = resolve_leg_tile(p->cfg.legendre, stage, p->npair, nf, t);
    if (p->cfg.legendre == TSPHERE_LEG_BF16X3 && r.impl == 4) {
      if (!(r.nf == 1 || r.nf == 2 || r.nf == 4) || r.k != 1 || r.nwaves < 1 || r.nwaves > 8 ||
          64LL * r.nwaves < p->npair)
        return fail(TSPHERE_ERR_INVALID,
                    "split-BF16 block tile nf=%d nwaves=%d k=%d: need nf in {1,2,4}, k 1, "
                    "64*nwaves >= %d latitude pairs (<= 8 waves)", r.nf, r.nwaves, r.k, p->npair);
    } else if (p->cfg.legendre == TSPHERE_LEG_BF16X3) {
      if (!(r.ctc == 1 || r.ctc == 2 || (stage == 0 && r.ctc == 4)))
        return fail(TSPHERE_ERR_INVALID, "MFMA tile: ctc must be 1, 2 (or 4 for synthesis)");
    } else if (r.impl == 5) {
      const bool rows_ok = r.rows == 1 || (r.rows == 2 && r.nf == 8 && stage == 0);
      if (!(r.nf == 8 || r.nf == 16) || r.k != 1 || !rows_ok || r.nwaves < 1 ||
          r.nwaves > 8 || 64LL * r.nwaves < p->npair)
        return fail(TSPHERE_ERR_INVALID,
                    "per-parity product tile nf=%d nwaves=%d k=%d rows=%d: need nf 8 or 16, k 1, "
                    "rows 1 (2 for the 8-field synthesis), 64*nwaves >= %d latitude pairs "
                    "(<= 8 waves)", r.nf, r.nwaves, r.k, r.rows, p->npair);
    } else {
      const bool kok = r.impl >= 2 ? (r.k == 1 || r.k == 2 || r.k == 4)
                                   : (r.k == 1 || r.k == 2 || r.k == 4 || r.k == 8);
      const bool nfok = r.nf == 1 || r.nf == 2 || r.nf == 4 || (r.nf == 8 && r.impl == 3 && r.k <= 2);
      if (!nfok || r.nwaves < 1 || r.nwaves > (r.impl == 3 ? 8 : 4) || !kok ||
          64LL * r.nwaves * r.k < p->npair)
        return fail(TSPHERE_ERR_INVALID,
                    "FP32 tile impl=%d nf=%d nwaves=%d k=%d: need nf in {1,2,4} (or 8 for impl 3 "
                    "with k <= 2), nwaves in 1..4 (1..8 for impl 3), k in {1,2,4,8} (VALU) or "
                    "{1,2,4} (matrix cores), 64*nwaves*k >= %d latitude pairs",
                    r.impl, r.nf, r.nwaves, r.k, p->npair);
    }
  }
  p->tile[stage] = t;
  return ok();
}

int tsphere_plan_get_tile(tsphere_plan p, int stage, int nfield, tsphere_leg_tile* out) {
  if (!p || !out) return fail(TSPHERE_ERR_INVALID, "plan_get_tile: NULL argument");
  if (stage != 0 && stage != 1) return fail(TSPHERE_ERR_INVALID, "stage must be 0 or 1");
  if (nfield < 1) return fail(TSPHERE_ERR_INVALID, "nfield must be >= 1");
  *out = from_tile(resolve_leg_tile(p->cfg.legendre, stage, p->npair, nfield, p->tile[stage]));
  return ok();
}

// Candidate tiles of one stage: FP32 -- every fields-per-block choice (eight
// only for the block-matrix kernels, with at most two chains per lane) with
// every wave count at the smallest k that covers the pairs, and the
// per-parity products with 8 and 16 fields; MFMA -- the column tiles per
// block.
static std::vector<LegTile> tile_candidates(const Plan* p, int stage, int nfield) {
  std::vector<LegTile> c;
  if (p->cfg.legendre == TSPHERE_LEG_BF16X3) {
    const int ct = (2 * nfield + 15) / 16;
    for (int ctc : {1, 2, 4})
      if ((stage == 0 || ctc <= 2) && (ctc == 1 || ctc <= ct)) c.push_back(LegTile{0, 0, 0, ctc, 0, 0});
    if (p->npair <= kB4MaxPairs)  // 4x4x4 blocks: fields per workgroup x rows
      for (int nf : {1, 2, 4})
        if (nf == 1 || nf <= nfield || (nf == 4 && nfield == 3))
          for (int rows : {1, 2}) c.push_back(LegTile{nf, (p->npair + 63) / 64, 1, 0, 4, rows});
    return c;
  }
  for (int impl = 1; impl <= 3; impl++)
    for (int nf : {1, 2, 4, 8}) {
      if (nf > 1 && nf > nfield && !(nf == 4 && nfield == 3)) continue;
      if (nf == 8 && impl != 3) continue;  // eight fields: block-matrix kernels only
      for (int nw = 1; nw <= (impl == 3 ? 8 : 4); nw += (nw >= 4 ? 4 : 1)) {
        const int kmax = nf == 8 ? 2 : impl >= 2 ? 4 : 8;
        int k = 1;
        while (k <= kmax && 64 * nw * k < p->npair) k *= 2;
        if (k > kmax) continue;
        if (nw > 1 && 64 * (nw - 1) * k >= p->npair) continue;  // a wave with no pairs
        c.push_back(LegTile{nf, nw, k, 0, impl, 2});
        if (impl != 2) c.push_back(LegTile{nf, nw, k, 0, impl, 1});
      }
    }
  // per-parity matrix products: 8 or 16 fields per workgroup, a wave per 64 pairs
  if (p->npair <= 64 * 8)
    for (int nf : {8, 16})
      if (nf <= nfield) {
        c.push_back(LegTile{nf, (p->npair + 63) / 64, 1, 0, 5, 1});
        // two consecutive rows per workgroup: the 8-field synthesis
        if (nf == 8 && stage == 0) c.push_back(LegTile{8, (p->npair + 63) / 64, 1, 0, 5, 2});
      }
  return c;
}

int tsphere_plan_tune(tsphere_plan p, int nfield, int reps, tsphere_stream stream,
                      tsphere_leg_tile* best, float* best_ms) {
  if (!p) return fail(TSPHERE_ERR_INVALID, "plan is NULL");
  if (nfield < 1 || nfield > p->cfg.max_fields)
    return fail(TSPHERE_ERR_INVALID, "nfield=%d outside [1, max_fields=%d]", nfield,
                p->cfg.max_fields);
  if (reps < 1) reps = 1;
  hipStream_t s = S(stream);
  const long long qn = 2 * p->nlm;
  float* q = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc((void**)&q, sizeof(float) * qn * nfield);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_fill(q, 0.01f, qn * nfield, s);
  if (e == hipSuccess) e = launch_fill(reinterpret_cast<float*>(p->d_fm), 0.01f,
                                       2LL * nfield * (p->cfg.mmax + 1) * p->cfg.nlat, s);
  const Fields F = strided(q, 4 * qn);
  int rc = TSPHERE_OK;
  for (int stage = 0; stage < 2 && e == hipSuccess; stage++) {
    LegTile chosen{};
    float chosen_ms = 0.f;
    for (const LegTile& t : tile_candidates(p, stage, nfield)) {
      auto launch = [&]() -> hipError_t {
        if (p->cfg.legendre == TSPHERE_LEG_BF16X3 && t.impl == 4)
          return stage == 0 ? launch_leg_synth_b4(p->tablesz64(), nfield, F, p->d_fm, t, s)
                            : launch_leg_analys_b4(p->tablesz64(), nfield, p->d_fm, F, t, s);
        if (p->cfg.legendre == TSPHERE_LEG_BF16X3)
          return stage == 0 ? launch_leg_synth_mfma(p->tables64(), nfield, F, p->d_fm, t, s)
                            : launch_leg_analys_mfma(p->tables64(), nfield, p->d_fm, F, t, s);
        if (p->cfg.legendre == TSPHERE_LEG_F64REC)
          return stage == 0 ? launch_leg_synth_f64r(p->tablesz64(), nfield, F, p->d_fm, t, s)
                            : launch_leg_analys_f64r(p->tablesz64(), nfield, p->d_fm, F, t, s);
        return stage == 0 ? launch_leg_synth(p->tables(), nfield, F, p->d_fm, t, s)
                          : launch_leg_analys(p->tables(), nfield, p->d_fm, F, t, s);
      };
      if ((e = launch()) != hipSuccess) break;  // warm-up (and code load)
      if ((e = hipEventRecord(e0, s)) != hipSuccess) break;
      for (int r = 0; r < reps && e == hipSuccess; r++) e = launch();
      if (e == hipSuccess) e = hipEventRecord(e1, s);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) break;
      ms /= reps;
      if (chosen_ms == 0.f || ms < chosen_ms) {
        chosen = t;
        chosen_ms = ms;
      }
    }
    if (e != hipSuccess) break;
    p->tile[stage] = chosen;
    if (best) best[stage] = from_tile(resolve_leg_tile(p->cfg.legendre, stage, p->npair, nfield,
                                                       chosen));
    if (best_ms) best_ms[stage] = chosen_ms;
  }
  if (e != hipSuccess) rc = hip_fail(e, "tsphere_plan_tune");
  // cleanup: the first error (if any) is already recorded
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (q) {
    (void)hipStreamSynchronize(s);
    (void)hipFree(q);
  }
  return rc == TSPHERE_OK ? ok() : rc;
}

int tsphere_synth(tsphere_plan p, const float* qlm, float* spat, tsphere_stream s) {
  return tsphere_synth_batch(p, 1, qlm, 0, spat, 0, s);
}

int tsphere_analys(tsphere_plan p, const float* spat, float* qlm, tsphere_stream s) {
  return tsphere_analys_batch(p, 1, spat, 0, qlm, 0, s);
}

// ------------------------------------------------------ hipRTC modules
struct tsphere_module_t {
  hipModule_t mod = nullptr;
};
struct tsphere_kernel_t {
  hipFunction_t fn = nullptr;
};

static int rtc_compile(const char* source, const char* options, std::vector<char>& code) {
  if (!source) return fail(TSPHERE_ERR_INVALID, "source is NULL");
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, source, "tsphere_generated.hip", 0, nullptr, nullptr) !=
      HIPRTC_SUCCESS)
    return fail(TSPHERE_ERR_COMPILE, "hiprtcCreateProgram failed");
  // Element-wise kernels are HBM-bound, so contraction buys nothing; keeping
  // every f32 operation separately rounded makes them reproducible against
  // an unfused fp32 evaluation of the same expression.
  std::vector<std::string> opts = {"--offload-arch=gfx950",
                                   "-O3",
                                   "-ffp-contract=off",
                                   "-fhip-fp32-correctly-rounded-divide-sqrt",
                                   "-Xclang",
                                   "-target-feature",
                                   "-Xclang",
                                   "-packed-fp32-ops"};
  if (options) {
    std::string o(options);
    size_t pos = 0;
    while (pos <= o.size()) {
      size_t nl = o.find('\n', pos);
      if (nl == std::string::npos) nl = o.size();
      if (nl > pos) opts.push_back(o.substr(pos, nl - pos));
      pos = nl + 1;
    }
  }
  std::vector<const char*> argv;
  for (auto& s : opts) argv.push_back(s.c_str());
  hiprtcResult r = hiprtcCompileProgram(prog, (int)argv.size(), argv.data());
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    return fail(TSPHERE_ERR_COMPILE, "hipRTC compile failed (%s):\n%s", hiprtcGetErrorString(r),
                log.c_str());
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  code.resize(n);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  return TSPHERE_OK;
}

int tsphere_rtc_check(const char* source, const char* options) {
  std::vector<char> code;
  if (int r = rtc_compile(source, options, code)) return r;
  return ok();
}

int tsphere_module_compile(const char* source, const char* options, tsphere_module* out) {
  if (!out) return fail(TSPHERE_ERR_INVALID, "out is NULL");
  *out = nullptr;
  std::vector<char> code;
  if (int r = rtc_compile(source, options, code)) return r;
  hipModule_t mod;
  TS_HIP(hipModuleLoadData(&mod, code.data()), "hipModuleLoadData");
  *out = new tsphere_module_t{mod};
  return ok();
}

int tsphere_module_destroy(tsphere_module m) {
  if (m) {
    hipError_t e = m->mod ? hipModuleUnload(m->mod) : hipSuccess;
    delete m;
    if (e != hipSuccess) return hip_fail(e, "hipModuleUnload");
  }
  return ok();
}

int tsphere_module_kernel(tsphere_module m, const char* name, tsphere_kernel* out) {
  if (!m || !name || !out) return fail(TSPHERE_ERR_INVALID, "module_kernel: NULL argument");
  hipFunction_t fn;
  hipError_t e = hipModuleGetFunction(&fn, m->mod, name);
  if (e != hipSuccess) return fail(TSPHERE_ERR_INVALID, "no kernel named '%s' in module", name);
  *out = new tsphere_kernel_t{fn};
  return ok();
}

int tsphere_kernel_destroy(tsphere_kernel k) {
  delete k;
  return ok();
}

int tsphere_kernel_launch(tsphere_kernel k, long long count, int nptr, void* const* ptrs,
                          int nscalar, const float* scalars, tsphere_stream s) {
  if (!k) return fail(TSPHERE_ERR_INVALID, "kernel is NULL");
  if (count < 0 || nptr < 1 || nptr > 64 || nscalar < 0 || nscalar > 64)
    return fail(TSPHERE_ERR_INVALID, "kernel_launch: bad count/nptr/nscalar");
  if (!ptrs || (nscalar && !scalars)) return fail(TSPHERE_ERR_INVALID, "kernel_launch: NULL array");
  if (count == 0) return ok();
  // Pack (long long count, pointers..., floats...) with natural alignment.
  alignas(8) unsigned char args[8 + 8 * 64 + 4 * 64];
  size_t off = 0;
  std::memcpy(args + off, &count, 8);
  off += 8;
  for (int i = 0; i < nptr; i++) {
    std::memcpy(args + off, &ptrs[i], 8);
    off += 8;
  }
  for (int i = 0; i < nscalar; i++) {
    std::memcpy(args + off, &scalars[i], 4);
    off += 4;
  }
  size_t size = off;
  void* config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, args, HIP_LAUNCH_PARAM_BUFFER_SIZE, &size,
                    HIP_LAUNCH_PARAM_END};
  // Each thread handles 4 consecutive elements (float4 traffic).
  const long long threads = (count + 3) / 4;
  const unsigned blocks = (unsigned)((threads + 255) / 256);
  TS_HIP(hipModuleLaunchKernel(k->fn, blocks, 1, 1, 256, 1, 1, 0, S(s), nullptr, config),
         "hipModuleLaunchKernel");
  return ok();
}

int tsphere_kernel_launch_grid(tsphere_kernel k, unsigned gx, unsigned gy, unsigned block,
                               int nptr, void* const* ptrs, int nscalar, const float* scalars,
                               tsphere_stream s) {
  if (!k) return fail(TSPHERE_ERR_INVALID, "kernel is NULL");
  if (nptr < 0 || nptr > 64 || nscalar < 0 || nscalar > 64 || block < 1 || block > 1024 ||
      gx < 1 || gy < 1)
    return fail(TSPHERE_ERR_INVALID, "kernel_launch_grid: bad grid/block/nptr/nscalar");
  if ((nptr && !ptrs) || (nscalar && !scalars))
    return fail(TSPHERE_ERR_INVALID, "kernel_launch_grid: NULL array");
  // Pack (pointers..., floats...) with natural alignment.
  alignas(8) unsigned char args[8 * 64 + 4 * 64];
  size_t off = 0;
  for (int i = 0; i < nptr; i++) {
    std::memcpy(args + off, &ptrs[i], 8);
    off += 8;
  }
  for (int i = 0; i < nscalar; i++) {
    std::memcpy(args + off, &scalars[i], 4);
    off += 4;
  }
  size_t size = off;
  void* config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, args, HIP_LAUNCH_PARAM_BUFFER_SIZE, &size,
                    HIP_LAUNCH_PARAM_END};
  TS_HIP(hipModuleLaunchKernel(k->fn, gx, gy, 1, block, 1, 1, 0, S(s), nullptr, config),
         "hipModuleLaunchKernel");
  return ok();
}

// ---------------------------------------------------------------- graphs
struct tsphere_graph_t {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int tsphere_graph_begin(tsphere_stream s) {
  if (!s) return fail(TSPHERE_ERR_INVALID, "graph capture needs a created stream, not NULL");
  TS_HIP(hipStreamBeginCapture(S(s), hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
  return ok();
}

int tsphere_graph_end(tsphere_stream s, tsphere_graph* out) {
  if (!s || !out) return fail(TSPHERE_ERR_INVALID, "graph_end: NULL argument");
  *out = nullptr;
  hipGraph_t g = nullptr;
  TS_HIP(hipStreamEndCapture(S(s), &g), "hipStreamEndCapture");
  hipGraphExec_t x = nullptr;
  hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return hip_fail(e, "hipGraphInstantiate");
  }
  // move the graph to the device now (enqueued on s), so that its first
  // launch costs what every later one does
  e = hipGraphUpload(x, S(s));
  if (e != hipSuccess) {
    (void)hipGraphExecDestroy(x);
    (void)hipGraphDestroy(g);
    return hip_fail(e, "hipGraphUpload");
  }
  *out = new tsphere_graph_t{g, x};
  return ok();
}

int tsphere_graph_launch(tsphere_graph g, tsphere_stream s) {
  if (!g) return fail(TSPHERE_ERR_INVALID, "graph is NULL");
  TS_HIP(hipGraphLaunch(g->exec, S(s)), "hipGraphLaunch");
  return ok();
}

int tsphere_graph_destroy(tsphere_graph g) {
  if (g) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
  }
  return ok();
}

}  // extern "C"
