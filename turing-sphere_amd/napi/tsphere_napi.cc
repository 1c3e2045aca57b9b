// N-API binding of libtsphere (include/tsphere.h) for the reference's
// TypeScript host: the calls its ShtPlan (src/sht/sht.ts) and ModelPlan /
// GpuModel (src/mgpu/plan.ts, model.ts) make on a WebGPU device, made here on
// the MI355X through the C ABI.  Handles and device pointers travel as
// BigInt.  Errors throw: `ShtConfigError` for an invalid transform
// configuration (layout.ts:34-43 messages), `TsphereError` otherwise.
//
// Built by turing-sphere_amd/build.py when node's headers are present (no
// node-gyp needed): g++ -shared -fPIC -I<node include> against libtsphere.so,
// with an rpath to the package directory.
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "tsphere.h"

namespace {

napi_value throw_error(napi_env env, int rc, const char* what) {
  const char* name = rc == TSPHERE_ERR_INVALID ? "ShtConfigError" : "TsphereError";
  std::string msg = std::string(what) + ": " + tsphere_last_error();
  napi_value m, err, nm, code;
  napi_create_string_utf8(env, msg.c_str(), NAPI_AUTO_LENGTH, &m);
  napi_create_error(env, nullptr, m, &err);
  napi_create_string_utf8(env, name, NAPI_AUTO_LENGTH, &nm);
  napi_set_named_property(env, err, "name", nm);
  napi_create_int32(env, rc, &code);
  napi_set_named_property(env, err, "code", code);
  napi_throw(env, err);
  return nullptr;
}

napi_value type_error(napi_env env, const char* msg) {
  napi_throw_type_error(env, nullptr, msg);
  return nullptr;
}

struct Args {
  size_t argc = 8;
  napi_value v[8];
  explicit Args(napi_env env, napi_callback_info info) {
    napi_get_cb_info(env, info, &argc, v, nullptr, nullptr);
  }
};

bool get_ptr(napi_env env, napi_value v, void** out) {  // BigInt (or 0 / undefined = NULL)
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_undefined || t == napi_null) {
    *out = nullptr;
    return true;
  }
  if (t != napi_bigint) return false;
  uint64_t u = 0;
  bool lossless = false;
  napi_get_value_bigint_uint64(env, v, &u, &lossless);
  *out = reinterpret_cast<void*>(static_cast<uintptr_t>(u));
  return lossless;
}

napi_value make_ptr(napi_env env, const void* p) {
  napi_value out;
  napi_create_bigint_uint64(env, static_cast<uint64_t>(reinterpret_cast<uintptr_t>(p)), &out);
  return out;
}

double get_num(napi_env env, napi_value v, double dflt = 0) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t != napi_number) return dflt;
  double d = dflt;
  napi_get_value_double(env, v, &d);
  return d;
}

int get_int_prop(napi_env env, napi_value obj, const char* key, int dflt) {
  bool has = false;
  napi_has_named_property(env, obj, key, &has);
  if (!has) return dflt;
  napi_value v;
  napi_get_named_property(env, obj, key, &v);
  return static_cast<int>(get_num(env, v, dflt));
}

napi_value num(napi_env env, double d) {
  napi_value v;
  napi_create_double(env, d, &v);
  return v;
}

// ------------------------------------------------------------------ basics
napi_value Version(napi_env env, napi_callback_info) { return num(env, tsphere_version()); }

napi_value DeviceCount(napi_env env, napi_callback_info) {
  int n = 0;
  if (int rc = tsphere_device_count(&n)) return throw_error(env, rc, "deviceCount");
  return num(env, n);
}

napi_value StreamCreate(napi_env env, napi_callback_info) {
  tsphere_stream s = nullptr;
  if (int rc = tsphere_stream_create(&s)) return throw_error(env, rc, "streamCreate");
  return make_ptr(env, s);
}

napi_value StreamSync(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void* s = nullptr;
  if (a.argc >= 1 && !get_ptr(env, a.v[0], &s)) return type_error(env, "stream must be a BigInt");
  if (int rc = tsphere_stream_sync(s)) return throw_error(env, rc, "streamSync");
  return nullptr;
}

// ------------------------------------------------------ transform plans
// planCreate({lmax, mmax, nlat, nphi, fourier: 'auto'|'fft'|'dft', legendre: 'fp32'|'bf16x3',
//             maxFields}) -- ShtPlan.create (sht.ts:166-183)
napi_value PlanCreate(napi_env env, napi_callback_info info) {
  Args a(env, info);
  if (a.argc < 1) return type_error(env, "planCreate(cfg)");
  tsphere_sht_config c{};
  c.lmax = get_int_prop(env, a.v[0], "lmax", -1);
  c.mmax = get_int_prop(env, a.v[0], "mmax", c.lmax);
  c.nlat = get_int_prop(env, a.v[0], "nlat", -1);
  c.nphi = get_int_prop(env, a.v[0], "nphi", -1);
  c.max_fields = get_int_prop(env, a.v[0], "maxFields", 1);
  auto str_prop = [&](const char* key, const char* dflt) {
    bool has = false;
    napi_has_named_property(env, a.v[0], key, &has);
    if (!has) return std::string(dflt);
    napi_value v;
    napi_get_named_property(env, a.v[0], key, &v);
    char buf[32] = {0};
    size_t n = 0;
    if (napi_get_value_string_utf8(env, v, buf, sizeof buf, &n) != napi_ok) return std::string(dflt);
    return std::string(buf, n);
  };
  const std::string f = str_prop("fourier", "auto"), l = str_prop("legendre", "fp32");
  c.fourier = f == "fft" ? TSPHERE_FOURIER_FFT : f == "dft" ? TSPHERE_FOURIER_DFT
                                                            : TSPHERE_FOURIER_AUTO;
  if (l == "fp32")
    c.legendre = TSPHERE_LEG_FP32;
  else if (l == "f64rec")
    c.legendre = TSPHERE_LEG_F64REC;
  else if (l == "bf16x3")
    c.legendre = TSPHERE_LEG_BF16X3;
  else
    return type_error(env, "planCreate: legendre must be 'fp32', 'f64rec' or 'bf16x3'");
  tsphere_plan p = nullptr;
  if (int rc = tsphere_plan_create(&c, &p)) return throw_error(env, rc, "ShtPlan.create");
  return make_ptr(env, p);
}

napi_value PlanQuery(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void* p = nullptr;
  if (a.argc < 1 || !get_ptr(env, a.v[0], &p)) return type_error(env, "planQuery(plan)");
  tsphere_plan_info in{};
  if (int rc = tsphere_plan_query(static_cast<tsphere_plan>(p), &in))
    return throw_error(env, rc, "planQuery");
  napi_value o, s;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "lmax", num(env, in.lmax));
  napi_set_named_property(env, o, "mmax", num(env, in.mmax));
  napi_set_named_property(env, o, "nlat", num(env, in.nlat));
  napi_set_named_property(env, o, "nphi", num(env, in.nphi));
  napi_set_named_property(env, o, "nlm", num(env, static_cast<double>(in.nlm)));
  napi_create_string_utf8(env, in.fourier == TSPHERE_FOURIER_FFT ? "fft" : "dft",
                          NAPI_AUTO_LENGTH, &s);
  napi_set_named_property(env, o, "fourierMode", s);
  napi_set_named_property(env, o, "legLat", num(env, in.leg_lat));
  napi_set_named_property(env, o, "maxFields", num(env, in.max_fields));
  return o;
}

napi_value PlanDestroy(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void* p = nullptr;
  if (a.argc < 1 || !get_ptr(env, a.v[0], &p)) return type_error(env, "planDestroy(plan)");
  tsphere_plan_destroy(static_cast<tsphere_plan>(p));
  return nullptr;
}

// synth(plan, qlmPtr, spatPtr, stream?) / analys(plan, spatPtr, qlmPtr, stream?):
// encodeSynthInto / encodeAnalysInto (sht.ts:326-353) -- no implicit sync
template <bool SYNTH>
napi_value Transform(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void *p = nullptr, *src = nullptr, *dst = nullptr, *s = nullptr;
  if (a.argc < 3 || !get_ptr(env, a.v[0], &p) || !get_ptr(env, a.v[1], &src) ||
      !get_ptr(env, a.v[2], &dst) || (a.argc >= 4 && !get_ptr(env, a.v[3], &s)))
    return type_error(env, SYNTH ? "synth(plan, qlm, spat, stream?)"
                                 : "analys(plan, spat, qlm, stream?)");
  int rc = SYNTH ? tsphere_synth(static_cast<tsphere_plan>(p), static_cast<const float*>(src),
                                 static_cast<float*>(dst), s)
                 : tsphere_analys(static_cast<tsphere_plan>(p), static_cast<const float*>(src),
                                  static_cast<float*>(dst), s);
  if (rc) return throw_error(env, rc, SYNTH ? "synth" : "analys");
  return nullptr;
}

// ------------------------------------------------------------ device memory
napi_value Malloc(napi_env env, napi_callback_info info) {
  Args a(env, info);
  const double bytes = a.argc >= 1 ? get_num(env, a.v[0], -1) : -1;
  if (bytes < 0) return type_error(env, "malloc(bytes)");
  void* p = nullptr;
  if (int rc = tsphere_malloc(static_cast<size_t>(bytes), &p)) return throw_error(env, rc, "malloc");
  return make_ptr(env, p);
}

napi_value Free(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void* p = nullptr;
  if (a.argc < 1 || !get_ptr(env, a.v[0], &p)) return type_error(env, "free(ptr)");
  if (int rc = tsphere_free(p)) return throw_error(env, rc, "free");
  return nullptr;
}

// upload(ptr, Float32Array, stream?) -- queue.writeBuffer; returns after the copy
napi_value Upload(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void *p = nullptr, *s = nullptr;
  bool is_ta = false;
  if (a.argc < 2 || !get_ptr(env, a.v[0], &p)) return type_error(env, "upload(ptr, Float32Array)");
  napi_is_typedarray(env, a.v[1], &is_ta);
  if (!is_ta) return type_error(env, "upload: data must be a Float32Array");
  napi_typedarray_type t;
  size_t len = 0;
  void* data = nullptr;
  napi_get_typedarray_info(env, a.v[1], &t, &len, &data, nullptr, nullptr);
  if (t != napi_float32_array) return type_error(env, "upload: data must be a Float32Array");
  if (a.argc >= 3 && !get_ptr(env, a.v[2], &s))
    return type_error(env, "stream must be a BigInt (or undefined)");
  if (int rc = tsphere_memcpy_htod(p, data, 4 * len, s)) return throw_error(env, rc, "upload");
  if (int rc = tsphere_stream_sync(s)) return throw_error(env, rc, "upload");
  return nullptr;
}

// download(ptr, count, stream?) -> Float32Array (blocking: the readback of model.ts:261)
napi_value Download(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void *p = nullptr, *s = nullptr;
  const double n = a.argc >= 2 ? get_num(env, a.v[1], -1) : -1;
  if (a.argc < 2 || !get_ptr(env, a.v[0], &p) || n < 0)
    return type_error(env, "download(ptr, count, stream?)");
  if (a.argc >= 3 && !get_ptr(env, a.v[2], &s))
    return type_error(env, "stream must be a BigInt (or undefined)");
  napi_value ab, out;
  void* data = nullptr;
  napi_create_arraybuffer(env, 4 * static_cast<size_t>(n), &data, &ab);
  if (int rc = tsphere_memcpy_dtoh(data, p, 4 * static_cast<size_t>(n), s))
    return throw_error(env, rc, "download");
  napi_create_typedarray(env, napi_float32_array, static_cast<size_t>(n), ab, 0, &out);
  return out;
}

// ------------------------------------------------- generated kernels, graphs
napi_value ModuleCompile(napi_env env, napi_callback_info info) {
  Args a(env, info);
  size_t n = 0;
  if (a.argc < 1 || napi_get_value_string_utf8(env, a.v[0], nullptr, 0, &n) != napi_ok)
    return type_error(env, "moduleCompile(source)");
  std::vector<char> src(n + 1);
  napi_get_value_string_utf8(env, a.v[0], src.data(), n + 1, &n);
  tsphere_module m = nullptr;
  if (int rc = tsphere_module_compile(src.data(), nullptr, &m))
    return throw_error(env, rc, "moduleCompile");
  return make_ptr(env, m);
}

napi_value ModuleKernel(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void* m = nullptr;
  char name[256] = {0};
  size_t n = 0;
  if (a.argc < 2 || !get_ptr(env, a.v[0], &m) ||
      napi_get_value_string_utf8(env, a.v[1], name, sizeof name, &n) != napi_ok)
    return type_error(env, "moduleKernel(module, name)");
  tsphere_kernel k = nullptr;
  if (int rc = tsphere_module_kernel(static_cast<tsphere_module>(m), name, &k))
    return throw_error(env, rc, "moduleKernel");
  return make_ptr(env, k);
}

// kernelLaunch(kernel, count, [ptr...], [scalar...], stream?) -- scalars by value
napi_value KernelLaunch(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void *k = nullptr, *s = nullptr;
  if (a.argc < 4 || !get_ptr(env, a.v[0], &k)) return type_error(env, "kernelLaunch(k, n, ptrs, scalars)");
  const double count = get_num(env, a.v[1], -1);
  uint32_t np = 0, ns = 0;
  napi_get_array_length(env, a.v[2], &np);
  napi_get_array_length(env, a.v[3], &ns);
  std::vector<void*> ptrs(np);
  std::vector<float> sc(ns);
  for (uint32_t i = 0; i < np; i++) {
    napi_value e;
    napi_get_element(env, a.v[2], i, &e);
    if (!get_ptr(env, e, &ptrs[i])) return type_error(env, "kernelLaunch: pointers must be BigInt");
  }
  for (uint32_t i = 0; i < ns; i++) {
    napi_value e;
    napi_get_element(env, a.v[3], i, &e);
    sc[i] = static_cast<float>(get_num(env, e));
  }
  if (a.argc >= 5 && !get_ptr(env, a.v[4], &s))
    return type_error(env, "stream must be a BigInt (or undefined)");
  if (int rc = tsphere_kernel_launch(static_cast<tsphere_kernel>(k), static_cast<long long>(count),
                                     static_cast<int>(np), ptrs.data(), static_cast<int>(ns),
                                     sc.data(), s))
    return throw_error(env, rc, "kernelLaunch");
  return nullptr;
}

// Arrays of BigInt pointers / numbers from a JS array.
bool get_ptrs(napi_env env, napi_value arr, std::vector<void*>& out) {
  uint32_t n = 0;
  if (napi_get_array_length(env, arr, &n) != napi_ok) return false;
  out.resize(n);
  for (uint32_t i = 0; i < n; i++) {
    napi_value e;
    napi_get_element(env, arr, i, &e);
    if (!get_ptr(env, e, &out[i])) return false;
  }
  return true;
}

// Integers from a JS array: every entry a number with no fractional part.
bool get_ints(napi_env env, napi_value arr, std::vector<int>& out) {
  uint32_t n = 0;
  if (napi_get_array_length(env, arr, &n) != napi_ok) return false;
  out.resize(n);
  for (uint32_t i = 0; i < n; i++) {
    napi_value e;
    napi_get_element(env, arr, i, &e);
    napi_valuetype t;
    napi_typeof(env, e, &t);
    if (t != napi_number) return false;
    double d = 0;
    napi_get_value_double(env, e, &d);
    if (!(d >= -2147483648.0 && d <= 2147483647.0) || d != static_cast<double>(static_cast<int>(d)))
      return false;
    out[i] = static_cast<int>(d);
  }
  return true;
}

bool get_floats(napi_env env, napi_value arr, std::vector<float>& out) {
  uint32_t n = 0;
  if (napi_get_array_length(env, arr, &n) != napi_ok) return false;
  out.resize(n);
  for (uint32_t i = 0; i < n; i++) {
    napi_value e;
    napi_get_element(env, arr, i, &e);
    out[i] = static_cast<float>(get_num(env, e));
  }
  return true;
}

// kernelLaunchGrid(kernel, gx, gy, block, [ptr...], [scalar...], stream?)
napi_value KernelLaunchGrid(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void *k = nullptr, *s = nullptr;
  std::vector<void*> ptrs;
  std::vector<float> sc;
  if (a.argc < 6 || !get_ptr(env, a.v[0], &k) || !get_ptrs(env, a.v[4], ptrs) ||
      !get_floats(env, a.v[5], sc))
    return type_error(env, "kernelLaunchGrid(k, gx, gy, block, ptrs, scalars)");
  if (a.argc >= 7 && !get_ptr(env, a.v[6], &s))
    return type_error(env, "stream must be a BigInt (or undefined)");
  if (int rc = tsphere_kernel_launch_grid(
          static_cast<tsphere_kernel>(k), static_cast<unsigned>(get_num(env, a.v[1], 0)),
          static_cast<unsigned>(get_num(env, a.v[2], 0)),
          static_cast<unsigned>(get_num(env, a.v[3], 0)), static_cast<int>(ptrs.size()),
          ptrs.data(), static_cast<int>(sc.size()), sc.data(), s))
    return throw_error(env, rc, "kernelLaunchGrid");
  return nullptr;
}

// legSynth(plan, [qlm...], fm, stream?) / legAnalys(plan, fm, [qlm...], stream?)
template <bool SYNTH>
napi_value LegStage(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void *p = nullptr, *fm = nullptr, *s = nullptr;
  std::vector<void*> q;
  const int qi = SYNTH ? 1 : 2, fi = SYNTH ? 2 : 1;
  if (a.argc < 3 || !get_ptr(env, a.v[0], &p) || !get_ptrs(env, a.v[qi], q) ||
      !get_ptr(env, a.v[fi], &fm))
    return type_error(env, SYNTH ? "legSynth(plan, qlmPtrs, fm)" : "legAnalys(plan, fm, qlmPtrs)");
  if (a.argc >= 4 && !get_ptr(env, a.v[3], &s))
    return type_error(env, "stream must be a BigInt (or undefined)");
  const int n = static_cast<int>(q.size());
  const int rc =
      SYNTH ? tsphere_leg_synth_multi(static_cast<tsphere_plan>(p), n,
                                      reinterpret_cast<const float* const*>(q.data()),
                                      static_cast<float*>(fm), s)
            : tsphere_leg_analys_multi(static_cast<tsphere_plan>(p), n,
                                       static_cast<const float*>(fm),
                                       reinterpret_cast<float* const*>(q.data()), s);
  if (rc) return throw_error(env, rc, SYNTH ? "legSynth" : "legAnalys");
  return nullptr;
}

// legImexSynth(plan, [analysed...], [state...], lam, [s1...], [s2...], fm, stream?):
// each state field's implicit step X = (X + s1 A) / (1 + s2 lam) from its
// analysed values A, then the Legendre synthesis of the new state into fm
// (the planner's "fold" form of a step, after the grid op and legAnalys)
napi_value LegImexSynth(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void *p = nullptr, *lam = nullptr, *fm = nullptr, *s = nullptr;
  std::vector<void*> an, st;
  std::vector<float> s1, s2;
  if (a.argc < 7 || !get_ptr(env, a.v[0], &p) || !get_ptrs(env, a.v[1], an) ||
      !get_ptrs(env, a.v[2], st) || !get_ptr(env, a.v[3], &lam) || !get_floats(env, a.v[4], s1) ||
      !get_floats(env, a.v[5], s2) || !get_ptr(env, a.v[6], &fm))
    return type_error(env, "legImexSynth(plan, analysedPtrs, statePtrs, lam, s1, s2, fm)");
  const size_t n = st.size();
  if (an.size() != n || s1.size() != n || s2.size() != n)
    return type_error(env, "legImexSynth: one analysed pointer, s1 and s2 per state field");
  if (a.argc >= 8 && !get_ptr(env, a.v[7], &s))
    return type_error(env, "stream must be a BigInt (or undefined)");
  if (int rc = tsphere_leg_imex_synth(static_cast<tsphere_plan>(p), static_cast<int>(n),
                                      reinterpret_cast<const float* const*>(an.data()),
                                      reinterpret_cast<float* const*>(st.data()),
                                      static_cast<const float*>(lam), s1.data(), s2.data(),
                                      static_cast<float*>(fm), s))
    return throw_error(env, rc, "legImexSynth");
  return nullptr;
}

// legSpectralImex(plan, fm, [state...], lam, [s1...], [s2...], [species...], stream?):
// the Legendre analysis of fm, each field's implicit step into state
// species[f] and the Legendre synthesis of the new state into fm, one launch
// (the planner's default "spectral" form of a step, after the grid op)
napi_value LegSpectralImex(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void *p = nullptr, *fm = nullptr, *lam = nullptr, *s = nullptr;
  std::vector<void*> st;
  std::vector<float> s1, s2;
  std::vector<int> spc;
  if (a.argc < 7 || !get_ptr(env, a.v[0], &p) || !get_ptr(env, a.v[1], &fm) ||
      !get_ptrs(env, a.v[2], st) || !get_ptr(env, a.v[3], &lam) || !get_floats(env, a.v[4], s1) ||
      !get_floats(env, a.v[5], s2))
    return type_error(env, "legSpectralImex(plan, fm, statePtrs, lam, s1, s2, species)");
  if (!get_ints(env, a.v[6], spc))
    return type_error(env, "legSpectralImex: species must be integer state indices");
  const size_t n = st.size();
  if (s1.size() != n || s2.size() != n || spc.size() != n)
    return type_error(env, "legSpectralImex: one s1, s2 and species entry per state field");
  if (a.argc >= 8 && !get_ptr(env, a.v[7], &s))
    return type_error(env, "stream must be a BigInt (or undefined)");
  if (int rc = tsphere_leg_spectral_imex(static_cast<tsphere_plan>(p), static_cast<int>(n),
                                         static_cast<float*>(fm),
                                         reinterpret_cast<float* const*>(st.data()),
                                         static_cast<const float*>(lam), s1.data(), s2.data(),
                                         spc.data(), s))
    return throw_error(env, rc, "legSpectralImex");
  return nullptr;
}

// fourierTable(plan) -> device pointer of the plan's twiddle table
napi_value FourierTable(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void* p = nullptr;
  if (a.argc < 1 || !get_ptr(env, a.v[0], &p)) return type_error(env, "fourierTable(plan)");
  const void* t = nullptr;
  if (int rc = tsphere_plan_fourier_table(static_cast<tsphere_plan>(p), &t))
    return throw_error(env, rc, "fourierTable");
  return make_ptr(env, t);
}

napi_value GraphBegin(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void* s = nullptr;
  if (a.argc < 1 || !get_ptr(env, a.v[0], &s)) return type_error(env, "graphBegin(stream)");
  if (int rc = tsphere_graph_begin(s)) return throw_error(env, rc, "graphBegin");
  return nullptr;
}

napi_value GraphEnd(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void* s = nullptr;
  if (a.argc < 1 || !get_ptr(env, a.v[0], &s)) return type_error(env, "graphEnd(stream)");
  tsphere_graph g = nullptr;
  if (int rc = tsphere_graph_end(s, &g)) return throw_error(env, rc, "graphEnd");
  return make_ptr(env, g);
}

napi_value GraphLaunch(napi_env env, napi_callback_info info) {
  Args a(env, info);
  void *g = nullptr, *s = nullptr;
  if (a.argc < 2 || !get_ptr(env, a.v[0], &g) || !get_ptr(env, a.v[1], &s))
    return type_error(env, "graphLaunch(graph, stream)");
  if (int rc = tsphere_graph_launch(static_cast<tsphere_graph>(g), s))
    return throw_error(env, rc, "graphLaunch");
  return nullptr;
}

}  // namespace

NAPI_MODULE_INIT() {
  const napi_property_descriptor d[] = {
      {"version", nullptr, Version, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"deviceCount", nullptr, DeviceCount, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"streamCreate", nullptr, StreamCreate, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"streamSync", nullptr, StreamSync, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"planCreate", nullptr, PlanCreate, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"planQuery", nullptr, PlanQuery, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"planDestroy", nullptr, PlanDestroy, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"synth", nullptr, Transform<true>, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"analys", nullptr, Transform<false>, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"malloc", nullptr, Malloc, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"free", nullptr, Free, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"upload", nullptr, Upload, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"download", nullptr, Download, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"moduleCompile", nullptr, ModuleCompile, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"moduleKernel", nullptr, ModuleKernel, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"kernelLaunch", nullptr, KernelLaunch, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"kernelLaunchGrid", nullptr, KernelLaunchGrid, nullptr, nullptr, nullptr, napi_default,
       nullptr},
      {"legSynth", nullptr, LegStage<true>, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"legAnalys", nullptr, LegStage<false>, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"legImexSynth", nullptr, LegImexSynth, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"legSpectralImex", nullptr, LegSpectralImex, nullptr, nullptr, nullptr, napi_default,
       nullptr},
      {"fourierTable", nullptr, FourierTable, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"graphBegin", nullptr, GraphBegin, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"graphEnd", nullptr, GraphEnd, nullptr, nullptr, nullptr, napi_default, nullptr},
      {"graphLaunch", nullptr, GraphLaunch, nullptr, nullptr, nullptr, napi_default, nullptr},
  };
  napi_define_properties(env, exports, sizeof d / sizeof *d, d);
  return exports;
}
