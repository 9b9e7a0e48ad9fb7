// gpugo.hip -- libgpugo.so: C ABI (include/gpugo.h), objective-independent
// kernels (interval op batch, select, compact, box materialisation) and the
// host side of the branch-and-bound loop.  Objective-specific kernels live in
// hipRTC modules built from text by gpugo_amd.emit.
//
// Reference correspondence (paths relative to GPUGO_1.1/):
//   k_ia_eval        IA device functions ("Updated Interval Arithmetic
//                    Operations/IA.py":57-570)
//   k_select         list prune + argmin pop + completeRegion + cycle update
//                    (1-Ackley/ackley50.py:1154-1220, :1270-1370, :820-948),
//                    generalised to the K smallest boxes per step
//   module kernel    sampling + ruleOutGPU (1-Ackley/ackley50.py:602-766)
//   k_compact        constructFlagArr + mask compaction (:768-810, :1217-1242)
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpugo.h"
#include "ia.hpp"
#include "bnb_args.hpp"

using namespace gpugo;

#define GPUGO_VERSION 100  // 0.1.0

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int set_err(int code, const char *fmt, ...) {
  char buf[2048];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHECK(x)                                                              \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess)                                                        \
      return set_err(e_ == hipErrorOutOfMemory ? GPUGO_E_NOMEM : GPUGO_E_HIP,    \
                     "%s: %s", #x, hipGetErrorString(e_));                       \
  } while (0)

extern "C" const char *gpugo_last_error(void) { return g_err.c_str(); }
extern "C" int gpugo_version(void) { return GPUGO_VERSION; }

extern "C" int gpugo_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

static int static_arena(int dev, unsigned char **out);

extern "C" int gpugo_init(int device) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return set_err(GPUGO_E_NODEV, "no HIP device available");
  if (device < 0 || device >= n) return set_err(GPUGO_E_INVALID, "device %d out of range [0,%d)", device, n);
  HIPCHECK(hipSetDevice(device));
  HIPCHECK(hipFree(nullptr));
  unsigned char *arena = nullptr;
  return static_arena(device, &arena);
}

// ---------------------------------------------------------------------------
// batched interval ops
// ---------------------------------------------------------------------------
template <typename T, int R>
__global__ __launch_bounds__(256) void k_ia_eval(int op, long long n, const double *__restrict__ alo,
                                                 const double *__restrict__ ahi, const double *__restrict__ blo,
                                                 const double *__restrict__ bhi, double s,
                                                 double *__restrict__ olo, double *__restrict__ ohi) {
  typedef IA<T, R> A;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  iv<T> a = mk<T>(to_dn<T>(alo[i]), to_up<T>(ahi[i]));
  iv<T> b = blo ? mk<T>(to_dn<T>(blo[i]), to_up<T>(bhi[i])) : mk<T>(T(0), T(0));
  // scalar operand as an interval (exact for FP64; outward for FP32)
  iv<T> S = mk<T>(to_dn<T>(s), to_up<T>(s));
  bool s_exact = S.lo == S.hi;
  iv<T> r;
  switch (op) {
    case GPUGO_OP_ADD: r = A::add(a, b); break;
    case GPUGO_OP_SUB: r = A::sub(a, b); break;
    case GPUGO_OP_MUL: r = A::mul(a, b); break;
    case GPUGO_OP_DIV: r = A::div(a, b); break;
    case GPUGO_OP_NEG: r = A::neg(a); break;
    case GPUGO_OP_SCALE: r = s_exact ? A::scale(a, S.lo) : A::mul(a, S); break;
    case GPUGO_OP_ADDC: r = s_exact ? A::addc(a, S.lo) : A::add(a, S); break;
    case GPUGO_OP_RSUBC: r = s_exact ? A::rsubc(S.lo, a) : A::sub(S, a); break;
    case GPUGO_OP_DIVC: r = s_exact ? A::divc(a, S.lo) : A::div(a, S); break;
    case GPUGO_OP_SQR: r = A::sqr(a); break;
    case GPUGO_OP_POWN: r = A::pown(a, (int)s); break;
    case GPUGO_OP_SQRT: r = A::sqrt(a); break;
    case GPUGO_OP_EXP: r = A::exp(a); break;
    case GPUGO_OP_SIN: r = A::sin(a); break;
    case GPUGO_OP_COS: r = A::cos(a); break;
    default: r = mk<T>(T(0), T(0)); break;
  }
  olo[i] = (double)r.lo;
  ohi[i] = (double)r.hi;
}

static bool is_unary(int op) {
  return op == GPUGO_OP_NEG || op == GPUGO_OP_SQR || op == GPUGO_OP_SQRT || op == GPUGO_OP_EXP ||
         op == GPUGO_OP_SIN || op == GPUGO_OP_COS;
}
static bool is_binary(int op) {
  return op == GPUGO_OP_ADD || op == GPUGO_OP_SUB || op == GPUGO_OP_MUL || op == GPUGO_OP_DIV;
}

static int check_intervals(const double *lo, const double *hi, int64_t n, const char *name) {
  for (int64_t i = 0; i < n; ++i)
    if (!(lo[i] <= hi[i]))
      return set_err(GPUGO_E_DOMAIN, "%s[%lld] = [%g, %g] is NaN or inverted", name, (long long)i, lo[i], hi[i]);
  return GPUGO_OK;
}

struct DevBuf {
  void *p = nullptr;
  bool own = true;  // false: points into the static arena (never freed)
  ~DevBuf() { if (p && own) (void)hipFree(p); }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 8); }
  template <typename T> T *as() const { return (T *)p; }
};

extern "C" int gpugo_ia_eval(int op, int prec, int rounding, int64_t n, const double *alo, const double *ahi,
                             const double *blo, const double *bhi, double s, double *olo, double *ohi) {
  if (op < 0 || op >= GPUGO_OP_COUNT) return set_err(GPUGO_E_INVALID, "unknown op %d", op);
  if (prec != GPUGO_PREC_F64 && prec != GPUGO_PREC_F32) return set_err(GPUGO_E_INVALID, "unknown precision %d", prec);
  if (rounding != GPUGO_ROUND_EXACT && rounding != GPUGO_ROUND_FAST && rounding != GPUGO_ROUND_DIGITS15)
    return set_err(GPUGO_E_INVALID, "unknown rounding %d", rounding);
  if (rounding == GPUGO_ROUND_DIGITS15 && prec != GPUGO_PREC_F64)
    return set_err(GPUGO_E_INVALID, "15-digit rounding is FP64 only");
  if (n < 0) return set_err(GPUGO_E_INVALID, "negative n");
  if (n == 0) return GPUGO_OK;
  if (!alo || !ahi || !olo || !ohi) return set_err(GPUGO_E_INVALID, "null pointer");
  if (is_binary(op) && (!blo || !bhi)) return set_err(GPUGO_E_INVALID, "binary op needs b");
  if (std::isnan(s)) return set_err(GPUGO_E_DOMAIN, "scalar is NaN");
  if ((op == GPUGO_OP_DIVC) && s == 0.0) return set_err(GPUGO_E_DOMAIN, "division by scalar zero");
  if (op == GPUGO_OP_POWN && (s < 0 || s > 1024 || s != std::floor(s)))
    return set_err(GPUGO_E_DOMAIN, "pown exponent must be an integer in [0,1024]");
  int rc = check_intervals(alo, ahi, n, "a");
  if (rc) return rc;
  if (is_binary(op)) {
    rc = check_intervals(blo, bhi, n, "b");
    if (rc) return rc;
  }
  size_t bytes = sizeof(double) * (size_t)n;
  DevBuf d_alo, d_ahi, d_blo, d_bhi, d_olo, d_ohi;
  HIPCHECK(d_alo.alloc(bytes));
  HIPCHECK(d_ahi.alloc(bytes));
  HIPCHECK(d_olo.alloc(bytes));
  HIPCHECK(d_ohi.alloc(bytes));
  HIPCHECK(hipMemcpy(d_alo.p, alo, bytes, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d_ahi.p, ahi, bytes, hipMemcpyHostToDevice));
  const double *pb_lo = nullptr, *pb_hi = nullptr;
  if (is_binary(op)) {
    HIPCHECK(d_blo.alloc(bytes));
    HIPCHECK(d_bhi.alloc(bytes));
    HIPCHECK(hipMemcpy(d_blo.p, blo, bytes, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_bhi.p, bhi, bytes, hipMemcpyHostToDevice));
    pb_lo = d_blo.as<double>();
    pb_hi = d_bhi.as<double>();
  }
  dim3 grid((unsigned)((n + 255) / 256)), block(256);
  const double *a0 = d_alo.as<double>(), *a1 = d_ahi.as<double>();
  double *o0 = d_olo.as<double>(), *o1 = d_ohi.as<double>();
  if (rounding == GPUGO_ROUND_DIGITS15)
    hipLaunchKernelGGL((k_ia_eval<double, ROUND_DIGITS15>), grid, block, 0, 0, op, (long long)n, a0, a1, pb_lo, pb_hi, s, o0, o1);
  else if (prec == GPUGO_PREC_F64 && rounding == GPUGO_ROUND_EXACT)
    hipLaunchKernelGGL((k_ia_eval<double, ROUND_EXACT>), grid, block, 0, 0, op, (long long)n, a0, a1, pb_lo, pb_hi, s, o0, o1);
  else if (prec == GPUGO_PREC_F64)
    hipLaunchKernelGGL((k_ia_eval<double, ROUND_FAST>), grid, block, 0, 0, op, (long long)n, a0, a1, pb_lo, pb_hi, s, o0, o1);
  else if (rounding == GPUGO_ROUND_EXACT)
    hipLaunchKernelGGL((k_ia_eval<float, ROUND_EXACT>), grid, block, 0, 0, op, (long long)n, a0, a1, pb_lo, pb_hi, s, o0, o1);
  else
    hipLaunchKernelGGL((k_ia_eval<float, ROUND_FAST>), grid, block, 0, 0, op, (long long)n, a0, a1, pb_lo, pb_hi, s, o0, o1);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpy(olo, d_olo.p, bytes, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(ohi, d_ohi.p, bytes, hipMemcpyDeviceToHost));
  return GPUGO_OK;
}

// ---------------------------------------------------------------------------
// modules
// ---------------------------------------------------------------------------
struct gpugo_module_s {
  hipModule_t mod = nullptr;
  hipFunction_t bound[3] = {nullptr, nullptr, nullptr};     // [rounding]
  hipFunction_t bound_mv[3] = {nullptr, nullptr, nullptr};  // [rounding], mean-value variant
  hipFunction_t eval[3] = {nullptr, nullptr, nullptr};
  hipFunction_t eval32[2] = {nullptr, nullptr};  // FP32 value (precision F32 modules)
  hipFunction_t sample32[2] = {nullptr, nullptr};  // FP32 sample pass of a solve's first step
  hipFunction_t bound_g[2] = {nullptr, nullptr};     // term-parallel variants (split modules)
  hipFunction_t bound_mv_g[2] = {nullptr, nullptr};
  int split_g = 0;                                    // lanes per box of those variants
  // the step's select / compact / root kernels compiled into the module
  // (used instead of the library's copies: no code-object switch per launch)
  hipFunction_t select_fn = nullptr, compact_fn = nullptr, root_fn = nullptr;
  hipFunction_t compact_select_fn = nullptr;  // compaction fused with the next step's select
  hipFunction_t persist[2] = {nullptr, nullptr};     // persistent step loops (split modules)
  hipFunction_t persist_mv[2] = {nullptr, nullptr};
  // windowed modules (large d): rest-sum kernel per rounding, rest slots
  hipFunction_t prep[3] = {nullptr, nullptr, nullptr};
  hipFunction_t prep_fold[3] = {nullptr, nullptr, nullptr};  // rest sums split over workgroups
  int nrest = 0;
  // term tables (windowed modules): stencil width of each rest slot and of
  // the gradient template; tab bit 0: tables, bit 1: gradient table
  int tab = 0, tab_w[8] = {0, 0, 0, 0, 0, 0, 0, 0}, tab_gw = 1, tab_parts = 0;
  int dim = 0, nparams = 0, abi = 0;
  bool f32 = false;                // has the FP32 value code list (precision F32)
  std::vector<signed char> psign;  // declared coefficient signs
  std::vector<char> code;
  // batch modules (several families, runtime d <= dim)
  bool batch = false, batch_mv = false;
  hipFunction_t batch_bound[2] = {nullptr, nullptr};
  int nfam = 0, pmax = 0;
  std::vector<int> fam_nparams;
  std::vector<signed char> fam_psign;  // [nfam][pmax]
};

#define GPUGO_MODULE_ABI GPUGO_MODULE_ABI_V

static int batch_module_finish(gpugo_module_t m, hipFunction_t info) {
  m->batch = true;
  DevBuf d;
  HIPCHECK(d.alloc(64 * sizeof(int)));
  void *args[] = {&d.p};
  HIPCHECK(hipModuleLaunchKernel(info, 1, 1, 1, 1, 1, 1, 0, nullptr, args, nullptr));
  int h[64];
  HIPCHECK(hipMemcpy(h, d.p, sizeof h, hipMemcpyDeviceToHost));
  m->dim = h[0];
  m->nfam = h[1];
  m->abi = h[2];
  m->pmax = h[4];
  if (m->abi != GPUGO_MODULE_ABI)
    return set_err(GPUGO_E_INVALID, "module ABI %d, library expects %d", m->abi, GPUGO_MODULE_ABI);
  if (m->nfam < 1 || m->nfam > 59 || m->dim < 1 || m->pmax < 1)
    return set_err(GPUGO_E_INVALID, "malformed batch module (dmax %d, families %d)", m->dim, m->nfam);
  m->fam_nparams.assign(h + 5, h + 5 + m->nfam);
  m->batch_mv = (h[3] & 4) != 0;
  if (h[3] & 1) HIPCHECK(hipModuleGetFunction(&m->batch_bound[0], m->mod, "gpugo_batch_bound_exact"));
  if (h[3] & 2) HIPCHECK(hipModuleGetFunction(&m->batch_bound[1], m->mod, "gpugo_batch_bound_fast"));
  hipFunction_t fs;
  HIPCHECK(hipModuleGetFunction(&fs, m->mod, "gpugo_batch_param_signs"));
  DevBuf ds;
  const size_t ns = (size_t)m->nfam * m->pmax;
  HIPCHECK(ds.alloc(ns));
  void *sargs[] = {&ds.p};
  HIPCHECK(hipModuleLaunchKernel(fs, 1, 1, 1, 256, 1, 1, 0, nullptr, sargs, nullptr));
  m->fam_psign.assign(ns, 0);
  HIPCHECK(hipMemcpy(m->fam_psign.data(), ds.p, ns, hipMemcpyDeviceToHost));
  return GPUGO_OK;
}

static int module_finish(gpugo_module_t m) {
  HIPCHECK(hipModuleLoadData(&m->mod, m->code.data()));
  hipFunction_t info;
  if (hipModuleGetFunction(&info, m->mod, "gpugo_info") != hipSuccess) {
    (void)hipGetLastError();  // the lookup's error must not surface at a later launch check
    if (hipModuleGetFunction(&info, m->mod, "gpugo_batch_info") != hipSuccess) {
      (void)hipGetLastError();
      return set_err(GPUGO_E_INVALID, "code object is neither an objective module nor a batch module");
    }
    return batch_module_finish(m, info);
  }
  DevBuf d;
  int h[32];
  memset(h, 0, sizeof h);
  HIPCHECK(d.alloc(sizeof h));
  HIPCHECK(hipMemset(d.p, 0, sizeof h));
  void *args[] = {&d.p};
  HIPCHECK(hipModuleLaunchKernel(info, 1, 1, 1, 1, 1, 1, 0, nullptr, args, nullptr));
  HIPCHECK(hipMemcpy(h, d.p, sizeof h, hipMemcpyDeviceToHost));
  m->dim = h[0];
  m->nparams = h[1];
  m->abi = h[2];
  m->f32 = h[4] != 0;
  m->split_g = h[5];
  m->nrest = h[6];
  if (m->nrest > 0 && m->nrest <= 8 && (h[7] & 1)) {
    m->tab = h[7] & 7;
    for (int s2 = 0; s2 < m->nrest; ++s2) m->tab_w[s2] = h[8 + s2];
    m->tab_gw = h[16];
    m->tab_parts = (h[7] & 4) ? h[17] : 0;
  }
  if (hipModuleGetFunction(&m->select_fn, m->mod, "gpugo_select") != hipSuccess ||
      hipModuleGetFunction(&m->compact_fn, m->mod, "gpugo_compact") != hipSuccess ||
      hipModuleGetFunction(&m->root_fn, m->mod, "gpugo_root") != hipSuccess ||
      hipModuleGetFunction(&m->compact_select_fn, m->mod, "gpugo_compact_select") != hipSuccess) {
    (void)hipGetLastError();
    m->select_fn = m->compact_fn = m->root_fn = m->compact_select_fn = nullptr;
  }
  if (m->split_g > 1) {
    if (h[3] & 1) {
      HIPCHECK(hipModuleGetFunction(&m->bound_g[0], m->mod, "gpugo_bound_exact_g"));
      HIPCHECK(hipModuleGetFunction(&m->bound_mv_g[0], m->mod, "gpugo_bound_exact_mv_g"));
    }
    if (h[3] & 2) {
      HIPCHECK(hipModuleGetFunction(&m->bound_g[1], m->mod, "gpugo_bound_fast_g"));
      HIPCHECK(hipModuleGetFunction(&m->bound_mv_g[1], m->mod, "gpugo_bound_fast_mv_g"));
    }
    const char *pn[2][2] = {{"gpugo_persist_exact_g", "gpugo_persist_exact_mv_g"},
                            {"gpugo_persist_fast_g", "gpugo_persist_fast_mv_g"}};
    for (int r = 0; r < 2; ++r) {
      if (!(h[3] & (1 << r))) continue;
      if (hipModuleGetFunction(&m->persist[r], m->mod, pn[r][0]) != hipSuccess ||
          hipModuleGetFunction(&m->persist_mv[r], m->mod, pn[r][1]) != hipSuccess) {
        (void)hipGetLastError();
        m->persist[r] = m->persist_mv[r] = nullptr;
      }
    }
  }
  if (m->f32 && (h[3] & 1)) {
    HIPCHECK(hipModuleGetFunction(&m->eval32[0], m->mod, "gpugo_eval32_exact"));
    HIPCHECK(hipModuleGetFunction(&m->sample32[0], m->mod, "gpugo_sample32_exact"));
  }
  if (m->f32 && (h[3] & 2)) {
    HIPCHECK(hipModuleGetFunction(&m->eval32[1], m->mod, "gpugo_eval32_fast"));
    HIPCHECK(hipModuleGetFunction(&m->sample32[1], m->mod, "gpugo_sample32_fast"));
  }
  if (m->abi != GPUGO_MODULE_ABI)
    return set_err(GPUGO_E_INVALID, "module ABI %d, library expects %d", m->abi, GPUGO_MODULE_ABI);
  const char *rname[3] = {"exact", "fast", "digits15"};
  for (int r = 0; r < 3; ++r) {
    if (!(h[3] & (1 << r))) continue;
    char nm[64];
    snprintf(nm, sizeof nm, "gpugo_bound_%s", rname[r]);
    HIPCHECK(hipModuleGetFunction(&m->bound[r], m->mod, nm));
    snprintf(nm, sizeof nm, "gpugo_eval_%s", rname[r]);
    HIPCHECK(hipModuleGetFunction(&m->eval[r], m->mod, nm));
    if (m->nrest > 0) {
      // windowed module: rest sums per parent record, no mean-value variant
      snprintf(nm, sizeof nm, "gpugo_prep_%s", rname[r]);
      HIPCHECK(hipModuleGetFunction(&m->prep[r], m->mod, nm));
      snprintf(nm, sizeof nm, "gpugo_prep_fold_%s", rname[r]);
      HIPCHECK(hipModuleGetFunction(&m->prep_fold[r], m->mod, nm));
    } else {
      snprintf(nm, sizeof nm, "gpugo_bound_%s_mv", rname[r]);
      HIPCHECK(hipModuleGetFunction(&m->bound_mv[r], m->mod, nm));
    }
  }
  // declared coefficient signs
  m->psign.assign(m->nparams, 0);
  if (m->nparams > 0) {
    hipFunction_t fs;
    HIPCHECK(hipModuleGetFunction(&fs, m->mod, "gpugo_param_signs"));
    DevBuf ds;
    HIPCHECK(ds.alloc(m->nparams));
    void *sargs[] = {&ds.p};
    HIPCHECK(hipModuleLaunchKernel(fs, 1, 1, 1, 256, 1, 1, 0, nullptr, sargs, nullptr));
    HIPCHECK(hipMemcpy(m->psign.data(), ds.p, m->nparams, hipMemcpyDeviceToHost));
  }
  return GPUGO_OK;
}

static int rtc_compile(const char *source, const char *name, const char *extra_options, std::vector<char> &code) {
  hiprtcProgram prog;
  hiprtcResult r = hiprtcCreateProgram(&prog, source, name ? name : "gpugo_objective.hip", 0, nullptr, nullptr);
  if (r != HIPRTC_SUCCESS) return set_err(GPUGO_E_COMPILE, "hiprtcCreateProgram: %s", hiprtcGetErrorString(r));
  std::vector<std::string> opts = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                                   "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  if (extra_options && *extra_options) {
    std::string e(extra_options), tok;
    for (size_t i = 0; i <= e.size(); ++i) {
      if (i == e.size() || e[i] == ' ') {
        if (!tok.empty()) opts.push_back(tok);
        tok.clear();
      } else
        tok.push_back(e[i]);
    }
  }
  std::vector<const char *> copts;
  for (auto &o : opts) copts.push_back(o.c_str());
  r = hiprtcCompileProgram(prog, (int)copts.size(), copts.data());
  if (r != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls + 1, '\0');
    hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    return set_err(GPUGO_E_COMPILE, "hipRTC compile failed: %s", log.c_str());
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  code.resize(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  return GPUGO_OK;
}

static thread_local std::vector<char> g_rtc_code;

extern "C" int gpugo_rtc_compile(const char *source, const char *name, const char *extra_options,
                                 size_t *code_size) {
  if (!source || !code_size) return set_err(GPUGO_E_INVALID, "null pointer");
  g_rtc_code.clear();
  int rc = rtc_compile(source, name, extra_options, g_rtc_code);
  *code_size = g_rtc_code.size();
  return rc;
}

extern "C" int gpugo_rtc_code_copy(void *buf, size_t size) {
  if (!buf) return set_err(GPUGO_E_INVALID, "null pointer");
  if (size < g_rtc_code.size()) return set_err(GPUGO_E_INVALID, "buffer too small");
  memcpy(buf, g_rtc_code.data(), g_rtc_code.size());
  return GPUGO_OK;
}

extern "C" int gpugo_module_compile(const char *source, const char *name, const char *extra_options,
                                    gpugo_module_t *out) {
  if (!source || !out) return set_err(GPUGO_E_INVALID, "null pointer");
  *out = nullptr;
  auto *m = new gpugo_module_s;
  int rc = rtc_compile(source, name, extra_options, m->code);
  if (rc) {
    delete m;
    return rc;
  }
  rc = module_finish(m);
  if (rc) {
    if (m->mod) (void)hipModuleUnload(m->mod);
    delete m;
    return rc;
  }
  *out = m;
  return GPUGO_OK;
}

extern "C" int gpugo_module_load(const void *code_object, size_t size, gpugo_module_t *out) {
  if (!code_object || !out || size == 0) return set_err(GPUGO_E_INVALID, "null pointer or empty code");
  *out = nullptr;
  auto *m = new gpugo_module_s;
  m->code.assign((const char *)code_object, (const char *)code_object + size);
  int rc = module_finish(m);
  if (rc) {
    if (m->mod) (void)hipModuleUnload(m->mod);
    delete m;
    return rc;
  }
  *out = m;
  return GPUGO_OK;
}

extern "C" int gpugo_module_code(gpugo_module_t m, void *buf, size_t *size) {
  if (!m || !size) return set_err(GPUGO_E_INVALID, "null pointer");
  if (buf) {
    if (*size < m->code.size()) return set_err(GPUGO_E_INVALID, "buffer too small");
    memcpy(buf, m->code.data(), m->code.size());
  }
  *size = m->code.size();
  return GPUGO_OK;
}

extern "C" void gpugo_module_destroy(gpugo_module_t m) {
  if (!m) return;
  if (m->mod) (void)hipModuleUnload(m->mod);
  delete m;
}

extern "C" int gpugo_window_ranges(int64_t n, const int32_t *args, int32_t *out) {
  if (n < 0 || (n > 0 && (!args || !out))) return set_err(GPUGO_E_INVALID, "null pointer or negative n");
  for (int64_t i = 0; i < n; ++i) {
    const int32_t *a = args + 6 * i;
    const int win = a[0], c = a[1], d = a[2], klo = a[3], khi = a[4], nt = a[5];
    if (d < 1 || c < 1 || c > d || win < 0 || win >= d || klo > khi || nt < 0)
      return set_err(GPUGO_E_INVALID, "row %lld: need 0 <= win < d, 1 <= c <= d, klo <= khi, nt >= 0",
                     (long long)i);
    int lo[2] = {0, 0}, hi[2] = {0, 0};
    const int r = win_ranges(win, c, d, klo, khi, nt, lo, hi);
    int32_t *o = out + 5 * i;
    o[0] = r;
    o[1] = lo[0];
    o[2] = hi[0];
    o[3] = lo[1];
    o[4] = hi[1];
  }
  return GPUGO_OK;
}

static int check_module_problem(gpugo_module_t m, int32_t d, const double *params, int64_t n_params) {
  if (!m) return set_err(GPUGO_E_INVALID, "null module");
  if (m->batch) return set_err(GPUGO_E_INVALID, "batch module: use gpugo_batch_*");
  if (d != m->dim) return set_err(GPUGO_E_INVALID, "module is compiled for d=%d, got d=%d", m->dim, d);
  if (n_params != m->nparams)
    return set_err(GPUGO_E_INVALID, "module expects %d parameters, got %lld", m->nparams, (long long)n_params);
  if (n_params > 0 && !params) return set_err(GPUGO_E_INVALID, "null params");
  for (int64_t i = 0; i < n_params; ++i) {
    if (!std::isfinite(params[i])) return set_err(GPUGO_E_DOMAIN, "params[%lld] is not finite", (long long)i);
    const int s = i < (int64_t)m->psign.size() ? m->psign[i] : 0;
    if ((s > 0 && params[i] < 0) || (s < 0 && params[i] > 0))
      return set_err(GPUGO_E_DOMAIN, "params[%lld] = %g violates the sign the module was compiled for (%s)",
                     (long long)i, params[i], s > 0 ? ">= 0" : "<= 0");
  }
  return GPUGO_OK;
}

extern "C" int gpugo_eval_boxes(gpugo_module_t m, int32_t d, const double *params, int64_t n_params, int64_t n,
                                const double *lo, const double *hi, const double *dom_lo, const double *dom_hi,
                                int32_t n_samples, int32_t rounding, double *flo, double *fhi, double *glo,
                                double *ghi, int32_t *kkt, double *ub) {
  int rc = check_module_problem(m, d, params, n_params);
  if (rc) return rc;
  if (rounding != GPUGO_ROUND_EXACT && rounding != GPUGO_ROUND_FAST && rounding != GPUGO_ROUND_DIGITS15)
    return set_err(GPUGO_E_INVALID, "unknown rounding %d", rounding);
  if (!m->eval[rounding]) return set_err(GPUGO_E_INVALID, "module has no variant for rounding %d", rounding);
  if (n < 0) return set_err(GPUGO_E_INVALID, "negative n");
  if (n == 0) return GPUGO_OK;
  if (!lo || !hi) return set_err(GPUGO_E_INVALID, "null box pointer");
  if ((kkt || glo || ghi) && (!dom_lo || !dom_hi) && kkt) return set_err(GPUGO_E_INVALID, "kkt needs dom_lo/dom_hi");
  if (ub && n_samples < 1) return set_err(GPUGO_E_INVALID, "n_samples must be >= 1");
  rc = check_intervals(lo, hi, n * d, "box");
  if (rc) return rc;
  const size_t nb = sizeof(double) * (size_t)n * d;
  DevBuf d_lo, d_hi, d_par, d_dlo, d_dhi, d_flo, d_fhi, d_glo, d_ghi, d_ub, d_kkt;
  HIPCHECK(d_lo.alloc(nb));
  HIPCHECK(d_hi.alloc(nb));
  HIPCHECK(hipMemcpy(d_lo.p, lo, nb, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d_hi.p, hi, nb, hipMemcpyHostToDevice));
  HIPCHECK(d_par.alloc(sizeof(double) * (size_t)(n_params > 0 ? n_params : 1)));
  if (n_params > 0) HIPCHECK(hipMemcpy(d_par.p, params, sizeof(double) * n_params, hipMemcpyHostToDevice));
  HIPCHECK(d_dlo.alloc(sizeof(double) * d));
  HIPCHECK(d_dhi.alloc(sizeof(double) * d));
  if (dom_lo && dom_hi) {
    HIPCHECK(hipMemcpy(d_dlo.p, dom_lo, sizeof(double) * d, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_dhi.p, dom_hi, sizeof(double) * d, hipMemcpyHostToDevice));
  }
  EvalArgs a;
  memset(&a, 0, sizeof a);
  a.params = d_par.as<double>();
  a.lo = d_lo.as<double>();
  a.hi = d_hi.as<double>();
  a.dom_lo = d_dlo.as<double>();
  a.dom_hi = d_dhi.as<double>();
  const size_t n1 = sizeof(double) * (size_t)n;
  if (flo || fhi) {
    HIPCHECK(d_flo.alloc(n1));
    HIPCHECK(d_fhi.alloc(n1));
    a.flo = d_flo.as<double>();
    a.fhi = d_fhi.as<double>();
  }
  if (glo || ghi) {
    HIPCHECK(d_glo.alloc(nb));
    HIPCHECK(d_ghi.alloc(nb));
    a.glo = d_glo.as<double>();
    a.ghi = d_ghi.as<double>();
  }
  if (ub) {
    HIPCHECK(d_ub.alloc(n1));
    a.ub = d_ub.as<double>();
  }
  if (kkt) {
    HIPCHECK(d_kkt.alloc(sizeof(int) * (size_t)n));
    a.kkt = d_kkt.as<int>();
  }
  a.n = n;
  a.d = d;
  a.n_samples = n_samples;
  void *args[] = {&a};
  unsigned grid = (unsigned)((n + 63) / 64);
  HIPCHECK(hipModuleLaunchKernel(m->eval[rounding], grid, 1, 1, 64, 1, 1, 0, nullptr, args, nullptr));
  HIPCHECK(hipDeviceSynchronize());
  if (flo) HIPCHECK(hipMemcpy(flo, d_flo.p, n1, hipMemcpyDeviceToHost));
  if (fhi) HIPCHECK(hipMemcpy(fhi, d_fhi.p, n1, hipMemcpyDeviceToHost));
  if (glo) HIPCHECK(hipMemcpy(glo, d_glo.p, nb, hipMemcpyDeviceToHost));
  if (ghi) HIPCHECK(hipMemcpy(ghi, d_ghi.p, nb, hipMemcpyDeviceToHost));
  if (ub) HIPCHECK(hipMemcpy(ub, d_ub.p, n1, hipMemcpyDeviceToHost));
  if (kkt) HIPCHECK(hipMemcpy(kkt, d_kkt.p, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost));
  return GPUGO_OK;
}

extern "C" int gpugo_eval_boxes_f32(gpugo_module_t m, int32_t d, const double *params, int64_t n_params, int64_t n,
                                    const double *lo, const double *hi, int32_t rounding, double *flo, double *fhi) {
  int rc = check_module_problem(m, d, params, n_params);
  if (rc) return rc;
  if (rounding != GPUGO_ROUND_EXACT && rounding != GPUGO_ROUND_FAST)
    return set_err(GPUGO_E_INVALID, "unknown rounding %d", rounding);
  if (!m->eval32[rounding]) return set_err(GPUGO_E_INVALID, "module has no FP32 code list for this rounding");
  if (n < 0) return set_err(GPUGO_E_INVALID, "negative n");
  if (n == 0) return GPUGO_OK;
  if (!lo || !hi || !flo || !fhi) return set_err(GPUGO_E_INVALID, "null pointer");
  rc = check_intervals(lo, hi, n * d, "box");
  if (rc) return rc;
  const size_t nb = sizeof(double) * (size_t)n * d, n1 = sizeof(double) * (size_t)n;
  std::vector<float> p32(2 * (n_params > 0 ? n_params : 1), 0.0f);
  for (int64_t i = 0; i < n_params; ++i) {
    const double v = params[i];
    float f = (float)v;
    p32[2 * i] = (double)f > v ? std::nextafter(f, -__builtin_huge_valf()) : f;
    p32[2 * i + 1] = (double)f < v ? std::nextafter(f, __builtin_huge_valf()) : f;
  }
  DevBuf d_lo, d_hi, d_p32, d_flo, d_fhi;
  HIPCHECK(d_lo.alloc(nb));
  HIPCHECK(d_hi.alloc(nb));
  HIPCHECK(d_p32.alloc(sizeof(float) * p32.size()));
  HIPCHECK(d_flo.alloc(n1));
  HIPCHECK(d_fhi.alloc(n1));
  HIPCHECK(hipMemcpy(d_lo.p, lo, nb, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d_hi.p, hi, nb, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(d_p32.p, p32.data(), sizeof(float) * p32.size(), hipMemcpyHostToDevice));
  EvalArgs a;
  memset(&a, 0, sizeof a);
  a.lo = d_lo.as<double>();
  a.hi = d_hi.as<double>();
  a.ub = (double *)d_p32.p;  // coefficient intervals (iv<float>)
  a.flo = d_flo.as<double>();
  a.fhi = d_fhi.as<double>();
  a.n = n;
  a.d = d;
  void *args[] = {&a};
  HIPCHECK(hipModuleLaunchKernel(m->eval32[rounding], (unsigned)((n + 63) / 64), 1, 1, 64, 1, 1, 0, nullptr, args,
                                 nullptr));
  HIPCHECK(hipDeviceSynchronize());
  HIPCHECK(hipMemcpy(flo, d_flo.p, n1, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(fhi, d_fhi.p, n1, hipMemcpyDeviceToHost));
  return GPUGO_OK;
}

#include "bnb_step.hpp"

__global__ __launch_bounds__(SEL_T) void k_select(SelectArgs a) { select_body(a); }
__global__ __launch_bounds__(256) void k_compact(CompactArgs a) { compact_body(a, blockIdx.x, gridDim.x); }
__global__ __launch_bounds__(64) void k_root(RootArgs a) { root_body(a); }

// ---------------------------------------------------------------------------
// batches: one workgroup per slot for select and compaction
// ---------------------------------------------------------------------------
struct BatchArgs {
  SlotInfo *info;
  DevState *st;
  double *q_lb;
  unsigned *q_par, *q_pos;
  double *plo, *phi, *pw;
  int *pwin;
  double *params, *dom_lo, *dom_hi;
  long long *wave_off;
  long long pcap, qcap;
  int dmax, pmax, slots;
  int grow;  // 1: a larger tier exists; slots that would outgrow this one stop with GPUGO_ST_GROW
  int cb;    // compaction workgroups per slot
  int grow_age;  // tier 0 with a tier above: slots that have run this many steps move up (0: never)
  // priority bands: [slots][band_cap] entries and [slots] states
  double *band_lb;
  unsigned *band_par, *band_pos, *band_idx;
  BandState *band;
  long long band_cap;
};

// entries of a slot's priority band
#define GPUGO_BAND_CAP 16384

__device__ __forceinline__ BandView slot_band(const BatchArgs &b, int slot) {
  BandView v;
  const long long o = (long long)slot * b.band_cap;
  v.lb = b.band_lb + o;
  v.par = b.band_par + o;
  v.pos = b.band_pos + o;
  v.idx = b.band_idx + o;
  v.st = b.band + slot;
  v.cap = b.band_cap;
  return v;
}

__global__ __launch_bounds__(SEL_T) void kb_select(BatchArgs b) {
  const int slot = blockIdx.x;
  const SlotInfo I = b.info[slot];
  if (!I.active) return;
  DevState *st = b.st + slot;
  const long long qoff = ((long long)slot * 2 + (st->compactions & 1)) * b.qcap;
  const long long poff = (long long)slot * b.pcap;
  SelectArgs sa;
  sa.band = BandView{};  // the slot's band is passed to select_body below
  sa.st = st;
  sa.q_lb = b.q_lb + qoff;
  sa.q_par = b.q_par + qoff;
  sa.q_pos = b.q_pos + qoff;
  sa.plo = b.plo + poff * b.dmax;
  sa.phi = b.phi + poff * b.dmax;
  sa.pw = b.pw + poff * b.dmax;
  sa.pwin = b.pwin + poff;
  sa.out_lo = sa.out_hi = sa.out_lb = nullptr;
  sa.out_cap = 0;
  sa.p_cap = b.pcap;
  sa.d = I.d;
  sa.wid = I.wid;
  sa.c = I.c;
  sa.k_select = I.k_select;
  sa.tol_f = I.tol_f;
  sa.tol_x = 0.0;
  sa.max_iters = I.max_iters;
  sa.pstride = b.dmax;
  sa.ext_n = 0;
  sa.pext = nullptr;
  sa.tie_xor = 0u;
  sa.nwmin = 0;
  sa.wmin = nullptr;
  sa.lazy = 0;
  sa.grow = b.grow;
  sa.q_cap = b.qcap;
  sa.W = I.W;
  sa.grow_age = b.grow_age;
  const BandView bv = slot_band(b, slot);
  select_body(sa, &bv);
}

// The slot's queue is compacted by b.cb workgroups (grid-stride over the
// queue, CMP_U entries per thread and one atomic per workgroup pass for the
// output position); the last workgroup of the slot to finish switches the
// live buffer.  Larger tiers (longer queues) use more workgroups per slot.
__global__ __launch_bounds__(256) void kb_compact(BatchArgs b) {
  __shared__ unsigned wcnt[CMP_U * 4];
  __shared__ unsigned long long s_kmin, s_kmax, s_base;
  const int cb = b.cb;
  const int slot = blockIdx.x / cb, sub = blockIdx.x % cb;
  const SlotInfo I = b.info[slot];
  if (!I.active) return;
  DevState *st = b.st + slot;
  if (st->status != GPUGO_ST_RUNNING && !st->compact_final) return;
  const int cur = (int)(st->compactions & 1);
  const long long in = ((long long)slot * 2 + cur) * b.qcap, out = ((long long)slot * 2 + (cur ^ 1)) * b.qcap;
  long long n = st->q_count;
  if (n > b.qcap) n = b.qcap;
  const double ub = key_value(st->ub_key);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  unsigned long long kmin = ~0ull, kmax = 0ull;
  if (threadIdx.x == 0) {
    s_kmin = ~0ull;
    s_kmax = 0ull;
  }
  // kept entries below the band threshold also go to the slot's band
  const BandView bv = slot_band(b, slot);
  const double bg = bv.st->g;
  const long long span = 256LL * CMP_U;
  for (long long base = (long long)sub * span; base < n; base += (long long)cb * span) {
    double lb[CMP_U];
#pragma unroll
    for (int u = 0; u < CMP_U; ++u) {
      const long long i = base + u * 256 + threadIdx.x;
      lb[u] = i < n ? b.q_lb[in + i] : 0.0;
    }
    unsigned kept = 0u, pre[CMP_U];
#pragma unroll
    for (int u = 0; u < CMP_U; ++u) {
      const long long i = base + u * 256 + threadIdx.x;
      const bool keep = i < n && lb[u] <= ub;
      if (keep) {
        const unsigned long long k = order_key(lb[u]);
        kmin = k < kmin ? k : kmin;
        kmax = k > kmax ? k : kmax;
        kept |= 1u << u;
      }
      const unsigned long long mask = __ballot(keep);
      pre[u] = __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
      if (lane == 0) wcnt[u * 4 + w] = (unsigned)__popcll(mask);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      // exclusive prefix over (u, wavefront): the output order of the pass
      unsigned t = 0;
      for (int q = 0; q < CMP_U * 4; ++q) {
        const unsigned c = wcnt[q];
        wcnt[q] = t;
        t += c;
      }
      s_base = t ? atomicAdd((unsigned long long *)&st->compact_count, (unsigned long long)t) : 0ull;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CMP_U; ++u) {
      const bool k = (kept >> u) & 1u;
      const bool inband = k && lb[u] < bg;
      const unsigned long long bmask = __ballot(inband);
      unsigned long long bbase = 0;
      if (bmask && lane == 0)
        bbase = atomicAdd((unsigned long long *)&bv.st->n, (unsigned long long)__popcll(bmask));
      bbase = (unsigned long long)__shfl((long long)bbase, 0, 64);
      if (k) {
        const long long i = base + u * 256 + threadIdx.x;
        const long long orel = (long long)s_base + wcnt[u * 4 + w] + pre[u];
        const unsigned par = b.q_par[in + i], pos = b.q_pos[in + i];
        b.q_lb[out + orel] = lb[u];
        b.q_par[out + orel] = par;
        b.q_pos[out + orel] = pos;
        if (inband) {
          const unsigned long long bo =
              bbase + __builtin_amdgcn_mbcnt_hi((unsigned)(bmask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bmask, 0u));
          if (bo < (unsigned long long)bv.cap) {
            const long long e = (long long)slot * b.band_cap + (long long)bo;
            b.band_lb[e] = lb[u];
            b.band_par[e] = par;
            b.band_pos[e] = pos;
            b.band_idx[e] = (unsigned)orel;
          }
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long t0 = __shfl_xor(kmin, o, 64), t1 = __shfl_xor(kmax, o, 64);
    kmin = t0 < kmin ? t0 : kmin;
    kmax = t1 > kmax ? t1 : kmax;
  }
  if (lane == 0 && kmin <= kmax) {
    atomicMin(&s_kmin, kmin);
    atomicMax(&s_kmax, kmax);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (s_kmin <= s_kmax) {
      atomicMin(&st->lbmin_key, s_kmin);
      atomicMax(&st->lbmax_key, s_kmax);
    }
    __threadfence();
    if (atomicAdd(&st->cblocks, 1) == cb - 1) {
      // every workgroup of this slot has read `cur` and written its part
      __threadfence();
      if (n > st->queue_peak) st->queue_peak = n;
      st->compactions += 1;
      st->cblocks = 0;
    }
  }
}

// Work list of the step: per running slot, iterations/bounded as the
// single-problem kernel counts them, and the exclusive scan of its
// wavefront count ceil(n_children / 64) into wave_off[0..slots].
__device__ __forceinline__ long long slot_waves(const BatchArgs &b, int s) {
  const SlotInfo &I = b.info[s];
  if (!I.active) return 0;
  const DevState *st = b.st + s;
  if (st->status != GPUGO_ST_RUNNING || st->n_split <= 0) return 0;
  return ((long long)st->n_split * (long long)I.W + 63) >> 6;
}

__global__ __launch_bounds__(1024) void kb_plan(BatchArgs b) {
  __shared__ long long sh[17];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int per = (b.slots + 1023) / 1024;
  const int s0 = tid * per, s1 = (s0 + per < b.slots) ? s0 + per : b.slots;
  long long local = 0;
  for (int s = s0; s < s1; ++s) {
    const long long nw = slot_waves(b, s);
    if (nw) {
      DevState *st = b.st + s;
      st->iterations += 1;
      st->bounded += (long long)st->n_split * (long long)b.info[s].W;
    }
    local += nw;
  }
  long long inc = local;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const long long t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) sh[w] = inc;
  __syncthreads();
  if (w == 0) {
    const long long x = lane < 16 ? sh[lane] : 0;
    long long xi = x;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      const long long t = __shfl_up(xi, o, 64);
      if (lane >= o) xi += t;
    }
    if (lane < 16) sh[lane] = xi - x;
    if (lane == 15) sh[16] = xi;
  }
  __syncthreads();
  long long run = sh[w] + inc - local;
  for (int s = s0; s < s1; ++s) {
    b.wave_off[s] = run;
    run += slot_waves(b, s);
  }
  if (tid == 0) b.wave_off[b.slots] = sh[16];
}

// Admission: one workgroup per record writes the slot's description,
// coefficients, initial box, root parent record and fresh state.
__global__ __launch_bounds__(64) void kb_admit(BatchArgs b, const char *recs, long long rec_bytes) {
  const AdmitHead *h = (const AdmitHead *)(recs + (long long)blockIdx.x * rec_bytes);
  const int s = h->slot;
  const double *pr = (const double *)(h + 1);
  const double *lo = pr + b.pmax, *hi = lo + b.dmax, *pw = hi + b.dmax;
  const long long root = (long long)s * b.pcap;
  for (int i = threadIdx.x; i < b.pmax; i += 64) b.params[(long long)s * b.pmax + i] = pr[i];
  for (int j = threadIdx.x; j < b.dmax; j += 64) {
    b.dom_lo[(long long)s * b.dmax + j] = lo[j];
    b.dom_hi[(long long)s * b.dmax + j] = hi[j];
    b.plo[root * b.dmax + j] = lo[j];
    b.phi[root * b.dmax + j] = hi[j];
    b.pw[root * b.dmax + j] = pw[j];
  }
  if (threadIdx.x == 0) {
    b.pwin[root] = 0;
    b.st[s] = h->st;
    b.info[s] = h->info;
    b.band[s].g = -__builtin_huge_val();
    b.band[s].n = 0;
  }
}

// Move slot s of batch `src` (stopped in GPUGO_ST_GROW) into slot t of batch
// `dst`, a tier with larger queues and arenas over the same module: the
// description, coefficients, initial box, the parent records in use, the
// live (compacted) queue buffer and the state, which resumes as running.
// Grid-stride over all copied elements; the state and description are
// written by thread 0 of workgroup 0.
__global__ __launch_bounds__(256) void kb_move(BatchArgs src, BatchArgs dst, int s, int t) {
  const DevState S = src.st[s];
  const long long n_rec = S.parents_used, n_q = S.compact_count;
  const int D = src.dmax;
  const int buf = (int)(S.compactions & 1);
  const long long qs = ((long long)s * 2 + buf) * src.qcap, qt = ((long long)t * 2 + buf) * dst.qcap;
  const long long rs = (long long)s * src.pcap, rt = (long long)t * dst.pcap;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (long long i = i0; i < n_rec * D; i += stride) {
    dst.plo[rt * D + i] = src.plo[rs * D + i];
    dst.phi[rt * D + i] = src.phi[rs * D + i];
    dst.pw[rt * D + i] = src.pw[rs * D + i];
  }
  for (long long i = i0; i < n_rec; i += stride) dst.pwin[rt + i] = src.pwin[rs + i];
  for (long long i = i0; i < n_q; i += stride) {
    dst.q_lb[qt + i] = src.q_lb[qs + i];
    dst.q_par[qt + i] = src.q_par[qs + i];
    dst.q_pos[qt + i] = src.q_pos[qs + i];
  }
  for (long long i = i0; i < src.pmax; i += stride)
    dst.params[(long long)t * dst.pmax + i] = src.params[(long long)s * src.pmax + i];
  for (long long i = i0; i < D; i += stride) {
    dst.dom_lo[(long long)t * D + i] = src.dom_lo[(long long)s * D + i];
    dst.dom_hi[(long long)t * D + i] = src.dom_hi[(long long)s * D + i];
  }
  if (i0 == 0) {
    DevState z = S;
    z.status = GPUGO_ST_RUNNING;
    z.overflow = 0;
    z.compact_final = 0;
    z.cblocks = 0;
    z.fresh = 0;
    dst.st[t] = z;
    dst.info[t] = src.info[s];
    dst.band[t].g = -__builtin_huge_val();
    dst.band[t].n = 0;
  }
}

// ---------------------------------------------------------------------------
// smallest live lower bound (used when a run stops without a certificate)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_minlb(const DevState *st, const double *lb, long long n,
                                                unsigned long long *out) {
  __shared__ unsigned long long s_min;
  if (threadIdx.x == 0) s_min = order_key(__builtin_huge_val());
  __syncthreads();
  unsigned long long m = order_key(__builtin_huge_val());
  for (long long i = threadIdx.x; i < n; i += blockDim.x) {
    double v = lb[i];
    if (v == v) {
      unsigned long long k = order_key(v);
      if (k < m) m = k;
    }
  }
  atomicMin(&s_min, m);
  __syncthreads();
  if (threadIdx.x == 0) *out = s_min;
}

// materialise queue entries as boxes (gpugo_boxes_copy)
__global__ __launch_bounds__(256) void k_materialize(const double *plo, const double *phi, const double *pw,
                                                     const int *pwin, const unsigned *q_par, const unsigned *q_pos,
                                                     const long long *order, long long n, int d, int wid, int c,
                                                     double *lo, double *hi) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  long long e = order[i];
  unsigned par = q_par[e], pos = q_pos[e];
  decode_box(plo + (long long)par * d, phi + (long long)par * d, pw + (long long)par * d, pwin[par], d, wid, c, pos,
             lo + i * d, hi + i * d);
}

// ---------------------------------------------------------------------------
// problems
// ---------------------------------------------------------------------------
// Workgroups of the two compaction launches; each leaves its smallest key
// in p->wmin for a one-box select.
#define GPUGO_COMPACT_WG 512   // gpugo_compact, 256 threads
#ifndef GPUGO_FUSED_WG
#define GPUGO_FUSED_WG 128     // gpugo_compact_select, 1024 threads
#endif
#define GPUGO_LAZY_FULL_WG 512 // the same as the full compaction of a lazy (one box per step) queue:
                               // two workgroups per CU for the queues that grow large
#define GPUGO_LAZY_WG 256      // gpugo_compact_select as the lazy tail, 1024 threads
#define GPUGO_WMIN_MAX GPUGO_COMPACT_WG
#define GPUGO_SBAND_CAP 16384  // priority band of a single solve with k_select > 1 (entries)
#define GPUGO_POLL_STATE_OFF 64  // bytes from the status words to the pinned state copies
#define GPUGO_CHUNK_LOG2 13    // lazy queue: 8192 entries per chunk minimum
#define GPUGO_LAZY_MAX_EVERY 256
#define GPUGO_TAB_MAX (1 << 20) // term-table entries per parent record, at most
#define GPUGO_POLL_PENDING (-1) // polled status word before the device writes it
static_assert(GPUGO_LAZY_WG <= GPUGO_WMIN_MAX && GPUGO_LAZY_FULL_WG <= GPUGO_WMIN_MAX,
              "wmin holds one key per compaction or lazy-tail workgroup");

struct Queue {
  DevBuf lb, par, pos;
  double *plb() const { return lb.as<double>(); }
  unsigned *ppar() const { return par.as<unsigned>(); }
  unsigned *ppos() const { return pos.as<unsigned>(); }
};

struct gpugo_problem_s {
  gpugo_module_t m = nullptr;
  gpugo_problem_desc desc;
  std::vector<double> lo, hi, params;
  int c = 0;
  unsigned W = 0;
  long long q_cap = 0, p_cap = 0, out_cap = 0;
  long long q_req = 0;  // queue capacity asked for (q_cap adds a lazy queue's stale budget)
  DevBuf d_lo, d_hi, d_params, d_params32, plo, phi, pw, pwin, out_lo, out_hi, out_lb, st, minkey;
  DevBuf bar;          // persistent step loop: grid barrier [2] and error flag
  DevBuf pext;         // windowed modules: rest sums per parent record [p_cap][ext_n]
  DevBuf ptab, ptab32; // windowed modules: term tables of a step's records [k_select][tab_n] (FP64 / FP32)
  int tab_n = 0, tab_n32 = 0;
  DevBuf rpart;        // windowed modules at large d: rest-sum partials [k_select][pairs][rsplit]
  int rsplit = 1;
  DevBuf wmin;         // k_select 1: per-compaction-workgroup smallest key [GPUGO_WMIN_MAX][3]
  DevBuf cmin;         // lazy queue: smallest key per chunk of 2^GPUGO_CHUNK_LOG2 entries, [3] each
  // k_select > 1 over an eagerly compacted queue: the priority band (entries
  // below a threshold, rebuilt by each compaction) the select reads instead
  // of the whole queue; band_cap 0: none
  DevBuf band_lb, band_par, band_pos, band_idx, band_st;
  long long band_cap = 0;
  // One box per step: the queue is compacted every compact_every steps; the
  // steps between leave consumed and ruled-out entries in place and refresh
  // chunk minima instead (lazy tails).
  bool lazy = false;
  int compact_every = 1;
  int since_compact = 0;
  long long lazy_steps = 0;  // lazy tails enqueued (they keep the queue in its buffer)
  int ext_n = 0;
  int persist_blocks = 0;
  Queue q[2];
  hipStream_t stream = nullptr;
  bool started = false;
  bool selected_ahead = false;  // the last compaction also ran the next step's select
  long long steps_enqueued = 0;
  // Upper bound of the queue entries (stale ones included) before the next
  // step's appends: each step appends at most (boxes split) * W.  Sizes the
  // compaction grids, so a small queue does not pay for 512 workgroups.
  long long q_hi = 0;
  int last_nwg = GPUGO_COMPACT_WG;  // workgroups of the last compaction (its wmin entries)
  DevState hs;  // last state read back
  // state polling: pinned copies of DevState and an event per slot
  // [2] status words, then at byte GPUGO_POLL_STATE_OFF [2] copies of the
  // final state; pinned, written by the compaction of a batch's last step
  int *hpoll = nullptr;
  int *dpoll = nullptr;   // its device address
  hipEvent_t poll_ev[2] = {nullptr, nullptr};
  // optional timing of the split+bound launches
  bool timing = false;
  std::vector<hipEvent_t> tev;   // pairs (start, stop) recorded since the last read
  size_t tev_used = 0;
  double bound_ms = 0;
  long long bound_launches = 0;
};

static int timing_collect(gpugo_problem_t p) {
  for (size_t i = 0; i + 1 < p->tev_used; i += 2) {
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, p->tev[i], p->tev[i + 1]));
    p->bound_ms += ms;
  }
  p->tev_used = 0;
  return GPUGO_OK;
}

static int ipow_checked(int wid, int c, unsigned *out) {
  unsigned long long w = 1;
  for (int k = 0; k < c; ++k) {
    w *= (unsigned long long)wid;
    if (w > 0xffffffffULL) return 0;
  }
  *out = (unsigned)w;
  return 1;
}

// Shape of a problem's solve state: split window and children per box, queue
// and parent-arena capacities (desc's, else the defaults given), output
// arena, per-record payload of windowed modules and the lazy-queue schedule.
static int problem_shape(gpugo_problem_s *p, const gpugo_problem_desc *desc, long long q_default,
                         long long p_default) {
  const int d = desc->d;
  p->c = desc->cycle < d ? desc->cycle : d;
  if (!ipow_checked(desc->wid, p->c, &p->W)) return set_err(GPUGO_E_INVALID, "wid^cycle must be < 2^32");
  long long qc = desc->max_queue > 0 ? desc->max_queue : q_default;
  if (qc > 0xfffffff0LL) qc = 0xfffffff0LL;
  // windowed modules: (1 + n_samples) * nrest rest-sum intervals per record
  p->ext_n = p->m->nrest > 0 ? 2 * p->m->nrest * (1 + desc->n_samples) : 0;
  p->q_cap = qc;
  p->q_req = qc;
  // term tables of a windowed module (the module's tab_layout): per point
  // (the record, then its n_samples diagonal points) NT entries, slot s
  // (c + 2(w_s - 1)) terms x wid^w_s digit combinations; then the gradient
  // table or the gradient parts table.  Left out when large (wide grids) or
  // switched off.
  // rest sums over more than 4096 terms: chunks of about 2048 terms per
  // workgroup, folded in chunk order by the prep fold kernel
  p->rsplit = (p->m->nrest > 0 && d + 2 > 4096) ? (d + 2 + 2047) / 2048 : 1;
  p->tab_n = p->tab_n32 = 0;
  if (p->m->tab && !(desc->flags & GPUGO_FLAG_NO_TABLES)) {
    auto ipow = [](long long b, int e) {
      long long r = 1;
      for (int i = 0; i < e; ++i) r *= b;
      return r;
    };
    long long nt = 0;
    for (int s2 = 0; s2 < p->m->nrest; ++s2)
      nt += (long long)(p->c + 2 * (p->m->tab_w[s2] - 1)) * ipow(desc->wid, p->m->tab_w[s2]);
    long long tot = (1LL + desc->n_samples) * nt;
    const long long gtab = (long long)(p->c + 2 * (p->m->tab_gw - 1)) * ipow(desc->wid, p->m->tab_gw);
    if (p->m->tab & 2) tot += gtab;
    // parts table: per gradient part, d entries then the window components'
    tot += (long long)p->m->tab_parts * (desc->d + gtab);
    if (tot <= GPUGO_TAB_MAX) {
      p->tab_n = (int)tot;
      p->tab_n32 = desc->precision == GPUGO_PREC_F32 ? (int)((1LL + desc->n_samples) * nt) : 0;
    }
  }
  p->p_cap = desc->max_parents > 0 ? desc->max_parents : p_default;
  // output boxes (x-width rule) up to 2^16 or 2^24 coordinates
  p->out_cap = std::max(64LL, std::min(1LL << 16, (1LL << 24) / d));
  if (p->desc.max_iters <= 0) p->desc.max_iters = 1LL << 40;
  p->lazy = false;
  p->compact_every = 1;
  if (desc->k_select == 1 && !(desc->flags & (GPUGO_FLAG_PERSIST | GPUGO_FLAG_EAGER_COMPACT))) {
    // A compaction leaves the live entries, which an eagerly compacted
    // queue holds at the same step (<= qc, or that queue had overflowed);
    // the compact_every steps up to the next one append at most W each.
    // Allocating that much beyond qc means a lazy queue never overflows
    // where an eager one would not (the eighth of qc bounds the extra).
    const long long every = std::min((long long)GPUGO_LAZY_MAX_EVERY, qc / 8 / (long long)p->W);
    if (every >= 2 && qc + every * (long long)p->W <= 0xfffffff0LL) {
      p->lazy = true;
      p->compact_every = (int)every;
      p->q_cap = qc + every * (long long)p->W;
    }
  }
  p->band_cap = desc->k_select > 1 && !(desc->flags & GPUGO_FLAG_NO_BAND) ? GPUGO_SBAND_CAP : 0;
  return GPUGO_OK;
}

// The device buffers of a problem, in one order for both owners:
// alloc(buf, bytes, name) either allocates (gpugo_problem_create) or places
// the buffer in the static arena (gpugo_solve_static).
template <class Alloc>
static int problem_buffers(gpugo_problem_s *p, Alloc &&alloc) {
  const long long d = p->desc.d, np = p->desc.n_params > 0 ? p->desc.n_params : 1;
  const long long pc = p->p_cap, qc = p->q_cap;
  int rc;
#define PBUF(buf, bytes) \
  if ((rc = alloc(buf, (size_t)(bytes), #buf)) != GPUGO_OK) return rc
  PBUF(p->st, sizeof(DevState));
  PBUF(p->d_lo, sizeof(double) * d);
  PBUF(p->d_hi, sizeof(double) * d);
  PBUF(p->d_params, sizeof(double) * np);
  PBUF(p->d_params32, 2 * sizeof(float) * np);
  PBUF(p->plo, sizeof(double) * d * pc);
  PBUF(p->phi, sizeof(double) * d * pc);
  PBUF(p->pw, sizeof(double) * d * pc);
  PBUF(p->pwin, sizeof(int) * pc);
  PBUF(p->out_lo, sizeof(double) * d * p->out_cap);
  PBUF(p->out_hi, sizeof(double) * d * p->out_cap);
  PBUF(p->out_lb, sizeof(double) * p->out_cap);
  PBUF(p->minkey, 64);
  PBUF(p->bar, 64);
  if (p->ext_n > 0) PBUF(p->pext, sizeof(double) * p->ext_n * pc);
  if (p->tab_n > 0) PBUF(p->ptab, 16 * (size_t)p->tab_n * p->desc.k_select);
  if (p->tab_n32 > 0) PBUF(p->ptab32, 8 * (size_t)p->tab_n32 * p->desc.k_select);
  if (p->rsplit > 1)
    PBUF(p->rpart, 16 * (size_t)p->rsplit * (size_t)(1 + p->desc.n_samples) * p->m->nrest * p->desc.k_select);
  PBUF(p->wmin, sizeof(unsigned long long) * 3 * GPUGO_WMIN_MAX);
  if (p->lazy) {
    const long long nch = (qc + (1LL << GPUGO_CHUNK_LOG2) - 1) >> GPUGO_CHUNK_LOG2;
    PBUF(p->cmin, sizeof(unsigned long long) * 3 * nch);
  }
  if (p->band_cap > 0) {
    PBUF(p->band_lb, sizeof(double) * p->band_cap);
    PBUF(p->band_par, sizeof(unsigned) * p->band_cap);
    PBUF(p->band_pos, sizeof(unsigned) * p->band_cap);
    PBUF(p->band_idx, sizeof(unsigned) * p->band_cap);
    PBUF(p->band_st, sizeof(BandState));
  }
  for (int b = 0; b < 2; ++b) {
    PBUF(p->q[b].lb, sizeof(double) * qc);
    PBUF(p->q[b].par, sizeof(unsigned) * qc);
    PBUF(p->q[b].pos, sizeof(unsigned) * qc);
  }
#undef PBUF
  return GPUGO_OK;
}

// Every check of a problem description against its module (shared by
// gpugo_problem_create and gpugo_solve_static).
static int problem_check(gpugo_module_t m, const gpugo_problem_desc *desc) {
  const int d = desc->d;
  if (d < 1) return set_err(GPUGO_E_INVALID, "d must be >= 1");
  if (d > GPUGO_DMAX_SELECT) return set_err(GPUGO_E_INVALID, "d must be <= %d", GPUGO_DMAX_SELECT);
  int rc = check_module_problem(m, d, desc->params, desc->n_params);
  if (rc) return rc;
  if (!desc->lo || !desc->hi) return set_err(GPUGO_E_INVALID, "null box");
  rc = check_intervals(desc->lo, desc->hi, d, "box");
  if (rc) return rc;
  for (int j = 0; j < d; ++j)
    if (!std::isfinite(desc->lo[j]) || !std::isfinite(desc->hi[j]))
      return set_err(GPUGO_E_DOMAIN, "box coordinate %d is not finite", j);
  if (desc->wid < 2 || desc->wid > 64) return set_err(GPUGO_E_INVALID, "wid must be in [2,64]");
  if (desc->cycle < 1) return set_err(GPUGO_E_INVALID, "cycle must be >= 1");
  if (desc->k_select < 1 || desc->k_select > SEL_KMAX)
    return set_err(GPUGO_E_INVALID, "k_select must be in [1,%d]", SEL_KMAX);
  if (desc->n_samples < 1 || desc->n_samples > 64) return set_err(GPUGO_E_INVALID, "n_samples must be in [1,64]");
  if (desc->precision != GPUGO_PREC_F64 && desc->precision != GPUGO_PREC_F32)
    return set_err(GPUGO_E_INVALID, "unknown precision %d", desc->precision);
  if (desc->precision == GPUGO_PREC_F32 && !m->f32)
    return set_err(GPUGO_E_INVALID, "module has no FP32 code list (emit it with f32=True)");
  if (desc->flags & ~(GPUGO_FLAG_MEAN_VALUE | GPUGO_FLAG_SPLIT | GPUGO_FLAG_PERSIST | GPUGO_FLAG_EAGER_COMPACT |
                      GPUGO_FLAG_NO_TABLES | GPUGO_FLAG_NO_BAND))
    return set_err(GPUGO_E_INVALID, "unknown flags 0x%x", desc->flags);
  if ((desc->flags & GPUGO_FLAG_SPLIT) && m->split_g < 2)
    return set_err(GPUGO_E_INVALID, "module has no term-parallel kernels (emit it with split=G)");
  if ((desc->flags & GPUGO_FLAG_PERSIST) && !(desc->flags & GPUGO_FLAG_SPLIT))
    return set_err(GPUGO_E_INVALID, "GPUGO_FLAG_PERSIST needs GPUGO_FLAG_SPLIT");
  if (desc->rounding == GPUGO_ROUND_DIGITS15 &&
      (desc->flags & (GPUGO_FLAG_SPLIT | GPUGO_FLAG_PERSIST) || desc->precision != GPUGO_PREC_F64))
    return set_err(GPUGO_E_INVALID, "15-digit rounding has no split, persistent or FP32 form");
  if ((desc->flags & GPUGO_FLAG_PERSIST) && !m->persist[desc->rounding == GPUGO_ROUND_FAST ? 1 : 0])
    return set_err(GPUGO_E_INVALID, "module has no persistent step loop for this rounding");
  if ((desc->flags & GPUGO_FLAG_SPLIT) && desc->precision == GPUGO_PREC_F32)
    return set_err(GPUGO_E_INVALID, "the term-parallel kernels have no FP32 pruning pass");
  if (desc->rounding != GPUGO_ROUND_EXACT && desc->rounding != GPUGO_ROUND_FAST &&
      desc->rounding != GPUGO_ROUND_DIGITS15)
    return set_err(GPUGO_E_INVALID, "unknown rounding %d", desc->rounding);
  if (!m->bound[desc->rounding])
    return set_err(GPUGO_E_INVALID, "module has no variant for rounding %d", desc->rounding);
  if (m->nrest > 0) {
    // windowed (large-d) module: natural-extension bound, diagonal samples
    if (desc->flags & (GPUGO_FLAG_MEAN_VALUE | GPUGO_FLAG_SPLIT | GPUGO_FLAG_PERSIST))
      return set_err(GPUGO_E_INVALID, "windowed modules take no mean-value, split or persistent flags");
  }
  if (std::isnan(desc->tol_f) || std::isnan(desc->tol_x) || std::isnan(desc->init_ub))
    return set_err(GPUGO_E_DOMAIN, "NaN tolerance or init_ub");
  if (!(desc->tol_f > 0) && !(desc->tol_x > 0)) return set_err(GPUGO_E_INVALID, "set tol_f or tol_x");
  return GPUGO_OK;
}

extern "C" int gpugo_problem_create(gpugo_module_t m, const gpugo_problem_desc *desc, gpugo_problem_t *out) {
  if (!desc || !out) return set_err(GPUGO_E_INVALID, "null pointer");
  *out = nullptr;
  int rc = problem_check(m, desc);
  if (rc) return rc;
  const int d = desc->d;
  auto *p = new gpugo_problem_s;
  p->m = m;
  p->desc = *desc;
  p->lo.assign(desc->lo, desc->lo + d);
  p->hi.assign(desc->hi, desc->hi + d);
  p->params.assign(desc->params, desc->params + desc->n_params);
  p->desc.lo = p->lo.data();
  p->desc.hi = p->hi.data();
  p->desc.params = p->params.data();
  auto fail = [&](int code) {
    delete p;
    return code;
  };
  // default arenas: parent records up to 2^20 or 8 GiB
  const long long rec_bytes = 24LL * d + 4 + 8LL * (m->nrest > 0 ? 2 * m->nrest * (1 + desc->n_samples) : 0);
  rc = problem_shape(p, desc, 1LL << 24, std::min(1LL << 20, (8LL << 30) / rec_bytes));
  if (rc) return fail(rc);
  hipError_t e;
  rc = problem_buffers(p, [](DevBuf &buf, size_t bytes, const char *name) {
    const hipError_t err = buf.alloc(bytes);
    if (err != hipSuccess)
      return set_err(err == hipErrorOutOfMemory ? GPUGO_E_NOMEM : GPUGO_E_HIP, "alloc %s: %s", name,
                     hipGetErrorString(err));
    return (int)GPUGO_OK;
  });
  if (rc) return fail(rc);
  if ((e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking)) != hipSuccess)
    return fail(set_err(GPUGO_E_HIP, "stream: %s", hipGetErrorString(e)));
  if ((e = hipMemcpy(p->d_lo.p, p->lo.data(), sizeof(double) * d, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(p->d_hi.p, p->hi.data(), sizeof(double) * d, hipMemcpyHostToDevice)) != hipSuccess ||
      (desc->n_params > 0 &&
       (e = hipMemcpy(p->d_params.p, p->params.data(), sizeof(double) * desc->n_params, hipMemcpyHostToDevice)) !=
           hipSuccess))
    return fail(set_err(GPUGO_E_HIP, "upload: %s", hipGetErrorString(e)));
  {
    // coefficients as FP32 intervals rounded outward (FP32 pruning pass)
    std::vector<float> p32(2 * (desc->n_params > 0 ? desc->n_params : 1), 0.0f);
    for (int64_t i = 0; i < desc->n_params; ++i) {
      const double v = desc->params[i];
      float f = (float)v;
      p32[2 * i] = (double)f > v ? std::nextafter(f, -__builtin_huge_valf()) : f;
      p32[2 * i + 1] = (double)f < v ? std::nextafter(f, __builtin_huge_valf()) : f;
    }
    if ((e = hipMemcpy(p->d_params32.p, p32.data(), sizeof(float) * p32.size(), hipMemcpyHostToDevice)) !=
        hipSuccess)
      return fail(set_err(GPUGO_E_HIP, "upload: %s", hipGetErrorString(e)));
  }
  // coherent (fine-grained): the compaction's system-scope store of the
  // status reaches host memory without depending on HIP_HOST_COHERENT
  if ((e = hipHostMalloc((void **)&p->hpoll, GPUGO_POLL_STATE_OFF + 2 * sizeof(DevState),
                         hipHostMallocMapped | hipHostMallocCoherent)) != hipSuccess)
    return fail(set_err(GPUGO_E_NOMEM, "pinned poll buffer: %s", hipGetErrorString(e)));
  if ((e = hipHostGetDevicePointer((void **)&p->dpoll, p->hpoll, 0)) != hipSuccess)
    return fail(set_err(GPUGO_E_HIP, "poll buffer device address: %s", hipGetErrorString(e)));
  for (int b = 0; b < 2; ++b)
    if ((e = hipEventCreateWithFlags(&p->poll_ev[b], hipEventDisableTiming)) != hipSuccess)
      return fail(set_err(GPUGO_E_HIP, "event: %s", hipGetErrorString(e)));
  memset(&p->hs, 0, sizeof p->hs);
  if (desc->flags & GPUGO_FLAG_PERSIST) {
    // every workgroup of the cooperative launch must be resident
    hipFunction_t fn = (desc->flags & GPUGO_FLAG_MEAN_VALUE) ? m->persist_mv[desc->rounding] : m->persist[desc->rounding];
    int occ = 0, dev = 0, cus = 0;
    if ((e = hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, 1024, 0)) != hipSuccess ||
        (e = hipGetDevice(&dev)) != hipSuccess ||
        (e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
      return fail(set_err(GPUGO_E_HIP, "occupancy: %s", hipGetErrorString(e)));
    if (occ < 1) return fail(set_err(GPUGO_E_HIP, "persistent step loop cannot be resident"));
    p->persist_blocks = occ * cus;
  }
  *out = p;
  return GPUGO_OK;
}

extern "C" void gpugo_problem_destroy(gpugo_problem_t p) {
  if (!p) return;
  if (p->stream) {
    (void)hipStreamSynchronize(p->stream);
    (void)hipStreamDestroy(p->stream);
  }
  for (int b = 0; b < 2; ++b)
    if (p->poll_ev[b]) (void)hipEventDestroy(p->poll_ev[b]);
  for (auto ev : p->tev) (void)hipEventDestroy(ev);
  if (p->hpoll) (void)hipHostFree(p->hpoll);
  delete p;
}

extern "C" int gpugo_problem_reset(gpugo_problem_t p) {
  if (!p) return set_err(GPUGO_E_INVALID, "null problem");
  HIPCHECK(hipStreamSynchronize(p->stream));
  p->started = false;
  p->selected_ahead = false;
  p->steps_enqueued = 0;
  p->since_compact = 0;
  p->lazy_steps = 0;
  p->tev_used = 0;
  p->bound_ms = 0;
  p->bound_launches = 0;
  memset(&p->hs, 0, sizeof p->hs);
  return GPUGO_OK;
}

extern "C" int gpugo_problem_timing(gpugo_problem_t p, int enable) {
  if (!p) return set_err(GPUGO_E_INVALID, "null problem");
  HIPCHECK(hipStreamSynchronize(p->stream));
  int rc = timing_collect(p);
  if (rc) return rc;
  p->timing = enable != 0;
  return GPUGO_OK;
}

// The root state is written on the device by one kernel (the initial box is
// already resident), so a solve starts without a host round trip; the host
// keeps the same initial state for its own bookkeeping.
static int root_init(gpugo_problem_t p) {
  DevState s;
  memset(&s, 0, sizeof s);
  s.ub_key = order_key(p->desc.init_ub);
  s.ub_snap = p->desc.init_ub;
  s.lb_global = -__builtin_huge_val();
  s.out_min_lb = __builtin_huge_val();
  s.width = __builtin_huge_val();
  s.parents_used = 1;
  s.n_split = 1;
  s.status = GPUGO_ST_RUNNING;
  s.lbmin_key = ~0ull;
  s.lbmax_key = 0ull;
  p->hs = s;
  p->q_hi = 0;
  p->last_nwg = GPUGO_COMPACT_WG;
  RootArgs ra;
  ra.st = p->st.as<DevState>();
  ra.plo = p->plo.as<double>();
  ra.phi = p->phi.as<double>();
  ra.pw = p->pw.as<double>();
  ra.pwin = p->pwin.as<int>();
  ra.lo = p->d_lo.as<double>();
  ra.hi = p->d_hi.as<double>();
  ra.init_ub = p->desc.init_ub;
  ra.d = p->desc.d;
  ra.wid = p->desc.wid;
  ra.band = p->band_cap > 0 ? p->band_st.as<BandState>() : nullptr;
  if (p->m->root_fn) {
    void *args[] = {&ra};
    HIPCHECK(hipModuleLaunchKernel(p->m->root_fn, 1, 1, 1, 64, 1, 1, 0, p->stream, args, nullptr));
  } else {
    hipLaunchKernelGGL(k_root, dim3(1), dim3(64), 0, p->stream, ra);
    HIPCHECK(hipGetLastError());
  }
  return GPUGO_OK;
}

// Pinned copy k of a stopped problem's final state (host and device address).
static DevState *poll_state_h(gpugo_problem_t p, int k) {
  return reinterpret_cast<DevState *>(reinterpret_cast<char *>(p->hpoll) + GPUGO_POLL_STATE_OFF) + k;
}
static DevState *poll_state_d(gpugo_problem_t p, int k) {
  return reinterpret_cast<DevState *>(reinterpret_cast<char *>(p->dpoll) + GPUGO_POLL_STATE_OFF) + k;
}

// The problem's priority band (st null when it keeps none).
static BandView band_view(gpugo_problem_t p) {
  BandView v{};
  if (p->band_cap > 0) {
    v.lb = p->band_lb.as<double>();
    v.par = p->band_par.as<unsigned>();
    v.pos = p->band_pos.as<unsigned>();
    v.idx = p->band_idx.as<unsigned>();
    v.st = p->band_st.as<BandState>();
    v.cap = p->band_cap;
  }
  return v;
}

// Select of the step that reads queue buffer q (the K smallest boxes,
// stop tests, new parent records); nwg: workgroups of the compaction
// that wrote q.
static SelectArgs select_args(gpugo_problem_t p, const Queue &q, int nwg) {
  SelectArgs sa;
  memset(&sa, 0, sizeof sa);
  sa.st = p->st.as<DevState>();
  sa.q_lb = q.plb();
  sa.q_par = q.ppar();
  sa.q_pos = q.ppos();
  sa.plo = p->plo.as<double>();
  sa.phi = p->phi.as<double>();
  sa.pw = p->pw.as<double>();
  sa.pwin = p->pwin.as<int>();
  sa.out_lo = p->out_lo.as<double>();
  sa.out_hi = p->out_hi.as<double>();
  sa.out_lb = p->out_lb.as<double>();
  sa.out_cap = p->out_cap;
  sa.p_cap = p->p_cap;
  sa.d = p->desc.d;
  sa.wid = p->desc.wid;
  sa.c = p->c;
  sa.k_select = p->desc.k_select;
  sa.tol_f = p->desc.tol_f;
  sa.tol_x = p->desc.tol_x;
  sa.max_iters = 0;
  sa.pstride = p->desc.d;
  sa.ext_n = p->ext_n;
  sa.pext = p->pext.as<double>();
  sa.tie_xor = p->m->nrest > 0 ? ~0u : 0u;  // windowed (large-d): newest parent first among ties
  // one box per step: the previous compaction's per-workgroup smallest keys
  sa.wmin = p->desc.k_select == 1 ? p->wmin.as<unsigned long long>() : nullptr;
  sa.nwmin = sa.wmin ? nwg : 0;
  sa.lazy = p->lazy ? 1 : 0;
  sa.grow = 0;
  sa.q_cap = p->q_cap;
  sa.W = p->W;
  sa.grow_age = 0;
  sa.band = band_view(p);
  return sa;
}

// One step: select (unless the previous compaction ran it), the windowed
// modules' rest sums, split+bound, compact (or, for a lazy queue, refresh
// the chunk minima).  poll: pinned status word written after the step (or
// null).  more: another step follows in this call, so the compaction may
// run its select.  A call's last step always compacts.  compact_only: the
// problem has stopped and only the stop step's final compaction is left
// (the select, rest sums and split+bound of a stopped problem do nothing).
static int enqueue_step(gpugo_problem_t p, int *poll, bool more, bool compact_only = false) {
  const int d = p->desc.d;
  // where a fused compaction + select that stops the problem copies its state
  DevState *pstate = poll ? poll_state_d(p, (int)(poll - p->dpoll)) : nullptr;
  // every compaction moves the queue to the other buffer, a lazy tail does not
  const int cur = (int)((p->steps_enqueued - p->lazy_steps) & 1);
  Queue &A = p->q[cur], &B = p->q[cur ^ 1];
  DevState *st = p->st.as<DevState>();
  if (!compact_only && p->steps_enqueued > 0 && !p->selected_ahead) {
    SelectArgs sa = select_args(p, A, p->last_nwg);  // after a plain compaction
    if (p->m->select_fn) {
      void *sargs[] = {&sa};
      HIPCHECK(hipModuleLaunchKernel(p->m->select_fn, 1, 1, 1, SEL_T, 1, 1, 0, p->stream, sargs, nullptr));
    } else {
      hipLaunchKernelGGL(k_select, dim3(1), dim3(SEL_T), 0, p->stream, sa);
      HIPCHECK(hipGetLastError());
    }
  }
  p->selected_ahead = false;
  BoundArgs ba;
  memset(&ba, 0, sizeof ba);
  ba.plo = p->plo.as<double>();
  ba.phi = p->phi.as<double>();
  ba.pw = p->pw.as<double>();
  ba.pwin = p->pwin.as<int>();
  ba.params = p->d_params.as<double>();
  ba.dom_lo = p->d_lo.as<double>();
  ba.dom_hi = p->d_hi.as<double>();
  ba.q_lb = A.plb();
  ba.q_par = A.ppar();
  ba.q_pos = A.ppos();
  ba.st = st;
  ba.q_cap = p->q_cap;
  ba.d = d;
  ba.wid = p->desc.wid;
  ba.c = p->c;
  ba.n_samples = p->desc.n_samples;
  ba.W = p->W;
  ba.inv_wid = 1.0 / (double)p->desc.wid;
  ba.half_inv_wid = 0.5 / (double)p->desc.wid;
  ba.flags = p->desc.flags | (p->desc.precision == GPUGO_PREC_F32 ? GPUGO_FLAG_F32_PRUNE : 0);
  ba.params32 = p->d_params32.p;
  ba.prest = p->pext.as<double>();
  ba.nrest = p->m->nrest;
  ba.ptab = p->ptab.p;
  ba.ptab32 = p->ptab32.p;
  ba.tab_n = p->tab_n;
  ba.tab_n32 = p->tab_n32;
  ba.rpart = p->rpart.p;
  ba.rsplit = p->rsplit;
  if (p->m->nrest > 0 && !compact_only) {
    // rest sums of the parent records this step splits: one workgroup per
    // (point, template sum) pair and parent, then the workgroups that fill
    // each record's term tables (256 entries each)
    const unsigned nb = p->steps_enqueued == 0 ? 1u : (unsigned)p->desc.k_select;
    const unsigned npairs = (unsigned)((1 + p->desc.n_samples) * p->m->nrest);
    const unsigned ntab = (unsigned)((p->tab_n + 255) / 256);
    void *args[] = {&ba};
    HIPCHECK(hipModuleLaunchKernel(p->m->prep[p->desc.rounding], npairs * (unsigned)p->rsplit + ntab, nb, 1, 256, 1,
                                   1, 0, p->stream, args, nullptr));
    if (p->rsplit > 1)
      HIPCHECK(hipModuleLaunchKernel(p->m->prep_fold[p->desc.rounding], (npairs + 63) / 64, nb, 1, 64, 1, 1, 0,
                                     p->stream, args, nullptr));
  }
  if (!compact_only && p->steps_enqueued == 0 && p->desc.precision == GPUGO_PREC_F32 && p->desc.rounding < 2 &&
      p->m->sample32[p->desc.rounding]) {
    // first step of an FP32-prune solve: sample every child in FP32 first,
    // so the FP32 value test has an incumbent in this step too
    long long blocks = ((long long)p->W + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    void *args[] = {&ba};
    HIPCHECK(hipModuleLaunchKernel(p->m->sample32[p->desc.rounding], (unsigned)blocks, 1, 1, 256, 1, 1, 0, p->stream,
                                   args, nullptr));
  }
  if (!compact_only) {
    const bool split = (p->desc.flags & GPUGO_FLAG_SPLIT) != 0;
    long long maxk = p->steps_enqueued == 0 ? 1 : p->desc.k_select;
    long long nthreads = maxk * (long long)p->W * (split ? p->m->split_g : 1);
    long long blocks = (nthreads + 255) / 256;
    if (blocks > 16384) blocks = 16384;
    void *args[] = {&ba};
    const bool mv = (p->desc.flags & GPUGO_FLAG_MEAN_VALUE) != 0;
    hipFunction_t fn = split ? (mv ? p->m->bound_mv_g[p->desc.rounding] : p->m->bound_g[p->desc.rounding])
                             : (mv ? p->m->bound_mv[p->desc.rounding] : p->m->bound[p->desc.rounding]);
    if (p->timing) {
      // events taken from the kernel's own dispatch (start and end of its
      // execution), not recorded around it: the launch gap is excluded
      while (p->tev.size() < p->tev_used + 2) {
        hipEvent_t ev;
        HIPCHECK(hipEventCreate(&ev));
        p->tev.push_back(ev);
      }
      hipEvent_t e0 = p->tev[p->tev_used++], e1 = p->tev[p->tev_used++];
      HIPCHECK(hipExtModuleLaunchKernel(fn, (uint32_t)(blocks * 256), 1, 1, 256, 1, 1, 0, p->stream, args, nullptr,
                                        e0, e1, 0));
    } else {
      HIPCHECK(hipModuleLaunchKernel(fn, (unsigned)blocks, 1, 1, 256, 1, 1, 0, p->stream, args, nullptr));
    }
  }
  CompactArgs ca;
  ca.st = st;
  ca.lb_in = A.plb();
  ca.par_in = A.ppar();
  ca.pos_in = A.ppos();
  ca.lb_out = B.plb();
  ca.par_out = B.ppar();
  ca.pos_out = B.ppos();
  ca.cap = p->q_cap;
  ca.poll = poll;
  ca.tie_xor = p->m->nrest > 0 ? ~0u : 0u;
  ca.wmin = p->desc.k_select == 1 ? p->wmin.as<unsigned long long>() : nullptr;
  ca.cmin = nullptr;
  ca.chunk_log2 = GPUGO_CHUNK_LOG2;
  ca.pad_c = 0;
  ca.band = band_view(p);
  const bool lazy_tail = p->lazy && more && p->m->compact_select_fn && p->steps_enqueued > 0 &&
                         p->since_compact + 1 < p->compact_every;
  // entries the compaction may read: the bound before this step plus its appends
  {
    const long long maxk = p->steps_enqueued == 0 ? 1 : p->desc.k_select;
    const long long n_hi = p->q_hi + maxk * (long long)p->W;
    p->q_hi = n_hi < p->q_cap ? n_hi : p->q_cap;
  }
  // workgroups for up to q_hi entries: a full pass covers SEL_T * 8 entries
  // per workgroup (CMP_U = 8), a lazy tail one chunk per workgroup
  auto grid_for = [&](long long per_wg, int max_wg) {
    const long long n = (p->q_hi + per_wg - 1) / per_wg;
    return (int)(n < 1 ? 1 : (n > max_wg ? max_wg : n));
  };
  if (lazy_tail) {
    // the queue stays in A; workgroups refresh the changed chunk minima,
    // the last one selects from A
    const int nwg = grid_for(1LL << GPUGO_CHUNK_LOG2, GPUGO_LAZY_WG);
    ca.cmin = p->cmin.as<unsigned long long>();
    SelectArgs sa = select_args(p, A, nwg);
    ca.poll = nullptr;
    void *fargs[] = {&ca, &sa, &poll, &pstate};
    HIPCHECK(hipModuleLaunchKernel(p->m->compact_select_fn, nwg, 1, 1, SEL_T, 1, 1, 0, p->stream, fargs,
                                   nullptr));
    p->selected_ahead = true;
    ++p->lazy_steps;
    ++p->since_compact;
    ++p->steps_enqueued;
    return GPUGO_OK;
  }
  p->since_compact = 0;
  if (more && p->m->compact_select_fn) {
    // workgroups of 1024 threads compact; the last one selects from B
    const int nwg = grid_for(8LL * SEL_T, p->lazy ? GPUGO_LAZY_FULL_WG : GPUGO_FUSED_WG);
    SelectArgs sa = select_args(p, B, nwg);
    ca.poll = nullptr;
    void *fargs[] = {&ca, &sa, &poll, &pstate};
    HIPCHECK(hipModuleLaunchKernel(p->m->compact_select_fn, nwg, 1, 1, SEL_T, 1, 1, 0, p->stream, fargs, nullptr));
    p->selected_ahead = true;
  } else {
    const int nwg = grid_for(8LL * 256, GPUGO_COMPACT_WG);
    p->last_nwg = nwg;
    if (p->m->compact_fn) {
      void *cargs[] = {&ca};
      HIPCHECK(hipModuleLaunchKernel(p->m->compact_fn, nwg, 1, 1, 256, 1, 1, 0, p->stream, cargs, nullptr));
    } else {
      hipLaunchKernelGGL(k_compact, dim3(nwg), dim3(256), 0, p->stream, ca);
      HIPCHECK(hipGetLastError());
    }
  }
  ++p->steps_enqueued;
  return GPUGO_OK;
}

static int read_state(gpugo_problem_t p) {
  HIPCHECK(hipMemcpyAsync(&p->hs, p->st.p, sizeof(DevState), hipMemcpyDeviceToHost, p->stream));
  HIPCHECK(hipStreamSynchronize(p->stream));
  return GPUGO_OK;
}

static int live_buffer(gpugo_problem_t p) { return (int)(p->hs.compactions & 1); }

static int fill_result(gpugo_problem_t p, gpugo_result *res, double ms, bool max_iters_hit) {
  DevState &s = p->hs;
  memset(res, 0, sizeof *res);
  double ub = key_value(s.ub_key);
  double lb = s.lb_global;
  if (s.status == GPUGO_ST_RUNNING || s.status == GPUGO_ST_OVERFLOW) {
    // recompute the smallest live lower bound exactly
    const int lbuf = live_buffer(p);
    long long n = s.compact_count;
    unsigned long long *d_key = p->minkey.as<unsigned long long>();
    hipLaunchKernelGGL(k_minlb, dim3(1), dim3(1024), 0, p->stream, p->st.as<DevState>(), p->q[lbuf].plb(), n, d_key);
    HIPCHECK(hipGetLastError());
    unsigned long long k;
    HIPCHECK(hipMemcpyAsync(&k, d_key, sizeof k, hipMemcpyDeviceToHost, p->stream));
    HIPCHECK(hipStreamSynchronize(p->stream));
    lb = key_value(k);
    if (s.out_min_lb < lb) lb = s.out_min_lb;
    if (n == 0 && s.n_out == 0) lb = -__builtin_huge_val();
  }
  res->lb = lb;
  res->ub = ub;
  {
    // ub - lb rounded up, as the device's stop test rounds it: the nearest
    // difference, one ulp up when its rounding error (two-sum) is positive
    double w = ub - lb;
    if (std::isfinite(w)) {
      const double bv = w - ub, err = (ub - (w - bv)) + (-lb - bv);
      if (err > 0) w = std::nextafter(w, __builtin_huge_val());
    }
    res->width = (s.status == GPUGO_ST_CERTIFIED || s.status == GPUGO_ST_X_CONVERGED) ? s.width : w;
  }
  res->status = s.status;
  res->overflow = s.overflow & 3;
  if (s.status == GPUGO_ST_RUNNING && max_iters_hit) res->status = GPUGO_ST_MAX_ITERS;
  res->iterations = s.iterations;
  res->boxes_bounded = s.bounded;
  res->pruned_value = s.pruned_value;
  res->pruned_kkt = s.pruned_kkt;
  res->queue_size = s.compact_count;
  res->queue_peak = s.queue_peak;
  res->n_out = s.n_out;
  res->ms = ms;
  return GPUGO_OK;
}

// Up to `iters` steps in one cooperative launch of the persistent step
// loop (it stops by itself when the problem stops).  With timing on, the
// launch is bracketed by events and counted as the bound time.
static int persist_steps(gpugo_problem_t p, int64_t iters) {
  const int d = p->desc.d;
  const bool mv = (p->desc.flags & GPUGO_FLAG_MEAN_VALUE) != 0;
  hipFunction_t fn = mv ? p->m->persist_mv[p->desc.rounding] : p->m->persist[p->desc.rounding];
  DevState *st = p->st.as<DevState>();
  PersistArgs pa;
  memset(&pa, 0, sizeof pa);
  BoundArgs &ba = pa.ba;
  ba.plo = p->plo.as<double>();
  ba.phi = p->phi.as<double>();
  ba.pw = p->pw.as<double>();
  ba.pwin = p->pwin.as<int>();
  ba.params = p->d_params.as<double>();
  ba.dom_lo = p->d_lo.as<double>();
  ba.dom_hi = p->d_hi.as<double>();
  ba.st = st;
  ba.q_cap = p->q_cap;
  ba.d = d;
  ba.wid = p->desc.wid;
  ba.c = p->c;
  ba.n_samples = p->desc.n_samples;
  ba.W = p->W;
  ba.inv_wid = 1.0 / (double)p->desc.wid;
  ba.half_inv_wid = 0.5 / (double)p->desc.wid;
  ba.flags = p->desc.flags;
  ba.params32 = p->d_params32.p;
  SelectArgs &sa = pa.sa;
  sa.st = st;
  sa.plo = p->plo.as<double>();
  sa.phi = p->phi.as<double>();
  sa.pw = p->pw.as<double>();
  sa.pwin = p->pwin.as<int>();
  sa.out_lo = p->out_lo.as<double>();
  sa.out_hi = p->out_hi.as<double>();
  sa.out_lb = p->out_lb.as<double>();
  sa.out_cap = p->out_cap;
  sa.p_cap = p->p_cap;
  sa.d = d;
  sa.wid = p->desc.wid;
  sa.c = p->c;
  sa.k_select = p->desc.k_select;
  sa.tol_f = p->desc.tol_f;
  sa.tol_x = p->desc.tol_x;
  sa.max_iters = 0;
  sa.pstride = d;
  sa.ext_n = 0;
  sa.pext = nullptr;
  sa.tie_xor = 0u;
  sa.nwmin = 0;
  sa.wmin = nullptr;
  sa.lazy = 0;
  sa.grow = 0;
  sa.q_cap = p->q_cap;
  sa.W = p->W;
  sa.grow_age = 0;
  sa.band = band_view(p);
  for (int b = 0; b < 2; ++b) {
    pa.q_lb[b] = p->q[b].plb();
    pa.q_par[b] = p->q[b].ppar();
    pa.q_pos[b] = p->q[b].ppos();
  }
  pa.bar = p->bar.as<unsigned>();
  pa.err = (int *)(p->bar.as<unsigned>() + 2);
  pa.step0 = p->steps_enqueued;
  pa.max_steps = iters;
  pa.cap = p->q_cap;
  pa.nblocks = p->persist_blocks;
  HIPCHECK(hipMemsetAsync(p->bar.p, 0, 16, p->stream));
  if (p->timing) {
    while (p->tev.size() < p->tev_used + 2) {
      hipEvent_t ev;
      HIPCHECK(hipEventCreate(&ev));
      p->tev.push_back(ev);
    }
    HIPCHECK(hipEventRecord(p->tev[p->tev_used++], p->stream));
  }
  void *args[] = {&pa};
  HIPCHECK(hipModuleLaunchCooperativeKernel(fn, (unsigned)pa.nblocks, 1, 1, 1024, 1, 1, 0, p->stream, args));
  if (p->timing) HIPCHECK(hipEventRecord(p->tev[p->tev_used++], p->stream));
  int err = 0;
  HIPCHECK(hipMemcpyAsync(&err, pa.err, sizeof err, hipMemcpyDeviceToHost, p->stream));
  HIPCHECK(hipStreamSynchronize(p->stream));
  if (err) return set_err(GPUGO_E_HIP, "persistent step loop: a grid barrier wait gave up");
  p->steps_enqueued += iters;  // a launch that returns running ran all of them
  // the steps this launch ran: the device counts them in iterations; the
  // parity of the next launch's live buffer follows the compactions count
  return GPUGO_OK;
}

// The status word of a batch of steps: the host spins on the pinned word
// (the batch's last compaction stores it with a system-scope store) instead
// of blocking in hipEventSynchronize, whose wake-up left the GPU idle for
// tens of microseconds between batches.  Every few hundred spins it asks the
// batch's event; if that completed without a word (not expected), the
// status is read from the device state.
static int wait_poll(gpugo_problem_t p, int slot, int *out) {
  for (unsigned it = 1;; ++it) {
    const int v = __atomic_load_n(&p->hpoll[slot], __ATOMIC_ACQUIRE);
    if (v != GPUGO_POLL_PENDING) {
      *out = v;
      return GPUGO_OK;
    }
    if ((it & 511u) == 0u) {
      const hipError_t e = hipEventQuery(p->poll_ev[slot]);
      if (e == hipSuccess) {
        const int w = __atomic_load_n(&p->hpoll[slot], __ATOMIC_ACQUIRE);
        if (w != GPUGO_POLL_PENDING) {
          *out = w;
          return GPUGO_OK;
        }
        int st = 0;
        HIPCHECK(hipMemcpy(&st, &p->st.as<DevState>()->status, sizeof st, hipMemcpyDeviceToHost));
        *out = st;
        return GPUGO_OK;
      }
      if (e != hipErrorNotReady) return set_err(GPUGO_E_HIP, "event query: %s", hipGetErrorString(e));
    }
    __builtin_ia32_pause();
  }
}

static int run_steps(gpugo_problem_t p, int64_t iters, gpugo_result *res, bool solve) {
  if (!p || !res) return set_err(GPUGO_E_INVALID, "null pointer");
  auto t0 = std::chrono::steady_clock::now();
  if (!p->started) {
    int rc = root_init(p);
    if (rc) return rc;
    p->started = true;
  }
  const long long it0 = p->hs.iterations;
  int final_copy = -1;  // slot whose pinned state copy holds the stopped problem's final state
  if (p->hs.status == GPUGO_ST_RUNNING && (p->desc.flags & GPUGO_FLAG_PERSIST) && iters > 0) {
    int rc = persist_steps(p, iters);
    if (rc) return rc;
  } else if (p->hs.status == GPUGO_ST_RUNNING && !p->hpoll) {
    // no pinned status word (static-arena solves): batches of steps, then a
    // copy of the status to host memory and a stream synchronisation
    const int64_t B = 8;
    int64_t done = 0;
    while (done < iters) {
      const int64_t batch = B < iters - done ? B : iters - done;
      for (int64_t i = 0; i < batch; ++i) {
        int rc = enqueue_step(p, nullptr, done + i + 1 < iters);
        if (rc) return rc;
      }
      done += batch;
      int status = GPUGO_ST_RUNNING;
      HIPCHECK(hipMemcpyAsync(&status, &p->st.as<DevState>()->status, sizeof status, hipMemcpyDeviceToHost,
                              p->stream));
      HIPCHECK(hipStreamSynchronize(p->stream));
      if (status != GPUGO_ST_RUNNING) break;
    }
  } else if (p->hs.status == GPUGO_ST_RUNNING) {
    // Steps are enqueued in batches; the compaction of a batch's last step
    // writes the status into a pinned slot (no copy kernel).  The host waits
    // only for the previous batch, so it stays one batch ahead of the device; steps
    // enqueued after a stop are no-ops on the device.  Exactly `iters`
    // steps are enqueued at most.
    // The first step of a solve runs alone and the host waits for it: a
    // problem that certifies in its first step then enqueues no steps that
    // would find it stopped (one host round trip for the others).
    const int64_t B = 2;
    int64_t done = 0;
    int slot = 0, pending = 0;
    while (done < iters) {
      const bool first = p->steps_enqueued == 0;
      int64_t batch = first ? 1 : (B < iters - done ? B : iters - done);
      __atomic_store_n(&p->hpoll[slot], GPUGO_POLL_PENDING, __ATOMIC_RELEASE);
      for (int64_t i = 0; i < batch; ++i) {
        int rc = enqueue_step(p, i + 1 == batch ? p->dpoll + slot : nullptr, done + i + 1 < iters);
        if (rc) return rc;
      }
      done += batch;
      HIPCHECK(hipEventRecord(p->poll_ev[slot], p->stream));
      ++pending;
      if (pending == 2 || first) {
        const int old = first ? slot : slot ^ 1;
        int pv;
        int rc = wait_poll(p, old, &pv);
        if (rc) return rc;
        --pending;
        if (pv != GPUGO_ST_RUNNING) {
          // the stop's final compaction still to run (a large queue) and no
          // later step in flight to run it: enqueue it alone
          if (pending == 0 && (pv & GPUGO_POLL_FINAL)) {
            int rc = enqueue_step(p, nullptr, false, true);
            if (rc) return rc;
          } else if (pending == 0 && (pv & GPUGO_POLL_STATE)) {
            final_copy = old;
          } else if (pending == 1) {
            // the newer batch (steps that find the problem stopped, or the
            // one that runs its final compaction) ends with a word too
            int pv2;
            int rc = wait_poll(p, slot, &pv2);
            if (rc) return rc;
            --pending;
            if (pv2 & GPUGO_POLL_STATE) final_copy = slot;
          }
          break;
        }
      }
      slot ^= 1;
    }
  }
  int rc;
  if (final_copy >= 0) {
    // the device wrote the final state before the status word; nothing is
    // left in flight but the end of that launch
    memcpy(&p->hs, poll_state_h(p, final_copy), sizeof(DevState));
    HIPCHECK(hipStreamSynchronize(p->stream));
  } else {
    rc = read_state(p);
    if (rc) return rc;
  }
  if (p->timing) {
    // working launches come first in the recorded order
    long long nwork = p->hs.iterations - it0;
    size_t keep = (size_t)(2 * nwork) < p->tev_used ? (size_t)(2 * nwork) : p->tev_used;
    p->tev_used = keep;
    rc = timing_collect(p);
    if (rc) return rc;
    p->bound_launches += nwork;
  }
  p->tev_used = 0;
  double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  bool hit = solve && p->hs.status == GPUGO_ST_RUNNING;
  rc = fill_result(p, res, ms, hit);
  if (rc) return rc;
  res->bound_ms = p->bound_ms;
  res->bound_launches = p->bound_launches;
  if (p->hs.overflow & 1) return set_err(GPUGO_E_CAPACITY, "box queue capacity (%lld) exceeded", p->q_req);
  if (p->hs.overflow & 2) return set_err(GPUGO_E_CAPACITY, "parent arena capacity (%lld) exceeded", p->p_cap);
  return GPUGO_OK;
}

extern "C" int gpugo_step(gpugo_problem_t p, int64_t iters, gpugo_result *res) {
  if (iters < 0) return set_err(GPUGO_E_INVALID, "negative iters");
  return run_steps(p, iters, res, false);
}

extern "C" int gpugo_solve(gpugo_problem_t p, gpugo_result *res) {
  if (!p) return set_err(GPUGO_E_INVALID, "null problem");
  long long left = p->desc.max_iters - p->hs.iterations;
  if (left < 0) left = 0;
  return run_steps(p, left, res, true);
}

// ---------------------------------------------------------------------------
// static arena: a solve that allocates nothing
// ---------------------------------------------------------------------------
// Device memory of gpugo_solve_static: one array per GPU, held from the
// library's load until the process ends.  Host code never uses its name as
// an address: the device address comes from hipGetSymbolAddress, once per
// device (gpugo_init, else the first static solve on that device).
#define GPUGO_STATIC_BYTES GPUGO_STATIC_ARENA_BYTES
#define GPUGO_STATIC_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[GPUGO_STATIC_BYTES];
static void *g_scratch_dev[GPUGO_STATIC_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int static_arena(int dev, unsigned char **out) {
  if (dev < 0 || dev >= GPUGO_STATIC_MAX_DEVICES) return set_err(GPUGO_E_INVALID, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_err(GPUGO_E_HIP, "static arena address: %s", hipGetErrorString(e));
    g_scratch_dev[dev] = ptr;
  }
  *out = (unsigned char *)g_scratch_dev[dev];
  return GPUGO_OK;
}

// outward FP32 intervals of the FP64 coefficients (FP32 pruning pass)
__global__ __launch_bounds__(256) void k_params32(const double *p, float *out, long long n) {
  for (long long i = threadIdx.x; i < n; i += blockDim.x) {
    out[2 * i] = to_dn<float>(p[i]);
    out[2 * i + 1] = to_up<float>(p[i]);
  }
}

extern "C" int gpugo_solve_static(gpugo_module_t m, const gpugo_problem_desc *desc, void *stream,
                                  gpugo_result *res) {
  if (!desc || !res) return set_err(GPUGO_E_INVALID, "null pointer");
  int rc = problem_check(m, desc);
  if (rc) return rc;
  if (desc->flags & GPUGO_FLAG_PERSIST)
    return set_err(GPUGO_E_INVALID, "the persistent step loop needs gpugo_problem_create");
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  unsigned char *base = nullptr;
  rc = static_arena(dev, &base);
  if (rc) return rc;
  gpugo_problem_s p;  // its buffers point into the arena; nothing to free
  p.m = m;
  p.desc = *desc;
  const int d = desc->d;
  // defaults sized for the arena: 2^20 queue entries, 2^15 parent records
  rc = problem_shape(&p, desc, 1LL << 20, 1LL << 15);
  if (rc) return rc;
  size_t off = 0;
  rc = problem_buffers(&p, [&](DevBuf &buf, size_t bytes, const char *) {
    buf.p = nullptr;
    buf.own = false;
    if (off + bytes <= GPUGO_STATIC_BYTES) buf.p = base + off;
    off = (off + bytes + 255) & ~(size_t)255;
    return (int)GPUGO_OK;
  });
  if (rc) return rc;
  if (off > GPUGO_STATIC_BYTES)
    return set_err(GPUGO_E_CAPACITY,
                   "the static arena holds %llu bytes; this problem needs %llu (queue %lld entries, %lld parent "
                   "records): lower max_queue / max_parents",
                   (unsigned long long)GPUGO_STATIC_BYTES, (unsigned long long)off, p.q_cap, p.p_cap);
  p.stream = (hipStream_t)stream;
  const long long np = desc->n_params;
  HIPCHECK(hipMemcpyAsync(p.d_lo.p, desc->lo, sizeof(double) * d, hipMemcpyHostToDevice, p.stream));
  HIPCHECK(hipMemcpyAsync(p.d_hi.p, desc->hi, sizeof(double) * d, hipMemcpyHostToDevice, p.stream));
  if (np > 0) {
    HIPCHECK(hipMemcpyAsync(p.d_params.p, desc->params, sizeof(double) * np, hipMemcpyHostToDevice, p.stream));
    hipLaunchKernelGGL(k_params32, dim3(1), dim3(256), 0, p.stream, p.d_params.as<double>(),
                       (float *)p.d_params32.p, np);
    HIPCHECK(hipGetLastError());
  }
  memset(&p.hs, 0, sizeof p.hs);
  rc = run_steps(&p, p.desc.max_iters, res, true);
  // the arena is reused by the next call: nothing may still run on it
  const hipError_t e = hipStreamSynchronize(p.stream);
  if (!rc && e != hipSuccess) rc = set_err(GPUGO_E_HIP, "hipStreamSynchronize: %s", hipGetErrorString(e));
  p.stream = nullptr;
  return rc;
}

extern "C" int gpugo_boxes_copy(gpugo_problem_t p, int64_t max, double *lo, double *hi, double *lb, int64_t *n) {
  if (!p || !n) return set_err(GPUGO_E_INVALID, "null pointer");
  if (max < 0) return set_err(GPUGO_E_INVALID, "negative max");
  *n = 0;
  if (!p->started) return GPUGO_OK;
  int rc = read_state(p);
  if (rc) return rc;
  const int d = p->desc.d;
  const long long nlive = p->hs.compact_count;
  if (nlive == 0) {
    // output boxes of the x-width rule
    long long no = p->hs.n_out < p->out_cap ? p->hs.n_out : p->out_cap;
    *n = no;
    long long m = no < max ? no : max;
    if (m > 0) {
      if (lo) HIPCHECK(hipMemcpy(lo, p->out_lo.p, sizeof(double) * d * m, hipMemcpyDeviceToHost));
      if (hi) HIPCHECK(hipMemcpy(hi, p->out_hi.p, sizeof(double) * d * m, hipMemcpyDeviceToHost));
      if (lb) HIPCHECK(hipMemcpy(lb, p->out_lb.p, sizeof(double) * m, hipMemcpyDeviceToHost));
    }
    return GPUGO_OK;
  }
  *n = nlive;
  long long m = nlive < max ? nlive : max;
  if (m == 0) return GPUGO_OK;
  const Queue &Q = p->q[live_buffer(p)];
  std::vector<double> hlb(nlive);
  std::vector<unsigned> hpar(nlive), hpos(nlive);
  HIPCHECK(hipMemcpy(hlb.data(), Q.lb.p, sizeof(double) * nlive, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(hpar.data(), Q.par.p, sizeof(unsigned) * nlive, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(hpos.data(), Q.pos.p, sizeof(unsigned) * nlive, hipMemcpyDeviceToHost));
  std::vector<long long> order(nlive);
  for (long long i = 0; i < nlive; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](long long x, long long y) {
    if (hlb[x] != hlb[y]) return hlb[x] < hlb[y];
    if (hpar[x] != hpar[y]) return hpar[x] < hpar[y];
    return hpos[x] < hpos[y];
  });
  order.resize(m);
  DevBuf d_order, d_lo, d_hi;
  HIPCHECK(d_order.alloc(sizeof(long long) * m));
  HIPCHECK(d_lo.alloc(sizeof(double) * d * m));
  HIPCHECK(d_hi.alloc(sizeof(double) * d * m));
  HIPCHECK(hipMemcpy(d_order.p, order.data(), sizeof(long long) * m, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_materialize, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, p->stream, p->plo.as<double>(),
                     p->phi.as<double>(), p->pw.as<double>(), p->pwin.as<int>(), Q.ppar(), Q.ppos(),
                     d_order.as<long long>(), m, d, p->desc.wid, p->c, d_lo.as<double>(), d_hi.as<double>());
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(p->stream));
  if (lo) HIPCHECK(hipMemcpy(lo, d_lo.p, sizeof(double) * d * m, hipMemcpyDeviceToHost));
  if (hi) HIPCHECK(hipMemcpy(hi, d_hi.p, sizeof(double) * d * m, hipMemcpyDeviceToHost));
  if (lb)
    for (long long i = 0; i < m; ++i) lb[i] = hlb[order[i]];
  return GPUGO_OK;
}

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
static_assert(sizeof(AdmitHead) % 8 == 0, "admission records must keep doubles aligned");

struct BatchHostSlot {
  int running = 0;
  int growing = 0;  // stopped in GPUGO_ST_GROW, waiting for a slot in the next tier
  int64_t tag = 0;
  long long admit_poll = 0;  // polls with a smaller sequence number predate the admission
  std::chrono::steady_clock::time_point t0;
};

struct gpugo_batch_s {
  gpugo_module_t m = nullptr;
  gpugo_batch_desc desc;
  int dmax = 0, pmax = 0, slots = 0, bound_grid = 0;
  long long qcap = 0, pcap = 0, rec_bytes = 0;
  DevBuf info, st, plo, phi, pw, pwin, params, dom_lo, dom_hi, q_lb, q_par, q_pos, wave_off, recs;
  DevBuf band_lb, band_par, band_pos, band_idx, band;
  hipStream_t stream = nullptr;
  BatchArgs ba;
  BatchBoundArgs bb;
  hipFunction_t bound = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // admission staging: two pinned buffers, each reusable once its copy ran
  char *stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  bool stage_busy[2] = {false, false};
  int cur_stage = 0, n_rec = 0;
  std::vector<int> free_list;
  std::vector<BatchHostSlot> hs;
  long long running = 0;
  // state polling: pinned copies of all slot states, at most two in flight
  DevState *hpoll[2] = {nullptr, nullptr};
  hipEvent_t pev[2] = {nullptr, nullptr};
  long long poll_seq[2] = {0, 0};
  int pend_first = 0, pend_n = 0;
  long long polls = 0;
  std::vector<std::pair<int64_t, gpugo_result>> done;
  size_t done_head = 0;
  // growth tiers: tier k (k >= 1) has GPUGO_GROW_FACTOR^k times the per-slot
  // queue and arena of the batch the caller created (tier 0) and
  // slots >> (GPUGO_GROW_SLOT_SHIFT * k) slots (at least 1).  Tier 0 owns the
  // others; finished problems of every tier are reported by tier 0.
  int level = 0;
  gpugo_batch_s *root = nullptr;
  std::vector<gpugo_batch_s *> tiers;  // tier 0 only: tiers 1..grow
  std::vector<int> grow_wait;          // slots stopped in GPUGO_ST_GROW, not yet moved
  hipEvent_t ev_ready = nullptr;       // recorded on this tier's stream before it is the source of moves
  hipEvent_t ev_moved = nullptr;       // recorded on this tier's stream after moves into it
  void *kp_bb[1] = {nullptr};          // kernel argument array of the split+bound launches
};

#define GPUGO_GROW_MAX 3
#define GPUGO_GROW_FACTOR 8
#define GPUGO_GROW_SLOT_SHIFT 4

// description of growth tier k of a batch with description d
static gpugo_batch_desc tier_desc(const gpugo_batch_desc &d, long long q0, long long p0, int k) {
  gpugo_batch_desc t = d;
  long long f = 1;
  for (int i = 0; i < k; ++i) f *= GPUGO_GROW_FACTOR;
  t.slots = std::max(1, d.slots >> (GPUGO_GROW_SLOT_SHIFT * k));
  t.max_queue = std::min(q0 * f, 0xfffffff0LL);
  t.max_parents = std::min(p0 * f, 0xfffffff0LL);
  t.grow = 0;
  return t;
}

static int batch_sizes(gpugo_module_t m, const gpugo_batch_desc *desc, long long *qcap, long long *pcap,
                       long long *rec_bytes, int64_t *bytes) {
  if (!m || !desc) return set_err(GPUGO_E_INVALID, "null pointer");
  if (!m->batch) return set_err(GPUGO_E_INVALID, "not a batch module");
  if (desc->slots < 1 || desc->slots > (1 << 20)) return set_err(GPUGO_E_INVALID, "slots must be in [1, 2^20]");
  if (desc->steps_per_launch < 1 || desc->steps_per_launch > 1024)
    return set_err(GPUGO_E_INVALID, "steps_per_launch must be in [1, 1024]");
  if (desc->rounding != GPUGO_ROUND_EXACT && desc->rounding != GPUGO_ROUND_FAST)
    return set_err(GPUGO_E_INVALID, "unknown rounding %d", desc->rounding);
  if (!m->batch_bound[desc->rounding])
    return set_err(GPUGO_E_INVALID, "module has no %s variant", desc->rounding ? "fast" : "exact");
  if (desc->max_queue < 0 || desc->max_parents < 0 || desc->bound_blocks < 0)
    return set_err(GPUGO_E_INVALID, "negative capacity");
  if (desc->grow < 0 || desc->grow > GPUGO_GROW_MAX)
    return set_err(GPUGO_E_INVALID, "grow must be in [0, %d]", GPUGO_GROW_MAX);
  if (desc->promote_age < 0) return set_err(GPUGO_E_INVALID, "promote_age must be >= 0");
  const long long q = desc->max_queue ? desc->max_queue : 65536, p = desc->max_parents ? desc->max_parents : 16384;
  if (q > 0xfffffff0LL || p > 0xfffffff0LL) return set_err(GPUGO_E_INVALID, "per-slot capacity too large");
  const long long D = m->dim, PM = m->pmax;
  const long long rb = (long long)sizeof(AdmitHead) + 8 * (PM + 3 * D);
  long long total = 0;
  for (int k = 0; k <= desc->grow; ++k) {
    const gpugo_batch_desc t = tier_desc(*desc, q, p, k);
    const long long S = t.slots, tq = t.max_queue, tp = t.max_parents;
    total += S * (long long)(sizeof(SlotInfo) + sizeof(DevState)) + S * tp * (3 * D * 8 + 4) + S * PM * 8 +
             S * 2 * D * 8 + S * 2 * tq * 16 + (S + 1) * 8 + S * rb +
             S * (GPUGO_BAND_CAP * 20LL + (long long)sizeof(BandState));
  }
  *qcap = q;
  *pcap = p;
  *rec_bytes = rb;
  *bytes = total;
  return GPUGO_OK;
}

extern "C" int gpugo_batch_bytes(gpugo_module_t m, const gpugo_batch_desc *desc, int64_t *bytes) {
  if (!bytes) return set_err(GPUGO_E_INVALID, "null pointer");
  long long q, p, rb;
  return batch_sizes(m, desc, &q, &p, &rb, bytes);
}

extern "C" int gpugo_mem_info(int64_t *free_bytes, int64_t *total_bytes) {
  if (!free_bytes || !total_bytes) return set_err(GPUGO_E_INVALID, "null pointer");
  size_t f = 0, t = 0;
  HIPCHECK(hipMemGetInfo(&f, &t));
  *free_bytes = (int64_t)f;
  *total_bytes = (int64_t)t;
  return GPUGO_OK;
}

static int batch_enqueue_step(gpugo_batch_t b) {
  hipLaunchKernelGGL(kb_select, dim3((unsigned)b->slots), dim3(SEL_T), 0, b->stream, b->ba);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(kb_plan, dim3(1), dim3(1024), 0, b->stream, b->ba);
  HIPCHECK(hipGetLastError());
  // the argument array belongs to the batch: a captured launch may keep it
  // for the life of the graph
  HIPCHECK(hipModuleLaunchKernel(b->bound, (unsigned)b->bound_grid, 1, 1, 256, 1, 1, 0, b->stream, b->kp_bb,
                                 nullptr));
  hipLaunchKernelGGL(kb_compact, dim3((unsigned)b->slots * b->ba.cb), dim3(256), 0, b->stream, b->ba);
  HIPCHECK(hipGetLastError());
  return GPUGO_OK;
}

extern "C" void gpugo_batch_destroy(gpugo_batch_t b) {
  if (!b) return;
  for (auto *t : b->tiers) gpugo_batch_destroy(t);
  b->tiers.clear();
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->ev_ready) (void)hipEventDestroy(b->ev_ready);
  if (b->ev_moved) (void)hipEventDestroy(b->ev_moved);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (int k = 0; k < 2; ++k) {
    if (b->stage[k]) (void)hipHostFree(b->stage[k]);
    if (b->hpoll[k]) (void)hipHostFree(b->hpoll[k]);
    if (b->stage_ev[k]) (void)hipEventDestroy(b->stage_ev[k]);
    if (b->pev[k]) (void)hipEventDestroy(b->pev[k]);
  }
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// One tier of a batch: its slots, pools, stream, polling buffers and (with
// use_graph) the captured launch.  grow_next: a larger tier exists above it.
static int batch_create_level(gpugo_module_t m, const gpugo_batch_desc *desc, int level, int grow_next,
                              gpugo_batch_t *out) {
  if (!out) return set_err(GPUGO_E_INVALID, "null pointer");
  *out = nullptr;
  long long qcap, pcap, rb;
  int64_t total;
  int rc = batch_sizes(m, desc, &qcap, &pcap, &rb, &total);
  if (rc) return rc;
  auto *b = new gpugo_batch_s;
  b->m = m;
  b->desc = *desc;
  b->level = level;
  b->dmax = m->dim;
  b->pmax = m->pmax;
  b->slots = desc->slots;
  b->qcap = qcap;
  b->pcap = pcap;
  b->rec_bytes = rb;
  b->bound = m->batch_bound[desc->rounding];
  auto fail = [&](int code) {
    gpugo_batch_destroy(b);
    return code;
  };
  const long long S = b->slots, D = b->dmax;
  hipError_t e;
#define BALLOC(buf, bytes)                                                                     \
  if ((e = (buf).alloc(bytes)) != hipSuccess)                                                  \
    return fail(set_err(e == hipErrorOutOfMemory ? GPUGO_E_NOMEM : GPUGO_E_HIP, "alloc %s: %s", \
                        #buf, hipGetErrorString(e)))
  BALLOC(b->info, sizeof(SlotInfo) * S);
  BALLOC(b->st, sizeof(DevState) * S);
  BALLOC(b->plo, sizeof(double) * S * pcap * D);
  BALLOC(b->phi, sizeof(double) * S * pcap * D);
  BALLOC(b->pw, sizeof(double) * S * pcap * D);
  BALLOC(b->pwin, sizeof(int) * S * pcap);
  BALLOC(b->params, sizeof(double) * S * b->pmax);
  BALLOC(b->dom_lo, sizeof(double) * S * D);
  BALLOC(b->dom_hi, sizeof(double) * S * D);
  BALLOC(b->q_lb, sizeof(double) * S * 2 * qcap);
  BALLOC(b->q_par, sizeof(unsigned) * S * 2 * qcap);
  BALLOC(b->q_pos, sizeof(unsigned) * S * 2 * qcap);
  BALLOC(b->wave_off, sizeof(long long) * (S + 1));
  BALLOC(b->recs, (size_t)(rb * S));
  BALLOC(b->band_lb, sizeof(double) * S * GPUGO_BAND_CAP);
  BALLOC(b->band_par, sizeof(unsigned) * S * GPUGO_BAND_CAP);
  BALLOC(b->band_pos, sizeof(unsigned) * S * GPUGO_BAND_CAP);
  BALLOC(b->band_idx, sizeof(unsigned) * S * GPUGO_BAND_CAP);
  BALLOC(b->band, sizeof(BandState) * S);
#undef BALLOC
  if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess)
    return fail(set_err(GPUGO_E_HIP, "stream: %s", hipGetErrorString(e)));
  if ((e = hipEventCreateWithFlags(&b->ev_ready, hipEventDisableTiming)) != hipSuccess ||
      (e = hipEventCreateWithFlags(&b->ev_moved, hipEventDisableTiming)) != hipSuccess)
    return fail(set_err(GPUGO_E_HIP, "event: %s", hipGetErrorString(e)));
  for (int k = 0; k < 2; ++k) {
    if ((e = hipHostMalloc((void **)&b->stage[k], (size_t)(rb * S), hipHostMallocDefault)) != hipSuccess ||
        (e = hipHostMalloc((void **)&b->hpoll[k], sizeof(DevState) * S, hipHostMallocDefault)) != hipSuccess)
      return fail(set_err(GPUGO_E_NOMEM, "pinned buffers: %s", hipGetErrorString(e)));
    if ((e = hipEventCreateWithFlags(&b->stage_ev[k], hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->pev[k], hipEventDisableTiming)) != hipSuccess)
      return fail(set_err(GPUGO_E_HIP, "event: %s", hipGetErrorString(e)));
  }
  // all slots inactive
  {
    std::vector<SlotInfo> zi(S);
    memset(zi.data(), 0, sizeof(SlotInfo) * S);
    if ((e = hipMemcpy(b->info.p, zi.data(), sizeof(SlotInfo) * S, hipMemcpyHostToDevice)) != hipSuccess)
      return fail(set_err(GPUGO_E_HIP, "upload: %s", hipGetErrorString(e)));
    std::vector<DevState> zs(S);
    memset(zs.data(), 0, sizeof(DevState) * S);
    for (auto &z : zs) z.status = GPUGO_ST_EXHAUSTED;
    if ((e = hipMemcpy(b->st.p, zs.data(), sizeof(DevState) * S, hipMemcpyHostToDevice)) != hipSuccess)
      return fail(set_err(GPUGO_E_HIP, "upload: %s", hipGetErrorString(e)));
    std::vector<BandState> zb(S);
    for (auto &z : zb) {
      z.g = -__builtin_huge_val();
      z.n = 0;
    }
    if ((e = hipMemcpy(b->band.p, zb.data(), sizeof(BandState) * S, hipMemcpyHostToDevice)) != hipSuccess)
      return fail(set_err(GPUGO_E_HIP, "upload: %s", hipGetErrorString(e)));
  }
  BatchArgs &a = b->ba;
  a.info = b->info.as<SlotInfo>();
  a.st = b->st.as<DevState>();
  a.q_lb = b->q_lb.as<double>();
  a.q_par = b->q_par.as<unsigned>();
  a.q_pos = b->q_pos.as<unsigned>();
  a.plo = b->plo.as<double>();
  a.phi = b->phi.as<double>();
  a.pw = b->pw.as<double>();
  a.pwin = b->pwin.as<int>();
  a.params = b->params.as<double>();
  a.dom_lo = b->dom_lo.as<double>();
  a.dom_hi = b->dom_hi.as<double>();
  a.wave_off = b->wave_off.as<long long>();
  a.pcap = pcap;
  a.qcap = qcap;
  a.dmax = b->dmax;
  a.pmax = b->pmax;
  a.slots = b->slots;
  a.grow = grow_next;
  // compaction workgroups per slot: 8, and 8 times more per growth tier
  a.cb = 8 << (3 * std::min(level, 2));
  a.grow_age = (level == 0 && grow_next) ? desc->promote_age : 0;
  a.band_lb = b->band_lb.as<double>();
  a.band_par = b->band_par.as<unsigned>();
  a.band_pos = b->band_pos.as<unsigned>();
  a.band_idx = b->band_idx.as<unsigned>();
  a.band = b->band.as<BandState>();
  a.band_cap = GPUGO_BAND_CAP;
  BatchBoundArgs &g = b->bb;
  memset(&g, 0, sizeof g);
  g.info = a.info;
  g.st = a.st;
  g.plo = a.plo;
  g.phi = a.phi;
  g.pw = a.pw;
  g.pwin = a.pwin;
  g.params = a.params;
  g.dom_lo = a.dom_lo;
  g.dom_hi = a.dom_hi;
  g.q_lb = a.q_lb;
  g.q_par = a.q_par;
  g.q_pos = a.q_pos;
  g.wave_off = a.wave_off;
  g.pcap = pcap;
  g.qcap = qcap;
  g.slots = b->slots;
  g.pmax = b->pmax;
  if (desc->bound_blocks > 0) {
    b->bound_grid = desc->bound_blocks;
  } else {
    int dev = 0, cus = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    b->bound_grid = 8 * (cus > 0 ? cus : 256);
  }
  b->hs.resize(S);
  b->free_list.reserve(S);
  for (long long s = S - 1; s >= 0; --s) b->free_list.push_back((int)s);
  b->kp_bb[0] = &b->bb;
  if (desc->use_graph) {
    if ((e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
      return fail(set_err(GPUGO_E_HIP, "graph capture: %s", hipGetErrorString(e)));
    for (int k = 0; k < desc->steps_per_launch; ++k) {
      rc = batch_enqueue_step(b);
      if (rc) {
        hipGraph_t gtmp = nullptr;
        (void)hipStreamEndCapture(b->stream, &gtmp);
        if (gtmp) (void)hipGraphDestroy(gtmp);
        return fail(rc);
      }
    }
    if ((e = hipStreamEndCapture(b->stream, &b->graph)) != hipSuccess)
      return fail(set_err(GPUGO_E_HIP, "graph capture: %s", hipGetErrorString(e)));
    if ((e = hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0)) != hipSuccess)
      return fail(set_err(GPUGO_E_HIP, "graph instantiate: %s", hipGetErrorString(e)));
  }
  *out = b;
  return GPUGO_OK;
}

extern "C" int gpugo_batch_create(gpugo_module_t m, const gpugo_batch_desc *desc, gpugo_batch_t *out) {
  if (!out || !desc) return set_err(GPUGO_E_INVALID, "null pointer");
  *out = nullptr;
  gpugo_batch_t b = nullptr;
  int rc = batch_create_level(m, desc, 0, desc->grow > 0, &b);
  if (rc) return rc;
  b->root = b;
  const long long q0 = b->qcap, p0 = b->pcap;
  for (int k = 1; k <= desc->grow; ++k) {
    const gpugo_batch_desc td = tier_desc(*desc, q0, p0, k);
    gpugo_batch_t t = nullptr;
    rc = batch_create_level(m, &td, k, k < desc->grow, &t);
    if (rc) {
      gpugo_batch_destroy(b);
      return rc;
    }
    t->root = b;
    b->tiers.push_back(t);
  }
  *out = b;
  return GPUGO_OK;
}

// the tier above b (nullptr for the last)
static gpugo_batch_t next_tier(gpugo_batch_t b) {
  gpugo_batch_t r = b->root;
  return b->level < (int)r->tiers.size() ? r->tiers[b->level] : nullptr;
}

// problems admitted into the batch and not yet reported, over all tiers
static long long total_running(gpugo_batch_t r) {
  long long n = r->running;
  for (auto *t : r->tiers) n += t->running;
  return n;
}

extern "C" int gpugo_batch_admit(gpugo_batch_t b, int32_t family, const gpugo_problem_desc *desc, int64_t tag,
                                 int32_t *slot_out) {
  if (!b || !desc) return set_err(GPUGO_E_INVALID, "null pointer");
  gpugo_module_t m = b->m;
  if (family < 0 || family >= m->nfam) return set_err(GPUGO_E_INVALID, "family %d out of range [0,%d)", family, m->nfam);
  const int d = desc->d;
  if (d < 1 || d > b->dmax) return set_err(GPUGO_E_INVALID, "d must be in [1,%d] for this batch module", b->dmax);
  const int np = m->fam_nparams[family];
  if (desc->n_params != np)
    return set_err(GPUGO_E_INVALID, "family %d expects %d parameters (dmax layout), got %lld", family, np,
                   (long long)desc->n_params);
  if (np > 0 && !desc->params) return set_err(GPUGO_E_INVALID, "null params");
  for (int i = 0; i < np; ++i) {
    const double v = desc->params[i];
    if (!std::isfinite(v)) return set_err(GPUGO_E_DOMAIN, "params[%d] is not finite", i);
    const int s = m->fam_psign[(size_t)family * m->pmax + i];
    if ((s > 0 && v < 0) || (s < 0 && v > 0))
      return set_err(GPUGO_E_DOMAIN, "params[%d] = %g violates the sign the module was compiled for", i, v);
  }
  if (!desc->lo || !desc->hi) return set_err(GPUGO_E_INVALID, "null box");
  int rc = check_intervals(desc->lo, desc->hi, d, "box");
  if (rc) return rc;
  for (int j = 0; j < d; ++j)
    if (!std::isfinite(desc->lo[j]) || !std::isfinite(desc->hi[j]))
      return set_err(GPUGO_E_DOMAIN, "box coordinate %d is not finite", j);
  if (desc->wid < 2 || desc->wid > 64) return set_err(GPUGO_E_INVALID, "wid must be in [2,64]");
  if (desc->cycle < 1) return set_err(GPUGO_E_INVALID, "cycle must be >= 1");
  if (desc->k_select < 1 || desc->k_select > SEL_KMAX)
    return set_err(GPUGO_E_INVALID, "k_select must be in [1,%d]", SEL_KMAX);
  if (desc->n_samples < 1 || desc->n_samples > 64) return set_err(GPUGO_E_INVALID, "n_samples must be in [1,64]");
  if (std::isnan(desc->tol_f) || std::isnan(desc->init_ub)) return set_err(GPUGO_E_DOMAIN, "NaN tolerance or init_ub");
  if (!(desc->tol_f > 0)) return set_err(GPUGO_E_INVALID, "batch problems need tol_f > 0");
  if (desc->flags & ~GPUGO_FLAG_MEAN_VALUE) return set_err(GPUGO_E_INVALID, "unknown flags 0x%x", desc->flags);
  if ((desc->flags & GPUGO_FLAG_MEAN_VALUE) && !m->batch_mv)
    return set_err(GPUGO_E_INVALID, "batch module built without the mean-value bound");
  const int c = desc->cycle < d ? desc->cycle : d;
  unsigned W = 0;
  if (!ipow_checked(desc->wid, c, &W) || W > (1u << 24))
    return set_err(GPUGO_E_INVALID, "wid^cycle must be <= 2^24 in a batch");
  if ((long long)desc->k_select * W > b->qcap)
    return set_err(GPUGO_E_INVALID, "k_select * wid^cycle = %lld exceeds the per-slot queue (%lld)",
                   (long long)desc->k_select * W, b->qcap);
  if (b->free_list.empty()) return set_err(GPUGO_E_CAPACITY, "no free slot");
  // staging buffer: wait until its previous copy has run
  const int k = b->cur_stage;
  if (b->n_rec == 0 && b->stage_busy[k]) {
    HIPCHECK(hipEventSynchronize(b->stage_ev[k]));
    b->stage_busy[k] = false;
  }
  const int s = b->free_list.back();
  char *rec = b->stage[k] + (long long)b->n_rec * b->rec_bytes;
  memset(rec, 0, (size_t)b->rec_bytes);
  AdmitHead *h = (AdmitHead *)rec;
  h->slot = s;
  SlotInfo &I = h->info;
  I.active = 1;
  I.family = family;
  I.d = d;
  I.c = c;
  I.wid = desc->wid;
  I.k_select = desc->k_select;
  I.n_samples = desc->n_samples;
  I.flags = desc->flags & GPUGO_FLAG_MEAN_VALUE;
  I.W = W;
  I.max_iters = desc->max_iters > 0 ? desc->max_iters : (1LL << 40);
  I.tol_f = desc->tol_f;
  I.inv_wid = 1.0 / (double)desc->wid;
  I.half_inv_wid = 0.5 / (double)desc->wid;
  DevState &z = h->st;
  z.ub_key = order_key(desc->init_ub);
  z.ub_snap = desc->init_ub;
  z.lb_global = -__builtin_huge_val();
  z.out_min_lb = __builtin_huge_val();
  z.width = __builtin_huge_val();
  z.parents_used = 1;
  z.first_parent = 0;
  z.n_split = 1;
  z.status = GPUGO_ST_RUNNING;
  z.lbmin_key = ~0ull;
  z.lbmax_key = 0ull;
  z.fresh = 1;
  double *pr = (double *)(h + 1);
  double *lo = pr + b->pmax, *hi = lo + b->dmax, *pw = hi + b->dmax;
  for (int i = 0; i < np; ++i) pr[i] = desc->params[i];
  for (int j = 0; j < d; ++j) {
    lo[j] = desc->lo[j];
    hi[j] = desc->hi[j];
    pw[j] = (desc->hi[j] - desc->lo[j]) / (double)desc->wid;
  }
  ++b->n_rec;
  b->free_list.pop_back();
  BatchHostSlot &H = b->hs[s];
  H.running = 1;
  H.tag = tag;
  H.admit_poll = b->polls;
  H.t0 = std::chrono::steady_clock::now();
  ++b->running;
  if (slot_out) *slot_out = s;
  return GPUGO_OK;
}

static int batch_flush(gpugo_batch_t b) {
  if (b->n_rec == 0) return GPUGO_OK;
  const int k = b->cur_stage;
  HIPCHECK(hipMemcpyAsync(b->recs.p, b->stage[k], (size_t)(b->n_rec * b->rec_bytes), hipMemcpyHostToDevice,
                          b->stream));
  HIPCHECK(hipEventRecord(b->stage_ev[k], b->stream));
  b->stage_busy[k] = true;
  hipLaunchKernelGGL(kb_admit, dim3((unsigned)b->n_rec), dim3(64), 0, b->stream, b->ba, (const char *)b->recs.p,
                     b->rec_bytes);
  HIPCHECK(hipGetLastError());
  b->n_rec = 0;
  b->cur_stage ^= 1;
  return GPUGO_OK;
}

// result of a finished slot from its final state
static void batch_result(const DevState &s, gpugo_result *r, double ms) {
  memset(r, 0, sizeof *r);
  r->lb = s.lb_global;
  r->ub = key_value(s.ub_key);
  r->width = s.width;
  r->status = s.status;
  r->overflow = s.overflow & 3;
  r->iterations = s.iterations;
  r->boxes_bounded = s.bounded;
  r->pruned_value = s.pruned_value;
  r->pruned_kkt = s.pruned_kkt;
  r->queue_size = s.compact_count;
  r->queue_peak = s.queue_peak;
  r->n_out = s.n_out;
  r->ms = ms;
}

// wait for the oldest poll in flight and collect the slots it shows finished
static int batch_collect_one(gpugo_batch_t b) {
  if (b->pend_n == 0) return GPUGO_OK;
  const int k = b->pend_first;
  HIPCHECK(hipEventSynchronize(b->pev[k]));
  const DevState *ps = b->hpoll[k];
  const long long seq = b->poll_seq[k];
  const auto now = std::chrono::steady_clock::now();
  for (int s = 0; s < b->slots; ++s) {
    BatchHostSlot &H = b->hs[s];
    if (!H.running || H.admit_poll > seq || ps[s].status == GPUGO_ST_RUNNING) continue;
    if (ps[s].status == GPUGO_ST_GROW) {
      // frozen until it moves into the next tier (its later steps are no-ops)
      if (!H.growing) {
        H.growing = 1;
        b->grow_wait.push_back(s);
      }
      continue;
    }
    gpugo_result r;
    batch_result(ps[s], &r, std::chrono::duration<double, std::milli>(now - H.t0).count());
    b->root->done.emplace_back(H.tag, r);
    H.running = 0;
    b->free_list.push_back(s);
    --b->running;
  }
  b->pend_first ^= 1;
  --b->pend_n;
  return GPUGO_OK;
}

// Move the slots of every tier that stopped in GPUGO_ST_GROW into free
// slots of the tier above.  The copy runs on the destination tier's stream,
// so no launch of that tier overlaps it (its launches read every slot's
// state), after the source tier's work so far (the frozen state is final);
// the source tier waits for the copy before its next work, which may write
// the freed slot (an admission, or a move into it from the tier below).  A
// slot that finds no free slot above stays frozen and moves on a later call.
static int batch_process_moves(gpugo_batch_t r) {
  std::vector<gpugo_batch_t> lv{r};
  lv.insert(lv.end(), r->tiers.begin(), r->tiers.end());
  for (size_t k = 0; k + 1 < lv.size(); ++k) {
    gpugo_batch_t X = lv[k], Y = lv[k + 1];
    bool moved = false;
    while (!X->grow_wait.empty() && !Y->free_list.empty()) {
      const int s = X->grow_wait.front(), t = Y->free_list.back();
      X->grow_wait.erase(X->grow_wait.begin());
      Y->free_list.pop_back();
      if (!moved) {
        HIPCHECK(hipEventRecord(X->ev_ready, X->stream));
        HIPCHECK(hipStreamWaitEvent(Y->stream, X->ev_ready, 0));
      }
      hipLaunchKernelGGL(kb_move, dim3(1024), dim3(256), 0, Y->stream, X->ba, Y->ba, s, t);
      HIPCHECK(hipGetLastError());
      BatchHostSlot &HX = X->hs[s], &HY = Y->hs[t];
      HY.running = 1;
      HY.growing = 0;
      HY.tag = HX.tag;
      HY.t0 = HX.t0;
      HY.admit_poll = Y->polls;  // polls taken before the move show the slot's old state
      ++Y->running;
      HX.running = 0;
      HX.growing = 0;
      X->free_list.push_back(s);
      --X->running;
      moved = true;
    }
    if (moved) {
      HIPCHECK(hipEventRecord(Y->ev_moved, Y->stream));
      HIPCHECK(hipStreamWaitEvent(X->stream, Y->ev_moved, 0));
    }
  }
  return GPUGO_OK;
}

// One launch of tier b (its graph, or steps_per_launch steps), then a copy of
// all slot states to pinned memory; at most two copies are in flight, the
// older one is collected first.
// One launch of a tier, then a copy of its slots' states to host memory.
// block: keep at most one launch in flight (wait for the older one); else
// the caller has made sure that fewer than two are in flight.
static int batch_launch_one(gpugo_batch_t b, bool block = true) {
  int rc;
  if (b->gexec) {
    HIPCHECK(hipGraphLaunch(b->gexec, b->stream));
  } else {
    for (int k = 0; k < b->desc.steps_per_launch; ++k) {
      rc = batch_enqueue_step(b);
      if (rc) return rc;
    }
  }
  if (b->pend_n == 2) {
    rc = batch_collect_one(b);
    if (rc) return rc;
  }
  const int k = (b->pend_first + b->pend_n) & 1;
  HIPCHECK(hipMemcpyAsync(b->hpoll[k], b->st.p, sizeof(DevState) * b->slots, hipMemcpyDeviceToHost, b->stream));
  HIPCHECK(hipEventRecord(b->pev[k], b->stream));
  b->poll_seq[k] = b->polls++;
  ++b->pend_n;
  if (block && b->pend_n == 2) {
    rc = batch_collect_one(b);
    if (rc) return rc;
  }
  return GPUGO_OK;
}

// Collect the launches of a tier whose state copies have landed, oldest
// first, without waiting for the others.
static int batch_collect_ready(gpugo_batch_t b) {
  while (b->pend_n > 0) {
    const hipError_t e = hipEventQuery(b->pev[b->pend_first]);
    if (e == hipErrorNotReady) return GPUGO_OK;
    if (e != hipSuccess) return set_err(GPUGO_E_HIP, "event query: %s", hipGetErrorString(e));
    const int rc = batch_collect_one(b);
    if (rc) return rc;
  }
  return GPUGO_OK;
}

extern "C" int gpugo_batch_advance(gpugo_batch_t b, int64_t *launched) {
  if (!b) return set_err(GPUGO_E_INVALID, "null batch");
  if (b->level != 0) return set_err(GPUGO_E_INVALID, "not a batch handle");
  int rc = batch_flush(b);
  if (rc) return rc;
  std::vector<gpugo_batch_t> lv{b};
  lv.insert(lv.end(), b->tiers.begin(), b->tiers.end());
  for (auto *x : lv) {
    rc = batch_collect_ready(x);
    if (rc) return rc;
  }
  int64_t n = 0;
  for (auto *x : lv) {
    if (x->running > 0 && x->pend_n < 2) {
      rc = batch_launch_one(x, false);
      if (rc) return rc;
      ++n;
    }
  }
  rc = batch_process_moves(b);
  if (rc) return rc;
  if (launched) *launched = n;
  return GPUGO_OK;
}

extern "C" int gpugo_batch_run(gpugo_batch_t b, int64_t launches) {
  if (!b) return set_err(GPUGO_E_INVALID, "null batch");
  if (b->level != 0) return set_err(GPUGO_E_INVALID, "not a batch handle");
  if (launches < 0) return set_err(GPUGO_E_INVALID, "negative launches");
  int rc = batch_flush(b);
  if (rc) return rc;
  for (int64_t i = 0; i < launches && total_running(b) > 0; ++i) {
    // every tier with problems in it advances by one launch on its own stream
    if (b->running > 0) {
      rc = batch_launch_one(b);
      if (rc) return rc;
    }
    for (auto *t : b->tiers) {
      if (t->running > 0) {
        rc = batch_launch_one(t);
        if (rc) return rc;
      }
    }
    rc = batch_process_moves(b);
    if (rc) return rc;
  }
  return GPUGO_OK;
}

extern "C" int gpugo_batch_sync(gpugo_batch_t b) {
  if (!b) return set_err(GPUGO_E_INVALID, "null batch");
  std::vector<gpugo_batch_t> lv{b};
  lv.insert(lv.end(), b->tiers.begin(), b->tiers.end());
  for (auto *x : lv) {
    while (x->pend_n > 0) {
      int rc = batch_collect_one(x);
      if (rc) return rc;
    }
    HIPCHECK(hipStreamSynchronize(x->stream));
  }
  return GPUGO_OK;
}

extern "C" int gpugo_batch_results(gpugo_batch_t b, int64_t max, int64_t *tags, gpugo_result *res, int64_t *n) {
  if (!b || !n) return set_err(GPUGO_E_INVALID, "null pointer");
  if (max < 0) return set_err(GPUGO_E_INVALID, "negative max");
  int64_t k = 0;
  while (k < max && b->done_head < b->done.size()) {
    if (tags) tags[k] = b->done[b->done_head].first;
    if (res) res[k] = b->done[b->done_head].second;
    ++k;
    ++b->done_head;
  }
  if (b->done_head == b->done.size()) {
    b->done.clear();
    b->done_head = 0;
  }
  *n = k;
  return GPUGO_OK;
}

extern "C" int gpugo_batch_counts(gpugo_batch_t b, int64_t *running, int64_t *free_slots, int64_t *pending) {
  if (!b) return set_err(GPUGO_E_INVALID, "null batch");
  if (running) *running = total_running(b);
  if (free_slots) *free_slots = (int64_t)b->free_list.size();
  if (pending) *pending = (int64_t)(b->done.size() - b->done_head);
  return GPUGO_OK;
}
