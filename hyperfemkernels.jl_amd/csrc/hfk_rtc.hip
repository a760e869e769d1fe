// Pointwise constitutive laws compiled at run time with hiprtc (the role of the generated PointForm of
// src/codegen.jl:166-198 and its compile cache keyed by signature, src/tree.jl:270-283): the host passes HIP
// source of a law -- a struct with
//   __device__ void setup(const double* F, double lam, double mu);    // F[c + 3J] = dx_c/dX_J
//   __device__ double stress(int c, int J) const;                      // first Piola-Kirchhoff P_cJ
//   __device__ double tangent(int c, int J, int k, int L) const;      // dP_cJ / dF_kL
// (the interface of the built-in laws) -- which is compiled into two pointwise kernels: the stored tangent
// D = dV iJt' T iJt in the operator's layout (prepare!), and the residual's point coefficients
// dV sum_J iJt[J,K] P_cJ (residual!).  The built-in kernels stage F, iJt, dV and mu per point and contract the
// coefficients.  Compiled code objects are cached per process by a hash of the full source.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstdint>
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "hfk_op.hpp"
#include "hfk_solver.hpp"

namespace hfk {
namespace {

const char* kPrelude = R"HFK(
__device__ inline double hfk_det3(const double* a) {
  return a[0] * (a[4] * a[8] - a[7] * a[5]) - a[3] * (a[1] * a[8] - a[7] * a[2]) + a[6] * (a[1] * a[5] - a[4] * a[2]);
}
__device__ inline void hfk_inv3(const double* a, double* r, double d) {
  const double id = 1.0 / d;
  r[0] = (a[4] * a[8] - a[7] * a[5]) * id;
  r[3] = -(a[3] * a[8] - a[6] * a[5]) * id;
  r[6] = (a[3] * a[7] - a[6] * a[4]) * id;
  r[1] = -(a[1] * a[8] - a[7] * a[2]) * id;
  r[4] = (a[0] * a[8] - a[6] * a[2]) * id;
  r[7] = -(a[0] * a[7] - a[6] * a[1]) * id;
  r[2] = (a[1] * a[5] - a[4] * a[2]) * id;
  r[5] = -(a[0] * a[5] - a[3] * a[2]) * id;
  r[8] = (a[0] * a[4] - a[3] * a[1]) * id;
}
)HFK";

// st[(batch * 20 + e) * nslot + slot]: F (0..8), iJt (9..17), dV (18), mu (19)
const char* kWrappers = R"HFK(
extern "C" __global__ void hfk_rtc_tangent(long long npts, int nslot, const double* __restrict__ st, double lam,
                                           void* __restrict__ out, int storage) {
  const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const long long batch = p / nslot;
  const int slot = (int)(p - batch * nslot);
  const double* s = st + batch * 20 * nslot + slot;
  double F[9], iJt[9];
  for (int e = 0; e < 9; ++e) {
    F[e] = s[(long long)e * nslot];
    iJt[e] = s[(long long)(9 + e) * nslot];
  }
  const double dV = s[18LL * nslot], mu = s[19LL * nslot];
  HFK_LAW law;
  law.setup(F, lam, mu);
  const bool sym = (storage & 1) != 0, f32 = storage >= 2;
  const int ne = sym ? 45 : 81;
  const long long base = batch * ne * nslot + slot;
  for (int k = 0; k < 3; ++k)
    for (int c = 0; c < 3; ++c) {
      if (sym && c > k) continue;
      double Tb[9], W[9];
      for (int L = 0; L < 3; ++L)
        for (int J = 0; J < 3; ++J) Tb[J + 3 * L] = law.tangent(c, J, k, L);
      for (int L = 0; L < 3; ++L)
        for (int K = 0; K < 3; ++K) {
          double v = 0;
          for (int J = 0; J < 3; ++J) v += iJt[J + 3 * K] * Tb[J + 3 * L];
          W[K + 3 * L] = v;
        }
      for (int M = 0; M < 3; ++M)
        for (int K = 0; K < 3; ++K) {
          double v = 0;
          for (int L = 0; L < 3; ++L) v += W[K + 3 * L] * iJt[L + 3 * M];
          const int i = K + 3 * c, j = M + 3 * k;
          long long idx;
          if (!sym) idx = base + (long long)(i + 9 * j) * nslot;
          else if (c < k || K <= M) idx = base + (long long)(j * (j + 1) / 2 + i) * nslot;
          else continue;
          if (f32) static_cast<float*>(out)[idx] = (float)(dV * v);
          else static_cast<double*>(out)[idx] = dV * v;
        }
    }
}

extern "C" __global__ void hfk_rtc_coefs(long long npts, int nslot, const double* __restrict__ st, double lam,
                                         double* __restrict__ coef) {
  const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const long long batch = p / nslot;
  const int slot = (int)(p - batch * nslot);
  const double* s = st + batch * 20 * nslot + slot;
  double F[9], iJt[9];
  for (int e = 0; e < 9; ++e) {
    F[e] = s[(long long)e * nslot];
    iJt[e] = s[(long long)(9 + e) * nslot];
  }
  const double dV = s[18LL * nslot], mu = s[19LL * nslot];
  HFK_LAW law;
  law.setup(F, lam, mu);
  for (int c = 0; c < 3; ++c) {
    double Pc[3];
    for (int J = 0; J < 3; ++J) Pc[J] = law.stress(c, J);
    for (int K = 0; K < 3; ++K) {
      double v = 0;
      for (int J = 0; J < 3; ++J) v += iJt[J + 3 * K] * Pc[J];
      coef[(batch * 9 + 3 * c + K) * nslot + slot] = dV * v;
    }
  }
}
)HFK";

uint64_t fnv1a(const std::string& s) {
  uint64_t h = 1469598103934665603ull;
  for (unsigned char ch : s) {
    h ^= ch;
    h *= 1099511628211ull;
  }
  return h;
}

std::mutex g_cache_mutex;
std::map<uint64_t, std::vector<char>> g_code_cache;  // compiled code objects by source hash

std::string full_source(const char* source, const char* name) {
  return std::string(kPrelude) + "\n#define HFK_LAW " + name + "\n" + source + "\n" + kWrappers;
}

// hiprtc for gfx950 (no device needed); false with the compiler log on failure
bool compile(const std::string& full, std::vector<char>& code, std::string& log) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, full.c_str(), "hfk_law.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
    log = "hiprtcCreateProgram failed";
    return false;
  }
  // packed FP32 instructions stay off, as in the library's own build
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult cr = hiprtcCompileProgram(prog, 6, opts);
  size_t ls = 0;
  hiprtcGetProgramLogSize(prog, &ls);
  log.assign(ls, '\0');
  if (ls) hiprtcGetProgramLog(prog, &log[0]);
  if (cr == HIPRTC_SUCCESS) {
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
  }
  hiprtcDestroyProgram(&prog);
  return cr == HIPRTC_SUCCESS;
}

bool valid_identifier(const char* s) {
  if (!s || !*s) return false;
  for (const char* p = s; *p; ++p) {
    const char ch = *p;
    const bool ok = (ch >= 'a' && ch <= 'z') || (ch >= 'A' && ch <= 'Z') || ch == '_' || (p != s && ch >= '0' && ch <= '9');
    if (!ok) return false;
  }
  return true;
}

}  // namespace
}  // namespace hfk

using namespace hfk;

struct hfk_law {
  hfk_ctx* ctx = nullptr;
  hipModule_t mod = nullptr;
  hipFunction_t ftan = nullptr, fcoef = nullptr;
  uint64_t key = 0;
  bool from_cache = false;
  ~hfk_law() {
    if (!mod) return;
    // launches of this module may still be queued on any stream of the device
    if (ctx && hipSetDevice(ctx->device) == hipSuccess) (void)hipDeviceSynchronize();
    (void)hipModuleUnload(mod);
  }
};

namespace hfk {

int rtc_state(hfk_op* op, const double* x, hipStream_t s) {
  const size_t n = (size_t)op->nbatches * op->C * op->np * kStateEntries;
  HIPCHK(op->ctx, op->rtc_state.alloc(n * sizeof(double)));
  HIPCHK(op->ctx, launch_point_state(op->dev(), x, op->rtc_state.as<double>(), BatchRange{0, (int)op->nbatches}, s));
  return HFK_OK;
}

int rtc_prepare(hfk_op* op, int storage, void* T, const double* x, hipStream_t s) {
  int rc = rtc_state(op, x, s);
  if (rc) return rc;
  long long npts = (long long)op->nbatches * op->C * op->np;
  int nslot = op->C * op->np;
  const double* st = op->rtc_state.as<double>();
  double lam = op->lam;
  void* out = T;
  int stg = storage;
  void* args[] = {&npts, &nslot, &st, &lam, &out, &stg};
  const unsigned grid = (unsigned)((npts + 255) / 256);
  HIPCHK(op->ctx, hipModuleLaunchKernel(op->rtc->ftan, grid, 1, 1, 256, 1, 1, 0, s, args, nullptr));
  return HFK_OK;
}

int rtc_residual(hfk_op* op, const double* x, double* b, hipStream_t s) {
  int rc = rtc_state(op, x, s);
  if (rc) return rc;
  long long npts = (long long)op->nbatches * op->C * op->np;
  int nslot = op->C * op->np;
  HIPCHK(op->ctx, op->rtc_coef.alloc((size_t)npts * 9 * sizeof(double)));
  const double* st = op->rtc_state.as<double>();
  double lam = op->lam;
  double* coef = op->rtc_coef.as<double>();
  void* args[] = {&npts, &nslot, &st, &lam, &coef};
  const unsigned grid = (unsigned)((npts + 255) / 256);
  HIPCHK(op->ctx, hipModuleLaunchKernel(op->rtc->fcoef, grid, 1, 1, 256, 1, 1, 0, s, args, nullptr));
  HIPCHK(op->ctx, launch_fill(b, op->nfree, 0.0, s));
  const OpDev d = op->dev();
  for (const BatchRange& r : op->color_batches) HIPCHK(op->ctx, launch_residual_coefs(d, coef, b, r, s));
  return HFK_OK;
}

}  // namespace hfk

extern "C" {

int hfk_law_create(hfk_ctx* ctx, const char* source, const char* name, hfk_law** out) {
  if (!ctx || !source || !out) return set_err(ctx, HFK_EARG, "HyperFEMKernels: null argument");
  if (!valid_identifier(name)) return set_err(ctx, HFK_EARG, "HyperFEMKernels: the law name must be a C identifier");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const std::string full = full_source(source, name);
  const uint64_t key = fnv1a(full);
  std::vector<char> code;
  bool cached = false;
  {
    std::lock_guard<std::mutex> lk(g_cache_mutex);
    auto it = g_code_cache.find(key);
    if (it != g_code_cache.end()) {
      code = it->second;
      cached = true;
    }
  }
  if (!cached) {
    std::string log;
    if (!compile(full, code, log)) return set_err(ctx, HFK_EARG, "HyperFEMKernels: the law does not compile: " + log);
    std::lock_guard<std::mutex> lk(g_cache_mutex);
    g_code_cache[key] = code;
  }
  std::unique_ptr<hfk_law> law(new hfk_law());
  law->ctx = ctx;
  law->key = key;
  law->from_cache = cached;
  HIPCHK(ctx, hipModuleLoadData(&law->mod, code.data()));
  HIPCHK(ctx, hipModuleGetFunction(&law->ftan, law->mod, "hfk_rtc_tangent"));
  HIPCHK(ctx, hipModuleGetFunction(&law->fcoef, law->mod, "hfk_rtc_coefs"));
  *out = law.release();
  return HFK_OK;
}

int hfk_law_check(const char* source, const char* name, char* log, int64_t log_size) {
  if (log && log_size > 0) log[0] = '\0';
  if (!source) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null source");
  if (!valid_identifier(name)) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: the law name must be a C identifier");
  std::vector<char> code;
  std::string msg;
  const bool ok = compile(full_source(source, name), code, msg);
  if (log && log_size > 0) {
    const size_t n = std::min<size_t>(msg.size(), (size_t)log_size - 1);
    std::memcpy(log, msg.data(), n);
    log[n] = '\0';
  }
  if (!ok) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: the law does not compile: " + msg);
  return HFK_OK;
}

int hfk_law_destroy(hfk_law* law) {
  delete law;
  return HFK_OK;
}

int hfk_law_info(hfk_law* law, uint64_t* key, int* from_cache) {
  if (!law) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null law");
  if (key) *key = law->key;
  if (from_cache) *from_cache = law->from_cache ? 1 : 0;
  return HFK_OK;
}

int hfk_op_set_law(hfk_op* op, hfk_law* law) {
  if (!op) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null operator");
  if (law && law->ctx != op->ctx) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: the law belongs to another context");
  op->rtc = law;
  op->storage = -1;  // the stored tangents belong to the previous law
  return HFK_OK;
}

}  // extern "C"
