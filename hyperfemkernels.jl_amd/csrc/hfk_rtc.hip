// Pointwise constitutive laws compiled at run time with hiprtc (the role of the generated PointForm of
// src/codegen.jl:166-198 and of its reuse per structural signature, src/tree.jl:244-283).  The host passes HIP
// source of a law -- a struct with
//   __device__ void setup(const double* F, double lam, double mu);    // F[c + 3J] = dx_c/dX_J
//   __device__ double stress(int c, int J) const;                      // first Piola-Kirchhoff P_cJ
//   __device__ double tangent(int c, int J, int k, int L) const;      // dP_cJ / dF_kL
// (the interface of the built-in laws).  The library's own prepare! and residual! kernels (hfk_point.hpp, embedded
// in the library as text) are instantiated with that law, so a law given at run time runs in the same fused
// kernels as a built-in one: gather, interpolation, geometry and the law in registers, no point data through HBM.
// One module per (Q, P) of the operators that use the law, compiled at first use.  Code objects are cached per
// process by a hash of the full program text.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "hfk_op.hpp"

// kTypesSrc, kPointSrc, kOseenFormSrc: the text of hfk_types.hpp, hfk_point.hpp and hfk_oseen_form.hpp (generated at
// build time)
#include "hfk_embedded.inc"

namespace hfk {
namespace {

const char* kPrelude = R"HFK(#include "hfk_point.hpp"
__device__ inline double hfk_det3(const double* a) { return hfk::det3(a); }
__device__ inline void hfk_inv3(const double* a, double* r, double d) { hfk::inv3(a, r, d); }
)HFK";

uint64_t fnv1a(const std::string& s) {
  uint64_t h = 1469598103934665603ull;
  for (unsigned char ch : s) {
    h ^= ch;
    h *= 1099511628211ull;
  }
  return h;
}

std::string law_source(const char* source) { return std::string(kPrelude) + "\n" + source + "\n"; }

// a small kernel that calls every member of the law: compiles in a fraction of a second, so a law is
// validated when it is created
std::string probe_program(const char* source, const char* name) {
  return law_source(source) + "extern \"C\" __global__ void hfk_law_probe(const double* F, double* out) {\n  " + name +
         " law;\n  law.setup(F, F[9], F[10]);\n  out[0] = law.stress(0, 0) + law.tangent(0, 0, 0, 0);\n}\n";
}

// a function of the physical coordinates, `double name(const double* x, const double* p)`, and the kernel that
// materialises it at the quadrature points of every cell of the operator's batches into mu_pts' layout
// [batch][cell*np + q]; padding cells (ids 0) take the constant `fill`
std::string coord_program(const char* source, const char* name) {
  return law_source(source) +
         "extern \"C\" __global__ void __launch_bounds__(256) hfk_coord_fill(hfk::OpDev op, const double* p, "
         "double fill, double* out) {\n"
         "  const int ns = op.C * op.np;\n"
         "  const int tid = threadIdx.x;\n"
         "  if (tid >= ns) return;\n"
         "  const size_t slot = (size_t)blockIdx.x * op.C + tid / op.np;\n"
         "  double v = fill;\n"
         "  if (op.ids[slot * op.nd] != 0) {\n"
         "    double x[3];\n"
         "    hfk::point_coordinates(op.tab, op.coords + slot * 24, tid % op.np, op.P, x);\n"
         "    v = " + std::string(name) + "(x, p);\n"
         "  }\n"
         "  out[(size_t)blockIdx.x * ns + tid] = v;\n"
         "}\n";
}

std::vector<std::string> kernel_names(const std::string& name, int Q, int P) {
  const std::string qp = std::to_string(Q) + ", " + std::to_string(P) + ", " + name;
  return {"hfk::k_prepare<" + qp + ", double, 0>", "hfk::k_prepare<" + qp + ", double, 1>",
          "hfk::k_prepare<" + qp + ", float, 0>", "hfk::k_prepare<" + qp + ", float, 1>",
          "hfk::k_residual<" + qp + ">", "hfk::k_energy<" + qp + ">"};
}

// the module of a law's kernels: its source and whether it defines an energy (a global the host reads)
std::string law_module(const char* source, const std::string& name) {
  return law_source(source) + "extern \"C\" __device__ int hfk_law_has_energy = hfk::HasEnergy<" + name + ">::value;\n";
}

// a pointwise function `double name(const double* Gu, const double* Gdu, const double* p)` wrapped for
// k_extremum, with a probe kernel that calls it
std::string pointfn_program(const char* source, const char* name) {
  return law_source(source) + "struct HfkPointFn {\n  __device__ static double eval(const double* a, const double* b, "
         "const double* p) { return " + std::string(name) + "(a, b, p); }\n};\n"
         "extern \"C\" __global__ void hfk_pointfn_probe(const double* g, double* out) { out[0] = "
         "HfkPointFn::eval(g, g + 9, g + 18); }\n";
}

std::vector<std::string> extremum_names(int Q, int P) {
  const std::string qp = std::to_string(Q) + ", " + std::to_string(P) + ", HfkPointFn, ";
  return {"hfk::k_extremum<" + qp + "1>", "hfk::k_extremum<" + qp + "2>"};
}

// hiprtc for gfx950 (no device needed); false with the compiler log on failure.  `names`: template
// instantiations to emit, their lowered names returned in `lowered`.
bool compile(const std::string& program, const std::vector<std::string>& names, std::vector<char>& code,
             std::vector<std::string>& lowered, std::string& log) {
  const char* hdr_src[] = {kTypesSrc, kPointSrc, kOseenFormSrc, kSkelSrc};
  const char* hdr_name[] = {"hfk_types.hpp", "hfk_point.hpp", "hfk_oseen_form.hpp", "hfk_skel.hpp"};
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, program.c_str(), "hfk_law.hip", 4, hdr_src, hdr_name) != HIPRTC_SUCCESS) {
    log = "hiprtcCreateProgram failed";
    return false;
  }
  for (const std::string& n : names) hiprtcAddNameExpression(prog, n.c_str());
  // the library's own flags: packed FP32 instructions stay off
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops"};
  const hiprtcResult cr = hiprtcCompileProgram(prog, 7, opts);
  size_t ls = 0;
  hiprtcGetProgramLogSize(prog, &ls);
  log.assign(ls, '\0');
  if (ls) hiprtcGetProgramLog(prog, &log[0]);
  bool ok = cr == HIPRTC_SUCCESS;
  if (ok) {
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    lowered.clear();
    for (const std::string& n : names) {
      const char* l = nullptr;
      if (hiprtcGetLoweredName(prog, n.c_str(), &l) != HIPRTC_SUCCESS || !l) {
        log = "no lowered name for " + n;
        ok = false;
        break;
      }
      lowered.push_back(l);
    }
  }
  hiprtcDestroyProgram(&prog);
  return ok;
}

// compiled code objects with their lowered kernel names, by a hash of the program and the instantiations
struct CacheEntry {
  std::vector<char> code;
  std::vector<std::string> lowered;
};
std::mutex g_cache_mutex;
std::map<uint64_t, CacheEntry> g_cache;

bool compile_cached(const std::string& program, const std::vector<std::string>& names, CacheEntry& out, bool& cached,
                    std::string& log) {
  std::string key_text = program;
  for (const std::string& n : names) key_text += "\n//" + n;
  const uint64_t key = fnv1a(key_text);
  {
    std::lock_guard<std::mutex> lk(g_cache_mutex);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) {
      out = it->second;
      cached = true;
      return true;
    }
  }
  cached = false;
  if (!compile(program, names, out.code, out.lowered, log)) return false;
  std::lock_guard<std::mutex> lk(g_cache_mutex);
  g_cache[key] = out;
  return true;
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

// the kernels of one (Q, P): prepare! per storage 0..3, residual!, the energy
struct RtcInstance {
  hipModule_t mod = nullptr;
  hipFunction_t prep[4] = {};
  hipFunction_t res = nullptr;
  hipFunction_t energy = nullptr;
  int has_energy = 0;
};

// the extremum kernels (min, max) of a pointwise function for one (Q, P)
struct FnInstance {
  hipModule_t mod = nullptr;
  hipFunction_t ext[2] = {};
};

struct hfk_pointfn {
  hfk_ctx* ctx = nullptr;
  std::string source, name;
  std::mutex m;
  std::map<int, FnInstance> inst;  // by Q * 16 + P
  ~hfk_pointfn() {
    if (inst.empty()) return;
    if (ctx && hipSetDevice(ctx->device) == hipSuccess) (void)hipDeviceSynchronize();
    for (auto& kv : inst) (void)hipModuleUnload(kv.second.mod);
  }
};

struct hfk_law {
  hfk_ctx* ctx = nullptr;
  std::string source, name;
  uint64_t key = 0;
  bool from_cache = false;
  std::mutex m;
  std::map<int, RtcInstance> inst;  // by Q * 16 + P
  ~hfk_law() {
    if (inst.empty()) return;
    // launches of these modules may still be queued on any stream of the device
    if (ctx && hipSetDevice(ctx->device) == hipSuccess) (void)hipDeviceSynchronize();
    for (auto& kv : inst) (void)hipModuleUnload(kv.second.mod);
  }
};

namespace hfk {
namespace {

int law_instance(hfk_op* op, RtcInstance** out) {
  hfk_law* law = op->rtc;
  std::lock_guard<std::mutex> lk(law->m);
  const int k = op->Q * 16 + op->P;
  auto it = law->inst.find(k);
  if (it != law->inst.end()) {
    *out = &it->second;
    return HFK_OK;
  }
  CacheEntry e;
  bool cached = false;
  std::string log;
  if (!compile_cached(law_module(law->source.c_str(), law->name), kernel_names(law->name, op->Q, op->P), e, cached, log))
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: the law's kernels do not compile: " + log);
  RtcInstance ri;
  HIPCHK(op->ctx, hipModuleLoadData(&ri.mod, e.code.data()));
  for (int s = 0; s < 4; ++s) HIPCHK(op->ctx, hipModuleGetFunction(&ri.prep[s], ri.mod, e.lowered[s].c_str()));
  HIPCHK(op->ctx, hipModuleGetFunction(&ri.res, ri.mod, e.lowered[4].c_str()));
  HIPCHK(op->ctx, hipModuleGetFunction(&ri.energy, ri.mod, e.lowered[5].c_str()));
  hipDeviceptr_t flag = nullptr;
  size_t flag_bytes = 0;
  HIPCHK(op->ctx, hipModuleGetGlobal(&flag, &flag_bytes, ri.mod, "hfk_law_has_energy"));
  HIPCHK(op->ctx, hipMemcpyDtoH(&ri.has_energy, flag, sizeof(int)));
  *out = &(law->inst[k] = ri);
  return HFK_OK;
}

int fn_instance(hfk_op* op, hfk_pointfn* fn, FnInstance** out) {
  std::lock_guard<std::mutex> lk(fn->m);
  const int k = op->Q * 16 + op->P;
  auto it = fn->inst.find(k);
  if (it != fn->inst.end()) {
    *out = &it->second;
    return HFK_OK;
  }
  CacheEntry e;
  bool cached = false;
  std::string log;
  if (!compile_cached(pointfn_program(fn->source.c_str(), fn->name.c_str()), extremum_names(op->Q, op->P), e, cached, log))
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: the extremum kernels do not compile: " + log);
  FnInstance fi;
  HIPCHK(op->ctx, hipModuleLoadData(&fi.mod, e.code.data()));
  for (int m = 0; m < 2; ++m) HIPCHK(op->ctx, hipModuleGetFunction(&fi.ext[m], fi.mod, e.lowered[m].c_str()));
  *out = &(fn->inst[k] = fi);
  return HFK_OK;
}

}  // namespace

int rtc_prepare(hfk_op* op, int storage, void* T, const double* x, hipStream_t s) {
  if (storage >= 4)
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: the factorised tangent storage needs a built-in neo-Hookean or "
                                      "linear law; a compiled law stores full or symmetric tangents");
  RtcInstance* ri = nullptr;
  int rc = law_instance(op, &ri);
  if (rc) return rc;
  OpDev d = op->dev();
  int b0 = 0;
  void* args[] = {&d, &T, &x, &b0};
  HIPCHK(op->ctx, hipModuleLaunchKernel(ri->prep[storage], (unsigned)op->nbatches, 1, 1, kThreads, 1, 1, 0, s, args,
                                        nullptr));
  return HFK_OK;
}

int rtc_residual(hfk_op* op, const double* x, double* b, hipStream_t s) {
  RtcInstance* ri = nullptr;
  int rc = law_instance(op, &ri);
  if (rc) return rc;
  HIPCHK(op->ctx, launch_fill(b, op->nfree, 0.0, s));
  OpDev d = op->dev();
  for (const BatchRange& r : op->color_batches) {
    int b0 = r.first;
    void* args[] = {&d, &x, &b, &b0};
    HIPCHK(op->ctx, hipModuleLaunchKernel(ri->res, (unsigned)r.count, 1, 1, kThreads, 1, 1, 0, s, args, nullptr));
  }
  return HFK_OK;
}

bool rtc_law_loaded(hfk_op* op) {
  hfk_law* law = op->rtc;
  std::lock_guard<std::mutex> lk(law->m);
  return law->inst.count(op->Q * 16 + op->P) > 0;
}

std::string oseen_form_program(const char* source) {
  return std::string("#include \"hfk_oseen_form.hpp\"\n") + source + "\n";
}

std::vector<std::string> oseen_form_names(const std::string& n) {
  return {"hfk::k_oseen_form_jac<" + n + ", false>", "hfk::k_oseen_form_jac<" + n + ", true>",
          "hfk::k_oseen_form_res<" + n + ", false>", "hfk::k_oseen_form_res<" + n + ", true>"};
}

int rtc_oseen_form(hfk_ctx* ctx, const char* source, const char* name, OseenFormKernels* out) {
  if (!source || !valid_identifier(name))
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: a point form needs its source and a C identifier as its name");
  const std::string program = oseen_form_program(source);
  const std::vector<std::string> names = oseen_form_names(name);
  CacheEntry e;
  bool cached = false;
  std::string log;
  if (!compile_cached(program, names, e, cached, log))
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: the point form does not compile: " + log);
  HIPCHK(ctx, hipModuleLoadData(&out->mod, e.code.data()));
  for (int a = 0; a < 2; ++a) {
    HIPCHK(ctx, hipModuleGetFunction(&out->jac[a], out->mod, e.lowered[a].c_str()));
    HIPCHK(ctx, hipModuleGetFunction(&out->res[a], out->mod, e.lowered[2 + a].c_str()));
  }
  return HFK_OK;
}

// the program and kernel instantiations of a face form with ns sides (2: skeleton, 1: boundary) for (Q, P); false
// when the name is not an identifier
bool skel_program(int Q, int P, int ns, const char* source, const char* name, std::string& program,
                  std::vector<std::string>& names) {
  // the built-in forms: the gradient-jump penalty (skeleton), the follower pressure (boundary)
  std::string form = ns == 2 ? "hfk::HfkGradJump" : "hfk::HfkFollowerPressure";
  if (source) {
    if (!valid_identifier(name)) return false;
    form = name;
  }
  program = std::string("#include \"hfk_skel.hpp\"\n") + (source ? source : "") + "\n";
  const std::string qp = std::to_string(Q) + ", " + std::to_string(P) + ", " + form;
  const std::string n = std::to_string(ns);
  names = {"hfk::k_skel<" + qp + ", 0, " + n + ">", "hfk::k_skel<" + qp + ", 1, " + n + ">",
           "hfk::k_skel<" + qp + ", 2, " + n + ">", "hfk::k_skel_jac<" + qp + ", " + n + ">"};
  return true;
}

int rtc_skel(hfk_ctx* ctx, int Q, int P, int ns, const char* source, const char* name, SkelData* sk) {
  std::string program;
  std::vector<std::string> names;
  if (!skel_program(Q, P, ns, source, name, program, names))
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: a facet form needs its source and a C identifier as its name");
  CacheEntry e;
  bool cached = false;
  std::string log;
  if (!compile_cached(program, names, e, cached, log))
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: the facet form does not compile: " + log);
  HIPCHK(ctx, hipModuleLoadData(&sk->mod, e.code.data()));
  HIPCHK(ctx, hipModuleGetFunction(&sk->res, sk->mod, e.lowered[0].c_str()));
  HIPCHK(ctx, hipModuleGetFunction(&sk->jv, sk->mod, e.lowered[1].c_str()));
  HIPCHK(ctx, hipModuleGetFunction(&sk->energy, sk->mod, e.lowered[2].c_str()));
  HIPCHK(ctx, hipModuleGetFunction(&sk->jac, sk->mod, e.lowered[3].c_str()));
  return HFK_OK;
}

int rtc_energy(hfk_op* op, const double* x, double* part, hipStream_t s) {
  RtcInstance* ri = nullptr;
  int rc = law_instance(op, &ri);
  if (rc) return rc;
  if (!ri->has_energy)
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: the compiled law defines no `double energy() const`");
  OpDev d = op->dev();
  void* args[] = {&d, &x, &part};
  HIPCHK(op->ctx, hipModuleLaunchKernel(ri->energy, (unsigned)op->nbatches, 1, 1, kThreads, 1, 1, 0, s, args, nullptr));
  return HFK_OK;
}

}  // namespace hfk

extern "C" {

int hfk_pointfn_create(hfk_ctx* ctx, const char* source, const char* name, hfk_pointfn** out) {
  if (!ctx || !source || !out) return set_err(ctx, HFK_EARG, "HyperFEMKernels: null argument");
  if (!valid_identifier(name)) return set_err(ctx, HFK_EARG, "HyperFEMKernels: the function name must be a C identifier");
  CacheEntry e;
  bool cached = false;
  std::string log;
  if (!compile_cached(pointfn_program(source, name), {}, e, cached, log))
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: the pointwise function does not compile: " + log);
  std::unique_ptr<hfk_pointfn> fn(new hfk_pointfn());
  fn->ctx = ctx;
  fn->source = source;
  fn->name = name;
  *out = fn.release();
  return HFK_OK;
}

int hfk_pointfn_destroy(hfk_pointfn* fn) {
  delete fn;
  return HFK_OK;
}

int hfk_extremum(hfk_op* op, hfk_pointfn* fn, const double* x, const double* dx, const double* dx_dir,
                 const double* params, int64_t nparams, int reduce, double* value) {
  if (!op || !op->ctx) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null operator");
  if (!fn || !x || !dx || !value || (nparams > 0 && !params) || nparams < 0)
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null argument");
  if (fn->ctx != op->ctx) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: the function belongs to another context");
  if (reduce != 0 && reduce != 1) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: reduce must be 0 (min) or 1 (max)");
  HIPCHK(op->ctx, hipSetDevice(op->ctx->device));
  FnInstance* fi = nullptr;
  int rc = fn_instance(op, fn, &fi);
  if (rc) return rc;
  hipStream_t s = op->ctx->stream;
  const size_t nf = (size_t)op->nfree, nd = dx_dir ? (size_t)op->ndir : 0, np = (size_t)std::max<int64_t>(nparams, 0);
  // staging: x, dx, dx_dir, params, result
  DevBuf buf;
  HIPCHK(op->ctx, buf.alloc((2 * nf + nd + np + 1) * sizeof(double)));
  double* xd = buf.as<double>();
  double* dxd = xd + nf;
  double* ddd = dxd + nf;
  double* pd = ddd + nd;
  double* res = pd + np;
  HIPCHK(op->ctx, hipMemcpyAsync(xd, x, nf * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(op->ctx, hipMemcpyAsync(dxd, dx, nf * sizeof(double), hipMemcpyHostToDevice, s));
  if (nd) HIPCHK(op->ctx, hipMemcpyAsync(ddd, dx_dir, nd * sizeof(double), hipMemcpyHostToDevice, s));
  if (np) HIPCHK(op->ctx, hipMemcpyAsync(pd, params, np * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(op->ctx, op->part.alloc((size_t)std::max<int64_t>(op->nbatches, 1) * sizeof(double)));
  OpDev d = op->dev();
  const double* xa = xd;
  const double* dxa = dxd;
  const double* dda = nd ? ddd : nullptr;
  const double* pa = pd;
  double* part = op->part.as<double>();
  void* args[] = {&d, &xa, &dxa, &dda, &pa, &part};
  HIPCHK(op->ctx, hipModuleLaunchKernel(fi->ext[reduce], (unsigned)op->nbatches, 1, 1, kThreads, 1, 1, 0, s, args, nullptr));
  HIPCHK(op->ctx, launch_reduce_partials(part, op->nbatches, reduce + 1, res, s));
  HIPCHK(op->ctx, hipMemcpyAsync(value, res, sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(op->ctx, hipStreamSynchronize(s));
  return HFK_OK;
}

int hfk_law_create(hfk_ctx* ctx, const char* source, const char* name, hfk_law** out) {
  if (!ctx || !source || !out) return set_err(ctx, HFK_EARG, "HyperFEMKernels: null argument");
  if (!valid_identifier(name)) return set_err(ctx, HFK_EARG, "HyperFEMKernels: the law name must be a C identifier");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  CacheEntry e;
  bool cached = false;
  std::string log;
  const std::string probe = probe_program(source, name);
  if (!compile_cached(probe, {}, e, cached, log))
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: the law does not compile: " + log);
  std::unique_ptr<hfk_law> law(new hfk_law());
  law->ctx = ctx;
  law->source = source;
  law->name = name;
  law->key = fnv1a(probe);
  law->from_cache = cached;
  *out = law.release();
  return HFK_OK;
}

int hfk_law_check(const char* source, const char* name, char* log, int64_t log_size) {
  if (log && log_size > 0) log[0] = '\0';
  if (!source) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null source");
  if (!valid_identifier(name)) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: the law name must be a C identifier");
  std::vector<char> code;
  std::vector<std::string> lowered;
  std::string msg;
  const bool ok = compile(probe_program(source, name), {}, code, lowered, msg);
  if (log && log_size > 0) {
    const size_t n = std::min<size_t>(msg.size(), (size_t)log_size - 1);
    std::memcpy(log, msg.data(), n);
    log[n] = '\0';
  }
  if (!ok) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: the law does not compile: " + msg);
  return HFK_OK;
}

int hfk_oseen_form_check(const char* source, const char* name, char* log, int64_t log_size) {
  if (!source || !valid_identifier(name)) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: invalid form source or name");
  CacheEntry e;
  bool cached = false;
  std::string msg;
  const bool ok = compile_cached(oseen_form_program(source), oseen_form_names(name), e, cached, msg);
  if (log && log_size > 0) {
    const size_t n = std::min<size_t>(msg.size(), (size_t)log_size - 1);
    std::memcpy(log, msg.data(), n);
    log[n] = '\0';
  }
  return ok ? HFK_OK : set_err(nullptr, HFK_EARG, "HyperFEMKernels: the point form does not compile: " + msg);
}

int hfk_skel_form_check(int Q, int P, int sides, const char* source, const char* name, char* log, int64_t log_size) {
  std::string program, msg;
  std::vector<std::string> names;
  if (Q < 2 || Q > 4 || P < 2 || P > 4 || (sides != 1 && sides != 2) ||
      !skel_program(Q, P, sides, source, name, program, names))
    return set_err(nullptr, HFK_EARG, "HyperFEMKernels: invalid (Q, P) or facet form name");
  CacheEntry e;
  bool cached = false;
  const bool ok = compile_cached(program, names, e, cached, msg);
  if (log && log_size > 0) {
    const size_t n = std::min<size_t>(msg.size(), (size_t)log_size - 1);
    std::memcpy(log, msg.data(), n);
    log[n] = '\0';
  }
  return ok ? HFK_OK : set_err(nullptr, HFK_EARG, "HyperFEMKernels: the facet form does not compile: " + msg);
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

int hfk_op_set_mu_function(hfk_op* op, const char* source, const char* name, const double* params, int64_t nparams) {
  if (!op || !op->ctx) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null operator");
  if (!source) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: null source");
  if (!valid_identifier(name)) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: the function name must be a C identifier");
  if (nparams < 0 || (nparams > 0 && !params)) return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: invalid parameters");
  HIPCHK(op->ctx, hipSetDevice(op->ctx->device));
  CacheEntry e;
  bool cached = false;
  std::string log;
  if (!compile_cached(coord_program(source, name), {}, e, cached, log))
    return set_err(op->ctx, HFK_EARG, "HyperFEMKernels: the function of x does not compile: " + log);
  hipStream_t s = op->ctx->stream;
  DevBuf p;
  HIPCHK(op->ctx, p.alloc((size_t)std::max<int64_t>(nparams, 1) * sizeof(double)));
  if (nparams > 0) HIPCHK(op->ctx, hipMemcpyAsync(p.p, params, (size_t)nparams * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(op->ctx, op->mu_pts.alloc((size_t)op->nbatches * op->C * op->np * sizeof(double)));
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
  HIPCHK(op->ctx, hipModuleLoadData(&mod, e.code.data()));
  hipError_t err = hipModuleGetFunction(&fn, mod, "hfk_coord_fill");
  if (err == hipSuccess) {
    OpDev d = op->dev();
    double* pp = p.as<double>();
    double fill = op->mu;
    double* out = op->mu_pts.as<double>();
    void* args[] = {&d, &pp, &fill, &out};
    err = hipModuleLaunchKernel(fn, (unsigned)op->nbatches, 1, 1, kThreads, 1, 1, 0, s, args, nullptr);
  }
  if (err == hipSuccess) err = hipStreamSynchronize(s);
  (void)hipModuleUnload(mod);
  HIPCHK(op->ctx, err);
  op->has_mu_field = true;
  return HFK_OK;
}

int hfk_op_mu_field(hfk_op* op, double* mu_points) {
  if (!op || !op->ctx || !mu_points) return set_err(op ? op->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  HIPCHK(op->ctx, hipSetDevice(op->ctx->device));
  const size_t nslot = (size_t)op->nbatches * op->C;
  if (!op->has_mu_field) {
    std::fill(mu_points, mu_points + (size_t)op->ncells * op->np, op->mu);
    return HFK_OK;
  }
  std::vector<double> h(nslot * op->np);
  HIPCHK(op->ctx, hipStreamSynchronize(op->ctx->stream));
  HIPCHK(op->ctx, hipMemcpy(h.data(), op->mu_pts.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t sl = 0; sl < nslot; ++sl) {
    const int64_t cell = op->batch_cell[sl];
    if (cell >= 0) std::memcpy(&mu_points[(size_t)cell * op->np], &h[sl * op->np], op->np * sizeof(double));
  }
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
