// Learned coarse correction: a 4-level 3D convolutional U-Net on the vertex lattice of the assembled Q1 level,
// BF16 weights and activations on MFMA with FP32 accumulation, trained on the device by the kernels below
// (forward, backward data, backward weights, Adam).  It approximates the coarse inverse: z = (s/h) N(r/s) with
// s the RMS of r and h the cell size, and is used inside the V-cycle as the coarse correction (alone, or as the
// initial guess of the coarse PCG).  The Krylov tolerance, not the network, fixes the answer.
//
// Tensors are channels-last on a lattice of N^3 vertices (x fastest): BF16 activations, FP32 gradients, each a
// view (pointer, channel count, channel stride) so that the concatenations of the up path are read in place.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "hfk_net.hpp"
#include "hfk_op.hpp"
#include "hfk_solver.hpp"

using namespace hfk;

namespace hfk {
namespace {

typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

struct View {  // channels-last tensor view
  void* p;
  int C;       // channels of the view
  int stride;  // channels per voxel of the storage
};

__device__ __forceinline__ bool neighbour(int v, int tap, int N, int& nb) {
  const int x = v % N, y = (v / N) % N, z = v / (N * N);
  const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
  const int xx = x + dx, yy = y + dy, zz = z + dz;
  if (xx < 0 || yy < 0 || zz < 0 || xx >= N || yy >= N || zz >= N) return false;
  nb = xx + N * (yy + N * zz);
  return true;
}

// 8 consecutive BF16 channels (16 bytes; channel offsets and strides are multiples of 8)
__device__ __forceinline__ bf16x8 load8(const unsigned short* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const s16x8*>(p));
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, s16x8{0, 0, 0, 0, 0, 0, 0, 0}); }

// Y[v][co] = act( sum_tap sum_ci X[v + off(tap)][ci] W[tap][co][ci] ), X = [X1 | X2] (zero outside the lattice).
// Implicit GEMM on gfx950's v_mfma_f32_16x16x32_bf16: M = voxels, N = output channels, K = 27 x input channels
// in steps of 32 (a 16-channel view fills the lower half of K, the upper half is zero).  One work-group computes 16
// voxels x all output channels (NT tiles of 16); its 4 waves split the 27 taps (wave w takes taps w, w + 4, ...)
// and their partial sums are added through LDS in wave order.  On the small lattices of the coarse levels
// (5^3 .. 33^3 voxels) this gives 4x the waves of a one-wave-per-tile mapping and a quarter of the dependent MFMA
// chain.  Lane l: A row m = l % 16 (a voxel) with k = 8 (l / 16) + i, B column n = l % 16 (an output channel) with
// the same k (A and B share the k map, so the product does not depend on the hardware's order of k); D rows
// 4 (l / 16) + i, column l % 16.
// OUT: 0 = BF16 with ReLU, 1 = FP32 (overwrite), 2 = FP32 accumulate.
template <int NT, int OUT>
__global__ __launch_bounds__(256) void k_conv(int N, View X1, View X2, const unsigned short* __restrict__ W, View Y) {
  __shared__ f32x4 part[4][NT][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nvox = N * N * N;
  const int v0 = blockIdx.x * 16;  // the grid has ceil(nvox / 16) work-groups: v0 < nvox
  const int m = lane & 15, kg = lane >> 4;
  const int vm = v0 + m;
  const int Cin = X1.C + (X2.p ? X2.C : 0);
  const int Cout = NT * 16;
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int tap = wave; tap < 27; tap += 4) {
    int nb = 0;
    const bool in = vm < nvox && neighbour(vm, tap, N, nb);
    const unsigned short* wt = W + (size_t)tap * Cout * Cin;
    for (int part_ = 0; part_ < 2; ++part_) {
      const View& X = part_ == 0 ? X1 : X2;
      if (!X.p) continue;
      const int cbase = part_ == 0 ? 0 : X1.C;
      const unsigned short* xp = static_cast<const unsigned short*>(X.p) + (size_t)nb * X.stride + 8 * kg;
      for (int c0 = 0; c0 < X.C; c0 += 32) {
        const bool kin = c0 + 8 * kg < X.C;  // false for the upper half of K past a 16-channel remainder
        const bf16x8 a = in && kin ? load8(xp + c0) : zero8();
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const bf16x8 b = kin ? load8(wt + (size_t)(t * 16 + m) * Cin + cbase + c0 + 8 * kg) : zero8();
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) part[wave][t][lane] = acc[t];
  __syncthreads();
  for (int t = wave; t < NT; t += 4) {
    const f32x4 sum = ((part[0][t][lane] + part[1][t][lane]) + part[2][t][lane]) + part[3][t][lane];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int v = v0 + 4 * kg + i;
      if (v >= nvox) continue;
      const int co = t * 16 + m;
      const float val = sum[i];
      if (OUT == 0) {
        static_cast<unsigned short*>(Y.p)[(size_t)v * Y.stride + co] = f2bf(val > 0.f ? val : 0.f);
      } else if (OUT == 1) {
        static_cast<float*>(Y.p)[(size_t)v * Y.stride + co] = val;
      } else {
        static_cast<float*>(Y.p)[(size_t)v * Y.stride + co] += val;
      }
    }
  }
}

// dW[tap][co][ci] += sum_v G[v][co] X[v + off(tap)][ci]: per (tap, co tile, ci tile) a GEMM over the voxels
// (K = voxels, 32 per v_mfma_f32_16x16x32_bf16).  A (M = co, K = v): lane m = l % 16 -> co, k = 8 (l / 16) + i ->
// voxel; B (K = v, N = ci) with the same k map.
// Each wave reduces a chunk of kVoxWave voxels and writes its 16 x 16 tile to P[tile][chunk][256];
// k_wgrad_reduce then adds the chunks of every weight in chunk order, so the gradient does not depend on the order
// in which the waves finish (training is reproducible).
constexpr int kVoxWave = 512;
__global__ __launch_bounds__(256) void k_conv_wgrad(int N, View G, View X1, View X2, float* __restrict__ P,
                                                    int ntco, int ntci) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nvox = N * N * N;
  const int tile = blockIdx.y;
  const int tap = tile / (ntco * ntci), r = tile % (ntco * ntci), tco = r / ntci, tci = r % ntci;
  const int vstart = (blockIdx.x * 4 + wave) * kVoxWave;
  if (vstart >= nvox) return;
  const int m = lane & 15, kg = lane >> 4;
  const int Cin = X1.C + (X2.p ? X2.C : 0);
  const int ci = tci * 16 + m;
  const View& X = ci < X1.C ? X1 : X2;
  const int cx = ci < X1.C ? ci : ci - X1.C;
  const int co = tco * 16 + m;
  const unsigned short* g = static_cast<const unsigned short*>(G.p);
  const unsigned short* x = static_cast<const unsigned short*>(X.p);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const int vend = min(vstart + kVoxWave, nvox);
  for (int vb = vstart; vb < vend; vb += 32) {
    s16x8 a, b;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int v = vb + 8 * kg + i;
      int nb = 0;
      const bool ok = v < vend;
      a[i] = ok ? (short)g[(size_t)v * G.stride + co] : (short)0;
      b[i] = (ok && neighbour(v, tap, N, nb)) ? (short)x[(size_t)nb * X.stride + cx] : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), acc,
                                                   0, 0, 0);
  }
  // D rows 4 (l / 16) + i -> co, column l % 16 -> ci
  const int nchunks = (nvox + kVoxWave - 1) / kVoxWave;
  float* pt = P + ((size_t)tile * nchunks + (blockIdx.x * 4 + wave)) * 256;
#pragma unroll
  for (int i = 0; i < 4; ++i) pt[(4 * kg + i) * 16 + m] = acc[i];
}

// dW[tap][co][ci] += sum over the chunks, in order, of P[tile][chunk][row * 16 + col]; one thread per weight
__global__ __launch_bounds__(256) void k_wgrad_reduce(const float* __restrict__ P, int nchunks, int ntiles, int ntco,
                                                      int ntci, int Cin, float* __restrict__ dW) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  if (g >= ntiles * 256) return;
  const int tile = g >> 8, e = g & 255;
  const int tap = tile / (ntco * ntci), r = tile % (ntco * ntci), tco = r / ntci, tci = r % ntci;
  const float* pt = P + (size_t)tile * nchunks * 256 + e;
  float sum = 0.f;
  for (int c = 0; c < nchunks; ++c) sum += pt[(size_t)c * 256];
  const int row = e >> 4, col = e & 15;
  dW[((size_t)tap * (ntco * 16) + tco * 16 + row) * Cin + tci * 16 + col] += sum;
}

// Wt[tap][ci][co] = W[26 - tap][co][ci] (the adjoint of the zero-padded correlation)
__global__ void k_weight_adjoint(const unsigned short* __restrict__ W, unsigned short* __restrict__ Wt, int Cout, int Cin) {
  const int n = 27 * Cout * Cin;
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < n; g += gridDim.x * blockDim.x) {
    const int tap = g / (Cout * Cin), r = g % (Cout * Cin), co = r / Cin, ci = r % Cin;
    Wt[((size_t)(26 - tap) * Cin + ci) * Cout + co] = W[g];
  }
}

// G = bf16(dY * (A > 0)) (relu: true) or bf16(dY)
__global__ void k_relu_back(int64_t nvox, View dY, View A, View G, int relu) {
  const int64_t n = nvox * G.C;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = g / G.C;
    const int c = (int)(g % G.C);
    float d = static_cast<const float*>(dY.p)[v * dY.stride + c];
    if (relu && !(bf2f(static_cast<const unsigned short*>(A.p)[v * A.stride + c]) > 0.f)) d = 0.f;
    static_cast<unsigned short*>(G.p)[v * G.stride + c] = f2bf(d);
  }
}

// lattice transfers between N (fine, odd) and Nc = (N + 1) / 2: coarse node i sits on fine node 2i.
// restriction (full weighting, weights 1/4, 1/2, 1/4 per direction) and trilinear prolongation, BF16 forward;
// their adjoints in FP32 for the backward pass.
__device__ __forceinline__ float w_r(int d) { return d == 0 ? 0.5f : 0.25f; }
__device__ __forceinline__ float w_p(int d) { return d == 0 ? 1.0f : 0.5f; }

__global__ void k_restrict_bf(int Nf, View F, View Cv) {
  const int Nc = (Nf + 1) / 2;
  const int64_t n = (int64_t)Nc * Nc * Nc * Cv.C;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t vc = g / Cv.C;
    const int c = (int)(g % Cv.C);
    const int x = (int)(vc % Nc), y = (int)((vc / Nc) % Nc), z = (int)(vc / ((int64_t)Nc * Nc));
    float s = 0.f;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int fx = 2 * x + dx, fy = 2 * y + dy, fz = 2 * z + dz;
          if (fx < 0 || fy < 0 || fz < 0 || fx >= Nf || fy >= Nf || fz >= Nf) continue;
          const int64_t vf = fx + (int64_t)Nf * (fy + (int64_t)Nf * fz);
          s += w_r(dx) * w_r(dy) * w_r(dz) * bf2f(static_cast<const unsigned short*>(F.p)[vf * F.stride + c]);
        }
    static_cast<unsigned short*>(Cv.p)[vc * Cv.stride + c] = f2bf(s);
  }
}

// fine += R' coarse (FP32)
__global__ void k_restrict_T(int Nf, View Cv, View F) {
  const int Nc = (Nf + 1) / 2;
  const int64_t n = (int64_t)Nf * Nf * Nf * F.C;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t vf = g / F.C;
    const int c = (int)(g % F.C);
    const int f[3] = {(int)(vf % Nf), (int)((vf / Nf) % Nf), (int)(vf / ((int64_t)Nf * Nf))};
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      lo[d] = f[d] / 2;                     // smallest i with 2i >= f - 1
      hi[d] = min((f[d] + 1) / 2, Nc - 1);  // largest i with 2i <= f + 1
    }
    float s = 0.f;
    for (int iz = lo[2]; iz <= hi[2]; ++iz)
      for (int iy = lo[1]; iy <= hi[1]; ++iy)
        for (int ix = lo[0]; ix <= hi[0]; ++ix) {
          const float w = w_r(f[0] - 2 * ix) * w_r(f[1] - 2 * iy) * w_r(f[2] - 2 * iz);
          s += w * static_cast<const float*>(Cv.p)[(ix + (int64_t)Nc * (iy + (int64_t)Nc * iz)) * Cv.stride + c];
        }
    static_cast<float*>(F.p)[vf * F.stride + c] += s;
  }
}

__device__ __forceinline__ float ld_act(const unsigned short* p, int64_t i) { return bf2f(p[i]); }
__device__ __forceinline__ float ld_act(const float* p, int64_t i) { return p[i]; }
__device__ __forceinline__ void st_act(unsigned short* p, int64_t i, float v) { p[i] = f2bf(v); }
__device__ __forceinline__ void st_act(float* p, int64_t i, float v) { p[i] = v; }

// fine = P coarse (T: BF16 or FP32): fine node f takes the coarse nodes i with |f - 2i| <= 1, weights 1 and 1/2
template <typename T>
__global__ void k_prolong(int Nf, View Cv, View F) {
  const int Nc = (Nf + 1) / 2;
  const int64_t n = (int64_t)Nf * Nf * Nf * F.C;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t vf = g / F.C;
    const int c = (int)(g % F.C);
    const int f[3] = {(int)(vf % Nf), (int)((vf / Nf) % Nf), (int)(vf / ((int64_t)Nf * Nf))};
    int lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      lo[d] = f[d] / 2;
      hi[d] = min((f[d] + 1) / 2, Nc - 1);
    }
    float s = 0.f;
    for (int iz = lo[2]; iz <= hi[2]; ++iz)
      for (int iy = lo[1]; iy <= hi[1]; ++iy)
        for (int ix = lo[0]; ix <= hi[0]; ++ix) {
          const float w = w_p(f[0] - 2 * ix) * w_p(f[1] - 2 * iy) * w_p(f[2] - 2 * iz);
          s += w * ld_act(static_cast<const T*>(Cv.p), (ix + (int64_t)Nc * (iy + (int64_t)Nc * iz)) * Cv.stride + c);
        }
    st_act(static_cast<T*>(F.p), vf * F.stride + c, s);
  }
}

// coarse = P' fine (FP32, overwrite)
__global__ void k_prolong_T(int Nf, View F, View Cv) {
  const int Nc = (Nf + 1) / 2;
  const int64_t n = (int64_t)Nc * Nc * Nc * Cv.C;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t vc = g / Cv.C;
    const int c = (int)(g % Cv.C);
    const int x = (int)(vc % Nc), y = (int)((vc / Nc) % Nc), z = (int)(vc / ((int64_t)Nc * Nc));
    float s = 0.f;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int fx = 2 * x + dx, fy = 2 * y + dy, fz = 2 * z + dz;
          if (fx < 0 || fy < 0 || fz < 0 || fx >= Nf || fy >= Nf || fz >= Nf) continue;
          const int64_t vf = fx + (int64_t)Nf * (fy + (int64_t)Nf * fz);
          s += w_p(dx) * w_p(dy) * w_p(dz) * static_cast<const float*>(F.p)[vf * F.stride + c];
        }
    static_cast<float*>(Cv.p)[vc * Cv.stride + c] = s;
  }
}

// input of the network: x[v][c] = r[dof(v, c)] / scale (c < 3), the auxiliary vertex field (c = 3, e.g. the
// normalised log shear modulus; 0 when absent), 0 elsewhere (16 channels)
__global__ void k_pack(int64_t nvox, const int32_t* __restrict__ node_dofs, const double* __restrict__ r,
                       const double* __restrict__ scale, const float* __restrict__ aux, unsigned short* __restrict__ X) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nvox * 16; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = g / 16;
    const int c = (int)(g % 16);
    float val = 0.f;
    if (c < 3) {
      const int id = node_dofs[v * 3 + c];
      if (id > 0) val = (float)(r[id - 1] / scale[0]);
    } else if (c == 3 && aux) {
      val = aux[v];
    }
    X[g] = f2bf(val);
  }
}

// w[dof] = out[v][c] (the network's output in its own units) on the free dofs
__global__ void k_out_to_vec(int64_t nvox, const int32_t* __restrict__ node_dofs, const float* __restrict__ out,
                             double* __restrict__ w) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nvox * 3; g += (int64_t)gridDim.x * blockDim.x) {
    const int id = node_dofs[g];
    if (id > 0) w[id - 1] = (double)out[(g / 3) * 16 + g % 3];
  }
}

// dOut[v][c] = gw[dof(v, c)] * scale_factor (c < 3, free), 0 elsewhere
__global__ void k_vec_to_gout(int64_t nvox, const int32_t* __restrict__ node_dofs, const double* __restrict__ gw,
                              double sc, float* __restrict__ dOut) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nvox * 16; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = g / 16;
    const int c = (int)(g % 16);
    float val = 0.f;
    if (c < 3) {
      const int id = node_dofs[v * 3 + c];
      if (id > 0) val = (float)(gw[id - 1] * sc);
    }
    dOut[g] = val;
  }
}

// z[dof] = out[v][c] * scale * factor on free dofs
__global__ void k_unpack(int64_t nvox, const int32_t* __restrict__ node_dofs, const float* __restrict__ out,
                         const double* __restrict__ scale, double factor, double* __restrict__ z) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nvox * 3; g += (int64_t)gridDim.x * blockDim.x) {
    const int id = node_dofs[g];
    if (id > 0) z[id - 1] = (double)out[(g / 3) * 16 + g % 3] * scale[0] * factor;
  }
}

// scale = sqrt(r.r / n)  (from a dot product already computed into s[0])
__global__ void k_rms(const double* __restrict__ rr, int64_t n, double* __restrict__ scale) {
  const double v = sqrt(rr[0] / (double)n);
  scale[0] = v > 0 ? v : 1.0;
}

// loss of one sample: sum over free (v, c < 3) of (out - t)^2 / sum t^2 with t = e / (scale * factor); gradient
// dOut = 2 (out - t) / sum t^2; partial sums per block (deterministic), the host adds them
__global__ __launch_bounds__(256) void k_loss(int64_t nvox, const int32_t* __restrict__ node_dofs,
                                              const float* __restrict__ out, const double* __restrict__ e,
                                              const double* __restrict__ scale, double factor,
                                              const double* __restrict__ tt, float* __restrict__ dOut,
                                              double* __restrict__ part) {
  __shared__ double sh[256];
  double s = 0;
  const double inv = 1.0 / (scale[0] * factor);
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < nvox * 16; g += (int64_t)gridDim.x * blockDim.x) {
    const int64_t v = g / 16;
    const int c = (int)(g % 16);
    float grad = 0.f;
    if (c < 3) {
      const int id = node_dofs[v * 3 + c];
      if (id > 0) {
        const double t = e[id - 1] * inv;
        const double d = (double)out[g] - t;
        s += d * d;
        grad = (float)(2.0 * d / tt[0]);
      }
    }
    dOut[g] = grad;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] / tt[0];
}

// Adam (Kingma & Ba 2015) on the FP32 master weights; BF16 copy refreshed; gradient cleared
__global__ void k_adam(int64_t n, float* __restrict__ w, float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, unsigned short* __restrict__ wb, float lr, float b1, float b2, float eps,
                       float c1, float c2, float gscale) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i] * gscale;
    m1[i] = b1 * m1[i] + (1.f - b1) * gi;
    m2[i] = b2 * m2[i] + (1.f - b2) * gi * gi;
    w[i] -= lr * (m1[i] / c1) / (sqrtf(m2[i] / c2) + eps);
    wb[i] = f2bf(w[i]);
    g[i] = 0.f;
  }
}

__global__ void k_to_bf(int64_t n, const float* __restrict__ w, unsigned short* __restrict__ wb) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    wb[i] = f2bf(w[i]);
}

inline dim3 grid1(int64_t n) {
  int64_t b = (n + 255) / 256;
  return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 8192)));
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------
// host side

namespace {

struct Layer {
  int level, cin1, cin2, cout;
  bool relu;
  size_t woff;  // offset of the layer's weights in the flat parameter vector
};

// The U-Net: channels 16, 32, 64, 64 on levels 0..3 (see the header of this file)
const Layer kLayers[] = {
    {0, 16, 0, 16, true, 0},  {0, 16, 0, 16, true, 0},  {1, 16, 0, 32, true, 0},  {1, 32, 0, 32, true, 0},
    {2, 32, 0, 64, true, 0},  {3, 64, 0, 64, true, 0},  {3, 64, 0, 64, true, 0},  {2, 64, 64, 64, true, 0},
    {1, 32, 64, 32, true, 0}, {0, 16, 32, 16, true, 0}, {0, 16, 0, 16, false, 0},
};
constexpr int kNL = 11;

}  // namespace
}  // namespace hfk

struct hfk_net {
  hfk_ctx* ctx = nullptr;
  int N[4] = {0, 0, 0, 0};
  int64_t nvox[4] = {0, 0, 0, 0};
  std::vector<hfk::Layer> layers;
  size_t nparam = 0;
  hfk::DevBuf w32, wbf, wt, grad, m1, m2;  // parameters (FP32 master, BF16), adjoint weights, gradient, Adam moments
  hfk::DevBuf node_dofs;                   // [v][3] signed coarse dof ids on the input lattice
  int64_t ncoarse = 0;
  // input lattice: the coarse space's vertices (Nin^3).  With coarsen = 1 the U-Net runs on every other vertex
  // (N[0] = (Nin + 1) / 2): its input is the full-weighting restriction of the packed input and its output is
  // prolongated trilinearly back to the input lattice; otherwise N[0] = Nin.
  int coarsen = 0, Nin = 0;
  int64_t nvin = 0;
  hfk::DevBuf xin, outf, g_outf;
  double h = 1.0;                          // cell size (output factor 1/h)
  // activations (BF16) and gradients (FP32)
  hfk::DevBuf x0, a0, s0, r1, a2, s1, r2, s2, r3, a5, a6, u2, a7, u1, a8, u0, a9, out;
  hfk::DevBuf g_out, g_a9, g_cat0, g_a8tmp, g_cat1, g_a7tmp, g_cat2, g_a6tmp, g_a5, g_r3, g_r2, g_a2, g_r1, g_a0, gbf;
  hfk::DevBuf scal, part, dots;
  hfk::DevBuf wpart;  // per-chunk weight-gradient tiles of one layer (k_conv_wgrad -> k_wgrad_reduce)
  hfk::DevBuf aux;  // auxiliary input channel per vertex (FP32)
  bool has_aux = false;
  int step = 0;
};

namespace hfk {
namespace {

// views
inline View vb(const DevBuf& b, int C, int stride) { return View{b.p, C, stride}; }
inline View vb(const DevBuf& b, int C, int stride, int off, bool f32) {
  return View{f32 ? (void*)(b.as<float>() + off) : (void*)(b.as<unsigned short>() + off), C, stride};
}
const View kNone{nullptr, 0, 0};

hipError_t conv(hfk_net* nt, int li, View X1, View X2, const unsigned short* W, View Y, int out, hipStream_t s) {
  const Layer& L = nt->layers[li];
  const int N = nt->N[L.level];
  const int64_t nv = nt->nvox[L.level];
  const dim3 g((unsigned)((nv + 15) / 16)), b(256);
  const int nt16 = Y.C / 16;
#define LAUNCH(NT_)                                                                                   \
  if (out == 0) hipLaunchKernelGGL((k_conv<NT_, 0>), g, b, 0, s, N, X1, X2, W, Y);                     \
  else if (out == 1) hipLaunchKernelGGL((k_conv<NT_, 1>), g, b, 0, s, N, X1, X2, W, Y);                \
  else hipLaunchKernelGGL((k_conv<NT_, 2>), g, b, 0, s, N, X1, X2, W, Y);
  switch (nt16) {
    case 1: LAUNCH(1) break;
    case 2: LAUNCH(2) break;
    case 3: LAUNCH(3) break;
    case 4: LAUNCH(4) break;
    case 6: LAUNCH(6) break;
    case 8: LAUNCH(8) break;
    default: return hipErrorInvalidValue;
  }
#undef LAUNCH
  return hipGetLastError();
}

// forward pass of one sample already packed into x0; output FP32 in nt->out
int forward(hfk_net* nt, hipStream_t s) {
  const unsigned short* W = nt->wbf.as<unsigned short>();
  auto Wl = [&](int li) { return W + nt->layers[li].woff; };
  hfk_ctx* ctx = nt->ctx;
  HIPCHK(ctx, conv(nt, 0, vb(nt->x0, 16, 16), kNone, Wl(0), vb(nt->a0, 16, 16), 0, s));
  HIPCHK(ctx, conv(nt, 1, vb(nt->a0, 16, 16), kNone, Wl(1), vb(nt->s0, 16, 16), 0, s));
  hipLaunchKernelGGL(k_restrict_bf, grid1(nt->nvox[1] * 16), dim3(256), 0, s, nt->N[0], vb(nt->s0, 16, 16), vb(nt->r1, 16, 16));
  HIPCHK(ctx, conv(nt, 2, vb(nt->r1, 16, 16), kNone, Wl(2), vb(nt->a2, 32, 32), 0, s));
  HIPCHK(ctx, conv(nt, 3, vb(nt->a2, 32, 32), kNone, Wl(3), vb(nt->s1, 32, 32), 0, s));
  hipLaunchKernelGGL(k_restrict_bf, grid1(nt->nvox[2] * 32), dim3(256), 0, s, nt->N[1], vb(nt->s1, 32, 32), vb(nt->r2, 32, 32));
  HIPCHK(ctx, conv(nt, 4, vb(nt->r2, 32, 32), kNone, Wl(4), vb(nt->s2, 64, 64), 0, s));
  hipLaunchKernelGGL(k_restrict_bf, grid1(nt->nvox[3] * 64), dim3(256), 0, s, nt->N[2], vb(nt->s2, 64, 64), vb(nt->r3, 64, 64));
  HIPCHK(ctx, conv(nt, 5, vb(nt->r3, 64, 64), kNone, Wl(5), vb(nt->a5, 64, 64), 0, s));
  HIPCHK(ctx, conv(nt, 6, vb(nt->a5, 64, 64), kNone, Wl(6), vb(nt->a6, 64, 64), 0, s));
  hipLaunchKernelGGL(k_prolong<unsigned short>, grid1(nt->nvox[2] * 64), dim3(256), 0, s, nt->N[2], vb(nt->a6, 64, 64), vb(nt->u2, 64, 64));
  HIPCHK(ctx, conv(nt, 7, vb(nt->s2, 64, 64), vb(nt->u2, 64, 64), Wl(7), vb(nt->a7, 64, 64), 0, s));
  hipLaunchKernelGGL(k_prolong<unsigned short>, grid1(nt->nvox[1] * 64), dim3(256), 0, s, nt->N[1], vb(nt->a7, 64, 64), vb(nt->u1, 64, 64));
  HIPCHK(ctx, conv(nt, 8, vb(nt->s1, 32, 32), vb(nt->u1, 64, 64), Wl(8), vb(nt->a8, 32, 32), 0, s));
  hipLaunchKernelGGL(k_prolong<unsigned short>, grid1(nt->nvox[0] * 32), dim3(256), 0, s, nt->N[0], vb(nt->a8, 32, 32), vb(nt->u0, 32, 32));
  HIPCHK(ctx, conv(nt, 9, vb(nt->s0, 16, 16), vb(nt->u0, 32, 32), Wl(9), vb(nt->a9, 16, 16), 0, s));
  HIPCHK(ctx, conv(nt, 10, vb(nt->a9, 16, 16), kNone, Wl(10), vb(nt->out, 16, 16), 1, s));
  HIPCHK(ctx, hipGetLastError());
  return HFK_OK;
}

// gradient of one conv layer: G = bf16(dY * relu'(Y)); dW += corr(X, G); dX (+)= convT(G)
int conv_backward(hfk_net* nt, int li, View dY, View Yact, View X1, View X2, View dX, int dx_out, hipStream_t s) {
  const Layer& L = nt->layers[li];
  hfk_ctx* ctx = nt->ctx;
  const int N = nt->N[L.level];
  const int64_t nv = nt->nvox[L.level];
  const int cin = L.cin1 + L.cin2;
  View G{nt->gbf.p, L.cout, L.cout};
  hipLaunchKernelGGL(k_relu_back, grid1(nv * L.cout), dim3(256), 0, s, nv, dY, Yact, G, L.relu ? 1 : 0);
  const int ntco = L.cout / 16, ntci = cin / 16;
  const int ntiles = 27 * ntco * ntci;
  const int nchunks = (int)((nv + kVoxWave - 1) / kVoxWave);
  const dim3 gw((unsigned)((nv + 4 * kVoxWave - 1) / (4 * kVoxWave)), (unsigned)ntiles);
  hipLaunchKernelGGL(k_conv_wgrad, gw, dim3(256), 0, s, N, G, X1, X2, nt->wpart.as<float>(), ntco, ntci);
  hipLaunchKernelGGL(k_wgrad_reduce, dim3((unsigned)ntiles), dim3(256), 0, s, nt->wpart.as<float>(), nchunks, ntiles,
                     ntco, ntci, cin, nt->grad.as<float>() + L.woff);
  HIPCHK(ctx, hipGetLastError());
  if (dX.p) {
    const size_t nw = (size_t)27 * L.cout * cin;
    hipLaunchKernelGGL(k_weight_adjoint, grid1((int64_t)nw), dim3(256), 0, s, nt->wbf.as<unsigned short>() + L.woff,
                       nt->wt.as<unsigned short>(), L.cout, cin);
    // the adjoint conv: input G (cout channels), weights Wt [tap][cin][cout], output dX (cin channels)
    const dim3 g((unsigned)((nv + 15) / 16)), b(256);
    const int nt16 = cin / 16;
#define LAUNCH(NT_)                                                                                                   \
  if (dx_out == 1) hipLaunchKernelGGL((k_conv<NT_, 1>), g, b, 0, s, N, G, kNone, nt->wt.as<unsigned short>(), dX);    \
  else hipLaunchKernelGGL((k_conv<NT_, 2>), g, b, 0, s, N, G, kNone, nt->wt.as<unsigned short>(), dX);
    switch (nt16) {
      case 1: LAUNCH(1) break;
      case 2: LAUNCH(2) break;
      case 3: LAUNCH(3) break;
      case 4: LAUNCH(4) break;
      case 6: LAUNCH(6) break;
      case 8: LAUNCH(8) break;
      default: return set_err(ctx, HFK_ESTATE, "HyperFEMKernels: unsupported channel count");
    }
#undef LAUNCH
    HIPCHK(ctx, hipGetLastError());
  }
  return HFK_OK;
}

// backward pass from nt->g_out (FP32 dLoss/dOut); accumulates into nt->grad
int backward(hfk_net* nt, hipStream_t s) {
  int rc;
  const View none = kNone;
  auto f32 = [](const DevBuf& b, int C, int stride, int off = 0) { return View{(void*)(b.as<float>() + off), C, stride}; };
  auto b16 = [](const DevBuf& b, int C, int stride, int off = 0) {
    return View{(void*)(b.as<unsigned short>() + off), C, stride};
  };
  // layer 10 (linear): dY = g_out; X = a9
  if ((rc = conv_backward(nt, 10, f32(nt->g_out, 16, 16), b16(nt->out, 16, 16), b16(nt->a9, 16, 16), none,
                          f32(nt->g_a9, 16, 16), 1, s)))
    return rc;
  // layer 9: X = [s0 | u0] -> d[s0 | u0] in g_cat0 (48 channels)
  if ((rc = conv_backward(nt, 9, f32(nt->g_a9, 16, 16), b16(nt->a9, 16, 16), b16(nt->s0, 16, 16), b16(nt->u0, 32, 32),
                          f32(nt->g_cat0, 48, 48), 1, s)))
    return rc;
  // u0 = P a8: d a8 = P' d u0
  hipLaunchKernelGGL(k_prolong_T, grid1(nt->nvox[1] * 32), dim3(256), 0, s, nt->N[0], f32(nt->g_cat0, 32, 48, 16),
                     f32(nt->g_a8tmp, 32, 32));
  // layer 8: X = [s1 | u1] -> g_cat1 (96)
  if ((rc = conv_backward(nt, 8, f32(nt->g_a8tmp, 32, 32), b16(nt->a8, 32, 32), b16(nt->s1, 32, 32), b16(nt->u1, 64, 64),
                          f32(nt->g_cat1, 96, 96), 1, s)))
    return rc;
  hipLaunchKernelGGL(k_prolong_T, grid1(nt->nvox[2] * 64), dim3(256), 0, s, nt->N[1], f32(nt->g_cat1, 64, 96, 32),
                     f32(nt->g_a7tmp, 64, 64));
  // layer 7: X = [s2 | u2] -> g_cat2 (128)
  if ((rc = conv_backward(nt, 7, f32(nt->g_a7tmp, 64, 64), b16(nt->a7, 64, 64), b16(nt->s2, 64, 64), b16(nt->u2, 64, 64),
                          f32(nt->g_cat2, 128, 128), 1, s)))
    return rc;
  hipLaunchKernelGGL(k_prolong_T, grid1(nt->nvox[3] * 64), dim3(256), 0, s, nt->N[2], f32(nt->g_cat2, 64, 128, 64),
                     f32(nt->g_a6tmp, 64, 64));
  // layers 6, 5 (level 3)
  if ((rc = conv_backward(nt, 6, f32(nt->g_a6tmp, 64, 64), b16(nt->a6, 64, 64), b16(nt->a5, 64, 64), none,
                          f32(nt->g_a5, 64, 64), 1, s)))
    return rc;
  if ((rc = conv_backward(nt, 5, f32(nt->g_a5, 64, 64), b16(nt->a5, 64, 64), b16(nt->r3, 64, 64), none,
                          f32(nt->g_r3, 64, 64), 1, s)))
    return rc;
  // r3 = R s2: d s2 (channels 0..63 of g_cat2) += R' d r3
  hipLaunchKernelGGL(k_restrict_T, grid1(nt->nvox[2] * 64), dim3(256), 0, s, nt->N[2], f32(nt->g_r3, 64, 64),
                     f32(nt->g_cat2, 64, 128, 0));
  // layer 4: Y = s2
  if ((rc = conv_backward(nt, 4, f32(nt->g_cat2, 64, 128, 0), b16(nt->s2, 64, 64), b16(nt->r2, 32, 32), none,
                          f32(nt->g_r2, 32, 32), 1, s)))
    return rc;
  hipLaunchKernelGGL(k_restrict_T, grid1(nt->nvox[1] * 32), dim3(256), 0, s, nt->N[1], f32(nt->g_r2, 32, 32),
                     f32(nt->g_cat1, 32, 96, 0));
  // layers 3, 2 (level 1): Y3 = s1, Y2 = a2
  if ((rc = conv_backward(nt, 3, f32(nt->g_cat1, 32, 96, 0), b16(nt->s1, 32, 32), b16(nt->a2, 32, 32), none,
                          f32(nt->g_a2, 32, 32), 1, s)))
    return rc;
  if ((rc = conv_backward(nt, 2, f32(nt->g_a2, 32, 32), b16(nt->a2, 32, 32), b16(nt->r1, 16, 16), none,
                          f32(nt->g_r1, 16, 16), 1, s)))
    return rc;
  hipLaunchKernelGGL(k_restrict_T, grid1(nt->nvox[0] * 16), dim3(256), 0, s, nt->N[0], f32(nt->g_r1, 16, 16),
                     f32(nt->g_cat0, 16, 48, 0));
  // layers 1, 0 (level 0): Y1 = s0, Y0 = a0
  if ((rc = conv_backward(nt, 1, f32(nt->g_cat0, 16, 48, 0), b16(nt->s0, 16, 16), b16(nt->a0, 16, 16), none,
                          f32(nt->g_a0, 16, 16), 1, s)))
    return rc;
  if ((rc = conv_backward(nt, 0, f32(nt->g_a0, 16, 16), b16(nt->a0, 16, 16), b16(nt->x0, 16, 16), none, none, 1, s)))
    return rc;
  HIPCHK(nt->ctx, hipGetLastError());
  return HFK_OK;
}

// pack r (coarse free vector, device) into x0 with its RMS as scale (scal[1]); scal[0] holds r.r
int pack_input(hfk_net* nt, const double* r, hipStream_t s) {
  double* sc = nt->scal.as<double>();
  HIPCHK(nt->ctx, vec_dots(nt->ncoarse, r, r, nullptr, nullptr, nullptr, nullptr, nt->dots.as<double>(), sc, s));
  hipLaunchKernelGGL(k_rms, dim3(1), dim3(1), 0, s, sc, nt->ncoarse, sc + 1);
  hipLaunchKernelGGL(k_pack, grid1(nt->nvin * 16), dim3(256), 0, s, nt->nvin, nt->node_dofs.as<int32_t>(), r, sc + 1,
                     nt->has_aux ? nt->aux.as<float>() : nullptr, (nt->coarsen ? nt->xin : nt->x0).as<unsigned short>());
  if (nt->coarsen)
    hipLaunchKernelGGL(k_restrict_bf, grid1(nt->nvox[0] * 16), dim3(256), 0, s, nt->Nin, vb(nt->xin, 16, 16),
                       vb(nt->x0, 16, 16));
  HIPCHK(nt->ctx, hipGetLastError());
  return HFK_OK;
}

// the network's output on the input lattice (FP32, 16-channel rows; channels 0..2 meaningful)
const float* output_on_input(hfk_net* nt, hipStream_t s) {
  if (!nt->coarsen) return nt->out.as<float>();
  hipLaunchKernelGGL(k_prolong<float>, grid1(nt->nvin * 3), dim3(256), 0, s, nt->Nin, View{nt->out.p, 3, 16},
                     View{nt->outf.p, 3, 16});
  return nt->outf.as<float>();
}

// dLoss/dOut on the input lattice (written by the caller into input_gradient(nt)) -> nt->g_out
float* input_gradient(hfk_net* nt) { return (nt->coarsen ? nt->g_outf : nt->g_out).as<float>(); }
void gradient_to_network(hfk_net* nt, hipStream_t s) {
  if (nt->coarsen)
    hipLaunchKernelGGL(k_prolong_T, grid1(nt->nvox[0] * 16), dim3(256), 0, s, nt->Nin, View{nt->g_outf.p, 16, 16},
                       View{nt->g_out.p, 16, 16});
}

int net_apply(hfk_net* nt, const double* r, double* z, hipStream_t s) {
  int rc;
  if ((rc = pack_input(nt, r, s))) return rc;
  if ((rc = forward(nt, s))) return rc;
  HIPCHK(nt->ctx, launch_fill(z, nt->ncoarse, 0.0, s));
  const float* o = output_on_input(nt, s);
  hipLaunchKernelGGL(k_unpack, grid1(nt->nvin * 3), dim3(256), 0, s, nt->nvin, nt->node_dofs.as<int32_t>(), o,
                     nt->scal.as<double>() + 1, 1.0 / nt->h, z);
  HIPCHK(nt->ctx, hipGetLastError());
  return HFK_OK;
}

// loss and gradient of one sample (r, e on the device); returns the loss through *loss (blocking read)
int sample_grad(hfk_net* nt, const double* r, const double* e, hipStream_t s, double* loss) {
  int rc;
  if ((rc = pack_input(nt, r, s))) return rc;
  if ((rc = forward(nt, s))) return rc;
  double* sc = nt->scal.as<double>();
  // tt = sum (e / (scale / h))^2 = (e.e) h^2 / scale^2 ; computed on the device
  HIPCHK(nt->ctx, vec_dots(nt->ncoarse, e, e, nullptr, nullptr, nullptr, nullptr, nt->dots.as<double>(), sc + 2, s));
  std::vector<double> hs(3);
  HIPCHK(nt->ctx, hipMemcpyAsync(hs.data(), sc, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(nt->ctx, hipStreamSynchronize(s));
  const double factor = 1.0 / nt->h;
  const double tt = hs[2] / (hs[1] * factor * hs[1] * factor);
  HIPCHK(nt->ctx, hipMemcpyAsync(sc + 3, &tt, sizeof(double), hipMemcpyHostToDevice, s));
  const int nb = 512;
  const float* o = output_on_input(nt, s);
  hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, s, nt->nvin, nt->node_dofs.as<int32_t>(), o, e, sc + 1, factor,
                     sc + 3, input_gradient(nt), nt->part.as<double>());
  gradient_to_network(nt, s);
  std::vector<double> part(nb);
  HIPCHK(nt->ctx, hipMemcpyAsync(part.data(), nt->part.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(nt->ctx, hipStreamSynchronize(s));
  double L = 0;
  for (double v : part) L += v;
  *loss = L;
  return backward(nt, s);
}

int net_alloc(hfk_net* nt) {
  hfk_ctx* ctx = nt->ctx;
  auto bf = [&](DevBuf& b, int lvl, int C) { return b.alloc((size_t)nt->nvox[lvl] * C * sizeof(unsigned short)); };
  auto fl = [&](DevBuf& b, int lvl, int C) { return b.alloc((size_t)nt->nvox[lvl] * C * sizeof(float)); };
  HIPCHK(ctx, bf(nt->x0, 0, 16));
  HIPCHK(ctx, bf(nt->a0, 0, 16));
  HIPCHK(ctx, bf(nt->s0, 0, 16));
  HIPCHK(ctx, bf(nt->r1, 1, 16));
  HIPCHK(ctx, bf(nt->a2, 1, 32));
  HIPCHK(ctx, bf(nt->s1, 1, 32));
  HIPCHK(ctx, bf(nt->r2, 2, 32));
  HIPCHK(ctx, bf(nt->s2, 2, 64));
  HIPCHK(ctx, bf(nt->r3, 3, 64));
  HIPCHK(ctx, bf(nt->a5, 3, 64));
  HIPCHK(ctx, bf(nt->a6, 3, 64));
  HIPCHK(ctx, bf(nt->u2, 2, 64));
  HIPCHK(ctx, bf(nt->a7, 2, 64));
  HIPCHK(ctx, bf(nt->u1, 1, 64));
  HIPCHK(ctx, bf(nt->a8, 1, 32));
  HIPCHK(ctx, bf(nt->u0, 0, 32));
  HIPCHK(ctx, bf(nt->a9, 0, 16));
  HIPCHK(ctx, fl(nt->out, 0, 16));
  HIPCHK(ctx, fl(nt->g_out, 0, 16));
  HIPCHK(ctx, fl(nt->g_a9, 0, 16));
  HIPCHK(ctx, fl(nt->g_cat0, 0, 48));
  HIPCHK(ctx, fl(nt->g_a8tmp, 1, 32));
  HIPCHK(ctx, fl(nt->g_cat1, 1, 96));
  HIPCHK(ctx, fl(nt->g_a7tmp, 2, 64));
  HIPCHK(ctx, fl(nt->g_cat2, 2, 128));
  HIPCHK(ctx, fl(nt->g_a6tmp, 3, 64));
  HIPCHK(ctx, fl(nt->g_a5, 3, 64));
  HIPCHK(ctx, fl(nt->g_r3, 3, 64));
  HIPCHK(ctx, fl(nt->g_r2, 2, 32));
  HIPCHK(ctx, fl(nt->g_a2, 1, 32));
  HIPCHK(ctx, fl(nt->g_r1, 1, 16));
  HIPCHK(ctx, fl(nt->g_a0, 0, 16));
  HIPCHK(ctx, nt->gbf.alloc((size_t)nt->nvox[0] * 64 * sizeof(unsigned short)));
  if (nt->coarsen) {
    HIPCHK(ctx, nt->xin.alloc((size_t)nt->nvin * 16 * sizeof(unsigned short)));
    HIPCHK(ctx, nt->outf.alloc((size_t)nt->nvin * 16 * sizeof(float)));
    HIPCHK(ctx, nt->g_outf.alloc((size_t)nt->nvin * 16 * sizeof(float)));
    HIPCHK(ctx, hipMemset(nt->outf.p, 0, (size_t)nt->nvin * 16 * sizeof(float)));
  }
  size_t maxw = 0, maxp = 0;
  for (const Layer& L : nt->layers) {
    maxw = std::max(maxw, (size_t)27 * L.cout * (L.cin1 + L.cin2));
    const size_t nchunks = (size_t)((nt->nvox[L.level] + kVoxWave - 1) / kVoxWave);
    maxp = std::max(maxp, (size_t)27 * (L.cout / 16) * ((L.cin1 + L.cin2) / 16) * nchunks * 256);
  }
  HIPCHK(ctx, nt->wt.alloc(maxw * sizeof(unsigned short)));
  HIPCHK(ctx, nt->wpart.alloc(maxp * sizeof(float)));
  HIPCHK(ctx, nt->scal.alloc(8 * sizeof(double)));
  HIPCHK(ctx, nt->part.alloc(512 * sizeof(double)));
  HIPCHK(ctx, nt->dots.alloc(dot_partial_doubles() * sizeof(double)));
  return HFK_OK;
}

}  // namespace

int net_apply_dev(hfk_net* nt, const double* r, double* z, hipStream_t s) { return net_apply(nt, r, z, s); }
int64_t net_ncoarse(const hfk_net* nt) { return nt->ncoarse; }
double net_h(const hfk_net* nt) { return nt->h; }

int net_forward_vec(hfk_net* nt, const double* r, double* w, hipStream_t s) {
  int rc;
  if ((rc = pack_input(nt, r, s))) return rc;
  if ((rc = forward(nt, s))) return rc;
  HIPCHK(nt->ctx, launch_fill(w, nt->ncoarse, 0.0, s));
  const float* o = output_on_input(nt, s);
  hipLaunchKernelGGL(k_out_to_vec, grid1(nt->nvin * 3), dim3(256), 0, s, nt->nvin, nt->node_dofs.as<int32_t>(), o, w);
  HIPCHK(nt->ctx, hipGetLastError());
  return HFK_OK;
}

int net_backward_vec(hfk_net* nt, const double* gw, double sc, hipStream_t s) {
  hipLaunchKernelGGL(k_vec_to_gout, grid1(nt->nvin * 16), dim3(256), 0, s, nt->nvin, nt->node_dofs.as<int32_t>(), gw, sc,
                     input_gradient(nt));
  gradient_to_network(nt, s);
  HIPCHK(nt->ctx, hipGetLastError());
  return backward(nt, s);
}

int net_zero_grad(hfk_net* nt, hipStream_t s) {
  HIPCHK(nt->ctx, hipMemsetAsync(nt->grad.p, 0, nt->nparam * sizeof(float), s));
  return HFK_OK;
}

int net_copy_grad(hfk_net* nt, float* out, hipStream_t s) {
  HIPCHK(nt->ctx, hipMemcpyAsync(out, nt->grad.p, nt->nparam * sizeof(float), hipMemcpyDeviceToHost, s));
  HIPCHK(nt->ctx, hipStreamSynchronize(s));
  return HFK_OK;
}

int net_adam(hfk_net* nt, double lr, int batch, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  nt->step++;
  const float c1 = 1.f - std::pow(b1, (float)nt->step), c2 = 1.f - std::pow(b2, (float)nt->step);
  hipLaunchKernelGGL(k_adam, grid1((int64_t)nt->nparam), dim3(256), 0, s, (int64_t)nt->nparam, nt->w32.as<float>(),
                     nt->grad.as<float>(), nt->m1.as<float>(), nt->m2.as<float>(), nt->wbf.as<unsigned short>(), (float)lr,
                     b1, b2, eps, c1, c2, 1.0f / (float)batch);
  HIPCHK(nt->ctx, hipGetLastError());
  return HFK_OK;
}

}  // namespace hfk

extern "C" {

int hfk_net_create(hfk_ctx* ctx, int n_lattice, int coarsen, const int32_t* node_dofs, int64_t ncoarse, double h,
                   uint64_t seed, hfk_net** out) {
  if (!ctx || !out || !node_dofs) return set_err(ctx, HFK_EARG, "HyperFEMKernels: null argument");
  if (coarsen != 0 && coarsen != 1) return set_err(ctx, HFK_EARG, "HyperFEMKernels: coarsen must be 0 or 1");
  if (coarsen && (n_lattice % 2) == 0)
    return set_err(ctx, HFK_EARG, "HyperFEMKernels: a coarsened network needs an odd lattice size");
  const int n0 = coarsen ? (n_lattice + 1) / 2 : n_lattice;
  if (n0 < 9 || (n0 % 2) == 0 || ((n0 + 1) / 2) % 2 == 0 || ((((n0 + 1) / 2) + 1) / 2) % 2 == 0)
    return set_err(ctx, HFK_EARG,
                   "HyperFEMKernels: the network lattice must be 2^k m + 1 for three halvings (e.g. 17, 33, 65) and "
                   "at least 9");
  if (!(h > 0) || ncoarse <= 0) return set_err(ctx, HFK_EARG, "HyperFEMKernels: invalid cell size or dof count");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<hfk_net> nt(new hfk_net());
  nt->ctx = ctx;
  nt->coarsen = coarsen;
  nt->Nin = n_lattice;
  nt->nvin = (int64_t)n_lattice * n_lattice * n_lattice;
  nt->N[0] = n0;
  for (int l = 1; l < 4; ++l) nt->N[l] = (nt->N[l - 1] + 1) / 2;
  for (int l = 0; l < 4; ++l) nt->nvox[l] = (int64_t)nt->N[l] * nt->N[l] * nt->N[l];
  nt->ncoarse = ncoarse;
  nt->h = h;
  nt->layers.assign(kLayers, kLayers + kNL);
  size_t off = 0;
  for (Layer& L : nt->layers) {
    L.woff = off;
    off += (size_t)27 * L.cout * (L.cin1 + L.cin2);
  }
  nt->nparam = off;
  // He initialisation (fan_in = 27 cin, gain 2 for ReLU; the last layer 1/fan_in), the output rows >= 3 of the
  // last layer and the input columns >= 3 of the first layer are zero (padding channels)
  std::vector<float> w(off, 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  for (int li = 0; li < kNL; ++li) {
    const Layer& L = nt->layers[li];
    const int cin = L.cin1 + L.cin2;
    const float sd = std::sqrt((L.relu ? 2.0f : 1.0f) / (27.0f * (li == 0 ? 4 : cin)));
    for (int tap = 0; tap < 27; ++tap)
      for (int co = 0; co < L.cout; ++co)
        for (int ci = 0; ci < cin; ++ci) {
          float v = sd * nd(rng);
          if (li == 0 && ci >= 4) v = 0.f;
          if (li == kNL - 1 && co >= 3) v = 0.f;
          w[L.woff + ((size_t)tap * L.cout + co) * cin + ci] = v;
        }
  }
  HIPCHK(ctx, nt->w32.alloc(off * sizeof(float)));
  HIPCHK(ctx, hipMemcpy(nt->w32.p, w.data(), off * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(ctx, nt->wbf.alloc(off * sizeof(unsigned short)));
  for (auto* b : {&nt->grad, &nt->m1, &nt->m2}) {
    HIPCHK(ctx, b->alloc(off * sizeof(float)));
    HIPCHK(ctx, hipMemset(b->p, 0, off * sizeof(float)));
  }
  hipLaunchKernelGGL(k_to_bf, grid1((int64_t)off), dim3(256), 0, ctx->stream, (int64_t)off, nt->w32.as<float>(),
                     nt->wbf.as<unsigned short>());
  HIPCHK(ctx, nt->node_dofs.alloc((size_t)nt->nvin * 3 * sizeof(int32_t)));
  HIPCHK(ctx, hipMemcpy(nt->node_dofs.p, node_dofs, (size_t)nt->nvin * 3 * sizeof(int32_t), hipMemcpyHostToDevice));
  int rc = net_alloc(nt.get());
  if (rc) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  *out = nt.release();
  return HFK_OK;
}

int hfk_net_destroy(hfk_net* nt) {
  delete nt;
  return HFK_OK;
}

int hfk_net_set_aux(hfk_net* nt, const float* aux) {
  if (!nt) return set_err(nullptr, HFK_EARG, "HyperFEMKernels: null network");
  if (!aux) {
    nt->has_aux = false;
    return HFK_OK;
  }
  HIPCHK(nt->ctx, nt->aux.alloc((size_t)nt->nvin * sizeof(float)));
  HIPCHK(nt->ctx, hipMemcpy(nt->aux.p, aux, (size_t)nt->nvin * sizeof(float), hipMemcpyHostToDevice));
  nt->has_aux = true;
  return HFK_OK;
}

int64_t hfk_net_num_params(const hfk_net* nt) { return nt ? (int64_t)nt->nparam : -1; }

int hfk_net_get_weights(hfk_net* nt, float* w) {
  if (!nt || !w) return set_err(nt ? nt->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  HIPCHK(nt->ctx, hipMemcpy(w, nt->w32.p, nt->nparam * sizeof(float), hipMemcpyDeviceToHost));
  return HFK_OK;
}

int hfk_net_set_weights(hfk_net* nt, const float* w) {
  if (!nt || !w) return set_err(nt ? nt->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  HIPCHK(nt->ctx, hipMemcpy(nt->w32.p, w, nt->nparam * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_to_bf, grid1((int64_t)nt->nparam), dim3(256), 0, nt->ctx->stream, (int64_t)nt->nparam,
                     nt->w32.as<float>(), nt->wbf.as<unsigned short>());
  HIPCHK(nt->ctx, hipStreamSynchronize(nt->ctx->stream));
  return HFK_OK;
}

int hfk_net_apply(hfk_net* nt, const double* r, double* z) {
  if (!nt || !r || !z) return set_err(nt ? nt->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  hfk_ctx* ctx = nt->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  DevBuf rd, zd;
  HIPCHK(ctx, rd.alloc(nt->ncoarse * sizeof(double)));
  HIPCHK(ctx, zd.alloc(nt->ncoarse * sizeof(double)));
  hipStream_t s = ctx->stream;
  HIPCHK(ctx, hipMemcpyAsync(rd.p, r, nt->ncoarse * sizeof(double), hipMemcpyHostToDevice, s));
  int rc = net_apply(nt, rd.as<double>(), zd.as<double>(), s);
  if (rc) return rc;
  HIPCHK(ctx, hipMemcpyAsync(z, zd.p, nt->ncoarse * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(ctx, hipStreamSynchronize(s));
  return HFK_OK;
}

int hfk_net_apply_dev(hfk_net* nt, const double* r_dev, double* z_dev, void* stream) {
  if (!nt || !r_dev || !z_dev) return set_err(nt ? nt->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  HIPCHK(nt->ctx, hipSetDevice(nt->ctx->device));
  return net_apply(nt, r_dev, z_dev, stream ? (hipStream_t)stream : nt->ctx->stream);
}

int hfk_net_gradient(hfk_net* nt, const double* r, const double* e, int nsamples, float* grad, double* loss) {
  if (!nt || !r || !e || nsamples < 1) return set_err(nt ? nt->ctx : nullptr, HFK_EARG, "HyperFEMKernels: null argument");
  hfk_ctx* ctx = nt->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int64_t n = nt->ncoarse;
  DevBuf rd, ed;
  HIPCHK(ctx, rd.alloc(n * sizeof(double)));
  HIPCHK(ctx, ed.alloc(n * sizeof(double)));
  HIPCHK(ctx, hipMemsetAsync(nt->grad.p, 0, nt->nparam * sizeof(float), s));
  double total = 0;
  for (int k = 0; k < nsamples; ++k) {
    HIPCHK(ctx, hipMemcpyAsync(rd.p, r + (size_t)k * n, n * sizeof(double), hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ed.p, e + (size_t)k * n, n * sizeof(double), hipMemcpyHostToDevice, s));
    double L;
    int rc = sample_grad(nt, rd.as<double>(), ed.as<double>(), s, &L);
    if (rc) return rc;
    total += L;
  }
  if (grad) HIPCHK(ctx, hipMemcpyAsync(grad, nt->grad.p, nt->nparam * sizeof(float), hipMemcpyDeviceToHost, s));
  HIPCHK(ctx, hipStreamSynchronize(s));
  if (loss) *loss = total;
  return HFK_OK;
}

// Adam training on (r_k, e_k) pairs resident on the device ([k][ncoarse] each): `steps` steps of `batch` samples
// drawn in a fixed pseudo-random order; the mean loss of every step is written to history (length steps)
int hfk_net_train_dev(hfk_net* nt, const double* r_dev, const double* e_dev, int nsamples, int steps, int batch,
                      double lr, uint64_t seed, double* history, void* stream) {
  if (!nt || !r_dev || !e_dev || nsamples < 1 || steps < 0 || batch < 1 || !(lr > 0))
    return set_err(nt ? nt->ctx : nullptr, HFK_EARG, "HyperFEMKernels: invalid training arguments");
  hfk_ctx* ctx = nt->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
  std::mt19937_64 rng(seed);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int st = 0; st < steps; ++st) {
    double L = 0;
    for (int b = 0; b < batch; ++b) {
      const int k = (int)(rng() % (uint64_t)nsamples);
      double l;
      int rc = sample_grad(nt, r_dev + (size_t)k * nt->ncoarse, e_dev + (size_t)k * nt->ncoarse, s, &l);
      if (rc) return rc;
      L += l;
    }
    nt->step++;
    const float c1 = 1.f - std::pow(b1, (float)nt->step), c2 = 1.f - std::pow(b2, (float)nt->step);
    hipLaunchKernelGGL(k_adam, grid1((int64_t)nt->nparam), dim3(256), 0, s, (int64_t)nt->nparam, nt->w32.as<float>(),
                       nt->grad.as<float>(), nt->m1.as<float>(), nt->m2.as<float>(), nt->wbf.as<unsigned short>(),
                       (float)lr, b1, b2, eps, c1, c2, 1.0f / (float)batch);
    HIPCHK(ctx, hipGetLastError());
    if (history) history[st] = L / batch;
  }
  HIPCHK(ctx, hipStreamSynchronize(s));
  return HFK_OK;
}

}  // extern "C"
