// Internal declarations shared by the host side of libhfk (hfk_capi.hip) and the kernel launchers
// (hfk_kernels.hip).  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "hfk_types.hpp"

namespace hfk {

// Geometry of the launch of one colour (or of all batches when the kernel has no scatter).
struct BatchRange {
  int first;  // first batch
  int count;  // number of batches
};



// Launchers (hfk_kernels.hip).  All are asynchronous on `s`.
hipError_t launch_prepare(const OpDev& op, int storage, void* tangents, const double* x, BatchRange r, hipStream_t s);
// Single-launch mode of the product: colfc[2c] = first batch and colfc[2c+1] = batch count of colour c (device),
// gx = neighbourhoods per XCD, grid = 8 * gx * ncol work-groups.
struct ApplyMap {
  const int* colfc;
  int ncol;
  int gx;
  int grid;
};

// y[id] += alpha * (J dx)[id] over the batches r of one colour (m == null), or over every batch in one launch
// with atomic adds (m != null); the caller zeroes y first for a plain product
hipError_t launch_apply(const OpDev& op, int storage, const void* tangents, const double* dx, double* y, BatchRange r,
                        hipStream_t s, double alpha = 1.0, const ApplyMap* m = nullptr);
// nodal 3x3 blocks of the tangent, accumulated into B[blk*9 + ci + 3cj] over the batches r of one colour
hipError_t launch_blocks(const OpDev& op, int storage, const void* tangents, const int32_t* block_of, double* B,
                         BatchRange r, hipStream_t s);
hipError_t launch_residual(const OpDev& op, const double* x, double* b, BatchRange r, hipStream_t s);
// strain energy of the built-in law per batch into part[nbatches] (k_energy)
hipError_t launch_energy(const OpDev& op, const double* x, double* part, int64_t nbatches, hipStream_t s);
// boundary traction load of the faces `faces[0..nfaces)` (indices into face_slot/face_local/traction) added into f
hipError_t launch_face_load(const OpDev& op, const int64_t* face_slot, const int32_t* face_local, const double* traction,
                            const int32_t* faces, int64_t nfaces, int64_t nfree, double* f, hipStream_t s);
// Robin faces: face mass matrices M (nfaces x Q^4, k included) and the terms built from them, per launch group
hipError_t launch_face_mass(const OpDev& op, const int64_t* face_slot, const int32_t* face_local, const double* kf,
                            int64_t nfaces, double* M, hipStream_t s);
hipError_t launch_face_apply(const OpDev& op, const int64_t* face_slot, const int32_t* face_local, const double* M,
                             const int32_t* faces, int64_t n, const double* uf, const double* ud, double* y, double alpha,
                             hipStream_t s);
hipError_t launch_face_blocks(const OpDev& op, const int64_t* face_slot, const int32_t* face_local, const double* M,
                              const int32_t* faces, int64_t n, const int32_t* block_of, double* B, hipStream_t s);
hipError_t launch_face_assemble(const OpDev& op, const int64_t* face_slot, const int32_t* face_local, const double* M,
                                const int32_t* faces, int64_t n, double* nzval, hipStream_t s);
hipError_t launch_face_energy(const OpDev& op, const int64_t* face_slot, const int32_t* face_local, const double* M,
                              int64_t nfaces, const double* x, double* part, hipStream_t s);
// y -= a x
hipError_t launch_sub_scaled(double* y, const double* x, double a, int64_t n, hipStream_t s);
// part[0..nblocks) = scale * (fixed-order partial sums of a . b)
hipError_t launch_dot_partials(const double* a, const double* b, int64_t n, double scale, double* part, int nblocks,
                               hipStream_t s);
// out[0] = sum (mode 0), min (1) or max (2) of part[0..n) in a fixed order
hipError_t launch_reduce_partials(const double* part, int64_t n, int mode, double* out, hipStream_t s);
hipError_t launch_assemble(const OpDev& op, const double* tangents64, double* nzval, BatchRange r, hipStream_t s);
hipError_t launch_fill(double* y, int64_t n, double v, hipStream_t s);
hipError_t launch_expand_tangents(const OpDev& op, int storage, const void* tangents, double* D, int64_t nbatches,
                                  hipStream_t s);

// number of stored entries and bytes per entry of a storage mode (0 full64, 1 sym64, 2 full32, 3 sym32, 4 fac64,
// 5 fac32)
inline int store_entries(int storage) {
  return storage >= 6 ? 11 : (storage >= 4 ? 17 : ((storage == 1 || storage == 3) ? 45 : 81));
}
inline int store_bytes(int storage) { return (storage == 2 || storage == 3 || storage == 5 || storage == 7) ? 4 : 8; }
inline bool store_fp32(int storage) { return storage == 2 || storage == 3 || storage == 5 || storage == 7; }
// the factorised storages (an isotropic law with a factorised tangent); the affine ones need parallelepiped cells
inline bool store_factorised(int storage) { return storage >= 4 && storage <= 7; }
inline bool store_affine(int storage) { return storage == 6 || storage == 7; }

}  // namespace hfk
