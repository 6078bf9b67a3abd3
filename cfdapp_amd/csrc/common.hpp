// Internal definitions shared by the HIP translation units of libcfdhip.
//
// One context = one device + one HIP stream + the resident state of one
// structured Cartesian case: the mesh plan, the fields, the assembled systems
// and the Krylov work vectors. Nothing here crosses the C ABI
// (include/cfdhip.h is the boundary).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cfdhip.h"

namespace cfdhip {

// Library errors are C++ exceptions internally; every C entry point catches
// them, records the message on the context and returns the code.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define CFDHIP_CHECK(expr)                                                                  \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess)                                                                   \
      throw ::cfdhip::Error(CFDHIP_ERR_RUNTIME,                                             \
                            std::string(#expr) + ": " + hipGetErrorString(_e));             \
  } while (0)

inline void require(bool ok, const std::string& msg, int code = CFDHIP_ERR_INVALID) {
  if (!ok) throw Error(code, msg);
}

// Move-only device allocation.
template <class T>
struct DeviceArray {
  T* ptr = nullptr;
  int64_t count = 0;
  DeviceArray() = default;
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : ptr(o.ptr), count(o.count) {
    o.ptr = nullptr;
    o.count = 0;
  }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr;
      count = o.count;
      o.ptr = nullptr;
      o.count = 0;
    }
    return *this;
  }
  ~DeviceArray() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
  }
  // Keeps the allocation when it is already large enough (no reallocation
  // across outer iterations, as the reference asserts for its buffers).
  void ensure(int64_t n) {
    if (n <= count && ptr) return;
    release();
    if (n > 0) CFDHIP_CHECK(hipMalloc(&ptr, sizeof(T) * static_cast<size_t>(n)));
    count = n;
  }
  T* get() const { return ptr; }
};

// Boundary patch constants. On a uniform Cartesian mesh every face of a patch
// has the same geometry; the host plan builder evaluates the reference
// formulas for every boundary face and refuses the mesh if any patch is not
// uniform, so kernels can take these per patch.
struct PatchGeom {
  int vkind;          // CFDHIP_WALL ...
  double vval[3];     // prescribed velocity (MovingWall / Inlet)
  int pkind;          // CFDHIP_FIXED_GRADIENT / CFDHIP_FIXED_VALUE
  double pval;        // gradient or value
  double S[3];        // stored (outward) area vector, signed zeros included
  double area;        // |S|
  double bdist;       // owner centroid -> face centroid distance
  // MeshGeometry::boundaryInwardStencil + boundaryFaceDiffusionTerms (wall stencil)
  int stencil;
  double cP, cF, cB;
  double deltaP[3], deltaF[3];
  int deltaZero;      // both tangential transfers exactly zero
  // boundaryConsistentFaceValue (Green-Gauss opposite-face correction)
  int consistent;
  double back, farDist, w, offset[3];
};

// Device view of the plan, passed by value to every kernel.
struct View {
  int dim;
  int64_t nx, ny, nz, n, nf, nfx, offY, offZ;
  int64_t ncell[3];
  int64_t stride[3];
  // Per-axis face tables, 5 blocks of (ncell[a] + 1): dPf | dNf | dSum | dPN | dC
  // where for face index f between cells f-1 and f: dPf/dNf = owner/neighbour
  // centroid -> face distance, dSum = dPf + dNf, dPN = |dC|, dC = center(f) -
  // center(f-1). Entries 0 and ncell hold the boundary-face distances in dPf.
  const double* tab[3];
  double area[3];     // |S| of an internal face per axis
  double S[3][3];     // internal face area vector per axis (owner -> neighbour)
  double vol;
  double rho;
  int transfer;       // boundaryTransferNeeded: run the 4 extra Green-Gauss sweeps
  int npatch;
  PatchGeom patch[6];
};

enum TableBlock { kDPf = 0, kDNf = 1, kDSum = 2, kDPN = 3, kDC = 4 };

__host__ __device__ inline double tabv(const View& v, int a, int block, int64_t f) {
  return v.tab[a][block * (v.ncell[a] + 1) + f];
}

__host__ __device__ inline void cellIndex(const View& v, int64_t c, int64_t idx[3]) {
  idx[0] = c % v.nx;
  const int64_t r = c / v.nx;
  idx[1] = r % v.ny;
  idx[2] = r / v.ny;
}

// Face id of the face on axis a whose along-axis face index is idx[a].
__host__ __device__ inline int64_t faceId(const View& v, int a, const int64_t idx[3]) {
  if (a == 0) return (idx[2] * v.ny + idx[1]) * (v.nx + 1) + idx[0];
  if (a == 1) return v.offY + (idx[2] * (v.ny + 1) + idx[1]) * v.nx + idx[0];
  return v.offZ + (idx[2] * v.ny + idx[1]) * v.nx + idx[0];
}

// Inverse of faceId: axis and (i, j, k) with the face index in idx[axis].
__host__ __device__ inline int decodeFace(const View& v, int64_t f, int64_t idx[3]) {
  if (f < v.offY) {
    idx[0] = f % (v.nx + 1);
    const int64_t r = f / (v.nx + 1);
    idx[1] = r % v.ny;
    idx[2] = r / v.ny;
    return 0;
  }
  if (v.dim == 2 || f < v.offZ) {
    const int64_t g = f - v.offY;
    idx[0] = g % v.nx;
    const int64_t r = g / v.nx;
    idx[1] = r % (v.ny + 1);
    idx[2] = r / (v.ny + 1);
    return 1;
  }
  const int64_t g = f - v.offZ;
  idx[0] = g % v.nx;
  const int64_t r = g / v.nx;
  idx[1] = r % v.ny;
  idx[2] = r / v.ny;
  return 2;
}

__host__ __device__ inline int64_t cellOf(const View& v, const int64_t idx[3]) {
  return idx[0] + v.nx * (idx[1] + v.ny * idx[2]);
}

// z-guarded dot product of include/cfd/core/Vector3.hpp:78-82.
__host__ __device__ inline double dot3(const double a[3], const double b[3]) {
  const double inPlane = (a[0] * b[0]) + (a[1] * b[1]);
  const double normal = a[2] * b[2];
  return normal == 0.0 ? inPlane : inPlane + normal;
}

constexpr int kBlock = 256;
inline dim3 blocksFor(int64_t count, int block = kBlock) {
  int64_t b = (count + block - 1) / block;
  return dim3(static_cast<unsigned>(b < 1 ? 1 : b));
}

}  // namespace cfdhip
