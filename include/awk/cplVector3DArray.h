// cplVector3DArray.h -- device vector field as three SoA planes, the reference's
// cplVector3DArray (CUDACommon/inc/cudaVector3DArray.h:32-70): x, y, z plane
// pointers, n elements, nAlign the plane stride of one allocation.  Planes may
// also be any three separate device buffers (nAlign = n then).
#ifndef AWK_CPLVECTOR3DARRAY_H
#define AWK_CPLVECTOR3DARRAY_H

#include <hip/hip_runtime.h>

#include "awk_types.h"

struct cplVector3DArray {
  float *x = nullptr, *y = nullptr, *z = nullptr;
  unsigned int n = 0;       // number of elements
  unsigned int nAlign = 0;  // plane stride when the planes share one allocation
  cplVector3DArray() {}
  cplVector3DArray(float *px, float *py, float *pz, unsigned int count)
      : x(px), y(py), z(pz), n(count), nAlign(count) {}
  unsigned int size() const { return n; }
  unsigned int capacity() const { return nAlign; }
  bool isContinous() const { return n == nAlign; }
  float *elementArray(int i) { return i == 0 ? x : i == 1 ? y : i == 2 ? z : nullptr; }
  const float *elementArray(int i) const { return i == 0 ? x : i == 1 ? y : i == 2 ? z : nullptr; }
};

// allocateDeviceVector3DArray / freeDeviceVector3DArray (cudaVector3DArray.cu):
// one allocation, planes at x, x + n, x + 2n.
inline void allocateDeviceVector3DArray(cplVector3DArray &a, unsigned int n) {
  float *p = nullptr;
  if (hipMalloc(&p, 3 * (size_t)n * sizeof(float)) != hipSuccess)
    throw AtlasWerksException(__FILE__, __LINE__, "device allocation failed");
  a = cplVector3DArray(p, p + n, p + 2 * (size_t)n, n);
}
inline void freeDeviceVector3DArray(cplVector3DArray &a) {
  if (a.x) (void)hipFree(a.x);
  a = cplVector3DArray();
}

#endif
