// CUDAUtilities.h -- host AoS VectorField <-> device SoA cplVector3DArray copies,
// CUDAUtilities::CopyVectorFieldFromDevice / CopyVectorFieldToDevice
// (Base/CUDAUtilities.h:19-30, CUDAUtilities.cxx:22-72).  The host field moves
// as one block and is transposed on the device (awk_vf_aos_to_soa /
// awk_vf_soa_to_aos) rather than one component at a time on the host.  Both calls
// are synchronous, as the reference's are.  useHostArraySize: the element count
// comes from the host field (else from the device field); tmp is accepted for
// signature compatibility and not needed.
#ifndef AWK_CUDAUTILITIES_H
#define AWK_CUDAUTILITIES_H

#include "cplVector3DArray.h"

class CUDAUtilities {
 public:
  static void CopyVectorFieldFromDevice(const cplVector3DArray &deviceVF, VectorField &hostVF,
                                        bool useHostArraySize = false, float *tmp = nullptr) {
    (void)tmp;
    const size_t n = useHostArraySize ? hostVF.getNumElements() : deviceVF.size();
    if (hostVF.getNumElements() < n)
      throw AtlasWerksException(__FILE__, __LINE__, "CopyVectorFieldFromDevice: host field too small");
    if (!n) return;
    float *stage = nullptr;
    if (hipMalloc(&stage, 3 * n * sizeof(float)) != hipSuccess)
      throw AtlasWerksException(__FILE__, __LINE__, "CopyVectorFieldFromDevice: staging allocation failed");
    awk_status st = awk_vf_soa_to_aos(deviceVF.x, deviceVF.y, deviceVF.z, stage, (long long)n, nullptr);
    hipError_t e = st == AWK_OK ? hipMemcpy(reinterpret_cast<float *>(hostVF.getDataPointer(0)), stage,
                                            3 * n * sizeof(float), hipMemcpyDeviceToHost)
                                : hipSuccess;
    (void)hipFree(stage);
    AWK_CHECK(st);
    if (e != hipSuccess) throw AtlasWerksException(__FILE__, __LINE__, hipGetErrorString(e));
  }
  static void CopyVectorFieldToDevice(const VectorField &hostVF, cplVector3DArray &deviceVF,
                                      bool useHostArraySize = false, float *tmp = nullptr) {
    (void)tmp;
    const size_t n = useHostArraySize ? hostVF.getNumElements() : deviceVF.size();
    if (hostVF.getNumElements() < n)
      throw AtlasWerksException(__FILE__, __LINE__, "CopyVectorFieldToDevice: host field too small");
    if (!n) return;
    float *stage = nullptr;
    if (hipMalloc(&stage, 3 * n * sizeof(float)) != hipSuccess)
      throw AtlasWerksException(__FILE__, __LINE__, "CopyVectorFieldToDevice: staging allocation failed");
    hipError_t e = hipMemcpy(stage, reinterpret_cast<const float *>(hostVF.getDataPointer(0)),
                             3 * n * sizeof(float), hipMemcpyHostToDevice);
    awk_status st = e == hipSuccess ? awk_vf_aos_to_soa(stage, deviceVF.x, deviceVF.y, deviceVF.z,
                                                        (long long)n, nullptr)
                                    : AWK_OK;
    if (st == AWK_OK && e == hipSuccess) e = hipStreamSynchronize(nullptr);
    (void)hipFree(stage);
    AWK_CHECK(st);
    if (e != hipSuccess) throw AtlasWerksException(__FILE__, __LINE__, hipGetErrorString(e));
  }
};

#endif
