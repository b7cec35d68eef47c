// KernelFactory.h -- KernelFactory::NewGPUKernel (Algorithms/Common/KernelFactory.cxx:
// 45-60): the GPU operator a KernelParam selects.  This build provides the
// DiffOper (FluidKernelFFT) operator; a MultiGaussian parameter is rejected with
// the reference's exception type, as it rejects unknown kernel types.
#ifndef AWK_KERNELFACTORY_H
#define AWK_KERNELFACTORY_H

#include "DiffOperGPU.h"

class KernelFactory {
 public:
  static KernelInterfaceGPU *NewGPUKernel(const KernelParam &param) {
    if (param.IsDiffOperParam()) return new DiffOperGPU();
    if (param.IsMultiGaussianKernelParam())
      throw AtlasWerksException(__FILE__, __LINE__,
                                "Error, the MultiGaussian kernel is outside this GPU build");
    throw AtlasWerksException(__FILE__, __LINE__, "Error, Unknown kernel type");
  }
};

#endif
