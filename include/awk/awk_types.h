// awk_types.h -- the small vocabulary types the C++ shims share with the
// reference API (AtlasWerks Base/AtlasWerksTypes.h, DataTypes/Vector3D.h,
// Base/AtlasWerksException.h).  Define AWK_USE_REFERENCE_TYPES to take the
// reference's own Vector3D / AtlasWerksException instead (drop-in builds).
#ifndef AWK_TYPES_H
#define AWK_TYPES_H

#include <stdexcept>
#include <string>

#include "../awk_hip.h"

#ifdef AWK_USE_REFERENCE_TYPES
#include "AtlasWerksException.h"
#include "Vector3D.h"
typedef Vector3D<int> Vector3Di;
typedef Vector3D<float> Vector3Df;
#else
template <class T>
struct Vector3D {
  T x, y, z;
  Vector3D() : x(0), y(0), z(0) {}
  Vector3D(T a, T b, T c) : x(a), y(b), z(c) {}
  T productOfElements() const { return x * y * z; }
};
typedef Vector3D<int> Vector3Di;
typedef Vector3D<float> Vector3Df;
typedef Vector3D<unsigned int> SizeType;
typedef Vector3D<float> SpacingType;

// AtlasWerksException(file, line, msg) (Base/AtlasWerksException.h:30-58)
class AtlasWerksException : public std::runtime_error {
 public:
  AtlasWerksException(const char *file, int line, const std::string &msg)
      : std::runtime_error(std::string(file) + ":" + std::to_string(line) + ": " + msg) {}
};
#endif

// BackgroundStrategy (CUDACommon/inc/cudaHField3DUtils.h; HField3DUtils.h)
enum BackgroundStrategy {
  BACKGROUND_STRATEGY_ID = AWK_BG_ID,
  BACKGROUND_STRATEGY_ZERO = AWK_BG_ZERO,
  BACKGROUND_STRATEGY_PARTIAL_ID = AWK_BG_PARTIAL_ID,
  BACKGROUND_STRATEGY_PARTIAL_ZERO = AWK_BG_PARTIAL_ZERO,
  BACKGROUND_STRATEGY_CLAMP = AWK_BG_CLAMP,
  BACKGROUND_STRATEGY_WRAP = AWK_BG_WRAP,
  BACKGROUND_STRATEGY_VAL = AWK_BG_VAL
};

// C-ABI status -> exception, the reference's error convention for
// configuration errors (DiffOperGPU.cxx:127-161, cudaHField3DUtils.cu:306-307).
inline void awkThrowOnError(awk_status st, const char *file, int line) {
  if (st != AWK_OK) throw AtlasWerksException(file, line, awk_last_error());
}
#define AWK_CHECK(call) awkThrowOnError((call), __FILE__, __LINE__)

typedef awk_stream_t awkStream_t;

#endif
