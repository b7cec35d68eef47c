// awk_types.h -- the vocabulary types the C++ shims share with the reference API
// (Base/AtlasWerksTypes.h, DataTypes/Vector3D.h, DataTypes/Array3D.h,
// DataTypes/Image.h, Base/AtlasWerksException.h).
//
// Two builds:
//  * standalone (default): minimal stand-ins with the reference's names and the
//    members the drivers use (host images and AoS vector fields, x fastest);
//  * drop-in (-DAWK_USE_REFERENCE_TYPES): the reference's own headers supply
//    these types, and the shims' GPU classes compile against them.
#ifndef AWK_TYPES_H
#define AWK_TYPES_H

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "../awk_hip.h"

#ifdef AWK_USE_REFERENCE_TYPES
#include "AtlasWerksException.h"
#include "AtlasWerksTypes.h"
typedef Vector3D<int> Vector3Di;
typedef Vector3D<float> Vector3Df;
#else
typedef float Real;

template <class T>
struct Vector3D {
  T x, y, z;
  Vector3D() : x(0), y(0), z(0) {}
  Vector3D(T a, T b, T c) : x(a), y(b), z(c) {}
  T productOfElements() const { return x * y * z; }
  T &operator[](int i) { return i == 0 ? x : i == 1 ? y : z; }
  const T &operator[](int i) const { return i == 0 ? x : i == 1 ? y : z; }
  bool operator==(const Vector3D &o) const { return x == o.x && y == o.y && z == o.z; }
  bool operator!=(const Vector3D &o) const { return !(*this == o); }
};
typedef Vector3D<int> Vector3Di;
typedef Vector3D<float> Vector3Df;
typedef Vector3D<unsigned int> SizeType;
typedef Vector3D<Real> SpacingType;
typedef Vector3D<Real> OriginType;

// Array3D<T> (DataTypes/Array3D.h): a host volume, x fastest.
template <class T>
class Array3D {
 public:
  Array3D() {}
  explicit Array3D(const SizeType &size) { resize(size); }
  void resize(const SizeType &size) {
    mSize = size;
    mData.assign((size_t)size.x * size.y * size.z, T());
  }
  const SizeType &getSize() const { return mSize; }
  unsigned int getNumElements() const { return (unsigned int)mData.size(); }
  T *getDataPointer(unsigned int i = 0) { return mData.data() + i; }
  const T *getDataPointer(unsigned int i = 0) const { return mData.data() + i; }
  T &operator()(unsigned x, unsigned y, unsigned z) {
    return mData[((size_t)z * mSize.y + y) * mSize.x + x];
  }
  const T &operator()(unsigned x, unsigned y, unsigned z) const {
    return mData[((size_t)z * mSize.y + y) * mSize.x + x];
  }
  T &operator()(unsigned i) { return mData[i]; }
  const T &operator()(unsigned i) const { return mData[i]; }
  void fill(const T &v) { mData.assign(mData.size(), v); }

 protected:
  SizeType mSize;
  std::vector<T> mData;
};

// Image<T> (DataTypes/Image.h): an Array3D with origin and spacing.
template <class T>
class Image : public Array3D<T> {
 public:
  Image() : mSpacing(1, 1, 1) {}
  explicit Image(const SizeType &size, const OriginType &origin = OriginType(),
                 const SpacingType &spacing = SpacingType(1, 1, 1))
      : Array3D<T>(size), mOrigin(origin), mSpacing(spacing) {}
  const OriginType &getOrigin() const { return mOrigin; }
  const SpacingType &getSpacing() const { return mSpacing; }
  void setOrigin(const OriginType &o) { mOrigin = o; }
  void setSpacing(const SpacingType &s) { mSpacing = s; }

 private:
  OriginType mOrigin;
  SpacingType mSpacing;
};
typedef Image<Real> RealImage;
// VectorField: AoS Vector3D<float> per voxel (AtlasWerksTypes.h)
typedef Array3D<Vector3D<float> > VectorField;

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
