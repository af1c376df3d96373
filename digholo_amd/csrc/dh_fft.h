// Complex helpers and small DFT butterflies shared by the FFT kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace dh {

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
template <bool INV>
__device__ __forceinline__ float2 twiddle(const float2* tw, int k) {
  float2 w = tw[k];
  if (INV) w.y = -w.y;
  return w;
}

// ---------------------------------------------------------------------------
// small DFT butterflies, X[k] = sum_r v[r] exp(-/+ 2 pi i r k / R)
// ---------------------------------------------------------------------------
template <bool INV>
__device__ __forceinline__ void dft2(float2* v) {
  float2 a = v[0];
  v[0] = cadd(a, v[1]);
  v[1] = csub(a, v[1]);
}
template <bool INV>
__device__ __forceinline__ void dft4(float2* v) {
  float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
  float2 t2 = cadd(v[1], v[3]), t3 = mul_mi<INV>(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[2] = csub(t0, t2);
  v[1] = cadd(t1, t3);
  v[3] = csub(t1, t3);
}
template <bool INV>
__device__ __forceinline__ void dft8(float2* v) {
  const float s = 0.70710678118654752440f;
  float2 a[4] = {v[0], v[2], v[4], v[6]};
  float2 b[4] = {v[1], v[3], v[5], v[7]};
  dft4<INV>(a);
  dft4<INV>(b);
  // b[k] *= exp(-/+ i pi k / 4)
  float2 b1 = INV ? make_float2(s * (b[1].x - b[1].y), s * (b[1].x + b[1].y))
                  : make_float2(s * (b[1].x + b[1].y), s * (b[1].y - b[1].x));
  float2 b2 = mul_mi<INV>(b[2]);
  float2 b3 = INV ? make_float2(-s * (b[3].x + b[3].y), s * (b[3].x - b[3].y))
                  : make_float2(s * (b[3].y - b[3].x), -s * (b[3].x + b[3].y));
  v[0] = cadd(a[0], b[0]);
  v[4] = csub(a[0], b[0]);
  v[1] = cadd(a[1], b1);
  v[5] = csub(a[1], b1);
  v[2] = cadd(a[2], b2);
  v[6] = csub(a[2], b2);
  v[3] = cadd(a[3], b3);
  v[7] = csub(a[3], b3);
}
template <bool INV>
__device__ __forceinline__ void dft3(float2* v) {
  const float s3 = 0.86602540378443864676f;
  float2 t = cadd(v[1], v[2]);
  float2 u = mul_mi<INV>(cscale(csub(v[1], v[2]), s3));
  float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
  v[0] = cadd(v[0], t);
  v[1] = cadd(m, u);
  v[2] = csub(m, u);
}
template <bool INV>
__device__ __forceinline__ void dft5(float2* v) {
  const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
  const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
  float2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
  float2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  float2 a1 = make_float2(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
  float2 a2 = make_float2(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
  float2 b1 = mul_mi<INV>(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
  float2 b2 = mul_mi<INV>(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
  v[0] = make_float2(v[0].x + t1.x + t2.x, v[0].y + t1.y + t2.y);
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}

template <int R, bool INV>
__device__ __forceinline__ void dftR(float2* v) {
  if constexpr (R == 2) dft2<INV>(v);
  else if constexpr (R == 3) dft3<INV>(v);
  else if constexpr (R == 4) dft4<INV>(v);
  else if constexpr (R == 5) dft5<INV>(v);
  else dft8<INV>(v);
}

}  // namespace dh
