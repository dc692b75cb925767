// b2g_math.h -- single-precision 2D algebra for the device kernels.
//
// The operations and their evaluation order follow include/box2d/b2_math.h of
// the reference (b2Vec2, b2Rot, b2Transform, b2Mat22), so that kernels built
// with -ffp-contract=off produce the same single-precision results as the
// reference on the same inputs. sin/cos/sqrt are the HIP device library's.
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

#define B2G_INLINE __device__ __host__ __forceinline__

namespace b2g {

// b2_common.h:110-182 / b2_settings.h constants (meters, seconds)
constexpr float kEpsilon = FLT_EPSILON;
constexpr float kPi = 3.14159265359f;
constexpr float kLinearSlop = 0.005f;
constexpr float kAngularSlop = 2.0f / 180.0f * kPi;
constexpr float kPolygonRadius = 2.0f * kLinearSlop;
constexpr float kMaxLinearCorrection = 0.2f;
constexpr float kMaxAngularCorrection = 8.0f / 180.0f * kPi;
constexpr float kMaxTranslation = 2.0f;
constexpr float kMaxTranslationSq = kMaxTranslation * kMaxTranslation;
constexpr float kMaxRotation = 0.5f * kPi;
constexpr float kMaxRotationSq = kMaxRotation * kMaxRotation;
constexpr float kBaumgarte = 0.2f;

struct vec2 {
  float x, y;
};

B2G_INLINE vec2 mk(float x, float y) { return vec2{x, y}; }
B2G_INLINE vec2 operator+(vec2 a, vec2 b) { return vec2{a.x + b.x, a.y + b.y}; }
B2G_INLINE vec2 operator-(vec2 a, vec2 b) { return vec2{a.x - b.x, a.y - b.y}; }
B2G_INLINE vec2 operator-(vec2 a) { return vec2{-a.x, -a.y}; }
B2G_INLINE vec2 operator*(float s, vec2 a) { return vec2{s * a.x, s * a.y}; }
B2G_INLINE float dot(vec2 a, vec2 b) { return a.x * b.x + a.y * b.y; }
B2G_INLINE float cross(vec2 a, vec2 b) { return a.x * b.y - a.y * b.x; }
B2G_INLINE vec2 cross(vec2 a, float s) { return vec2{s * a.y, -s * a.x}; }
B2G_INLINE vec2 cross(float s, vec2 a) { return vec2{-s * a.y, s * a.x}; }
B2G_INLINE float length_sq(vec2 a) { return a.x * a.x + a.y * a.y; }
B2G_INLINE float dist_sq(vec2 a, vec2 b) {
  vec2 c = a - b;
  return dot(c, c);
}
B2G_INLINE float minf(float a, float b) { return a < b ? a : b; }
B2G_INLINE float maxf(float a, float b) { return a > b ? a : b; }
B2G_INLINE float clampf(float a, float lo, float hi) { return maxf(lo, minf(a, hi)); }
B2G_INLINE float absf(float a) { return a > 0.0f ? a : -a; }
B2G_INLINE vec2 vmin(vec2 a, vec2 b) { return vec2{minf(a.x, b.x), minf(a.y, b.y)}; }
B2G_INLINE vec2 vmax(vec2 a, vec2 b) { return vec2{maxf(a.x, b.x), maxf(a.y, b.y)}; }

B2G_INLINE float length(vec2 a) { return sqrtf(a.x * a.x + a.y * a.y); }

// b2Vec2::Normalize: leaves vectors shorter than epsilon unchanged
B2G_INLINE vec2 normalize(vec2 v) {
  const float l = length(v);
  if (l < kEpsilon) return v;
  const float inv = 1.0f / l;
  return vec2{v.x * inv, v.y * inv};
}

struct rot2 {
  float s, c;
};

__device__ __forceinline__ rot2 rot_of(float angle) {
  rot2 r;
  sincosf(angle, &r.s, &r.c);
  return r;
}
B2G_INLINE vec2 mul(rot2 q, vec2 v) { return vec2{q.c * v.x - q.s * v.y, q.s * v.x + q.c * v.y}; }
B2G_INLINE vec2 mul_t(rot2 q, vec2 v) { return vec2{q.c * v.x + q.s * v.y, -q.s * v.x + q.c * v.y}; }
B2G_INLINE rot2 mul_t(rot2 q, rot2 r) { return rot2{q.c * r.s - q.s * r.c, q.c * r.c + q.s * r.s}; }

struct xform {
  vec2 p;
  rot2 q;
};

B2G_INLINE vec2 mul(const xform& t, vec2 v) {
  const float x = (t.q.c * v.x - t.q.s * v.y) + t.p.x;
  const float y = (t.q.s * v.x + t.q.c * v.y) + t.p.y;
  return vec2{x, y};
}
B2G_INLINE vec2 mul_t(const xform& t, vec2 v) {
  const float px = v.x - t.p.x;
  const float py = v.y - t.p.y;
  return vec2{t.q.c * px + t.q.s * py, -t.q.s * px + t.q.c * py};
}
B2G_INLINE xform mul_t(const xform& a, const xform& b) {
  xform c;
  c.q = mul_t(a.q, b.q);
  c.p = mul_t(a.q, b.p - a.p);
  return c;
}
B2G_INLINE xform xform_of(float4 f) { return xform{vec2{f.x, f.y}, rot2{f.z, f.w}}; }

// b2Mat22 with the reference's GetInverse / Solve
struct mat22 {
  vec2 ex, ey;
};
B2G_INLINE mat22 inverse(const mat22& m) {
  const float a = m.ex.x, b = m.ey.x, c = m.ex.y, d = m.ey.y;
  float det = a * d - b * c;
  if (det != 0.0f) det = 1.0f / det;
  mat22 r;
  r.ex.x = det * d;
  r.ey.x = -det * b;
  r.ex.y = -det * c;
  r.ey.y = det * a;
  return r;
}
B2G_INLINE vec2 solve(const mat22& m, vec2 b) {
  const float a11 = m.ex.x, a12 = m.ey.x, a21 = m.ex.y, a22 = m.ey.y;
  float det = a11 * a22 - a12 * a21;
  if (det != 0.0f) det = 1.0f / det;
  return vec2{det * (a22 * b.x - a12 * b.y), det * (a11 * b.y - a21 * b.x)};
}
B2G_INLINE vec2 mul(const mat22& a, vec2 v) {
  return vec2{a.ex.x * v.x + a.ey.x * v.y, a.ex.y * v.x + a.ey.y * v.y};
}

B2G_INLINE float as_float(int32_t i) { return __builtin_bit_cast(float, i); }
B2G_INLINE int32_t as_int(float f) { return __builtin_bit_cast(int32_t, f); }
B2G_INLINE uint32_t as_uint(float f) { return __builtin_bit_cast(uint32_t, f); }
B2G_INLINE float uint_as_float(uint32_t u) { return __builtin_bit_cast(float, u); }

}  // namespace b2g
