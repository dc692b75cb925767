// restate.cpp -- CPU restatement of the b2World::Step hot path (test infrastructure).
//
// See restate.h for scope and ordering. Arithmetic follows the reference's
// single-precision expressions operation by operation (compiled with
// -ffp-contract=off, so no fused multiply-adds are introduced), which makes
// the narrowphase and the shape set-up bit-identical to the reference on the
// same inputs. Citations are file:line in the reference tree.
#include "restate.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Math (include/box2d/b2_math.h)
// ---------------------------------------------------------------------------
const float kEps = FLT_EPSILON;                         // b2_common.h:40
const float kPi = 3.14159265359f;                       // b2_common.h:41
const float kLinearSlop = 0.005f;                       // b2_common.h:127
const float kAngularSlop = 2.0f / 180.0f * kPi;         // b2_common.h:132
const float kPolygonRadius = 2.0f * kLinearSlop;        // b2_common.h:137
const float kMaxLinearCorrection = 0.2f;                // b2_common.h:150
const float kMaxAngularCorrection = 8.0f / 180.0f * kPi;
const float kTimeToSleep = 0.5f;                          // b2_common.h:176
const float kLinearSleepTolerance = 0.01f * 1.0f;         // b2_common.h:179 (lengthUnitsPerMeter 1)
const float kAngularSleepTolerance = 2.0f / 180.0f * 3.14159265359f;  // b2_common.h:182 (b2_pi)
const float kMaxTranslation = 2.0f;
const float kMaxTranslationSq = kMaxTranslation * kMaxTranslation;
const float kMaxRotation = 0.5f * kPi;
const float kMaxRotationSq = kMaxRotation * kMaxRotation;
const float kBaumgarte = 0.2f;

struct V2 {
  float x, y;
};
inline V2 v2(float x, float y) { return V2{x, y}; }
inline V2 operator+(V2 a, V2 b) { return V2{a.x + b.x, a.y + b.y}; }
inline V2 operator-(V2 a, V2 b) { return V2{a.x - b.x, a.y - b.y}; }
inline V2 operator-(V2 a) { return V2{-a.x, -a.y}; }
inline V2 operator*(float s, V2 a) { return V2{s * a.x, s * a.y}; }
inline float dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
inline float cross(V2 a, V2 b) { return a.x * b.y - a.y * b.x; }
inline V2 cross(V2 a, float s) { return V2{s * a.y, -s * a.x}; }
inline V2 cross(float s, V2 a) { return V2{-s * a.y, s * a.x}; }
inline float len(V2 a) { return std::sqrt(a.x * a.x + a.y * a.y); }
inline float len_sq(V2 a) { return a.x * a.x + a.y * a.y; }
inline float dist_sq(V2 a, V2 b) { V2 c = a - b; return dot(c, c); }
inline V2 vmin(V2 a, V2 b) { return V2{a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y}; }
inline V2 vmax(V2 a, V2 b) { return V2{a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y}; }
inline float fmin_(float a, float b) { return a < b ? a : b; }
inline float fmax_(float a, float b) { return a > b ? a : b; }
inline float clampf(float a, float lo, float hi) { return fmax_(lo, fmin_(a, hi)); }
inline float fabs_(float a) { return a > 0.0f ? a : -a; }
// b2Vec2::Normalize (b2_math.h:117-129)
inline V2 normalize(V2 v) {
  float l = len(v);
  if (l < kEps) return v;
  float inv = 1.0f / l;
  return V2{v.x * inv, v.y * inv};
}

struct Rot {
  float s, c;
};
inline Rot rot(float angle) { return Rot{std::sin(angle), std::cos(angle)}; }  // b2_math.h:313-318
inline V2 mul(Rot q, V2 v) { return V2{q.c * v.x - q.s * v.y, q.s * v.x + q.c * v.y}; }
inline V2 mulT(Rot q, V2 v) { return V2{q.c * v.x + q.s * v.y, -q.s * v.x + q.c * v.y}; }
inline Rot mulT(Rot q, Rot r) { return Rot{q.c * r.s - q.s * r.c, q.c * r.c + q.s * r.s}; }

struct Xf {
  V2 p;
  Rot q;
};
inline V2 mul(const Xf& T, V2 v) {  // b2_math.h:591-597
  float x = (T.q.c * v.x - T.q.s * v.y) + T.p.x;
  float y = (T.q.s * v.x + T.q.c * v.y) + T.p.y;
  return V2{x, y};
}
inline V2 mulT(const Xf& T, V2 v) {  // b2_math.h:599-607
  float px = v.x - T.p.x;
  float py = v.y - T.p.y;
  float x = (T.q.c * px + T.q.s * py);
  float y = (-T.q.s * px + T.q.c * py);
  return V2{x, y};
}
inline Xf mulT(const Xf& A, const Xf& B) {  // b2_math.h:621-627
  Xf C;
  C.q = mulT(A.q, B.q);
  C.p = mulT(A.q, B.p - A.p);
  return C;
}
inline Xf xf_from(const float* f) { return Xf{V2{f[0], f[1]}, Rot{f[2], f[3]}}; }

struct Mat22 {
  V2 ex, ey;
  // b2Mat22::GetInverse (b2_math.h:224-236)
  Mat22 inverse() const {
    float a = ex.x, b = ey.x, c = ex.y, d = ey.y;
    float det = a * d - b * c;
    if (det != 0.0f) det = 1.0f / det;
    Mat22 B;
    B.ex.x = det * d;
    B.ey.x = -det * b;
    B.ex.y = -det * c;
    B.ey.y = det * a;
    return B;
  }
  // b2Mat22::Solve (b2_math.h:240-252)
  V2 solve(V2 b) const {
    float a11 = ex.x, a12 = ey.x, a21 = ex.y, a22 = ey.y;
    float det = a11 * a22 - a12 * a21;
    if (det != 0.0f) det = 1.0f / det;
    return V2{det * (a22 * b.x - a12 * b.y), det * (a11 * b.y - a21 * b.x)};
  }
};
inline V2 mul(const Mat22& A, V2 v) { return V2{A.ex.x * v.x + A.ey.x * v.y, A.ex.y * v.x + A.ey.y * v.y}; }

// ---------------------------------------------------------------------------
// Shapes (src/collision/b2_polygon_shape.cpp, b2_circle_shape.cpp, b2_edge_shape.cpp)
// ---------------------------------------------------------------------------
struct Poly {
  int count;
  V2 v[8], n[8];
  V2 centroid;
  float radius;
};

void set_box(Poly* p, float hx, float hy) {  // b2_polygon_shape.cpp:36-48
  p->count = 4;
  p->v[0] = v2(-hx, -hy); p->v[1] = v2(hx, -hy); p->v[2] = v2(hx, hy); p->v[3] = v2(-hx, hy);
  p->n[0] = v2(0.0f, -1.0f); p->n[1] = v2(1.0f, 0.0f); p->n[2] = v2(0.0f, 1.0f); p->n[3] = v2(-1.0f, 0.0f);
  p->centroid = v2(0.0f, 0.0f);
}

void set_box(Poly* p, float hx, float hy, V2 center, float angle) {  // :50-73
  set_box(p, hx, hy);
  p->centroid = center;
  Xf xf{center, rot(angle)};
  for (int i = 0; i < 4; ++i) {
    p->v[i] = mul(xf, p->v[i]);
    p->n[i] = mul(xf.q, p->n[i]);
  }
}

V2 centroid_of(const V2* vs, int count) {  // :75-111
  V2 c = v2(0.0f, 0.0f);
  float area = 0.0f;
  V2 s = vs[0];
  const float inv3 = 1.0f / 3.0f;
  for (int i = 0; i < count; ++i) {
    V2 p1 = vs[0] - s;
    V2 p2 = vs[i] - s;
    V2 p3 = i + 1 < count ? vs[i + 1] - s : vs[0] - s;
    V2 e1 = p2 - p1;
    V2 e2 = p3 - p1;
    float D = cross(e1, e2);
    float tri = 0.5f * D;
    area += tri;
    c = c + (tri * inv3) * (p1 + p2 + p3);
  }
  c = (1.0f / area) * c + s;
  return c;
}

void set_hull(Poly* p, const V2* pts, int count) {  // :113-243 (weld, gift wrap, normals)
  if (count < 3) { set_box(p, 1.0f, 1.0f); return; }
  int n = count < 8 ? count : 8;
  V2 ps[8];
  int temp = 0;
  for (int i = 0; i < n; ++i) {
    V2 v = pts[i];
    bool unique = true;
    for (int j = 0; j < temp; ++j) {
      if (dist_sq(v, ps[j]) < ((0.5f * kLinearSlop) * (0.5f * kLinearSlop))) { unique = false; break; }
    }
    if (unique) ps[temp++] = v;
  }
  n = temp;
  if (n < 3) { set_box(p, 1.0f, 1.0f); return; }
  int i0 = 0;
  float x0 = ps[0].x;
  for (int i = 1; i < n; ++i) {
    float x = ps[i].x;
    if (x > x0 || (x == x0 && ps[i].y < ps[i0].y)) { i0 = i; x0 = x; }
  }
  int hull[8];
  int m = 0, ih = i0;
  for (;;) {
    if (m >= 8) break;  // cannot happen for n <= 8 distinct points
    hull[m] = ih;
    int ie = 0;
    for (int j = 1; j < n; ++j) {
      if (ie == ih) { ie = j; continue; }
      V2 r = ps[ie] - ps[hull[m]];
      V2 v = ps[j] - ps[hull[m]];
      float c = cross(r, v);
      if (c < 0.0f) ie = j;
      if (c == 0.0f && len_sq(v) > len_sq(r)) ie = j;
    }
    ++m;
    ih = ie;
    if (ie == i0) break;
  }
  if (m < 3) { set_box(p, 1.0f, 1.0f); return; }
  p->count = m;
  for (int i = 0; i < m; ++i) p->v[i] = ps[hull[i]];
  for (int i = 0; i < m; ++i) {
    int i2 = i + 1 < m ? i + 1 : 0;
    V2 edge = p->v[i2] - p->v[i];
    p->n[i] = normalize(cross(edge, 1.0f));
  }
  p->centroid = centroid_of(p->v, m);
}

struct MassData {
  float mass;
  V2 center;
  float I;
};

MassData poly_mass(const Poly& p, float density) {  // b2_polygon_shape.cpp:390-464
  V2 center = v2(0.0f, 0.0f);
  float area = 0.0f, I = 0.0f;
  V2 s = p.v[0];
  const float k_inv3 = 1.0f / 3.0f;
  for (int i = 0; i < p.count; ++i) {
    V2 e1 = p.v[i] - s;
    V2 e2 = i + 1 < p.count ? p.v[i + 1] - s : p.v[0] - s;
    float D = cross(e1, e2);
    float tri = 0.5f * D;
    area += tri;
    center = center + (tri * k_inv3) * (e1 + e2);
    float ex1 = e1.x, ey1 = e1.y, ex2 = e2.x, ey2 = e2.y;
    float intx2 = ex1 * ex1 + ex2 * ex1 + ex2 * ex2;
    float inty2 = ey1 * ey1 + ey2 * ey1 + ey2 * ey2;
    I += (0.25f * k_inv3 * D) * (intx2 + inty2);
  }
  MassData md;
  md.mass = density * area;
  center = (1.0f / area) * center;
  md.center = center + s;
  md.I = density * I;
  md.I += md.mass * (dot(md.center, md.center) - dot(center, center));
  return md;
}

MassData circle_mass(float r, V2 p, float density) {  // b2_circle_shape.cpp:98-105
  MassData md;
  md.mass = density * kPi * r * r;
  md.center = p;
  md.I = md.mass * (0.5f * r * r + dot(p, p));
  return md;
}

void shape_aabb(const b2g_shape& s, const Xf& xf, float* out) {
  if (s.type == B2G_CIRCLE) {  // b2_circle_shape.cpp:91-96
    V2 p = xf.p + mul(xf.q, v2(s.centroid[0], s.centroid[1]));
    out[0] = p.x - s.radius; out[1] = p.y - s.radius;
    out[2] = p.x + s.radius; out[3] = p.y + s.radius;
  } else if (s.type == B2G_EDGE) {  // b2_edge_shape.cpp:156-167
    V2 a = mul(xf, v2(s.vertices[2], s.vertices[3]));
    V2 b = mul(xf, v2(s.vertices[4], s.vertices[5]));
    V2 lo = vmin(a, b), hi = vmax(a, b);
    out[0] = lo.x - s.radius; out[1] = lo.y - s.radius;
    out[2] = hi.x + s.radius; out[3] = hi.y + s.radius;
  } else {  // b2_polygon_shape.cpp:373-388
    V2 lo = mul(xf, v2(s.vertices[0], s.vertices[1]));
    V2 hi = lo;
    for (int i = 1; i < s.count; ++i) {
      V2 v = mul(xf, v2(s.vertices[2 * i], s.vertices[2 * i + 1]));
      lo = vmin(lo, v);
      hi = vmax(hi, v);
    }
    out[0] = lo.x - s.radius; out[1] = lo.y - s.radius;
    out[2] = hi.x + s.radius; out[3] = hi.y + s.radius;
  }
}

// ---------------------------------------------------------------------------
// Narrowphase (src/collision/b2_collide_*.cpp, b2_collision.cpp)
// ---------------------------------------------------------------------------
struct Feature {
  uint8_t indexA, indexB, typeA, typeB;
};
inline uint32_t fkey(Feature f) {  // b2ContactID union: cf bytes in memory order
  return (uint32_t)f.indexA | ((uint32_t)f.indexB << 8) | ((uint32_t)f.typeA << 16) |
         ((uint32_t)f.typeB << 24);
}
const uint8_t kVertex = 0, kFace = 1;

struct MPoint {
  V2 lp;
  float ni, ti;
  uint32_t id;
};
struct Manifold {
  MPoint pts[2];
  V2 ln, lp;
  int type;  // 0 circles, 1 faceA, 2 faceB
  int count;
};

struct ClipV {
  V2 v;
  Feature id;
};

Poly poly_of(const b2g_shape& s) {
  Poly p;
  p.count = s.count;
  p.radius = s.radius;
  p.centroid = v2(s.centroid[0], s.centroid[1]);
  for (int i = 0; i < s.count; ++i) {
    p.v[i] = v2(s.vertices[2 * i], s.vertices[2 * i + 1]);
    p.n[i] = v2(s.normals[2 * i], s.normals[2 * i + 1]);
  }
  return p;
}

void collide_circles(Manifold* m, const b2g_shape& A, const Xf& xfA, const b2g_shape& B,
                     const Xf& xfB) {  // b2_collide_circle.cpp:27-53
  m->count = 0;
  V2 pA = mul(xfA, v2(A.centroid[0], A.centroid[1]));
  V2 pB = mul(xfB, v2(B.centroid[0], B.centroid[1]));
  V2 d = pB - pA;
  float distSqr = dot(d, d);
  float radius = A.radius + B.radius;
  if (distSqr > radius * radius) return;
  m->type = 0;
  m->lp = v2(A.centroid[0], A.centroid[1]);
  m->ln = v2(0.0f, 0.0f);
  m->count = 1;
  m->pts[0].lp = v2(B.centroid[0], B.centroid[1]);
  m->pts[0].id = 0;
}

void collide_poly_circle(Manifold* m, const b2g_shape& A, const Xf& xfA, const b2g_shape& B,
                         const Xf& xfB) {  // b2_collide_circle.cpp:55-158
  m->count = 0;
  V2 cp = v2(B.centroid[0], B.centroid[1]);
  V2 c = mul(xfB, cp);
  V2 cLocal = mulT(xfA, c);
  int normalIndex = 0;
  float separation = -FLT_MAX;
  float radius = A.radius + B.radius;
  int vc = A.count;
  Poly P = poly_of(A);
  for (int i = 0; i < vc; ++i) {
    float s = dot(P.n[i], cLocal - P.v[i]);
    if (s > radius) return;
    if (s > separation) { separation = s; normalIndex = i; }
  }
  int i1 = normalIndex, i2 = i1 + 1 < vc ? i1 + 1 : 0;
  V2 v1 = P.v[i1], vv2 = P.v[i2];
  if (separation < kEps) {
    m->count = 1; m->type = 1;
    m->ln = P.n[normalIndex];
    m->lp = 0.5f * (v1 + vv2);
    m->pts[0].lp = cp; m->pts[0].id = 0;
    return;
  }
  float u1 = dot(cLocal - v1, vv2 - v1);
  float u2 = dot(cLocal - vv2, v1 - vv2);
  if (u1 <= 0.0f) {
    if (dist_sq(cLocal, v1) > radius * radius) return;
    m->count = 1; m->type = 1;
    m->ln = normalize(cLocal - v1);
    m->lp = v1;
    m->pts[0].lp = cp; m->pts[0].id = 0;
  } else if (u2 <= 0.0f) {
    if (dist_sq(cLocal, vv2) > radius * radius) return;
    m->count = 1; m->type = 1;
    m->ln = normalize(cLocal - vv2);
    m->lp = vv2;
    m->pts[0].lp = cp; m->pts[0].id = 0;
  } else {
    V2 faceCenter = 0.5f * (v1 + vv2);
    float s = dot(cLocal - faceCenter, P.n[i1]);
    if (s > radius) return;
    m->count = 1; m->type = 1;
    m->ln = P.n[i1];
    m->lp = faceCenter;
    m->pts[0].lp = cp; m->pts[0].id = 0;
  }
}

// b2ClipSegmentToLine (b2_collision.cpp:205-237)
int clip_segment(ClipV vOut[2], const ClipV vIn[2], V2 normal, float offset, int vertexIndexA) {
  int count = 0;
  float d0 = dot(normal, vIn[0].v) - offset;
  float d1 = dot(normal, vIn[1].v) - offset;
  if (d0 <= 0.0f) vOut[count++] = vIn[0];
  if (d1 <= 0.0f) vOut[count++] = vIn[1];
  if (d0 * d1 < 0.0f) {
    float interp = d0 / (d0 - d1);
    vOut[count].v = vIn[0].v + interp * (vIn[1].v - vIn[0].v);
    vOut[count].id.indexA = (uint8_t)vertexIndexA;
    vOut[count].id.indexB = vIn[0].id.indexB;
    vOut[count].id.typeA = kVertex;
    vOut[count].id.typeB = kFace;
    ++count;
  }
  return count;
}

// b2FindMaxSeparation (b2_collide_polygon.cpp:27-66)
float find_max_separation(int* edgeIndex, const Poly& p1, const Xf& xf1, const Poly& p2,
                          const Xf& xf2) {
  Xf xf = mulT(xf2, xf1);
  int best = 0;
  float maxSep = -FLT_MAX;
  for (int i = 0; i < p1.count; ++i) {
    V2 n = mul(xf.q, p1.n[i]);
    V2 v1 = mul(xf, p1.v[i]);
    float si = FLT_MAX;
    for (int j = 0; j < p2.count; ++j) {
      float sij = dot(n, p2.v[j] - v1);
      if (sij < si) si = sij;
    }
    if (si > maxSep) { maxSep = si; best = i; }
  }
  *edgeIndex = best;
  return maxSep;
}

// b2FindIncidentEdge (b2_collide_polygon.cpp:68-111)
void find_incident_edge(ClipV c[2], const Poly& p1, const Xf& xf1, int edge1, const Poly& p2,
                        const Xf& xf2) {
  V2 normal1 = mulT(xf2.q, mul(xf1.q, p1.n[edge1]));
  int index = 0;
  float minDot = FLT_MAX;
  for (int i = 0; i < p2.count; ++i) {
    float d = dot(normal1, p2.n[i]);
    if (d < minDot) { minDot = d; index = i; }
  }
  int i1 = index, i2 = i1 + 1 < p2.count ? i1 + 1 : 0;
  c[0].v = mul(xf2, p2.v[i1]);
  c[0].id = Feature{(uint8_t)edge1, (uint8_t)i1, kFace, kVertex};
  c[1].v = mul(xf2, p2.v[i2]);
  c[1].id = Feature{(uint8_t)edge1, (uint8_t)i2, kFace, kVertex};
}

// b2CollidePolygons (b2_collide_polygon.cpp:120-243)
void collide_polygons(Manifold* m, const b2g_shape& A, const Xf& xfA, const b2g_shape& B,
                      const Xf& xfB) {
  m->count = 0;
  Poly polyA = poly_of(A), polyB = poly_of(B);
  float totalRadius = polyA.radius + polyB.radius;
  int edgeA = 0;
  float sepA = find_max_separation(&edgeA, polyA, xfA, polyB, xfB);
  if (sepA > totalRadius) return;
  int edgeB = 0;
  float sepB = find_max_separation(&edgeB, polyB, xfB, polyA, xfA);
  if (sepB > totalRadius) return;
  const Poly *poly1, *poly2;
  Xf xf1, xf2;
  int edge1;
  uint8_t flip;
  const float k_tol = 0.1f * kLinearSlop;
  if (sepB > sepA + k_tol) {
    poly1 = &polyB; poly2 = &polyA; xf1 = xfB; xf2 = xfA; edge1 = edgeB; m->type = 2; flip = 1;
  } else {
    poly1 = &polyA; poly2 = &polyB; xf1 = xfA; xf2 = xfB; edge1 = edgeA; m->type = 1; flip = 0;
  }
  ClipV incident[2];
  find_incident_edge(incident, *poly1, xf1, edge1, *poly2, xf2);
  int count1 = poly1->count;
  int iv1 = edge1, iv2 = edge1 + 1 < count1 ? edge1 + 1 : 0;
  V2 v11 = poly1->v[iv1], v12 = poly1->v[iv2];
  V2 localTangent = normalize(v12 - v11);
  V2 localNormal = cross(localTangent, 1.0f);
  V2 planePoint = 0.5f * (v11 + v12);
  V2 tangent = mul(xf1.q, localTangent);
  V2 normal = cross(tangent, 1.0f);
  v11 = mul(xf1, v11);
  v12 = mul(xf1, v12);
  float frontOffset = dot(normal, v11);
  float sideOffset1 = -dot(tangent, v11) + totalRadius;
  float sideOffset2 = dot(tangent, v12) + totalRadius;
  ClipV cp1[2], cp2[2];
  int np = clip_segment(cp1, incident, -tangent, sideOffset1, iv1);
  if (np < 2) return;
  np = clip_segment(cp2, cp1, tangent, sideOffset2, iv2);
  if (np < 2) return;
  m->ln = localNormal;
  m->lp = planePoint;
  int pc = 0;
  for (int i = 0; i < 2; ++i) {
    float separation = dot(normal, cp2[i].v) - frontOffset;
    if (separation <= totalRadius) {
      MPoint* cp = m->pts + pc;
      cp->lp = mulT(xf2, cp2[i].v);
      Feature f = cp2[i].id;
      if (flip) f = Feature{f.indexB, f.indexA, f.typeB, f.typeA};
      cp->id = fkey(f);
      ++pc;
    }
  }
  m->count = pc;
}

// b2CollideEdgeAndCircle (b2_collide_edge.cpp:31-165)
void collide_edge_circle(Manifold* m, const b2g_shape& E, const Xf& xfA, const b2g_shape& C,
                         const Xf& xfB) {
  m->count = 0;
  V2 cp = v2(C.centroid[0], C.centroid[1]);
  V2 Q = mulT(xfA, mul(xfB, cp));
  V2 A = v2(E.vertices[2], E.vertices[3]), B = v2(E.vertices[4], E.vertices[5]);
  V2 e = B - A;
  V2 n = v2(e.y, -e.x);
  float offset = dot(n, Q - A);
  bool oneSided = E.count != 0;
  if (oneSided && offset < 0.0f) return;
  float u = dot(e, B - Q);
  float v = dot(e, Q - A);
  float radius = E.radius + C.radius;
  Feature cf;
  cf.indexB = 0;
  cf.typeB = kVertex;
  if (v <= 0.0f) {
    V2 P = A;
    V2 d = Q - P;
    if (dot(d, d) > radius * radius) return;
    if (oneSided) {
      V2 A1 = v2(E.vertices[0], E.vertices[1]);
      V2 B1 = A;
      V2 e1 = B1 - A1;
      float u1 = dot(e1, B1 - Q);
      if (u1 > 0.0f) return;
    }
    cf.indexA = 0; cf.typeA = kVertex;
    m->count = 1; m->type = 0; m->ln = v2(0.0f, 0.0f); m->lp = P;
    m->pts[0].id = fkey(cf); m->pts[0].lp = cp;
    return;
  }
  if (u <= 0.0f) {
    V2 P = B;
    V2 d = Q - P;
    if (dot(d, d) > radius * radius) return;
    if (oneSided) {
      V2 B2 = v2(E.vertices[6], E.vertices[7]);
      V2 A2 = B;
      V2 e2 = B2 - A2;
      float vv = dot(e2, Q - A2);
      if (vv > 0.0f) return;
    }
    cf.indexA = 1; cf.typeA = kVertex;
    m->count = 1; m->type = 0; m->ln = v2(0.0f, 0.0f); m->lp = P;
    m->pts[0].id = fkey(cf); m->pts[0].lp = cp;
    return;
  }
  float den = dot(e, e);
  V2 P = (1.0f / den) * (u * A + v * B);
  V2 d = Q - P;
  if (dot(d, d) > radius * radius) return;
  if (offset < 0.0f) n = v2(-n.x, -n.y);
  n = normalize(n);
  cf.indexA = 0; cf.typeA = kFace;
  m->count = 1; m->type = 1; m->ln = n; m->lp = A;
  m->pts[0].id = fkey(cf); m->pts[0].lp = cp;
}

struct EPAxis {
  V2 normal;
  int type;  // 0 unknown, 1 edgeA, 2 edgeB
  int index;
  float separation;
};

// b2CollideEdgeAndPolygon (b2_collide_edge.cpp:205-524)
void collide_edge_polygon(Manifold* m, const b2g_shape& E, const Xf& xfA, const b2g_shape& Bs,
                          const Xf& xfB) {
  m->count = 0;
  Poly polyB = poly_of(Bs);
  Xf xf = mulT(xfA, xfB);
  V2 centroidB = mul(xf, polyB.centroid);
  V2 v1 = v2(E.vertices[2], E.vertices[3]), vv2 = v2(E.vertices[4], E.vertices[5]);
  V2 edge1 = normalize(vv2 - v1);
  V2 normal1 = v2(edge1.y, -edge1.x);
  float offset1 = dot(normal1, centroidB - v1);
  bool oneSided = E.count != 0;
  if (oneSided && offset1 < 0.0f) return;
  V2 tv[8], tn[8];
  int tc = polyB.count;
  for (int i = 0; i < tc; ++i) {
    tv[i] = mul(xf, polyB.v[i]);
    tn[i] = mul(xf.q, polyB.n[i]);
  }
  float radius = polyB.radius + E.radius;
  // b2ComputeEdgeSeparation (:205-239)
  EPAxis edgeAxis;
  edgeAxis.type = 1; edgeAxis.index = -1; edgeAxis.separation = -FLT_MAX; edgeAxis.normal = v2(0.0f, 0.0f);
  {
    V2 axes[2] = {normal1, -normal1};
    for (int j = 0; j < 2; ++j) {
      float sj = FLT_MAX;
      for (int i = 0; i < tc; ++i) {
        float si = dot(axes[j], tv[i] - v1);
        if (si < sj) sj = si;
      }
      if (sj > edgeAxis.separation) { edgeAxis.index = j; edgeAxis.separation = sj; edgeAxis.normal = axes[j]; }
    }
  }
  if (edgeAxis.separation > radius) return;
  // b2ComputePolygonSeparation (:241-267)
  EPAxis polyAxis;
  polyAxis.type = 0; polyAxis.index = -1; polyAxis.separation = -FLT_MAX; polyAxis.normal = v2(0.0f, 0.0f);
  for (int i = 0; i < tc; ++i) {
    V2 n = -tn[i];
    float s1 = dot(n, tv[i] - v1);
    float s2 = dot(n, tv[i] - vv2);
    float s = fmin_(s1, s2);
    if (s > polyAxis.separation) { polyAxis.type = 2; polyAxis.index = i; polyAxis.separation = s; polyAxis.normal = n; }
  }
  if (polyAxis.separation > radius) return;
  const float k_relativeTol = 0.98f, k_absoluteTol = 0.001f;
  EPAxis primary;
  if (polyAxis.separation - radius > k_relativeTol * (edgeAxis.separation - radius) + k_absoluteTol)
    primary = polyAxis;
  else
    primary = edgeAxis;
  if (oneSided) {  // :332-387 smooth collision
    V2 edge0 = normalize(v1 - v2(E.vertices[0], E.vertices[1]));
    V2 normal0 = v2(edge0.y, -edge0.x);
    bool convex1 = cross(edge0, edge1) >= 0.0f;
    V2 edge2 = normalize(v2(E.vertices[6], E.vertices[7]) - vv2);
    V2 normal2 = v2(edge2.y, -edge2.x);
    bool convex2 = cross(edge1, edge2) >= 0.0f;
    const float sinTol = 0.1f;
    bool side1 = dot(primary.normal, edge1) <= 0.0f;
    if (side1) {
      if (convex1) {
        if (cross(primary.normal, normal0) > sinTol) return;
      } else {
        primary = edgeAxis;
      }
    } else {
      if (convex2) {
        if (cross(normal2, primary.normal) > sinTol) return;
      } else {
        primary = edgeAxis;
      }
    }
  }
  ClipV clip[2];
  int ri1, ri2;
  V2 rv1, rv2, rnormal, side1n, side2n;
  if (primary.type == 1) {
    m->type = 1;
    int best = 0;
    float bestValue = dot(primary.normal, tn[0]);
    for (int i = 1; i < tc; ++i) {
      float value = dot(primary.normal, tn[i]);
      if (value < bestValue) { bestValue = value; best = i; }
    }
    int i1 = best, i2 = i1 + 1 < tc ? i1 + 1 : 0;
    clip[0].v = tv[i1];
    clip[0].id = Feature{0, (uint8_t)i1, kFace, kVertex};
    clip[1].v = tv[i2];
    clip[1].id = Feature{0, (uint8_t)i2, kFace, kVertex};
    ri1 = 0; ri2 = 1; rv1 = v1; rv2 = vv2; rnormal = primary.normal;
    side1n = -edge1; side2n = edge1;
  } else {
    m->type = 2;
    clip[0].v = vv2;
    clip[0].id = Feature{1, (uint8_t)primary.index, kVertex, kFace};
    clip[1].v = v1;
    clip[1].id = Feature{0, (uint8_t)primary.index, kVertex, kFace};
    ri1 = primary.index;
    ri2 = ri1 + 1 < tc ? ri1 + 1 : 0;
    rv1 = tv[ri1]; rv2 = tv[ri2]; rnormal = tn[ri1];
    side1n = v2(rnormal.y, -rnormal.x);
    side2n = -side1n;
  }
  float sideOffset1 = dot(side1n, rv1);
  float sideOffset2 = dot(side2n, rv2);
  ClipV cp1[2], cp2[2];
  int np = clip_segment(cp1, clip, side1n, sideOffset1, ri1);
  if (np < 2) return;
  np = clip_segment(cp2, cp1, side2n, sideOffset2, ri2);
  if (np < 2) return;
  if (primary.type == 1) {
    m->ln = rnormal; m->lp = rv1;
  } else {
    m->ln = polyB.n[ri1]; m->lp = polyB.v[ri1];
  }
  int pc = 0;
  for (int i = 0; i < 2; ++i) {
    float separation = dot(rnormal, cp2[i].v - rv1);
    if (separation <= radius) {
      MPoint* cp = m->pts + pc;
      if (primary.type == 1) {
        cp->lp = mulT(xf, cp2[i].v);
        cp->id = fkey(cp2[i].id);
      } else {
        cp->lp = cp2[i].v;
        Feature f = cp2[i].id;
        cp->id = fkey(Feature{f.indexB, f.indexA, f.typeB, f.typeA});
      }
      ++pc;
    }
  }
  m->count = pc;
}

// dispatch table of b2Contact (b2_contact.cpp:48-56); returns false when no
// function exists for (typeA, typeB)
bool evaluate(Manifold* m, const b2g_shape& A, const Xf& xfA, const b2g_shape& B, const Xf& xfB) {
  const int ta = A.type, tb = B.type;
  if (ta == B2G_CIRCLE && tb == B2G_CIRCLE) collide_circles(m, A, xfA, B, xfB);
  else if (ta == B2G_POLYGON && tb == B2G_CIRCLE) collide_poly_circle(m, A, xfA, B, xfB);
  else if (ta == B2G_POLYGON && tb == B2G_POLYGON) collide_polygons(m, A, xfA, B, xfB);
  else if (ta == B2G_EDGE && tb == B2G_CIRCLE) collide_edge_circle(m, A, xfA, B, xfB);
  else if (ta == B2G_EDGE && tb == B2G_POLYGON) collide_edge_polygon(m, A, xfA, B, xfB);
  else return false;
  return true;
}

inline bool has_function(int ta, int tb) {
  return (ta == B2G_CIRCLE && tb == B2G_CIRCLE) || (ta == B2G_POLYGON && tb == B2G_CIRCLE) ||
         (ta == B2G_POLYGON && tb == B2G_POLYGON) || (ta == B2G_EDGE && tb == B2G_CIRCLE) ||
         (ta == B2G_EDGE && tb == B2G_POLYGON);
}

struct WorldManifold {
  V2 normal;
  V2 points[2];
  float seps[2];
};

// b2WorldManifold::Initialize (b2_collision.cpp:26-90)
void world_manifold(WorldManifold* wm, const Manifold& m, const Xf& xfA, float rA, const Xf& xfB,
                    float rB) {
  if (m.count == 0) return;
  if (m.type == 0) {
    wm->normal = v2(1.0f, 0.0f);
    V2 pA = mul(xfA, m.lp);
    V2 pB = mul(xfB, m.pts[0].lp);
    if (dist_sq(pA, pB) > kEps * kEps) wm->normal = normalize(pB - pA);
    V2 cA = pA + rA * wm->normal;
    V2 cB = pB - rB * wm->normal;
    wm->points[0] = 0.5f * (cA + cB);
    wm->seps[0] = dot(cB - cA, wm->normal);
  } else if (m.type == 1) {
    wm->normal = mul(xfA.q, m.ln);
    V2 planePoint = mul(xfA, m.lp);
    for (int i = 0; i < m.count; ++i) {
      V2 clipPoint = mul(xfB, m.pts[i].lp);
      V2 cA = clipPoint + (rA - dot(clipPoint - planePoint, wm->normal)) * wm->normal;
      V2 cB = clipPoint - rB * wm->normal;
      wm->points[i] = 0.5f * (cA + cB);
      wm->seps[i] = dot(cB - cA, wm->normal);
    }
  } else {
    wm->normal = mul(xfB.q, m.ln);
    V2 planePoint = mul(xfB, m.lp);
    for (int i = 0; i < m.count; ++i) {
      V2 clipPoint = mul(xfA, m.pts[i].lp);
      V2 cB = clipPoint + (rB - dot(clipPoint - planePoint, wm->normal)) * wm->normal;
      V2 cA = clipPoint - rA * wm->normal;
      wm->points[i] = 0.5f * (cA + cB);
      wm->seps[i] = dot(cA - cB, wm->normal);
    }
    wm->normal = -wm->normal;
  }
}

void manifold_to_record(const Manifold& m, b2g_contact* o) {
  o->type = m.type;
  o->point_count = m.count;
  o->local_normal[0] = m.ln.x; o->local_normal[1] = m.ln.y;
  o->local_point[0] = m.lp.x; o->local_point[1] = m.lp.y;
  for (int k = 0; k < 2; ++k) {
    if (k < m.count) {
      o->point_local[2 * k] = m.pts[k].lp.x;
      o->point_local[2 * k + 1] = m.pts[k].lp.y;
      o->normal_impulse[k] = m.pts[k].ni;
      o->tangent_impulse[k] = m.pts[k].ti;
      o->id[k] = m.pts[k].id;
    } else {
      o->point_local[2 * k] = o->point_local[2 * k + 1] = 0.0f;
      o->normal_impulse[k] = o->tangent_impulse[k] = 0.0f;
      o->id[k] = 0;
    }
  }
}

void record_to_manifold(const b2g_contact& o, Manifold* m) {
  std::memset(m, 0, sizeof(*m));
  m->type = o.type;
  m->count = o.point_count;
  m->ln = v2(o.local_normal[0], o.local_normal[1]);
  m->lp = v2(o.local_point[0], o.local_point[1]);
  for (int k = 0; k < 2; ++k) {
    m->pts[k].lp = v2(o.point_local[2 * k], o.point_local[2 * k + 1]);
    m->pts[k].ni = o.normal_impulse[k];
    m->pts[k].ti = o.tangent_impulse[k];
    m->pts[k].id = o.id[k];
  }
}

// ---------------------------------------------------------------------------
// World
// ---------------------------------------------------------------------------
struct Body {
  int type;
  Xf xf;
  V2 c, lc;
  float a;
  V2 v;
  float w;
  float mass, invMass, invI;
  float linDamp, angDamp, gravScale;
  V2 force;
  float torque;
  // sleep (b2_body.h:541-543; flags e_awakeFlag / e_autoSleepFlag): zero means
  // awake and allowed to sleep, the b2BodyDef defaults
  bool asleep;
  bool no_autosleep;
  float sleep_time;
};

struct Contact {
  int a, b;  // shape indices in contact order
  uint64_t key;
  bool touching;
  Manifold m;
  float friction, restitution, threshold, tangentSpeed;
};

struct Joint {
  b2g_revolute r;
};

struct VelPoint {
  V2 rA, rB;
  float ni, nm, ti, tm, bias;
};
struct VelC {
  VelPoint p[2];
  V2 normal;
  Mat22 normalMass, K;
  int ia, ib;
  float mA, mB, iA, iB;
  float friction, restitution, threshold, tangentSpeed;
  int count;
  int contact;
};
struct PosC {
  V2 lps[2], ln, lp;
  int ia, ib;
  float mA, mB;
  V2 lcA, lcB;
  float iA, iB;
  int type;
  float rA, rB;
  int count;
};
struct JointSolver {  // b2RevoluteJoint solver temporaries
  int ia, ib;
  V2 lcA, lcB;
  float mA, mB, iA, iB;
  V2 rA, rB;
  Mat22 K;
  float axialMass;
  float angle;
};

}  // namespace

struct orc_world {
  V2 gravity;
  bool warm;
  bool allow_sleep = false;  // b2World::m_allowSleep
  float inv_dt0 = 0.0f;
  bool new_contacts = true;
  std::vector<Body> bodies;
  std::vector<b2g_shape> shapes;
  std::vector<Joint> joints;
  std::vector<Contact> contacts;  // sorted by key
  std::unordered_map<uint64_t, int> colour_of;
  // order in which the next Collide walks the contacts (keys); empty = the
  // parallel walk described in collide()
  std::vector<uint64_t> collide_order;
  std::vector<int32_t> joint_colour;
};

namespace {

uint64_t pair_key(int i, int j) {
  uint32_t lo = (uint32_t)(i < j ? i : j), hi = (uint32_t)(i < j ? j : i);
  return ((uint64_t)lo << 32) | hi;
}

// b2ContactFilter::ShouldCollide (src/dynamics/b2_world_callbacks.cpp:28-40)
bool filter_ok(const b2g_shape& A, const b2g_shape& B) {
  if (A.group == B.group && A.group != 0) return A.group > 0;
  return (A.mask & B.category) != 0 && (A.category & B.mask) != 0;
}

// b2Body::ShouldCollide (src/dynamics/b2_body.cpp:459-480)
bool bodies_collide(const orc_world* w, int ba, int bb) {
  if (w->bodies[ba].type != B2G_DYNAMIC && w->bodies[bb].type != B2G_DYNAMIC) return false;
  for (const Joint& j : w->joints) {
    if ((j.r.flags & 1) != 0) continue;
    if ((j.r.body_a == ba && j.r.body_b == bb) || (j.r.body_a == bb && j.r.body_b == ba)) return false;
  }
  return true;
}

void sync_transform(Body& b) {  // b2_body.h:957-961
  b.xf.q = rot(b.a);
  b.xf.p = b.c - mul(b.xf.q, b.lc);
}

// b2Body::SetAwake (include/box2d/b2_body.h:719-740)
void set_awake(Body& b, bool flag) {
  if (b.type == B2G_STATIC) return;
  b.sleep_time = 0.0f;
  b.asleep = !flag;
  if (!flag) {
    b.v = v2(0.0f, 0.0f);
    b.w = 0.0f;
    b.force = v2(0.0f, 0.0f);
    b.torque = 0.0f;
  }
}

// Pair set of b2BroadPhase::UpdateAndQuery + QueryCallback: sort-and-sweep on x
// with the inclusive AABB test of b2TestOverlap (include/box2d/b2_collision.h:270-276).
void find_pairs_impl(int n, const float* ab, const int32_t* sbody, const int32_t* stype,
                     const int32_t* filt, const int32_t* btype,
                     const std::vector<std::pair<int, int>>* nocollide,
                     std::vector<uint64_t>* out) {
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(), [&](int i, int j) {
    if (ab[4 * i] != ab[4 * j]) return ab[4 * i] < ab[4 * j];
    return i < j;
  });
  for (int ii = 0; ii < n; ++ii) {
    const int i = order[ii];
    const float* A = ab + 4 * i;
    for (int jj = ii + 1; jj < n; ++jj) {
      const int j = order[jj];
      const float* B = ab + 4 * j;
      if (B[0] > A[2]) break;
      bool overlap = (A[2] >= B[0]) & (A[0] <= B[2]) & (A[3] >= B[1]) & (A[1] <= B[3]);
      if (!overlap) continue;
      const int ba = sbody[i], bb = sbody[j];
      if (ba == bb) continue;
      if (btype[ba] != B2G_DYNAMIC && btype[bb] != B2G_DYNAMIC) continue;
      if (!has_function(stype[i], stype[j]) && !has_function(stype[j], stype[i])) continue;
      if (filt) {
        const int32_t* fa = filt + 3 * i;
        const int32_t* fb = filt + 3 * j;
        bool ok;
        if (fa[2] == fb[2] && fa[2] != 0) ok = fa[2] > 0;
        else ok = (fa[1] & fb[0]) != 0 && (fa[0] & fb[1]) != 0;
        if (!ok) continue;
      }
      if (nocollide) {
        bool blocked = false;
        for (auto& p : *nocollide) {
          if ((p.first == ba && p.second == bb) || (p.first == bb && p.second == ba)) { blocked = true; break; }
        }
        if (blocked) continue;
      }
      out->push_back(pair_key(i, j));
    }
  }
  std::sort(out->begin(), out->end());
}

// Broadphase at the end of b2World::Solve (b2_world.cpp:663-669): refresh AABBs
// from the current transforms, find the pair set, keep existing contacts
// (persist) and create new ones (b2ContactManager::QueryCallback,
// b2_contact_manager.cpp:136-188; b2Contact ctor b2_contact.cpp:96-124).
void broadphase(orc_world* w) {
  const int n = (int)w->shapes.size();
  std::vector<float> ab(4 * (size_t)n);
  std::vector<int32_t> sbody(n), stype(n), filt(3 * (size_t)n), btype(w->bodies.size());
  for (int i = 0; i < n; ++i) {
    const b2g_shape& s = w->shapes[i];
    shape_aabb(s, w->bodies[s.body].xf, &ab[4 * (size_t)i]);
    sbody[i] = s.body;
    stype[i] = s.type;
    filt[3 * i] = s.category; filt[3 * i + 1] = s.mask; filt[3 * i + 2] = s.group;
  }
  for (size_t b = 0; b < w->bodies.size(); ++b) btype[b] = w->bodies[b].type;
  std::vector<std::pair<int, int>> nocollide;
  for (const Joint& j : w->joints)
    if ((j.r.flags & 1) == 0) nocollide.push_back({j.r.body_a, j.r.body_b});
  std::vector<uint64_t> keys;
  find_pairs_impl(n, ab.data(), sbody.data(), stype.data(), filt.data(), btype.data(),
                  nocollide.empty() ? nullptr : &nocollide, &keys);
  std::vector<Contact> next;
  next.reserve(keys.size());
  size_t oi = 0;
  // a contact that leaves the list with points wakes its bodies
  // (b2Contact::Destroy, src/dynamics/b2_contact.cpp:79-91)
  auto destroyed = [&](const Contact& c) {
    if (c.m.count > 0 && !w->shapes[c.a].is_sensor && !w->shapes[c.b].is_sensor) {
      set_awake(w->bodies[w->shapes[c.a].body], true);
      set_awake(w->bodies[w->shapes[c.b].body], true);
    }
  };
  for (uint64_t k : keys) {
    while (oi < w->contacts.size() && w->contacts[oi].key < k) destroyed(w->contacts[oi++]);
    if (oi < w->contacts.size() && w->contacts[oi].key == k) {
      next.push_back(w->contacts[oi++]);
      continue;
    }
    const int i = (int)(k >> 32), j = (int)(k & 0xffffffffu);
    Contact c;
    std::memset(&c, 0, sizeof(c));
    const b2g_shape &si = w->shapes[i], &sj = w->shapes[j];
    if (has_function(si.type, sj.type)) { c.a = i; c.b = j; } else { c.a = j; c.b = i; }
    c.key = k;
    c.touching = false;
    c.m.count = 0;
    const b2g_shape &A = w->shapes[c.a], &B = w->shapes[c.b];
    c.friction = std::sqrt(A.friction * B.friction);                        // b2_contact.h:42-45
    c.restitution = A.restitution > B.restitution ? A.restitution : B.restitution;
    c.threshold = A.restitution_threshold < B.restitution_threshold ? A.restitution_threshold
                                                                      : B.restitution_threshold;
    c.tangentSpeed = 0.0f;
    next.push_back(c);
  }
  for (; oi < w->contacts.size(); ++oi) destroyed(w->contacts[oi]);
  w->contacts.swap(next);
}

// b2ContactManager::Collide + b2Contact::Update (b2_contact_manager.cpp:66-116,
// b2_contact.cpp:126-210): new manifold, impulse carry-over by feature key.
void collide(orc_world* w) {
  // A contact is updated when one of its bodies is awake and not static
  // (b2_contact_manager.cpp:100-106). The reference walks its list in order,
  // so a body woken by an earlier contact's touching change already counts
  // for later contacts of the same walk. With a walk order given
  // (collide_order, the reference's list order) that is what happens here;
  // without one the flags are those at the start of Collide and a wake-up
  // counts from the next step on (as on the device, where the contacts are
  // updated in parallel).
  const bool walk = !w->collide_order.empty();
  std::vector<char> active0(w->bodies.size());
  for (size_t i = 0; i < w->bodies.size(); ++i)
    active0[i] = !w->bodies[i].asleep && w->bodies[i].type != B2G_STATIC;
  std::vector<int> visit;
  visit.reserve(w->contacts.size());
  if (walk) {
    std::unordered_map<uint64_t, int> at;
    for (int i = 0; i < (int)w->contacts.size(); ++i) at[w->contacts[i].key] = i;
    std::vector<char> seen(w->contacts.size(), 0);
    for (uint64_t k : w->collide_order) {
      auto it = at.find(k);
      if (it != at.end() && !seen[it->second]) { visit.push_back(it->second); seen[it->second] = 1; }
    }
    for (int i = 0; i < (int)w->contacts.size(); ++i)
      if (!seen[i]) visit.push_back(i);
    w->collide_order.clear();
  } else {
    for (int i = 0; i < (int)w->contacts.size(); ++i) visit.push_back(i);
  }
  for (int ci : visit) {
    Contact& c = w->contacts[ci];
    const b2g_shape &A = w->shapes[c.a], &B = w->shapes[c.b];
    const Body &ba = w->bodies[A.body], &bb = w->bodies[B.body];
    const bool activeA = walk ? (!ba.asleep && ba.type != B2G_STATIC) : active0[A.body];
    const bool activeB = walk ? (!bb.asleep && bb.type != B2G_STATIC) : active0[B.body];
    if (!activeA && !activeB) continue;
    Manifold old = c.m;
    evaluate(&c.m, A, ba.xf, B, bb.xf);
    for (int i = 0; i < c.m.count; ++i) {
      MPoint* mp2 = c.m.pts + i;
      mp2->ni = 0.0f;
      mp2->ti = 0.0f;
      for (int j = 0; j < old.count; ++j) {
        if (old.pts[j].id == mp2->id) {
          mp2->ni = old.pts[j].ni;
          mp2->ti = old.pts[j].ti;
          break;
        }
      }
    }
    const bool was = c.touching;
    c.touching = c.m.count > 0;
    if (c.touching != was) {  // b2_contact.cpp:183-188
      set_awake(w->bodies[A.body], true);
      set_awake(w->bodies[B.body], true);
    }
  }
}

struct UF {
  std::vector<int> p;
  explicit UF(int n) : p(n) { std::iota(p.begin(), p.end(), 0); }
  int find(int x) {
    while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
    return x;
  }
  void unite(int a, int b) {
    a = find(a); b = find(b);
    if (a == b) return;
    if (a < b) p[b] = a; else p[a] = b;
  }
};

void solve(orc_world* w, float dt, int vi, int pi) {
  const float h = dt;
  const float inv_dt = dt > 0.0f ? 1.0f / dt : 0.0f;
  const float dtRatio = w->inv_dt0 * dt;
  const int nb = (int)w->bodies.size();
  std::vector<V2> pc(nb), vv(nb);
  std::vector<float> pa(nb), vw(nb);
  // Constraint candidates (touching, enabled, non-sensor contacts; joints
  // with a non-static body)
  std::vector<int> cl;
  for (int i = 0; i < (int)w->contacts.size(); ++i) {
    const Contact& c = w->contacts[i];
    if (!c.touching) continue;
    if (w->shapes[c.a].is_sensor || w->shapes[c.b].is_sensor) continue;
    cl.push_back(i);
  }
  std::vector<int> jl;
  for (int i = 0; i < (int)w->joints.size(); ++i) {
    const Joint& j = w->joints[i];
    if (w->bodies[j.r.body_a].type == B2G_STATIC && w->bodies[j.r.body_b].type == B2G_STATIC) continue;
    jl.push_back(i);
  }
  // Islands: connected components through non-static bodies (b2_world.cpp:526-647)
  UF uf(nb);
  auto link = [&](int a, int b) {
    if (w->bodies[a].type != B2G_STATIC && w->bodies[b].type != B2G_STATIC) uf.unite(a, b);
  };
  for (int ci : cl) link(w->shapes[w->contacts[ci].a].body, w->shapes[w->contacts[ci].b].body);
  for (int ji : jl) link(w->joints[ji].r.body_a, w->joints[ji].r.body_b);
  auto island_of = [&](int ba, int bb) {
    return w->bodies[ba].type != B2G_STATIC ? uf.find(ba) : uf.find(bb);
  };
  // Only islands with an awake body are built and solved (the seeds,
  // b2_world.cpp:526-552); the depth-first search wakes every body it
  // reaches without resetting its sleep timer (:573-574).
  std::vector<char> isl_awake(nb, 0), active(nb, 0);
  for (int i = 0; i < nb; ++i)
    if (w->bodies[i].type != B2G_STATIC && !w->bodies[i].asleep) isl_awake[uf.find(i)] = 1;
  for (int i = 0; i < nb; ++i) {
    Body& b = w->bodies[i];
    if (b.type == B2G_STATIC || !isl_awake[uf.find(i)]) continue;
    active[i] = 1;
    b.asleep = false;
  }
  {
    std::vector<int> keep;
    for (int ci : cl) {
      const Contact& c = w->contacts[ci];
      if (isl_awake[island_of(w->shapes[c.a].body, w->shapes[c.b].body)]) keep.push_back(ci);
    }
    cl.swap(keep);
    keep.clear();
    for (int ji : jl)
      if (isl_awake[island_of(w->joints[ji].r.body_a, w->joints[ji].r.body_b)]) keep.push_back(ji);
    jl.swap(keep);
  }
  // ordered by colour then index
  auto colour = [&](int ci) -> int {
    auto it = w->colour_of.find(w->contacts[ci].key);
    return it == w->colour_of.end() ? 0 : it->second;
  };
  std::stable_sort(cl.begin(), cl.end(), [&](int x, int y) { return colour(x) < colour(y); });
  if (!w->joint_colour.empty())
    std::stable_sort(jl.begin(), jl.end(), [&](int x, int y) { return w->joint_colour[x] < w->joint_colour[y]; });
  // Integrate velocities and damping (b2_island.cpp:257-298; SolveOrphan :187-206)
  for (int i = 0; i < nb; ++i) {
    Body& b = w->bodies[i];
    V2 v = b.v;
    float wv = b.w;
    if (b.type == B2G_DYNAMIC && active[i]) {
      v = v + h * b.invMass * ((b.gravScale * b.mass) * w->gravity + b.force);
      wv += h * b.invI * b.torque;
      v = V2{v.x / (1.0f + h * b.linDamp), v.y / (1.0f + h * b.linDamp)};
      wv = wv / (1.0f + h * b.angDamp);
    }
    pc[i] = b.c;
    pa[i] = b.a;
    vv[i] = v;
    vw[i] = wv;
  }

  // b2ContactSolver::Initialize + InitializeVelocityConstraints (b2_contact_solver.cpp:65-258)
  const int nc = (int)cl.size();
  std::vector<VelC> vcs(nc);
  std::vector<PosC> pcs(nc);
  std::vector<int> c_island(nc);
  for (int k = 0; k < nc; ++k) {
    Contact& c = w->contacts[cl[k]];
    const b2g_shape &SA = w->shapes[c.a], &SB = w->shapes[c.b];
    const Body &bA = w->bodies[SA.body], &bB = w->bodies[SB.body];
    VelC& vc = vcs[k];
    PosC& p = pcs[k];
    c_island[k] = island_of(SA.body, SB.body);
    vc.contact = cl[k];
    vc.friction = c.friction; vc.restitution = c.restitution; vc.threshold = c.threshold;
    vc.tangentSpeed = c.tangentSpeed;
    vc.ia = SA.body; vc.ib = SB.body;
    vc.mA = bA.invMass; vc.mB = bB.invMass; vc.iA = bA.invI; vc.iB = bB.invI;
    vc.count = c.m.count;
    p.ia = SA.body; p.ib = SB.body;
    p.mA = bA.invMass; p.mB = bB.invMass; p.lcA = bA.lc; p.lcB = bB.lc;
    p.iA = bA.invI; p.iB = bB.invI;
    p.ln = c.m.ln; p.lp = c.m.lp; p.count = c.m.count;
    p.rA = SA.radius; p.rB = SB.radius; p.type = c.m.type;
    for (int j = 0; j < c.m.count; ++j) {
      if (w->warm) {
        vc.p[j].ni = dtRatio * c.m.pts[j].ni;
        vc.p[j].ti = dtRatio * c.m.pts[j].ti;
      } else {
        vc.p[j].ni = 0.0f;
        vc.p[j].ti = 0.0f;
      }
      p.lps[j] = c.m.pts[j].lp;
    }
    // position-dependent part (:155-258)
    V2 cA = pc[vc.ia], cB = pc[vc.ib];
    float aA = pa[vc.ia], aB = pa[vc.ib];
    V2 vA = vv[vc.ia], vB = vv[vc.ib];
    float wA = vw[vc.ia], wB = vw[vc.ib];
    Xf xfA, xfB;
    xfA.q = rot(aA);
    xfB.q = rot(aB);
    xfA.p = cA - mul(xfA.q, p.lcA);
    xfB.p = cB - mul(xfB.q, p.lcB);
    WorldManifold wm;
    world_manifold(&wm, c.m, xfA, p.rA, xfB, p.rB);
    vc.normal = wm.normal;
    V2 tangent = cross(vc.normal, 1.0f);
    for (int j = 0; j < vc.count; ++j) {
      VelPoint& vcp = vc.p[j];
      vcp.rA = wm.points[j] - cA;
      vcp.rB = wm.points[j] - cB;
      float rnA = cross(vcp.rA, vc.normal);
      float rnB = cross(vcp.rB, vc.normal);
      float kNormal = vc.mA + vc.mB + vc.iA * rnA * rnA + vc.iB * rnB * rnB;
      vcp.nm = kNormal > 0.0f ? 1.0f / kNormal : 0.0f;
      float rtA = cross(vcp.rA, tangent);
      float rtB = cross(vcp.rB, tangent);
      float kTangent = vc.mA + vc.mB + vc.iA * rtA * rtA + vc.iB * rtB * rtB;
      vcp.tm = kTangent > 0.0f ? 1.0f / kTangent : 0.0f;
      vcp.bias = 0.0f;
      float vRel = dot(vc.normal, vB + cross(wB, vcp.rB) - (vA + cross(wA, vcp.rA)));
      if (vRel < -vc.threshold) vcp.bias = -vc.restitution * vRel;
    }
    if (vc.count == 2) {
      VelPoint &v1 = vc.p[0], &v2p = vc.p[1];
      float rn1A = cross(v1.rA, vc.normal), rn1B = cross(v1.rB, vc.normal);
      float rn2A = cross(v2p.rA, vc.normal), rn2B = cross(v2p.rB, vc.normal);
      float k11 = vc.mA + vc.mB + vc.iA * rn1A * rn1A + vc.iB * rn1B * rn1B;
      float k22 = vc.mA + vc.mB + vc.iA * rn2A * rn2A + vc.iB * rn2B * rn2B;
      float k12 = vc.mA + vc.mB + vc.iA * rn1A * rn2A + vc.iB * rn1B * rn2B;
      const float k_maxConditionNumber = 1000.0f;
      if (k11 * k11 < k_maxConditionNumber * (k11 * k22 - k12 * k12)) {
        vc.K.ex = v2(k11, k12);
        vc.K.ey = v2(k12, k22);
        vc.normalMass = vc.K.inverse();
      } else {
        vc.count = 1;
      }
    }
  }
  // WarmStart (:260-299)
  if (w->warm) {
    for (int k = 0; k < nc; ++k) {
      VelC& vc = vcs[k];
      V2 vA = vv[vc.ia], vB = vv[vc.ib];
      float wA = vw[vc.ia], wB = vw[vc.ib];
      V2 normal = vc.normal, tangent = cross(normal, 1.0f);
      for (int j = 0; j < vc.count; ++j) {
        V2 P = vc.p[j].ni * normal + vc.p[j].ti * tangent;
        wA -= vc.iA * cross(vc.p[j].rA, P);
        vA = vA - vc.mA * P;
        wB += vc.iB * cross(vc.p[j].rB, P);
        vB = vB + vc.mB * P;
      }
      vv[vc.ia] = vA; vw[vc.ia] = wA;
      vv[vc.ib] = vB; vw[vc.ib] = wB;
    }
  }
  // b2RevoluteJoint::InitVelocityConstraints (b2_revolute_joint.cpp:73-165)
  const int nj = (int)jl.size();
  std::vector<JointSolver> js(nj);
  std::vector<int> j_island(nj);
  for (int k = 0; k < nj; ++k) {
    b2g_revolute& r = w->joints[jl[k]].r;
    JointSolver& s = js[k];
    const Body &bA = w->bodies[r.body_a], &bB = w->bodies[r.body_b];
    j_island[k] = island_of(r.body_a, r.body_b);
    s.ia = r.body_a; s.ib = r.body_b;
    s.lcA = bA.lc; s.lcB = bB.lc;
    s.mA = bA.invMass; s.mB = bB.invMass; s.iA = bA.invI; s.iB = bB.invI;
    float aA = pa[s.ia], aB = pa[s.ib];
    V2 vA = vv[s.ia], vB = vv[s.ib];
    float wA = vw[s.ia], wB = vw[s.ib];
    Rot qA = rot(aA), qB = rot(aB);
    s.rA = mul(qA, v2(r.local_anchor_a[0], r.local_anchor_a[1]) - s.lcA);
    s.rB = mul(qB, v2(r.local_anchor_b[0], r.local_anchor_b[1]) - s.lcB);
    float mA = s.mA, mB = s.mB, iA = s.iA, iB = s.iB;
    s.K.ex.x = mA + mB + s.rA.y * s.rA.y * iA + s.rB.y * s.rB.y * iB;
    s.K.ey.x = -s.rA.y * s.rA.x * iA - s.rB.y * s.rB.x * iB;
    s.K.ex.y = s.K.ey.x;
    s.K.ey.y = mA + mB + s.rA.x * s.rA.x * iA + s.rB.x * s.rB.x * iB;
    s.axialMass = iA + iB;
    bool fixedRotation;
    if (s.axialMass > 0.0f) { s.axialMass = 1.0f / s.axialMass; fixedRotation = false; }
    else fixedRotation = true;
    s.angle = aB - aA - r.reference_angle;
    if ((r.flags & 2) == 0 || fixedRotation) { r.lower_impulse = 0.0f; r.upper_impulse = 0.0f; }
    if ((r.flags & 4) == 0 || fixedRotation) r.motor_impulse = 0.0f;
    if (w->warm) {
      r.impulse[0] *= dtRatio; r.impulse[1] *= dtRatio;
      r.motor_impulse *= dtRatio; r.lower_impulse *= dtRatio; r.upper_impulse *= dtRatio;
      float axialImpulse = r.motor_impulse + r.lower_impulse - r.upper_impulse;
      V2 P = v2(r.impulse[0], r.impulse[1]);
      vA = vA - mA * P;
      wA -= iA * (cross(s.rA, P) + axialImpulse);
      vB = vB + mB * P;
      wB += iB * (cross(s.rB, P) + axialImpulse);
    } else {
      r.impulse[0] = r.impulse[1] = 0.0f;
      r.motor_impulse = r.lower_impulse = r.upper_impulse = 0.0f;
    }
    vv[s.ia] = vA; vw[s.ia] = wA;
    vv[s.ib] = vB; vw[s.ib] = wB;
  }
  // Velocity iterations (b2_island.cpp:327-338)
  for (int it = 0; it < vi; ++it) {
    for (int k = 0; k < nj; ++k) {  // b2RevoluteJoint::SolveVelocityConstraints (:167-243)
      b2g_revolute& r = w->joints[jl[k]].r;
      JointSolver& s = js[k];
      V2 vA = vv[s.ia], vB = vv[s.ib];
      float wA = vw[s.ia], wB = vw[s.ib];
      float mA = s.mA, mB = s.mB, iA = s.iA, iB = s.iB;
      bool fixedRotation = (iA + iB == 0.0f);
      if ((r.flags & 4) && !fixedRotation) {
        float Cdot = wB - wA - r.motor_speed;
        float impulse = -s.axialMass * Cdot;
        float oldImpulse = r.motor_impulse;
        float maxImpulse = h * r.max_motor_torque;
        r.motor_impulse = clampf(r.motor_impulse + impulse, -maxImpulse, maxImpulse);
        impulse = r.motor_impulse - oldImpulse;
        wA -= iA * impulse;
        wB += iB * impulse;
      }
      if ((r.flags & 2) && !fixedRotation) {
        {
          float C = s.angle - r.lower_angle;
          float Cdot = wB - wA;
          float impulse = -s.axialMass * (Cdot + fmax_(C, 0.0f) * inv_dt);
          float oldImpulse = r.lower_impulse;
          r.lower_impulse = fmax_(r.lower_impulse + impulse, 0.0f);
          impulse = r.lower_impulse - oldImpulse;
          wA -= iA * impulse;
          wB += iB * impulse;
        }
        {
          float C = r.upper_angle - s.angle;
          float Cdot = wA - wB;
          float impulse = -s.axialMass * (Cdot + fmax_(C, 0.0f) * inv_dt);
          float oldImpulse = r.upper_impulse;
          r.upper_impulse = fmax_(r.upper_impulse + impulse, 0.0f);
          impulse = r.upper_impulse - oldImpulse;
          wA += iA * impulse;
          wB -= iB * impulse;
        }
      }
      {
        V2 Cdot = vB + cross(wB, s.rB) - vA - cross(wA, s.rA);
        V2 impulse = s.K.solve(-Cdot);
        r.impulse[0] += impulse.x;
        r.impulse[1] += impulse.y;
        vA = vA - mA * impulse;
        wA -= iA * cross(s.rA, impulse);
        vB = vB + mB * impulse;
        wB += iB * cross(s.rB, impulse);
      }
      vv[s.ia] = vA; vw[s.ia] = wA;
      vv[s.ib] = vB; vw[s.ib] = wB;
    }
    for (int k = 0; k < nc; ++k) {  // b2ContactSolver::SolveVelocityConstraints (:301-639)
      VelC& vc = vcs[k];
      float mA = vc.mA, iA = vc.iA, mB = vc.mB, iB = vc.iB;
      V2 vA = vv[vc.ia], vB = vv[vc.ib];
      float wA = vw[vc.ia], wB = vw[vc.ib];
      V2 normal = vc.normal;
      V2 tangent = cross(normal, 1.0f);
      for (int j = 0; j < vc.count; ++j) {
        VelPoint& vcp = vc.p[j];
        V2 dv = vB + cross(wB, vcp.rB) - (vA + cross(wA, vcp.rA));
        float vt = dot(dv, tangent) - vc.tangentSpeed;
        float lambda = vcp.tm * (-vt);
        float maxFriction = vc.friction * vcp.ni;
        float newImpulse = clampf(vcp.ti + lambda, -maxFriction, maxFriction);
        lambda = newImpulse - vcp.ti;
        vcp.ti = newImpulse;
        V2 P = lambda * tangent;
        vA = vA - mA * P;
        wA -= iA * cross(vcp.rA, P);
        vB = vB + mB * P;
        wB += iB * cross(vcp.rB, P);
      }
      if (vc.count == 1) {
        VelPoint& vcp = vc.p[0];
        V2 dv = vB + cross(wB, vcp.rB) - (vA + cross(wA, vcp.rA));
        float vn = dot(dv, normal);
        float lambda = -vcp.nm * (vn - vcp.bias);
        float newImpulse = fmax_(vcp.ni + lambda, 0.0f);
        lambda = newImpulse - vcp.ni;
        vcp.ni = newImpulse;
        V2 P = lambda * normal;
        vA = vA - mA * P;
        wA -= iA * cross(vcp.rA, P);
        vB = vB + mB * P;
        wB += iB * cross(vcp.rB, P);
      } else {
        VelPoint &cp1 = vc.p[0], &cp2 = vc.p[1];
        V2 a = v2(cp1.ni, cp2.ni);
        V2 dv1 = vB + cross(wB, cp1.rB) - (vA + cross(wA, cp1.rA));
        V2 dv2 = vB + cross(wB, cp2.rB) - (vA + cross(wA, cp2.rA));
        float vn1 = dot(dv1, normal);
        float vn2 = dot(dv2, normal);
        V2 b = v2(vn1 - cp1.bias, vn2 - cp2.bias);
        b = b - mul(vc.K, a);
        V2 x;
        bool solved = false;
        for (;;) {
          x = -mul(vc.normalMass, b);  // case 1
          if (x.x >= 0.0f && x.y >= 0.0f) { solved = true; break; }
          x.x = -cp1.nm * b.x;  // case 2
          x.y = 0.0f;
          vn2 = vc.K.ex.y * x.x + b.y;
          if (x.x >= 0.0f && vn2 >= 0.0f) { solved = true; break; }
          x.x = 0.0f;  // case 3
          x.y = -cp2.nm * b.y;
          vn1 = vc.K.ey.x * x.y + b.x;
          if (x.y >= 0.0f && vn1 >= 0.0f) { solved = true; break; }
          x.x = 0.0f;  // case 4
          x.y = 0.0f;
          vn1 = b.x;
          vn2 = b.y;
          if (vn1 >= 0.0f && vn2 >= 0.0f) { solved = true; break; }
          break;
        }
        if (solved) {
          V2 d = x - a;
          V2 P1 = d.x * normal;
          V2 P2 = d.y * normal;
          vA = vA - mA * (P1 + P2);
          wA -= iA * (cross(cp1.rA, P1) + cross(cp2.rA, P2));
          vB = vB + mB * (P1 + P2);
          wB += iB * (cross(cp1.rB, P1) + cross(cp2.rB, P2));
          cp1.ni = x.x;
          cp2.ni = x.y;
        }
      }
      vv[vc.ia] = vA; vw[vc.ia] = wA;
      vv[vc.ib] = vB; vw[vc.ib] = wB;
    }
  }
  // StoreImpulses (:641-657): only vc.count points
  for (int k = 0; k < nc; ++k) {
    VelC& vc = vcs[k];
    Manifold& m = w->contacts[vc.contact].m;
    for (int j = 0; j < vc.count; ++j) {
      m.pts[j].ni = vc.p[j].ni;
      m.pts[j].ti = vc.p[j].ti;
    }
  }
  // Integrate positions with the translation/rotation clamp (b2_island.cpp:341-385)
  for (int i = 0; i < nb; ++i) {
    if (!active[i]) continue;
    V2 v = vv[i];
    float wv = vw[i];
    V2 translation = h * v;
    if (dot(translation, translation) > kMaxTranslationSq) {
      float ratio = kMaxTranslation / len(translation);
      v = ratio * v;
    }
    float rotation = h * wv;
    if (rotation * rotation > kMaxRotationSq) {
      float ratio = kMaxRotation / fabs_(rotation);
      wv *= ratio;
    }
    pc[i] = pc[i] + h * v;
    pa[i] += h * wv;
    vv[i] = v;
    vw[i] = wv;
  }
  // Position iterations with per-island early exit (b2_island.cpp:387-427)
  std::unordered_map<int, bool> solved;
  std::unordered_map<int, float> minSep;
  std::unordered_map<int, bool> jointsOk;
  for (int k = 0; k < nc; ++k) solved[c_island[k]] = false;
  for (int k = 0; k < nj; ++k) solved[j_island[k]] = false;
  for (int it = 0; it < pi; ++it) {
    minSep.clear();
    jointsOk.clear();
    for (int k = 0; k < nc; ++k) {  // SolvePositionConstraints (:711-787)
      if (solved[c_island[k]]) continue;
      PosC& p = pcs[k];
      V2 cA = pc[p.ia], cB = pc[p.ib];
      float aA = pa[p.ia], aB = pa[p.ib];
      float& ms = minSep.emplace(c_island[k], 0.0f).first->second;
      for (int j = 0; j < p.count; ++j) {
        Xf xfA, xfB;
        xfA.q = rot(aA);
        xfB.q = rot(aB);
        xfA.p = cA - mul(xfA.q, p.lcA);
        xfB.p = cB - mul(xfB.q, p.lcB);
        V2 normal, point;
        float separation;
        if (p.type == 0) {  // b2PositionSolverManifold::Initialize (:659-708)
          V2 pointA = mul(xfA, p.lp);
          V2 pointB = mul(xfB, p.lps[0]);
          normal = normalize(pointB - pointA);
          point = 0.5f * (pointA + pointB);
          separation = dot(pointB - pointA, normal) - p.rA - p.rB;
        } else if (p.type == 1) {
          normal = mul(xfA.q, p.ln);
          V2 planePoint = mul(xfA, p.lp);
          V2 clipPoint = mul(xfB, p.lps[j]);
          separation = dot(clipPoint - planePoint, normal) - p.rA - p.rB;
          point = clipPoint;
        } else {
          normal = mul(xfB.q, p.ln);
          V2 planePoint = mul(xfB, p.lp);
          V2 clipPoint = mul(xfA, p.lps[j]);
          separation = dot(clipPoint - planePoint, normal) - p.rA - p.rB;
          point = clipPoint;
          normal = -normal;
        }
        V2 rA = point - cA, rB = point - cB;
        ms = fmin_(ms, separation);
        float C = clampf(kBaumgarte * (separation + kLinearSlop), -kMaxLinearCorrection, 0.0f);
        float rnA = cross(rA, normal), rnB = cross(rB, normal);
        float K = p.mA + p.mB + p.iA * rnA * rnA + p.iB * rnB * rnB;
        float impulse = K > 0.0f ? -C / K : 0.0f;
        V2 P = impulse * normal;
        cA = cA - p.mA * P;
        aA -= p.iA * cross(rA, P);
        cB = cB + p.mB * P;
        aB += p.iB * cross(rB, P);
      }
      pc[p.ia] = cA; pa[p.ia] = aA;
      pc[p.ib] = cB; pa[p.ib] = aB;
    }
    for (int k = 0; k < nj; ++k) {  // b2RevoluteJoint::SolvePositionConstraints (:245-321)
      if (solved[j_island[k]]) continue;
      b2g_revolute& r = w->joints[jl[k]].r;
      JointSolver& s = js[k];
      V2 cA = pc[s.ia], cB = pc[s.ib];
      float aA = pa[s.ia], aB = pa[s.ib];
      Rot qA = rot(aA), qB = rot(aB);
      float angularError = 0.0f, positionError = 0.0f;
      bool fixedRotation = (s.iA + s.iB == 0.0f);
      if ((r.flags & 2) && !fixedRotation) {
        float angle = aB - aA - r.reference_angle;
        float C = 0.0f;
        if (fabs_(r.upper_angle - r.lower_angle) < 2.0f * kAngularSlop)
          C = clampf(angle - r.lower_angle, -kMaxAngularCorrection, kMaxAngularCorrection);
        else if (angle <= r.lower_angle)
          C = clampf(angle - r.lower_angle + kAngularSlop, -kMaxAngularCorrection, 0.0f);
        else if (angle >= r.upper_angle)
          C = clampf(angle - r.upper_angle - kAngularSlop, 0.0f, kMaxAngularCorrection);
        float limitImpulse = -s.axialMass * C;
        aA -= s.iA * limitImpulse;
        aB += s.iB * limitImpulse;
        angularError = fabs_(C);
      }
      {
        qA = rot(aA);
        qB = rot(aB);
        V2 rA = mul(qA, v2(r.local_anchor_a[0], r.local_anchor_a[1]) - s.lcA);
        V2 rB = mul(qB, v2(r.local_anchor_b[0], r.local_anchor_b[1]) - s.lcB);
        V2 C = cB + rB - cA - rA;
        positionError = len(C);
        float mA = s.mA, mB = s.mB, iA = s.iA, iB = s.iB;
        Mat22 K;
        K.ex.x = mA + mB + iA * rA.y * rA.y + iB * rB.y * rB.y;
        K.ex.y = -iA * rA.x * rA.y - iB * rB.x * rB.y;
        K.ey.x = K.ex.y;
        K.ey.y = mA + mB + iA * rA.x * rA.x + iB * rB.x * rB.x;
        V2 impulse = -K.solve(C);
        cA = cA - mA * impulse;
        aA -= iA * cross(rA, impulse);
        cB = cB + mB * impulse;
        aB += iB * cross(rB, impulse);
      }
      pc[s.ia] = cA; pa[s.ia] = aA;
      pc[s.ib] = cB; pa[s.ib] = aB;
      bool ok = positionError <= kLinearSlop && angularError <= kAngularSlop;
      auto itj = jointsOk.emplace(j_island[k], true).first;
      itj->second = itj->second && ok;
    }
    // per-island early exit: contactsOkay && jointsOkay (b2_island.cpp:403-408)
    for (auto& kv : solved) {
      if (kv.second) continue;
      auto ms = minSep.find(kv.first);
      bool contactsOk = ms == minSep.end() ? true : ms->second >= -3.0f * kLinearSlop;
      auto jo = jointsOk.find(kv.first);
      bool jOk = jo == jointsOk.end() ? true : jo->second;
      if (contactsOk && jOk) kv.second = true;
    }
  }
  // Write back (b2_island.cpp:430-438)
  for (int i = 0; i < nb; ++i) {
    Body& b = w->bodies[i];
    if (!active[i]) continue;
    b.c = pc[i];
    b.a = pa[i];
    b.v = vv[i];
    b.w = vw[i];
    sync_transform(b);
  }
  if (!w->allow_sleep) return;
  // Sleep (b2_island.cpp:444-483; orphans :233-250). An orphan is a body
  // without any contact or joint (b2_world.cpp:545-549); it runs its own
  // timer, and its seed flags set the awake flag again right after
  // SolveOrphan, so an orphan whose timer runs out has its velocities and
  // forces cleared and stays awake.
  const float linTolSqr = kLinearSleepTolerance * kLinearSleepTolerance;
  const float angTolSqr = kAngularSleepTolerance * kAngularSleepTolerance;
  std::vector<int> ncontact(nb, 0);
  std::vector<char> hasjoint(nb, 0);
  for (const Contact& c : w->contacts) {
    ++ncontact[w->shapes[c.a].body];
    ++ncontact[w->shapes[c.b].body];
  }
  for (const Joint& j : w->joints) hasjoint[j.r.body_a] = hasjoint[j.r.body_b] = 1;
  std::vector<float> min_sleep(nb, FLT_MAX);
  for (int i = 0; i < nb; ++i) {
    if (!active[i]) continue;
    Body& b = w->bodies[i];
    const bool slow = !b.no_autosleep && !(b.w * b.w > angTolSqr) && !(dot(b.v, b.v) > linTolSqr);
    if (ncontact[i] == 0 && !hasjoint[i]) {
      if (!slow) {
        b.sleep_time = 0.0f;
      } else {
        b.sleep_time += h;
        if (b.sleep_time >= kTimeToSleep) set_awake(b, false);
      }
      b.asleep = false;
      continue;
    }
    float& m = min_sleep[uf.find(i)];
    if (!slow) {
      b.sleep_time = 0.0f;
      m = 0.0f;
    } else {
      b.sleep_time += h;
      m = fmin_(m, b.sleep_time);
    }
  }
  // positionSolved: the island's early exit happened; an island without
  // constraints exits at its first position iteration
  std::vector<char> has_cons(nb, 0);
  for (int k = 0; k < nc; ++k) has_cons[c_island[k]] = 1;
  for (int k = 0; k < nj; ++k) has_cons[j_island[k]] = 1;
  for (int i = 0; i < nb; ++i) {
    if (!active[i] || (ncontact[i] == 0 && !hasjoint[i])) continue;
    const int r = uf.find(i);
    const bool position_solved = has_cons[r] ? solved[r] : pi > 0;
    if (min_sleep[r] >= kTimeToSleep && position_solved) set_awake(w->bodies[i], false);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C entry points
// ---------------------------------------------------------------------------
extern "C" {

orc_world* orc_world_from_scene(const b2g_scene* s, int warm_starting) {
  orc_world* w = new orc_world();
  w->gravity = v2(s->gravity[0], s->gravity[1]);
  w->warm = warm_starting != 0;
  w->bodies.resize(s->body_count);
  for (int i = 0; i < s->body_count; ++i) {  // b2Body ctor (b2_body.cpp:29-122)
    Body& b = w->bodies[i];
    std::memset(&b, 0, sizeof(b));
    b.type = s->body_type[i];
    b.xf.p = v2(s->body_pos[2 * i], s->body_pos[2 * i + 1]);
    b.xf.q = rot(s->body_angle[i]);
    b.lc = v2(0.0f, 0.0f);
    b.c = b.xf.p;
    b.a = s->body_angle[i];
    if (s->body_vel) {
      b.v = v2(s->body_vel[3 * i], s->body_vel[3 * i + 1]);
      b.w = s->body_vel[3 * i + 2];
    }
    if (s->body_flags) {  // b2BodyDef::allowSleep, ::awake (b2_body.cpp:52-61)
      b.no_autosleep = (s->body_flags[i] & 4) != 0;
      b.asleep = (s->body_flags[i] & 8) != 0 && b.type != B2G_STATIC;
    }
    if (s->body_damping) {
      b.linDamp = s->body_damping[2 * i];
      b.angDamp = s->body_damping[2 * i + 1];
    }
    b.gravScale = s->body_gravity_scale ? s->body_gravity_scale[i] : 1.0f;
  }
  // fixtures + b2Body::ResetMassData after each CreateFixture (b2_body.cpp:230-269, 352-416)
  w->shapes.resize(s->fixture_count);
  for (int i = 0; i < s->fixture_count; ++i) {
    b2g_shape& r = w->shapes[i];
    std::memset(&r, 0, sizeof(r));
    const float* g = s->fix_geom + 17 * i;
    const int kind = s->fix_kind[i];
    r.body = s->fix_body[i];
    if (kind == 0) {
      r.type = B2G_CIRCLE;
      r.radius = g[0];
      r.centroid[0] = g[1];
      r.centroid[1] = g[2];
    } else if (kind == 1) {
      r.type = B2G_EDGE;
      r.radius = kPolygonRadius;
      r.count = g[0] != 0.0f ? 1 : 0;
      if (r.count) {
        for (int k = 0; k < 8; ++k) r.vertices[k] = g[1 + k];
      } else {
        r.vertices[2] = g[3]; r.vertices[3] = g[4]; r.vertices[4] = g[5]; r.vertices[5] = g[6];
      }
    } else {
      Poly p;
      if (kind == 2) {
        V2 pts[8];
        int n = (int)g[0];
        for (int k = 0; k < n; ++k) pts[k] = v2(g[1 + 2 * k], g[2 + 2 * k]);
        set_hull(&p, pts, n);
      } else if (g[5] != 0.0f) {
        set_box(&p, g[0], g[1], v2(g[2], g[3]), g[4]);
      } else {
        set_box(&p, g[0], g[1]);
      }
      r.type = B2G_POLYGON;
      r.radius = kPolygonRadius;
      r.count = p.count;
      r.centroid[0] = p.centroid.x; r.centroid[1] = p.centroid.y;
      for (int k = 0; k < p.count; ++k) {
        r.vertices[2 * k] = p.v[k].x; r.vertices[2 * k + 1] = p.v[k].y;
        r.normals[2 * k] = p.n[k].x; r.normals[2 * k + 1] = p.n[k].y;
      }
    }
    const float* mat = s->fix_material + 4 * i;
    r.friction = mat[1];
    r.restitution = mat[2];
    r.restitution_threshold = mat[3];
    if (s->fix_filter) {
      r.category = s->fix_filter[4 * i] & 0xFFFF;
      r.mask = s->fix_filter[4 * i + 1] & 0xFFFF;
      r.group = (int16_t)s->fix_filter[4 * i + 2];
      r.is_sensor = s->fix_filter[4 * i + 3];
    } else {
      r.category = 1; r.mask = 0xFFFF; r.group = 0; r.is_sensor = 0;
    }
  }
  // b2Body::ResetMassData after every fixture with density > 0; the fixture
  // list is traversed newest first (CreateFixture prepends, b2_body.cpp:244-246).
  std::vector<std::vector<int>> so_far(s->body_count);
  for (int fi = 0; fi < s->fixture_count; ++fi) {
    const int bi = s->fix_body[fi];
    so_far[bi].push_back(fi);
    if (!(s->fix_material[4 * fi] > 0.0f)) continue;
    Body& b = w->bodies[bi];
    float m = 0.0f, In = 0.0f;
    V2 lc = v2(0.0f, 0.0f);
    b.invMass = 0.0f;
    b.invI = 0.0f;
    if (b.type != B2G_DYNAMIC) {  // static / kinematic (b2_body.cpp:362-369)
      b.mass = 0.0f;
      b.lc = v2(0.0f, 0.0f);
      b.c = b.xf.p;
      continue;
    }
    const std::vector<int>& fl = so_far[bi];
    for (int k = (int)fl.size() - 1; k >= 0; --k) {
      const int fj = fl[k];
      const float density = s->fix_material[4 * fj];
      if (density == 0.0f) continue;
      const b2g_shape& r = w->shapes[fj];
      MassData md;
      if (r.type == B2G_CIRCLE) md = circle_mass(r.radius, v2(r.centroid[0], r.centroid[1]), density);
      else if (r.type == B2G_POLYGON) md = poly_mass(poly_of(r), density);
      else { md.mass = 0.0f; md.center = v2(0.0f, 0.0f); md.I = 0.0f; }
      m += md.mass;
      lc = lc + md.mass * md.center;
      In += md.I;
    }
    if (m > 0.0f) {
      b.invMass = 1.0f / m;
      lc = b.invMass * lc;
    }
    const bool fixedRot = s->body_flags && (s->body_flags[bi] & 1);
    if (In > 0.0f && !fixedRot) {
      In -= m * dot(lc, lc);
      b.invI = 1.0f / In;
    }
    b.mass = m;
    V2 oldCenter = b.c;
    b.lc = lc;
    b.c = mul(b.xf, b.lc);
    b.v = b.v + cross(b.w, b.c - oldCenter);
  }
  // revolute joints (b2RevoluteJointDef::Initialize, b2_revolute_joint.cpp:41-48)
  w->joints.resize(s->joint_count);
  for (int i = 0; i < s->joint_count; ++i) {
    b2g_revolute& r = w->joints[i].r;
    std::memset(&r, 0, sizeof(r));
    r.body_a = s->joint_bodies[2 * i];
    r.body_b = s->joint_bodies[2 * i + 1];
    const float* p = s->joint_params + 6 * i;
    V2 anchor = v2(p[0], p[1]);
    V2 la = mulT(w->bodies[r.body_a].xf, anchor);
    V2 lb = mulT(w->bodies[r.body_b].xf, anchor);
    r.local_anchor_a[0] = la.x; r.local_anchor_a[1] = la.y;
    r.local_anchor_b[0] = lb.x; r.local_anchor_b[1] = lb.y;
    r.reference_angle = w->bodies[r.body_b].a - w->bodies[r.body_a].a;
    r.lower_angle = p[2]; r.upper_angle = p[3];
    r.motor_speed = p[4]; r.max_motor_torque = p[5];
    r.flags = s->joint_flags[i];
  }
  w->new_contacts = true;
  return w;
}

void orc_world_destroy(orc_world* w) { delete w; }

void orc_world_get_shapes(orc_world* w, b2g_shape* out) {
  std::memcpy(out, w->shapes.data(), w->shapes.size() * sizeof(b2g_shape));
}

int orc_world_body_count(orc_world* w) { return (int)w->bodies.size(); }

void orc_world_get_bodies(orc_world* w, b2g_body* out) {
  for (size_t i = 0; i < w->bodies.size(); ++i) {
    const Body& b = w->bodies[i];
    b2g_body& o = out[i];
    std::memset(&o, 0, sizeof(o));
    o.type = b.type;
    o.xf[0] = b.xf.p.x; o.xf[1] = b.xf.p.y; o.xf[2] = b.xf.q.s; o.xf[3] = b.xf.q.c;
    o.c[0] = b.c.x; o.c[1] = b.c.y; o.a = b.a;
    o.local_center[0] = b.lc.x; o.local_center[1] = b.lc.y;
    o.v[0] = b.v.x; o.v[1] = b.v.y; o.w = b.w;
    o.mass = b.mass; o.inv_mass = b.invMass; o.inv_i = b.invI;
    o.linear_damping = b.linDamp; o.angular_damping = b.angDamp; o.gravity_scale = b.gravScale;
    o.force[0] = b.force.x; o.force[1] = b.force.y; o.torque = b.torque;
    o.flags = (b.asleep ? B2G_BODY_ASLEEP : 0) | (b.no_autosleep ? B2G_BODY_NO_AUTOSLEEP : 0);
    o.sleep_time = b.sleep_time;
  }
}

void orc_world_set_bodies(orc_world* w, const b2g_body* in) {
  for (size_t i = 0; i < w->bodies.size(); ++i) {
    Body& b = w->bodies[i];
    const b2g_body& o = in[i];
    b.type = o.type;
    b.xf.p = v2(o.xf[0], o.xf[1]); b.xf.q = Rot{o.xf[2], o.xf[3]};
    b.c = v2(o.c[0], o.c[1]); b.a = o.a;
    b.lc = v2(o.local_center[0], o.local_center[1]);
    b.v = v2(o.v[0], o.v[1]); b.w = o.w;
    b.mass = o.mass; b.invMass = o.inv_mass; b.invI = o.inv_i;
    b.linDamp = o.linear_damping; b.angDamp = o.angular_damping; b.gravScale = o.gravity_scale;
    b.force = v2(o.force[0], o.force[1]); b.torque = o.torque;
    b.asleep = (o.flags & B2G_BODY_ASLEEP) != 0 && b.type != B2G_STATIC;
    b.no_autosleep = (o.flags & B2G_BODY_NO_AUTOSLEEP) != 0;
    b.sleep_time = o.sleep_time;
  }
}

void orc_world_get_joints(orc_world* w, b2g_revolute* out) {
  for (size_t i = 0; i < w->joints.size(); ++i) out[i] = w->joints[i].r;
}

void orc_world_set_joints(orc_world* w, const b2g_revolute* in) {
  for (size_t i = 0; i < w->joints.size(); ++i) w->joints[i].r = in[i];
}

int orc_world_contact_count(orc_world* w) { return (int)w->contacts.size(); }

void orc_world_get_contacts(orc_world* w, b2g_contact* out) {
  for (size_t i = 0; i < w->contacts.size(); ++i) {
    const Contact& c = w->contacts[i];
    b2g_contact& o = out[i];
    std::memset(&o, 0, sizeof(o));
    o.shape_a = c.a; o.shape_b = c.b;
    o.touching = c.touching ? 1 : 0;
    manifold_to_record(c.m, &o);
    o.friction = c.friction; o.restitution = c.restitution;
    o.restitution_threshold = c.threshold; o.tangent_speed = c.tangentSpeed;
  }
}

void orc_world_set_contacts(orc_world* w, int n, const b2g_contact* in) {
  w->contacts.resize(n);
  for (int i = 0; i < n; ++i) {
    Contact& c = w->contacts[i];
    const b2g_contact& o = in[i];
    c.a = o.shape_a; c.b = o.shape_b;
    c.key = pair_key(c.a, c.b);
    c.touching = o.touching != 0;
    record_to_manifold(o, &c.m);
    c.friction = o.friction; c.restitution = o.restitution;
    c.threshold = o.restitution_threshold; c.tangentSpeed = o.tangent_speed;
  }
  std::sort(w->contacts.begin(), w->contacts.end(),
            [](const Contact& x, const Contact& y) { return x.key < y.key; });
  w->new_contacts = false;
}

void orc_world_set_inv_dt0(orc_world* w, float inv_dt0) { w->inv_dt0 = inv_dt0; }

void orc_world_set_allow_sleep(orc_world* w, int flag) { w->allow_sleep = flag != 0; }



void orc_world_set_colours(orc_world* w, int n, const uint64_t* keys, const int32_t* colours,
                           const int32_t* joint_colours) {
  w->colour_of.clear();
  for (int i = 0; i < n; ++i) w->colour_of[keys[i]] = colours[i];
  w->joint_colour.clear();
  if (joint_colours) w->joint_colour.assign(joint_colours, joint_colours + w->joints.size());
}

void orc_world_set_collide_order(orc_world* w, int n, const uint64_t* keys) {
  w->collide_order.assign(keys, keys + n);
}

// b2World::Step (b2_world.cpp:1108-1171) with continuous physics off.
void orc_world_step(orc_world* w, float dt, int vi, int pi, unsigned flags) {
  if (w->new_contacts || (flags & B2G_STEP_FIND_NEW_CONTACTS)) {
    broadphase(w);
    w->new_contacts = false;
  }
  collide(w);
  if (dt > 0.0f) {
    solve(w, dt, vi, pi);
    if (!(flags & B2G_STEP_SKIP_BROADPHASE)) broadphase(w);
    w->inv_dt0 = 1.0f / dt;
  }
  for (Body& b : w->bodies) {  // ClearForces (:1173-1180)
    b.force = v2(0.0f, 0.0f);
    b.torque = 0.0f;
  }
}

void orc_collide(const b2g_shape* a, const float* xfa, const b2g_shape* b, const float* xfb,
                 b2g_contact* out) {
  Manifold m;
  std::memset(&m, 0, sizeof(m));
  std::memset(out, 0, sizeof(*out));
  if (!evaluate(&m, *a, xf_from(xfa), *b, xf_from(xfb))) {
    out->point_count = -1;
    return;
  }
  manifold_to_record(m, out);
  out->touching = m.count > 0;
}

void orc_world_manifold(const b2g_contact* c, const float* xfa, float ra, const float* xfb,
                        float rb, float* out) {
  Manifold m;
  record_to_manifold(*c, &m);
  WorldManifold wm;
  std::memset(&wm, 0, sizeof(wm));
  world_manifold(&wm, m, xf_from(xfa), ra, xf_from(xfb), rb);
  out[0] = wm.normal.x; out[1] = wm.normal.y;
  out[2] = wm.points[0].x; out[3] = wm.points[0].y;
  out[4] = wm.points[1].x; out[5] = wm.points[1].y;
  out[6] = wm.seps[0]; out[7] = wm.seps[1];
}

void orc_shape_aabb(const b2g_shape* s, const float* xf, float* out) {
  shape_aabb(*s, xf_from(xf), out);
}

long long orc_find_pairs(int n, const float* aabbs, const int32_t* shape_body,
                         const int32_t* shape_type, const int32_t* filter, int body_count,
                         const int32_t* body_type, long long cap, uint64_t* out_keys) {
  (void)body_count;
  std::vector<uint64_t> keys;
  find_pairs_impl(n, aabbs, shape_body, shape_type, filter, body_type, nullptr, &keys);
  for (long long i = 0; i < (long long)keys.size() && i < cap; ++i) out_keys[i] = keys[i];
  return (long long)keys.size();
}

void orc_block_solve(const float* Kin, const float* bin, float* a_io, int* case_taken) {
  Mat22 K;
  K.ex = v2(Kin[0], Kin[1]);
  K.ey = v2(Kin[1], Kin[2]);
  Mat22 nm = K.inverse();
  V2 a = v2(a_io[0], a_io[1]);
  V2 b = v2(bin[0], bin[1]);
  b = b - mul(K, a);
  float m1 = Kin[0] > 0.0f ? 1.0f / Kin[0] : 0.0f;
  float m2 = Kin[2] > 0.0f ? 1.0f / Kin[2] : 0.0f;
  V2 x = -mul(nm, b);
  *case_taken = 0;
  if (x.x >= 0.0f && x.y >= 0.0f) { *case_taken = 1; }
  else {
    x = v2(-m1 * b.x, 0.0f);
    float vn2 = K.ex.y * x.x + b.y;
    if (x.x >= 0.0f && vn2 >= 0.0f) { *case_taken = 2; }
    else {
      x = v2(0.0f, -m2 * b.y);
      float vn1 = K.ey.x * x.y + b.x;
      if (x.y >= 0.0f && vn1 >= 0.0f) { *case_taken = 3; }
      else {
        x = v2(0.0f, 0.0f);
        if (b.x >= 0.0f && b.y >= 0.0f) *case_taken = 4;
      }
    }
  }
  if (*case_taken) { a_io[0] = x.x; a_io[1] = x.y; }
}

}  // extern "C"
