// b2g_collide.h -- per-lane narrowphase: contact manifolds for the five shape
// pairs the reference dispatches (src/dynamics/b2_contact.cpp:48-56).
//
// One lane evaluates one contact. Geometry is held in registers (at most 8
// vertices/normals per polygon); every array access uses a compile-time index
// (unrolled loops with count guards, select chains for data-dependent indices)
// so nothing spills to scratch memory. Arithmetic mirrors the reference
// operation by operation (see b2g_math.h).
#pragma once

#include "b2g_math.h"

namespace b2g {

enum : int { kCircle = 0, kEdge = 1, kPolygon = 2 };
enum : int { kManCircles = 0, kManFaceA = 1, kManFaceB = 2 };
constexpr int kMaxVerts = 8;

// Shape geometry. Device layout per shape: 9 float4
//   g[0] = (radius, centroid.x, centroid.y, 0)          circle: centroid = m_p
//   g[1..4] = vertices 0..7 (two per float4)            edge: v0, v1, v2, v3
//   g[5..8] = normals 0..7
struct Geo {
  int type;
  int count;  // polygon vertex count; edge: oneSided flag
  float radius;
  vec2 centroid;
  vec2 v[kMaxVerts];
  vec2 n[kMaxVerts];
};

__device__ __forceinline__ void load_geo(Geo& g, int type, int count, const float4* __restrict__ src) {
  g.type = type;
  g.count = count;
  const float4 h = src[0];
  g.radius = h.x;
  g.centroid = mk(h.y, h.z);
  if (type == kCircle) return;
  const int nq = type == kEdge ? 2 : 4;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k < nq) {
      const float4 a = src[1 + k];
      g.v[2 * k] = mk(a.x, a.y);
      g.v[2 * k + 1] = mk(a.z, a.w);
    }
  }
  if (type != kPolygon) return;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float4 a = src[5 + k];
    g.n[2 * k] = mk(a.x, a.y);
    g.n[2 * k + 1] = mk(a.z, a.w);
  }
}

// value of a[i] for a data-dependent i, as a chain of register selects. The
// empty asm after each select is an optimisation barrier: without it the
// compiler folds the chain back into a[i], which forces the arrays into
// scratch memory.
__device__ __forceinline__ vec2 pick(const vec2 (&a)[kMaxVerts], int i) {
  vec2 r = a[0];
#pragma unroll
  for (int k = 1; k < kMaxVerts; ++k) {
    r.x = (k == i) ? a[k].x : r.x;
    r.y = (k == i) ? a[k].y : r.y;
    asm volatile("" : "+v"(r.x), "+v"(r.y));
  }
  return r;
}

// b2ContactFeature packed as b2ContactID::key (bytes indexA, indexB, typeA, typeB)
__device__ __forceinline__ uint32_t feature(uint32_t ia, uint32_t ib, uint32_t ta, uint32_t tb) {
  return (ia & 0xffu) | ((ib & 0xffu) << 8) | (ta << 16) | (tb << 24);
}
__device__ __forceinline__ uint32_t feature_swap(uint32_t f) {
  // swap (indexA, indexB) and (typeA, typeB)
  return ((f & 0xffu) << 8) | ((f >> 8) & 0xffu) | (((f >> 16) & 0xffu) << 24) | (((f >> 24) & 0xffu) << 16);
}
constexpr uint32_t kFeatVertex = 0, kFeatFace = 1;

struct Man {
  vec2 lp0, lp1;  // points[i].localPoint
  uint32_t id0, id1;
  vec2 ln, lp;    // localNormal, localPoint
  int type;
  int count;
};

struct ClipV {
  vec2 v;
  uint32_t id;
};

// b2ClipSegmentToLine (src/collision/b2_collision.cpp:205-237)
__device__ __forceinline__ int clip_segment(ClipV& o0, ClipV& o1, const ClipV& i0, const ClipV& i1,
                                            vec2 normal, float offset, int vertex_index_a) {
  int count = 0;
  const float d0 = dot(normal, i0.v) - offset;
  const float d1 = dot(normal, i1.v) - offset;
  if (d0 <= 0.0f) { o0 = i0; count = 1; }
  if (d1 <= 0.0f) {
    if (count == 0) o0 = i1; else o1 = i1;
    ++count;
  }
  if (d0 * d1 < 0.0f) {
    const float interp = d0 / (d0 - d1);
    ClipV c;
    c.v = i0.v + interp * (i1.v - i0.v);
    c.id = feature((uint32_t)vertex_index_a, (i0.id >> 8) & 0xffu, kFeatVertex, kFeatFace);
    if (count == 0) o0 = c; else o1 = c;
    ++count;
  }
  return count;
}

// b2CollideCircles (src/collision/b2_collide_circle.cpp:27-53)
__device__ __forceinline__ void collide_circles(Man& m, const Geo& a, const xform& xa, const Geo& b,
                                                const xform& xb) {
  m.count = 0;
  const vec2 pA = mul(xa, a.centroid);
  const vec2 pB = mul(xb, b.centroid);
  const vec2 d = pB - pA;
  const float dist_sqr = dot(d, d);
  const float radius = a.radius + b.radius;
  if (dist_sqr > radius * radius) return;
  m.type = kManCircles;
  m.lp = a.centroid;
  m.ln = mk(0.0f, 0.0f);
  m.count = 1;
  m.lp0 = b.centroid;
  m.id0 = 0;
}

// b2CollidePolygonAndCircle (src/collision/b2_collide_circle.cpp:55-158)
__device__ __forceinline__ void collide_polygon_circle(Man& m, const Geo& a, const xform& xa,
                                                       const Geo& b, const xform& xb) {
  m.count = 0;
  const vec2 c = mul(xb, b.centroid);
  const vec2 cl = mul_t(xa, c);
  int normal_index = 0;
  float separation = -FLT_MAX;
  const float radius = a.radius + b.radius;
  bool out = false;
#pragma unroll
  for (int i = 0; i < kMaxVerts; ++i) {
    if (i < a.count && !out) {
      const float s = dot(a.n[i], cl - a.v[i]);
      if (s > radius) out = true;
      else if (s > separation) { separation = s; normal_index = i; }
    }
  }
  if (out) return;
  const int i1 = normal_index;
  const int i2 = i1 + 1 < a.count ? i1 + 1 : 0;
  const vec2 v1 = pick(a.v, i1), v2 = pick(a.v, i2);
  const vec2 n1 = pick(a.n, i1);
  m.lp0 = b.centroid;
  m.id0 = 0;
  if (separation < kEpsilon) {
    m.count = 1;
    m.type = kManFaceA;
    m.ln = n1;
    m.lp = 0.5f * (v1 + v2);
    return;
  }
  const float u1 = dot(cl - v1, v2 - v1);
  const float u2 = dot(cl - v2, v1 - v2);
  if (u1 <= 0.0f) {
    if (dist_sq(cl, v1) > radius * radius) return;
    m.count = 1;
    m.type = kManFaceA;
    m.ln = normalize(cl - v1);
    m.lp = v1;
  } else if (u2 <= 0.0f) {
    if (dist_sq(cl, v2) > radius * radius) return;
    m.count = 1;
    m.type = kManFaceA;
    m.ln = normalize(cl - v2);
    m.lp = v2;
  } else {
    const vec2 face_center = 0.5f * (v1 + v2);
    const float s = dot(cl - face_center, n1);
    if (s > radius) return;
    m.count = 1;
    m.type = kManFaceA;
    m.ln = n1;
    m.lp = face_center;
  }
}

// b2FindMaxSeparation (src/collision/b2_collide_polygon.cpp:27-66)
__device__ __forceinline__ float find_max_separation(int& edge, const Geo& p1, const xform& x1,
                                                     const Geo& p2, const xform& x2) {
  const xform xf = mul_t(x2, x1);
  int best = 0;
  float max_sep = -FLT_MAX;
#pragma unroll
  for (int i = 0; i < kMaxVerts; ++i) {
    if (i < p1.count) {
      const vec2 n = mul(xf.q, p1.n[i]);
      const vec2 v1 = mul(xf, p1.v[i]);
      float si = FLT_MAX;
#pragma unroll
      for (int j = 0; j < kMaxVerts; ++j) {
        if (j < p2.count) {
          const float sij = dot(n, p2.v[j] - v1);
          if (sij < si) si = sij;
        }
      }
      if (si > max_sep) { max_sep = si; best = i; }
    }
  }
  edge = best;
  return max_sep;
}

// Reference face on p1 (edge1), incident edge on p2; the tail of
// b2CollidePolygons (src/collision/b2_collide_polygon.cpp:68-111, 164-243).
// Called with (a, b) or (b, a) so the roles never need a runtime copy.
__device__ __forceinline__ void clip_polygons(Man& m, const Geo& p1, const xform& x1, int edge1, const Geo& p2,
                                              const xform& x2, float total_radius, bool flip) {
  // b2FindIncidentEdge (:68-111)
  const vec2 normal1 = mul_t(x2.q, mul(x1.q, pick(p1.n, edge1)));
  int index = 0;
  float min_dot = FLT_MAX;
#pragma unroll
  for (int i = 0; i < kMaxVerts; ++i) {
    if (i < p2.count) {
      const float d = dot(normal1, p2.n[i]);
      if (d < min_dot) { min_dot = d; index = i; }
    }
  }
  const int j1 = index, j2 = j1 + 1 < p2.count ? j1 + 1 : 0;
  ClipV inc0, inc1;
  inc0.v = mul(x2, pick(p2.v, j1));
  inc0.id = feature((uint32_t)edge1, (uint32_t)j1, kFeatFace, kFeatVertex);
  inc1.v = mul(x2, pick(p2.v, j2));
  inc1.id = feature((uint32_t)edge1, (uint32_t)j2, kFeatFace, kFeatVertex);

  const int iv1 = edge1, iv2 = edge1 + 1 < p1.count ? edge1 + 1 : 0;
  vec2 v11 = pick(p1.v, iv1), v12 = pick(p1.v, iv2);
  const vec2 local_tangent = normalize(v12 - v11);
  const vec2 local_normal = cross(local_tangent, 1.0f);
  const vec2 plane_point = 0.5f * (v11 + v12);
  const vec2 tangent = mul(x1.q, local_tangent);
  const vec2 normal = cross(tangent, 1.0f);
  v11 = mul(x1, v11);
  v12 = mul(x1, v12);
  const float front_offset = dot(normal, v11);
  const float side_offset1 = -dot(tangent, v11) + total_radius;
  const float side_offset2 = dot(tangent, v12) + total_radius;
  ClipV c1a, c1b, c2a, c2b;
  int np = clip_segment(c1a, c1b, inc0, inc1, -tangent, side_offset1, iv1);
  if (np < 2) return;
  np = clip_segment(c2a, c2b, c1a, c1b, tangent, side_offset2, iv2);
  if (np < 2) return;
  m.ln = local_normal;
  m.lp = plane_point;
  int pc = 0;
  const float s0 = dot(normal, c2a.v) - front_offset;
  if (s0 <= total_radius) {
    m.lp0 = mul_t(x2, c2a.v);
    m.id0 = flip ? feature_swap(c2a.id) : c2a.id;
    pc = 1;
  }
  const float s1 = dot(normal, c2b.v) - front_offset;
  if (s1 <= total_radius) {
    const vec2 lpt = mul_t(x2, c2b.v);
    const uint32_t id = flip ? feature_swap(c2b.id) : c2b.id;
    if (pc == 0) { m.lp0 = lpt; m.id0 = id; }
    else { m.lp1 = lpt; m.id1 = id; }
    ++pc;
  }
  m.count = pc;
}

// b2CollidePolygons (src/collision/b2_collide_polygon.cpp:120-163)
__device__ __forceinline__ void collide_polygons(Man& m, const Geo& a, const xform& xa, const Geo& b,
                                                 const xform& xb) {
  m.count = 0;
  const float total_radius = a.radius + b.radius;
  int edge_a = 0;
  const float sep_a = find_max_separation(edge_a, a, xa, b, xb);
  if (sep_a > total_radius) return;
  int edge_b = 0;
  const float sep_b = find_max_separation(edge_b, b, xb, a, xa);
  if (sep_b > total_radius) return;
  const float k_tol = 0.1f * kLinearSlop;
  if (sep_b > sep_a + k_tol) {
    m.type = kManFaceB;
    clip_polygons(m, b, xb, edge_b, a, xa, total_radius, true);
  } else {
    m.type = kManFaceA;
    clip_polygons(m, a, xa, edge_a, b, xb, total_radius, false);
  }
}

// b2CollideEdgeAndCircle (src/collision/b2_collide_edge.cpp:31-165)
__device__ __forceinline__ void collide_edge_circle(Man& m, const Geo& e, const xform& xa, const Geo& c,
                                                    const xform& xb) {
  m.count = 0;
  const vec2 Q = mul_t(xa, mul(xb, c.centroid));
  const vec2 A = e.v[1], B = e.v[2];
  const vec2 ed = B - A;
  vec2 n = mk(ed.y, -ed.x);
  const float offset = dot(n, Q - A);
  const bool one_sided = e.count != 0;
  if (one_sided && offset < 0.0f) return;
  const float u = dot(ed, B - Q);
  const float v = dot(ed, Q - A);
  const float radius = e.radius + c.radius;
  m.lp0 = c.centroid;
  if (v <= 0.0f) {  // region A
    const vec2 d = Q - A;
    if (dot(d, d) > radius * radius) return;
    if (one_sided) {
      const vec2 e1 = A - e.v[0];
      if (dot(e1, A - Q) > 0.0f) return;
    }
    m.count = 1;
    m.type = kManCircles;
    m.ln = mk(0.0f, 0.0f);
    m.lp = A;
    m.id0 = feature(0, 0, kFeatVertex, kFeatVertex);
    return;
  }
  if (u <= 0.0f) {  // region B
    const vec2 d = Q - B;
    if (dot(d, d) > radius * radius) return;
    if (one_sided) {
      const vec2 e2 = e.v[3] - B;
      if (dot(e2, Q - B) > 0.0f) return;
    }
    m.count = 1;
    m.type = kManCircles;
    m.ln = mk(0.0f, 0.0f);
    m.lp = B;
    m.id0 = feature(1, 0, kFeatVertex, kFeatVertex);
    return;
  }
  const float den = dot(ed, ed);  // region AB
  const vec2 P = (1.0f / den) * (u * A + v * B);
  const vec2 d = Q - P;
  if (dot(d, d) > radius * radius) return;
  if (offset < 0.0f) n = mk(-n.x, -n.y);
  n = normalize(n);
  m.count = 1;
  m.type = kManFaceA;
  m.ln = n;
  m.lp = A;
  m.id0 = feature(0, 0, kFeatFace, kFeatVertex);
}

// b2CollideEdgeAndPolygon (src/collision/b2_collide_edge.cpp:169-524)
__device__ __forceinline__ void collide_edge_polygon(Man& m, const Geo& e, const xform& xa,
                                                     const Geo& poly, const xform& xb) {
  m.count = 0;
  const xform xf = mul_t(xa, xb);
  const vec2 centroid_b = mul(xf, poly.centroid);
  const vec2 v1 = e.v[1], v2 = e.v[2];
  const vec2 edge1 = normalize(v2 - v1);
  const vec2 normal1 = mk(edge1.y, -edge1.x);
  const float offset1 = dot(normal1, centroid_b - v1);
  const bool one_sided = e.count != 0;
  if (one_sided && offset1 < 0.0f) return;
  vec2 tv[kMaxVerts], tn[kMaxVerts];
  const int tc = poly.count;
#pragma unroll
  for (int i = 0; i < kMaxVerts; ++i) {
    tv[i] = mul(xf, poly.v[i]);
    tn[i] = mul(xf.q, poly.n[i]);
  }
  const float radius = poly.radius + e.radius;
  // b2ComputeEdgeSeparation (:205-239)
  int edge_index = -1;
  float edge_sep = -FLT_MAX;
  vec2 edge_normal = mk(0.0f, 0.0f);
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const vec2 axis = j == 0 ? normal1 : -normal1;
    float sj = FLT_MAX;
#pragma unroll
    for (int i = 0; i < kMaxVerts; ++i) {
      if (i < tc) {
        const float si = dot(axis, tv[i] - v1);
        if (si < sj) sj = si;
      }
    }
    if (sj > edge_sep) { edge_index = j; edge_sep = sj; edge_normal = axis; }
  }
  (void)edge_index;
  if (edge_sep > radius) return;
  // b2ComputePolygonSeparation (:241-267)
  int poly_type = 0;  // 0 unknown, 2 edgeB
  int poly_index = -1;
  float poly_sep = -FLT_MAX;
  vec2 poly_normal = mk(0.0f, 0.0f);
#pragma unroll
  for (int i = 0; i < kMaxVerts; ++i) {
    if (i < tc) {
      const vec2 n = -tn[i];
      const float s1 = dot(n, tv[i] - v1);
      const float s2 = dot(n, tv[i] - v2);
      const float s = minf(s1, s2);
      if (s > poly_sep) { poly_type = 2; poly_index = i; poly_sep = s; poly_normal = n; }
    }
  }
  if (poly_sep > radius) return;
  const float k_relative_tol = 0.98f, k_absolute_tol = 0.001f;
  // primary axis: type 1 = edgeA, 2 = edgeB (0 unknown behaves as edgeB's branch)
  int p_type;
  int p_index;
  vec2 p_normal;
  if (poly_sep - radius > k_relative_tol * (edge_sep - radius) + k_absolute_tol) {
    p_type = poly_type; p_index = poly_index; p_normal = poly_normal;
  } else {
    p_type = 1; p_index = edge_index; p_normal = edge_normal;
  }
  if (one_sided) {  // smooth collision (:332-387)
    const vec2 edge0 = normalize(v1 - e.v[0]);
    const vec2 normal0 = mk(edge0.y, -edge0.x);
    const bool convex1 = cross(edge0, edge1) >= 0.0f;
    const vec2 edge2 = normalize(e.v[3] - v2);
    const vec2 normal2 = mk(edge2.y, -edge2.x);
    const bool convex2 = cross(edge1, edge2) >= 0.0f;
    const float sin_tol = 0.1f;
    const bool side1 = dot(p_normal, edge1) <= 0.0f;
    if (side1) {
      if (convex1) {
        if (cross(p_normal, normal0) > sin_tol) return;
      } else {
        p_type = 1; p_index = edge_index; p_normal = edge_normal;
      }
    } else {
      if (convex2) {
        if (cross(normal2, p_normal) > sin_tol) return;
      } else {
        p_type = 1; p_index = edge_index; p_normal = edge_normal;
      }
    }
  }
  ClipV k0, k1;
  int ri1, ri2;
  vec2 rv1, rv2, rnormal, side1n, side2n;
  if (p_type == 1) {
    m.type = kManFaceA;
    int best = 0;
    float best_value = dot(p_normal, tn[0]);
#pragma unroll
    for (int i = 1; i < kMaxVerts; ++i) {
      if (i < tc) {
        const float value = dot(p_normal, tn[i]);
        if (value < best_value) { best_value = value; best = i; }
      }
    }
    const int i1 = best, i2 = i1 + 1 < tc ? i1 + 1 : 0;
    k0.v = pick(tv, i1);
    k0.id = feature(0, (uint32_t)i1, kFeatFace, kFeatVertex);
    k1.v = pick(tv, i2);
    k1.id = feature(0, (uint32_t)i2, kFeatFace, kFeatVertex);
    ri1 = 0; ri2 = 1;
    rv1 = v1; rv2 = v2;
    rnormal = p_normal;
    side1n = -edge1;
    side2n = edge1;
  } else {
    m.type = kManFaceB;
    k0.v = v2;
    k0.id = feature(1, (uint32_t)p_index, kFeatVertex, kFeatFace);
    k1.v = v1;
    k1.id = feature(0, (uint32_t)p_index, kFeatVertex, kFeatFace);
    ri1 = p_index;
    ri2 = ri1 + 1 < tc ? ri1 + 1 : 0;
    rv1 = pick(tv, ri1);
    rv2 = pick(tv, ri2);
    rnormal = pick(tn, ri1);
    side1n = mk(rnormal.y, -rnormal.x);
    side2n = -side1n;
  }
  const float side_offset1 = dot(side1n, rv1);
  const float side_offset2 = dot(side2n, rv2);
  ClipV c1a, c1b, c2a, c2b;
  int np = clip_segment(c1a, c1b, k0, k1, side1n, side_offset1, ri1);
  if (np < 2) return;
  np = clip_segment(c2a, c2b, c1a, c1b, side2n, side_offset2, ri2);
  if (np < 2) return;
  if (p_type == 1) {
    m.ln = rnormal;
    m.lp = rv1;
  } else {
    m.ln = pick(poly.n, ri1);
    m.lp = pick(poly.v, ri1);
  }
  int pc = 0;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const ClipV& cv = i == 0 ? c2a : c2b;
    const float separation = dot(rnormal, cv.v - rv1);
    if (separation <= radius) {
      vec2 lpt;
      uint32_t id;
      if (p_type == 1) {
        lpt = mul_t(xf, cv.v);
        id = cv.id;
      } else {
        lpt = cv.v;
        id = feature_swap(cv.id);
      }
      if (pc == 0) { m.lp0 = lpt; m.id0 = id; }
      else { m.lp1 = lpt; m.id1 = id; }
      ++pc;
    }
  }
  m.count = pc;
}

// Dispatch (b2Contact::Evaluate via the function table); false if the
// ordered type pair has no collide function.
__device__ __forceinline__ bool evaluate(Man& m, const Geo& a, const xform& xa, const Geo& b,
                                         const xform& xb) {
  if (a.type == kCircle) {
    if (b.type != kCircle) return false;
    collide_circles(m, a, xa, b, xb);
  } else if (a.type == kPolygon) {
    if (b.type == kCircle) collide_polygon_circle(m, a, xa, b, xb);
    else if (b.type == kPolygon) collide_polygons(m, a, xa, b, xb);
    else return false;
  } else {
    if (b.type == kCircle) collide_edge_circle(m, a, xa, b, xb);
    else if (b.type == kPolygon) collide_edge_polygon(m, a, xa, b, xb);
    else return false;
  }
  return true;
}

// ---- ray casts: b2CircleShape::RayCast (src/collision/b2_circle_shape.cpp:56-86),
// b2EdgeShape::RayCast (b2_edge_shape.cpp:89-154), b2PolygonShape::RayCast
// (b2_polygon_shape.cpp:303-371); same operation order, so the hits match the
// host's bit for bit when compiled without contraction.
B2G_INLINE bool raycast_shape(const Geo& g, const xform& xf, vec2 in1, vec2 in2, float max_fraction,
                              float& fraction, vec2& normal) {
  if (g.type == kCircle) {
    const vec2 position = xf.p + mul(xf.q, g.centroid);
    const vec2 s = in1 - position;
    const float b = dot(s, s) - g.radius * g.radius;
    const vec2 r = in2 - in1;
    const float c = dot(s, r);
    const float rr = dot(r, r);
    const float sigma = c * c - rr * b;
    if (sigma < 0.0f || rr < kEpsilon) return false;
    float a = -(c + sqrtf(sigma));
    if (0.0f <= a && a <= max_fraction * rr) {
      a /= rr;
      fraction = a;
      normal = normalize(s + a * r);
      return true;
    }
    return false;
  }
  const vec2 p1 = mul_t(xf.q, in1 - xf.p);
  const vec2 p2 = mul_t(xf.q, in2 - xf.p);
  const vec2 d = p2 - p1;
  if (g.type == kEdge) {
    const vec2 v1 = g.v[1], v2 = g.v[2];
    const vec2 e = v2 - v1;
    const vec2 nrm = normalize(mk(e.y, -e.x));
    const float numerator = dot(nrm, v1 - p1);
    if (g.count != 0 && numerator > 0.0f) return false;
    const float denominator = dot(nrm, d);
    if (denominator == 0.0f) return false;
    const float t = numerator / denominator;
    if (t < 0.0f || max_fraction < t) return false;
    const vec2 q = p1 + t * d;
    const vec2 r = v2 - v1;
    const float rr = dot(r, r);
    if (rr == 0.0f) return false;
    const float s = dot(q - v1, r) / rr;
    if (s < 0.0f || 1.0f < s) return false;
    fraction = t;
    normal = numerator > 0.0f ? -mul(xf.q, nrm) : mul(xf.q, nrm);
    return true;
  }
  float lower = 0.0f, upper = max_fraction;
  int index = -1;
#pragma unroll
  for (int i = 0; i < kMaxVerts; ++i) {
    if (i < g.count) {
      const float numerator = dot(g.n[i], g.v[i] - p1);
      const float denominator = dot(g.n[i], d);
      if (denominator == 0.0f) {
        if (numerator < 0.0f) return false;
      } else if (denominator < 0.0f && numerator < lower * denominator) {
        lower = numerator / denominator;
        index = i;
      } else if (denominator > 0.0f && numerator < upper * denominator) {
        upper = numerator / denominator;
      }
      if (upper < lower) return false;
    }
  }
  if (index >= 0) {
    fraction = lower;
    normal = mul(xf.q, pick(g.n, index));
    return true;
  }
  return false;
}

B2G_INLINE bool has_collide_function(int ta, int tb) {
  return (ta == kCircle && tb == kCircle) || (ta == kPolygon && (tb == kCircle || tb == kPolygon)) ||
         (ta == kEdge && (tb == kCircle || tb == kPolygon));
}

}  // namespace b2g
