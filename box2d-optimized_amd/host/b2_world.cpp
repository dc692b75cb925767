// b2_world.cpp -- host side of the drop-in API: shape set-up, mass data,
// body/joint bookkeeping and the b2World::Step driver over the b2g C ABI.
//
// Set-up arithmetic (b2PolygonShape::Set, SetAsBox, ComputeMass, ComputeAABB,
// b2Body::ResetMassData, b2RevoluteJointDef::Initialize) follows the
// reference's single-precision expressions (files cited per function) so the
// device receives the same shapes and mass properties the reference computes.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <unordered_set>

#include "box2d/box2d.h"

const b2Vec2 b2Vec2_zero(0.0f, 0.0f);

// ---------------------------------------------------------------------------
// shapes
// ---------------------------------------------------------------------------
void b2CircleShape::ComputeAABB(b2AABB* aabb, const b2Transform& xf) const {  // b2_circle_shape.cpp:91-96
  b2Vec2 p = xf.p + b2Mul(xf.q, m_p);
  aabb->lowerBound.Set(p.x - m_radius, p.y - m_radius);
  aabb->upperBound.Set(p.x + m_radius, p.y + m_radius);
}

void b2CircleShape::ComputeMass(b2MassData* md, float density) const {  // :98-105
  md->mass = density * b2_pi * m_radius * m_radius;
  md->center = m_p;
  md->I = md->mass * (0.5f * m_radius * m_radius + b2Dot(m_p, m_p));
}

void b2EdgeShape::SetOneSided(const b2Vec2& v0, const b2Vec2& v1, const b2Vec2& v2, const b2Vec2& v3) {
  m_vertex0 = v0;
  m_vertex1 = v1;
  m_vertex2 = v2;
  m_vertex3 = v3;
  m_oneSided = true;
}

void b2EdgeShape::SetTwoSided(const b2Vec2& v1, const b2Vec2& v2) {
  m_vertex1 = v1;
  m_vertex2 = v2;
  m_oneSided = false;
}

void b2EdgeShape::ComputeAABB(b2AABB* aabb, const b2Transform& xf) const {  // b2_edge_shape.cpp:156-167
  b2Vec2 v1 = b2Mul(xf, m_vertex1);
  b2Vec2 v2 = b2Mul(xf, m_vertex2);
  b2Vec2 lower = b2Min(v1, v2);
  b2Vec2 upper = b2Max(v1, v2);
  b2Vec2 r(m_radius, m_radius);
  aabb->lowerBound = lower - r;
  aabb->upperBound = upper + r;
}

void b2EdgeShape::ComputeMass(b2MassData* md, float density) const {
  B2_NOT_USED(density);
  md->mass = 0.0f;
  md->center = 0.5f * (m_vertex1 + m_vertex2);
  md->I = 0.0f;
}

void b2PolygonShape::SetAsBox(float hx, float hy) {  // b2_polygon_shape.cpp:36-48
  m_count = 4;
  m_vertices[0].Set(-hx, -hy);
  m_vertices[1].Set(hx, -hy);
  m_vertices[2].Set(hx, hy);
  m_vertices[3].Set(-hx, hy);
  m_normals[0].Set(0.0f, -1.0f);
  m_normals[1].Set(1.0f, 0.0f);
  m_normals[2].Set(0.0f, 1.0f);
  m_normals[3].Set(-1.0f, 0.0f);
  m_centroid.SetZero();
}

void b2PolygonShape::SetAsBox(float hx, float hy, const b2Vec2& center, float angle) {  // :50-73
  SetAsBox(hx, hy);
  m_centroid = center;
  b2Transform xf;
  xf.p = center;
  xf.q.Set(angle);
  for (int32 i = 0; i < m_count; ++i) {
    m_vertices[i] = b2Mul(xf, m_vertices[i]);
    m_normals[i] = b2Mul(xf.q, m_normals[i]);
  }
}

static b2Vec2 polygon_centroid(const b2Vec2* vs, int32 count) {  // :75-111
  b2Vec2 c(0.0f, 0.0f);
  float area = 0.0f;
  const b2Vec2 s = vs[0];
  const float inv3 = 1.0f / 3.0f;
  for (int32 i = 0; i < count; ++i) {
    b2Vec2 p1 = vs[0] - s;
    b2Vec2 p2 = vs[i] - s;
    b2Vec2 p3 = i + 1 < count ? vs[i + 1] - s : vs[0] - s;
    b2Vec2 e1 = p2 - p1;
    b2Vec2 e2 = p3 - p1;
    float D = b2Cross(e1, e2);
    float tri = 0.5f * D;
    area += tri;
    c += tri * inv3 * (p1 + p2 + p3);
  }
  c = (1.0f / area) * c + s;
  return c;
}

// weld (half linear slop), gift-wrap hull, normals, centroid (:113-243)
void b2PolygonShape::Set(const b2Vec2* vertices, int32 count) {
  if (count < 3) {
    SetAsBox(1.0f, 1.0f);
    return;
  }
  int32 n = b2Min(count, b2_maxPolygonVertices);
  b2Vec2 ps[b2_maxPolygonVertices];
  int32 temp = 0;
  for (int32 i = 0; i < n; ++i) {
    b2Vec2 v = vertices[i];
    bool unique = true;
    for (int32 j = 0; j < temp; ++j) {
      if (b2DistanceSquared(v, ps[j]) < ((0.5f * b2_linearSlop) * (0.5f * b2_linearSlop))) {
        unique = false;
        break;
      }
    }
    if (unique) ps[temp++] = v;
  }
  n = temp;
  if (n < 3) {
    SetAsBox(1.0f, 1.0f);
    return;
  }
  int32 i0 = 0;
  float x0 = ps[0].x;
  for (int32 i = 1; i < n; ++i) {
    float x = ps[i].x;
    if (x > x0 || (x == x0 && ps[i].y < ps[i0].y)) {
      i0 = i;
      x0 = x;
    }
  }
  int32 hull[b2_maxPolygonVertices];
  int32 m = 0;
  int32 ih = i0;
  for (;;) {
    if (m >= b2_maxPolygonVertices) break;
    hull[m] = ih;
    int32 ie = 0;
    for (int32 j = 1; j < n; ++j) {
      if (ie == ih) {
        ie = j;
        continue;
      }
      b2Vec2 r = ps[ie] - ps[hull[m]];
      b2Vec2 v = ps[j] - ps[hull[m]];
      float c = b2Cross(r, v);
      if (c < 0.0f) ie = j;
      if (c == 0.0f && v.LengthSquared() > r.LengthSquared()) ie = j;
    }
    ++m;
    ih = ie;
    if (ie == i0) break;
  }
  if (m < 3) {
    SetAsBox(1.0f, 1.0f);
    return;
  }
  m_count = m;
  for (int32 i = 0; i < m; ++i) m_vertices[i] = ps[hull[i]];
  for (int32 i = 0; i < m; ++i) {
    int32 i2 = i + 1 < m ? i + 1 : 0;
    b2Vec2 edge = m_vertices[i2] - m_vertices[i];
    m_normals[i] = b2Cross(edge, 1.0f);
    m_normals[i].Normalize();
  }
  m_centroid = polygon_centroid(m_vertices, m);
}

void b2PolygonShape::ComputeAABB(b2AABB* aabb, const b2Transform& xf) const {  // :373-388
  b2Vec2 lower = b2Mul(xf, m_vertices[0]);
  b2Vec2 upper = lower;
  for (int32 i = 1; i < m_count; ++i) {
    b2Vec2 v = b2Mul(xf, m_vertices[i]);
    lower = b2Min(lower, v);
    upper = b2Max(upper, v);
  }
  b2Vec2 r(m_radius, m_radius);
  aabb->lowerBound = lower - r;
  aabb->upperBound = upper + r;
}

void b2PolygonShape::ComputeMass(b2MassData* md, float density) const {  // :390-464
  b2Vec2 center(0.0f, 0.0f);
  float area = 0.0f;
  float I = 0.0f;
  const b2Vec2 s = m_vertices[0];
  const float k_inv3 = 1.0f / 3.0f;
  for (int32 i = 0; i < m_count; ++i) {
    b2Vec2 e1 = m_vertices[i] - s;
    b2Vec2 e2 = i + 1 < m_count ? m_vertices[i + 1] - s : m_vertices[0] - s;
    float D = b2Cross(e1, e2);
    float tri = 0.5f * D;
    area += tri;
    center += tri * k_inv3 * (e1 + e2);
    float ex1 = e1.x, ey1 = e1.y;
    float ex2 = e2.x, ey2 = e2.y;
    float intx2 = ex1 * ex1 + ex2 * ex1 + ex2 * ex2;
    float inty2 = ey1 * ey1 + ey2 * ey1 + ey2 * ey2;
    I += (0.25f * k_inv3 * D) * (intx2 + inty2);
  }
  md->mass = density * area;
  center *= 1.0f / area;
  md->center = center + s;
  md->I = density * I;
  md->I += md->mass * (b2Dot(md->center, md->center) - b2Dot(center, center));
}

// b2WorldManifold::Initialize (src/collision/b2_collision.cpp:26-90)
void b2WorldManifold::Initialize(const b2Manifold* manifold, const b2Transform& xfA, float radiusA,
                                 const b2Transform& xfB, float radiusB) {
  if (manifold->pointCount == 0) return;
  switch (manifold->type) {
    case b2Manifold::e_circles: {
      normal.Set(1.0f, 0.0f);
      b2Vec2 pointA = b2Mul(xfA, manifold->localPoint);
      b2Vec2 pointB = b2Mul(xfB, manifold->points[0].localPoint);
      if (b2DistanceSquared(pointA, pointB) > b2_epsilon * b2_epsilon) {
        normal = pointB - pointA;
        normal.Normalize();
      }
      b2Vec2 cA = pointA + radiusA * normal;
      b2Vec2 cB = pointB - radiusB * normal;
      points[0] = 0.5f * (cA + cB);
      separations[0] = b2Dot(cB - cA, normal);
    } break;
    case b2Manifold::e_faceA: {
      normal = b2Mul(xfA.q, manifold->localNormal);
      b2Vec2 planePoint = b2Mul(xfA, manifold->localPoint);
      for (int32 i = 0; i < manifold->pointCount; ++i) {
        b2Vec2 clipPoint = b2Mul(xfB, manifold->points[i].localPoint);
        b2Vec2 cA = clipPoint + (radiusA - b2Dot(clipPoint - planePoint, normal)) * normal;
        b2Vec2 cB = clipPoint - radiusB * normal;
        points[i] = 0.5f * (cA + cB);
        separations[i] = b2Dot(cB - cA, normal);
      }
    } break;
    case b2Manifold::e_faceB: {
      normal = b2Mul(xfB.q, manifold->localNormal);
      b2Vec2 planePoint = b2Mul(xfB, manifold->localPoint);
      for (int32 i = 0; i < manifold->pointCount; ++i) {
        b2Vec2 clipPoint = b2Mul(xfA, manifold->points[i].localPoint);
        b2Vec2 cB = clipPoint + (radiusB - b2Dot(clipPoint - planePoint, normal)) * normal;
        b2Vec2 cA = clipPoint - radiusA * normal;
        points[i] = 0.5f * (cA + cB);
        separations[i] = b2Dot(cA - cB, normal);
      }
      normal = -normal;
    } break;
  }
}

// ---------------------------------------------------------------------------
// fixtures / bodies
// ---------------------------------------------------------------------------
const b2AABB& b2Fixture::GetAABB() const {
  m_shape->ComputeAABB(&m_aabb, m_body->GetTransform());
  return m_aabb;
}

b2Body::b2Body(const b2BodyDef* bd, b2World* world, int32 index) {  // b2_body.cpp:29-122
  m_world = world;
  m_index = index;
  m_type = bd->type;
  m_xf.p = bd->position;
  m_xf.q.Set(bd->angle);
  m_localCenter.SetZero();
  m_c = m_xf.p;
  m_a = bd->angle;
  m_linearVelocity = bd->linearVelocity;
  m_angularVelocity = bd->angularVelocity;
  m_force.SetZero();
  m_torque = 0.0f;
  m_mass = m_invMass = m_I = m_invI = 0.0f;
  m_linearDamping = bd->linearDamping;
  m_angularDamping = bd->angularDamping;
  m_gravityScale = bd->gravityScale;
  m_fixedRotation = bd->fixedRotation;
  m_bullet = bd->bullet;
  m_sleepingAllowed = bd->allowSleep;
  m_awake = bd->awake && bd->type != b2_staticBody;
  m_enabled = bd->enabled;
  m_userData = bd->userData;
}

b2Body::~b2Body() {
  b2Fixture* f = m_fixtureList;
  while (f) {
    b2Fixture* n = f->m_next;
    delete f;
    f = n;
  }
}

void b2Body::MarkDirty() {
  m_world->m_dirty[m_index] = 1;
  m_world->m_anyDirty = true;
}

void b2Body::Sync() const { m_world->SyncBodies(); }

// b2Body::ResetMassData (src/dynamics/b2_body.cpp:352-416)
void b2Body::ResetMassData() {
  m_mass = 0.0f;
  m_invMass = 0.0f;
  m_I = 0.0f;
  m_invI = 0.0f;
  m_localCenter.SetZero();
  if (m_type == b2_staticBody || m_type == b2_kinematicBody) {
    m_c = m_xf.p;
    return;
  }
  b2Vec2 localCenter = b2Vec2_zero;
  for (b2Fixture* f = m_fixtureList; f; f = f->m_next) {
    if (f->m_density == 0.0f) continue;
    b2MassData md;
    f->GetMassData(&md);
    m_mass += md.mass;
    localCenter += md.mass * md.center;
    m_I += md.I;
  }
  if (m_mass > 0.0f) {
    m_invMass = 1.0f / m_mass;
    localCenter *= m_invMass;
  }
  if (m_I > 0.0f && !m_fixedRotation) {
    m_I -= m_mass * b2Dot(localCenter, localCenter);
    m_invI = 1.0f / m_I;
  } else {
    m_I = 0.0f;
    m_invI = 0.0f;
  }
  b2Vec2 oldCenter = m_c;
  m_localCenter = localCenter;
  m_c = b2Mul(m_xf, m_localCenter);
  m_linearVelocity += b2Cross(m_angularVelocity, m_c - oldCenter);
}

b2Fixture* b2Body::CreateFixture(const b2FixtureDef* def) {  // b2_body.cpp:230-269
  b2Assert(!m_world->IsLocked());
  if (m_world->IsLocked() || !def || !def->shape) return nullptr;
  Sync();
  b2Fixture* f = new b2Fixture();
  f->m_body = this;
  f->m_shape = def->shape->Clone();
  f->m_density = def->density;
  f->m_friction = def->friction;
  f->m_restitution = def->restitution;
  f->m_restitutionThreshold = def->restitutionThreshold;
  f->m_filter = def->filter;
  f->m_isSensor = def->isSensor;
  f->m_userData = def->userData;
  f->m_index = (int32)m_world->m_fixtures.size();
  f->m_next = m_fixtureList;
  m_fixtureList = f;
  ++m_fixtureCount;
  m_world->m_fixtures.push_back(f);
  m_world->m_refilter.push_back(0);
  if (f->m_isSensor) ++m_world->m_sensorFixtures;
  if (f->m_density > 0.0f) ResetMassData();
  MarkDirty();
  m_world->m_newContacts = true;
  return f;
}

b2Fixture* b2Body::CreateFixture(const b2Shape* shape, float density) {
  b2FixtureDef def;
  def.shape = shape;
  def.density = density;
  return CreateFixture(&def);
}

// b2_body.cpp:129-226
void b2Body::SetType(b2BodyType type) {
  b2Assert(!m_world->IsLocked());
  if (m_world->IsLocked() || m_type == type) return;
  b2World* w = m_world;
  w->BeginEdit();
  const bool was_static = m_type == b2_staticBody;
  if (was_static || type == b2_staticBody) {  // statics sit at the tail of the body list
    if (m_prev) m_prev->m_next = m_next; else w->m_bodyListHead = m_next;
    if (m_next) m_next->m_prev = m_prev; else w->m_bodyListTail = m_prev;
    m_prev = m_next = nullptr;
    if (type == b2_staticBody) {
      m_prev = w->m_bodyListTail;
      if (w->m_bodyListTail) w->m_bodyListTail->m_next = this; else w->m_bodyListHead = this;
      w->m_bodyListTail = this;
    } else {
      m_next = w->m_bodyListHead;
      if (w->m_bodyListHead) w->m_bodyListHead->m_prev = this; else w->m_bodyListTail = this;
      w->m_bodyListHead = this;
    }
  }
  m_type = type;
  ResetMassData();
  if (m_type == b2_staticBody) {
    m_linearVelocity.SetZero();
    m_angularVelocity = 0.0f;
    m_awake = false;
  } else {
    m_awake = true;  // SetAwake(true)
    m_sleepTime = 0.0f;
  }
  m_force.SetZero();
  m_torque = 0.0f;
  // the body's contacts are destroyed (EndContact for touching ones)
  const b2Body* self = this;
  w->RemoveContactsIf(
      [](const b2World* wd, const b2g_contact& c, const void* arg) {
        const b2Body* b = static_cast<const b2Body*>(arg);
        return wd->m_fixtures[c.shape_a]->m_body == b || wd->m_fixtures[c.shape_b]->m_body == b;
      },
      self);
  MarkDirty();
}

// b2_body.cpp:280-350
void b2Body::DestroyFixture(b2Fixture* fixture) {
  if (fixture == nullptr) return;
  b2Assert(!m_world->IsLocked());
  if (m_world->IsLocked()) return;
  b2Assert(fixture->m_body == this);
  b2Fixture** node = &m_fixtureList;
  bool found = false;
  while (*node) {
    if (*node == fixture) {
      *node = fixture->m_next;
      found = true;
      break;
    }
    node = &(*node)->m_next;
  }
  b2Assert(found);
  if (!found) return;
  b2World* w = m_world;
  w->BeginEdit();
  const int32 index = fixture->m_index;
  w->RemoveContactsIf(
      [](const b2World*, const b2g_contact& c, const void* arg) {
        const int32 k = *static_cast<const int32*>(arg);
        return c.shape_a == k || c.shape_b == k;
      },
      &index);
  if (fixture->m_isSensor) --w->m_sensorFixtures;
  w->RemoveFixtureIndex(index);
  fixture->m_body = nullptr;
  fixture->m_next = nullptr;
  delete fixture;
  --m_fixtureCount;
  ResetMassData();
  MarkDirty();
}

// b2_body.cpp:418-455
void b2Body::SetMassData(const b2MassData* massData) {
  b2Assert(!m_world->IsLocked());
  if (m_world->IsLocked() || m_type != b2_dynamicBody) return;
  Sync();
  m_invMass = 0.0f;
  m_I = 0.0f;
  m_invI = 0.0f;
  m_mass = massData->mass;
  if (m_mass <= 0.0f) m_mass = 1.0f;
  m_invMass = 1.0f / m_mass;
  if (massData->I > 0.0f && !m_fixedRotation) {
    m_I = massData->I - m_mass * b2Dot(massData->center, massData->center);
    b2Assert(m_I > 0.0f);
    m_invI = 1.0f / m_I;
  }
  b2Vec2 oldCenter = m_c;
  m_localCenter = massData->center;
  m_c = b2Mul(m_xf, m_localCenter);
  m_linearVelocity += b2Cross(m_angularVelocity, m_c - oldCenter);
  MarkDirty();
}

// include/box2d/b2_body.h:719-740 (either way the sleep timer restarts)
void b2Body::SetAwake(bool flag) {
  if (m_type == b2_staticBody) return;
  Sync();
  m_sleepTime = 0.0f;
  if (flag) {
    m_awake = true;
  } else {
    m_awake = false;
    m_linearVelocity.SetZero();
    m_angularVelocity = 0.0f;
    m_force.SetZero();
    m_torque = 0.0f;
  }
  MarkDirty();
}

// include/box2d/b2_body.h:747-758
void b2Body::SetSleepingAllowed(bool flag) {
  Sync();
  m_sleepingAllowed = flag;
  if (!flag) SetAwake(true);
  MarkDirty();
}

bool b2Body::IsAwake() const {
  Sync();  // the device puts bodies to sleep and wakes them
  return m_awake;
}

// b2_body.cpp:512-558: enabling lets the next Step find the body's contacts
void b2Body::SetEnabled(bool flag) {
  b2Assert(!m_world->IsLocked());
  if (flag == m_enabled) return;
  m_enabled = flag;
  m_world->m_disabledBodies += flag ? -1 : 1;
  if (flag) m_world->m_newContacts = true;
}

// b2_body.cpp:560-580
void b2Body::SetFixedRotation(bool flag) {
  if (flag == m_fixedRotation) return;
  Sync();
  m_fixedRotation = flag;
  m_angularVelocity = 0.0f;
  ResetMassData();
  MarkDirty();
}

void b2Body::SetTransform(const b2Vec2& position, float angle) {  // b2_body.cpp:486-504
  b2Assert(!m_world->IsLocked());
  if (m_world->IsLocked()) return;
  Sync();
  m_xf.q.Set(angle);
  m_xf.p = position;
  m_c = b2Mul(m_xf, m_localCenter);
  m_a = angle;
  MarkDirty();
  m_world->m_newContacts = true;
}

const b2Transform& b2Body::GetTransform() const { Sync(); return m_xf; }
const b2Vec2& b2Body::GetPosition() const { Sync(); return m_xf.p; }
float b2Body::GetAngle() const { Sync(); return m_a; }
const b2Vec2& b2Body::GetWorldCenter() const { Sync(); return m_c; }
const b2Vec2& b2Body::GetLinearVelocity() const { Sync(); return m_linearVelocity; }
float b2Body::GetAngularVelocity() const { Sync(); return m_angularVelocity; }

void b2Body::GetMassData(b2MassData* data) const {
  data->mass = m_mass;
  data->I = m_I + m_mass * b2Dot(m_localCenter, m_localCenter);
  data->center = m_localCenter;
}

void b2Body::SetLinearVelocity(const b2Vec2& v) {
  if (m_type == b2_staticBody) return;
  Sync();
  m_linearVelocity = v;
  MarkDirty();
}

void b2Body::SetAngularVelocity(float w) {
  if (m_type == b2_staticBody) return;
  Sync();
  m_angularVelocity = w;
  MarkDirty();
}

// The appliers wake a sleeping body when asked and do nothing to a body that
// stays asleep (include/box2d/b2_body.h:803-905).
void b2Body::ApplyForce(const b2Vec2& force, const b2Vec2& point, bool wake) {
  if (m_type != b2_dynamicBody) return;
  if (wake && !m_awake) SetAwake(true);
  if (!m_awake) return;
  Sync();
  m_force += force;
  m_torque += b2Cross(point - m_c, force);
  MarkDirty();
}

void b2Body::ApplyForceToCenter(const b2Vec2& force, bool wake) {
  if (m_type != b2_dynamicBody) return;
  if (wake && !m_awake) SetAwake(true);
  if (!m_awake) return;
  Sync();
  m_force += force;
  MarkDirty();
}

void b2Body::ApplyTorque(float torque, bool wake) {
  if (m_type != b2_dynamicBody) return;
  if (wake && !m_awake) SetAwake(true);
  if (!m_awake) return;
  Sync();
  m_torque += torque;
  MarkDirty();
}

void b2Body::ApplyLinearImpulse(const b2Vec2& impulse, const b2Vec2& point, bool wake) {
  if (m_type != b2_dynamicBody) return;
  if (wake && !m_awake) SetAwake(true);
  if (!m_awake) return;
  Sync();
  m_linearVelocity += m_invMass * impulse;
  m_angularVelocity += m_invI * b2Cross(point - m_c, impulse);
  MarkDirty();
}

void b2Body::ApplyLinearImpulseToCenter(const b2Vec2& impulse, bool wake) {
  if (m_type != b2_dynamicBody) return;
  if (wake && !m_awake) SetAwake(true);
  if (!m_awake) return;
  Sync();
  m_linearVelocity += m_invMass * impulse;
  MarkDirty();
}

void b2Body::ApplyAngularImpulse(float impulse, bool wake) {
  if (m_type != b2_dynamicBody) return;
  if (wake && !m_awake) SetAwake(true);
  if (!m_awake) return;
  Sync();
  m_angularVelocity += m_invI * impulse;
  MarkDirty();
}

void b2Body::ToRecord(b2g_body* r) const {
  std::memset(r, 0, sizeof(*r));
  r->type = (int32_t)m_type;
  r->world = 0;
  r->xf[0] = m_xf.p.x; r->xf[1] = m_xf.p.y; r->xf[2] = m_xf.q.s; r->xf[3] = m_xf.q.c;
  r->c[0] = m_c.x; r->c[1] = m_c.y; r->a = m_a;
  r->local_center[0] = m_localCenter.x; r->local_center[1] = m_localCenter.y;
  r->v[0] = m_linearVelocity.x; r->v[1] = m_linearVelocity.y; r->w = m_angularVelocity;
  r->mass = m_mass; r->inv_mass = m_invMass; r->inv_i = m_invI;
  r->linear_damping = m_linearDamping; r->angular_damping = m_angularDamping; r->gravity_scale = m_gravityScale;
  r->force[0] = m_force.x; r->force[1] = m_force.y; r->torque = m_torque;
  r->flags = (m_awake || m_type == b2_staticBody ? 0 : B2G_BODY_ASLEEP) | (m_sleepingAllowed ? 0 : B2G_BODY_NO_AUTOSLEEP);
  r->sleep_time = m_sleepTime;
}

void b2Body::FromRecord(const b2g_body& r) {
  m_xf.p.Set(r.xf[0], r.xf[1]);
  m_xf.q.s = r.xf[2];
  m_xf.q.c = r.xf[3];
  m_c.Set(r.c[0], r.c[1]);
  m_a = r.a;
  m_linearVelocity.Set(r.v[0], r.v[1]);
  m_angularVelocity = r.w;
  m_force.Set(r.force[0], r.force[1]);
  m_torque = r.torque;
  m_awake = m_type == b2_staticBody || (r.flags & B2G_BODY_ASLEEP) == 0;
  m_sleepTime = r.sleep_time;
}

void b2Fixture::MarkShapeDirty() {
  if (m_body) m_body->m_world->m_shapesDirty = true;
}

// b2_fixture.cpp:142-183
void b2Fixture::SetFilterData(const b2Filter& filter) {
  m_filter = filter;
  Refilter();
}

void b2Fixture::Refilter() {
  if (m_body == nullptr) return;
  b2World* w = m_body->m_world;
  w->m_refilter[m_index] = 1;
  w->m_anyRefilter = true;
  w->m_shapesDirty = true;
}

void b2Fixture::SetSensor(bool sensor) {  // b2_fixture.cpp:185-194
  if (sensor == m_isSensor) return;
  if (m_body) {
    m_body->SetAwake(true);
    m_body->m_world->m_sensorFixtures += sensor ? 1 : -1;
  }
  m_isSensor = sensor;
}

void b2Fixture::SetDensity(float density) { m_density = density; }  // ResetMassData applies it
void b2Fixture::SetFriction(float friction) { m_friction = friction; MarkShapeDirty(); }
void b2Fixture::SetRestitution(float restitution) { m_restitution = restitution; MarkShapeDirty(); }
void b2Fixture::SetRestitutionThreshold(float threshold) { m_restitutionThreshold = threshold; MarkShapeDirty(); }

// src/dynamics/b2_world_callbacks.cpp:28-40
bool b2ContactFilter::ShouldCollide(b2Fixture* fixtureA, b2Fixture* fixtureB) {
  const b2Filter& filterA = fixtureA->GetFilterData();
  const b2Filter& filterB = fixtureB->GetFilterData();
  if (filterA.groupIndex == filterB.groupIndex && filterA.groupIndex != 0) return filterA.groupIndex > 0;
  return (filterA.maskBits & filterB.categoryBits) != 0 && (filterA.categoryBits & filterB.maskBits) != 0;
}

void b2ContactListener::PreSolve(b2Contact* contact, const b2Manifold* oldManifold) {
  B2_NOT_USED(contact);
  B2_NOT_USED(oldManifold);
}

void b2ContactListener::PostSolve(b2Contact* contact, const b2ContactImpulse* impulse) {
  B2_NOT_USED(contact);
  B2_NOT_USED(impulse);
}

// Whether an object's class overrides a virtual function: the function a
// virtual call on the object reaches, against the one it reaches on the base
// class (GNU C++'s bound member function pointer, taken in this file only).
#pragma GCC diagnostic push
#pragma GCC diagnostic ignored "-Wpmf-conversions"
static bool overrides_pre_solve(b2ContactListener* l) {
  static b2ContactListener base;
  typedef void (*Fn)(b2ContactListener*, b2Contact*, const b2Manifold*);
  return (Fn)(l->*(&b2ContactListener::PreSolve)) != (Fn)(base.*(&b2ContactListener::PreSolve));
}
static bool overrides_post_solve(b2ContactListener* l) {
  static b2ContactListener base;
  typedef void (*Fn)(b2ContactListener*, b2Contact*, const b2ContactImpulse*);
  return (Fn)(l->*(&b2ContactListener::PostSolve)) != (Fn)(base.*(&b2ContactListener::PostSolve));
}
static bool overrides_should_collide(b2ContactFilter* f) {
  static b2ContactFilter base;
  typedef bool (*Fn)(b2ContactFilter*, b2Fixture*, b2Fixture*);
  return (Fn)(f->*(&b2ContactFilter::ShouldCollide)) != (Fn)(base.*(&b2ContactFilter::ShouldCollide));
}
#pragma GCC diagnostic pop

// ---------------------------------------------------------------------------
// joints
// ---------------------------------------------------------------------------
void b2RevoluteJointDef::Initialize(b2Body* bA, b2Body* bB, const b2Vec2& anchor) {  // b2_revolute_joint.cpp:41-48
  bodyA = bA;
  bodyB = bB;
  localAnchorA = bodyA->GetLocalPoint(anchor);
  localAnchorB = bodyB->GetLocalPoint(anchor);
  referenceAngle = bodyB->GetAngle() - bodyA->GetAngle();
}

b2RevoluteJoint::b2RevoluteJoint(const b2RevoluteJointDef* def) {
  m_type = e_revoluteJoint;
  m_bodyA = def->bodyA;
  m_bodyB = def->bodyB;
  m_collideConnected = def->collideConnected;
  m_localAnchorA = def->localAnchorA;
  m_localAnchorB = def->localAnchorB;
  m_referenceAngle = def->referenceAngle;
  m_enableLimit = def->enableLimit;
  m_enableMotor = def->enableMotor;
  m_lowerAngle = def->lowerAngle;
  m_upperAngle = def->upperAngle;
  m_motorSpeed = def->motorSpeed;
  m_maxMotorTorque = def->maxMotorTorque;
  m_impulse.SetZero();
  m_motorImpulse = m_lowerImpulse = m_upperImpulse = 0.0f;
}

void b2RevoluteJoint::Sync() const { m_world->SyncJoints(); }

b2Vec2 b2RevoluteJoint::GetReactionForce(float inv_dt) const {  // b2_revolute_joint.cpp (GetReactionForce)
  Sync();
  b2Vec2 P(m_impulse.x, m_impulse.y);
  return inv_dt * P;
}

float b2RevoluteJoint::GetReactionTorque(float inv_dt) const {
  Sync();
  return inv_dt * (m_motorImpulse + m_lowerImpulse - m_upperImpulse);
}

float b2RevoluteJoint::GetMotorTorque(float inv_dt) const {
  Sync();
  return inv_dt * m_motorImpulse;
}

float b2RevoluteJoint::GetJointAngle() const {
  return m_bodyB->GetAngle() - m_bodyA->GetAngle() - m_referenceAngle;
}

// ---------------------------------------------------------------------------
// contacts
// ---------------------------------------------------------------------------
void b2Contact::GetWorldManifold(b2WorldManifold* wm) const {
  const b2Body* bodyA = m_fixtureA->GetBody();
  const b2Body* bodyB = m_fixtureB->GetBody();
  wm->Initialize(&m_manifold, bodyA->GetTransform(), m_fixtureA->GetShape()->m_radius, bodyB->GetTransform(),
                 m_fixtureB->GetShape()->m_radius);
}

// ---------------------------------------------------------------------------
// world
// ---------------------------------------------------------------------------
b2World::b2World(const b2Vec2& gravity) : m_gravity(gravity) { std::memset(&m_profile, 0, sizeof(m_profile)); }

b2World::~b2World() {
  if (m_batch) b2g_batch_destroy(m_batch);
  for (b2Joint* j : m_joints) delete j;
  for (b2Body* b : m_bodies) delete b;
}

b2Body* b2World::CreateBody(const b2BodyDef* def) {  // b2_world.cpp:140-176
  b2Assert(!IsLocked());
  if (IsLocked()) return nullptr;
  SyncBodies();
  b2Body* b = new b2Body(def, this, (int32)m_bodies.size());
  m_bodies.push_back(b);
  m_dirty.push_back(1);
  m_anyDirty = true;
  if (!b->m_enabled) ++m_disabledBodies;
  // statics at the tail, others at the head of the body list
  if (b->m_type == b2_staticBody) {
    b->m_prev = m_bodyListTail;
    if (m_bodyListTail) m_bodyListTail->m_next = b; else m_bodyListHead = b;
    m_bodyListTail = b;
  } else {
    b->m_next = m_bodyListHead;
    if (m_bodyListHead) m_bodyListHead->m_prev = b; else m_bodyListTail = b;
    m_bodyListHead = b;
  }
  return b;
}

// b2_world.cpp:178-266
void b2World::DestroyBody(b2Body* b) {
  b2Assert(!m_bodies.empty());
  b2Assert(!IsLocked());
  if (IsLocked() || b == nullptr || b->m_world != this) return;
  BeginEdit();
  // attached joints first
  for (size_t q = 0; q < m_joints.size();) {
    b2Joint* j = m_joints[q];
    if (j->m_bodyA == b || j->m_bodyB == b) {
      if (m_destructionListener) m_destructionListener->SayGoodbye(j);
      DestroyJoint(j);
    } else {
      ++q;
    }
  }
  // then the contacts (EndContact for touching ones)
  RemoveContactsIf(
      [](const b2World* w, const b2g_contact& c, const void* arg) {
        const b2Body* body = static_cast<const b2Body*>(arg);
        return w->m_fixtures[c.shape_a]->m_body == body || w->m_fixtures[c.shape_b]->m_body == body;
      },
      b);
  // then the fixtures
  b2Fixture* f = b->m_fixtureList;
  while (f) {
    b2Fixture* f0 = f;
    f = f->m_next;
    if (m_destructionListener) m_destructionListener->SayGoodbye(f0);
    if (f0->m_isSensor) --m_sensorFixtures;
    RemoveFixtureIndex(f0->m_index);
    delete f0;
    b->m_fixtureList = f;
    --b->m_fixtureCount;
  }
  b->m_fixtureList = nullptr;
  b->m_fixtureCount = 0;
  if (!b->m_enabled) --m_disabledBodies;
  if (b->m_prev) b->m_prev->m_next = b->m_next; else m_bodyListHead = b->m_next;
  if (b->m_next) b->m_next->m_prev = b->m_prev; else m_bodyListTail = b->m_prev;
  const int32 index = b->m_index;
  m_bodies.erase(m_bodies.begin() + index);
  m_dirty.erase(m_dirty.begin() + index);
  for (size_t i = index; i < m_bodies.size(); ++i) m_bodies[i]->m_index = (int32)i;
  delete b;
}

// b2_world.cpp:326-418. A joint with collideConnected == false kept its
// bodies from colliding; the next broadphase may now pair them.
void b2World::DestroyJoint(b2Joint* j) {
  b2Assert(!IsLocked());
  if (IsLocked() || j == nullptr) return;
  BeginEdit();
  if (j->m_prev) j->m_prev->m_next = j->m_next; else m_jointList = j->m_next;
  if (j->m_next) j->m_next->m_prev = j->m_prev;
  j->m_bodyA->SetAwake(true);
  j->m_bodyB->SetAwake(true);
  const int32 index = j->m_index;
  m_joints.erase(m_joints.begin() + index);
  for (size_t i = index; i < m_joints.size(); ++i) m_joints[i]->m_index = (int32)i;
  delete j;
}

void b2World::SetAllowSleeping(bool flag) {
  if (flag == m_allowSleep) return;
  m_allowSleep = flag;
  if (!m_allowSleep)
    for (b2Body* b : m_bodies) b->SetAwake(true);
}

void b2World::SetContactListener(b2ContactListener* listener) {
  m_listener = listener;
  m_listenerPreSolve = listener != nullptr && overrides_pre_solve(listener);
  m_listenerPostSolve = listener != nullptr && overrides_post_solve(listener);
}

// Everything that changes the contact list or renumbers records outside a
// step starts here: the device's bodies, joint impulses and contacts (with
// their colours, so the solve order of the survivors is kept) come to the
// host once; EnsureDevice sends them back renumbered.
void b2World::BeginEdit() {
  if (m_editing) return;
  m_editing = true;
  m_editContacts.clear();
  m_editColours.clear();
  if (!m_batch) return;
  SyncBodies();
  SyncJoints();
  const int n = b2g_download_contacts(m_batch, 0, 0, nullptr);
  if (n > 0) {
    m_editContacts.resize(n);
    m_editColours.resize(n);
    if (b2g_download_contacts(m_batch, 0, n, m_editContacts.data()) != B2G_OK ||
        b2g_download_contact_colours(m_batch, 0, n, m_editColours.data()) != B2G_OK) {
      m_editContacts.clear();
      m_editColours.clear();
    }
  }
  m_contactsValid = false;
}

// b2ContactManager::Destroy (b2_contact_manager.cpp:48-62): EndContact for a
// touching contact, then it leaves the list
void b2World::RemoveContactsIf(bool (*pred)(const b2World*, const b2g_contact&, const void*), const void* arg) {
  size_t k = 0;
  for (size_t i = 0; i < m_editContacts.size(); ++i) {
    const b2g_contact& c = m_editContacts[i];
    if (pred(this, c, arg)) {
      if (m_listener && c.touching) {
        b2Contact hc;
        ContactFromRecord(c, &hc);
        m_listener->EndContact(&hc);
      }
      continue;
    }
    m_editContacts[k] = c;
    if (k < m_editColours.size() && i < m_editColours.size()) m_editColours[k] = m_editColours[i];
    ++k;
  }
  m_editContacts.resize(k);
  if (m_editColours.size() > k) m_editColours.resize(k);
  m_contactsValid = false;
}

// fixture `index` is gone: later fixtures move down one, and so do the
// contacts' shape indices (no contact names the fixture any more)
void b2World::RemoveFixtureIndex(int32 index) {
  m_fixtures.erase(m_fixtures.begin() + index);
  m_refilter.erase(m_refilter.begin() + index);
  for (size_t i = index; i < m_fixtures.size(); ++i) m_fixtures[i]->m_index = (int32)i;
  for (b2g_contact& c : m_editContacts) {
    if (c.shape_a > index) --c.shape_a;
    if (c.shape_b > index) --c.shape_b;
  }
  m_shapesDirty = true;
}

// b2ContactManager::Collide's re-filtering of flagged contacts
// (b2_contact_manager.cpp:76-92): b2Body::ShouldCollide (b2_body.cpp:459-480)
// and the filter, at the start of the step
void b2World::ApplyRefilter() {
  m_anyRefilter = false;
  BeginEdit();
  RemoveContactsIf(
      [](const b2World* w, const b2g_contact& c, const void*) {
        b2Fixture* fa = w->m_fixtures[c.shape_a];
        b2Fixture* fb = w->m_fixtures[c.shape_b];
        if (!w->m_refilter[c.shape_a] && !w->m_refilter[c.shape_b]) return false;
        const b2Body* ba = fa->m_body;
        const b2Body* bb = fb->m_body;
        if (ba->m_type != b2_dynamicBody && bb->m_type != b2_dynamicBody) return true;
        for (const b2Joint* j : w->m_joints) {
          const bool links = (j->m_bodyA == ba && j->m_bodyB == bb) || (j->m_bodyA == bb && j->m_bodyB == ba);
          if (links && !j->m_collideConnected) return true;
        }
        b2ContactFilter def;
        return !def.ShouldCollide(fa, fb);
      },
      nullptr);
  std::fill(m_refilter.begin(), m_refilter.end(), 0);
}

const char* b2World::UnsupportedReason() const {
  if (m_continuousPhysics) return "continuous physics (TOI) is not run on the device: call SetContinuousPhysics(false)";
  if (m_disabledBodies > 0) return "a body is disabled (SetEnabled(false) or created with enabled = false)";
  if (m_sensorFixtures > 0) return "sensor fixtures are not run on the device";
  if (m_contactFilter && overrides_should_collide(m_contactFilter))
    return "a contact filter that overrides ShouldCollide cannot run on the device";
  if (m_listenerPreSolve) return "a contact listener that overrides PreSolve cannot run on the device";
  return nullptr;
}

b2Joint* b2World::CreateJoint(const b2JointDef* def) {  // b2_world.cpp:268-324
  b2Assert(!IsLocked());
  if (IsLocked() || !def) return nullptr;
  if (def->type != e_revoluteJoint) {
    b2Assert(false && "only revolute joints run on the device path");
    return nullptr;
  }
  SyncJoints();
  b2RevoluteJoint* j = new b2RevoluteJoint(static_cast<const b2RevoluteJointDef*>(def));
  j->m_world = this;
  j->m_index = (int32)m_joints.size();
  j->m_next = m_jointList;
  if (m_jointList) m_jointList->m_prev = j;
  m_jointList = j;
  m_joints.push_back(j);
  if (!def->collideConnected) m_newContacts = true;  // existing contacts are re-filtered
  return j;
}

int b2World::EnsureDevice() {
  const int nb = (int)m_bodies.size(), ns = (int)m_fixtures.size(), nj = (int)m_joints.size();
  const bool fits = m_batch && nb <= m_cap.bodies && ns <= m_cap.shapes && nj <= m_cap.joints;
  if (fits && !m_editing) return B2G_OK;
  // the contact list (with warm-start impulses and colours) survives a move
  // to a larger batch and a renumbering
  std::vector<b2g_contact> keep;
  std::vector<int32_t> keep_colours;
  if (m_editing) {
    keep.swap(m_editContacts);
    keep_colours.swap(m_editColours);
  } else if (m_batch) {
    SyncBodies();
    SyncJoints();
    const int n = b2g_download_contacts(m_batch, 0, 0, nullptr);
    if (n > 0) {
      keep.resize(n);
      keep_colours.resize(n);
      int rc = b2g_download_contacts(m_batch, 0, n, keep.data());
      if (!rc) rc = b2g_download_contact_colours(m_batch, 0, n, keep_colours.data());
      if (rc) return rc;
    }
  }
  m_editing = false;
  int rc;
  if (fits) {
    if ((rc = b2g_set_counts(m_batch, nb, ns, nj))) return rc;
  } else {
    if (m_batch) b2g_batch_destroy(m_batch);
    m_batch = nullptr;
    m_cap.worlds = 1;
    m_cap.bodies = std::max(64, nb * 2);
    m_cap.shapes = std::max(64, ns * 2);
    m_cap.joints = std::max(16, nj * 2);
    m_cap.contacts = std::max(4096, std::min(ns * 16, (1 << 24) - 1));
    if ((rc = b2g_batch_create(&m_batch, &m_cap))) return rc;
    b2g_world_params wp{{m_gravity.x, m_gravity.y}, 1.0f / 60.0f, 0.0f, m_warmStarting ? 1 : 0, 1, m_allowSleep ? 1 : 0};
    if ((rc = b2g_set_worlds(m_batch, 1, nullptr, nullptr, &wp))) return rc;
  }
  m_uploadedBodies = m_uploadedShapes = m_uploadedJoints = 0;
  std::fill(m_dirty.begin(), m_dirty.end(), 1);
  m_anyDirty = true;
  if (!keep.empty()) {
    // bodies/shapes must exist before contacts refer to them
    if ((rc = Upload())) return rc;
    if ((rc = b2g_upload_contacts(m_batch, (int32_t)keep.size(), keep.data()))) return rc;
    if (keep_colours.size() == keep.size() &&
        (rc = b2g_upload_contact_colours(m_batch, (int32_t)keep_colours.size(), keep_colours.data())))
      return rc;
  }
  m_contactsValid = false;
  return B2G_OK;
}

int b2World::Prepare() {
  int rc = EnsureDevice();
  if (!rc) rc = Upload();
  if (!rc) {
    b2g_world_params wp{{m_gravity.x, m_gravity.y}, 1.0f / 60.0f, 0.0f, m_warmStarting ? 1 : 0, 1, m_allowSleep ? 1 : 0};
    rc = b2g_set_world_params(m_batch, 0, 1, &wp);
  }
  return rc;
}

int b2World::Upload() {
  int rc = B2G_OK;
  const int nb = (int)m_bodies.size();
  if (m_anyDirty) {
    std::vector<b2g_body> rec;
    int i = 0;
    while (i < nb) {
      if (!m_dirty[i]) { ++i; continue; }
      int j = i;
      while (j < nb && m_dirty[j]) ++j;
      rec.resize(j - i);
      for (int k = i; k < j; ++k) m_bodies[k]->ToRecord(&rec[k - i]);
      if ((rc = b2g_upload_bodies(m_batch, i, j - i, rec.data()))) return rc;
      for (int k = i; k < j; ++k) m_dirty[k] = 0;
      i = j;
    }
    m_anyDirty = false;
    m_uploadedBodies = nb;
  }
  const int ns = (int)m_fixtures.size();
  if (m_shapesDirty) {  // a fixture's filter or material changed, or fixtures moved
    m_uploadedShapes = 0;
    m_shapesDirty = false;
  }
  if (ns > m_uploadedShapes) {
    std::vector<b2g_shape> rec(ns - m_uploadedShapes);
    for (int i = m_uploadedShapes; i < ns; ++i) {
      FixtureToRecord(m_fixtures[i], &rec[i - m_uploadedShapes]);
    }
    if ((rc = b2g_upload_shapes(m_batch, m_uploadedShapes, ns - m_uploadedShapes, rec.data()))) return rc;
    m_uploadedShapes = ns;
  }
  const int nj = (int)m_joints.size();
  if (nj > m_uploadedJoints) {
    std::vector<b2g_revolute> rec(nj - m_uploadedJoints);
    for (int i = m_uploadedJoints; i < nj; ++i) {
      JointToRecord(m_joints[i], &rec[i - m_uploadedJoints]);
    }
    if ((rc = b2g_upload_joints(m_batch, m_uploadedJoints, nj - m_uploadedJoints, rec.data()))) return rc;
    m_uploadedJoints = nj;
  }
  return rc;
}

void b2World::FixtureToRecord(const b2Fixture* f, b2g_shape* out) {
  b2g_shape& s = *out;
  std::memset(&s, 0, sizeof(s));
  const b2Shape* sh = f->m_shape;
  s.type = (int32_t)sh->GetType();
  s.body = f->m_body->m_index;
  s.radius = sh->m_radius;
  if (s.type == B2G_CIRCLE) {
    const b2CircleShape* c = static_cast<const b2CircleShape*>(sh);
    s.centroid[0] = c->m_p.x;
    s.centroid[1] = c->m_p.y;
  } else if (s.type == B2G_EDGE) {
    const b2EdgeShape* e = static_cast<const b2EdgeShape*>(sh);
    s.count = e->m_oneSided ? 1 : 0;
    const b2Vec2 v[4] = {e->m_vertex0, e->m_vertex1, e->m_vertex2, e->m_vertex3};
    for (int k = 0; k < 4; ++k) {
      s.vertices[2 * k] = v[k].x;
      s.vertices[2 * k + 1] = v[k].y;
    }
  } else {
    const b2PolygonShape* p = static_cast<const b2PolygonShape*>(sh);
    s.count = p->m_count;
    s.centroid[0] = p->m_centroid.x;
    s.centroid[1] = p->m_centroid.y;
    for (int k = 0; k < p->m_count; ++k) {
      s.vertices[2 * k] = p->m_vertices[k].x;
      s.vertices[2 * k + 1] = p->m_vertices[k].y;
      s.normals[2 * k] = p->m_normals[k].x;
      s.normals[2 * k + 1] = p->m_normals[k].y;
    }
  }
  s.friction = f->m_friction;
  s.restitution = f->m_restitution;
  s.restitution_threshold = f->m_restitutionThreshold;
  s.category = f->m_filter.categoryBits;
  s.mask = f->m_filter.maskBits;
  s.group = f->m_filter.groupIndex;
  s.is_sensor = f->m_isSensor ? 1 : 0;
}

void b2World::JointToRecord(const b2Joint* joint, b2g_revolute* out) {
  const b2RevoluteJoint* j = static_cast<const b2RevoluteJoint*>(joint);
  b2g_revolute& r = *out;
  std::memset(&r, 0, sizeof(r));
  r.body_a = j->m_bodyA->m_index;
  r.body_b = j->m_bodyB->m_index;
  r.local_anchor_a[0] = j->m_localAnchorA.x; r.local_anchor_a[1] = j->m_localAnchorA.y;
  r.local_anchor_b[0] = j->m_localAnchorB.x; r.local_anchor_b[1] = j->m_localAnchorB.y;
  r.reference_angle = j->m_referenceAngle;
  r.lower_angle = j->m_lowerAngle; r.upper_angle = j->m_upperAngle;
  r.motor_speed = j->m_motorSpeed; r.max_motor_torque = j->m_maxMotorTorque;
  r.flags = (j->m_collideConnected ? 1 : 0) | (j->m_enableLimit ? 2 : 0) | (j->m_enableMotor ? 4 : 0);
  r.impulse[0] = j->m_impulse.x; r.impulse[1] = j->m_impulse.y;
  r.motor_impulse = j->m_motorImpulse; r.lower_impulse = j->m_lowerImpulse; r.upper_impulse = j->m_upperImpulse;
}

int b2World::ExportRecords(int32_t world, int32_t body_offset, b2g_body* bodies, b2g_shape* shapes,
                           b2g_revolute* joints) const {
  SyncBodies();
  SyncJoints();
  for (size_t i = 0; i < m_bodies.size(); ++i) {
    m_bodies[i]->ToRecord(&bodies[i]);
    bodies[i].world = world;
  }
  for (size_t i = 0; i < m_fixtures.size(); ++i) {
    FixtureToRecord(m_fixtures[i], &shapes[i]);
    shapes[i].body += body_offset;
  }
  for (size_t i = 0; i < m_joints.size(); ++i) {
    JointToRecord(m_joints[i], &joints[i]);
    joints[i].body_a += body_offset;
    joints[i].body_b += body_offset;
  }
  return B2G_OK;
}

void b2World::SyncBodies() const {
  if (m_bodiesValid || !m_batch) return;
  b2World* self = const_cast<b2World*>(this);
  const int nb = (int)m_bodies.size();
  std::vector<b2g_body> rec(nb);
  if (nb && b2g_download_bodies(m_batch, 0, nb, rec.data()) == B2G_OK) {
    for (int i = 0; i < nb; ++i)
      if (!m_dirty[i]) m_bodies[i]->FromRecord(rec[i]);
  }
  self->m_bodiesValid = true;
}

void b2World::SyncJoints() const {
  if (m_jointsValid || !m_batch) return;
  const int nj = (int)m_joints.size();
  std::vector<b2g_revolute> rec(nj);
  if (nj && b2g_download_joints(m_batch, 0, nj, rec.data()) == B2G_OK) {
    for (int i = 0; i < nj; ++i) {
      b2RevoluteJoint* j = static_cast<b2RevoluteJoint*>(m_joints[i]);
      j->m_impulse.Set(rec[i].impulse[0], rec[i].impulse[1]);
      j->m_motorImpulse = rec[i].motor_impulse;
      j->m_lowerImpulse = rec[i].lower_impulse;
      j->m_upperImpulse = rec[i].upper_impulse;
    }
  }
  const_cast<b2World*>(this)->m_jointsValid = true;
}

void b2World::QueryAABB(b2QueryCallback* callback, const b2AABB& aabb) {
  if (!callback || m_fixtures.empty()) return;
  if ((m_status = EnsureDevice()) || (m_status = Upload())) return;
  const float box[4] = {aabb.lowerBound.x, aabb.lowerBound.y, aabb.upperBound.x, aabb.upperBound.y};
  const int32_t world = 0;
  std::vector<int32_t> out(64);
  int32_t n = 0;
  for (;;) {
    if ((m_status = b2g_query_aabbs(m_batch, 1, &world, box, (int32_t)out.size(), out.data(), &n))) return;
    if (n <= (int32_t)out.size()) break;
    out.resize(n);
  }
  for (int32_t i = 0; i < n; ++i) callback->ReportFixture(m_fixtures[out[i]]);
}

void b2World::RayCast(b2RayCastCallback* callback, const b2Vec2& point1, const b2Vec2& point2) {
  if (!callback || m_fixtures.empty()) return;
  if ((m_status = EnsureDevice()) || (m_status = Upload())) return;
  const float seg[4] = {point1.x, point1.y, point2.x, point2.y};
  const int32_t world = 0;
  std::vector<b2g_ray_hit> hits(64);
  int32_t n = 0;
  for (;;) {
    if ((m_status = b2g_raycast(m_batch, 1, &world, seg, (int32_t)hits.size(), hits.data(), &n))) return;
    if (n <= (int32_t)hits.size()) break;
    hits.resize(n);
  }
  hits.resize(n);
  std::sort(hits.begin(), hits.end(), [](const b2g_ray_hit& a, const b2g_ray_hit& b) {
    return a.fraction != b.fraction ? a.fraction < b.fraction : a.shape < b.shape;
  });
  // the reference's clipping (include/box2d/b2_broad_phase.h:696-712), in fraction order
  float max_fraction = 1.0f;
  for (const b2g_ray_hit& h : hits) {
    if (h.fraction > max_fraction) break;
    const float f = h.fraction;
    const b2Vec2 point = (1.0f - f) * point1 + f * point2;
    const float value = callback->ReportFixture(m_fixtures[h.shape], point, b2Vec2(h.normal[0], h.normal[1]), f);
    if (value == 0.0f) return;
    if (value > 0.0f) max_fraction = value;
  }
}

void b2World::ContactFromRecord(const b2g_contact& r, b2Contact* out) const {
  b2Contact& c = *out;
  c.m_fixtureA = m_fixtures[r.shape_a];
  c.m_fixtureB = m_fixtures[r.shape_b];
  c.m_touching = r.touching != 0;
  b2Manifold& m = c.m_manifold;
  m.type = (b2Manifold::Type)r.type;
  m.pointCount = r.point_count;
  m.localNormal.Set(r.local_normal[0], r.local_normal[1]);
  m.localPoint.Set(r.local_point[0], r.local_point[1]);
  for (int k = 0; k < 2; ++k) {
    m.points[k].localPoint.Set(r.point_local[2 * k], r.point_local[2 * k + 1]);
    m.points[k].normalImpulse = r.normal_impulse[k];
    m.points[k].tangentImpulse = r.tangent_impulse[k];
    m.points[k].id.key = r.id[k];
  }
  c.m_friction = r.friction;
  c.m_restitution = r.restitution;
  c.m_restitutionThreshold = r.restitution_threshold;
  c.m_tangentSpeed = r.tangent_speed;
}

void b2World::SyncContacts() {
  if (m_contactsValid) return;
  m_contactsValid = true;
  m_contacts.clear();
  if (m_editing) {  // the edited list has not gone back to the device yet
    m_contacts.resize(m_editContacts.size());
    for (size_t i = 0; i < m_editContacts.size(); ++i) ContactFromRecord(m_editContacts[i], &m_contacts[i]);
    for (size_t i = 0; i + 1 < m_contacts.size(); ++i) m_contacts[i].m_next = &m_contacts[i + 1];
    return;
  }
  if (!m_batch) return;
  const int n = b2g_download_contacts(m_batch, 0, 0, nullptr);
  if (n <= 0) return;
  std::vector<b2g_contact> rec(n);
  if (b2g_download_contacts(m_batch, 0, n, rec.data()) != B2G_OK) return;
  m_contacts.resize(n);
  for (int i = 0; i < n; ++i) ContactFromRecord(rec[i], &m_contacts[i]);
  for (int i = 0; i + 1 < n; ++i) m_contacts[i].m_next = &m_contacts[i + 1];
}

b2Contact* b2World::GetContactListStart() {
  SyncContacts();
  return m_contacts.empty() ? nullptr : &m_contacts[0];
}

int32 b2World::GetContactCount() const {
  if (m_editing) return (int32)m_editContacts.size();
  if (!m_batch) return 0;
  b2g_counts c;
  if (b2g_get_counts(m_batch, &c) != B2G_OK) return 0;
  return c.contacts;
}

// b2World::Step (src/dynamics/b2_world.cpp:1108-1171): Collide, Solve with the
// broadphase at its end, ClearForces; with dt <= 0, Collide and ClearForces
// only (and dtRatio's inv_dt0 is kept). A world that needs a feature the
// device does not run is refused before anything changes.
void b2World::Step(float dt, int32 velocityIterations, int32 positionIterations, int32 particleIterations) {
  B2_NOT_USED(particleIterations);  // no particle systems exist here
  if (const char* why = UnsupportedReason()) {
    m_status = B2G_ERR_UNSUPPORTED;
    m_statusMessage = why;
    if (m_lastRefusal != why) std::fprintf(stderr, "b2World::Step refused: %s\n", why);
    m_lastRefusal = why;
    b2Assert(false && "b2World::Step: a feature of this world does not run on the device");
    return;
  }
  m_lastRefusal.clear();
  const auto t0 = std::chrono::steady_clock::now();
  m_locked = true;
  m_status = B2G_OK;
  if (m_anyRefilter) ApplyRefilter();
  m_status = EnsureDevice();
  if (!m_status) m_status = Upload();
  if (!m_status) {
    const float dtRatio = m_inv_dt0 * dt;
    b2g_world_params wp{{m_gravity.x, m_gravity.y}, dt, dtRatio, m_warmStarting ? 1 : 0, 1, m_allowSleep ? 1 : 0};
    m_status = b2g_set_world_params(m_batch, 0, 1, &wp);
  }
  if (!m_status) {
    uint32_t flags = m_profiling ? B2G_STEP_PROFILE : 0u;
    if (m_newContacts) flags |= B2G_STEP_FIND_NEW_CONTACTS;
    if (m_listener) flags |= B2G_STEP_EVENTS;
    if (dt <= 0.0f) flags |= B2G_STEP_COLLIDE_ONLY;
    m_status = b2g_batch_step(m_batch, velocityIterations, positionIterations, flags);
    if (!m_status && m_profiling) {
      b2g_profile p;
      b2g_get_profile(m_batch, &p);
      m_profile.collide = p.collide;
      m_profile.solve = p.solve;
      m_profile.broadphase = p.broadphase;
    }
  }
  if (m_status) {
    m_statusMessage = b2g_last_error();
    std::fprintf(stderr, "b2World::Step (device): %s\n", m_statusMessage.c_str());
    b2Assert(false && "device step failed");
  }
  m_newContacts = false;
  if (dt > 0.0f) m_inv_dt0 = 1.0f / dt;
  m_bodiesValid = false;
  m_jointsValid = false;
  m_contactsValid = false;
  for (b2Body* b : m_bodies) {  // ClearForces (host mirror; the device cleared its copy)
    b->m_force.SetZero();
    b->m_torque = 0.0f;
  }
  if (m_listener && !m_status) DeliverEvents();
  m_locked = false;
  m_profile.step = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// The listener calls of the step just taken, in the reference's phase order:
// Begin/EndContact from Collide (b2_contact.cpp:197-203), PostSolve for every
// contact the solver stored impulses for (b2_island.cpp:621-647), then
// EndContact for touching contacts the broadphase destroyed
// (b2_contact_manager.cpp:48-51).
void b2World::DeliverEvents() {
  int32_t n = 0;
  std::vector<b2g_contact_event> ev;
  if (b2g_download_events(m_batch, 0, nullptr, &n) == B2G_OK && n > 0) {
    ev.resize(n);
    if (b2g_download_events(m_batch, n, ev.data(), &n) != B2G_OK) ev.clear();
  }
  for (const b2g_contact_event& e : ev) {
    if (e.kind != B2G_EVENT_BEGIN && e.kind != B2G_EVENT_END) continue;
    b2Contact c;
    ContactFromRecord(e.contact, &c);
    if (e.kind == B2G_EVENT_BEGIN) m_listener->BeginContact(&c);
    else m_listener->EndContact(&c);
  }
  if (m_listenerPostSolve) {
    auto post = [&](b2Contact* c, int32 points) {
      b2ContactImpulse imp;
      imp.count = points;
      for (int k = 0; k < b2_maxManifoldPoints; ++k) {
        imp.normalImpulses[k] = k < points ? c->m_manifold.points[k].normalImpulse : 0.0f;
        imp.tangentImpulses[k] = k < points ? c->m_manifold.points[k].tangentImpulse : 0.0f;
      }
      m_listener->PostSolve(c, &imp);
    };
    SyncContacts();
    std::vector<int32_t> points(m_contacts.size(), 0);
    if (!points.empty() &&
        b2g_download_solver_points(m_batch, 0, (int32_t)points.size(), points.data()) != B2G_OK)
      points.assign(points.size(), 0);
    for (size_t i = 0; i < m_contacts.size(); ++i)
      if (points[i] > 0) post(&m_contacts[i], points[i]);
    for (const b2g_contact_event& e : ev) {
      if (e.kind != B2G_EVENT_DESTROYED || e.solver_points <= 0) continue;
      b2Contact c;
      ContactFromRecord(e.contact, &c);
      post(&c, e.solver_points);
    }
  }
  for (const b2g_contact_event& e : ev) {
    if (e.kind != B2G_EVENT_DESTROYED) continue;
    b2Contact c;
    ContactFromRecord(e.contact, &c);
    m_listener->EndContact(&c);
  }
}

// ---------------------------------------------------------------------------
// scene loader (C ABI): builds a b2World through the API above
// ---------------------------------------------------------------------------
struct b2a_world {
  b2World* world;
  // the world's objects in creation order (as the device numbers them)
  std::vector<b2Body*> bodies;
  std::vector<b2Fixture*> fixtures;
  std::vector<b2Joint*> joints;
};

// Builds a b2World from a scene description through the public API
// (CreateBody / CreateFixture / CreateJoint), as user code would.
static b2a_world* world_from_scene(const b2g_scene* s, int32_t warm_starting) {
  b2a_world* w = new b2a_world();
  w->world = new b2World(b2Vec2(s->gravity[0], s->gravity[1]));
  w->world->SetContinuousPhysics(false);
  w->world->SetAllowSleeping(false);
  w->world->SetWarmStarting(warm_starting != 0);
  w->bodies.resize(s->body_count);
  for (int i = 0; i < s->body_count; ++i) {
    b2BodyDef bd;
    bd.type = (b2BodyType)s->body_type[i];
    bd.position.Set(s->body_pos[2 * i], s->body_pos[2 * i + 1]);
    bd.angle = s->body_angle[i];
    if (s->body_vel) {
      bd.linearVelocity.Set(s->body_vel[3 * i], s->body_vel[3 * i + 1]);
      bd.angularVelocity = s->body_vel[3 * i + 2];
    }
    if (s->body_damping) {
      bd.linearDamping = s->body_damping[2 * i];
      bd.angularDamping = s->body_damping[2 * i + 1];
    }
    if (s->body_gravity_scale) bd.gravityScale = s->body_gravity_scale[i];
    if (s->body_flags) {
      bd.fixedRotation = (s->body_flags[i] & 1) != 0;
      bd.bullet = (s->body_flags[i] & 2) != 0;
      bd.allowSleep = (s->body_flags[i] & 4) == 0;
      bd.awake = (s->body_flags[i] & 8) == 0;
    }
    w->bodies[i] = w->world->CreateBody(&bd);
  }
  for (int i = 0; i < s->fixture_count; ++i) {
    const float* g = s->fix_geom + 17 * i;
    b2CircleShape circle;
    b2EdgeShape edge;
    b2PolygonShape poly;
    const b2Shape* shape = nullptr;
    switch (s->fix_kind[i]) {
      case 0:
        circle.m_radius = g[0];
        circle.m_p.Set(g[1], g[2]);
        shape = &circle;
        break;
      case 1:
        if (g[0] != 0.0f)
          edge.SetOneSided(b2Vec2(g[1], g[2]), b2Vec2(g[3], g[4]), b2Vec2(g[5], g[6]), b2Vec2(g[7], g[8]));
        else
          edge.SetTwoSided(b2Vec2(g[3], g[4]), b2Vec2(g[5], g[6]));
        shape = &edge;
        break;
      case 2: {
        b2Vec2 pts[b2_maxPolygonVertices];
        const int n = (int)g[0];
        for (int k = 0; k < n && k < b2_maxPolygonVertices; ++k) pts[k].Set(g[1 + 2 * k], g[2 + 2 * k]);
        poly.Set(pts, n);
        shape = &poly;
        break;
      }
      default:
        if (g[5] != 0.0f) poly.SetAsBox(g[0], g[1], b2Vec2(g[2], g[3]), g[4]);
        else poly.SetAsBox(g[0], g[1]);
        shape = &poly;
        break;
    }
    b2FixtureDef fd;
    fd.shape = shape;
    fd.density = s->fix_material[4 * i + 0];
    fd.friction = s->fix_material[4 * i + 1];
    fd.restitution = s->fix_material[4 * i + 2];
    fd.restitutionThreshold = s->fix_material[4 * i + 3];
    if (s->fix_filter) {
      fd.filter.categoryBits = (uint16)s->fix_filter[4 * i + 0];
      fd.filter.maskBits = (uint16)s->fix_filter[4 * i + 1];
      fd.filter.groupIndex = (int16)s->fix_filter[4 * i + 2];
      fd.isSensor = s->fix_filter[4 * i + 3] != 0;
    }
    w->fixtures.push_back(w->bodies[s->fix_body[i]]->CreateFixture(&fd));
  }
  for (int i = 0; i < s->joint_count; ++i) {
    b2RevoluteJointDef jd;
    const float* p = s->joint_params + 6 * i;
    jd.Initialize(w->bodies[s->joint_bodies[2 * i]], w->bodies[s->joint_bodies[2 * i + 1]], b2Vec2(p[0], p[1]));
    const int f = s->joint_flags[i];
    jd.collideConnected = (f & 1) != 0;
    jd.enableLimit = (f & 2) != 0;
    jd.enableMotor = (f & 4) != 0;
    jd.lowerAngle = p[2];
    jd.upperAngle = p[3];
    jd.motorSpeed = p[4];
    jd.maxMotorTorque = p[5];
    w->joints.push_back(w->world->CreateJoint(&jd));
  }
  return w;
}

extern "C" {

int b2a_world_from_scene(b2a_world** out, const b2g_scene* s, int32_t warm_starting) {
  if (!out || !s) return B2G_ERR_INVALID;
  *out = world_from_scene(s, warm_starting);
  return B2G_OK;
}

int b2a_scene_records(const b2g_scene* s, int32_t world, int32_t body_offset, b2g_body* bodies,
                      b2g_shape* shapes, b2g_revolute* joints) {
  if (!s || !bodies || (s->fixture_count && !shapes) || (s->joint_count && !joints)) return B2G_ERR_INVALID;
  b2a_world* w = world_from_scene(s, 1);
  const int rc = w->world->ExportRecords(world, body_offset, bodies, shapes, joints);
  b2a_world_destroy(w);
  return rc;
}

int b2a_batch_from_scenes(b2g_batch** out, int32_t count, const b2g_scene* scenes, int32_t contacts_per_shape,
                          int32_t warm_starting) {
  if (!out || !scenes || count < 1 || contacts_per_shape < 1) return B2G_ERR_INVALID;
  int64_t nb = 0, ns = 0, nj = 0;
  for (int i = 0; i < count; ++i) {
    nb += scenes[i].body_count;
    ns += scenes[i].fixture_count;
    nj += scenes[i].joint_count;
  }
  if (nb >= (1ll << 31) || ns * contacts_per_shape >= (1ll << 28)) return B2G_ERR_CAPACITY;  // b2g_batch_create's limit
  std::vector<b2g_body> bodies(nb);
  std::vector<b2g_shape> shapes(ns);
  std::vector<b2g_revolute> joints(nj);
  std::vector<b2g_world_params> params(count);
  std::vector<int32_t> body_begin(count + 1), shape_begin(count + 1);
  int64_t ob = 0, os = 0, oj = 0;
  for (int i = 0; i < count; ++i) {
    const b2g_scene& s = scenes[i];
    body_begin[i] = (int32_t)ob;
    shape_begin[i] = (int32_t)os;
    int rc = b2a_scene_records(&s, i, (int32_t)ob, bodies.data() + ob, shapes.data() + os, joints.data() + oj);
    if (rc) return rc;
    params[i] = b2g_world_params{{s.gravity[0], s.gravity[1]}, 1.0f / 60.0f, 0.0f, warm_starting ? 1 : 0, 1};
    ob += s.body_count;
    os += s.fixture_count;
    oj += s.joint_count;
  }
  body_begin[count] = (int32_t)ob;
  shape_begin[count] = (int32_t)os;
  b2g_capacity cap{count, (int32_t)std::max<int64_t>(nb, 1), (int32_t)std::max<int64_t>(ns, 1),
                   (int32_t)std::max<int64_t>(nj, 1), (int32_t)std::max<int64_t>(ns * contacts_per_shape, 1024)};
  b2g_batch* b = nullptr;
  int rc = b2g_batch_create(&b, &cap);
  if (rc) return rc;
  if (!rc) rc = b2g_set_worlds(b, count, body_begin.data(), shape_begin.data(), params.data());
  if (!rc) rc = b2g_upload_bodies(b, 0, (int32_t)nb, bodies.data());
  if (!rc && ns) rc = b2g_upload_shapes(b, 0, (int32_t)ns, shapes.data());
  if (!rc && nj) rc = b2g_upload_joints(b, 0, (int32_t)nj, joints.data());
  if (rc) {
    b2g_batch_destroy(b);
    return rc;
  }
  *out = b;
  return B2G_OK;
}

int b2a_world_destroy(b2a_world* w) {
  if (!w) return B2G_OK;
  delete w->world;
  delete w;
  return B2G_OK;
}

int b2a_world_step(b2a_world* w, float dt, int32_t vi, int32_t pi) {
  if (!w) return B2G_ERR_INVALID;
  w->world->Step(dt, vi, pi);
  return w->world->GetLastStatus();
}

int b2a_world_get_bodies(b2a_world* w, float* out) {
  if (!w) return B2G_ERR_INVALID;
  for (size_t i = 0; i < w->bodies.size(); ++i) {
    b2Body* b = w->bodies[i];
    const b2Vec2 c = b->GetWorldCenter();
    const b2Vec2 v = b->GetLinearVelocity();
    const b2Transform& xf = b->GetTransform();
    float* o = out + 10 * i;
    o[0] = c.x; o[1] = c.y; o[2] = b->GetAngle();
    o[3] = v.x; o[4] = v.y; o[5] = b->GetAngularVelocity();
    o[6] = xf.p.x; o[7] = xf.p.y; o[8] = xf.q.s; o[9] = xf.q.c;
  }
  return B2G_OK;
}

int b2a_world_contact_count(b2a_world* w, int32_t* out) {
  if (!w || !out) return B2G_ERR_INVALID;
  *out = w->world->GetContactCount();
  return B2G_OK;
}

int b2a_world_get_contacts(b2a_world* w, int32_t capacity, b2g_contact* out, int32_t* count) {
  if (!w) return B2G_ERR_INVALID;
  b2g_batch* b = w->world->GetDeviceBatch();
  if (!b) {
    if (count) *count = 0;
    return B2G_OK;
  }
  const int n = b2g_download_contacts(b, 0, 0, nullptr);
  if (n < 0) return B2G_ERR_HIP;
  if (count) *count = n;
  if (out && n > 0) return b2g_download_contacts(b, 0, std::min(n, (int)capacity), out);
  return B2G_OK;
}

b2g_batch* b2a_world_batch(b2a_world* w) { return w ? w->world->GetDeviceBatch() : nullptr; }

int b2a_world_destroy_body(b2a_world* w, int32_t body) {
  if (!w || body < 0 || body >= (int32_t)w->bodies.size()) return B2G_ERR_INVALID;
  b2Body* b = w->bodies[body];
  // its fixtures and joints go with it
  for (size_t i = 0; i < w->fixtures.size();) {
    if (w->fixtures[i]->GetBody() == b) w->fixtures.erase(w->fixtures.begin() + i); else ++i;
  }
  for (size_t i = 0; i < w->joints.size();) {
    if (w->joints[i]->GetBodyA() == b || w->joints[i]->GetBodyB() == b) w->joints.erase(w->joints.begin() + i);
    else ++i;
  }
  w->bodies.erase(w->bodies.begin() + body);
  w->world->DestroyBody(b);
  return B2G_OK;
}

int b2a_world_destroy_joint(b2a_world* w, int32_t joint) {
  if (!w || joint < 0 || joint >= (int32_t)w->joints.size()) return B2G_ERR_INVALID;
  b2Joint* j = w->joints[joint];
  w->joints.erase(w->joints.begin() + joint);
  w->world->DestroyJoint(j);
  return B2G_OK;
}

int b2a_world_destroy_fixture(b2a_world* w, int32_t fixture) {
  if (!w || fixture < 0 || fixture >= (int32_t)w->fixtures.size()) return B2G_ERR_INVALID;
  b2Fixture* f = w->fixtures[fixture];
  w->fixtures.erase(w->fixtures.begin() + fixture);
  f->GetBody()->DestroyFixture(f);
  return B2G_OK;
}

int b2a_world_set_filter(b2a_world* w, int32_t fixture, int32_t category, int32_t mask, int32_t group) {
  if (!w || fixture < 0 || fixture >= (int32_t)w->fixtures.size()) return B2G_ERR_INVALID;
  b2Filter f;
  f.categoryBits = (uint16)category;
  f.maskBits = (uint16)mask;
  f.groupIndex = (int16)group;
  w->fixtures[fixture]->SetFilterData(f);
  return B2G_OK;
}

int b2a_world_set_allow_sleep(b2a_world* w, int32_t flag) {
  if (!w) return B2G_ERR_INVALID;
  w->world->SetAllowSleeping(flag != 0);
  return B2G_OK;
}

int b2a_world_body_count(b2a_world* w, int32_t* out) {
  if (!w || !out) return B2G_ERR_INVALID;
  *out = (int32_t)w->bodies.size();
  return B2G_OK;
}

int b2a_world_prepare(b2a_world* w) {
  if (!w) return B2G_ERR_INVALID;
  return w->world->Prepare();
}

}  // extern "C"
