// box2d.h -- drop-in C++ API of the MI355X engine.
//
// Class and member names, signatures and semantics follow the reference's
// public headers (include/box2d/b2_world.h, b2_body.h, b2_fixture.h,
// b2_shape.h, b2_polygon_shape.h, b2_circle_shape.h, b2_edge_shape.h,
// b2_revolute_joint.h, b2_contact.h, b2_collision.h, b2_world_callbacks.h,
// b2_math.h, b2_time_step.h) for the parts a b2World::Step user touches, so
// code written against the reference compiles against this header. Step()
// runs on the GPU through the b2g C ABI (include/b2g.h); body state lives in
// device memory between steps and host getters read a lazily refreshed copy.
//
// Device-path limits. Step() refuses to run, leaves the world unchanged,
// sets GetLastStatus() to B2G_ERR_UNSUPPORTED, prints the reason to stderr and
// fires b2Assert (compiled out under NDEBUG, as in the reference's Release
// build) while any of these is in effect: continuous physics (TOI, on by
// default in b2World), a disabled body, a sensor fixture, a contact filter
// that overrides ShouldCollide, or a contact listener that overrides
// PreSolve. Sleeping (on by default) runs on the device. Only revolute joints are
// created (CreateJoint of another type returns nullptr). Begin/EndContact and
// PostSolve are delivered after each Step from the device's record of it.
#ifndef B2AMD_BOX2D_H
#define B2AMD_BOX2D_H

#include <cassert>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "b2g.h"

#define b2Assert(A) assert(A)
#define B2_NOT_USED(x) ((void)(x))
#define b2_maxFloat FLT_MAX
#define b2_epsilon FLT_EPSILON
#define b2_pi 3.14159265359f
#define b2_maxManifoldPoints 2
#define b2_maxPolygonVertices 8
#define b2_linearSlop 0.005f
#define b2_angularSlop (2.0f / 180.0f * b2_pi)
#define b2_polygonRadius (2.0f * b2_linearSlop)

typedef signed char int8;
typedef signed short int16;
typedef signed int int32;
typedef unsigned char uint8;
typedef unsigned short uint16;
typedef unsigned int uint32;
typedef float float32;

// ---------------------------------------------------------------------------
// math (include/box2d/b2_math.h)
// ---------------------------------------------------------------------------
struct b2Vec2 {
  b2Vec2() {}
  b2Vec2(float xIn, float yIn) : x(xIn), y(yIn) {}
  void SetZero() { x = 0.0f; y = 0.0f; }
  void Set(float x_, float y_) { x = x_; y = y_; }
  b2Vec2 operator-() const { return b2Vec2(-x, -y); }
  void operator+=(const b2Vec2& v) { x += v.x; y += v.y; }
  void operator-=(const b2Vec2& v) { x -= v.x; y -= v.y; }
  void operator*=(float a) { x *= a; y *= a; }
  float Length() const { return sqrtf(x * x + y * y); }
  float LengthSquared() const { return x * x + y * y; }
  float Normalize() {
    float length = Length();
    if (length < b2_epsilon) return 0.0f;
    float inv = 1.0f / length;
    x *= inv;
    y *= inv;
    return length;
  }
  bool IsValid() const { return std::isfinite(x) && std::isfinite(y); }
  b2Vec2 Skew() const { return b2Vec2(-y, x); }
  float x, y;
};

inline b2Vec2 operator+(const b2Vec2& a, const b2Vec2& b) { return b2Vec2(a.x + b.x, a.y + b.y); }
inline b2Vec2 operator-(const b2Vec2& a, const b2Vec2& b) { return b2Vec2(a.x - b.x, a.y - b.y); }
inline b2Vec2 operator*(float s, const b2Vec2& a) { return b2Vec2(s * a.x, s * a.y); }
inline bool operator==(const b2Vec2& a, const b2Vec2& b) { return a.x == b.x && a.y == b.y; }
inline bool operator!=(const b2Vec2& a, const b2Vec2& b) { return a.x != b.x || a.y != b.y; }
inline float b2Dot(const b2Vec2& a, const b2Vec2& b) { return a.x * b.x + a.y * b.y; }
inline float b2Cross(const b2Vec2& a, const b2Vec2& b) { return a.x * b.y - a.y * b.x; }
inline b2Vec2 b2Cross(const b2Vec2& a, float s) { return b2Vec2(s * a.y, -s * a.x); }
inline b2Vec2 b2Cross(float s, const b2Vec2& a) { return b2Vec2(-s * a.y, s * a.x); }
inline float b2DistanceSquared(const b2Vec2& a, const b2Vec2& b) {
  b2Vec2 c = a - b;
  return b2Dot(c, c);
}
inline float b2Distance(const b2Vec2& a, const b2Vec2& b) { return (a - b).Length(); }
template <typename T>
inline T b2Abs(T a) { return a > T(0) ? a : -a; }
template <typename T>
inline T b2Min(T a, T b) { return a < b ? a : b; }
template <typename T>
inline T b2Max(T a, T b) { return a > b ? a : b; }
template <typename T>
inline T b2Clamp(T a, T low, T high) { return b2Max(low, b2Min(a, high)); }
inline b2Vec2 b2Min(const b2Vec2& a, const b2Vec2& b) { return b2Vec2(b2Min(a.x, b.x), b2Min(a.y, b.y)); }
inline b2Vec2 b2Max(const b2Vec2& a, const b2Vec2& b) { return b2Vec2(b2Max(a.x, b.x), b2Max(a.y, b.y)); }
extern const b2Vec2 b2Vec2_zero;

struct b2Rot {
  b2Rot() {}
  explicit b2Rot(float angle) { Set(angle); }
  void Set(float angle) { s = sinf(angle); c = cosf(angle); }
  void SetIdentity() { s = 0.0f; c = 1.0f; }
  float GetAngle() const { return atan2f(s, c); }
  b2Vec2 GetXAxis() const { return b2Vec2(c, s); }
  b2Vec2 GetYAxis() const { return b2Vec2(-s, c); }
  float s, c;
};
inline b2Vec2 b2Mul(const b2Rot& q, const b2Vec2& v) { return b2Vec2(q.c * v.x - q.s * v.y, q.s * v.x + q.c * v.y); }
inline b2Vec2 b2MulT(const b2Rot& q, const b2Vec2& v) { return b2Vec2(q.c * v.x + q.s * v.y, -q.s * v.x + q.c * v.y); }

struct b2Transform {
  b2Transform() {}
  b2Transform(const b2Vec2& position, const b2Rot& rotation) : p(position), q(rotation) {}
  void SetIdentity() { p.SetZero(); q.SetIdentity(); }
  void Set(const b2Vec2& position, float angle) { p = position; q.Set(angle); }
  b2Vec2 p;
  b2Rot q;
};
inline b2Vec2 b2Mul(const b2Transform& T, const b2Vec2& v) {
  float x = (T.q.c * v.x - T.q.s * v.y) + T.p.x;
  float y = (T.q.s * v.x + T.q.c * v.y) + T.p.y;
  return b2Vec2(x, y);
}
inline b2Vec2 b2MulT(const b2Transform& T, const b2Vec2& v) {
  float px = v.x - T.p.x;
  float py = v.y - T.p.y;
  return b2Vec2(T.q.c * px + T.q.s * py, -T.q.s * px + T.q.c * py);
}

struct b2Mat22 {
  b2Mat22() {}
  b2Mat22(const b2Vec2& c1, const b2Vec2& c2) : ex(c1), ey(c2) {}
  b2Vec2 ex, ey;
};

struct b2AABB {
  b2Vec2 lowerBound, upperBound;
  b2Vec2 GetCenter() const { return 0.5f * (lowerBound + upperBound); }
  b2Vec2 GetExtents() const { return 0.5f * (upperBound - lowerBound); }
};
inline bool b2TestOverlap(const b2AABB& a, const b2AABB& b) {
  return (a.upperBound.x >= b.lowerBound.x) & (a.lowerBound.x <= b.upperBound.x) &
         (a.upperBound.y >= b.lowerBound.y) & (a.lowerBound.y <= b.upperBound.y);
}

// ---------------------------------------------------------------------------
// time step / profile (include/box2d/b2_time_step.h)
// ---------------------------------------------------------------------------
struct b2Profile {
  float step, collide, solve, solveInit, solveVelocity, solvePosition, broadphase, solveTOI;
};

// ---------------------------------------------------------------------------
// collision (include/box2d/b2_collision.h)
// ---------------------------------------------------------------------------
struct b2ContactFeature {
  enum Type { e_vertex = 0, e_face = 1 };
  uint8 indexA, indexB, typeA, typeB;
};
union b2ContactID {
  b2ContactFeature cf;
  uint32 key;
};
struct b2ManifoldPoint {
  b2Vec2 localPoint;
  float normalImpulse;
  float tangentImpulse;
  b2ContactID id;
};
struct b2Manifold {
  enum Type { e_circles, e_faceA, e_faceB };
  b2ManifoldPoint points[b2_maxManifoldPoints];
  b2Vec2 localNormal;
  b2Vec2 localPoint;
  Type type;
  int32 pointCount;
};
struct b2WorldManifold {
  void Initialize(const b2Manifold* manifold, const b2Transform& xfA, float radiusA, const b2Transform& xfB,
                  float radiusB);
  b2Vec2 normal;
  b2Vec2 points[b2_maxManifoldPoints];
  float separations[b2_maxManifoldPoints];
};
struct b2MassData {
  float mass;
  b2Vec2 center;
  float I;
};

// ---------------------------------------------------------------------------
// shapes (include/box2d/b2_shape.h and the three shape headers)
// ---------------------------------------------------------------------------
class b2Shape {
 public:
  enum Type { e_circle = 0, e_edge = 1, e_polygon = 2, e_chain = 3, e_typeCount = 4 };
  virtual ~b2Shape() {}
  virtual b2Shape* Clone() const = 0;
  virtual void ComputeAABB(b2AABB* aabb, const b2Transform& xf) const = 0;
  virtual void ComputeMass(b2MassData* massData, float density) const = 0;
  Type GetType() const { return m_type; }
  Type m_type;
  float m_radius;
};

class b2CircleShape : public b2Shape {
 public:
  b2CircleShape() { m_type = e_circle; m_radius = 0.0f; m_p.SetZero(); }
  b2Shape* Clone() const override { return new b2CircleShape(*this); }
  void ComputeAABB(b2AABB* aabb, const b2Transform& xf) const override;
  void ComputeMass(b2MassData* massData, float density) const override;
  b2Vec2 m_p;
};

class b2EdgeShape : public b2Shape {
 public:
  b2EdgeShape() {
    m_type = e_edge;
    m_radius = b2_polygonRadius;
    m_vertex0.SetZero(); m_vertex1.SetZero(); m_vertex2.SetZero(); m_vertex3.SetZero();
    m_oneSided = false;
  }
  void SetOneSided(const b2Vec2& v0, const b2Vec2& v1, const b2Vec2& v2, const b2Vec2& v3);
  void SetTwoSided(const b2Vec2& v1, const b2Vec2& v2);
  b2Shape* Clone() const override { return new b2EdgeShape(*this); }
  void ComputeAABB(b2AABB* aabb, const b2Transform& xf) const override;
  void ComputeMass(b2MassData* massData, float density) const override;
  b2Vec2 m_vertex1, m_vertex2;
  b2Vec2 m_vertex0, m_vertex3;
  bool m_oneSided;
};

class b2PolygonShape : public b2Shape {
 public:
  b2PolygonShape() { m_type = e_polygon; m_radius = b2_polygonRadius; m_count = 0; m_centroid.SetZero(); }
  b2Shape* Clone() const override { return new b2PolygonShape(*this); }
  void Set(const b2Vec2* points, int32 count);
  void SetAsBox(float hx, float hy);
  void SetAsBox(float hx, float hy, const b2Vec2& center, float angle);
  void ComputeAABB(b2AABB* aabb, const b2Transform& xf) const override;
  void ComputeMass(b2MassData* massData, float density) const override;
  b2Vec2 m_centroid;
  b2Vec2 m_vertices[b2_maxPolygonVertices];
  b2Vec2 m_normals[b2_maxPolygonVertices];
  int32 m_count;
};

// ---------------------------------------------------------------------------
// fixtures, bodies, joints, contacts, world
// ---------------------------------------------------------------------------
class b2Body;
class b2Fixture;
class b2World;
class b2Contact;
class b2Joint;

struct b2BodyUserData { b2BodyUserData() { pointer = 0; } uintptr_t pointer; };
struct b2FixtureUserData { b2FixtureUserData() { pointer = 0; } uintptr_t pointer; };
struct b2JointUserData { b2JointUserData() { pointer = 0; } uintptr_t pointer; };

struct b2Filter {
  b2Filter() { categoryBits = 0x0001; maskBits = 0xFFFF; groupIndex = 0; }
  uint16 categoryBits;
  uint16 maskBits;
  int16 groupIndex;
};

struct b2FixtureDef {
  b2FixtureDef() {
    shape = nullptr; friction = 0.2f; restitution = 0.0f; restitutionThreshold = 1.0f;
    density = 0.0f; isSensor = false;
  }
  const b2Shape* shape;
  b2FixtureUserData userData;
  float friction;
  float restitution;
  float restitutionThreshold;
  float density;
  bool isSensor;
  b2Filter filter;
};

class b2Fixture {
 public:
  b2Shape::Type GetType() const { return m_shape->GetType(); }
  b2Shape* GetShape() { return m_shape; }
  const b2Shape* GetShape() const { return m_shape; }
  bool IsSensor() const { return m_isSensor; }
  const b2Filter& GetFilterData() const { return m_filter; }
  b2Body* GetBody() { return m_body; }
  const b2Body* GetBody() const { return m_body; }
  b2Fixture* GetNext() { return m_next; }
  b2FixtureUserData& GetUserData() { return m_userData; }
  float GetDensity() const { return m_density; }
  float GetFriction() const { return m_friction; }
  float GetRestitution() const { return m_restitution; }
  float GetRestitutionThreshold() const { return m_restitutionThreshold; }
  void GetMassData(b2MassData* massData) const { m_shape->ComputeMass(massData, m_density); }
  const b2AABB& GetAABB() const;
  // b2_fixture.cpp:142-183: contacts of this fixture are re-filtered at the
  // next Step's Collide (destroyed, with EndContact, if the pair no longer
  // collides); new pairs appear at the next broadphase
  void SetFilterData(const b2Filter& filter);
  void Refilter();
  // sensors do not run on the device: Step refuses while one exists
  void SetSensor(bool sensor);
  // the material of contacts created from now on (b2_fixture.h:189-216)
  void SetDensity(float density);
  void SetFriction(float friction);
  void SetRestitution(float restitution);
  void SetRestitutionThreshold(float threshold);

 private:
  friend class b2Body;
  friend class b2World;
  friend class b2Contact;
  void MarkShapeDirty();
  b2Fixture() {}
  ~b2Fixture() { delete m_shape; }
  b2Body* m_body = nullptr;
  b2Fixture* m_next = nullptr;
  b2Shape* m_shape = nullptr;
  float m_density = 0.0f, m_friction = 0.2f, m_restitution = 0.0f, m_restitutionThreshold = 1.0f;
  b2Filter m_filter;
  bool m_isSensor = false;
  b2FixtureUserData m_userData;
  int32 m_index = -1;  // creation index in the world (device shape slot)
  mutable b2AABB m_aabb;
};

enum b2BodyType { b2_staticBody = 0, b2_kinematicBody, b2_dynamicBody };

struct b2BodyDef {
  b2BodyDef() {
    position.Set(0.0f, 0.0f); angle = 0.0f; linearVelocity.Set(0.0f, 0.0f); angularVelocity = 0.0f;
    linearDamping = 0.0f; angularDamping = 0.0f; allowSleep = true; awake = true; fixedRotation = false;
    bullet = false; type = b2_staticBody; enabled = true; gravityScale = 1.0f;
  }
  b2BodyType type;
  b2Vec2 position;
  float angle;
  b2Vec2 linearVelocity;
  float angularVelocity;
  float linearDamping;
  float angularDamping;
  bool allowSleep, awake, fixedRotation, bullet, enabled;
  b2BodyUserData userData;
  float gravityScale;
};

class b2Body {
 public:
  b2Fixture* CreateFixture(const b2FixtureDef* def);
  b2Fixture* CreateFixture(const b2Shape* shape, float density);
  void SetTransform(const b2Vec2& position, float angle);
  const b2Transform& GetTransform() const;
  const b2Vec2& GetPosition() const;
  float GetAngle() const;
  const b2Vec2& GetWorldCenter() const;
  const b2Vec2& GetLocalCenter() const { return m_localCenter; }
  void SetLinearVelocity(const b2Vec2& v);
  const b2Vec2& GetLinearVelocity() const;
  void SetAngularVelocity(float omega);
  float GetAngularVelocity() const;
  void ApplyForce(const b2Vec2& force, const b2Vec2& point, bool wake);
  void ApplyForceToCenter(const b2Vec2& force, bool wake);
  void ApplyTorque(float torque, bool wake);
  void ApplyLinearImpulse(const b2Vec2& impulse, const b2Vec2& point, bool wake);
  void ApplyLinearImpulseToCenter(const b2Vec2& impulse, bool wake);
  void ApplyAngularImpulse(float impulse, bool wake);
  float GetMass() const { return m_mass; }
  float GetInertia() const { return m_I + m_mass * b2Dot(m_localCenter, m_localCenter); }
  void GetMassData(b2MassData* data) const;
  b2Vec2 GetWorldPoint(const b2Vec2& localPoint) const { return b2Mul(GetTransform(), localPoint); }
  b2Vec2 GetWorldVector(const b2Vec2& localVector) const { return b2Mul(GetTransform().q, localVector); }
  b2Vec2 GetLocalPoint(const b2Vec2& worldPoint) const { return b2MulT(GetTransform(), worldPoint); }
  b2Vec2 GetLocalVector(const b2Vec2& worldVector) const { return b2MulT(GetTransform().q, worldVector); }
  float GetLinearDamping() const { return m_linearDamping; }
  void SetLinearDamping(float d) { m_linearDamping = d; MarkDirty(); }
  float GetAngularDamping() const { return m_angularDamping; }
  void SetAngularDamping(float d) { m_angularDamping = d; MarkDirty(); }
  float GetGravityScale() const { return m_gravityScale; }
  void SetGravityScale(float s) { m_gravityScale = s; MarkDirty(); }
  b2BodyType GetType() const { return m_type; }
  // b2_body.cpp:129-226: destroys the body's contacts (EndContact for
  // touching ones), resets the mass, and a static body stops
  void SetType(b2BodyType type);
  // b2_body.cpp:280-350: destroys the fixture's contacts (EndContact for
  // touching ones) and resets the mass
  void DestroyFixture(b2Fixture* fixture);
  void SetMassData(const b2MassData* data);
  void ResetMassData();
  void SetBullet(bool flag) { m_bullet = flag; }
  bool IsBullet() const { return m_bullet; }
  // sleeping runs on the device (islands at rest fall asleep when the world
  // allows it); these read and set the device's flags
  void SetSleepingAllowed(bool flag);
  bool IsSleepingAllowed() const { return m_sleepingAllowed; }
  void SetAwake(bool flag);
  bool IsAwake() const;
  // disabled bodies are not run on the device: Step refuses while one exists
  void SetEnabled(bool flag);
  bool IsEnabled() const { return m_enabled; }
  void SetFixedRotation(bool flag);
  bool IsFixedRotation() const { return m_fixedRotation; }
  b2Fixture* GetFixtureList() { return m_fixtureList; }
  b2Body* GetNext() { return m_next; }
  b2BodyUserData& GetUserData() { return m_userData; }
  b2World* GetWorld() { return m_world; }
  int32 GetIndex() const { return m_index; }

 private:
  friend class b2World;
  friend class b2Fixture;
  friend class b2RevoluteJoint;
  b2Body(const b2BodyDef* bd, b2World* world, int32 index);
  ~b2Body();
  void MarkDirty();
  void Sync() const;
  void ToRecord(b2g_body* r) const;
  void FromRecord(const b2g_body& r);

  b2World* m_world;
  int32 m_index;
  b2BodyType m_type;
  b2Transform m_xf;
  b2Vec2 m_localCenter, m_c;
  float m_a;
  b2Vec2 m_linearVelocity;
  float m_angularVelocity;
  b2Vec2 m_force;
  float m_torque;
  float m_mass, m_invMass, m_I, m_invI;
  float m_linearDamping, m_angularDamping, m_gravityScale;
  bool m_fixedRotation;
  bool m_bullet = false, m_sleepingAllowed = true, m_awake = true, m_enabled = true;
  float m_sleepTime = 0.0f;
  b2Fixture* m_fixtureList = nullptr;
  int32 m_fixtureCount = 0;
  b2Body* m_prev = nullptr;
  b2Body* m_next = nullptr;
  b2BodyUserData m_userData;
};

enum b2JointType { e_unknownJoint, e_revoluteJoint };

struct b2JointDef {
  b2JointDef() { type = e_unknownJoint; bodyA = nullptr; bodyB = nullptr; collideConnected = false; }
  b2JointType type;
  b2JointUserData userData;
  b2Body* bodyA;
  b2Body* bodyB;
  bool collideConnected;
};

struct b2RevoluteJointDef : public b2JointDef {
  b2RevoluteJointDef() {
    type = e_revoluteJoint; localAnchorA.Set(0.0f, 0.0f); localAnchorB.Set(0.0f, 0.0f);
    referenceAngle = 0.0f; lowerAngle = 0.0f; upperAngle = 0.0f; maxMotorTorque = 0.0f; motorSpeed = 0.0f;
    enableLimit = false; enableMotor = false;
  }
  void Initialize(b2Body* bodyA, b2Body* bodyB, const b2Vec2& anchor);
  b2Vec2 localAnchorA, localAnchorB;
  float referenceAngle;
  bool enableLimit;
  float lowerAngle, upperAngle;
  bool enableMotor;
  float motorSpeed, maxMotorTorque;
};

class b2Joint {
 public:
  b2JointType GetType() const { return m_type; }
  b2Body* GetBodyA() { return m_bodyA; }
  b2Body* GetBodyB() { return m_bodyB; }
  virtual b2Vec2 GetReactionForce(float inv_dt) const = 0;
  virtual float GetReactionTorque(float inv_dt) const = 0;
  b2Joint* GetNext() { return m_next; }
  bool GetCollideConnected() const { return m_collideConnected; }
  virtual ~b2Joint() {}

 protected:
  friend class b2World;
  b2JointType m_type = e_unknownJoint;
  b2Body* m_bodyA = nullptr;
  b2Body* m_bodyB = nullptr;
  b2Joint* m_prev = nullptr;
  b2Joint* m_next = nullptr;
  bool m_collideConnected = false;
  int32 m_index = -1;
  b2World* m_world = nullptr;
};

class b2RevoluteJoint : public b2Joint {
 public:
  b2Vec2 GetReactionForce(float inv_dt) const override;
  float GetReactionTorque(float inv_dt) const override;
  const b2Vec2& GetLocalAnchorA() const { return m_localAnchorA; }
  const b2Vec2& GetLocalAnchorB() const { return m_localAnchorB; }
  float GetReferenceAngle() const { return m_referenceAngle; }
  float GetJointAngle() const;
  float GetMotorTorque(float inv_dt) const;

 private:
  friend class b2World;
  explicit b2RevoluteJoint(const b2RevoluteJointDef* def);
  void Sync() const;
  b2Vec2 m_localAnchorA, m_localAnchorB;
  float m_referenceAngle;
  bool m_enableLimit, m_enableMotor;
  float m_lowerAngle, m_upperAngle, m_motorSpeed, m_maxMotorTorque;
  mutable b2Vec2 m_impulse;
  mutable float m_motorImpulse, m_lowerImpulse, m_upperImpulse;
};

class b2Contact {
 public:
  b2Manifold* GetManifold() { return &m_manifold; }
  const b2Manifold* GetManifold() const { return &m_manifold; }
  void GetWorldManifold(b2WorldManifold* worldManifold) const;
  bool IsTouching() const { return m_touching; }
  bool IsEnabled() const { return true; }
  b2Contact* GetNext() { return m_next; }
  b2Fixture* GetFixtureA() { return m_fixtureA; }
  const b2Fixture* GetFixtureA() const { return m_fixtureA; }
  b2Fixture* GetFixtureB() { return m_fixtureB; }
  const b2Fixture* GetFixtureB() const { return m_fixtureB; }
  int32 GetChildIndexA() const { return 0; }
  int32 GetChildIndexB() const { return 0; }
  float GetFriction() const { return m_friction; }
  float GetRestitution() const { return m_restitution; }
  float GetRestitutionThreshold() const { return m_restitutionThreshold; }
  float GetTangentSpeed() const { return m_tangentSpeed; }

 private:
  friend class b2World;
  b2Fixture* m_fixtureA = nullptr;
  b2Fixture* m_fixtureB = nullptr;
  b2Manifold m_manifold;
  bool m_touching = false;
  float m_friction = 0.0f, m_restitution = 0.0f, m_restitutionThreshold = 0.0f, m_tangentSpeed = 0.0f;
  b2Contact* m_next = nullptr;
};

struct b2ContactImpulse {
  float normalImpulses[b2_maxManifoldPoints];
  float tangentImpulses[b2_maxManifoldPoints];
  int32 count;
};

// include/box2d/b2_world_callbacks.h:133-209. Begin/End and PostSolve are
// delivered after each Step from the device's record of it, in the
// reference's phase order (Collide's Begin/End, then PostSolve for every
// solved contact, then EndContact for touching contacts the broadphase
// destroyed); within a phase contacts come in shape-pair order. PreSolve
// cannot run on the device: a listener that overrides it makes Step refuse.
class b2ContactListener {
 public:
  virtual ~b2ContactListener() {}
  virtual void BeginContact(b2Contact* contact) { B2_NOT_USED(contact); }
  virtual void EndContact(b2Contact* contact) { B2_NOT_USED(contact); }
  // defined in the library (one definition, so the world can tell whether a
  // listener overrides them)
  virtual void PreSolve(b2Contact* contact, const b2Manifold* oldManifold);
  virtual void PostSolve(b2Contact* contact, const b2ContactImpulse* impulse);
};

// include/box2d/b2_world_callbacks.h:43-72: called before a joint or fixture
// is destroyed because its body is
class b2DestructionListener {
 public:
  virtual ~b2DestructionListener() {}
  virtual void SayGoodbye(b2Joint* joint) = 0;
  virtual void SayGoodbye(b2Fixture* fixture) = 0;
};

// include/box2d/b2_world_callbacks.h:78-111. The default ShouldCollide is the
// group/category/mask test (src/dynamics/b2_world_callbacks.cpp:28-40), which
// the device evaluates itself; a filter that overrides it cannot run there.
class b2ContactFilter {
 public:
  virtual ~b2ContactFilter() {}
  virtual bool ShouldCollide(b2Fixture* fixtureA, b2Fixture* fixtureB);
};

// include/box2d/b2_world_callbacks.h:213-264
class b2QueryCallback {
 public:
  virtual ~b2QueryCallback() {}
  virtual bool ReportFixture(b2Fixture* fixture) = 0;
};

class b2RayCastCallback {
 public:
  virtual ~b2RayCastCallback() {}
  // -1 ignore and continue, 0 terminate, fraction clip, 1 continue
  virtual float ReportFixture(b2Fixture* fixture, const b2Vec2& point, const b2Vec2& normal, float fraction) = 0;
};

class b2World {
 public:
  explicit b2World(const b2Vec2& gravity);
  ~b2World();
  b2World(const b2World&) = delete;
  b2World& operator=(const b2World&) = delete;

  void SetDestructionListener(b2DestructionListener* listener) { m_destructionListener = listener; }
  // a filter that overrides ShouldCollide makes Step refuse (see above)
  void SetContactFilter(b2ContactFilter* filter) { m_contactFilter = filter; }
  b2Body* CreateBody(const b2BodyDef* def);
  // b2_world.cpp:178-266: attached joints and fixtures go too (the
  // destruction listener hears of each), touching contacts end (EndContact)
  void DestroyBody(b2Body* body);
  b2Joint* CreateJoint(const b2JointDef* def);
  // b2_world.cpp:326-418: the bodies may collide again from the next
  // broadphase on
  void DestroyJoint(b2Joint* joint);
  // include/box2d/b2_world.h:120-142 (there are no particle systems here, so
  // particleIterations has nothing to step)
  void Step(float timeStep, int32 velocityIterations, int32 positionIterations, int32 particleIterations);
  void Step(float timeStep, int32 velocityIterations, int32 positionIterations) {
    Step(timeStep, velocityIterations, positionIterations, 1);
  }
  // fixtures whose AABB overlaps `aabb`, on the device tree (as in the
  // reference's broadphase, the callback's return value does not stop it)
  void QueryAABB(b2QueryCallback* callback, const b2AABB& aabb);
  // fixtures hit by the segment, reported nearest first; the callback's
  // return value clips (fraction), ignores (-1) or terminates (0)
  void RayCast(b2RayCastCallback* callback, const b2Vec2& point1, const b2Vec2& point2);
  void ClearForces() {}
  void SetContactListener(b2ContactListener* listener);
  b2Body* GetBodyList() { return m_bodyListHead; }
  b2Joint* GetJointList() { return m_jointList; }
  b2Contact* GetContactListStart();
  b2Contact* GetContactListEnd() { return nullptr; }
  // b2_world.cpp:474-489: turning sleeping off wakes every body
  void SetAllowSleeping(bool flag);
  bool GetAllowSleeping() const { return m_allowSleep; }
  void SetWarmStarting(bool flag) { m_warmStarting = flag; }
  bool GetWarmStarting() const { return m_warmStarting; }
  void SetContinuousPhysics(bool flag) { m_continuousPhysics = flag; }
  bool GetContinuousPhysics() const { return m_continuousPhysics; }
  void SetSubStepping(bool flag) { B2_NOT_USED(flag); }
  int32 GetBodyCount() const { return (int32)m_bodies.size(); }
  int32 GetJointCount() const { return (int32)m_joints.size(); }
  int32 GetContactCount() const;
  int32 GetProxyCount() const { return (int32)m_fixtures.size(); }
  void SetGravity(const b2Vec2& gravity) { m_gravity = gravity; }
  b2Vec2 GetGravity() const { return m_gravity; }
  bool IsLocked() const { return m_locked; }
  void SetAutoClearForces(bool flag) { B2_NOT_USED(flag); }
  bool GetAutoClearForces() const { return true; }
  // step and collide/solve/broadphase times; the stage split is measured on
  // the device only while profiling is on (events and a host wait per stage)
  const b2Profile& GetProfile() const { return m_profile; }
  void SetProfiling(bool flag) { m_profiling = flag; }
  // device handle (batch of one world)
  b2g_batch* GetDeviceBatch() { return m_batch; }
  // status of the last Step or device call (0 = ok) and its message
  int32 GetLastStatus() const { return m_status; }
  const char* GetLastErrorMessage() const { return m_statusMessage.c_str(); }
  // allocate device memory and upload everything created so far
  int Prepare();
  // device records of everything created so far, as this world would upload
  // them: bodies tagged with `world`, shape and joint body indices shifted by
  // `body_offset` (arrays sized GetBodyCount / GetProxyCount / GetJointCount)
  int ExportRecords(int32_t world, int32_t body_offset, b2g_body* bodies, b2g_shape* shapes,
                    b2g_revolute* joints) const;

 private:
  friend class b2Body;
  friend class b2RevoluteJoint;
  friend class b2Fixture;
  int EnsureDevice();
  int Upload();
  static void FixtureToRecord(const b2Fixture* f, b2g_shape* out);
  static void JointToRecord(const b2Joint* j, b2g_revolute* out);
  void SyncBodies() const;
  void SyncContacts();
  void ContactFromRecord(const b2g_contact& r, b2Contact* c) const;
  void SyncJoints() const;
  // destroying bodies, fixtures or joints, or re-filtering, edits the
  // contact list and renumbers records: the device state comes to the host
  // first and goes back, renumbered, before the next device call
  void BeginEdit();
  void RemoveContactsIf(bool (*pred)(const b2World*, const b2g_contact&, const void*), const void* arg);
  void RemoveFixtureIndex(int32 index);
  void ApplyRefilter();
  const char* UnsupportedReason() const;
  void DeliverEvents();

  b2Vec2 m_gravity;
  std::vector<b2Body*> m_bodies;        // creation order = device body index
  std::vector<b2Fixture*> m_fixtures;   // creation order = device shape index
  std::vector<b2Joint*> m_joints;
  b2Body* m_bodyListHead = nullptr;
  b2Body* m_bodyListTail = nullptr;
  b2Joint* m_jointList = nullptr;
  b2ContactListener* m_listener = nullptr;
  b2DestructionListener* m_destructionListener = nullptr;
  b2ContactFilter* m_contactFilter = nullptr;
  // host copy of the device contact list (with colours) while records are
  // being destroyed or re-filtered
  bool m_editing = false;
  std::vector<b2g_contact> m_editContacts;
  std::vector<int32_t> m_editColours;
  std::vector<char> m_refilter;  // per fixture: re-filter its contacts at the next Step
  bool m_anyRefilter = false;
  bool m_shapesDirty = false;
  std::string m_statusMessage;
  std::vector<b2Contact> m_contacts;
  bool m_contactsValid = true;
  b2g_batch* m_batch = nullptr;
  b2g_capacity m_cap{};
  int m_uploadedBodies = 0, m_uploadedShapes = 0, m_uploadedJoints = 0;
  std::vector<char> m_dirty;
  bool m_anyDirty = false;
  mutable bool m_bodiesValid = true;
  mutable bool m_jointsValid = true;
  bool m_newContacts = false;
  bool m_warmStarting = true, m_allowSleep = true, m_continuousPhysics = true, m_locked = false;
  bool m_profiling = false;
  // what Step must refuse (kept up to date by the mutators)
  int32 m_disabledBodies = 0, m_sensorFixtures = 0;
  bool m_listenerPreSolve = false, m_listenerPostSolve = false;
  std::string m_lastRefusal;
  float m_inv_dt0 = 0.0f;
  b2Profile m_profile;
  int32 m_status = 0;
};

#endif
