// ref_driver.cpp -- C entry points over the compiled reference (Box2D-optimized).
//
// Test infrastructure only. Built by oracle/Makefile together with the
// reference's own sources (read in place under the reference tree) into
// oracle/_ref/libbox2d_ref.so. It builds reference b2World objects from a
// b2g_scene through the reference's public API (b2World::CreateBody,
// b2Body::CreateFixture, b2World::CreateJoint), steps them and reads state
// back, so tests can use the reference itself as the parity anchor and
// bench.py can time the reference Step on the host.
#include "box2d/box2d.h"
#include "b2g.h"

#include <chrono>
#include <cstring>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

struct RefWorld {
  b2World world;
  std::vector<b2Body*> bodies;
  std::vector<b2Fixture*> fixtures;
  std::vector<b2RevoluteJoint*> joints;
  // Begin (1) / End (2) contact events of the last ref_world_step_ordered, in
  // the order the reference called its listener, keyed by fixture indices
  std::vector<int> ev_kind;
  std::vector<uint64_t> ev_key;
  int ev_collide = 0;  // leading events made by Collide; the rest by the broadphase's destroys
  // PostSolve calls of that step: contact key and b2ContactImpulse::count
  std::vector<uint64_t> ps_key;
  std::vector<int> ps_count;
  explicit RefWorld(const b2Vec2& g) : world(g) {}
};

// Builds the reference shape for fixture i of a scene (the same calls a user makes).
struct AnyShape {
  b2CircleShape circle;
  b2EdgeShape edge;
  b2PolygonShape polygon;
  b2Shape* get(int kind) {
    if (kind == B2G_CIRCLE) return &circle;
    if (kind == B2G_EDGE) return &edge;
    return &polygon;
  }
};

int make_shape(const b2g_scene* s, int i, AnyShape* out) {
  const int kind = s->fix_kind[i];
  const float* g = s->fix_geom + 17 * i;
  switch (kind) {
    case 0:
      out->circle.m_radius = g[0];
      out->circle.m_p.Set(g[1], g[2]);
      return B2G_CIRCLE;
    case 1:
      if (g[0] != 0.0f) {
        out->edge.SetOneSided(b2Vec2(g[1], g[2]), b2Vec2(g[3], g[4]), b2Vec2(g[5], g[6]),
                              b2Vec2(g[7], g[8]));
      } else {
        out->edge.SetTwoSided(b2Vec2(g[3], g[4]), b2Vec2(g[5], g[6]));
      }
      return B2G_EDGE;
    case 2: {
      int n = (int)g[0];
      b2Vec2 pts[b2_maxPolygonVertices];
      for (int k = 0; k < n; ++k) pts[k].Set(g[1 + 2 * k], g[2 + 2 * k]);
      out->polygon.Set(pts, n);
      return B2G_POLYGON;
    }
    default:
      if (g[5] != 0.0f) {
        out->polygon.SetAsBox(g[0], g[1], b2Vec2(g[2], g[3]), g[4]);
      } else {
        out->polygon.SetAsBox(g[0], g[1]);
      }
      return B2G_POLYGON;
  }
}

void shape_to_record(const b2Shape* sh, b2g_shape* r) {
  std::memset(r, 0, sizeof(*r));
  r->type = (int)sh->GetType();
  r->radius = sh->m_radius;
  if (sh->GetType() == b2Shape::e_circle) {
    const b2CircleShape* c = (const b2CircleShape*)sh;
    r->centroid[0] = c->m_p.x;
    r->centroid[1] = c->m_p.y;
  } else if (sh->GetType() == b2Shape::e_edge) {
    const b2EdgeShape* e = (const b2EdgeShape*)sh;
    r->count = e->m_oneSided ? 1 : 0;
    const b2Vec2 v[4] = {e->m_vertex0, e->m_vertex1, e->m_vertex2, e->m_vertex3};
    for (int k = 0; k < 4; ++k) {
      r->vertices[2 * k] = v[k].x;
      r->vertices[2 * k + 1] = v[k].y;
    }
  } else {
    const b2PolygonShape* p = (const b2PolygonShape*)sh;
    r->count = p->m_count;
    r->centroid[0] = p->m_centroid.x;
    r->centroid[1] = p->m_centroid.y;
    for (int k = 0; k < p->m_count; ++k) {
      r->vertices[2 * k] = p->m_vertices[k].x;
      r->vertices[2 * k + 1] = p->m_vertices[k].y;
      r->normals[2 * k] = p->m_normals[k].x;
      r->normals[2 * k + 1] = p->m_normals[k].y;
    }
  }
}

// Reference shape objects rebuilt from a processed record (fields are public).
void record_to_shape(const b2g_shape* r, AnyShape* out) {
  if (r->type == B2G_CIRCLE) {
    out->circle.m_radius = r->radius;
    out->circle.m_p.Set(r->centroid[0], r->centroid[1]);
  } else if (r->type == B2G_EDGE) {
    out->edge.m_radius = r->radius;
    out->edge.m_vertex0.Set(r->vertices[0], r->vertices[1]);
    out->edge.m_vertex1.Set(r->vertices[2], r->vertices[3]);
    out->edge.m_vertex2.Set(r->vertices[4], r->vertices[5]);
    out->edge.m_vertex3.Set(r->vertices[6], r->vertices[7]);
    out->edge.m_oneSided = r->count != 0;
  } else {
    out->polygon.m_radius = r->radius;
    out->polygon.m_count = r->count;
    out->polygon.m_centroid.Set(r->centroid[0], r->centroid[1]);
    for (int k = 0; k < r->count; ++k) {
      out->polygon.m_vertices[k].Set(r->vertices[2 * k], r->vertices[2 * k + 1]);
      out->polygon.m_normals[k].Set(r->normals[2 * k], r->normals[2 * k + 1]);
    }
  }
}

b2Transform to_xf(const float* x) {
  b2Transform t;
  t.p.Set(x[0], x[1]);
  t.q.s = x[2];
  t.q.c = x[3];
  return t;
}

// Without points the manifold's type, normal and point carry no meaning: the
// reference leaves them unset in a new contact (b2Contact::b2Contact sets only
// pointCount, src/dynamics/b2_contact.cpp:118) and the collide functions
// return before writing them when they find nothing, so they are recorded as 0.
void manifold_to_record(const b2Manifold& m, b2g_contact* out) {
  const bool has = m.pointCount > 0;
  out->type = has ? (int)m.type : 0;
  out->point_count = m.pointCount;
  out->local_normal[0] = has ? m.localNormal.x : 0.0f;
  out->local_normal[1] = has ? m.localNormal.y : 0.0f;
  out->local_point[0] = has ? m.localPoint.x : 0.0f;
  out->local_point[1] = has ? m.localPoint.y : 0.0f;
  for (int k = 0; k < 2; ++k) {
    if (k < m.pointCount) {
      out->point_local[2 * k] = m.points[k].localPoint.x;
      out->point_local[2 * k + 1] = m.points[k].localPoint.y;
      out->normal_impulse[k] = m.points[k].normalImpulse;
      out->tangent_impulse[k] = m.points[k].tangentImpulse;
      out->id[k] = m.points[k].id.key;
    } else {
      out->point_local[2 * k] = out->point_local[2 * k + 1] = 0.0f;
      out->normal_impulse[k] = out->tangent_impulse[k] = 0.0f;
      out->id[k] = 0;
    }
  }
}

}  // namespace

extern "C" {

void* ref_world_from_scene(const b2g_scene* s, int warm_starting, int allow_sleep,
                           int continuous) {
  RefWorld* rw = new RefWorld(b2Vec2(s->gravity[0], s->gravity[1]));
  rw->world.SetContinuousPhysics(continuous != 0);
  rw->world.SetAllowSleeping(allow_sleep != 0);
  rw->world.SetWarmStarting(warm_starting != 0);
  rw->bodies.resize(s->body_count);
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
    bd.userData.pointer = (uintptr_t)i;
    rw->bodies[i] = rw->world.CreateBody(&bd);
  }
  rw->fixtures.resize(s->fixture_count);
  for (int i = 0; i < s->fixture_count; ++i) {
    AnyShape sh;
    int t = make_shape(s, i, &sh);
    b2FixtureDef fd;
    fd.shape = sh.get(t);
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
    fd.userData.pointer = (uintptr_t)i;
    rw->fixtures[i] = rw->bodies[s->fix_body[i]]->CreateFixture(&fd);
  }
  rw->joints.resize(s->joint_count);
  for (int i = 0; i < s->joint_count; ++i) {
    b2RevoluteJointDef jd;
    const float* p = s->joint_params + 6 * i;
    jd.Initialize(rw->bodies[s->joint_bodies[2 * i]], rw->bodies[s->joint_bodies[2 * i + 1]],
                  b2Vec2(p[0], p[1]));
    const int f = s->joint_flags[i];
    jd.collideConnected = (f & 1) != 0;
    jd.enableLimit = (f & 2) != 0;
    jd.enableMotor = (f & 4) != 0;
    jd.lowerAngle = p[2];
    jd.upperAngle = p[3];
    jd.motorSpeed = p[4];
    jd.maxMotorTorque = p[5];
    rw->joints[i] = (b2RevoluteJoint*)rw->world.CreateJoint(&jd);
  }
  return rw;
}

void ref_world_destroy(void* w) { delete (RefWorld*)w; }

// Steps n times; ms_each (nullable) receives the wall time of every Step.
void ref_world_step(void* w, float dt, int vi, int pi, int n, double* ms_each) {
  RefWorld* rw = (RefWorld*)w;
  for (int i = 0; i < n; ++i) {
    auto t0 = std::chrono::steady_clock::now();
    rw->world.Step(dt, vi, pi);
    auto t1 = std::chrono::steady_clock::now();
    if (ms_each) ms_each[i] = std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
}

// 10 floats per body: sweep c.x c.y a, v.x v.y w, xf p.x p.y q.s q.c
void ref_world_get_bodies(void* w, float* out) {
  RefWorld* rw = (RefWorld*)w;
  for (size_t i = 0; i < rw->bodies.size(); ++i) {
    b2Body* b = rw->bodies[i];
    const b2Vec2 c = b->GetWorldCenter();
    const b2Vec2 v = b->GetLinearVelocity();
    const b2Transform& xf = b->GetTransform();
    float* o = out + 10 * i;
    o[0] = c.x; o[1] = c.y; o[2] = b->GetAngle();
    o[3] = v.x; o[4] = v.y; o[5] = b->GetAngularVelocity();
    o[6] = xf.p.x; o[7] = xf.p.y; o[8] = xf.q.s; o[9] = xf.q.c;
  }
}

// b2Body::IsAwake per body (the sleep timer itself is private to b2Body)
void ref_world_get_awake(void* w, int* awake) {
  RefWorld* rw = (RefWorld*)w;
  for (size_t i = 0; i < rw->bodies.size(); ++i) awake[i] = rw->bodies[i]->IsAwake() ? 1 : 0;
}

// 4 floats per body: mass, rotational inertia about the centre, localCenter
void ref_world_get_mass(void* w, float* out) {
  RefWorld* rw = (RefWorld*)w;
  for (size_t i = 0; i < rw->bodies.size(); ++i) {
    b2MassData md;
    rw->bodies[i]->GetMassData(&md);
    const b2Vec2 lc = rw->bodies[i]->GetLocalCenter();
    out[4 * i + 0] = md.mass;
    out[4 * i + 1] = md.I - md.mass * b2Dot(lc, lc);
    out[4 * i + 2] = lc.x;
    out[4 * i + 3] = lc.y;
  }
}

int ref_world_contact_count(void* w) { return ((RefWorld*)w)->world.GetContactCount(); }

int ref_world_get_contacts(void* w, int cap, b2g_contact* out) {
  RefWorld* rw = (RefWorld*)w;
  int n = 0;
  for (b2Contact* c = rw->world.GetContactListStart(); c != rw->world.GetContactListEnd();
       c = c->GetNext()) {
    if (n < cap) {
      b2g_contact* o = out + n;
      std::memset(o, 0, sizeof(*o));
      o->shape_a = (int)c->GetFixtureA()->GetUserData().pointer;
      o->shape_b = (int)c->GetFixtureB()->GetUserData().pointer;
      o->touching = c->IsTouching() ? 1 : 0;
      manifold_to_record(*c->GetManifold(), o);
      o->friction = c->GetFriction();
      o->restitution = c->GetRestitution();
      o->restitution_threshold = c->GetRestitutionThreshold();
      o->tangent_speed = c->GetTangentSpeed();
    }
    ++n;
  }
  return n;
}

// 3 floats per joint: linear impulse x, y and the summed axial impulse
void ref_world_get_joint_impulses(void* w, float* out) {
  RefWorld* rw = (RefWorld*)w;
  for (size_t i = 0; i < rw->joints.size(); ++i) {
    const b2Vec2 f = rw->joints[i]->GetReactionForce(1.0f);
    out[3 * i + 0] = f.x;
    out[3 * i + 1] = f.y;
    out[3 * i + 2] = rw->joints[i]->GetReactionTorque(1.0f);
  }
}

// Processed shape of scene fixture i, as b2PolygonShape::Set / SetAsBox leave it.
void ref_shape_from_scene(const b2g_scene* s, int i, b2g_shape* out) {
  AnyShape sh;
  int t = make_shape(s, i, &sh);
  shape_to_record(sh.get(t), out);
}

// Narrowphase of one pair through the reference's b2Collide* functions
// (dispatch as b2Contact::Create orders it: src/dynamics/b2_contact.cpp:48-77).
void ref_collide(const b2g_shape* a, const float* xfa, const b2g_shape* b, const float* xfb,
                 b2g_contact* out) {
  AnyShape sa, sb;
  record_to_shape(a, &sa);
  record_to_shape(b, &sb);
  b2Manifold m;
  std::memset(&m, 0, sizeof(m));
  const b2Transform ta = to_xf(xfa), tb = to_xf(xfb);
  std::memset(out, 0, sizeof(*out));
  const int A = a->type, B = b->type;
  if (A == B2G_CIRCLE && B == B2G_CIRCLE) {
    b2CollideCircles(&m, &sa.circle, ta, &sb.circle, tb);
  } else if (A == B2G_POLYGON && B == B2G_CIRCLE) {
    b2CollidePolygonAndCircle(&m, &sa.polygon, ta, &sb.circle, tb);
  } else if (A == B2G_POLYGON && B == B2G_POLYGON) {
    b2CollidePolygons(&m, &sa.polygon, ta, &sb.polygon, tb);
  } else if (A == B2G_EDGE && B == B2G_CIRCLE) {
    b2CollideEdgeAndCircle(&m, &sa.edge, ta, &sb.circle, tb);
  } else if (A == B2G_EDGE && B == B2G_POLYGON) {
    b2CollideEdgeAndPolygon(&m, &sa.edge, ta, &sb.polygon, tb);
  } else {
    out->point_count = -1;  // no collide function for this ordering
    return;
  }
  manifold_to_record(m, out);
  out->touching = m.pointCount > 0;
}

// b2WorldManifold::Initialize; out: normal(2), points(4), separations(2)
void ref_world_manifold(const b2g_contact* c, const float* xfa, float ra, const float* xfb,
                        float rb, float* out) {
  b2Manifold m;
  std::memset(&m, 0, sizeof(m));
  m.type = (b2Manifold::Type)c->type;
  m.pointCount = c->point_count;
  m.localNormal.Set(c->local_normal[0], c->local_normal[1]);
  m.localPoint.Set(c->local_point[0], c->local_point[1]);
  for (int k = 0; k < 2; ++k) m.points[k].localPoint.Set(c->point_local[2 * k], c->point_local[2 * k + 1]);
  b2WorldManifold wm;
  std::memset(&wm, 0, sizeof(wm));
  wm.Initialize(&m, to_xf(xfa), ra, to_xf(xfb), rb);
  out[0] = wm.normal.x; out[1] = wm.normal.y;
  out[2] = wm.points[0].x; out[3] = wm.points[0].y;
  out[4] = wm.points[1].x; out[5] = wm.points[1].y;
  out[6] = wm.separations[0]; out[7] = wm.separations[1];
}

// b2World::GetProfile of the last step (ms): step, collide, solve, broadphase
void ref_world_profile(void* w, float* out) {
  const b2Profile& p = ((RefWorld*)w)->world.GetProfile();
  out[0] = p.step; out[1] = p.collide; out[2] = p.solve; out[3] = p.broadphase;
}

}  // extern "C"

// One reference Step that also reports the order in which the reference's
// island solver visits constraints: b2World::Solve's DFS (src/dynamics/
// b2_world.cpp:526-659) replayed over the body list, the per-body contact
// arrays and joint lists as they stand before the step. The touching set at
// the solve is what Collide left: the contacts it updated and found touching
// (b2ContactListener::PreSolve, called for every touching non-sensor contact
// it updates, src/dynamics/b2_contact.cpp:206-208) and the contacts it did
// not update (both bodies asleep or static, b2_contact_manager.cpp:100-111)
// that were touching before the step. Islands are seeded from the bodies
// awake at the solve: awake before the step or woken by one of Collide's
// touching transitions.
// out_keys receives (min << 32 | max) fixture keys of the solved contacts in
// solve order; out_joints the scene indices of the solved joints in order.
namespace {
uint64_t fixture_key(b2Contact* c) {
  const uint64_t fa = c->GetFixtureA()->GetUserData().pointer;
  const uint64_t fb = c->GetFixtureB()->GetUserData().pointer;
  return ((fa < fb ? fa : fb) << 32) | (fa < fb ? fb : fa);
}

// b2ContactListener calls of one Step: BeginContact / EndContact as the
// reference makes them (Collide's touching transitions, src/dynamics/
// b2_contact.cpp:197-203; EndContact of a touching contact destroyed by the
// broadphase, b2_contact_manager.cpp:48-51), PreSolve's touching set and
// PostSolve's impulse counts (src/dynamics/b2_island.cpp:621-647). Solve
// reports every island before it looks for new contacts
// (src/dynamics/b2_world.cpp:640-666), so the events before the first
// PostSolve are Collide's.
struct TouchRecorder : public b2ContactListener {
  std::unordered_set<const b2Contact*> touching;
  // Collide's touching transitions: the contacts, and the bodies they woke
  // (b2_contact.cpp:183-188)
  std::unordered_set<const b2Contact*> changed;
  std::unordered_set<const b2Body*> woken;
  RefWorld* rw = nullptr;
  bool solved = false;
  void transition(b2Contact* c) {
    changed.insert(c);
    woken.insert(c->GetFixtureA()->GetBody());
    woken.insert(c->GetFixtureB()->GetBody());
  }
  void BeginContact(b2Contact* c) override {
    rw->ev_kind.push_back(1);
    rw->ev_key.push_back(fixture_key(c));
    if (!solved) rw->ev_collide = (int)rw->ev_kind.size();
    transition(c);
  }
  void EndContact(b2Contact* c) override {
    rw->ev_kind.push_back(2);
    rw->ev_key.push_back(fixture_key(c));
    if (!solved) rw->ev_collide = (int)rw->ev_kind.size();
    // Collide clears the touching flag before it calls EndContact; the
    // broadphase's destroy calls it on a contact still flagged touching
    // (b2_contact_manager.cpp:48-51)
    if (!c->IsTouching()) transition(c);
  }
  void PreSolve(b2Contact* c, const b2Manifold*) override { touching.insert(c); }
  void PostSolve(b2Contact* c, const b2ContactImpulse* impulse) override {
    solved = true;
    rw->ps_key.push_back(fixture_key(c));
    rw->ps_count.push_back(impulse->count);
  }
};
}  // namespace

extern "C" int ref_world_step_ordered(void* w, float dt, int vi, int pi, int cap,
                                      uint64_t* out_keys, int* out_nc, int* out_joints,
                                      int* out_nj) {
  RefWorld* rw = (RefWorld*)w;
  struct CSnap {
    uint64_t key;
    b2Body* a;
    b2Body* b;
  };
  std::unordered_map<const b2Contact*, CSnap> csnap;
  std::vector<b2Body*> order;
  std::unordered_map<b2Body*, std::vector<const b2Contact*>> bcontacts;
  std::unordered_map<b2Body*, std::vector<std::pair<const b2Joint*, b2Body*>>> bjoints;
  std::unordered_map<const b2Joint*, int> jindex;
  std::unordered_set<const b2Contact*> touch0;
  std::unordered_set<const b2Body*> awake0;
  for (size_t i = 0; i < rw->joints.size(); ++i) jindex[rw->joints[i]] = (int)i;
  for (b2Contact* c = rw->world.GetContactListStart(); c != rw->world.GetContactListEnd(); c = c->GetNext())
    if (c->IsTouching() && !c->GetFixtureA()->IsSensor() && !c->GetFixtureB()->IsSensor()) touch0.insert(c);
  for (b2Body* b = rw->world.GetBodyList(); b; b = b->GetNext()) {
    if (b->IsAwake()) awake0.insert(b);
    order.push_back(b);
    std::vector<const b2Contact*>& v = bcontacts[b];
    for (int i = 0; i < b->GetContactCount(); ++i) {
      b2Contact* c = b->GetContact(i);
      v.push_back(c);
      if (!csnap.count(c)) {
        uint64_t fa = c->GetFixtureA()->GetUserData().pointer;
        uint64_t fb = c->GetFixtureB()->GetUserData().pointer;
        uint64_t lo = fa < fb ? fa : fb, hi = fa < fb ? fb : fa;
        csnap[c] = CSnap{(lo << 32) | hi, c->GetFixtureA()->GetBody(), c->GetFixtureB()->GetBody()};
      }
    }
    for (b2JointEdge* je = b->GetJointList(); je; je = je->next)
      bjoints[b].push_back({je->joint, je->other});
  }
  TouchRecorder rec;
  rw->ev_kind.clear();
  rw->ev_key.clear();
  rw->ev_collide = 0;
  rw->ps_key.clear();
  rw->ps_count.clear();
  rec.rw = rw;
  rw->world.SetContactListener(&rec);
  rw->world.Step(dt, vi, pi);
  rw->world.SetContactListener(nullptr);

  std::unordered_set<const b2Body*> bflag;
  std::unordered_set<const b2Contact*> cflag;
  std::unordered_set<const b2Joint*> jflag;
  int nc = 0, nj = 0;
  std::vector<b2Body*> stack;
  auto touching_at_solve = [&](const b2Contact* c) {
    return rec.touching.count(c) != 0 || (touch0.count(c) != 0 && rec.changed.count(c) == 0);
  };
  for (b2Body* seed : order) {
    if (bflag.count(seed)) continue;
    if (!awake0.count(seed) && !rec.woken.count(seed)) continue;
    if (seed->GetType() == b2_staticBody) break;
    if (bcontacts[seed].empty() && bjoints[seed].empty()) {
      bflag.insert(seed);
      continue;
    }
    std::vector<b2Body*> island;
    stack.clear();
    stack.push_back(seed);
    bflag.insert(seed);
    while (!stack.empty()) {
      b2Body* b = stack.back();
      stack.pop_back();
      island.push_back(b);
      if (b->GetType() == b2_staticBody) continue;
      for (const b2Contact* c : bcontacts[b]) {
        if (cflag.count(c)) continue;
        if (!touching_at_solve(c)) continue;
        const CSnap& cs = csnap[c];
        if (nc < cap) out_keys[nc] = cs.key;
        ++nc;
        cflag.insert(c);
        b2Body* other = cs.a == b ? cs.b : cs.a;
        if (bflag.count(other)) continue;
        stack.push_back(other);
        bflag.insert(other);
      }
      for (auto& je : bjoints[b]) {
        if (jflag.count(je.first)) continue;
        if (out_joints) out_joints[nj] = jindex[je.first];
        ++nj;
        jflag.insert(je.first);
        if (bflag.count(je.second)) continue;
        stack.push_back(je.second);
        bflag.insert(je.second);
      }
    }
    for (b2Body* b : island)
      if (b->GetType() == b2_staticBody) bflag.erase(b);
  }
  *out_nc = nc;
  if (out_nj) *out_nj = nj;
  return nc <= cap ? 0 : 1;
}

// The contact events of the last ref_world_step_ordered: kinds (1 Begin,
// 2 End) and fixture keys, up to cap, and how many of them Collide made
// (the rest came from the broadphase's destroys); returns the count.
extern "C" int ref_world_last_events(void* w, int cap, int* kinds, uint64_t* keys, int* n_collide) {
  RefWorld* rw = (RefWorld*)w;
  const int n = (int)rw->ev_kind.size();
  for (int i = 0; i < n && i < cap; ++i) {
    kinds[i] = rw->ev_kind[i];
    keys[i] = rw->ev_key[i];
  }
  if (n_collide) *n_collide = rw->ev_collide;
  return n;
}

// PostSolve calls of the last ref_world_step_ordered: contact keys and
// impulse counts, up to cap; returns the count.
extern "C" int ref_world_last_postsolve(void* w, int cap, uint64_t* keys, int* counts) {
  RefWorld* rw = (RefWorld*)w;
  const int n = (int)rw->ps_key.size();
  for (int i = 0; i < n && i < cap; ++i) {
    keys[i] = rw->ps_key[i];
    counts[i] = rw->ps_count[i];
  }
  return n;
}

// 4 floats per fixture: the AABB the reference broadphase last used
// (b2Fixture::GetAABB, refreshed by b2BroadPhase::UpdateAndQuery for non-static
// fixtures; static ones keep the AABB computed at creation).
extern "C" void ref_world_get_aabbs(void* w, float* out) {
  RefWorld* rw = (RefWorld*)w;
  for (size_t i = 0; i < rw->fixtures.size(); ++i) {
    const b2AABB& a = rw->fixtures[i]->GetAABB();
    out[4 * i + 0] = a.lowerBound.x;
    out[4 * i + 1] = a.lowerBound.y;
    out[4 * i + 2] = a.upperBound.x;
    out[4 * i + 3] = a.upperBound.y;
  }
}

// ---- user queries through the reference's b2World (QueryAABB / RayCast) ----
namespace {
struct QueryCollect : public b2QueryCallback {
  std::unordered_map<const b2Fixture*, int>* index;
  std::vector<int> found;
  bool ReportFixture(b2Fixture* f) override {
    found.push_back((*index)[f]);
    return true;
  }
};

// every fixture the ray hits: return 1 keeps the segment unclipped
struct RayCollectAll : public b2RayCastCallback {
  std::unordered_map<const b2Fixture*, int>* index;
  std::vector<float> hits;  // index, fraction, normal.x, normal.y, point.x, point.y
  float ReportFixture(b2Fixture* f, const b2Vec2& p, const b2Vec2& n, float fraction) override {
    hits.insert(hits.end(), {(float)(*index)[f], fraction, n.x, n.y, p.x, p.y});
    return 1.0f;
  }
};

std::unordered_map<const b2Fixture*, int> fixture_index(RefWorld* rw) {
  std::unordered_map<const b2Fixture*, int> m;
  for (size_t i = 0; i < rw->fixtures.size(); ++i) m[rw->fixtures[i]] = (int)i;
  return m;
}
}  // namespace

// fixture indices whose AABB overlaps the box; returns the count (<= cap written)
extern "C" int ref_world_query_aabb(void* w, const float* box, int cap, int* out) {
  RefWorld* rw = (RefWorld*)w;
  auto idx = fixture_index(rw);
  QueryCollect cb;
  cb.index = &idx;
  b2AABB a;
  a.lowerBound.Set(box[0], box[1]);
  a.upperBound.Set(box[2], box[3]);
  rw->world.QueryAABB(&cb, a);
  for (size_t i = 0; i < cb.found.size() && (int)i < cap; ++i) out[i] = cb.found[i];
  return (int)cb.found.size();
}

// every hit of the segment (6 floats each); returns the count (<= cap written)
extern "C" int ref_world_raycast_all(void* w, const float* seg, int cap, float* out) {
  RefWorld* rw = (RefWorld*)w;
  auto idx = fixture_index(rw);
  RayCollectAll cb;
  cb.index = &idx;
  rw->world.RayCast(&cb, b2Vec2(seg[0], seg[1]), b2Vec2(seg[2], seg[3]));
  const int n = (int)cb.hits.size() / 6;
  for (int i = 0; i < n && i < cap; ++i)
    for (int k = 0; k < 6; ++k) out[6 * i + k] = cb.hits[6 * i + k];
  return n;
}
