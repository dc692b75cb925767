// unsupported_features.cpp -- user code that turns on, one at a time, what the
// device path does not run (continuous physics is b2World's default) and checks
// that b2World::Step refuses each one loudly: status B2G_ERR_UNSUPPORTED, a
// message, and a world left as it was. Features the device does serve (a
// PostSolve listener, a contact filter that keeps the default ShouldCollide,
// sleeping and bodies put to sleep) are not refused. Once everything is off the world steps.
//
// Prints one line per case ("<case>: <status> <message>") and exits 0 when
// every case behaved so, the last step included (it ran on a device, or
// reported B2G_ERR_NO_DEVICE without one); 1 otherwise.
#include <cstdio>
#include <cstring>

#include "box2d/box2d.h"

namespace {
struct PreSolveListener : public b2ContactListener {
  void PreSolve(b2Contact*, const b2Manifold*) override {}
};
struct PostSolveListener : public b2ContactListener {
  int calls = 0;
  void PostSolve(b2Contact*, const b2ContactImpulse*) override { ++calls; }
};
struct NoFilter : public b2ContactFilter {
  bool ShouldCollide(b2Fixture*, b2Fixture*) override { return false; }
};
struct PlainFilter : public b2ContactFilter {};

int failures = 0;

// one Step that must be refused, leaving the box where it was
void expect_refused(b2World& world, b2Body* box, const char* what) {
  const b2Vec2 p0 = box->GetPosition();
  world.Step(1.0f / 60.0f, 8, 3);
  const b2Vec2 p1 = box->GetPosition();
  const bool ok = world.GetLastStatus() == B2G_ERR_UNSUPPORTED && std::strlen(world.GetLastErrorMessage()) > 0 &&
                  p0.x == p1.x && p0.y == p1.y;
  std::printf("%s: %d %s%s\n", what, world.GetLastStatus(), world.GetLastErrorMessage(), ok ? "" : "  <-- expected a refusal");
  if (!ok) ++failures;
}

// one Step that must not be refused
void expect_stepped(b2World& world, const char* what) {
  world.Step(1.0f / 60.0f, 8, 3);
  const int st = world.GetLastStatus();
  const bool ok = st == B2G_OK || st == B2G_ERR_NO_DEVICE;
  std::printf("%s: %d %s%s\n", what, st, st ? world.GetLastErrorMessage() : "", ok ? "" : "  <-- expected a step");
  if (!ok) ++failures;
}
}  // namespace

int main() {
  b2World world(b2Vec2(0.0f, -10.0f));
  b2BodyDef groundDef;
  b2Body* ground = world.CreateBody(&groundDef);
  b2EdgeShape edge;
  edge.SetTwoSided(b2Vec2(-40.0f, 0.0f), b2Vec2(40.0f, 0.0f));
  ground->CreateFixture(&edge, 0.0f);
  b2BodyDef boxDef;
  boxDef.type = b2_dynamicBody;
  boxDef.position.Set(0.0f, 4.0f);
  b2Body* box = world.CreateBody(&boxDef);
  b2PolygonShape shape;
  shape.SetAsBox(0.5f, 0.5f);
  b2Fixture* fixture = box->CreateFixture(&shape, 1.0f);

  expect_refused(world, box, "default world (continuous physics on)");
  world.SetContinuousPhysics(false);
  expect_stepped(world, "sleeping on");
  box->SetAwake(false);
  expect_stepped(world, "a body put to sleep");
  box->SetAwake(true);
  world.SetAllowSleeping(false);
  expect_stepped(world, "sleeping off");
  box->SetEnabled(false);
  expect_refused(world, box, "a disabled body");
  box->SetEnabled(true);
  fixture->SetSensor(true);
  expect_refused(world, box, "a sensor fixture");
  fixture->SetSensor(false);

  PreSolveListener pre;
  world.SetContactListener(&pre);
  expect_refused(world, box, "a listener overriding PreSolve");
  NoFilter no_filter;
  world.SetContactListener(nullptr);
  world.SetContactFilter(&no_filter);
  expect_refused(world, box, "a filter overriding ShouldCollide");

  PlainFilter plain;
  world.SetContactFilter(&plain);
  PostSolveListener post;
  world.SetContactListener(&post);
  expect_stepped(world, "default filter and a PostSolve listener");
  return failures == 0 ? 0 : 1;
}
