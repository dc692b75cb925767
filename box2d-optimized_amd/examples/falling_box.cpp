// falling_box.cpp -- user code written against the b2World API, built
// against the drop-in header and libbox2d_amd.so instead of the reference
// library. A 2 x 2 m box falls onto a 100 x 20 m static ground box, as in the
// reference's hello-world unit test (unit-test/hello_world.cpp), and a
// contact listener counts BeginContact events and sums the normal impulses
// PostSolve reports in the last step.
//
// Prints one line: "x y angle begin_events contacts impulse", then exits 0
// when the box rests at y = 1.01 +- 0.01 (the reference test's bound) and the
// ground's last-step normal impulse carries its weight, m g dt within 1%
// (momentum balance of a body at rest); 1 otherwise, 2 when the device path
// is unavailable.
#include <cmath>
#include <cstdio>

#include "box2d/box2d.h"

namespace {
struct CountingListener : public b2ContactListener {
  int begins = 0;
  float normal_impulse = 0.0f;  // summed over the PostSolve calls of one step
  void BeginContact(b2Contact*) override { ++begins; }
  void PostSolve(b2Contact*, const b2ContactImpulse* impulse) override {
    for (int k = 0; k < impulse->count; ++k) normal_impulse += impulse->normalImpulses[k];
  }
};
}  // namespace

int main() {
  b2World world(b2Vec2(0.0f, -10.0f));
  // continuous physics and sleeping are on by default and do not run on the
  // device (Step would refuse); the reference's benchmark drivers turn
  // continuous physics off the same way
  world.SetContinuousPhysics(false);
  world.SetAllowSleeping(false);
  CountingListener listener;
  world.SetContactListener(&listener);

  b2BodyDef groundDef;
  groundDef.position.Set(0.0f, -10.0f);
  b2Body* ground = world.CreateBody(&groundDef);
  b2PolygonShape groundBox;
  groundBox.SetAsBox(50.0f, 10.0f);
  ground->CreateFixture(&groundBox, 0.0f);

  b2BodyDef boxDef;
  boxDef.type = b2_dynamicBody;
  boxDef.position.Set(0.0f, 4.0f);
  b2Body* box = world.CreateBody(&boxDef);
  b2PolygonShape dynamicBox;
  dynamicBox.SetAsBox(1.0f, 1.0f);
  b2FixtureDef fd;
  fd.shape = &dynamicBox;
  fd.density = 1.0f;
  fd.friction = 0.3f;
  box->CreateFixture(&fd);

  if (world.Prepare() != 0) {
    std::printf("device unavailable: %s\n", b2g_last_error());
    return 2;
  }
  const float timeStep = 1.0f / 60.0f;
  for (int i = 0; i < 60; ++i) {
    listener.normal_impulse = 0.0f;
    world.Step(timeStep, 6, 2);
    if (world.GetLastStatus() != 0) {
      std::printf("step failed: %s\n", world.GetLastErrorMessage());
      return 2;
    }
  }
  const b2Vec2 p = box->GetPosition();
  const float a = box->GetAngle();
  const float weight_impulse = box->GetMass() * 10.0f * timeStep;
  std::printf("%.6f %.6f %.6f %d %d %.6f\n", p.x, p.y, a, listener.begins, world.GetContactCount(),
              listener.normal_impulse);
  const bool ok = std::fabs(p.x) < 0.01f && std::fabs(p.y - 1.01f) < 0.01f && std::fabs(a) < 0.01f &&
                  listener.begins == 1 &&
                  std::fabs(listener.normal_impulse - weight_impulse) < 0.01f * weight_impulse;
  return ok ? 0 : 1;
}
