// restate.h -- CPU restatement of Box2D-optimized's b2World::Step hot path.
//
// TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as a checker. The product (libbox2d_amd.so)
// never includes or links this code.
//
// What it restates (paths relative to the reference tree):
//   shapes      b2PolygonShape::Set/SetAsBox/ComputeMass, b2CircleShape::ComputeMass,
//               ComputeAABB of the three shapes        src/collision/b2_*_shape.cpp
//   bodies      b2Body::ResetMassData, SynchronizeTransform, ShouldCollide
//                                                       src/dynamics/b2_body.cpp, b2_body.h:957
//   narrowphase b2Collide{Circles,PolygonAndCircle,Polygons,EdgeAndCircle,EdgeAndPolygon},
//               b2ClipSegmentToLine, b2WorldManifold::Initialize
//                                                       src/collision/b2_collide_*.cpp, b2_collision.cpp
//   contacts    b2Contact::Update impulse carry-over, mixing laws, b2ContactFilter
//                                                       src/dynamics/b2_contact.cpp:126-210
//   solver      b2Island::Solve integration/clamps/write-back, b2ContactSolver (all but TOI),
//               b2RevoluteJoint Init/SolveVelocity/SolvePosition
//                                                       src/dynamics/b2_island.cpp:187-484,
//                                                       src/dynamics/b2_contact_solver.cpp:65-787,
//                                                       src/dynamics/b2_revolute_joint.cpp:73-321
//   broadphase  the pair SET that b2BroadPhase::UpdateAndQuery + QueryCallback produce
//               (inclusive tight-AABB overlap, filtered), by sort-and-sweep
//                                                       include/box2d/b2_broad_phase.h:252-511
//
// Parity status of each part is stated in tests/test_oracle.py (checks against
// the compiled reference, closed forms and invariants).
//
// Constraint order. The reference solves each island's constraints in DFS
// order. The restatement solves the constraints of all islands in one
// sequence: joints then contacts in each velocity sweep, contacts then joints
// in each position sweep (b2_island.cpp:329-337, 389-401), ordered by a
// caller-supplied colour (ascending) and then by index. Constraints of one
// colour share no non-static body, so any order inside a colour gives the
// same result; with the GPU's colours the restatement reproduces the GPU's
// Gauss-Seidel sequence. Position iterations stop per island exactly as
// b2Island::Solve does (b2_island.cpp:389-408).
#ifndef ORACLE_RESTATE_H
#define ORACLE_RESTATE_H

#include "b2g.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_world orc_world;

orc_world* orc_world_from_scene(const b2g_scene* scene, int warm_starting);
void orc_world_destroy(orc_world* w);

/* shapes after Set/SetAsBox, with material/filter (out: fixture_count records) */
void orc_world_get_shapes(orc_world* w, b2g_shape* out);
/* bodies (b2g_body records, scene order) */
void orc_world_get_bodies(orc_world* w, b2g_body* out);
void orc_world_set_bodies(orc_world* w, const b2g_body* in);
void orc_world_get_joints(orc_world* w, b2g_revolute* out);
void orc_world_set_joints(orc_world* w, const b2g_revolute* in);
int orc_world_body_count(orc_world* w);
int orc_world_contact_count(orc_world* w);
/* contacts sorted by (min shape, max shape) */
void orc_world_get_contacts(orc_world* w, b2g_contact* out);
void orc_world_set_contacts(orc_world* w, int n, const b2g_contact* in);
/* dt ratio carried between steps (m_inv_dt0 * dt) */
void orc_world_set_inv_dt0(orc_world* w, float inv_dt0);
/* b2World::SetAllowSleeping (the bodies' sleep state travels in b2g_body) */
void orc_world_set_allow_sleep(orc_world* w, int flag);

/* Solver order for the next steps: colour of each touching contact, looked up
 * by key (min << 32 | max), and of each joint. NULL = all colour 0. */
void orc_world_set_colours(orc_world* w, int n, const uint64_t* contact_keys,
                           const int32_t* contact_colours, const int32_t* joint_colours);

/* The order in which the next step's Collide walks the contacts (keys), as
 * the reference walks its contact list, a wake-up counting at once for the
 * contacts after it; used by that one step only. Without it the contacts
 * are updated from the awake flags at the start of Collide. */
void orc_world_set_collide_order(orc_world* w, int n, const uint64_t* contact_keys);

/* flags: B2G_STEP_FIND_NEW_CONTACTS forces a broadphase before Collide;
 * B2G_STEP_SKIP_BROADPHASE omits the end-of-step broadphase. */
void orc_world_step(orc_world* w, float dt, int vel_iters, int pos_iters, unsigned flags);

/* ---- stage functions ---- */
void orc_collide(const b2g_shape* a, const float* xfa, const b2g_shape* b, const float* xfb,
                 b2g_contact* out);
void orc_world_manifold(const b2g_contact* c, const float* xfa, float ra, const float* xfb,
                        float rb, float* out /* normal 2, points 4, separations 2 */);
/* AABB (4 floats) of a shape at a transform */
void orc_shape_aabb(const b2g_shape* s, const float* xf, float* out);
/* sorted pair keys of overlapping, collidable shapes; returns the pair count
 * (keys written up to cap) */
long long orc_find_pairs(int n, const float* aabbs, const int32_t* shape_body,
                         const int32_t* shape_type, const int32_t* filter, int body_count,
                         const int32_t* body_type, long long cap, uint64_t* out_keys);
/* 2-point block LCP of b2ContactSolver::SolveVelocityConstraints
 * (b2_contact_solver.cpp:389-631) on one constraint; in/out impulses a[2] */
void orc_block_solve(const float* K /* k11 k12 k22 */, const float* b /* vn - bias */,
                     float* a, int* case_taken);

#ifdef __cplusplus
}
#endif

#endif
