/*
 * b2g.h -- C ABI of the MI355X b2World::Step engine (libbox2d_amd.so).
 *
 * Plain C: pointers, sizes and status codes; no C++ or HIP types cross this
 * boundary. The C++ drop-in classes (b2World, b2Body, b2Fixture, b2Contact,
 * b2RevoluteJoint, ...) in box2d-optimized_amd/host/box2d/ are implemented on
 * top of these entry points, the same way the reference's b2World drives its
 * b2ContactManager / b2Island / b2BroadPhase internals.
 *
 * Every entry point returns 0 (B2G_OK) or a B2G_ERR_* code; b2g_last_error()
 * gives the message of the last failure on the calling thread.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   b2g_batch_step           -> b2World::Step            include/box2d/b2_world.h:127-142,
 *                                                         src/dynamics/b2_world.cpp:1108-1171
 *   b2g_upload_bodies        -> b2World::CreateBody / b2Body setters
 *                                                         src/dynamics/b2_world.cpp:140-176,
 *                                                         include/box2d/b2_body.h:183-285
 *   b2g_upload_shapes        -> b2Body::CreateFixture + b2BroadPhase::Add
 *                                                         src/dynamics/b2_body.cpp:230-269,
 *                                                         src/collision/b2_broad_phase.cpp:118-180
 *   b2g_upload_joints        -> b2World::CreateJoint (revolute)
 *                                                         src/dynamics/b2_world.cpp:268-324
 *   b2g_download_bodies      -> b2Body::GetTransform/GetPosition/GetAngle/
 *                               GetLinearVelocity/GetAngularVelocity
 *                                                         include/box2d/b2_body.h:187-217
 *   b2g_download_contacts    -> b2World::GetContactList + b2Contact::GetManifold/IsTouching
 *                                                         include/box2d/b2_world.h:187-192,
 *                                                         include/box2d/b2_contact.h:71-78
 *   b2g_upload_contacts      -> (no direct counterpart: restores the contact list with
 *                               warm-start impulses, which b2World::Dump cannot)
 *                                                         src/dynamics/b2_world.cpp:1469-1531
 *   b2g_get_counts           -> b2World::GetBodyCount/GetContactCount/GetJointCount
 *                                                         include/box2d/b2_world.h:211-223
 *   b2g_collide_pairs        -> b2Contact::Evaluate -> b2Collide* (5 shape pairs)
 *                                                         src/dynamics/b2_contact.cpp:48-56,
 *                                                         include/box2d/b2_collision.h:227-250
 *   b2g_find_pairs           -> b2BroadPhase::UpdateAndQuery pair reports, filtered as
 *                               b2ContactManager::QueryCallback filters them
 *                                                         include/box2d/b2_broad_phase.h:252-343,
 *                                                         src/dynamics/b2_contact_manager.cpp:136-188
 */
#ifndef B2G_H
#define B2G_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
enum {
  B2G_OK = 0,
  B2G_ERR_HIP = 1,          /* a HIP runtime call failed */
  B2G_ERR_OOM = 2,          /* device allocation failed */
  B2G_ERR_CAPACITY = 3,     /* a fixed capacity was exceeded (pairs, colours, ...) */
  B2G_ERR_INVALID = 4,      /* bad handle, index or size */
  B2G_ERR_UNSUPPORTED = 5,  /* feature the device path does not implement (sensors, ...) */
  B2G_ERR_NO_DEVICE = 6     /* no HIP device visible */
};

/* ---- shape / body kinds (values match b2Shape::Type, b2BodyType) ------ */
enum { B2G_CIRCLE = 0, B2G_EDGE = 1, B2G_POLYGON = 2 };
enum { B2G_STATIC = 0, B2G_KINEMATIC = 1, B2G_DYNAMIC = 2 };

/* ---- scene description ---------------------------------------------------
 * The raw definitions a user hands to CreateBody / CreateFixture / CreateJoint,
 * as flat arrays. Index order is creation order. Used by the scene loader of
 * the drop-in shim (b2a_world_from_scene) and by test drivers.
 *
 * fix_kind / fix_geom (17 floats per fixture):
 *   0 circle : [0]=radius, [1..2]=center                 (b2CircleShape m_radius, m_p)
 *   1 edge   : [0]=oneSided(0/1), [1..8]=v0,v1,v2,v3     (b2EdgeShape::SetOneSided / SetTwoSided(v1,v2))
 *   2 polygon: [0]=count(3..8), [1..16]=points           (b2PolygonShape::Set: weld + hull)
 *   3 box    : [0]=hx, [1]=hy, [2..3]=center, [4]=angle, [5]=1 to use the
 *              centred form SetAsBox(hx,hy,center,angle), else SetAsBox(hx,hy)
 * fix_material (4 per fixture): density, friction, restitution, restitutionThreshold
 * fix_filter   (4 per fixture): categoryBits, maskBits, groupIndex, isSensor (NULL = b2Filter defaults)
 * joint_* : revolute joints built with b2RevoluteJointDef::Initialize(bodyA, bodyB, anchor)
 *   joint_params (6 per joint): anchor.x, anchor.y, lowerAngle, upperAngle, motorSpeed, maxMotorTorque
 *   joint_flags: bit0 collideConnected, bit1 enableLimit, bit2 enableMotor
 */
typedef struct b2g_scene {
  float gravity[2];
  int32_t body_count;
  const int32_t* body_type;        /* [n] B2G_STATIC/KINEMATIC/DYNAMIC */
  const float* body_pos;           /* [2n] */
  const float* body_angle;         /* [n] */
  const float* body_vel;           /* [3n] vx, vy, w; NULL = zero */
  const float* body_damping;       /* [2n] linear, angular; NULL = zero */
  const float* body_gravity_scale; /* [n]; NULL = 1 */
  const int32_t* body_flags;       /* [n] bit0 fixedRotation, bit1 bullet, bit2 !allowSleep,
                                      bit3 !awake (b2BodyDef); NULL = 0 */
  int32_t fixture_count;
  const int32_t* fix_body;         /* [m] */
  const int32_t* fix_kind;         /* [m] */
  const float* fix_geom;           /* [17m] */
  const float* fix_material;       /* [4m] */
  const int32_t* fix_filter;       /* [4m] or NULL */
  int32_t joint_count;
  const int32_t* joint_bodies;     /* [2j] */
  const float* joint_params;       /* [6j] */
  const int32_t* joint_flags;      /* [j] */
} b2g_scene;

/* ---- device-facing records (host side of the upload/download calls) ---- */

/* One rigid body: the fields of b2Body (include/box2d/b2_body.h:498-547) that
 * the step reads or writes. */
typedef struct b2g_body {
  int32_t type;           /* B2G_STATIC / B2G_KINEMATIC / B2G_DYNAMIC */
  int32_t world;          /* index of the world in the batch */
  float xf[4];            /* p.x, p.y, q.s, q.c */
  float c[2];             /* sweep.c (centre of mass, world) */
  float a;                /* sweep.a */
  float local_center[2];  /* sweep.localCenter */
  float v[2];             /* linear velocity */
  float w;                /* angular velocity */
  float mass;             /* m_mass (gravity is applied as m * g * invMass) */
  float inv_mass, inv_i;
  float linear_damping, angular_damping, gravity_scale;
  float force[2];
  float torque;
  int32_t flags;          /* B2G_BODY_ASLEEP: !b2Body::IsAwake(); B2G_BODY_NO_AUTOSLEEP: sleeping not
                             allowed (b2BodyDef::allowSleep false); 0 = awake, may sleep (the defaults) */
  float sleep_time;       /* m_sleepTime (include/box2d/b2_body.h:541-543) */
} b2g_body;
enum { B2G_BODY_ASLEEP = 1, B2G_BODY_NO_AUTOSLEEP = 2 };

/* One fixture with its processed shape (after b2PolygonShape::Set etc.). */
typedef struct b2g_shape {
  int32_t type;           /* B2G_CIRCLE / B2G_EDGE / B2G_POLYGON */
  int32_t body;           /* body index in the batch */
  int32_t count;          /* polygon vertex count; edge: oneSided flag */
  float radius;           /* m_radius (polygon skin, circle radius) */
  float centroid[2];      /* polygon m_centroid; circle m_p */
  float vertices[16];     /* polygon vertices; edge v0,v1,v2,v3 */
  float normals[16];      /* polygon edge normals */
  float friction, restitution, restitution_threshold;
  int32_t category, mask, group;
  int32_t is_sensor;
} b2g_shape;

/* One revolute joint (src/dynamics/b2_revolute_joint.cpp:52-71 fields). */
typedef struct b2g_revolute {
  int32_t body_a, body_b;
  float local_anchor_a[2], local_anchor_b[2];
  float reference_angle;
  float lower_angle, upper_angle;
  float motor_speed, max_motor_torque;
  int32_t flags;          /* bit0 collideConnected, bit1 enableLimit, bit2 enableMotor */
  /* solver state carried across steps (warm starting) */
  float impulse[2];
  float motor_impulse, lower_impulse, upper_impulse;
} b2g_revolute;

/* One contact with its manifold (b2Contact + b2Manifold,
 * include/box2d/b2_contact.h:65-217, include/box2d/b2_collision.h:44-117). */
typedef struct b2g_contact {
  int32_t shape_a, shape_b;   /* fixture indices, in the contact's own A/B order */
  int32_t touching;           /* b2Contact::IsTouching */
  int32_t type;               /* b2Manifold::Type: 0 circles, 1 faceA, 2 faceB */
  int32_t point_count;
  float local_normal[2];
  float local_point[2];
  float point_local[4];       /* points[0].localPoint, points[1].localPoint */
  float normal_impulse[2];
  float tangent_impulse[2];
  uint32_t id[2];             /* b2ContactID::key of each point */
  float friction, restitution, restitution_threshold, tangent_speed;
} b2g_contact;

typedef struct b2g_counts {
  int32_t worlds, bodies, shapes, joints;
  int32_t contacts;           /* b2World::GetContactCount: AABB-overlapping pairs */
  int32_t touching;           /* contacts with pointCount > 0 */
  int32_t colours;            /* constraint colours used by the last step */
  int32_t islands;            /* islands of the last step */
  int32_t pos_iters_run;      /* max position iterations executed by an island */
  int32_t status;             /* sticky device status (B2G_ERR_CAPACITY on overflow) */
  int32_t serial;             /* constraints the last step solved in the serial group */
  int32_t graph_steps;        /* steps replayed from a captured graph so far */
} b2g_counts;

/* Per-stage device time of the last step, in milliseconds (b2Profile,
 * include/box2d/b2_time_step.h:29-36; measured with HIP events when
 * B2G_STEP_PROFILE is set). */
typedef struct b2g_profile {
  float step, collide, solve, broadphase;
} b2g_profile;

/* ---- batch (a set of independent worlds sharing device arrays) ------- */
typedef struct b2g_batch b2g_batch;

typedef struct b2g_capacity {
  int32_t worlds, bodies, shapes, joints;
  int32_t contacts;           /* AABB-overlap pairs kept per step */
} b2g_capacity;

/* per-world parameters of one step */
typedef struct b2g_world_params {
  float gravity[2];
  float dt;
  float dt_ratio;             /* m_inv_dt0 * dt (first step: 0) */
  int32_t warm_starting;
  int32_t active;             /* 0 = skip this world in the launch */
  int32_t allow_sleep;        /* b2World::SetAllowSleeping (include/box2d/b2_world.h:195-197): bodies of
                                 islands at rest for b2_timeToSleep fall asleep (b2_island.cpp:444-483) */
} b2g_world_params;

enum {
  B2G_STEP_FIND_NEW_CONTACTS = 1u << 0,  /* run the broadphase before Collide (m_newContacts) */
  B2G_STEP_PROFILE = 1u << 1,            /* time each stage with HIP events (adds syncs) */
  B2G_STEP_ASYNC = 1u << 2,              /* return without waiting for the device */
  B2G_STEP_SKIP_BROADPHASE = 1u << 3,    /* omit the end-of-step broadphase (testing only) */
  B2G_STEP_EVENTS = 1u << 5,             /* record contact events (b2g_download_events) */
  B2G_STEP_COLLIDE_ONLY = 1u << 6,       /* b2World::Step with dt <= 0: the pending broadphase
                                            (B2G_STEP_FIND_NEW_CONTACTS), Collide and ClearForces only;
                                            no solve, so stored impulses and bodies are kept
                                            (src/dynamics/b2_world.cpp:1114-1141, 1164-1167) */
  B2G_STEP_GRAPH = 1u << 4               /* replay a captured hipGraph of the whole step: no host round
                                            trip; colour slots and launch sizes fixed at capture from the
                                            last eager step (re-captured after uploads). The first step
                                            after an upload runs eagerly. */
};

int b2g_device_count(int32_t* count);
const char* b2g_last_error(void);

int b2g_batch_create(b2g_batch** out, const b2g_capacity* cap);
int b2g_batch_destroy(b2g_batch* batch);

/* world ranges: bodies [body_begin, body_end) etc. are assigned by the caller */
int b2g_set_worlds(b2g_batch* batch, int32_t count, const int32_t* body_begin,
                   const int32_t* shape_begin, const b2g_world_params* params);
/* Parameters of worlds [first, first + count). The copy to the device is in
 * stream order with the steps (the call returns without waiting for it);
 * `params` may be reused as soon as the call returns. */
int b2g_set_world_params(b2g_batch* batch, int32_t first, int32_t count,
                         const b2g_world_params* params);

/* Live record counts after bodies, fixtures or joints were destroyed
 * (b2World::DestroyBody src/dynamics/b2_world.cpp:178-266, b2World::DestroyJoint
 * :326-418, b2Body::DestroyFixture src/dynamics/b2_body.cpp:280-350). Counts may
 * be anything up to the capacity; the caller then uploads the records that are
 * renumbered or new (the device keeps no tombstones: indices are positions in
 * the caller's creation order). The contact list is emptied, so the caller
 * re-uploads the surviving contacts (b2g_upload_contacts) with their colours;
 * the next broadphase queries every world. */
int b2g_set_counts(b2g_batch* batch, int32_t bodies, int32_t shapes, int32_t joints);

int b2g_upload_bodies(b2g_batch* batch, int32_t first, int32_t count, const b2g_body* bodies);
int b2g_upload_shapes(b2g_batch* batch, int32_t first, int32_t count, const b2g_shape* shapes);
int b2g_upload_joints(b2g_batch* batch, int32_t first, int32_t count, const b2g_revolute* joints);
/* replaces the whole contact list (sorted by the batch internally) */
int b2g_upload_contacts(b2g_batch* batch, int32_t count, const b2g_contact* contacts);

/* one b2World::Step for every active world of the batch */
int b2g_batch_step(b2g_batch* batch, int32_t velocity_iterations, int32_t position_iterations,
                   uint32_t flags);
int b2g_batch_wait(b2g_batch* batch);

int b2g_get_counts(b2g_batch* batch, b2g_counts* out);

/* Timing of every launch of one kernel during the following steps (-1 =
 * off): HIP events around each run of consecutive launches of that kernel
 * (the per-colour launches of one sweep, or a single launch). The contact
 * sweeps are the per-colour launches (not the launch shared by the small
 * trailing colours). b2g_kernel_time returns the summed device time, the
 * number of launches and the work items they covered (contacts, constraints
 * or shapes). */
enum {
  B2G_KERNEL_NARROWPHASE = 0,
  B2G_KERNEL_WARM_START = 1,
  B2G_KERNEL_CONTACT_VELOCITY = 2,
  B2G_KERNEL_CONTACT_POSITION = 3,
  B2G_KERNEL_PAIRS_WRITE = 4,
  B2G_KERNEL_SOLVER = 5
};
int b2g_time_kernel(b2g_batch* batch, int32_t kernel);

/* Solver layout and tuning. B2G_OPT_SOLVER: B2G_SOLVER_AUTO (default: one
 * workgroup per world when the batch has several joint-free worlds of at most
 * 8192 bodies laid out world after world; else colour-major sweeps),
 * B2G_SOLVER_COLOURS (one launch
 * per colour over all worlds), B2G_SOLVER_WORLDS (one workgroup solves a
 * whole world, barriers between colours; needs that layout; the same results
 * as the colour sweeps) or B2G_SOLVER_CELLS (constraints split by the cells
 * of two staggered grids and a small cross group, each cell solved by one
 * workgroup with its bodies in LDS: three launches per iteration and no host
 * read per step; another valid constraint order, reported by
 * b2g_download_contact_colours as phase * 65536 + colour; on request only).
 * B2G_OPT_WORLD_BLOCK: threads per world workgroup, 0 = autotune: the first
 * per-world steps time 64..1024 threads, each with the world's bodies in
 * global memory and staged in LDS (when the largest world fits 63 KB), one
 * step per candidate, and keep the fastest.
 * Read-only: B2G_OPT_TUNED_WORLD_BLOCK, B2G_OPT_LAST_SOLVER (1 colours,
 * 2 worlds, 3 cells), B2G_OPT_PER_WORLD_ELIGIBLE. No reference counterpart. */
enum { B2G_SOLVER_AUTO = 0, B2G_SOLVER_COLOURS = 1, B2G_SOLVER_WORLDS = 2, B2G_SOLVER_CELLS = 3 };
enum {
  B2G_OPT_SOLVER = 1,
  B2G_OPT_WORLD_BLOCK = 2,
  B2G_OPT_TUNED_WORLD_BLOCK = 3,
  B2G_OPT_LAST_SOLVER = 4,
  B2G_OPT_PER_WORLD_ELIGIBLE = 5,
  B2G_OPT_WORLD_LDS = 6,        /* -1 autotune (default), 0 off, 1 stage each world's bodies in LDS */
  B2G_OPT_TUNED_WORLD_LDS = 7,  /* read-only: whether the per-world solver stages bodies in LDS */
  /* 1 (default): a world that is not active in a step keeps its pairs and,
   * on the steps that refit the tree, its boxes and tree nodes, instead of
   * querying and refitting again (its bodies have not moved). The contact
   * list is the same either way; 0 queries every world each step. Any
   * upload, and a step with B2G_STEP_SKIP_BROADPHASE, makes the next
   * broadphase query every world. */
  B2G_OPT_KEEP_IDLE = 8,
  /* Pair search of the broadphase: 0 (default) a hashed uniform grid when no
   * world of the batch has more than 64 static shapes, else the LBVH; 1 always
   * the LBVH. The pair list is the same either way. B2G_OPT_LAST_PAIR_SEARCH
   * (read-only): 1 LBVH, 2 grid in the last broadphase. */
  B2G_OPT_PAIR_SEARCH = 9,
  B2G_OPT_LAST_PAIR_SEARCH = 10
};
int b2g_set_option(b2g_batch* batch, int32_t option, int32_t value);
int b2g_get_option(b2g_batch* batch, int32_t option, int32_t* value);
int b2g_kernel_time(b2g_batch* batch, double* total_ms, int64_t* launches, int64_t* items);
int b2g_get_profile(b2g_batch* batch, b2g_profile* out);
int b2g_download_bodies(b2g_batch* batch, int32_t first, int32_t count, b2g_body* out);
/* contacts are ordered by (min shape, max shape); out may be NULL to query the count */
int b2g_download_contacts(b2g_batch* batch, int32_t first, int32_t count, b2g_contact* out);
int b2g_download_joints(b2g_batch* batch, int32_t first, int32_t count, b2g_revolute* out);
/* constraint colour of each touching contact of the last step (-1 = not solved);
 * out has one entry per downloaded contact, same order */
int b2g_download_contact_colours(b2g_batch* batch, int32_t first, int32_t count, int32_t* out);
/* Points the last step's solver stored for each current contact: the
 * b2ContactImpulse count b2Island::Report hands to PostSolve
 * (src/dynamics/b2_island.cpp:621-647; 0 = not solved that step, else the
 * velocity constraint's pointCount, which the block solver's condition check
 * may have lowered to 1, src/dynamics/b2_contact_solver.cpp:244-255). The
 * impulses themselves are the manifold's (b2g_download_contacts), as
 * b2ContactSolver::StoreImpulses wrote them. Same order as the contacts. */
int b2g_download_solver_points(b2g_batch* batch, int32_t first, int32_t count, int32_t* out);

/* Sets the persisted colours of the current contacts (count = contact count,
 * in the key-sorted order b2g_download_contacts returns; -1 = uncoloured).
 * No reference counterpart: with b2g_upload_contacts it restores a solver
 * state exactly. */
int b2g_upload_contact_colours(b2g_batch* batch, int32_t count, const int32_t* colours);

/* Contact events of the last step taken with B2G_STEP_EVENTS: the calls
 * b2ContactListener::BeginContact / EndContact receive in b2Contact::Update
 * (src/dynamics/b2_contact.cpp:197-203) and b2ContactManager::Destroy
 * (src/dynamics/b2_contact_manager.cpp:48-51). Collide's events come first,
 * then the broadphase's, each group ordered by shape pair (the reference
 * orders them by its contact list). `contact` holds the manifold as the
 * listener would see it. *count = number of events (may exceed capacity). */
enum { B2G_EVENT_BEGIN = 1, B2G_EVENT_END = 2, B2G_EVENT_DESTROYED = 3 };
typedef struct b2g_contact_event {
  int32_t kind;
  int32_t solver_points;  /* B2G_EVENT_DESTROYED: points the step's solver stored for the contact
                             before the broadphase destroyed it (PostSolve's count); else 0 */
  b2g_contact contact;
} b2g_contact_event;
int b2g_download_events(b2g_batch* batch, int32_t capacity, b2g_contact_event* out, int32_t* count);

/* Queries on the tree of the last broadphase (rebuilt first when bodies or
 * shapes changed since). b2g_query_aabbs: shapes of world worlds[i] whose
 * AABB overlaps aabbs[4i..4i+3] (b2World::QueryAABB,
 * src/dynamics/b2_world.cpp:1193-1198). b2g_raycast: every shape of world
 * worlds[i] hit by the segment segments[4i..4i+3] = (p1, p2), with the hit's
 * fraction and normal (b2World::RayCast, src/dynamics/b2_world.cpp:1228-1238,
 * at maxFraction 1; a caller applies its callback's clipping in fraction
 * order). Up to `capacity` results per query go to out[i * capacity + k];
 * out_counts[i] is the full count (may exceed capacity). */
typedef struct b2g_ray_hit {
  int32_t shape;
  float fraction;
  float normal[2];
} b2g_ray_hit;
int b2g_query_aabbs(b2g_batch* batch, int32_t n, const int32_t* worlds, const float* aabbs, int32_t capacity,
                    int32_t* out_shapes, int32_t* out_counts);
int b2g_raycast(b2g_batch* batch, int32_t n, const int32_t* worlds, const float* segments, int32_t capacity,
                b2g_ray_hit* out_hits, int32_t* out_counts);

/* Binary snapshot of a batch: capacity, world params, bodies, shapes,
 * joints with their impulses, contacts with manifolds, impulses and colours.
 * A batch loaded from it continues bit for bit as the saved one would
 * (world params hold the last dt and dtRatio; the caller continues dtRatio
 * as inv_dt0 * dt). Covers what b2World::Dump
 * (src/dynamics/b2_world.cpp:1469-1531) writes plus the contact state it
 * loses. */
int b2g_snapshot_save(b2g_batch* batch, const char* path);
int b2g_snapshot_load(b2g_batch** out, const char* path);

/* ---- stage entry points (pure functions over caller arrays) ------------ */

/* Narrowphase: for n pairs (shape_a[i], shape_b[i]) with transforms
 * xf_a/xf_b (4 floats each: p.x, p.y, q.s, q.c), compute the manifold into
 * out[i] (shape_a/shape_b/touching/type/points filled; impulses zeroed). */
int b2g_collide_pairs(int32_t n, const b2g_shape* shape_a, const float* xf_a,
                      const b2g_shape* shape_b, const float* xf_b, b2g_contact* out);

/* b2g_collide_pairs for callers that cannot allocate device memory (code
 * that runs under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. Same arguments and
 * results as b2g_collide_pairs, bit for bit; the staging lives in a 4 MiB
 * device array that each GPU holds from the first call until the process
 * ends, so a call takes at most 4096 pairs (B2G_ERR_CAPACITY beyond). Runs on
 * the null stream of the current device and returns when `out` is written;
 * calls that overlap on one GPU share the array and race. */
int b2g_collide_pairs_noalloc(int32_t n, const b2g_shape* shape_a, const float* xf_a,
                              const b2g_shape* shape_b, const float* xf_b, b2g_contact* out);

/* Broadphase: all pairs of the n AABBs (4 floats each: lower.x, lower.y,
 * upper.x, upper.y) that overlap (inclusive test, b2TestOverlap) and whose
 * shapes may collide (different bodies, at least one dynamic body, filter
 * bits, no edge-edge). body_type/shape_body/filter describe the shapes.
 * Pairs are written sorted as (min << 32 | max) keys. */
int b2g_find_pairs(int32_t n, const float* aabbs, const int32_t* shape_body,
                   const int32_t* shape_type, const int32_t* filter /* 3 per shape, NULL ok */,
                   int32_t body_count, const int32_t* body_type,
                   int64_t max_pairs, uint64_t* out_keys, int64_t* out_count);

/* ---- scene loader of the drop-in shim (b2World built through its C++ API) */
typedef struct b2a_world b2a_world;
int b2a_world_from_scene(b2a_world** out, const b2g_scene* scene, int32_t warm_starting);
int b2a_world_destroy(b2a_world* world);
/* b2World::Step(dt, vel, pos) */
int b2a_world_step(b2a_world* world, float dt, int32_t velocity_iterations,
                   int32_t position_iterations);
/* bodies in scene order: 6 floats each: c.x, c.y, a, v.x, v.y, w (sweep state) */
int b2a_world_get_bodies(b2a_world* world, float* out);
int b2a_world_contact_count(b2a_world* world, int32_t* out);
int b2a_world_get_contacts(b2a_world* world, int32_t capacity, b2g_contact* out, int32_t* count);
b2g_batch* b2a_world_batch(b2a_world* world);
/* allocate the device batch and upload bodies/shapes/joints without stepping */
int b2a_world_prepare(b2a_world* world);
/* b2World::DestroyBody (src/dynamics/b2_world.cpp:178-266), b2World::DestroyJoint
 * (:326-418), b2Body::DestroyFixture (src/dynamics/b2_body.cpp:280-350) and
 * b2Fixture::SetFilterData (src/dynamics/b2_fixture.cpp:142-146) on the
 * world's objects, named by their index in the world's current creation order
 * (later objects move down one index when one is destroyed). */
int b2a_world_destroy_body(b2a_world* world, int32_t body);
int b2a_world_destroy_joint(b2a_world* world, int32_t joint);
int b2a_world_destroy_fixture(b2a_world* world, int32_t fixture);
int b2a_world_set_filter(b2a_world* world, int32_t fixture, int32_t category, int32_t mask, int32_t group);
int b2a_world_body_count(b2a_world* world, int32_t* out);
/* b2World::SetAllowSleeping (src/dynamics/b2_world.cpp:474-489) */
int b2a_world_set_allow_sleep(b2a_world* world, int32_t flag);

/* Device records of a scene, built through the same b2World API (polygon hull
 * and welding, mass data) without touching the device: bodies tagged with
 * `world`, shape/joint body indices shifted by `body_offset`. Output arrays
 * hold body_count / fixture_count / joint_count records. Host only. */
int b2a_scene_records(const b2g_scene* scene, int32_t world, int32_t body_offset, b2g_body* bodies,
                      b2g_shape* shapes, b2g_revolute* joints);
/* A batch of `count` independent worlds, world i built from scenes[i] as
 * b2a_world_from_scene would build it, laid out world after world; contact
 * capacity = contacts_per_shape x total shapes. The batched form of creating
 * `count` b2Worlds; no reference counterpart (the reference steps one world). */
int b2a_batch_from_scenes(b2g_batch** out, int32_t count, const b2g_scene* scenes,
                          int32_t contacts_per_shape, int32_t warm_starting);

/* ---- serving: continuous-batching request loop over a resident pool ------
 * No reference counterpart (the reference steps one world on the caller's
 * thread); it serves callers that submit world-step requests. */
typedef struct b2s_request {
  double arrival;             /* seconds from the start (realtime) or tick number (not realtime) */
  int32_t world;              /* world index in the pool */
  int32_t substeps;           /* steps to run, >= 1 */
  float dt;                   /* step length of every substep */
  int32_t reserved;
} b2s_request;

typedef struct b2s_result {
  double arrival, admitted, completed;   /* same clock as the request's arrival */
  int32_t ticks;                         /* ticks the request was in flight */
  int32_t reserved;
} b2s_result;

typedef struct b2s_config {
  int64_t memory_budget;      /* device working memory of the worlds in flight, bytes */
  int32_t velocity_iterations, position_iterations;
  uint32_t step_flags;        /* 0 or B2G_STEP_GRAPH */
  int32_t realtime;           /* 1: arrivals on the wall clock; 0: arrivals on the tick counter */
} b2s_config;

/* wall_s: the loop's wall time; busy_s: host time inside b2g_batch_step;
 * device_s: device time of the ticks (HIP events around each step; eager
 * ticks only, 0 with B2G_STEP_GRAPH). */
typedef struct b2s_stats {
  double wall_s, busy_s, mean_in_flight;
  int64_t ticks, world_steps, body_steps;
  int32_t max_in_flight, completed;
  double device_s;
} b2s_stats;

/* Runs requests (sorted by arrival) against `pool`, a batch whose worlds
 * have taken at least one step. world_bodies / world_bytes: dynamic bodies
 * and device working bytes of each world; gravity: 2 per world; inv_dt0:
 * per-world 1/dt of the last step taken (0 if none), updated in place.
 * Returns when every request has completed. */
int b2s_run(b2g_batch* pool, int32_t world_count, const int32_t* world_bodies, const int64_t* world_bytes,
            const float* gravity, float* inv_dt0, int32_t request_count, const b2s_request* requests,
            const b2s_config* config, b2s_result* results, b2s_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* B2G_H */
