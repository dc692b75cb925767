// b2g_kernels.h -- the kernels of one b2World::Step on device memory.
//
// Stage order follows b2World::Step (src/dynamics/b2_world.cpp:1108-1171):
//   narrowphase (b2ContactManager::Collide) -> solve (b2World::Solve /
//   b2Island::Solve with b2ContactSolver and b2RevoluteJoint) -> broadphase
//   (b2ContactManager::FindNewContacts at the end of Solve) -> ClearForces.
//
// The island solver's sequential Gauss-Seidel is replaced by graph-coloured
// Gauss-Seidel: constraints of one colour share no dynamic body and run in
// parallel; colours run in sequence, joints before contacts in velocity
// sweeps and contacts before joints in position sweeps, as b2Island::Solve
// orders them. Position iterations stop per island (union-find components)
// exactly as b2Island::Solve stops them.
#pragma once

#include "b2g_collide.h"
#include "b2g_prims.h"
#include <climits>

namespace b2g {

// ---- device records ---------------------------------------------------------
struct Joint {  // b2RevoluteJoint state (src/dynamics/b2_revolute_joint.cpp:52-71)
  int a, b, flags, world;
  float la_x, la_y, lb_x, lb_y;
  float ref, lower, upper, speed;
  float max_torque, imp_x, imp_y, motor;
  float lower_imp, upper_imp, pad0, pad1;
};
struct JointSolve {  // InitVelocityConstraints temporaries
  float ra_x, ra_y, rb_x, rb_y;
  float kxx, kxy, kyx, kyy;
  float axial, angle, ma, mb;
  float ia, ib, lca_x, lca_y;
  float lcb_x, lcb_y, pad0, pad1;
};

enum Counter : int {
  C_CONTACTS = 0,   // contacts in the current buffer
  C_PAIRS = 1,      // pairs found by the last broadphase (may exceed capacity)
  C_CONS = 2,       // touching contacts solved this step
  C_STATUS = 3,     // sticky error bits
  C_REMAIN = 4,     // constraints still uncoloured
  C_NCOLOURS = 5,   // contact colours used (incl. the serial group)
  C_MAXCOLOUR = 6,
  C_ROUND = 7,      // colouring round base (claim priorities carry base + local round)
  C_SERIAL = 8,     // constraints placed in the serial group by the last step
  C_PENDING = 9,    // entries of the pending-colour list of this step
  C_LEFT = 10,      // entries of the list of constraints a collecting round left uncoloured
  C_EVENTS = 11,    // contact events recorded this step
  C_HEAVY = 12,     // over-full shapes whose pair segments need sorting (this broadphase)
  C_PART_CROSS = 13,  // constraints of the cells solver's cross group this step
  C_COUNT = 16
};
enum Status : int { S_CONTACT_OVERFLOW = 1, S_STACK_OVERFLOW = 2, S_TREE_CORRUPT = 4, S_EVENT_OVERFLOW = 8 };
// contact event kinds: touching began / ended in Collide, touching contact
// destroyed by the broadphase (b2ContactManager::Destroy)
enum EventKind : int { kEventBegin = 1, kEventEnd = 2, kEventDestroyed = 3 };
constexpr int kEventWords = 6;  // float4 per event: (shapeA, shapeB, kind, -), manifold (4), material
enum BodyType : int { kStaticBody = 0, kKinematicBody = 1, kDynamicBody = 2 };
constexpr int kMaxColours = 64;           // bit-mask colours; colour 64 = serial group
constexpr int kSerialColour = kMaxColours;
constexpr int kColourSlots = kMaxColours + 1;
constexpr int kBodyTab = 32;
constexpr int kCntColOff = 32;  // the colour offsets sit this many ints into the counter block  // colours a body's warm-start gather covers

// all device arrays, passed by value to every kernel
struct Dev {
  int nb, ns, nj, nw;
  // bodies
  float4* xf;     // p.x, p.y, q.s, q.c
  float4* pos;    // sweep c.x, c.y, a
  float4* vel;    // v.x, v.y, w
  float4* prop;   // invMass, invI, localCenter.x, localCenter.y
  float4* prop2;  // mass, gravityScale, linearDamping, angularDamping
  float4* force;  // f.x, f.y, torque
  int* btype;
  int* bworld;
  // worlds
  float4* wpar;   // gravity.x, gravity.y, dt, dtRatio
  int* wflag;     // bit0 warm starting, bit1 active
  int* wshape0;   // first shape index of each world (world-local shape numbering)
  int* wshape1;   // one past the last shape index of each world
  int2* wbody;    // body range of each world (worlds laid out world after world)
  int* worder;    // worlds by decreasing size: the per-world solver's block order
  int* wact;      // the worlds active in the step, in the same order (mostly idle batches)
  int* wcol;      // per world: kColourSlots + 1 slot offsets of its colour groups
  int* wcolmax;   // per world: highest normal colour used this step (-1: none)
  int* wfresh;    // per world: colours of its last colouring from scratch (0: none yet)
  int* wrecolour; // per world: 1 = colour every constraint afresh this step
  // shapes
  int4* shdr;     // type, body, count, world
  float4* geo;    // 9 per shape
  float4* smat;   // friction, restitution, restitutionThreshold, density
  int4* sfilt;    // category, mask, group, sensor
  float4* aabb;
  // joints
  Joint* joints;
  JointSolve* jsol;
  int* jorder;    // joint indices in colour order
  int* wj_off;    // per world: first entry in wj_list (nw + 1)
  int* wj_list;   // joint indices grouped by world
  int* jisland;   // island root per colour-ordered joint
  int* bj_off;    // CSR body -> partner bodies of joints with collideConnected == false
  int* bj_other;
  // contacts (current / next)
  uint64_t* ckey; int2* cpair; float4* cman; float4* cmat; int* ccol;
  uint64_t* nkey; int2* npair; float4* nman; float4* nmat; int* ncol;
  int ccap;
  // constraints of this step
  int* touch;     // per contact: 1 if solved this step
  int* tpos;      // exclusive scan of touch
  int* clist;     // compacted contact indices
  int2* cbody;    // per compacted contact: its two bodies, a static body as ~index
  int* ccolour;   // colour per compacted constraint
  int* csorted;   // colour-sorted slot -> contact index
  int* col_count; int* col_off; int* col_cursor;
  float4* vc[10];
  float4* pc[5];
  int* cisland;
  int* cpend;     // constraints (compacted slots) without a colour at the start of this step
  int* cleft;     // those a collecting claim round left uncoloured (k_colour_finish's list)
  float4* events; // kEventWords per contact event (recorded when events_on)
  int events_on;
  int ecap;
  unsigned long long* bmask;
  int* btab;      // per body, kBodyTab entries: its constraints' slots in colour order (slot * 2 + side)
  unsigned long long* bslot;
  // islands
  int* parent;
  unsigned* sepbad;  // 3 * nb: per island and position iteration (mod 3), 1 when a contact of the
                     // island separated by less than -3 linearSlop in that iteration
  int* jok;          // 3 * nb: per island and position iteration (mod 3), 0 when a joint was out of slop
  int* solved;       // nb
  // broadphase
  uint64_t* mkey[2]; uint32_t* mval[2];
  int* sort_hist;
  int* node_child;   // 2 per internal node: >= 0 internal, < 0 leaf ~k
  int* node_parent;
  int* leaf_parent;
  int2* node_range;
  float4* node_rec;  // 3 per internal node: left child box, right child box, (left, right, split, -)
  int4* sfast;       // per shape: body, type | dynamic << 4 | joint partners << 5 | static << 6,
                     // category | mask << 16, group
  float4* sparse;    // (levels) x ns AABB table over sorted leaves; level 0 = leaf AABBs
  int sparse_levels;
  uint64_t* pslots;  // up to kPairSlots per shape: pairs found by the traversal, sorted
  int2* heavy;       // (first, count) of the pair segments of over-full shapes, to be sorted
  int* wleaf;        // 2 per world: first leaf, last leaf + 1
  int* wroot;        // per world: the internal node whose leaves are exactly the world's (-1: one leaf)
  unsigned* wbounds; // 4 per world (ordered floats): min x, min y, max x, max y
  int* pcount;       // per shape: pairs whose lower shape it is (the pair list's segment length)
  int* sinv;         // per shape: its leaf's sorted position in the last full tree build
  int* pfound;       // per shape: pairs its traversal found (own + those of a static partner);
                     // for a static shape, the fill cursor of its segment
  int* poff;         // exclusive scan
  int* poff_prev;    // the last broadphase's poff: shape s's segment of the current contact list is
                     // [poff_prev[s], poff_prev[s + 1]) (the contact list is that broadphase's pair list)
  int idle;          // 1: the current contact list is the last broadphase's pair list and no body,
                     // shape, joint or contact was uploaded since; a world that is not active then
                     // keeps its pairs, its boxes and its part of the tree (leaf order, nodes, table)
  int segs;          // 1: the current contact list is the last broadphase's pair list and the
                     // layout is world after world, so a world's contacts are [poff_prev[first shape],
                     // poff_prev[one past its last shape])
  uint64_t* skey;    // per shape: its Morton key at its world's last tree build
  uint64_t* xkey;    // second buffer of the active shapes' key sort (the first is mkey[1])
  uint32_t* xval;
  int* wsort;        // per world: its shape count if active, scanned into the first position of its
                     // keys among the active shapes' keys
  int* scan_tiles;
  int* cnt;
  // sleep (b2Body e_awakeFlag / m_sleepTime, b2_island.cpp:444-483)
  int sleep_on;       // this step applies the awake flags (a world allows sleeping, or a body sleeps)
  int* bsflag;        // per body: bit0 asleep, bit1 not allowed to sleep
  float* bstime;      // per body: sleep time
  int* bwake;         // per body: this step's Collide changed a touching state of it
  int* bactive;       // per body: in an island with an awake body this step
  int* bncontact;     // per body: contacts in the list during the solve (orphans have none)
  int* bjointed;      // per body: a joint attaches it
  int* isl_awake;     // per island root: the island holds an awake body
  unsigned* isl_min;  // per island root: least sleep time of its bodies (float bits, >= 0)
  int* isl_cons;      // per island root: the island has constraints this step
  // the cells solver of large worlds (b2g_partition.h)
  int part_on;        // this step solves by cells: contact_init leaves the colour tables alone
  float4* part_geo;   // 1: (lo.x, lo.y, 1/L, L) of the cells
  int* part_boff;     // 2 * kCells + 1: offsets of the cells' body lists (A cells, then B cells)
  int* part_bcur;     // 2 * kCells: fill cursors of those lists
  // the grid pair search (b2g_grid.h)
  int grid_T;         // buckets (a power of two)
  unsigned* grid_acc; // 1: ordered-float maximum reach of the non-static shapes
  int* gstatic;       // the static shapes, world after world
  int* gsoff;         // nw + 1: each world's range of gstatic
  int* gcount;        // grid_T + 1: shapes per bucket (back to 0 after the scatter)
  int* goff;          // grid_T + 1: bucket offsets, then the binned shape count
  int* gids;          // per binned shape, by bucket
  float4* gbox;       // their AABBs, in the same order
  int4* gfast;        // their pair-filter words, in the same order
  int* gworld;        // their worlds, in the same order
  int* gltab;         // nw + 1: each world's own bucket table in gloff (batches of small worlds)
  int* gloff;         // those tables: world w's T + 1 entries (T = gltab[w + 1] - gltab[w] - 1, a
                      // power of two) are bucket offsets into gids, then the end of its entries
  unsigned* part_acc; // 8: bounds of the dynamic bodies' centres (ordered floats), their count, -,
                      // and (words 6-7, 64 bits) the fixed-point sum of their footprints
  unsigned* part_sync;  // 4: finished-workgroup counters of the launches whose last workgroup goes on
  int* part_blist;    // 2 * nb: bodies of each cell
  int2* part_brank;   // per body: its slot in its A cell and in its B cell (-1: not dynamic)
  int* part_ccnt;     // kPartGroups: constraints per group (A cells, B cells, cross)
  int* part_bucket;   // 2 * kCells * kCellCons: compacted constraints of each cell
  int* part_cbucket;  // the cross group's compacted constraints
  int* part_coff;     // kPartGroups * (kColourSlots + 1): slot offsets of each group's colours
  unsigned long long* part_claim;  // per body: colouring claims of the cross group
  unsigned long long* part_used;   // per body: colours the cross group gave its constraints
  int* part_xcol;     // per cross-group constraint: its colour while colouring
  int2* sloc;         // per slot: LDS slots of its two bodies in its cell (-1: global memory)
};

// grid-stride helpers with an XCD-aware block order: consecutive work tiles
// go to the same XCD (blocks are dispatched round-robin over 8 XCDs), so
// neighbouring constraints/bodies share an L2.
__device__ __forceinline__ int xcd_block() {
  const int g = (int)gridDim.x;
  if (g < 16 || (g & 7)) return (int)blockIdx.x;
  const int per = g >> 3;
  return ((int)blockIdx.x & 7) * per + ((int)blockIdx.x >> 3);
}
#define B2G_FOR(i, begin, end)                                                     \
  for (int i = (begin) + xcd_block() * (int)blockDim.x + (int)threadIdx.x; i < (end); \
       i += (int)(gridDim.x * blockDim.x))

// Same iteration space as B2G_FOR, but every lane of a block stays in the
// loop until the block is done (lanes past `end` run with valid == false), so
// wave-wide shuffles inside the body see all 64 lanes.
#define B2G_FOR_ALL(i, valid, begin, end)                                                     \
  for (int base_##i = (begin) + xcd_block() * (int)blockDim.x; base_##i < (end);              \
       base_##i += (int)(gridDim.x * blockDim.x))                                             \
    for (int i = base_##i + (int)threadIdx.x, valid = (i < (end)); valid >= 0; valid = -1)

// one atomic per wave for a counter update from any subset of active lanes
__device__ __forceinline__ void wave_atomic_add(int* addr, int per_lane) {
  const uint64_t act = __ballot(1);
  const int leader = __ffsll((long long)act) - 1;
  if (lane_id() == leader) atomicAdd(addr, per_lane * __popcll(act));
}

// slot of this lane in an append to a device list, from the calling lanes
// (one atomic per wave); lanes get consecutive slots in lane order
__device__ __forceinline__ int wave_append(int* counter) {
  const uint64_t act = __ballot(1);
  const int leader = __ffsll((long long)act) - 1;
  int base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, __popcll(act));
  base = __shfl(base, leader, kWave);
  return base + __popcll(act & lanemask_lt());
}

// atomicMax of v into base[key] for the valid lanes, one atomic per wave when
// all valid lanes share the key. All 64 lanes must call this.
__device__ __forceinline__ void wave_atomic_max_int_keyed(int* base, int key, bool valid, int v) {
  const uint64_t act = __ballot(valid);
  if (act == 0) return;
  const int first = __ffsll((long long)act) - 1;
  const int key0 = __shfl(key, first, kWave);
  if (__ballot(valid && key != key0) == 0) {
    int m = valid ? v : INT_MIN;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) m = max(m, __shfl_xor(m, off, kWave));
    if (lane_id() == first) atomicMax(base + key0, m);
  } else if (valid) {
    atomicMax(base + key, v);
  }
}

__device__ __forceinline__ float wave_min_all(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off, kWave));
  return v;
}
__device__ __forceinline__ float wave_max_all(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, kWave));
  return v;
}

__device__ __forceinline__ vec2 v2of(float4 f) { return mk(f.x, f.y); }
__device__ __forceinline__ bool world_active(const Dev& d, int w) { return (d.wflag[w] & 2) != 0; }

// one contact event: shapes, kind, the manifold as stored, the material
__device__ __forceinline__ void record_event(const Dev& d, int2 pr, int kind, const float4* man, float4 mat) {
  const int slot = wave_append(&d.cnt[C_EVENTS]);
  if (slot >= d.ecap) {
    atomicOr(&d.cnt[C_STATUS], S_EVENT_OVERFLOW);
    return;
  }
  float4* e = d.events + (size_t)kEventWords * slot;
  e[0] = make_float4(__int_as_float(pr.x), __int_as_float(pr.y), __int_as_float(kind), 0.0f);
  e[1] = man[0];
  e[2] = man[1];
  e[3] = man[2];
  e[4] = man[3];
  e[5] = mat;
}

__global__ void k_events_reset(Dev d) {
  if (threadIdx.x == 0 && blockIdx.x == 0) d.cnt[C_EVENTS] = 0;
}

// ===========================================================================
// Velocity integration (b2_island.cpp:257-298, SolveOrphan :187-206)
// ===========================================================================
// the step's colour counters (block 0 of the launch that integrates velocities)
__device__ __forceinline__ void reset_colour_counters(const Dev& d) {
  for (int c = threadIdx.x; c < kColourSlots; c += blockDim.x) d.col_count[c] = 0;
  if (threadIdx.x == 0) {
    d.cnt[C_REMAIN] = 0;
    d.cnt[C_PENDING] = 0;
    d.cnt[C_LEFT] = 0;
    d.cnt[C_NCOLOURS] = 0;
    d.cnt[C_SERIAL] = 0;
  }
}
constexpr int kBodyAsleep = 1, kBodyNoAutoSleep = 2;
// a body that counts as awake for Collide: not static, awake at the start of
// the step (b2_contact_manager.cpp:100-106)
__device__ __forceinline__ bool body_awake0(const Dev& d, int b) {
  return d.btype[b] != kStaticBody && (d.bsflag[b] & kBodyAsleep) == 0;
}
__device__ __forceinline__ bool world_sleeps(const Dev& d, int w) { return (d.wflag[w] & 4) != 0; }

// integrate = 1: the step's per-body resets and the velocity integration;
// 0: the resets only (the integration follows the islands, which decide which
// bodies move when sleeping applies)
__device__ __forceinline__ void integrate_velocity(const Dev& d, int b, int integrate = 1) {
  {
    // a body of an idle world keeps its island label and colour mask (stale;
    // every reader of them checks the world first)
    const int w = d.bworld[b];
    if (!world_active(d, w)) return;
    if (integrate != 2) {
      d.parent[b] = b;
      d.bmask[b] = 0ull;
    }
    if (d.btype[b] != kDynamicBody || integrate == 0) return;
    const float4 wp = d.wpar[w];
    const float h = wp.z;
    const float4 v4 = d.vel[b], p = d.prop[b], p2 = d.prop2[b], f = d.force[b];
    // v += h * invMass * (gravityScale * mass * gravity + force)
    const float s1 = h * p.x;
    const float s2 = p2.y * p2.x;
    float vx = v4.x + s1 * (s2 * wp.x + f.x);
    float vy = v4.y + s1 * (s2 * wp.y + f.y);
    float w_ = v4.z + h * p.y * f.z;
    const float ld = 1.0f + h * p2.z;
    const float ad = 1.0f + h * p2.w;
    vx = vx / ld;
    vy = vy / ld;
    w_ = w_ / ad;
    d.vel[b] = make_float4(vx, vy, w_, 0.0f);
  }
}

// ===========================================================================
// Narrowphase: b2ContactManager::Collide + b2Contact::Update
// (src/dynamics/b2_contact_manager.cpp:66-116, src/dynamics/b2_contact.cpp:126-210)
// ===========================================================================
// Contact i's Collide step (b2Contact::Update)
__device__ __forceinline__ void narrow_contact(const Dev& d, int i) {
  const int2 pr = d.cpair[i];
  const int4 ha = d.shdr[pr.x], hb = d.shdr[pr.y];
  int solve = 0;
  if (world_active(d, ha.w) && d.sleep_on) {
    // b2Body::GetContactCount for the orphan test, and a contact between
    // two sleeping (or static) bodies keeps its manifold and touching state
    atomicAdd(&d.bncontact[ha.y], 1);
    atomicAdd(&d.bncontact[hb.y], 1);
    if (!body_awake0(d, ha.y) && !body_awake0(d, hb.y)) {
      float4* mp = d.cman + 4 * (size_t)i;
      int* packed = reinterpret_cast<int*>(mp + 3) + 2;
      const int w0 = *packed;
      *packed = w0 & 0xffff;  // solved by no step since
      d.touch[i] = ((w0 >> 8) & 0xff) > 0 ? 1 : 0;
      return;
    }
  }
  if (world_active(d, ha.w)) {
    Geo ga, gb;
    load_geo(ga, ha.x, ha.z, d.geo + 9 * (size_t)pr.x);
    load_geo(gb, hb.x, hb.z, d.geo + 9 * (size_t)pr.y);
    const xform xa = xform_of(d.xf[ha.y]);
    const xform xb = xform_of(d.xf[hb.y]);
    float4* mp = d.cman + 4 * (size_t)i;
    const float4 o0 = mp[0], o1 = mp[1], o3 = mp[3];
    const int old_count = (as_int(o3.z) >> 8) & 0xff;
    Man m;
    m.count = 0;
    m.type = (as_int(o3.z) & 0xff);
    const float4 o2 = mp[2];
    m.ln = mk(o2.x, o2.y);
    m.lp = mk(o2.z, o2.w);
    m.lp0 = mk(0.0f, 0.0f);
    m.lp1 = mk(0.0f, 0.0f);
    m.id0 = m.id1 = 0;
    evaluate(m, ga, xa, gb, xb);
    // match old feature keys to carry the impulses (b2_contact.cpp:158-181)
    float ni0 = 0.0f, ti0 = 0.0f, ni1 = 0.0f, ti1 = 0.0f;
    const uint32_t oid0 = as_uint(o3.x), oid1 = as_uint(o3.y);
    if (m.count > 0) {
      if (old_count > 0 && oid0 == m.id0) { ni0 = o0.z; ti0 = o0.w; }
      else if (old_count > 1 && oid1 == m.id0) { ni0 = o1.z; ti0 = o1.w; }
    }
    if (m.count > 1) {
      if (old_count > 0 && oid0 == m.id1) { ni1 = o0.z; ti1 = o0.w; }
      else if (old_count > 1 && oid1 == m.id1) { ni1 = o1.z; ti1 = o1.w; }
    }
    mp[0] = make_float4(m.lp0.x, m.lp0.y, ni0, ti0);
    mp[1] = make_float4(m.count > 1 ? m.lp1.x : 0.0f, m.count > 1 ? m.lp1.y : 0.0f, ni1, ti1);
    mp[2] = make_float4(m.ln.x, m.ln.y, m.lp.x, m.lp.y);
    mp[3] = make_float4(uint_as_float(m.count > 0 ? m.id0 : 0u), uint_as_float(m.count > 1 ? m.id1 : 0u),
                        as_float((m.type & 0xff) | (m.count << 8)), 0.0f);
    solve = m.count > 0 ? 1 : 0;
    // a touching transition wakes both bodies (b2_contact.cpp:183-188); the
    // wake-up applies after Collide, when the islands are seeded
    if (d.sleep_on && ((old_count > 0) != (m.count > 0))) {
      d.bwake[ha.y] = 1;
      d.bwake[hb.y] = 1;
    }
    // BeginContact / EndContact on touching transitions (b2_contact.cpp:197-203)
    if (d.events_on && ((old_count > 0) != (m.count > 0))) {
      const float4 man[4] = {mp[0], mp[1], mp[2], mp[3]};
      record_event(d, pr, m.count > 0 ? kEventBegin : kEventEnd, man, d.cmat[i]);
    }
    // a contact that is not solved this step gives up its colour: colours
    // kept across steps are then always a subset of the previous step's
    // valid colouring
    if (!solve) d.ccol[i] = -1;
  }
  d.touch[i] = solve;
}

// A block's lanes take its tile of kBlock contacts sorted by the pair's shape
// types (a counting sort in LDS), so a wave mostly runs one collide function
// instead of every one its contacts need, each masked; which lane takes a
// contact changes nothing in its result. The same launch integrates the
// bodies' velocities (b2Island::Solve, b2_island.cpp:257-298), which touches
// none of the contact data, and clears the step's colour counters. A step
// with dt <= 0 runs Collide alone (integrate = 0).
__global__ void __launch_bounds__(kBlock) k_narrowphase(Dev d, int integrate) {
  __shared__ int s_cnt[17];  // 9 ordered type pairs; [16]: the tile's contacts of active worlds
  __shared__ int s_perm[kBlock];
  if (blockIdx.x == 0) reset_colour_counters(d);
  const int n = d.cnt[C_CONTACTS];
  B2G_FOR_ALL(i, valid, 0, n) {
    if (threadIdx.x < 17) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    int key = -1, rank = 0;
    if (valid) {
      const int2 pr = d.cpair[i];
      const int4 ha = d.shdr[pr.x];
      if (world_active(d, ha.w)) {
        key = 3 * ha.x + d.shdr[pr.y].x;  // kCircle 0, kEdge 1, kPolygon 2
        rank = atomicAdd(&s_cnt[key], 1);
      } else {
        d.touch[i] = 0;  // a world that sits out the step: not solved, manifold kept
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int run = 0;
      for (int k = 0; k < 9; ++k) {
        const int c = s_cnt[k];
        s_cnt[k] = run;
        run += c;
      }
      s_cnt[16] = run;
    }
    __syncthreads();
    if (key >= 0) s_perm[s_cnt[key] + rank] = i;
    __syncthreads();
    if ((int)threadIdx.x < s_cnt[16]) narrow_contact(d, s_perm[threadIdx.x]);
    __syncthreads();  // s_cnt and s_perm are the next tile's
  }
  if (integrate) B2G_FOR(b, 0, d.nb) integrate_velocity(d, b, d.sleep_on ? 0 : 1);
}

// compacted list of solved contacts (order preserved)
__device__ __forceinline__ void compact_touching(const Dev& d, int i, int t) {
  d.clist[t] = i;
  const int2 pr = d.cpair[i];
  const int ba = d.shdr[pr.x].y, bb = d.shdr[pr.y].y;
  d.cbody[t] = make_int2(d.btype[ba] == kStaticBody ? ~ba : ba, d.btype[bb] == kStaticBody ? ~bb : bb);
}
__global__ void __launch_bounds__(kBlock) k_compact_touching(Dev d) {
  const int n = d.cnt[C_CONTACTS];
  B2G_FOR(i, 0, n) {
    if (d.touch[i]) compact_touching(d, i, d.tpos[i]);
  }
}

// The scan's apply pass over the touching flags (tile offsets folded in, as
// k_scan_apply_fold) with the compaction in the same launch: a solved
// contact's place in the compacted list is its own scan value.
__global__ void __launch_bounds__(kBlock) k_touch_scan_apply(Dev d, const int* __restrict__ tile_sums) {
  __shared__ int lds[kBlock / kWave];
  const int n = d.cnt[C_CONTACTS];
  const int tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles == 0 && blockIdx.x == 0 && threadIdx.x == 0) d.cnt[C_CONS] = 0;
  for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
    int part = 0;
    for (int u = threadIdx.x; u < t; u += blockDim.x) part += tile_sums[u];
    int offset;
    (void)block_exclusive_scan(part, lds, &offset);
    const int base = t * kScanTile + threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      v[k] = base + k < n ? d.touch[base + k] : 0;
      s += v[k];
    }
    int total;
    int ex = block_exclusive_scan(s, lds, &total) + offset;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      if (base + k < n) {
        d.tpos[base + k] = ex;
        if (v[k]) compact_touching(d, base + k, ex);
      }
      ex += v[k];
    }
    if (t == tiles - 1 && threadIdx.x == 0) d.cnt[C_CONS] = offset + total;
  }
}

// ===========================================================================
// Graph colouring (Jones-Plassmann rounds with per-body colour masks)
// Contacts keep their colour from the previous step while they stay
// touching; only newly touching contacts are coloured, each taking the lowest
// colour free at both of its dynamic bodies.
// ===========================================================================
__device__ __forceinline__ bool is_dyn(const Dev& d, int b) { return d.btype[b] == kDynamicBody; }

// Contacts that stay touching keep last step's colour if it is below the
// colour limit `K` of this step; the rest are coloured by claim rounds.
__global__ void __launch_bounds__(kBlock) k_colour_init(Dev d, int K) {
  const int n = d.cnt[C_CONS];
  B2G_FOR(k, 0, n) {
    const int c = d.clist[k];
    const int2 bp = d.cbody[k];  // the bodies, without the contact -> shape -> body chain
    const int ba = bp.x >= 0 ? bp.x : ~bp.x, bb = bp.y >= 0 ? bp.y : ~bp.y;
    const int col = d.ccol[c];
    if (col >= 0 && col < K && !d.wrecolour[d.bworld[ba]]) {
      const unsigned long long bit = 1ull << col;
      if (is_dyn(d, ba)) atomicOr(&d.bmask[ba], bit);
      if (is_dyn(d, bb)) atomicOr(&d.bmask[bb], bit);
      d.ccolour[k] = col;
    } else {
      d.ccolour[k] = -1;
      wave_atomic_add(&d.cnt[C_REMAIN], 1);
      d.cpend[wave_append(&d.cnt[C_PENDING])] = k;
    }
  }
}

// Colouring priority of a contact: the claim round in the top 16 bits and a
// pseudo-random, collision-free 48-bit rank below it. The rank is a bijection
// of the contact's world-local shape pair (24 + 24 bits), so two contacts
// never tie, and a world is coloured the same whether it is stepped alone or
// inside a batch.
__device__ __forceinline__ unsigned long long mix48(unsigned long long x) {
  constexpr unsigned long long kMask = (1ull << 48) - 1;
  x = (x * 0x9E3779B97F4Bull) & kMask;  // odd multiplier: bijective mod 2^48
  x ^= x >> 23;                          // xorshift: bijective on 48 bits
  x = (x * 0xD6E8FEB86659ull) & kMask;   // odd
  x ^= x >> 21;
  return x;
}
__device__ __forceinline__ unsigned long long colour_priority(const Dev& d, int round, int2 pr) {
  const int base = d.wshape0[d.shdr[pr.x].w];
  const unsigned long long local = ((unsigned long long)(uint32_t)(pr.x - base) << 24) |
                                   (unsigned long long)(uint32_t)(pr.y - base);
  return ((unsigned long long)(round & 0xffff) << 48) | mix48(local);
}

__global__ void __launch_bounds__(kBlock) k_colour_claim(Dev d, int local_round) {
  if (d.cnt[C_REMAIN] == 0) return;
  const int round = d.cnt[C_ROUND] + local_round;
  const int n = d.cnt[C_PENDING];
  B2G_FOR(q, 0, n) {
    const int k = d.cpend[q];
    if (d.ccolour[k] >= 0) continue;
    const int c = d.clist[k];
    const int2 pr = d.cpair[c];
    const int ba = d.shdr[pr.x].y, bb = d.shdr[pr.y].y;
    const unsigned long long v = colour_priority(d, round, pr);
    if (is_dyn(d, ba)) atomicMax(&d.bslot[ba], v);
    if (is_dyn(d, bb)) atomicMax(&d.bslot[bb], v);
  }
}

// `collect`: the round's losers also go to the list k_colour_finish walks
// (the last device-wide round before it), so its rounds do not walk the
// whole pending list again
__global__ void __launch_bounds__(kBlock) k_colour_resolve(Dev d, int local_round, int K, int collect) {
  if (d.cnt[C_REMAIN] == 0) return;
  const int round = d.cnt[C_ROUND] + local_round;
  const int n = d.cnt[C_PENDING];
  B2G_FOR(q, 0, n) {
    const int k = d.cpend[q];
    if (d.ccolour[k] >= 0) continue;
    const int c = d.clist[k];
    const int2 pr = d.cpair[c];
    const int ba = d.shdr[pr.x].y, bb = d.shdr[pr.y].y;
    const unsigned long long v = colour_priority(d, round, pr);
    const bool da = is_dyn(d, ba), db = is_dyn(d, bb);
    const bool win = (!da || d.bslot[ba] == v) && (!db || d.bslot[bb] == v);
    if (!win) {
      if (collect) d.cleft[wave_append(&d.cnt[C_LEFT])] = k;
      continue;
    }
    const unsigned long long used = (da ? d.bmask[ba] : 0ull) | (db ? d.bmask[bb] : 0ull);
    int col = kSerialColour;
    if (~used != 0ull) col = __ffsll((long long)~used) - 1;
    if (col >= K) col = kSerialColour;
    d.ccolour[k] = col;
    if (col < kMaxColours) {
      const unsigned long long bit = 1ull << col;
      if (da) atomicOr(&d.bmask[ba], bit);
      if (db) atomicOr(&d.bmask[bb], bit);
    }
    wave_atomic_add(&d.cnt[C_REMAIN], -1);
  }
}

__global__ void __launch_bounds__(kBlock) k_colour_count(Dev d) {
  __shared__ int hist[kColourSlots];
  for (int c = threadIdx.x; c < kColourSlots; c += blockDim.x) hist[c] = 0;
  if (blockIdx.x == 0)  // k_colour_scatter's placement cursors start at zero
    for (int c = threadIdx.x; c < kColourSlots; c += blockDim.x) d.col_cursor[c] = 0;
  __syncthreads();
  const int n = d.cnt[C_CONS];
  const bool per_world = d.nw > 1;  // one world: k_colour_scatter reads its top colour off the histogram
  B2G_FOR_ALL(k, valid, 0, n) {
    const int col = valid ? d.ccolour[k] : -1;  // < 0: still uncoloured, the host runs more rounds
    int w = 0;
    if (col >= 0) {
      atomicAdd(&hist[col], 1);
      const int c = d.clist[k];
      d.ccol[c] = col;
      if (per_world) w = d.bworld[d.cbody[k].x >= 0 ? d.cbody[k].x : ~d.cbody[k].x];
    }
    if (per_world) wave_atomic_max_int_keyed(d.wcolmax, w, col >= 0 && col < kMaxColours, col);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < kColourSlots; c += blockDim.x)
    if (hist[c]) atomicAdd(&d.col_count[c], hist[c]);
}

// Before the 16-bit claim-round tag wraps: empty the claim slots and restart
// the round base (fill kernel, so the step stays capturable).
__global__ void __launch_bounds__(kBlock) k_reset_colour_rounds(Dev d) {
  B2G_FOR(b, 0, d.nb) d.bslot[b] = 0ull;
  if (blockIdx.x == 0 && threadIdx.x == 0) d.cnt[C_ROUND] = 0;
}

__global__ void k_clear_pairs(Dev d) {
  if (blockIdx.x == 0 && threadIdx.x == 0) d.cnt[C_PAIRS] = 0;
}

// Per-world colouring upkeep, after a step's colouring is complete. Kept
// colours fragment as contacts come and go (a new constraint sees coloured
// neighbourhoods at both of its bodies); when a world uses kRecolourSlack
// more colours than its last colouring from scratch, its next step colours
// every constraint afresh. Each world decides from its own steps only.
constexpr int kRecolourSlack = 3;
// `used`: the colours the world used this step (its highest normal colour + 1)
__device__ __forceinline__ void world_colour_stats(const Dev& d, int w, int used) {
  d.wcolmax[w] = -1;
  if (d.wrecolour[w]) {
    d.wfresh[w] = used > 0 ? used : 1;
    d.wrecolour[w] = 0;
  } else if (used >= d.wfresh[w] + kRecolourSlack) {
    d.wrecolour[w] = 1;
  }
}

// The claim rounds of a captured step that the multi-workgroup rounds left
// unfinished, in one workgroup (barriers between the claim and resolve
// passes instead of launches), up to `rounds` more rounds: the same claims,
// priorities and colour choice as k_colour_claim / k_colour_resolve, over the
// list the last (collecting) device-wide round left. Which lane takes which
// constraint changes nothing: the winners of a round share no dynamic body.
// The claim words are read with atomic loads: other waves wrote them with
// atomics. Launched with one workgroup of kFinishThreads threads. With more
// than `max_left` constraints left it does nothing (a world colouring afresh
// is faster in device-wide rounds, which the eager host loop then runs).
constexpr int kFinishThreads = 1024;
__global__ void __launch_bounds__(kFinishThreads) k_colour_finish(Dev d, int local_round0, int rounds, int K,
                                                                  int max_left) {
  const int left0 = d.cnt[C_REMAIN];
  if (left0 == 0 || left0 > max_left) return;
  const int n = d.cnt[C_LEFT];
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  for (int r = 0; r < rounds; ++r) {
    const int round = d.cnt[C_ROUND] + local_round0 + r;
    for (int q = tid; q < n; q += nt) {
      const int k = d.cleft[q];
      if (d.ccolour[k] >= 0) continue;
      const int2 pr = d.cpair[d.clist[k]];
      const int ba = d.shdr[pr.x].y, bb = d.shdr[pr.y].y;
      const unsigned long long v = colour_priority(d, round, pr);
      if (is_dyn(d, ba)) atomicMax(&d.bslot[ba], v);
      if (is_dyn(d, bb)) atomicMax(&d.bslot[bb], v);
    }
    __syncthreads();
    int left = 0;
    for (int q = tid; q < n; q += nt) {
      const int k = d.cleft[q];
      if (d.ccolour[k] >= 0) continue;
      const int2 pr = d.cpair[d.clist[k]];
      const int ba = d.shdr[pr.x].y, bb = d.shdr[pr.y].y;
      const unsigned long long v = colour_priority(d, round, pr);
      const bool da = is_dyn(d, ba), db = is_dyn(d, bb);
      const bool win =
          (!da || __hip_atomic_load(&d.bslot[ba], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == v) &&
          (!db || __hip_atomic_load(&d.bslot[bb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == v);
      if (!win) {
        left = 1;
        continue;
      }
      const unsigned long long used =
          (da ? __hip_atomic_load(&d.bmask[ba], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0ull) |
          (db ? __hip_atomic_load(&d.bmask[bb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0ull);
      int col = kSerialColour;
      if (~used != 0ull) col = __ffsll((long long)~used) - 1;
      if (col >= K) col = kSerialColour;
      d.ccolour[k] = col;
      if (col < kMaxColours) {
        const unsigned long long bit = 1ull << col;
        if (da) atomicOr(&d.bmask[ba], bit);
        if (db) atomicOr(&d.bmask[bb], bit);
      }
      atomicAdd(&d.cnt[C_REMAIN], -1);
    }
    if (!__syncthreads_or(left)) break;
  }
}

// Constraints still uncoloured after the rounds a step allows go to the serial
// group (solved on one lane after the parallel colours: still a valid
// Gauss-Seidel order).
__global__ void __launch_bounds__(kBlock) k_colour_leftover(Dev d) {
  if (d.cnt[C_REMAIN] == 0) return;
  const int n = d.cnt[C_PENDING];
  B2G_FOR(q, 0, n) {
    const int k = d.cpend[q];
    if (d.ccolour[k] >= 0) continue;
    d.ccolour[k] = kSerialColour;
  }
}

// one wave: offsets of the colour groups (lane c scans colour c, lane 0 also
// the serial slot); resets the scatter cursors and advances the claim-round
// base past the rounds just run

// Each block reserves a contiguous range per colour with one atomic, then
// places its constraints inside it (order inside a colour does not matter:
// those constraints share no dynamic body). Every block first scans the
// colour counts into offsets itself (one wave, 65 entries); block 0 also
// publishes them with the colour counters.
__global__ void __launch_bounds__(kBlock) k_colour_scatter(Dev d, int rounds_run) {
  static_assert(kColourSlots == kWave + 1, "one lane per normal colour plus the serial slot");
  __shared__ int hist[kColourSlots];
  __shared__ int base[kColourSlots];
  __shared__ int off[kColourSlots + 1];
  if (threadIdx.x < kWave) {
    const int lane = (int)threadIdx.x;
    const int cnt = d.col_count[lane];
    const int inc = wave_inclusive_scan(cnt);
    off[lane] = inc - cnt;
    const uint64_t used = __ballot(cnt > 0);
    const int top = used ? 64 - __clzll((long long)used) : 0;
    const int serial = d.col_count[kSerialColour];
    if (lane == kWave - 1) {
      off[kSerialColour] = inc;
      off[kSerialColour + 1] = inc + serial;
    }
    if (blockIdx.x == 0) {
      d.col_off[lane] = inc - cnt;
      if (lane == kWave - 1) {
        d.col_off[kSerialColour] = inc;
        d.col_off[kSerialColour + 1] = inc + serial;
        d.cnt[C_NCOLOURS] = serial > 0 ? kSerialColour + 1 : top;
        d.cnt[C_SERIAL] = serial;
        d.cnt[C_ROUND] += rounds_run;
        if (d.nw == 1) d.wcolmax[0] = top - 1;  // the one world's highest normal colour
      }
    }
  }
  __syncthreads();
  const int n = d.cnt[C_CONS];
  for (int start = xcd_block() * kBlock; start < n; start += gridDim.x * kBlock) {
    for (int c = threadIdx.x; c < kColourSlots; c += blockDim.x) hist[c] = 0;
    __syncthreads();
    const int k = start + (int)threadIdx.x;
    const int col = k < n ? d.ccolour[k] : -1;
    const int local = col >= 0 ? atomicAdd(&hist[col], 1) : 0;
    __syncthreads();
    for (int c = threadIdx.x; c < kColourSlots; c += blockDim.x)
      base[c] = hist[c] ? off[c] + atomicAdd(&d.col_cursor[c], hist[c]) : 0;
    __syncthreads();
    if (col >= 0) d.csorted[base[col] + local] = d.clist[k];
    __syncthreads();
  }
}

// the serial group runs on one lane; order it by contact index so its
// Gauss-Seidel sequence is deterministic
__global__ void k_sort_serial_group(Dev d) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int b = d.col_off[kSerialColour], e = d.col_off[kSerialColour + 1];
  for (int a = b + 1; a < e; ++a) {
    const int v = d.csorted[a];
    int q = a - 1;
    while (q >= b && d.csorted[q] > v) {
      d.csorted[q + 1] = d.csorted[q];
      --q;
    }
    d.csorted[q + 1] = v;
  }
}

__global__ void __launch_bounds__(kBlock) k_colour_clear_counts(Dev d) {
  for (int c = threadIdx.x; c < kColourSlots; c += blockDim.x) d.col_count[c] = 0;
}

// ===========================================================================
// Islands: connected components over non-static bodies (b2World::Solve DFS,
// src/dynamics/b2_world.cpp:526-659), by lock-free union-find.
// ===========================================================================
// Parent links always point to a smaller index (the union links the larger
// root under the smaller), so every walk below strictly decreases and ends.
// Loads are agent-scope atomics: a root linked by another CU (possibly on
// another XCD) is seen at the next read, not served from a stale cache line.
__device__ __forceinline__ int uf_load(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// find with path halving (each visited node skips to its grandparent, still
// an ancestor in the same set)
__device__ __forceinline__ int uf_find(int* parent, int x) {
  for (;;) {
    const int p = uf_load(parent + x);
    if (p == x) return x;
    const int gp = uf_load(parent + p);
    if (gp == p) return p;
    __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = gp;
  }
}

__device__ __forceinline__ void uf_union(int* parent, int a, int b) {
  for (;;) {
    a = uf_find(parent, a);
    b = uf_find(parent, b);
    if (a == b) return;
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    if (atomicCAS(&parent[hi], hi, lo) == hi) return;
  }
}

__global__ void __launch_bounds__(kBlock) k_island_link(Dev d) {
  const int n = d.cnt[C_CONS];
  B2G_FOR(k, 0, n + d.nj) {
    if (k < n) {
      const int2 bp = d.cbody[k];
      if ((bp.x | bp.y) >= 0) uf_union(d.parent, bp.x, bp.y);
    } else {
      const Joint& j = d.joints[k - n];
      if (!world_active(d, j.world)) continue;
      if (d.btype[j.a] != kStaticBody && d.btype[j.b] != kStaticBody) uf_union(d.parent, j.a, j.b);
    }
  }
}

// read-only walk to the root (no halving): while the labels below are being
// written, a halving store could replace a finished label with a non-root
// ancestor
__device__ __forceinline__ int uf_root(const int* parent, int x) {
  for (;;) {
    const int p = uf_load(parent + x);
    if (p == x) return x;
    x = p;
  }
}

__global__ void __launch_bounds__(kBlock) k_island_label(Dev d) {
  B2G_FOR(b, 0, d.nb) {
    if (d.btype[b] != kStaticBody)
      __hip_atomic_store(d.parent + b, uf_root(d.parent, b), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// first index of a sorted array whose value is >= key
__device__ __forceinline__ int lower_bound_u64(const uint64_t* a, int n, uint64_t key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// World w's contacts in the current list, [first, end): two segment offsets
// while d.segs is set, else two binary searches of the sorted keys (the
// first contact whose lower shape is the world's first shape, and the first
// past its last shape).
__device__ __forceinline__ int2 world_contacts(const Dev& d, int w) {
  const int nc = d.cnt[C_CONTACTS];
  const int s0 = d.wshape0[w], s1 = d.wshape1[w];
  if (d.segs) {
    const int c0 = s0 < d.ns ? min(d.poff_prev[s0], nc) : nc;
    const int c1 = s1 < d.ns ? min(d.poff_prev[s1], nc) : nc;
    return make_int2(c0, c1);
  }
  return make_int2(lower_bound_u64(d.ckey, nc, (uint64_t)(uint32_t)s0 << 32),
                   lower_bound_u64(d.ckey, nc, (uint64_t)(uint32_t)s1 << 32));
}

// Islands of one world per workgroup, with the union-find forest in LDS
// (kCap >= the world's body count). Same forest rule as k_island_link (the
// larger root hooks under the smaller), so every island is labelled by its
// smallest body index, as there. The world's solved contacts are contiguous
// in the compacted list (they are those of its shapes, whose keys are
// contiguous in the sorted contact list); its joints come from wj_list.
__device__ __forceinline__ int lds_load(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ int lds_find(int* par, int x) {
  for (;;) {
    const int p = lds_load(par + x);
    if (p == x) return x;
    const int gp = lds_load(par + p);
    if (gp == p) return p;
    __hip_atomic_store(par + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    x = gp;
  }
}
__device__ __forceinline__ void lds_union(int* par, int a, int b) {
  for (;;) {
    a = lds_find(par, a);
    b = lds_find(par, b);
    if (a == b) return;
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    if (atomicCAS(&par[hi], hi, lo) == hi) return;
  }
}

template <int kCap>
__global__ void __launch_bounds__(1024) k_world_islands(Dev d) {
  __shared__ int par[kCap];
  __shared__ int range[2];
  const int w = d.worder[blockIdx.x];
  if (!world_active(d, w)) return;  // the whole workgroup leaves
  const int2 wb = d.wbody[w];
  const int b0 = wb.x, nbw = wb.y - wb.x;
  if (threadIdx.x == 0) {
    const int nc = d.cnt[C_CONTACTS], total = d.cnt[C_CONS];
    const int2 cr = world_contacts(d, w);
    range[0] = cr.x < nc ? d.tpos[cr.x] : total;
    range[1] = cr.y < nc ? d.tpos[cr.y] : total;
  }
  for (int i = threadIdx.x; i < nbw; i += blockDim.x) par[i] = i;
  __syncthreads();
  const int t0 = range[0], t1 = range[1];
  constexpr int kBatch = 8;  // body pairs loaded ahead of the unions (independent loads in flight)
  const int stride = (int)blockDim.x;
  for (int k0 = t0 + (int)threadIdx.x; k0 < t1; k0 += kBatch * stride) {
    int2 bp[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int k = k0 + u * stride;
      bp[u] = k < t1 ? d.cbody[k] : make_int2(-1, -1);
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const unsigned la = (unsigned)(bp[u].x - b0), lb = (unsigned)(bp[u].y - b0);
      if ((bp[u].x | bp[u].y) >= 0 && la < (unsigned)nbw && lb < (unsigned)nbw) lds_union(par, (int)la, (int)lb);
    }
  }
  for (int q = d.wj_off[w] + (int)threadIdx.x; q < d.wj_off[w + 1]; q += blockDim.x) {
    const Joint& j = d.joints[d.wj_list[q]];
    const unsigned la = (unsigned)(j.a - b0), lb = (unsigned)(j.b - b0);
    if (d.btype[j.a] != kStaticBody && d.btype[j.b] != kStaticBody && la < (unsigned)nbw && lb < (unsigned)nbw)
      lds_union(par, (int)la, (int)lb);
  }
  __syncthreads();
  // labels (a static body is never linked: its own root, as the step's reset left it)
  for (int i = threadIdx.x; i < nbw; i += blockDim.x) {
    int x = i;  // no writes to the forest remain: plain walk to the root
    for (int p = par[x]; p != x; p = par[x]) x = p;
    d.parent[b0 + i] = b0 + x;
  }
}

__device__ __forceinline__ int island_of(const Dev& d, int ba, int bb) {
  return d.btype[ba] != kStaticBody ? d.parent[ba] : d.parent[bb];
}

__device__ __forceinline__ void island_reset(const Dev& d, int isl) {
  d.solved[isl] = 0;
  d.sepbad[isl] = 0u;
  d.jok[isl] = 1;
}

// Per-island early exit (b2_island.cpp:403-408): an island stops after the
// first position iteration in which no contact separated by less than
// -3 linearSlop and every joint was within its slop. Each constraint of
// iteration `it` decides it from iteration it - 1's flags (and `solved`, set
// by the first constraint to find it), so no launch between the iterations
// is needed. The flags rotate over three slots: iteration it marks slot
// it % 3, reads slot (it - 1) % 3 and clears slot (it + 1) % 3 for the next.
__device__ __forceinline__ size_t island_slot(const Dev& d, int isl, int it) {
  return (size_t)(it % 3) * d.nb + isl;
}
__device__ __forceinline__ bool island_done(const Dev& d, int isl, int it) {
  // the three words are loaded together (one dependent round trip, not two)
  const size_t prev = island_slot(d, isl, it > 0 ? it - 1 : 0);
  const int solved = d.solved[isl];
  const unsigned bad = d.sepbad[prev];
  const int jok = d.jok[prev];
  if (solved) return true;
  if (it == 0) return false;
  if (bad == 0u && jok != 0) {
    d.solved[isl] = 1;
    return true;
  }
  return false;
}
__device__ __forceinline__ void island_rearm(const Dev& d, int isl, int it) {
  const size_t nxt = island_slot(d, isl, it + 1);
  d.sepbad[nxt] = 0u;
  d.jok[nxt] = 1;
}

// island of solved constraint slot k (k < n) or joint k - n, and that
// island's position state reset
__device__ __forceinline__ void island_assign(const Dev& d, int n, int k) {
  int isl;
  if (k < n) {
    const float4 v2i = d.vc[2][k];
    isl = island_of(d, as_int(v2i.x), as_int(v2i.y));
    d.cisland[k] = isl;
  } else {
    const int q = k - n;
    if (d.jisland[q] < 0) return;
    const Joint& j = d.joints[d.jorder[q]];
    isl = island_of(d, j.a, j.b);
    d.jisland[q] = isl;
  }
  island_reset(d, isl);
}

// ===========================================================================
// b2ContactSolver::Initialize + InitializeVelocityConstraints
// (src/dynamics/b2_contact_solver.cpp:65-258); one lane per solved contact,
// in colour-sorted slot order. Velocity data (10 float4) and position data
// (5 float4) are stored structure-of-arrays by slot.
//   vc0 normal.x normal.y friction tangentSpeed   vc1 mA mB iA iB
//   vc2 ia ib pointCount contact (int bits)         vc3/vc5 rA, rB of point 0/1
//   vc4/vc6 normalImpulse tangentImpulse normalMass tangentMass of point 0/1
//   vc7 bias0 bias1 K.ex.y K.ey.x                   vc8 normalMass (ex, ey)
//   vc9 K (ex, ey)
//   pc0 localPoint localNormal  pc1 localPoints[0..1]  pc2 localCenterA/B
//   pc3 mA mB iA iB             pc4 radiusA radiusB type pointCount
// ===========================================================================
// (6 waves per SIMD: 80 registers, measured slightly faster than 5)
__global__ void __launch_bounds__(kBlock, 6) k_contact_init(Dev d) {
  const int n = d.cnt[C_CONS];
  B2G_FOR(s, 0, n) {
    const int c = d.csorted[s];
    const int2 pr = d.cpair[c];
    const int4 ha = d.shdr[pr.x], hb = d.shdr[pr.y];
    const int ia = ha.y, ib = hb.y;
    const float4 pa = d.prop[ia], pb = d.prop[ib];
    const float mA = pa.x, iA = pa.y, mB = pb.x, iB = pb.y;
    const vec2 lcA = mk(pa.z, pa.w), lcB = mk(pb.z, pb.w);
    const float4 m0 = d.cman[4 * (size_t)c + 0], m1 = d.cman[4 * (size_t)c + 1];
    const float4 m2 = d.cman[4 * (size_t)c + 2], m3 = d.cman[4 * (size_t)c + 3];
    const int mtype = as_int(m3.z) & 0xff;
    const int count = (as_int(m3.z) >> 8) & 0xff;
    const float4 mat = d.cmat[c];
    const float rA = d.geo[9 * (size_t)pr.x].x, rB = d.geo[9 * (size_t)pr.y].x;
    const int w = ha.w;
    const float4 wp = d.wpar[w];
    const bool warm = (d.wflag[w] & 1) != 0;
    const float ratio = wp.w;

    float ni0 = 0.0f, ti0 = 0.0f, ni1 = 0.0f, ti1 = 0.0f;
    if (warm) {
      ni0 = ratio * m0.z;
      ti0 = ratio * m0.w;
      ni1 = ratio * m1.z;
      ti1 = ratio * m1.w;
    }
    const float4 posA = d.pos[ia], posB = d.pos[ib];
    const float4 velA = d.vel[ia], velB = d.vel[ib];
    const vec2 cA = mk(posA.x, posA.y), cB = mk(posB.x, posB.y);
    const vec2 vA = mk(velA.x, velA.y), vB = mk(velB.x, velB.y);
    const float wA = velA.z, wB = velB.z;
    xform xfA, xfB;
    xfA.q = rot_of(posA.z);
    xfB.q = rot_of(posB.z);
    xfA.p = cA - mul(xfA.q, lcA);
    xfB.p = cB - mul(xfB.q, lcB);
    // b2WorldManifold::Initialize (src/collision/b2_collision.cpp:26-90)
    vec2 normal, wp0 = mk(0.0f, 0.0f), wp1 = mk(0.0f, 0.0f);
    const vec2 ln = mk(m2.x, m2.y), lp = mk(m2.z, m2.w);
    const vec2 lp0 = mk(m0.x, m0.y), lp1 = mk(m1.x, m1.y);
    if (mtype == kManCircles) {
      normal = mk(1.0f, 0.0f);
      const vec2 pA = mul(xfA, lp);
      const vec2 pB = mul(xfB, lp0);
      if (dist_sq(pA, pB) > kEpsilon * kEpsilon) normal = normalize(pB - pA);
      const vec2 qA = pA + rA * normal;
      const vec2 qB = pB - rB * normal;
      wp0 = 0.5f * (qA + qB);
    } else if (mtype == kManFaceA) {
      normal = mul(xfA.q, ln);
      const vec2 plane = mul(xfA, lp);
      {
        const vec2 clip = mul(xfB, lp0);
        const vec2 qA = clip + (rA - dot(clip - plane, normal)) * normal;
        const vec2 qB = clip - rB * normal;
        wp0 = 0.5f * (qA + qB);
      }
      if (count > 1) {
        const vec2 clip = mul(xfB, lp1);
        const vec2 qA = clip + (rA - dot(clip - plane, normal)) * normal;
        const vec2 qB = clip - rB * normal;
        wp1 = 0.5f * (qA + qB);
      }
    } else {
      normal = mul(xfB.q, ln);
      const vec2 plane = mul(xfB, lp);
      {
        const vec2 clip = mul(xfA, lp0);
        const vec2 qB = clip + (rB - dot(clip - plane, normal)) * normal;
        const vec2 qA = clip - rA * normal;
        wp0 = 0.5f * (qA + qB);
      }
      if (count > 1) {
        const vec2 clip = mul(xfA, lp1);
        const vec2 qB = clip + (rB - dot(clip - plane, normal)) * normal;
        const vec2 qA = clip - rA * normal;
        wp1 = 0.5f * (qA + qB);
      }
      normal = -normal;
    }
    const vec2 tangent = cross(normal, 1.0f);
    vec2 rA0 = wp0 - cA, rB0 = wp0 - cB, rA1 = wp1 - cA, rB1 = wp1 - cB;
    float nm0, tm0, bias0 = 0.0f, nm1 = 0.0f, tm1 = 0.0f, bias1 = 0.0f;
    {
      const float rnA = cross(rA0, normal), rnB = cross(rB0, normal);
      const float kn = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
      nm0 = kn > 0.0f ? 1.0f / kn : 0.0f;
      const float rtA = cross(rA0, tangent), rtB = cross(rB0, tangent);
      const float kt = mA + mB + iA * rtA * rtA + iB * rtB * rtB;
      tm0 = kt > 0.0f ? 1.0f / kt : 0.0f;
      const float vrel = dot(normal, vB + cross(wB, rB0) - (vA + cross(wA, rA0)));
      if (vrel < -mat.z) bias0 = -mat.y * vrel;
    }
    if (count > 1) {
      const float rnA = cross(rA1, normal), rnB = cross(rB1, normal);
      const float kn = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
      nm1 = kn > 0.0f ? 1.0f / kn : 0.0f;
      const float rtA = cross(rA1, tangent), rtB = cross(rB1, tangent);
      const float kt = mA + mB + iA * rtA * rtA + iB * rtB * rtB;
      tm1 = kt > 0.0f ? 1.0f / kt : 0.0f;
      const float vrel = dot(normal, vB + cross(wB, rB1) - (vA + cross(wA, rA1)));
      if (vrel < -mat.z) bias1 = -mat.y * vrel;
    }
    int vcount = count, pcount = count;
    if (d.sleep_on) {  // a sleeping island is not built (b2_world.cpp:535-539): no points to solve
      const int isl = d.btype[ia] != kStaticBody ? d.parent[ia] : d.parent[ib];
      if (d.isl_awake[isl]) d.isl_cons[isl] = 1;
      else vcount = pcount = 0;
    }
    mat22 K = {mk(0.0f, 0.0f), mk(0.0f, 0.0f)}, NM = K;
    if (count == 2 && vcount == 2) {  // block solver preparation (:236-256)
      const float rn1A = cross(rA0, normal), rn1B = cross(rB0, normal);
      const float rn2A = cross(rA1, normal), rn2B = cross(rB1, normal);
      const float k11 = mA + mB + iA * rn1A * rn1A + iB * rn1B * rn1B;
      const float k22 = mA + mB + iA * rn2A * rn2A + iB * rn2B * rn2B;
      const float k12 = mA + mB + iA * rn1A * rn2A + iB * rn1B * rn2B;
      const float kMaxCond = 1000.0f;
      if (k11 * k11 < kMaxCond * (k11 * k22 - k12 * k12)) {
        K.ex = mk(k11, k12);
        K.ey = mk(k12, k22);
        NM = inverse(K);
      } else {
        vcount = 1;
      }
    }
    d.vc[0][s] = make_float4(normal.x, normal.y, mat.x, mat.w);
    d.vc[1][s] = make_float4(mA, mB, iA, iB);
    // point count, bit 8: the world warm-starts
    d.vc[2][s] = make_float4(as_float(ia), as_float(ib), as_float(vcount | (warm ? 256 : 0)), as_float(c));
    {  // this slot's place in each dynamic body's colour-ordered list (rank of
       // its colour among the body's colours)
      const int col = d.part_on ? -1 : d.ccol[c];
      if (col >= 0 && col < kBodyTab) {
        const unsigned long long below = (1ull << col) - 1ull;
        if (d.btype[ia] == kDynamicBody) d.btab[(size_t)ia * kBodyTab + __popcll(d.bmask[ia] & below)] = 2 * s;
        if (d.btype[ib] == kDynamicBody) d.btab[(size_t)ib * kBodyTab + __popcll(d.bmask[ib] & below)] = 2 * s + 1;
      }
    }
    d.vc[3][s] = make_float4(rA0.x, rA0.y, rB0.x, rB0.y);
    d.vc[4][s] = make_float4(ni0, ti0, nm0, tm0);
    d.vc[5][s] = make_float4(rA1.x, rA1.y, rB1.x, rB1.y);
    d.vc[6][s] = make_float4(ni1, ti1, nm1, tm1);
    d.vc[7][s] = make_float4(bias0, bias1, K.ex.y, K.ey.x);
    d.vc[8][s] = make_float4(NM.ex.x, NM.ex.y, NM.ey.x, NM.ey.y);
    d.vc[9][s] = make_float4(K.ex.x, K.ex.y, K.ey.x, K.ey.y);
    d.pc[0][s] = make_float4(lp.x, lp.y, ln.x, ln.y);
    d.pc[1][s] = make_float4(lp0.x, lp0.y, lp1.x, lp1.y);
    d.pc[2][s] = make_float4(lcA.x, lcA.y, lcB.x, lcB.y);
    d.pc[3][s] = make_float4(mA, mB, iA, iB);
    d.pc[4][s] = make_float4(rA, rB, as_float(mtype), as_float(pcount));
  }
}

__device__ __forceinline__ bool writes_body(const Dev& d, int b) { return d.btype[b] == kDynamicBody; }

// b2ContactSolver::WarmStart (:260-299) of slot s
// Body-state access for the constraint functions: global arrays, or a
// world's bodies staged in LDS (bodies [base, base + n) of one world).
struct GlobalBodies {
  float4* p;
  __device__ __forceinline__ float4* operator()(int i) const { return p + i; }
  __device__ __forceinline__ bool writes(const Dev& d, int i) const { return writes_body(d, i); }
};
struct LdsBodies {
  float4* p;
  int base;
  __device__ __forceinline__ float4* operator()(int i) const { return p + (i - base); }
  __device__ __forceinline__ bool writes(const Dev& d, int i) const { return writes_body(d, i); }
};

template <class BV, class Src>
__device__ __forceinline__ void warm_start_src(const Dev& d, BV vel, const Src& src) {
  {
    const float4 v2i = src.vc(2);
    const int ia = as_int(v2i.x), ib = as_int(v2i.y), count = as_int(v2i.z) & 0xff;
    if ((as_int(v2i.z) & 256) == 0) return;  // no warm starting in this world
    const float4 n4 = src.vc(0), m = src.vc(1);
    const vec2 normal = mk(n4.x, n4.y), tangent = cross(normal, 1.0f);
    float4 va = *vel(ia), vb = *vel(ib);
    vec2 vA = mk(va.x, va.y), vB = mk(vb.x, vb.y);
    float wA = va.z, wB = vb.z;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if (j < count) {
        const float4 r = src.vc(3 + 2 * j), im = src.vc(4 + 2 * j);
        const vec2 P = im.x * normal + im.y * tangent;
        wA -= m.z * cross(mk(r.x, r.y), P);
        vA = vA - m.x * P;
        wB += m.w * cross(mk(r.z, r.w), P);
        vB = vB + m.y * P;
      }
    }
    if (vel.writes(d, ia)) *vel(ia) = make_float4(vA.x, vA.y, wA, 0.0f);
    if (vel.writes(d, ib)) *vel(ib) = make_float4(vB.x, vB.y, wB, 0.0f);
  }
}

// b2ContactSolver::SolveVelocityConstraints (:301-639) for slots [begin, end)
// Where a constraint solve reads its slot's fields: straight from the
// constraint arrays (each load issued where the solve needs it), or from a
// copy loaded ahead into registers (the tail launch loads the next colour's
// constraints while the current colour is solved).
// A source also says where the solved impulses go and whether the island's
// position iterations are over (these two default to the constraint arrays
// and the island state in global memory).
__device__ __forceinline__ void put_impulses_global(const Dev& d, int s, float4 q0, float4 q1, int count) {
  d.vc[4][s] = q0;
  if (count > 1) d.vc[6][s] = q1;
}
__device__ __forceinline__ bool island_skip_global(const Dev& d, int isl, int it) {
  if (island_done(d, isl, it)) return true;
  island_rearm(d, isl, it);
  return false;
}
struct SlotFields {
  const Dev& d;
  int s;
  __device__ __forceinline__ float4 vc(int k) const { return d.vc[k][s]; }
  __device__ __forceinline__ float4 pc(int k) const { return d.pc[k][s]; }
  __device__ __forceinline__ void put_impulses(const Dev& dd, int ss, float4 q0, float4 q1, int n) const {
    put_impulses_global(dd, ss, q0, q1, n);
  }
  __device__ __forceinline__ bool island_skip(const Dev& dd, int isl, int it) const {
    return island_skip_global(dd, isl, it);
  }
};
template <class BV>
__device__ __forceinline__ void warm_start_t(const Dev& d, int s, BV vel) {
  warm_start_src(d, vel, SlotFields{d, s});
}
__device__ __forceinline__ void warm_start(const Dev& d, int s) { warm_start_t(d, s, GlobalBodies{d.vel}); }
struct VelocityRegs {
  float4 v[10];
  __device__ __forceinline__ float4 vc(int k) const { return v[k]; }
  __device__ __forceinline__ void put_impulses(const Dev& d, int s, float4 q0, float4 q1, int n) const {
    put_impulses_global(d, s, q0, q1, n);
  }
};
__device__ __forceinline__ VelocityRegs load_velocity_regs(const Dev& d, int s) {
  VelocityRegs r;
#pragma unroll
  for (int k = 0; k < 10; ++k) r.v[k] = d.vc[k][s];
  return r;
}
struct PositionRegs {
  float4 v2;
  float4 p[5];
  int isl;
  __device__ __forceinline__ float4 vc(int) const { return v2; }  // only field 2 is read
  __device__ __forceinline__ float4 pc(int k) const { return p[k]; }
  __device__ __forceinline__ bool island_skip(const Dev& d, int is, int it) const {
    return island_skip_global(d, is, it);
  }
};
__device__ __forceinline__ PositionRegs load_position_regs(const Dev& d, int s) {
  PositionRegs r;
  r.v2 = d.vc[2][s];
#pragma unroll
  for (int k = 0; k < 5; ++k) r.p[k] = d.pc[k][s];
  r.isl = d.cisland[s];
  return r;
}

template <class BV, class Src>
__device__ __forceinline__ void contact_velocity_src(const Dev& d, int s, BV vel, const Src& src) {
  const float4 v2i = src.vc(2);
  const int ia = as_int(v2i.x), ib = as_int(v2i.y), count = as_int(v2i.z) & 0xff;
  if (count == 0) return;  // a constraint of a sleeping island
  const float4 n4 = src.vc(0), m = src.vc(1);
  const float mA = m.x, mB = m.y, iA = m.z, iB = m.w;
  const vec2 normal = mk(n4.x, n4.y);
  const float friction = n4.z, tangent_speed = n4.w;
  float4 va = *vel(ia), vb = *vel(ib);
  vec2 vA = mk(va.x, va.y), vB = mk(vb.x, vb.y);
  float wA = va.z, wB = vb.z;
  const float4 r0 = src.vc(3), r1 = src.vc(5);
  float4 q0 = src.vc(4), q1 = src.vc(6);
  const vec2 rA0 = mk(r0.x, r0.y), rB0 = mk(r0.z, r0.w);
  const vec2 rA1 = mk(r1.x, r1.y), rB1 = mk(r1.z, r1.w);
  const vec2 tangent = cross(normal, 1.0f);
  // friction first (:324-351)
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    if (j < count) {
      float4& q = j == 0 ? q0 : q1;
      const vec2 rA = j == 0 ? rA0 : rA1, rB = j == 0 ? rB0 : rB1;
      const vec2 dv = vB + cross(wB, rB) - (vA + cross(wA, rA));
      const float vt = dot(dv, tangent) - tangent_speed;
      float lambda = q.w * (-vt);
      const float max_friction = friction * q.x;
      const float new_impulse = clampf(q.y + lambda, -max_friction, max_friction);
      lambda = new_impulse - q.y;
      q.y = new_impulse;
      const vec2 P = lambda * tangent;
      vA = vA - mA * P;
      wA -= iA * cross(rA, P);
      vB = vB + mB * P;
      wB += iB * cross(rB, P);
    }
  }
  const float4 b7 = src.vc(7);
  if (count == 1) {
    const vec2 dv = vB + cross(wB, rB0) - (vA + cross(wA, rA0));
    const float vn = dot(dv, normal);
    float lambda = -q0.z * (vn - b7.x);
    const float new_impulse = maxf(q0.x + lambda, 0.0f);
    lambda = new_impulse - q0.x;
    q0.x = new_impulse;
    const vec2 P = lambda * normal;
    vA = vA - mA * P;
    wA -= iA * cross(rA0, P);
    vB = vB + mB * P;
    wB += iB * cross(rB0, P);
  } else {
    // 2-point block LCP by total enumeration (:389-631)
    const float4 nm4 = src.vc(8), k4 = src.vc(9);
    const mat22 K = {mk(k4.x, k4.y), mk(k4.z, k4.w)};
    const mat22 NM = {mk(nm4.x, nm4.y), mk(nm4.z, nm4.w)};
    const vec2 a = mk(q0.x, q1.x);
    const vec2 dv1 = vB + cross(wB, rB0) - (vA + cross(wA, rA0));
    const vec2 dv2 = vB + cross(wB, rB1) - (vA + cross(wA, rA1));
    float vn1 = dot(dv1, normal);
    float vn2 = dot(dv2, normal);
    vec2 b = mk(vn1 - b7.x, vn2 - b7.y);
    b = b - mul(K, a);
    vec2 x;
    bool solved = false;
    x = -mul(NM, b);
    if (x.x >= 0.0f && x.y >= 0.0f) {
      solved = true;
    } else {
      x = mk(-q0.z * b.x, 0.0f);
      vn2 = K.ex.y * x.x + b.y;
      if (x.x >= 0.0f && vn2 >= 0.0f) {
        solved = true;
      } else {
        x = mk(0.0f, -q1.z * b.y);
        vn1 = K.ey.x * x.y + b.x;
        if (x.y >= 0.0f && vn1 >= 0.0f) {
          solved = true;
        } else {
          x = mk(0.0f, 0.0f);
          vn1 = b.x;
          vn2 = b.y;
          if (vn1 >= 0.0f && vn2 >= 0.0f) solved = true;
        }
      }
    }
    if (solved) {
      const vec2 dd = x - a;
      const vec2 P1 = dd.x * normal;
      const vec2 P2 = dd.y * normal;
      vA = vA - mA * (P1 + P2);
      wA -= iA * (cross(rA0, P1) + cross(rA1, P2));
      vB = vB + mB * (P1 + P2);
      wB += iB * (cross(rB0, P1) + cross(rB1, P2));
      q0.x = x.x;
      q1.x = x.y;
    }
  }
  src.put_impulses(d, s, q0, q1, count);
  if (vel.writes(d, ia)) *vel(ia) = make_float4(vA.x, vA.y, wA, 0.0f);
  if (vel.writes(d, ib)) *vel(ib) = make_float4(vB.x, vB.y, wB, 0.0f);
}
template <class BV>
__device__ __forceinline__ void contact_velocity_t(const Dev& d, int s, BV vel) {
  contact_velocity_src(d, s, vel, SlotFields{d, s});
}
__device__ __forceinline__ void contact_velocity(const Dev& d, int s) {
  contact_velocity_t(d, s, GlobalBodies{d.vel});
}

// Slots of a colour: given by the host when it knows them (one dependent
// load less at the start of every sweep), else read from the offsets.
__device__ __forceinline__ int2 colour_slots(const Dev& d, int colour, int2 given) {
  return given.x >= 0 ? given : make_int2(d.col_off[colour], d.col_off[colour + 1]);
}

__global__ void __launch_bounds__(kBlock) k_warm_start(Dev d, int colour, int2 slots) {
  const int2 r = colour_slots(d, colour, slots);
  const int begin = r.x, end = r.y;
  B2G_FOR(s, begin, end) warm_start(d, s);
}

__global__ void __launch_bounds__(kBlock) k_contact_velocity(Dev d, int colour, int2 slots) {
  const int2 r = colour_slots(d, colour, slots);
  const int begin = r.x, end = r.y;
  B2G_FOR(s, begin, end) contact_velocity(d, s);
}

// b2ContactSolver::StoreImpulses (:641-657): only the vc pointCount points.
// That count (the b2ContactImpulse count b2Island::Report hands to PostSolve,
// b2_island.cpp:621-647) goes to bits 16-23 of the manifold's packed word;
// the narrowphase rewrites the word every step, so the bits name the points
// the last step solved (0: not solved).
__device__ __forceinline__ void store_impulses(const Dev& d, int s) {
  {
    const float4 v2i = d.vc[2][s];
    const int count = as_int(v2i.z) & 0xff, c = as_int(v2i.w);
    const float4 q0 = d.vc[4][s];
    float4* mp = d.cman + 4 * (size_t)c;
    int* packed = reinterpret_cast<int*>(mp + 3) + 2;
    *packed = (*packed & 0xffff) | (count << 16);
    if (count == 0) return;  // not solved: the manifold keeps its impulses
    float4 m0 = mp[0];
    m0.z = q0.x;
    m0.w = q0.y;
    mp[0] = m0;
    if (count > 1) {
      const float4 q1 = d.vc[6][s];
      float4 m1 = mp[1];
      m1.z = q1.x;
      m1.w = q1.y;
      mp[1] = m1;
    }
  }
}

// ===========================================================================
// Position integration with the translation / rotation clamp
// (src/dynamics/b2_island.cpp:341-385)
// ===========================================================================
__device__ __forceinline__ void integrate_position(const Dev& d, int b) {
  {
    const int w = d.bworld[b];
    if (!world_active(d, w) || d.btype[b] == kStaticBody) return;
    if (d.sleep_on && !d.bactive[b]) return;  // a sleeping island does not move
    const float h = d.wpar[w].z;
    const float4 v4 = d.vel[b];
    vec2 v = mk(v4.x, v4.y);
    float wv = v4.z;
    const vec2 tr = h * v;
    if (dot(tr, tr) > kMaxTranslationSq) {
      const float ratio = kMaxTranslation / length(tr);
      v = ratio * v;
    }
    const float rotation = h * wv;
    if (rotation * rotation > kMaxRotationSq) {
      const float ratio = kMaxRotation / absf(rotation);
      wv *= ratio;
    }
    const float4 p = d.pos[b];
    const vec2 c = mk(p.x, p.y) + h * v;
    const float a = p.z + h * wv;
    d.pos[b] = make_float4(c.x, c.y, a, 0.0f);
    d.vel[b] = make_float4(v.x, v.y, wv, 0.0f);
  }
}

// StoreImpulses for every solved contact and position integration for every
// body, in one launch (the two touch disjoint data)
// after the velocity iterations: store impulses, integrate positions and
// (islands_ready) give every constraint its island for the position phase
__global__ void __launch_bounds__(kBlock) k_store_and_integrate(Dev d, int islands_ready) {
  const int n = d.cnt[C_CONS];
  const int nc = n + (islands_ready ? d.nj : 0);
  const int m = nc > d.nb ? nc : d.nb;
  B2G_FOR(k, 0, m) {
    if (k < n) store_impulses(d, k);
    if (k < d.nb) integrate_position(d, k);
    if (islands_ready && k < nc) island_assign(d, n, k);
  }
}

// b2ContactSolver::SolvePositionConstraints (:659-787) for slots [begin, end)
// returns the minimum separation seen (0 when the island is already solved)
template <class BP, class Src>
__device__ __forceinline__ float contact_position_src(const Dev& d, int isl, int it, BP pos, const Src& src) {
  // the constraint's loads are issued before the island test, so they are in
  // flight together with it
  const float4 v2i = src.vc(2);
  const float4 q0 = src.pc(0), q1 = src.pc(1), q2 = src.pc(2), q3 = src.pc(3), q4 = src.pc(4);
  const int ia = as_int(v2i.x), ib = as_int(v2i.y);
  const float4 pa = *pos(ia), pb = *pos(ib);  // in flight with the island test too
  if (as_int(q4.w) == 0) return 0.0f;  // a constraint of a sleeping island
  if (src.island_skip(d, isl, it)) return 0.0f;
  const vec2 lp = mk(q0.x, q0.y), ln = mk(q0.z, q0.w);
  const vec2 lcA = mk(q2.x, q2.y), lcB = mk(q2.z, q2.w);
  const float mA = q3.x, mB = q3.y, iA = q3.z, iB = q3.w;
  const float radA = q4.x, radB = q4.y;
  const int mtype = as_int(q4.z), count = as_int(q4.w);
  vec2 cA = mk(pa.x, pa.y), cB = mk(pb.x, pb.y);
  float aA = pa.z, aB = pb.z;
  float min_sep = 0.0f;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    if (j < count) {
      xform xfA, xfB;
      xfA.q = rot_of(aA);
      xfB.q = rot_of(aB);
      xfA.p = cA - mul(xfA.q, lcA);
      xfB.p = cB - mul(xfB.q, lcB);
      const vec2 lpj = j == 0 ? mk(q1.x, q1.y) : mk(q1.z, q1.w);
      vec2 normal, point;
      float separation;
      if (mtype == kManCircles) {
        const vec2 pA = mul(xfA, lp);
        const vec2 pB = mul(xfB, mk(q1.x, q1.y));
        normal = normalize(pB - pA);
        point = 0.5f * (pA + pB);
        separation = dot(pB - pA, normal) - radA - radB;
      } else if (mtype == kManFaceA) {
        normal = mul(xfA.q, ln);
        const vec2 plane = mul(xfA, lp);
        const vec2 clip = mul(xfB, lpj);
        separation = dot(clip - plane, normal) - radA - radB;
        point = clip;
      } else {
        normal = mul(xfB.q, ln);
        const vec2 plane = mul(xfB, lp);
        const vec2 clip = mul(xfA, lpj);
        separation = dot(clip - plane, normal) - radA - radB;
        point = clip;
        normal = -normal;
      }
      const vec2 rA = point - cA, rB = point - cB;
      min_sep = minf(min_sep, separation);
      const float C = clampf(kBaumgarte * (separation + kLinearSlop), -kMaxLinearCorrection, 0.0f);
      const float rnA = cross(rA, normal), rnB = cross(rB, normal);
      const float K = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
      const float impulse = K > 0.0f ? -C / K : 0.0f;
      const vec2 P = impulse * normal;
      cA = cA - mA * P;
      aA -= iA * cross(rA, P);
      cB = cB + mB * P;
      aB += iB * cross(rB, P);
    }
  }
  if (pos.writes(d, ia)) *pos(ia) = make_float4(cA.x, cA.y, aA, 0.0f);
  if (pos.writes(d, ib)) *pos(ib) = make_float4(cB.x, cB.y, aB, 0.0f);
  return min_sep;
}
template <class BP>
__device__ __forceinline__ float contact_position_t(const Dev& d, int s, int isl, int it, BP pos) {
  return contact_position_src(d, isl, it, pos, SlotFields{d, s});
}
__device__ __forceinline__ float contact_position(const Dev& d, int s, int isl, int it) {
  return contact_position_t(d, s, isl, it, GlobalBodies{d.pos});
}

// The island early exit (b2_island.cpp:403-408) asks whether the island's
// minimum separation this iteration (b2ContactSolver::SolvePositionConstraints,
// b2_contact_solver.cpp:676-779, starting from 0) is >= -3 linearSlop, that
// is, whether no contact separated by less. So a contact whose own minimum
// `m` is below marks its island, and settled contacts write nothing: no
// minimum is reduced across the island, which would put every block of a
// large island on one address. The marks of a wave collapse to one store
// when they share an island.
__device__ __forceinline__ bool separation_bad(float m) { return !(m >= -3.0f * kLinearSlop); }
__device__ __forceinline__ void mark_island_sep(const Dev& d, int it, int isl, bool valid, float m) {
  const bool bad = valid && separation_bad(m);
  const uint64_t act = __ballot(bad);
  if (act == 0) return;
  unsigned* base = d.sepbad + (size_t)(it % 3) * d.nb;
  const int first = __ffsll((long long)act) - 1;
  const int key0 = __shfl(isl, first, kWave);
  if (__ballot(bad && isl != key0) == 0) {
    if (lane_id() == first) __hip_atomic_store(base + key0, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else if (bad) {
    __hip_atomic_store(base + isl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__device__ __forceinline__ void contact_position_serial(const Dev& d, int begin, int end, int it) {
  for (int s = begin; s < end; ++s) {
    const int isl = d.cisland[s];
    const float m = contact_position(d, s, isl, it);
    if (separation_bad(m)) d.sepbad[island_slot(d, isl, it)] = 1u;
  }
}

__global__ void __launch_bounds__(kBlock) k_contact_position(Dev d, int colour, int2 slots, int it) {
  const int2 r = colour_slots(d, colour, slots);
  const int begin = r.x, end = r.y;
  if (blockDim.x == 1) {  // the serial colour group: one lane, slot order
    contact_position_serial(d, begin, end, it);
    return;
  }
  B2G_FOR_ALL(s, valid, begin, end) {
    const int isl = valid ? d.cisland[s] : 0;
    const float m = valid ? contact_position(d, s, isl, it) : 0.0f;
    mark_island_sep(d, it, isl, valid != 0, m);
  }
}

// b2ContactSolver::WarmStart as a gather: each dynamic body applies the
// impulses of its constraints itself, in colour order (its list in btab),
// which is the order in which the colour sweeps update it; the arithmetic per
// point is warm_start_t's, so the velocities are the same bit for bit. One
// launch instead of one per colour. Needs every colour below kBodyTab; the
// serial group (not in the body masks) follows in its own launch.
__global__ void __launch_bounds__(kBlock) k_warm_gather(Dev d) {
  B2G_FOR(b, 0, d.nb) {
    if (!world_active(d, d.bworld[b])) continue;  // its colour mask is a past step's
    const unsigned long long mask = d.bmask[b];  // only dynamic bodies have bits
    if (mask == 0ull) continue;
    const int n = __popcll(mask);
    const float4 v0 = d.vel[b];
    vec2 v = mk(v0.x, v0.y);
    float w = v0.z;
    bool touched = false;
    constexpr int kChunk = 4;  // constraints whose loads are issued together
    for (int r0 = 0; r0 < n; r0 += kChunk) {
      int slot[kChunk];
      float4 v2i[kChunk], n4[kChunk], mm[kChunk], r0p[kChunk], i0p[kChunk], r1p[kChunk], i1p[kChunk];
#pragma unroll
      for (int u = 0; u < kChunk; ++u) slot[u] = r0 + u < n ? d.btab[(size_t)b * kBodyTab + r0 + u] : -1;
#pragma unroll
      for (int u = 0; u < kChunk; ++u) {
        const int s = slot[u] >= 0 ? slot[u] >> 1 : 0;
        v2i[u] = d.vc[2][s];
        n4[u] = d.vc[0][s];
        mm[u] = d.vc[1][s];
        r0p[u] = d.vc[3][s];
        i0p[u] = d.vc[4][s];
        r1p[u] = d.vc[5][s];
        i1p[u] = d.vc[6][s];
      }
#pragma unroll
      for (int u = 0; u < kChunk; ++u) {
        if (slot[u] < 0) continue;
        if ((as_int(v2i[u].z) & 256) == 0) continue;  // no warm starting in this world
        touched = true;
        const bool side_b = (slot[u] & 1) != 0;
        const int count = as_int(v2i[u].z) & 0xff;
        const float4 m = mm[u];
        const vec2 normal = mk(n4[u].x, n4[u].y), tangent = cross(normal, 1.0f);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if (j < count) {
            const float4 rr = j == 0 ? r0p[u] : r1p[u], im = j == 0 ? i0p[u] : i1p[u];
            const vec2 P = im.x * normal + im.y * tangent;
            if (!side_b) {
              w -= m.z * cross(mk(rr.x, rr.y), P);
              v = v - m.x * P;
            } else {
              w += m.w * cross(mk(rr.z, rr.w), P);
              v = v + m.y * P;
            }
          }
        }
      }
    }
    if (touched) d.vel[b] = make_float4(v.x, v.y, w, 0.0f);
  }
}

// Several small colours in one launch: a single workgroup sweeps them in
// colour order with a workgroup barrier between colours. A launch costs more
// than a small colour's work, so the last colours of a sweep (a few hundred
// constraints or fewer each: one CU's bandwidth bounds a workgroup's round)
// share one launch; the order of the updates is the same as with a launch per
// colour, and so are the results. The colours'
// slot ranges come from the host when it knows them, else from the offsets.
constexpr int kTailThreads = 1024;
constexpr int kTailMax = 16;  // colours per tail launch
struct TailSlots {
  int c0, n, known;       // colours c0 .. c0 + n - 1 (n <= kTailMax when known)
  int off[kTailMax + 1];  // known: slot offsets of colours c0 .. c0 + n
};
enum TailOp : int { kTailWarm = 0, kTailVelocity = 1, kTailPosition = 2 };

template <int kOp>
__global__ void __launch_bounds__(kTailThreads) k_sweep_tail(Dev d, TailSlots t, int it, int serial) {
  const int nt = (int)blockDim.x, tid = (int)threadIdx.x;
  // the colours' slot offsets (given, or read on the device), and the
  // non-empty colours in order
  __shared__ int s_off[kColourSlots + 1];
  __shared__ int s_q[kColourSlots];
  __shared__ int s_nq, s_one;
  if (t.known) {
    if (tid == 0)
      for (int q = 0; q <= t.n; ++q) s_off[q] = t.off[q];  // (uniform index: the argument stays in registers)
  } else {
    for (int q = tid; q <= t.n; q += nt) s_off[q] = d.col_off[t.c0 + q];
  }
  __syncthreads();
  if (tid == 0) {
    int k = 0;
    bool one = kOp != kTailWarm;
    for (int q = 0; q < t.n; ++q) {
      const int len = s_off[q + 1] - s_off[q];
      if (len <= 0) continue;
      s_q[k++] = q;
      one = one && len <= nt;
    }
    s_nq = k;
    s_one = one ? 1 : 0;
  }
  __syncthreads();
  const int nq = s_nq;
  if (s_one) {
    // every colour fits one pass: lane tid takes slot off[q] + tid of colour
    // q, and loads the next colour's constraint while this one is solved
    // (the constraint fields do not depend on the bodies; only the bodies
    // wait for the barrier)
    auto slot_of = [&](int k) {
      const int q = s_q[k];
      const int s = s_off[q] + tid;
      return s < s_off[q + 1] ? s : -1;
    };
    if (kOp == kTailVelocity) {
      int s_next = nq > 0 ? slot_of(0) : -1;
      VelocityRegs nxt{};
      if (s_next >= 0) nxt = load_velocity_regs(d, s_next);
      for (int k = 0; k < nq; ++k) {
        const int s_cur = s_next;
        const VelocityRegs cur = nxt;
        if (k + 1 < nq) {
          s_next = slot_of(k + 1);
          if (s_next >= 0) nxt = load_velocity_regs(d, s_next);
        }
        if (s_cur >= 0) contact_velocity_src(d, s_cur, GlobalBodies{d.vel}, cur);
        __syncthreads();  // the next colour sees this colour's body updates
      }
    } else {
      int s_next = nq > 0 ? slot_of(0) : -1;
      PositionRegs nxt{};
      if (s_next >= 0) nxt = load_position_regs(d, s_next);
      for (int k = 0; k < nq; ++k) {
        const int s_cur = s_next;
        const PositionRegs cur = nxt;
        if (k + 1 < nq) {
          s_next = slot_of(k + 1);
          if (s_next >= 0) nxt = load_position_regs(d, s_next);
        }
        const bool valid = s_cur >= 0;
        const float m = valid ? contact_position_src(d, cur.isl, it, GlobalBodies{d.pos}, cur) : 0.0f;
        mark_island_sep(d, it, valid ? cur.isl : 0, valid, m);
        __syncthreads();
      }
    }
  } else {
    for (int k = 0; k < nq; ++k) {
      const int q = s_q[k];
      const int begin = s_off[q], end = s_off[q + 1];
      for (int base = begin; base < end; base += nt) {
        const int s = base + tid;
        const bool valid = s < end;
        if (kOp == kTailWarm) {
          if (valid) warm_start(d, s);
        } else if (kOp == kTailVelocity) {
          if (valid) contact_velocity(d, s);
        } else {
          const int isl = valid ? d.cisland[s] : 0;
          const float m = valid ? contact_position(d, s, isl, it) : 0.0f;
          mark_island_sep(d, it, isl, valid, m);
        }
      }
      __syncthreads();  // the next colour sees this colour's body updates
    }
  }
  if (serial && threadIdx.x == 0) {  // the serial group follows the last colour
    const int b = d.col_off[kSerialColour], e = d.col_off[kSerialColour + 1];
    if (kOp == kTailPosition) {
      contact_position_serial(d, b, e, it);
    } else {
      for (int s = b; s < e; ++s) {
        if (kOp == kTailWarm) warm_start(d, s); else contact_velocity(d, s);
      }
    }
  }
}

// ===========================================================================
// b2RevoluteJoint (src/dynamics/b2_revolute_joint.cpp:73-321), colour order q
// ===========================================================================
__global__ void __launch_bounds__(kBlock) k_joint_prepare(Dev d) {
  B2G_FOR(q, 0, d.nj) {
    const Joint& j = d.joints[d.jorder[q]];
    // solved or not; the island itself is filled in when the islands are assigned
    bool on = world_active(d, j.world) && (d.btype[j.a] != kStaticBody || d.btype[j.b] != kStaticBody);
    if (on && d.sleep_on) {  // a joint of a sleeping island is not solved
      const int isl = island_of(d, j.a, j.b);
      if (d.isl_awake[isl]) d.isl_cons[isl] = 1;
      else on = false;
    }
    d.jisland[q] = on ? 0 : -1;
  }
}

__global__ void __launch_bounds__(kBlock) k_joint_init(Dev d, int begin, int end) {
  B2G_FOR(q, begin, end) {
    if (d.jisland[q] < 0) continue;
    const int ji = d.jorder[q];
    Joint j = d.joints[ji];
    JointSolve s;
    const float4 pa = d.prop[j.a], pb = d.prop[j.b];
    s.lca_x = pa.z; s.lca_y = pa.w; s.lcb_x = pb.z; s.lcb_y = pb.w;
    s.ma = pa.x; s.mb = pb.x; s.ia = pa.y; s.ib = pb.y;
    const float aA = d.pos[j.a].z, aB = d.pos[j.b].z;
    const float4 va = d.vel[j.a], vb = d.vel[j.b];
    vec2 vA = mk(va.x, va.y), vB = mk(vb.x, vb.y);
    float wA = va.z, wB = vb.z;
    const rot2 qA = rot_of(aA), qB = rot_of(aB);
    const vec2 rA = mul(qA, mk(j.la_x, j.la_y) - mk(s.lca_x, s.lca_y));
    const vec2 rB = mul(qB, mk(j.lb_x, j.lb_y) - mk(s.lcb_x, s.lcb_y));
    s.ra_x = rA.x; s.ra_y = rA.y; s.rb_x = rB.x; s.rb_y = rB.y;
    const float mA = s.ma, mB = s.mb, iA = s.ia, iB = s.ib;
    s.kxx = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
    s.kyx = -rA.y * rA.x * iA - rB.y * rB.x * iB;
    s.kxy = s.kyx;
    s.kyy = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
    s.axial = iA + iB;
    bool fixed_rotation;
    if (s.axial > 0.0f) {
      s.axial = 1.0f / s.axial;
      fixed_rotation = false;
    } else {
      fixed_rotation = true;
    }
    s.angle = aB - aA - j.ref;
    if ((j.flags & 2) == 0 || fixed_rotation) {
      j.lower_imp = 0.0f;
      j.upper_imp = 0.0f;
    }
    if ((j.flags & 4) == 0 || fixed_rotation) j.motor = 0.0f;
    const float4 wp = d.wpar[j.world];
    if (d.wflag[j.world] & 1) {
      const float r = wp.w;
      j.imp_x *= r;
      j.imp_y *= r;
      j.motor *= r;
      j.lower_imp *= r;
      j.upper_imp *= r;
      const float axial_impulse = j.motor + j.lower_imp - j.upper_imp;
      const vec2 P = mk(j.imp_x, j.imp_y);
      vA = vA - mA * P;
      wA -= iA * (cross(rA, P) + axial_impulse);
      vB = vB + mB * P;
      wB += iB * (cross(rB, P) + axial_impulse);
    } else {
      j.imp_x = j.imp_y = 0.0f;
      j.motor = j.lower_imp = j.upper_imp = 0.0f;
    }
    d.joints[ji] = j;
    d.jsol[q] = s;
    if (writes_body(d, j.a)) d.vel[j.a] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (writes_body(d, j.b)) d.vel[j.b] = make_float4(vB.x, vB.y, wB, 0.0f);
  }
}

__global__ void __launch_bounds__(kBlock) k_joint_velocity(Dev d, int begin, int end) {
  B2G_FOR(q, begin, end) {
    if (d.jisland[q] < 0) continue;
    const int ji = d.jorder[q];
    Joint j = d.joints[ji];
    const JointSolve s = d.jsol[q];
    const float4 va = d.vel[j.a], vb = d.vel[j.b];
    vec2 vA = mk(va.x, va.y), vB = mk(vb.x, vb.y);
    float wA = va.z, wB = vb.z;
    const float mA = s.ma, mB = s.mb, iA = s.ia, iB = s.ib;
    const bool fixed_rotation = (iA + iB == 0.0f);
    const float4 wp = d.wpar[j.world];
    const float h = wp.z;
    const float inv_dt = h > 0.0f ? 1.0f / h : 0.0f;
    if ((j.flags & 4) && !fixed_rotation) {
      const float Cdot = wB - wA - j.speed;
      float impulse = -s.axial * Cdot;
      const float old = j.motor;
      const float max_impulse = h * j.max_torque;
      j.motor = clampf(j.motor + impulse, -max_impulse, max_impulse);
      impulse = j.motor - old;
      wA -= iA * impulse;
      wB += iB * impulse;
    }
    if ((j.flags & 2) && !fixed_rotation) {
      {
        const float C = s.angle - j.lower;
        const float Cdot = wB - wA;
        float impulse = -s.axial * (Cdot + maxf(C, 0.0f) * inv_dt);
        const float old = j.lower_imp;
        j.lower_imp = maxf(j.lower_imp + impulse, 0.0f);
        impulse = j.lower_imp - old;
        wA -= iA * impulse;
        wB += iB * impulse;
      }
      {
        const float C = j.upper - s.angle;
        const float Cdot = wA - wB;
        float impulse = -s.axial * (Cdot + maxf(C, 0.0f) * inv_dt);
        const float old = j.upper_imp;
        j.upper_imp = maxf(j.upper_imp + impulse, 0.0f);
        impulse = j.upper_imp - old;
        wA += iA * impulse;
        wB -= iB * impulse;
      }
    }
    {
      const vec2 rA = mk(s.ra_x, s.ra_y), rB = mk(s.rb_x, s.rb_y);
      const vec2 Cdot = vB + cross(wB, rB) - vA - cross(wA, rA);
      const mat22 K = {mk(s.kxx, s.kxy), mk(s.kyx, s.kyy)};
      const vec2 impulse = solve(K, -Cdot);
      j.imp_x += impulse.x;
      j.imp_y += impulse.y;
      vA = vA - mA * impulse;
      wA -= iA * cross(rA, impulse);
      vB = vB + mB * impulse;
      wB += iB * cross(rB, impulse);
    }
    d.joints[ji] = j;
    if (writes_body(d, j.a)) d.vel[j.a] = make_float4(vA.x, vA.y, wA, 0.0f);
    if (writes_body(d, j.b)) d.vel[j.b] = make_float4(vB.x, vB.y, wB, 0.0f);
  }
}

__global__ void __launch_bounds__(kBlock) k_joint_position(Dev d, int begin, int end, int it) {
  B2G_FOR(q, begin, end) {
    const int isl = d.jisland[q];
    if (isl < 0 || island_done(d, isl, it)) continue;
    island_rearm(d, isl, it);
    const Joint j = d.joints[d.jorder[q]];
    const JointSolve s = d.jsol[q];
    const float4 pa = d.pos[j.a], pb = d.pos[j.b];
    vec2 cA = mk(pa.x, pa.y), cB = mk(pb.x, pb.y);
    float aA = pa.z, aB = pb.z;
    float angular_error = 0.0f, position_error = 0.0f;
    const bool fixed_rotation = (s.ia + s.ib == 0.0f);
    if ((j.flags & 2) && !fixed_rotation) {
      const float angle = aB - aA - j.ref;
      float C = 0.0f;
      if (absf(j.upper - j.lower) < 2.0f * kAngularSlop)
        C = clampf(angle - j.lower, -kMaxAngularCorrection, kMaxAngularCorrection);
      else if (angle <= j.lower)
        C = clampf(angle - j.lower + kAngularSlop, -kMaxAngularCorrection, 0.0f);
      else if (angle >= j.upper)
        C = clampf(angle - j.upper - kAngularSlop, 0.0f, kMaxAngularCorrection);
      const float limit_impulse = -s.axial * C;
      aA -= s.ia * limit_impulse;
      aB += s.ib * limit_impulse;
      angular_error = absf(C);
    }
    {
      const rot2 qA = rot_of(aA), qB = rot_of(aB);
      const vec2 rA = mul(qA, mk(j.la_x, j.la_y) - mk(s.lca_x, s.lca_y));
      const vec2 rB = mul(qB, mk(j.lb_x, j.lb_y) - mk(s.lcb_x, s.lcb_y));
      const vec2 C = cB + rB - cA - rA;
      position_error = length(C);
      const float mA = s.ma, mB = s.mb, iA = s.ia, iB = s.ib;
      mat22 K;
      K.ex.x = mA + mB + iA * rA.y * rA.y + iB * rB.y * rB.y;
      K.ex.y = -iA * rA.x * rA.y - iB * rB.x * rB.y;
      K.ey.x = K.ex.y;
      K.ey.y = mA + mB + iA * rA.x * rA.x + iB * rB.x * rB.x;
      const vec2 impulse = -solve(K, C);
      cA = cA - mA * impulse;
      aA -= iA * cross(rA, impulse);
      cB = cB + mB * impulse;
      aB += iB * cross(rB, impulse);
    }
    if (writes_body(d, j.a)) d.pos[j.a] = make_float4(cA.x, cA.y, aA, 0.0f);
    if (writes_body(d, j.b)) d.pos[j.b] = make_float4(cB.x, cB.y, aB, 0.0f);
    const bool ok = position_error <= kLinearSlop && angular_error <= kAngularSlop;
    if (!ok) d.jok[island_slot(d, isl, it)] = 0;
  }
}

// Island of every solved contact slot and joint, and the island's position
// state reset (the islands come from a pass that runs beside the velocity
// phase; the position phase is the first to need them).
__global__ void __launch_bounds__(kBlock) k_island_assign(Dev d) {
  const int n = d.cnt[C_CONS];
  B2G_FOR(k, 0, n + d.nj) island_assign(d, n, k);
}


// write back: SynchronizeTransform (b2_body.h:957-961) and ClearForces
// (b2_world.cpp:1173-1180)
// (also the per-world colouring upkeep of the colour-sweep path, which may run
// any time between this step's colouring and the next)
// body b's transform from its final position, and whether the step moves it
__device__ __forceinline__ float4 synced_xf(const Dev& d, int b) {
  const float4 p = d.pos[b], pr = d.prop[b];
  const rot2 q = rot_of(p.z);
  const vec2 c = mk(p.x, p.y) - mul(q, mk(pr.z, pr.w));
  return make_float4(c.x, c.y, q.s, q.c);
}
__device__ __forceinline__ bool synced_body(const Dev& d, int b) {
  return world_active(d, d.bworld[b]) && d.btype[b] != kStaticBody;
}
__device__ __forceinline__ void finish_step(const Dev& d) {
  B2G_FOR(w, 0, d.nw) {
    if (world_active(d, w)) world_colour_stats(d, w, d.wcolmax[w] + 1);
  }
  B2G_FOR(b, 0, d.nb) {
    if (!world_active(d, d.bworld[b])) continue;
    if (d.btype[b] != kStaticBody && (!d.sleep_on || d.bactive[b])) d.xf[b] = synced_xf(d, b);
    d.force[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}
__global__ void __launch_bounds__(kBlock) k_finish_bodies(Dev d) { finish_step(d); }

// ===========================================================================
// Sleep (b2World::Solve's island seeds, b2_world.cpp:526-574; b2Island::Solve's
// sleep timers, b2_island.cpp:444-483; SolveOrphan's, :233-250), applied when
// a world allows sleeping or a body sleeps. The islands are the union-find
// labels of the step (every island gets its smallest body as the label).
// ===========================================================================
constexpr float kTimeToSleep = 0.5f;                                // b2_common.h:176
constexpr float kLinearSleepTolerance = 0.01f * 1.0f;               // b2_common.h:179
constexpr float kAngularSleepTolerance = 2.0f / 180.0f * kPi;       // b2_common.h:182

// before Collide: the step's per-body sleep records
__global__ void __launch_bounds__(kBlock) k_sleep_reset(Dev d) {
  B2G_FOR(b, 0, d.nb) {
    d.bncontact[b] = 0;
    d.bwake[b] = 0;
    d.bactive[b] = 0;
    d.isl_awake[b] = 0;
    d.isl_min[b] = 0x7f7fffffu;  // FLT_MAX
    d.isl_cons[b] = 0;
  }
}

// after the islands: Collide's wake-ups (b2Body::SetAwake(true) resets the
// timer), then every island with an awake body is seeded
__global__ void __launch_bounds__(kBlock) k_sleep_seed(Dev d) {
  B2G_FOR(b, 0, d.nb) {
    if (!world_active(d, d.bworld[b]) || d.btype[b] == kStaticBody) continue;
    if (d.bwake[b]) {
      d.bsflag[b] &= ~kBodyAsleep;
      d.bstime[b] = 0.0f;
    }
    if ((d.bsflag[b] & kBodyAsleep) == 0) d.isl_awake[d.parent[b]] = 1;
  }
}

// a step without a solve (dt <= 0): Collide's wake-ups only
__global__ void __launch_bounds__(kBlock) k_sleep_wake(Dev d) {
  B2G_FOR(b, 0, d.nb) {
    if (!world_active(d, d.bworld[b]) || d.btype[b] == kStaticBody || !d.bwake[b]) continue;
    d.bsflag[b] &= ~kBodyAsleep;
    d.bstime[b] = 0.0f;
  }
}

// the bodies of seeded islands are solved this step: the search wakes them
// (without a timer reset) and their velocities are integrated
__global__ void __launch_bounds__(kBlock) k_sleep_activate(Dev d) {
  B2G_FOR(b, 0, d.nb) {
    if (!world_active(d, d.bworld[b]) || d.btype[b] == kStaticBody) continue;
    if (!d.isl_awake[d.parent[b]]) continue;
    d.bactive[b] = 1;
    d.bsflag[b] &= ~kBodyAsleep;
    integrate_velocity(d, b, 2);
  }
}

// b2Body::SetAwake(false) (include/box2d/b2_body.h:719-740)
__device__ __forceinline__ void put_to_sleep(const Dev& d, int b) {
  d.bsflag[b] |= kBodyAsleep;
  d.bstime[b] = 0.0f;
  d.vel[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  d.force[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
__device__ __forceinline__ bool sleep_candidate(const Dev& d, int b) {
  return world_active(d, d.bworld[b]) && world_sleeps(d, d.bworld[b]) && d.btype[b] != kStaticBody &&
         d.bactive[b];
}
// an orphan has no contact and no joint (b2_world.cpp:545-549)
__device__ __forceinline__ bool orphan_body(const Dev& d, int b) { return d.bncontact[b] == 0 && !d.bjointed[b]; }

// after the position iterations: the sleep timers, and each island's least
// timer. An orphan runs its own: when it runs out, the body's velocities and
// forces are cleared, and the seed flags set right after SolveOrphan keep it
// awake (b2_world.cpp:547-549).
__global__ void __launch_bounds__(kBlock) k_sleep_times(Dev d) {
  const float lin_tol_sq = kLinearSleepTolerance * kLinearSleepTolerance;
  const float ang_tol_sq = kAngularSleepTolerance * kAngularSleepTolerance;
  B2G_FOR(b, 0, d.nb) {
    if (!sleep_candidate(d, b)) continue;
    const float4 v = d.vel[b];
    const bool slow = (d.bsflag[b] & kBodyNoAutoSleep) == 0 && !(v.z * v.z > ang_tol_sq) &&
                      !(v.x * v.x + v.y * v.y > lin_tol_sq);
    const float h = d.wpar[d.bworld[b]].z;
    if (orphan_body(d, b)) {
      if (!slow) {
        d.bstime[b] = 0.0f;
      } else {
        const float t = d.bstime[b] + h;
        d.bstime[b] = t;
        if (t >= kTimeToSleep) {
          put_to_sleep(d, b);
          d.bsflag[b] &= ~kBodyAsleep;
        }
      }
      continue;
    }
    const int r = d.parent[b];
    if (!slow) {
      d.bstime[b] = 0.0f;
      atomicMin(&d.isl_min[r], 0u);
    } else {
      const float t = d.bstime[b] + h;
      d.bstime[b] = t;
      atomicMin(&d.isl_min[r], __float_as_uint(t));  // non-negative floats order as their bits
    }
  }
}

// an island whose bodies all slept long enough and whose position iterations
// converged goes to sleep; an island without constraints converges at its
// first position iteration
__global__ void __launch_bounds__(kBlock) k_sleep_apply(Dev d, int pi) {
  B2G_FOR(b, 0, d.nb) {
    if (!sleep_candidate(d, b) || orphan_body(d, b)) continue;
    const int r = d.parent[b];
    bool solved;
    if (!d.isl_cons[r]) {
      solved = pi > 0;
    } else {
      const size_t last = island_slot(d, r, pi > 0 ? pi - 1 : 0);
      solved = d.solved[r] != 0 || (pi > 0 && d.sepbad[last] == 0u && d.jok[last] != 0);
    }
    if (__uint_as_float(d.isl_min[r]) >= kTimeToSleep && solved) put_to_sleep(d, b);
  }
}

// ClearForces alone (b2_world.cpp:1173-1180), for a step without a solve
__global__ void __launch_bounds__(kBlock) k_clear_forces(Dev d) {
  B2G_FOR(b, 0, d.nb) {
    if (world_active(d, d.bworld[b])) d.force[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}


// ===========================================================================
// Per-world solver: one workgroup runs a whole world's solve (b2Island::Solve,
// src/dynamics/b2_island.cpp:253-438, with the contact and position phases
// of b2ContactSolver) with workgroup barriers between colours instead of one
// launch per colour. Constraint slots are world-major (each world's
// constraints are contiguous, grouped by colour inside), so a world reads one
// contiguous slice. Results equal the colour-major path's: constraints of one
// colour share no body, the serial group keeps its per-world order, and the
// island minimum separation is an order-free atomic minimum.
// ===========================================================================
// Slots of world w's constraints (world-major, colour-minor): the touching
// contacts of its shapes are contiguous in the key-sorted contact list and
// keep their order in the compacted list. Also persists the colours and
// advances the claim-round base (block 0).
__global__ void __launch_bounds__(kBlock) k_world_colour_sort(Dev d, int rounds_run) {
  __shared__ int hist[kColourSlots];
  __shared__ int off[kColourSlots + 1];
  __shared__ int cur[kColourSlots];
  __shared__ int range[2];
  const int w = d.worder[blockIdx.x];
  if (!world_active(d, w)) {  // no touching contact (the narrowphase left them all untouched): no colours
    if (threadIdx.x == 0) {
      if (blockIdx.x == 0) d.cnt[C_ROUND] += rounds_run;
      d.wcolmax[w] = -1;
    }
    return;  // the whole block
  }
  if (threadIdx.x == 0) {
    if (blockIdx.x == 0) d.cnt[C_ROUND] += rounds_run;
    const int nc = d.cnt[C_CONTACTS], total = d.cnt[C_CONS];
    const int2 cr = world_contacts(d, w);
    range[0] = cr.x < nc ? d.tpos[cr.x] : total;
    range[1] = cr.y < nc ? d.tpos[cr.y] : total;
  }
  for (int c = threadIdx.x; c < kColourSlots; c += blockDim.x) {
    hist[c] = 0;
    cur[c] = 0;
  }
  __syncthreads();
  const int t0 = range[0], t1 = range[1];
  for (int k = t0 + (int)threadIdx.x; k < t1; k += blockDim.x) {
    int col = d.ccolour[k];
    if (col < 0) col = kSerialColour;  // unreachable once colouring has finished
    atomicAdd(&hist[col], 1);
    d.ccol[d.clist[k]] = col;
  }
  __syncthreads();
  static_assert(kMaxColours == kWave, "one lane per normal colour");
  if (threadIdx.x < kWave) {  // the first wave: offsets of the 64 normal colours, then the serial group
    const int c = (int)threadIdx.x;
    const int h = hist[c];
    const int inc = wave_inclusive_scan(h);
    off[c] = inc - h;
    const uint64_t used = __ballot(h != 0);
    const int topn = used ? 64 - __clzll((long long)used) : 0;  // normal colours only
    if (c == kWave - 1) {
      const int ser = hist[kSerialColour];
      off[kSerialColour] = inc;
      off[kColourSlots] = inc + ser;
      const int top = ser ? kColourSlots : topn;
      if (top) atomicMax(&d.cnt[C_NCOLOURS], top);
      if (ser) atomicAdd(&d.cnt[C_SERIAL], ser);
      world_colour_stats(d, w, topn);  // the colouring upkeep, for this world right away
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c <= kColourSlots; c += blockDim.x) d.wcol[w * (kColourSlots + 1) + c] = t0 + off[c];
  for (int k = t0 + (int)threadIdx.x; k < t1; k += blockDim.x) {
    int col = d.ccolour[k];
    if (col < 0) col = kSerialColour;
    d.csorted[t0 + off[col] + atomicAdd(&cur[col], 1)] = d.clist[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {  // the serial group in contact order
    const int b = t0 + off[kSerialColour], e = t0 + off[kSerialColour + 1];
    for (int a = b + 1; a < e; ++a) {
      const int v = d.csorted[a];
      int q = a - 1;
      while (q >= b && d.csorted[q] > v) {
        d.csorted[q + 1] = d.csorted[q];
        --q;
      }
      d.csorted[q + 1] = v;
    }
  }
}

// The solve of world worder[blockIdx.x]: warm start, velocity iterations,
// store impulses + integrate positions, position iterations with the
// per-island early exit, write-back. Worlds without joints only. With kLds
// the body velocities (then positions) of a world of at most `lds_bodies`
// bodies live in LDS during the sweeps (dynamic shared memory of 16 bytes per
// body); a larger world of the launch reads them from global memory. The
// arithmetic is the same either way.
// (At least 5 waves per SIMD: 96 registers and a few spills measured
// slightly faster than 4 waves without spills; the CU hides one world's
// latency behind the others'.)
template <bool kLds>
__global__ void __launch_bounds__(1024, 5) k_solve_worlds(Dev d, int vi, int pi, int first, const int* order,
                                                          int lds_bodies) {
  __shared__ int off[kColourSlots + 1];
  extern __shared__ float4 s_body[];
  const int w = order[first + blockIdx.x];  // worlds first .. first + gridDim.x - 1, largest first
  if (!world_active(d, w)) return;  // whole block
  for (int c = threadIdx.x; c <= kColourSlots; c += blockDim.x) off[c] = d.wcol[w * (kColourSlots + 1) + c];
  __syncthreads();
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  const int t0 = off[0], t1 = off[kColourSlots];
  const int2 wb = d.wbody[w];
  const bool lds = kLds && wb.y - wb.x <= lds_bodies;  // uniform across the block
  // one colour at a time; the serial group on one lane, in slot order
  auto sweep = [&](auto fn) {
    for (int c = 0; c < kColourSlots; ++c) {
      const int lo = off[c], hi = off[c + 1];
      if (hi <= lo) continue;  // uniform across the block
      if (c == kSerialColour) {
        if (tid == 0)
          for (int s = lo; s < hi; ++s) fn(s);
      } else {
        for (int s = lo + tid; s < hi; s += nt) fn(s);
      }
      __syncthreads();
    }
  };
  auto stage_in = [&](const float4* src) {
    for (int b = wb.x + tid; b < wb.y; b += nt) s_body[b - wb.x] = src[b];
    __syncthreads();
  };
  auto stage_out = [&](float4* dst) {
    for (int b = wb.x + tid; b < wb.y; b += nt) dst[b] = s_body[b - wb.x];
    __syncthreads();
  };
  if (t1 > t0) {
    if (lds) {
      stage_in(d.vel);
      const LdsBodies vel{s_body, wb.x};
      sweep([&](int s) { warm_start_t(d, s, vel); });
      for (int it = 0; it < vi; ++it) sweep([&](int s) { contact_velocity_t(d, s, vel); });
      stage_out(d.vel);
    } else {
      sweep([&](int s) { warm_start(d, s); });
      for (int it = 0; it < vi; ++it) sweep([&](int s) { contact_velocity(d, s); });
    }
  }
  for (int s = t0 + tid; s < t1; s += nt) store_impulses(d, s);
  for (int b = wb.x + tid; b < wb.y; b += nt) integrate_position(d, b);
  __syncthreads();
  if (t1 > t0) {
    if (lds) stage_in(d.pos);
    const LdsBodies lpos{s_body, wb.x};
    const GlobalBodies gpos{d.pos};
    for (int it = 0; it < pi; ++it) {
      for (int c = 0; c < kColourSlots; ++c) {
        const int lo = off[c], hi = off[c + 1];
        if (hi <= lo) continue;
        if (c == kSerialColour) {
          if (tid == 0)
            for (int s = lo; s < hi; ++s) {
              const int isl = d.cisland[s];
              const float m =
                  lds ? contact_position_t(d, s, isl, it, lpos) : contact_position_t(d, s, isl, it, gpos);
              if (separation_bad(m)) d.sepbad[island_slot(d, isl, it)] = 1u;
            }
        } else {
          for (int base = lo; base < hi; base += nt) {  // every lane of a wave takes part in the min
            const int s = base + tid;
            const bool valid = s < hi;
            const int isl = valid ? d.cisland[s] : 0;
            float m = 0.0f;
            if (valid) m = lds ? contact_position_t(d, s, isl, it, lpos) : contact_position_t(d, s, isl, it, gpos);
            mark_island_sep(d, it, isl, valid, m);
          }
        }
        __syncthreads();
      }
      // per-island early exit (b2_island.cpp:403-408) after this iteration,
      // and the slot of iteration it + 1 cleared (see island_done)
      for (int k = t0 + tid; k < t1; k += nt) {
        const int isl = d.cisland[k];
        const size_t nxt = island_slot(d, isl, it + 1);
        d.sepbad[nxt] = 0u;
        d.jok[nxt] = 1;
        if (d.solved[isl]) continue;
        const size_t cur = island_slot(d, isl, it);
        if (d.sepbad[cur] == 0u && d.jok[cur] != 0) d.solved[isl] = 1;
      }
      __syncthreads();
    }
    if (lds) stage_out(d.pos);
  }
  // write back: SynchronizeTransform and ClearForces
  for (int b = wb.x + tid; b < wb.y; b += nt) {
    if (d.btype[b] != kStaticBody) {
      const float4 p = d.pos[b], pr = d.prop[b];
      const rot2 q = rot_of(p.z);
      const vec2 c = mk(p.x, p.y) - mul(q, mk(pr.z, pr.w));
      d.xf[b] = make_float4(c.x, c.y, q.s, q.c);
    }
    d.force[b] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

// ===========================================================================
// Broadphase: LBVH over Morton-sorted tight AABBs (the pair set that
// b2BroadPhase::UpdateAndQuery reports, include/box2d/b2_broad_phase.h:252-511,
// filtered as b2ContactManager::QueryCallback filters it,
// src/dynamics/b2_contact_manager.cpp:136-188).
// ===========================================================================
// Worlds that sit out a step. With d.idle set, such a world has not moved
// since the last broadphase and nothing was uploaded since: its pairs are its
// segments of the current contact list, and its boxes, its leaf order (the
// keys of its last build, which the stable sort puts back in the same order),
// its internal nodes and its sparse-table entries all stand.
__device__ __forceinline__ bool idle_world(const Dev& d, int w) { return d.idle && !world_active(d, w); }
// Shape s's segment [first, end) of the current contact list, which is the
// last broadphase's pair list (valid while d.idle is set).
__device__ __forceinline__ int2 prev_segment(const Dev& d, int s) {
  const int n = d.cnt[C_CONTACTS];  // the current list (its count is replaced after the merge)
  const int first = min(d.poff_prev[s], n);
  const int end = s + 1 < d.ns ? min(d.poff_prev[s + 1], n) : n;
  return make_int2(first, end);
}

__global__ void __launch_bounds__(kBlock) k_world_bounds_reset(Dev d) {
  B2G_FOR(w, 0, d.nw) {
    d.wbounds[4 * w + 0] = 0xffffffffu;
    d.wbounds[4 * w + 1] = 0xffffffffu;
    d.wbounds[4 * w + 2] = 0u;
    d.wbounds[4 * w + 3] = 0u;
  }
}

// ComputeAABB of the three shapes (circle :91-96, polygon :373-388, edge :156-167)
__device__ __forceinline__ float4 shape_aabb(int type, int count, const float4* g, const xform& xf) {
  const float4 h = g[0];
  const float r = h.x;
  if (type == kCircle) {
    const vec2 p = xf.p + mul(xf.q, mk(h.y, h.z));
    return make_float4(p.x - r, p.y - r, p.x + r, p.y + r);
  }
  if (type == kEdge) {
    const float4 e = g[1], f = g[2];
    const vec2 v1 = mul(xf, mk(e.z, e.w));
    const vec2 v2 = mul(xf, mk(f.x, f.y));
    const vec2 lo = vmin(v1, v2), hi = vmax(v1, v2);
    return make_float4(lo.x - r, lo.y - r, hi.x + r, hi.y + r);
  }
  const float4 a0 = g[1];
  vec2 lo = mul(xf, mk(a0.x, a0.y));
  vec2 hi = lo;
#pragma unroll
  for (int i = 1; i < kMaxVerts; ++i) {
    if (i < count) {
      const float4 a = g[1 + (i >> 1)];
      const vec2 v = mul(xf, (i & 1) ? mk(a.z, a.w) : mk(a.x, a.y));
      lo = vmin(lo, v);
      hi = vmax(hi, v);
    }
  }
  return make_float4(lo.x - r, lo.y - r, hi.x + r, hi.y + r);
}

// Shape AABBs. For a full tree build (`bounds`), also the world bounds of the
// AABB centres (2 * centre, as b2BroadPhase::GetCenter2), which only the
// Morton keys use; they are reduced per wave, then per block (runs of equal
// worlds merge), so one world's shapes do not all hit the same four words.
// For a refit, each box also goes straight to its leaf of the sparse table
// (the leaf order is the last build's) with the per-broadphase resets of
// k_leaf_aabbs, which then has no launch of its own.
// With `finish`, the launch also ends the step (k_finish_bodies' work): the
// shapes then take their bodies' transforms from the final positions, as the
// body threads compute them, rather than read them while they are written.
__global__ void __launch_bounds__(kBlock) k_aabb(Dev d, int bounds, int finish) {
  if (!bounds && blockIdx.x == 0 && threadIdx.x == 0) d.cnt[C_HEAVY] = 0;
  if (finish) finish_step(d);
  constexpr int kWaves = kBlock / kWave;
  __shared__ int s_w[kWaves];
  __shared__ float s_b[kWaves][4];
  B2G_FOR_ALL(i, valid, 0, d.ns) {
    int w = 0;
    float cx = 0.0f, cy = 0.0f;
    bool live = valid;
    if (valid) {
      const int4 h = d.shdr[i];
      if (idle_world(d, h.w)) {
        // box, leaf and tree stand; the pair count is the old segment's length
        const int2 seg = prev_segment(d, i);
        d.pcount[i] = seg.y - seg.x;
        live = false;
      } else {
        const float4 xf = finish && synced_body(d, h.y) ? synced_xf(d, h.y) : d.xf[h.y];
        const float4 bb = shape_aabb(h.x, h.z, d.geo + 9 * (size_t)i, xform_of(xf));
        d.aabb[i] = bb;
        if (!bounds) {
          d.sparse[d.sinv[i]] = bb;
          d.pcount[i] = 0;
        }
        cx = bb.x + bb.z;
        cy = bb.y + bb.w;
        w = h.w;
      }
    }
    if (!bounds) continue;
    unsigned* wb = d.wbounds;
    const uint64_t act = __ballot(live);
    const int first = act ? __ffsll((long long)act) - 1 : 0;
    const int key0 = __shfl(w, first, kWave);
    const bool uniform = __ballot(live && w != key0) == 0;
    const float x0 = wave_min_all(live ? cx : FLT_MAX), y0 = wave_min_all(live ? cy : FLT_MAX);
    const float x1 = wave_max_all(live ? cx : -FLT_MAX), y1 = wave_max_all(live ? cy : -FLT_MAX);
    if (!uniform && live) {  // several worlds in this wave
      atomicMin(wb + 4 * w + 0, float_to_ordered(cx));
      atomicMin(wb + 4 * w + 1, float_to_ordered(cy));
      atomicMax(wb + 4 * w + 2, float_to_ordered(cx));
      atomicMax(wb + 4 * w + 3, float_to_ordered(cy));
    }
    if (lane_id() == 0) {
      s_w[wave_id()] = act && uniform ? key0 : -1;
      s_b[wave_id()][0] = x0;
      s_b[wave_id()][1] = y0;
      s_b[wave_id()][2] = x1;
      s_b[wave_id()][3] = y1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int k = -1;
      float m[4] = {FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX};
      for (int v = 0; v <= kWaves; ++v) {
        const int kv = v < kWaves ? s_w[v] : -1;
        if (k >= 0 && kv != k) {  // flush the run of world k
          atomicMin(wb + 4 * k + 0, float_to_ordered(m[0]));
          atomicMin(wb + 4 * k + 1, float_to_ordered(m[1]));
          atomicMax(wb + 4 * k + 2, float_to_ordered(m[2]));
          atomicMax(wb + 4 * k + 3, float_to_ordered(m[3]));
          m[0] = m[1] = FLT_MAX;
          m[2] = m[3] = -FLT_MAX;
        }
        k = kv;
        if (kv >= 0) {
          m[0] = fminf(m[0], s_b[v][0]);
          m[1] = fminf(m[1], s_b[v][1]);
          m[2] = fmaxf(m[2], s_b[v][2]);
          m[3] = fmaxf(m[3], s_b[v][3]);
        }
      }
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kBlock) k_world_bounds_from_aabbs(Dev d) {
  B2G_FOR(i, 0, d.ns) {
    const float4 bb = d.aabb[i];
    const float cx = bb.x + bb.z, cy = bb.y + bb.w;
    unsigned* wb = d.wbounds + 4 * d.shdr[i].w;
    atomicMin(&wb[0], float_to_ordered(cx));
    atomicMin(&wb[1], float_to_ordered(cy));
    atomicMax(&wb[2], float_to_ordered(cx));
    atomicMax(&wb[3], float_to_ordered(cy));
  }
}

__device__ __forceinline__ uint32_t spread15(uint32_t x) {
  x &= 0x7fffu;
  x = (x | (x << 8)) & 0x00ff00ffu;
  x = (x | (x << 4)) & 0x0f0f0f0fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}

// With `compact`, only the active worlds' shapes are sorted: each writes its
// key to its world's block of the active keys (mkey[1]; the layout is world
// after world), and k_place_active_leaves puts the sorted blocks back.
__global__ void __launch_bounds__(kBlock) k_morton(Dev d, int compact) {
  B2G_FOR(i, 0, d.ns) {
    const int4 h = d.shdr[i];
    if (idle_world(d, h.w)) {  // the key of its world's last build
      if (!compact) {
        d.mkey[0][i] = d.skey[i];
        d.mval[0][i] = (uint32_t)i;
      }
      continue;
    }
    const float4 bb = d.aabb[i];
    const unsigned* wb = d.wbounds + 4 * h.w;
    const float x0 = ordered_to_float(wb[0]), y0 = ordered_to_float(wb[1]);
    const float x1 = ordered_to_float(wb[2]), y1 = ordered_to_float(wb[3]);
    const float sx = x1 > x0 ? 32767.0f / (x1 - x0) : 0.0f;
    const float sy = y1 > y0 ? 32767.0f / (y1 - y0) : 0.0f;
    const float cx = bb.x + bb.z, cy = bb.y + bb.w;
    const uint32_t qx = (uint32_t)minf(maxf((cx - x0) * sx, 0.0f), 32767.0f);
    const uint32_t qy = (uint32_t)minf(maxf((cy - y0) * sy, 0.0f), 32767.0f);
    uint32_t code = spread15(qx) | (spread15(qy) << 1);
    // a shape spanning more than 1/8 of its world's extent (a ground, a
    // wall) goes into a subtree of its own next to the rest: its box would
    // otherwise enlarge every node on its path, and every query would walk
    // that path
    if ((bb.z - bb.x) * 8.0f > x1 - x0 || (bb.w - bb.y) * 8.0f > y1 - y0) code |= 1u << 30;
    const uint64_t key = ((uint64_t)(uint32_t)h.w << 32) | code;
    d.skey[i] = key;
    if (compact) {
      const int c = d.wsort[h.w] + (i - d.wshape0[h.w]);
      d.mkey[1][c] = key;
      d.mval[1][c] = (uint32_t)i;
    } else {
      d.mkey[0][i] = key;
      d.mval[0][i] = (uint32_t)i;
    }
  }
}

// per world: its shape count if active, else 0 (scanned next)
__global__ void __launch_bounds__(kBlock) k_active_world_sizes(Dev d) {
  B2G_FOR(w, 0, d.nw) d.wsort[w] = world_active(d, w) ? d.wshape1[w] - d.wshape0[w] : 0;
}

// The sorted active keys back into the sorted order: world w's leaves are
// the sorted positions [wleaf[2w], wleaf[2w + 1]), a range fixed while the
// layout stands, and the sort kept each world's keys together and in order.
__global__ void __launch_bounds__(kBlock) k_place_active_leaves(Dev d, int m) {
  B2G_FOR(k, 0, m) {
    const uint64_t key = d.mkey[1][k];
    const int w = (int)(key >> 32);
    const int pos = d.wleaf[2 * w] + (k - d.wsort[w]);
    if (pos < d.wleaf[2 * w] || pos >= d.wleaf[2 * w + 1]) continue;  // (cannot happen)
    d.mkey[0][pos] = key;
    d.mval[0][pos] = d.mval[1][k];
  }
}

// delta of Karras (2012): common prefix of (key, index) at sorted positions i, j
__device__ __forceinline__ int lbvh_delta(const uint64_t* __restrict__ k, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t x = k[i] ^ k[j];
  if (x == 0) return 64 + __clz((unsigned)(i ^ j));
  return __clzll((long long)x);
}

__global__ void __launch_bounds__(kBlock) k_lbvh_build(Dev d, const uint64_t* __restrict__ keys) {
  const int n = d.ns;
  B2G_FOR(i, 0, n - 1) {
    // A node whose neighbours on both sides are leaves of its own world lies
    // inside that world's subtree: world bits lead the key, so every
    // cross-world prefix is shorter than any prefix inside a world, and the
    // node depends only on its world's keys. An idle world's keys are those
    // of its last build, so its nodes stand. (The world's two end positions
    // hold its root and a node joining worlds; they are rebuilt.)
    if (d.idle && i > 0) {
      const uint64_t w = keys[i] >> 32;
      if ((keys[i - 1] >> 32) == w && (keys[i + 1] >> 32) == w && !world_active(d, (int)w)) continue;
    }
    const int dir = (lbvh_delta(keys, n, i, i + 1) - lbvh_delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = lbvh_delta(keys, n, i, i - dir);
    int lmax = 2;
    while (lbvh_delta(keys, n, i, i + lmax * dir) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
      if (lbvh_delta(keys, n, i, i + (l + t) * dir) > dmin) l += t;
    const int j = i + l * dir;
    const int dnode = lbvh_delta(keys, n, i, j);
    int s = 0;
    int t = l;
    do {
      t = (t + 1) >> 1;
      if (lbvh_delta(keys, n, i, i + (s + t) * dir) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * dir + (dir < 0 ? -1 : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const int left = (lo == gamma) ? ~gamma : gamma;
    const int right = (hi == gamma + 1) ? ~(gamma + 1) : gamma + 1;
    d.node_child[2 * i] = left;
    d.node_child[2 * i + 1] = right;
    d.node_range[i] = make_int2(lo, hi);
    if (left < 0) d.leaf_parent[~left] = i; else d.node_parent[left] = i;
    if (right < 0) d.leaf_parent[~right] = i; else d.node_parent[right] = i;
    if (i == 0) d.node_parent[0] = -1;
  }
}

// leaf ranges of each world in sorted order
__global__ void __launch_bounds__(kBlock) k_world_leaves(Dev d, const uint64_t* __restrict__ keys) {
  const int n = d.ns;
  B2G_FOR(k, 0, n) {
    const int w = (int)(keys[k] >> 32);
    if (idle_world(d, w)) continue;  // its range and root stand
    if (k == 0 || (int)(keys[k - 1] >> 32) != w) {
      d.wleaf[2 * w] = k;
      d.wroot[w] = -1;  // k_node_records finds the world's subtree
    }
    if (k == n - 1 || (int)(keys[k + 1] >> 32) != w) d.wleaf[2 * w + 1] = k + 1;
  }
}

// Node bounds without a bottom-up walk: every internal node covers a
// contiguous range of sorted leaves, so its AABB is a range union, answered in
// O(1) from a sparse table (level j holds the union of 2^j leaves starting at
// each position). Building the table is log2(n) element-wise passes (four per
// launch); no atomics, fences or inter-block ordering are involved.
__device__ __forceinline__ float4 aabb_union(float4 a, float4 b) {
  return make_float4(minf(a.x, b.x), minf(a.y, b.y), maxf(a.z, b.z), maxf(a.w, b.w));
}

// level 0 of the sparse table: leaf boxes in sorted order (a full build and a
// refit alike); also clears the pair counts, which static shapes' partners
// count up in k_lbvh_pairs, and the list of pair segments to sort
__global__ void __launch_bounds__(kBlock) k_leaf_aabbs(Dev d, const uint32_t* __restrict__ vals) {
  if (blockIdx.x == 0 && threadIdx.x == 0) d.cnt[C_HEAVY] = 0;
  B2G_FOR(k, 0, d.ns) {
    if (d.idle && !world_active(d, (int)(d.mkey[0][k] >> 32))) continue;  // leaf stands (k_aabb set its count)
    const uint32_t i = vals[k];
    d.sparse[k] = d.aabb[i];
    d.pcount[i] = 0;
    d.sinv[i] = k;
  }
}

// levels j .. j + 3 from level j - 1 (2, 4, 8 and 16 entries of it; indices
// past the end clamp to the last leaf, which the unions absorb)
constexpr int kSparseLevelsPerLaunch = 4;
__global__ void __launch_bounds__(kBlock) k_sparse_levels(Dev d, int j) {
  const int n = d.ns;
  const int h = 1 << (j - 1);
  const float4* prev = d.sparse + (size_t)(j - 1) * n;
  B2G_FOR(i, 0, n) {
    // an idle world's entries stand (those that reach past the world's last
    // leaf are never read)
    if (d.idle && !world_active(d, (int)(d.mkey[0][i] >> 32))) continue;
    // level j + lv is the union of 2^(lv + 1) entries of level j - 1, h apart
    float4 u = prev[i];
    int have = 1;
    for (int lv = 0; lv < kSparseLevelsPerLaunch && j + lv < d.sparse_levels; ++lv) {
      for (int k = have; k < 2 * have; ++k) u = aabb_union(u, prev[min(i + k * h, n - 1)]);
      have *= 2;
      d.sparse[(size_t)(j + lv) * n + i] = u;
    }
  }
}

// Node records: every internal node stores both children's boxes, the
// children (internal index, or ~shape index for a leaf) and its split, so a
// traversal step is one 48-byte read. Every node covers a contiguous range of
// sorted leaves, so a child's box is a range union answered in O(1) from the
// sparse table; a child spanning several worlds gets an unbounded box (the
// traversal's split test keeps queries inside their own world's leaf range).
__device__ __forceinline__ float4 range_box(const Dev& d, const uint64_t* keys, int n, int lo, int hi) {
  if ((keys[lo] >> 32) != (keys[hi] >> 32)) return make_float4(-FLT_MAX, -FLT_MAX, FLT_MAX, FLT_MAX);
  const int len = hi - lo + 1;
  const int lev = 31 - __clz(len);
  if (lev >= d.sparse_levels) return make_float4(-FLT_MAX, -FLT_MAX, FLT_MAX, FLT_MAX);  // not built: no bound
  const float4* t = d.sparse + (size_t)lev * n;
  return aabb_union(t[lo], t[hi - (1 << lev) + 1]);
}

__global__ void __launch_bounds__(kBlock) k_node_records(Dev d, const uint64_t* __restrict__ keys,
                                                         const uint32_t* __restrict__ vals) {
  const int n = d.ns;
  B2G_FOR(i, 0, n - 1) {
    const int2 r = d.node_range[i];
    if (d.idle) {  // the nodes inside an idle world stand
      const int w = (int)(keys[r.x] >> 32);
      if ((int)(keys[r.y] >> 32) == w && !world_active(d, w)) continue;
    }
    const int lc = d.node_child[2 * i], rc = d.node_child[2 * i + 1];
    const int split = lc < 0 ? ~lc : d.node_range[lc].y;
    const float4 lb = range_box(d, keys, n, r.x, split);
    const float4 rb = range_box(d, keys, n, split + 1, r.y);
    const int lcode = lc < 0 ? ~(int)vals[~lc] : lc;
    const int rcode = rc < 0 ? ~(int)vals[~rc] : rc;
    // a leaf child's static flag, so a query can pass over it without its filter word
    const int sflags = (lc < 0 && (d.sfast[~lcode].y & 64) ? 1 : 0) | (rc < 0 && (d.sfast[~rcode].y & 64) ? 2 : 0);
    float4* o = d.node_rec + 3 * (size_t)i;
    o[0] = lb;
    o[1] = rb;
    o[2] = make_float4(__int_as_float(lcode), __int_as_float(rcode), __int_as_float(split), __int_as_float(sflags));
    // keys sharing the world prefix form a subtree of the radix tree: its
    // root is where that world's queries start
    const int w = (int)(keys[r.x] >> 32);
    if (d.wleaf[2 * w] == r.x && d.wleaf[2 * w + 1] == r.y + 1) d.wroot[w] = i;
  }
}

// Per-shape filter word for the pair test (one 16-byte read per candidate).
__global__ void __launch_bounds__(kBlock) k_shape_filters(Dev d) {
  B2G_FOR(i, 0, d.ns) {
    const int4 h = d.shdr[i];
    const int4 f = d.sfilt[i];
    const int b = h.y;
    const int dyn = d.btype[b] == kDynamicBody ? 1 : 0;
    const int jp = d.bj_off[b + 1] > d.bj_off[b] ? 1 : 0;
    const int stat = d.btype[b] == kStaticBody ? 1 : 0;
    d.sfast[i] = make_int4(b, (h.x & 15) | (dyn << 4) | (jp << 5) | (stat << 6), (f.x & 0xffff) | ((f.y & 0xffff) << 16),
                           f.z);
  }
}

__device__ __forceinline__ bool overlap(float4 a, float4 b) {  // b2TestOverlap, inclusive
  return (a.z >= b.x) & (a.x <= b.z) & (a.w >= b.y) & (a.y <= b.w);
}

// b2Body::ShouldCollide (src/dynamics/b2_body.cpp:459-480) +
// b2ContactFilter::ShouldCollide (src/dynamics/b2_world_callbacks.cpp:28-40)
// + existence of a collide function (src/dynamics/b2_contact.cpp:48-77), on
// the packed per-shape filter words
__device__ __forceinline__ bool pair_allowed(const Dev& d, int4 fa, int4 fb) {
  if (fa.x == fb.x) return false;                          // same body
  if (((fa.y | fb.y) & 16) == 0) return false;             // neither body dynamic
  const int ta = fa.y & 15, tb = fb.y & 15;
  if (!has_collide_function(ta, tb) && !has_collide_function(tb, ta)) return false;
  if (fa.w == fb.w && fa.w != 0) {
    if (fa.w <= 0) return false;
  } else {
    const int cata = fa.z & 0xffff, maska = (int)((unsigned)fa.z >> 16);
    const int catb = fb.z & 0xffff, maskb = (int)((unsigned)fb.z >> 16);
    if ((maska & catb) == 0 || (cata & maskb) == 0) return false;
  }
  if ((fa.y | fb.y) & 32) {  // a joint with collideConnected == false between the bodies
    for (int e = d.bj_off[fa.x]; e < d.bj_off[fa.x + 1]; ++e)
      if (d.bj_other[e] == fb.x) return false;
  }
  return true;
}

constexpr int kStack = 64;
constexpr int kQueryBlock = 128;  // LDS traversal stack: kStack x kQueryBlock ints (32 KB)
// Pair slots per query lane (LDS): 16 slots (48 KB per block with the stack)
// re-traverse fewer shapes; 8 (40 KB) fit four blocks per CU instead of
// three, which pays once the queries fill the GPU several times over.
constexpr int kPairSlots = 16;
constexpr int kPairSlotsDense = 8;
constexpr int kDenseQueryShapes = 65536;  // from this many shapes on, 8 slots

// Pair ownership. Every pair is found by exactly one traversal: a shape
// queries for partners j > i and for static partners of any index, and a
// static shape does not query. A static shape (a ground, a wall) can overlap
// thousands of shapes; its partners each find it at the cost of one more
// leaf test, instead of one lane walking all of them. The pair list stays
// sorted by key (lower shape << 32 | higher shape): a key with a static lower
// shape s belongs to s's segment, whose length the finders count up
// (pcount[s]) and whose order k_sort_pair_segments restores.
//
// The traversal stack lives in LDS, thread-minor (stk[level * blockDim +
// thread]), so lanes at different depths do not conflict. `emit(key, s)` is
// called for every pair found; s >= 0 is the static lower shape of a key that
// is not the querying shape's own.
// Every internal node of a world's subtree covers leaves of that world only,
// and a node's index is one end of its leaf range, so the nodes a query
// walks lie in the world's leaf range: with a 16-bit stack (worlds of at most
// 65536 shapes) entries are offsets from the world's first leaf, and the
// stack takes half the LDS.
template <class S, class Emit>
__device__ __forceinline__ void lbvh_query(const Dev& d, int i, S* stk, Emit emit) {
  const int n = d.ns;
  if (n < 2) return;
  const int4 fi = d.sfast[i];
  if (fi.y & 64) return;  // static: found by its partners
  const float4 q = d.aabb[i];
  const int w = d.shdr[i].w;
  const int wlo = d.wleaf[2 * w], whi = d.wleaf[2 * w + 1] - 1;
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  const int root = d.wroot[w];
  if (root < 0) return;  // the world has no other shape
  const int base = sizeof(S) < sizeof(int) ? wlo : 0;
  // a 16-bit entry holds any node of the world's range; a node outside it
  // would mean an inconsistent tree, flagged rather than truncated
  auto fits = [&](int node) { return sizeof(S) >= sizeof(int) || (unsigned)(node - base) <= 0xffffu; };
  if (!fits(root)) {
    atomicOr(&d.cnt[C_STATUS], S_TREE_CORRUPT);
    return;
  }
  int sp = 0;
  stk[tid] = (S)(root - base);  // the world's own subtree: no walk through nodes spanning worlds
  sp = 1;
  int visits = 0;
  // test both children of one node record; leaves are emitted, overlapping
  // inner children pushed
  auto visit = [&](float4 lb, float4 rb, float4 cw) {
    const int split = __float_as_int(cw.z);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      // the child's leaf range must meet the world's: [lo, split] / [split + 1, hi]
      if (side == 0 ? split < wlo : split + 1 > whi) continue;
      if (!overlap(q, side == 0 ? lb : rb)) continue;
      const int code = __float_as_int(side == 0 ? cw.x : cw.y);
      if (code < 0) {
        const int j = ~code;
        if (j == i) continue;
        const bool stat_j = ((__float_as_int(cw.w) >> side) & 1) != 0;
        if (j < i && !stat_j) continue;  // j's own traversal finds it
        if (!pair_allowed(d, fi, d.sfast[j])) continue;
        if (j < i)
          emit(((uint64_t)(uint32_t)j << 32) | (uint32_t)i, j);
        else
          emit(((uint64_t)(uint32_t)i << 32) | (uint32_t)j, -1);
      } else if (!fits(code)) {
        atomicOr(&d.cnt[C_STATUS], S_TREE_CORRUPT);
      } else if (sp < kStack) {
        stk[sp++ * nt + tid] = (S)(code - base);
      } else {
        atomicOr(&d.cnt[C_STATUS], S_STACK_OVERFLOW);
      }
    }
  };
  // up to kTrip nodes per trip: their records are loaded before any is
  // tested, so each lane has several loads in flight on its dependent walk
  constexpr int kTrip = 4;
  while (sp > 0) {
    int nodes[kTrip];
    int take = 0;
#pragma unroll
    for (int u = 0; u < kTrip; ++u) {
      nodes[u] = sp > 0 ? base + (int)stk[--sp * nt + tid] : -1;
      take += nodes[u] >= 0 ? 1 : 0;
    }
    visits += take;
    if (visits > n) {  // a tree has n - 1 internal nodes
      atomicOr(&d.cnt[C_STATUS], S_TREE_CORRUPT);
      break;
    }
    float4 r0[kTrip], r1[kTrip], r2[kTrip];
#pragma unroll
    for (int u = 0; u < kTrip; ++u) {
      const float4* rec = d.node_rec + 3 * (size_t)(nodes[u] >= 0 ? nodes[u] : nodes[0]);
      r0[u] = rec[0];
      r1[u] = rec[1];
      r2[u] = rec[2];
    }
#pragma unroll
    for (int u = 0; u < kTrip; ++u)
      if (nodes[u] >= 0) visit(r0[u], r1[u], r2[u]);
  }
}

// Queries run in Morton order (lane t takes the shape at sorted position t):
// neighbouring lanes query neighbouring boxes and walk the same nodes. A lane
// keeps up to kSlots keys, sorted (keys of static lower partners sort
// first); the counts go on past kSlots, and k_pairs_compact traverses such a
// shape again.
template <int kSlots, class S>
__global__ void __launch_bounds__(kQueryBlock) k_lbvh_pairs(Dev d) {
  __shared__ S stk[kStack * kQueryBlock];
  __shared__ uint64_t slots[kSlots * kQueryBlock];  // thread-minor, like the stack
  B2G_FOR_ALL(t, valid, 0, d.ns) {
    if (valid) {
      // an idle world's shapes keep their segments (k_aabb set their counts)
      if (d.idle && !world_active(d, (int)(d.mkey[0][t] >> 32))) continue;
      const int i = (int)d.mval[0][t];
      const int tid = (int)threadIdx.x;
      int found = 0, own = 0;
      lbvh_query(d, i, stk, [&](uint64_t key, int s) {
        if (found < kSlots) {
          int b = found - 1;  // sorted insert
          while (b >= 0 && slots[b * kQueryBlock + tid] > key) {
            slots[(b + 1) * kQueryBlock + tid] = slots[b * kQueryBlock + tid];
            --b;
          }
          slots[(b + 1) * kQueryBlock + tid] = key;
        }
        ++found;
        if (s >= 0)
          atomicAdd(&d.pcount[s], 1);
        else
          ++own;
      });
      uint64_t* out = d.pslots + (size_t)i * kSlots;
      const int kept = found < kSlots ? found : kSlots;
      for (int k = 0; k < kept; ++k) out[k] = slots[k * kQueryBlock + tid];
      d.pfound[i] = found;  // a static shape: 0, its segment's fill cursor
      if (!(d.sfast[i].y & 64)) d.pcount[i] = own;
    }
  }
}

// Write each shape's pairs to the pair list: its own keys to its segment
// (poff), keys of a static lower shape s to s's segment at a cursor. A shape
// with more pairs than slots is traversed again straight into place.
// Segments filled in traversal or cursor order are queued for sorting.
template <int kSlots, class S>
__global__ void __launch_bounds__(kQueryBlock) k_pairs_compact(Dev d) {
  __shared__ S stk[kStack * kQueryBlock];
  B2G_FOR_ALL(t, valid, 0, d.ns) {
    if (valid) {
      // an idle world's block of pairs is copied whole (k_keep_idle_pairs)
      if (d.idle && !world_active(d, (int)(d.mkey[0][t] >> 32))) continue;
      const int i = (int)d.mval[0][t];
      const int own = d.pcount[i];
      const int base = d.poff[i];
      const bool stat = (d.sfast[i].y & 64) != 0;
      if (base + own > d.ccap) {
        atomicOr(&d.cnt[C_STATUS], S_CONTACT_OVERFLOW);
        continue;
      }
      auto foreign = [&](uint64_t key, int s) {
        const int at = d.poff[s] + atomicAdd(&d.pfound[s], 1);
        if (at < d.ccap) d.nkey[at] = key;  // poff[s] + pcount[s] <= ccap checked by s's lane
      };
      if (stat) {
        if (own > 1) d.heavy[atomicAdd(&d.cnt[C_HEAVY], 1)] = make_int2(base, own);
        continue;
      }
      const int found = d.pfound[i];
      if (found <= kSlots) {
        const uint64_t* src = d.pslots + (size_t)i * kSlots;
        int k = 0;
        for (int m = 0; m < found; ++m) {
          const uint64_t key = src[m];
          const int lo = (int)(key >> 32);
          if (lo == i)
            d.nkey[base + k++] = key;
          else
            foreign(key, lo);
        }
      } else {
        // over-full shape: append in traversal order, sorted afterwards by
        // k_sort_pair_segments (an insertion sort here would be quadratic)
        int k = 0;
        lbvh_query(d, i, stk, [&](uint64_t key, int s) {
          if (s >= 0)
            foreign(key, s);
          else if (k < own)
            d.nkey[base + k++] = key;
        });
        if (own > 1) d.heavy[atomicAdd(&d.cnt[C_HEAVY], 1)] = make_int2(base, own);
      }
    }
  }
}

// An idle world's pairs: its shapes are contiguous, so its segments form one
// block in the current list and one block in the new list (their lengths were
// kept), copied whole by one workgroup.
__global__ void __launch_bounds__(kBlock) k_keep_idle_pairs(Dev d) {
  const int n = d.cnt[C_CONTACTS];
  for (int w = blockIdx.x; w < d.nw; w += gridDim.x) {
    if (world_active(d, w)) continue;
    const int f = d.wshape0[w], l = d.wshape1[w];
    if (l <= f) continue;
    const int src = min(d.poff_prev[f], n);
    const int end = l < d.ns ? min(d.poff_prev[l], n) : n;
    const int dst = d.poff[f];
    const int len = end - src;
    if (dst + len > d.ccap) {
      if (threadIdx.x == 0) atomicOr(&d.cnt[C_STATUS], S_CONTACT_OVERFLOW);
      continue;
    }
    for (int k = (int)threadIdx.x; k < len; k += (int)blockDim.x) d.nkey[dst + k] = d.ckey[src + k];
  }
}

// ---- user queries on the tree of the last broadphase ---------------------
// b2World::QueryAABB (src/dynamics/b2_world.cpp:1193-1198 ->
// b2BroadPhase::QueryAABB, include/box2d/b2_broad_phase.h:622-642): shapes of
// world w whose tight AABB overlaps the box; a world with one shape reports
// it unconditionally (its root is that leaf). Results beyond `cap` are
// counted, not written.
__global__ void __launch_bounds__(kQueryBlock) k_query_aabbs(Dev d, int n, const float4* __restrict__ boxes,
                                                             const int* __restrict__ worlds, int cap,
                                                             int* __restrict__ out, int* __restrict__ counts) {
  __shared__ int stk[kStack * kQueryBlock];
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  B2G_FOR_ALL(qi, valid, 0, n) {
    if (valid) {
      const float4 q = boxes[qi];
      const int w = worlds[qi];
      const int wlo = d.wleaf[2 * w], whi = d.wleaf[2 * w + 1] - 1;
      int* o = out + (size_t)qi * cap;
      int found = 0;
      if (whi < wlo) {
        // no shapes in this world
      } else if (wlo == whi || d.ns < 2) {
        if (cap > 0) o[0] = (int)d.mval[0][wlo];
        found = 1;
      } else {
        int sp = 1;
        stk[tid] = 0;
        int visits = 0;
        while (sp > 0) {
          const int node = stk[--sp * nt + tid];
          if (++visits > d.ns) {
            atomicOr(&d.cnt[C_STATUS], S_TREE_CORRUPT);
            break;
          }
          const float4* rec = d.node_rec + 3 * (size_t)node;
          const float4 lb = rec[0], rb = rec[1], cw = rec[2];
          const int split = __float_as_int(cw.z);
#pragma unroll
          for (int side = 0; side < 2; ++side) {
            if (side == 0 ? split < wlo : split + 1 > whi) continue;
            if (!overlap(q, side == 0 ? lb : rb)) continue;
            const int code = __float_as_int(side == 0 ? cw.x : cw.y);
            if (code < 0) {
              if (found < cap) o[found] = ~code;
              ++found;
            } else if (sp < kStack) {
              stk[sp++ * nt + tid] = code;
            } else {
              atomicOr(&d.cnt[C_STATUS], S_STACK_OVERFLOW);
            }
          }
        }
      }
      counts[qi] = found;
    }
  }
}

// b2World::RayCast (src/dynamics/b2_world.cpp:1228-1238 -> b2BroadPhase::RayCast,
// include/box2d/b2_broad_phase.h:646-716) over the whole segment: every shape
// of world w the segment hits (fraction in [0, 1]), pruned as the reference
// prunes at maxFraction 1 (segment AABB and the separating axis of the
// segment). The caller applies its callback's clipping in fraction order.
// Hits: (shape, fraction, normal.x, normal.y) per slot.
__global__ void __launch_bounds__(kQueryBlock) k_raycast(Dev d, int n, const float4* __restrict__ segs,
                                                         const int* __restrict__ worlds, int cap,
                                                         float4* __restrict__ out, int* __restrict__ counts) {
  __shared__ int stk[kStack * kQueryBlock];
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  B2G_FOR_ALL(qi, valid, 0, n) {
    if (valid) {
      const float4 sg = segs[qi];
      const vec2 p1 = mk(sg.x, sg.y), p2 = mk(sg.z, sg.w);
      const int w = worlds[qi];
      const int wlo = d.wleaf[2 * w], whi = d.wleaf[2 * w + 1] - 1;
      const vec2 r = normalize(p2 - p1);
      const vec2 v = cross(1.0f, r);
      const vec2 abs_v = mk(absf(v.x), absf(v.y));
      const vec2 t = p1 + 1.0f * (p2 - p1);
      const float4 seg_box = make_float4(minf(p1.x, t.x), minf(p1.y, t.y), maxf(p1.x, t.x), maxf(p1.y, t.y));
      float4* o = out + (size_t)qi * cap;
      int found = 0;
      auto test_box = [&](float4 b) {
        if (!overlap(b, seg_box)) return false;
        const vec2 c = mk(0.5f * (b.x + b.z), 0.5f * (b.y + b.w));
        const vec2 h = mk(0.5f * (b.z - b.x), 0.5f * (b.w - b.y));
        const float separation = absf(dot(v, p1 - c)) - dot(abs_v, h);
        return !(separation > 0.0f);
      };
      auto test_leaf = [&](int j) {
        const int4 hj = d.shdr[j];
        Geo g;
        load_geo(g, hj.x, hj.z, d.geo + 9 * (size_t)j);
        float fraction = 0.0f;
        vec2 normal = mk(0.0f, 0.0f);
        if (raycast_shape(g, xform_of(d.xf[hj.y]), p1, p2, 1.0f, fraction, normal)) {
          if (found < cap) o[found] = make_float4(__int_as_float(j), fraction, normal.x, normal.y);
          ++found;
        }
      };
      if (whi < wlo) {
        // no shapes in this world
      } else if (wlo == whi || d.ns < 2) {
        if (test_box(d.sparse[wlo])) test_leaf((int)d.mval[0][wlo]);
      } else {
        int sp = 1;
        stk[tid] = 0;
        int visits = 0;
        while (sp > 0) {
          const int node = stk[--sp * nt + tid];
          if (++visits > d.ns) {
            atomicOr(&d.cnt[C_STATUS], S_TREE_CORRUPT);
            break;
          }
          const float4* rec = d.node_rec + 3 * (size_t)node;
          const float4 lb = rec[0], rb = rec[1], cw = rec[2];
          const int split = __float_as_int(cw.z);
#pragma unroll
          for (int side = 0; side < 2; ++side) {
            if (side == 0 ? split < wlo : split + 1 > whi) continue;
            if (!test_box(side == 0 ? lb : rb)) continue;
            const int code = __float_as_int(side == 0 ? cw.x : cw.y);
            if (code < 0) {
              test_leaf(~code);
            } else if (sp < kStack) {
              stk[sp++ * nt + tid] = code;
            } else {
              atomicOr(&d.cnt[C_STATUS], S_STACK_OVERFLOW);
            }
          }
        }
      }
      counts[qi] = found;
    }
  }
}

// Ascending sort of the pair segments of over-full shapes: one workgroup per
// segment, a bitonic network whose comparators all put the smaller key at the
// lower index (so padding past the end behaves as +infinity). Segments of up
// to kSortLds keys are sorted in LDS, longer ones in place in global memory.
constexpr int kSortLds = 4096;
__device__ __forceinline__ void bitonic_steps(uint64_t* a, int n, int p2) {
  for (int k = 2; k <= p2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < (p2 >> 1); t += blockDim.x) {
        // element pair of this comparator: first step of a merge mirrors
        const int lo = ((t / j) * 2 * j) + (t % j);
        const int hi = (j == (k >> 1)) ? (lo ^ (k - 1)) : (lo + j);
        const int a0 = lo < hi ? lo : hi, a1 = lo < hi ? hi : lo;
        if (a1 < n) {
          const uint64_t x = a[a0], y = a[a1];
          if (x > y) {
            a[a0] = y;
            a[a1] = x;
          }
        }
      }
      __syncthreads();
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_sort_pair_segments(Dev d) {
  __shared__ uint64_t buf[kSortLds];
  const int nseg = d.cnt[C_HEAVY];
  for (int sgi = blockIdx.x; sgi < nseg; sgi += gridDim.x) {
    const int2 sg = d.heavy[sgi];
    uint64_t* a = d.nkey + sg.x;
    const int n = sg.y;
    int p2 = 1;
    while (p2 < n) p2 <<= 1;
    if (n <= kSortLds) {
      for (int t = threadIdx.x; t < n; t += blockDim.x) buf[t] = a[t];
      __syncthreads();
      bitonic_steps(buf, n, p2);
      for (int t = threadIdx.x; t < n; t += blockDim.x) a[t] = buf[t];
      __syncthreads();
    } else {
      bitonic_steps(a, n, p2);
    }
  }
}

// New contact list: persisting pairs keep their contact (orientation,
// manifold, impulses, material, colour); new pairs start empty with the
// factory's orientation and the mixing laws (src/dynamics/b2_contact.cpp:58-124,
// include/box2d/b2_contact.h:42-60). Same-type pairs take A = lower index.
__global__ void __launch_bounds__(kBlock) k_merge_contacts(Dev d) {
  const int old_count = d.cnt[C_CONTACTS];
  const int total = d.cnt[C_PAIRS];
  const int n = total < d.ccap ? total : d.ccap;
  B2G_FOR(p, 0, n) {
    const uint64_t key = d.nkey[p];
    int lo = 0, hi = old_count;
    if (d.idle) {  // the old list is a pair list too: only the lower shape's segment can hold the key
      const int2 seg = prev_segment(d, (int)(key >> 32));
      lo = seg.x;
      hi = seg.y;
    }
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (d.ckey[mid] < key) lo = mid + 1; else hi = mid;
    }
    if (lo < old_count && d.ckey[lo] == key) {
      d.npair[p] = d.cpair[lo];
#pragma unroll
      for (int q = 0; q < 4; ++q) d.nman[4 * (size_t)p + q] = d.cman[4 * (size_t)lo + q];
      d.nmat[p] = d.cmat[lo];
      d.ncol[p] = d.ccol[lo];
    } else {
      const int i = (int)(key >> 32), j = (int)(key & 0xffffffffu);
      const int ti = d.shdr[i].x, tj = d.shdr[j].x;
      const int2 pr = has_collide_function(ti, tj) ? make_int2(i, j) : make_int2(j, i);
      d.npair[p] = pr;
      const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      d.nman[4 * (size_t)p + 0] = z;
      d.nman[4 * (size_t)p + 1] = z;
      d.nman[4 * (size_t)p + 2] = z;
      d.nman[4 * (size_t)p + 3] = z;
      const float4 ma = d.smat[pr.x], mb = d.smat[pr.y];
      d.nmat[p] = make_float4(sqrtf(ma.x * mb.x), ma.y > mb.y ? ma.y : mb.y, ma.z < mb.z ? ma.z : mb.z, 0.0f);
      d.ncol[p] = -1;
    }
  }
}

// EndContact for touching contacts the broadphase no longer reports
// (b2ContactManager::Destroy, src/dynamics/b2_contact_manager.cpp:48-51):
// old contacts whose key is missing from the new, sorted pair list.
// The same contacts wake their bodies when sleeping applies (b2Contact::Destroy,
// src/dynamics/b2_contact.cpp:79-91).
__global__ void __launch_bounds__(kBlock) k_dead_contact_events(Dev d) {
  const int old_count = d.cnt[C_CONTACTS];
  const int total = d.cnt[C_PAIRS];
  const int n = total < d.ccap ? total : d.ccap;
  B2G_FOR(i, 0, old_count) {
    const float4* mp = d.cman + 4 * (size_t)i;
    const float4 m3 = mp[3];
    if (((as_int(m3.z) >> 8) & 0xff) == 0) continue;  // not touching
    const uint64_t key = d.ckey[i];
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (d.nkey[mid] < key) lo = mid + 1; else hi = mid;
    }
    if (lo < n && d.nkey[lo] == key) continue;  // persists
    const int2 pr = d.cpair[i];
    if (d.events_on) {
      const float4 man[4] = {mp[0], mp[1], mp[2], m3};
      record_event(d, pr, kEventDestroyed, man, d.cmat[i]);
    }
    if (d.sleep_on) {
      for (int side = 0; side < 2; ++side) {
        const int b = d.shdr[side ? pr.y : pr.x].y;
        if (d.btype[b] == kStaticBody) continue;
        atomicAnd(&d.bsflag[b], ~kBodyAsleep);  // b2Body::SetAwake(true)
        d.bstime[b] = 0.0f;
      }
    }
  }
}

__global__ void k_finish_pairs(Dev d) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const int total = d.cnt[C_PAIRS];
    if (total > d.ccap) atomicOr(&d.cnt[C_STATUS], S_CONTACT_OVERFLOW);
    d.cnt[C_CONTACTS] = total < d.ccap ? total : d.ccap;
  }
}

// ---- standalone narrowphase over caller pairs (b2g_collide_pairs) -----------
__global__ void __launch_bounds__(kBlock) k_collide_pairs(int n, const int4* __restrict__ ha, const float4* __restrict__ ga,
                                                          const float4* __restrict__ xa, const int4* __restrict__ hb,
                                                          const float4* __restrict__ gb, const float4* __restrict__ xb,
                                                          float4* __restrict__ out /* 4 per pair */, int* __restrict__ ok) {
  B2G_FOR(i, 0, n) {
    Geo a, b;
    load_geo(a, ha[i].x, ha[i].z, ga + 9 * (size_t)i);
    load_geo(b, hb[i].x, hb[i].z, gb + 9 * (size_t)i);
    Man m;
    m.count = 0;
    m.type = 0;
    m.ln = m.lp = m.lp0 = m.lp1 = mk(0.0f, 0.0f);
    m.id0 = m.id1 = 0;
    const bool has = evaluate(m, a, xform_of(xa[i]), b, xform_of(xb[i]));
    ok[i] = has ? 1 : 0;
    out[4 * (size_t)i + 0] = make_float4(m.lp0.x, m.lp0.y, 0.0f, 0.0f);
    out[4 * (size_t)i + 1] = make_float4(m.count > 1 ? m.lp1.x : 0.0f, m.count > 1 ? m.lp1.y : 0.0f, 0.0f, 0.0f);
    out[4 * (size_t)i + 2] = make_float4(m.ln.x, m.ln.y, m.lp.x, m.lp.y);
    out[4 * (size_t)i + 3] = make_float4(uint_as_float(m.count > 0 ? m.id0 : 0u), uint_as_float(m.count > 1 ? m.id1 : 0u),
                                         as_float((m.type & 0xff) | (m.count << 8)), 0.0f);
  }
}

}  // namespace b2g
