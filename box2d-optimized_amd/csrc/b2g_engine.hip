// b2g_engine.hip -- device arenas, step orchestration and the b2g C ABI.
//
// A b2g_batch holds any number of independent worlds in shared
// structure-of-arrays device buffers (bodies, shapes, joints, contacts); a
// single world is a batch of one. One b2g_batch_step runs every stage of
// b2World::Step for all active worlds with one launch per stage (one per
// colour for the solver sweeps), so thousands of small worlds cost the same
// number of launches as one large world.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "b2g.h"
#include "b2g_kernels.h"
#include "b2g_grid.h"
#include "b2g_partition.h"

using namespace b2g;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return fail(B2G_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));          \
  } while (0)

template <typename T>
int dalloc(T** p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) return fail(B2G_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
  e = hipMemset(*p, 0, n * sizeof(T));
  if (e != hipSuccess) return fail(B2G_ERR_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
  return B2G_OK;
}

constexpr int kColourRoundsPerStep = 4;
// one world: two device-wide rounds, then the one-workgroup finishing pass
// (k_colour_finish), whose rounds cost a barrier each instead of two launches
constexpr int kColourRoundsOneWorld = 2;
constexpr int kFinishMaxLeft = 4096;  // more left (a world colouring afresh): device-wide rounds instead
// claim rounds of a captured step of several worlds: the colouring is
// incremental (a contact keeps its colour while it persists), so a replay
// colours only the new contacts in these multi-workgroup rounds (a round with
// nothing pending returns at once, so spare rounds cost only their launch); a
// burst of new contacts (or a world coloured afresh) is finished by up to
// kGraphFinishRounds rounds in one workgroup, and anything still uncoloured
// joins the serial group
constexpr int kGraphColourRounds = 4;
constexpr int kGraphFinishRounds = 64;

inline int grid_for(long long n, int max_blocks = 4096) {
  long long g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  // multiples of 8 let the kernels map consecutive tiles to one XCD
  if (g >= 16) g = (g + 7) & ~7LL;
  return (int)g;
}

// Launch shape of a grid-stride kernel (no LDS, no block-wide step) over n
// items. Up to kSmallLaunchItems items the launch is latency-bound: 64-lane
// workgroups spread it over as many CUs as possible, one wave each. Larger
// launches take kBlock lanes (64-lane workgroups, capped at 4096, would loop).
constexpr long long kSmallLaunchItems = 65536;
struct LaunchShape {
  dim3 grid, block;
};
inline LaunchShape shape_for(long long n, int max_blocks = 4096) {
  if (n <= kSmallLaunchItems) return {dim3(grid_for(n * kBlock / kWave, max_blocks)), dim3(kWave)};
  return {dim3(grid_for(n, max_blocks)), dim3(kBlock)};
}

}  // namespace

struct b2g_batch {
  b2g_capacity cap{};
  Dev d{};
  int dev_id = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[7] = {};
  hipStream_t side = nullptr;  // the island pass, beside the velocity phase
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  hipEvent_t counts_ev = nullptr;  // the colour path's early counter read-back
  int* pinned = nullptr;     // host mirror of counters + colour offsets
  // pinned staging of b2g_set_world_params (copied in stream order, no host
  // wait; params_ev guards the staging against refilling before its copy ran)
  float4* stage_wpar = nullptr;
  int* stage_wflag = nullptr;
  hipEvent_t params_ev = nullptr;
  // the active worlds of a mostly idle step, for the per-world solver's
  // launches (order_ev guards the staging as params_ev does)
  std::vector<int> h_worder;
  int* stage_order = nullptr;
  hipEvent_t order_ev = nullptr;
  // host mirrors
  std::vector<int> h_btype, h_bworld;
  std::vector<int> h_shape_body;
  std::vector<b2g_shape> h_shape_rec;  // uploaded shape records (snapshots)
  std::vector<int> h_wshape0;   // first shape of each world
  std::vector<int> h_wshape1;   // one past its last shape (world layout)
  std::vector<b2g_revolute> h_joints;
  std::vector<int> h_jcol_off;  // joint colour ranges (host colouring)
  std::vector<b2g_world_params> h_params;
  bool joints_dirty = false;
  bool filters_dirty = true;
  int est_contacts = 0, est_cons = 0;
  unsigned round = 0;
  b2g_profile prof{};
  // last counters
  int last_ncolours = 0, last_islands = 0, last_pos_iters = 0;
  // HIP-event timing of one selected kernel (b2g_time_kernel)
  int timed_kernel = -1;
  std::vector<hipEvent_t> tev;
  size_t tev_used = 0;
  std::vector<long long> t_items;
  std::vector<int> t_runs;  // launches per timed run
  double t_ms = 0.0;
  long long t_launches = 0, t_total_items = 0;
  static constexpr int kUntimed = -2;  // a launch that no kernel timing selects
  // Kernel timing: an event pair around a run of consecutive launches of the
  // timed class (the per-colour sweeps of one iteration, or a single launch),
  // divided over the run's launches. Back-to-back launches on the stream tile
  // the device timeline, so this is each launch's duration plus the pair's
  // own overhead spread over the run.
  void tbegin(int id) {
    if (id != timed_kernel) return;
    if (tev_used + 2 > tev.size()) {
      // timing only: no system-scope fence (a cache writeback and
      // invalidate at each record would be timed with the launches)
      hipEvent_t e0, e1;
      (void)hipEventCreateWithFlags(&e0, hipEventDisableSystemFence);
      (void)hipEventCreateWithFlags(&e1, hipEventDisableSystemFence);
      tev.push_back(e0);
      tev.push_back(e1);
    }
    (void)hipEventRecord(tev[tev_used], stream);
  }
  void tend(int id, long long items, int launches = 1) {
    if (id != timed_kernel) return;
    (void)hipEventRecord(tev[tev_used + 1], stream);
    tev_used += 2;
    t_items.push_back(items);
    t_runs.push_back(launches);
  }
  template <class K, class... Args>
  void launch(int id, long long items, K kernel, dim3 grid, dim3 block, uint32_t shmem, Args... args) {
    tbegin(id);
    hipLaunchKernelGGL(kernel, grid, block, shmem, stream, args...);
    tend(id, items);
  }
  void tcollect() {
    if (tev_used == 0) return;
    (void)hipStreamSynchronize(stream);
    for (size_t k = 0; k + 1 < tev_used; k += 2) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, tev[k], tev[k + 1]) == hipSuccess) t_ms += ms;
      t_launches += t_runs[k / 2];
      t_total_items += t_items[k / 2];
    }
    tev_used = 0;
    t_items.clear();
    t_runs.clear();
  }

  // captured steps (graph mode)
  struct GraphEntry {
    hipGraphExec_t exec;
    int parity, vi, pi, K;
    unsigned flags;
    int block;  // per-world workgroup size (0: colour-major sweeps)
  };
  std::vector<GraphEntry> graphs;
  int parity = 0;            // which contact buffer is current (flips with every broadphase)
  int graph_K = 0;           // colour slots per sweep of captured steps (0: no eager step yet)
  int rounds_pending = 0;    // claim rounds enqueued since the last round-base advance
  unsigned round_host = 0;   // host mirror of the device claim-round base
  long long graph_steps = 0;
  // per-world solver (one workgroup per world)
  bool layout_dirty = true;  // world ranges need recomputing
  bool per_world_ok = false; // joint-free, bodies and shapes laid out world after world
  bool worlds_contiguous = false;  // bodies and shapes laid out world after world
  int graph_tail_from = 0;         // first colour of the shared tail launch (from the last eager step)
  int graph_big_items = 0;         // largest colour with a launch of its own (from the last eager step)
  int max_world_bodies = 0;
  int max_world_shapes = 0;
  int big_worlds = 0;              // worlds of more than kBigWorldBodies bodies (first in worder)
  int max_small_world_bodies = 0;  // largest of the others
  int solver_mode = 0;       // B2G_SOLVER_AUTO / _COLOURS / _WORLDS
  bool keep_idle = true;     // B2G_OPT_KEEP_IDLE
  int pair_search = 0;       // B2G_OPT_PAIR_SEARCH
  int last_pair_search = 0;  // 1 LBVH, 2 grid
  int world_static_max = 0;  // most static shapes in one world
  bool grid_local_ok = false;  // every world fits a bucket table of its own (k_grid_bin_worlds)
  bool use_grid() const { return pair_search == 0 && world_static_max <= kGridStaticMax && d.ns >= 2; }
  bool asleep_uploaded = false;  // a body was uploaded asleep: the awake flags apply from then on
  bool sleep_now() const {       // this step applies the awake flags and sleep timers
    if (asleep_uploaded) return true;
    for (int w = 0; w < d.nw; ++w)
      if (h_params[w].active && h_params[w].allow_sleep) return true;
    return false;
  }
  int world_block = 0;       // threads per world workgroup (0: autotune)
  int tuned_block = 0;       // chosen by the autotuner (or world_block)
  bool tuned_lds = false;    // autotuner: stage world bodies in LDS
  int world_lds = -1;        // -1 autotune, 0 off, 1 on (when the worlds fit)
  int tune_step = 0;         // autotuner progress: two timed steps per candidate
  float tune_ms[16] = {};
  int last_path = 0;         // 1 colour-major sweeps, 2 per-world, 3 cells

  int alloc();
  void release();
  void swap_contact_buffers();
  void graphs_clear();
  void enqueue_collide_and_colour(int K, int rounds, bool collect_last = false);
  void enqueue_collide_and_colour_eager();
  // device-wide claim rounds of a captured step before the finishing
  // workgroup: as an eager step's (a single world's few new contacts finish
  // in the workgroup; a batch's worlds colouring afresh need the whole GPU)
  int graph_colour_rounds() const { return d.nw == 1 ? kColourRoundsOneWorld : kGraphColourRounds; }
  void enqueue_collide();
  bool use_cells() const;
  void enqueue_cells_solve(int vi, int pi, bool finish_in_broadphase);
  int cells_step(int vi, int pi, unsigned flags);
  int sleep_step(int vi, int pi, unsigned flags);

  void enqueue_colour_round(int K, bool collect = false);
  void enqueue_colour_groups();
  void enqueue_islands_and_init();
  void enqueue_islands_join();
  void enqueue_solve(int vi, int pi, const std::vector<int>& colours, const std::vector<int>& sizes,
                     const std::vector<int>& begins, int graph_grid, bool finish_in_broadphase);
  bool finish_pending = false;  // the step's body finish is left to the broadphase's AABB launch
  void flush_finish();
  int enqueue_graph_body(int vi, int pi, unsigned flags, int K);
  int check_status();
  int status_error() const;
  int collide_only_step(unsigned flags);
  hipEvent_t graph_ev = nullptr;  // the last graph replay (its counter copy) has finished
  bool graph_ev_live = false;
  int update_world_layout();
  int update_first_shapes();
  bool use_per_world() const;
  int world_block_now() const;
  bool world_lds_now() const;
  void launch_solve_worlds(int vi, int pi);
  int eager_step(int vi, int pi, unsigned flags);
  int graph_step(int vi, int pi, unsigned flags);
  int scan(const int* in, const int* n_dev, int n_host, int* out, int* total, long long n_bound);
  int radix_sort(int n, int bits, uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1);
  // worlds that are not active in the coming step (host view of the flags)
  int idle_worlds() const {
    int n = 0;
    for (int w = 0; w < d.nw; ++w) n += h_params[w].active ? 0 : 1;
    return n;
  }
  int build_tree(bool use_given_aabbs);
  int broadphase(bool use_given_aabbs);
  int grid_pairs();
  int finish_pair_list(bool use_given_aabbs);
  void refresh_filters();
  bool tree_stale = true;     // bodies or shapes changed since the last tree build
  // the contact list is the last broadphase's pair list (segments in d.poff_prev), and no body,
  // shape, joint or contact changed since: worlds that are not active may keep their pairs
  bool pairs_current = false;
  int refits_left = 0;        // steps that may refit the last tree before it is rebuilt
  int static_shapes = 0;      // shapes of static bodies (their pair segments are sorted every broadphase)
  int tree_ns = -1;           // shape count of the last full build
  bool capturing = false;     // a step graph is being captured: full builds only
  void* qbuf = nullptr;       // scratch of the user queries
  size_t qbuf_bytes = 0;
  int query_scratch(size_t bytes);
  int rebuild_joint_tables();
  int step(int vi, int pi, unsigned flags);
  int read_counters();
};

int b2g_batch::alloc() {
  const int NB = std::max(cap.bodies, 1), NS = std::max(cap.shapes, 1), NJ = std::max(cap.joints, 1);
  const int NW = std::max(cap.worlds, 1), NC = std::max(cap.contacts, 1);
  int rc;
#define A(ptr, n) \
  if ((rc = dalloc(&(ptr), (size_t)(n))) != B2G_OK) return rc
  A(d.xf, NB); A(d.pos, NB); A(d.vel, NB); A(d.prop, NB); A(d.prop2, NB); A(d.force, NB);
  A(d.btype, NB); A(d.bworld, NB);
  A(d.wpar, NW); A(d.wflag, NW); A(d.wsort, NW); A(d.wshape0, NW); A(d.wshape1, NW); A(d.wbody, NW); A(d.worder, NW); A(d.wact, NW);
  A(d.wcol, (size_t)NW * (kColourSlots + 1));
  A(d.wcolmax, NW); A(d.wfresh, NW); A(d.wrecolour, NW); A(d.wbounds, 4 * NW); A(d.wleaf, 2 * NW); A(d.wroot, NW);
  A(d.shdr, NS); A(d.geo, 9 * (size_t)NS); A(d.smat, NS); A(d.sfilt, NS); A(d.aabb, NS);
  A(d.joints, NJ); A(d.jsol, NJ); A(d.jorder, NJ); A(d.jisland, NJ); A(d.wj_off, NW + 1); A(d.wj_list, NJ);
  A(d.bj_off, NB + 1); A(d.bj_other, 2 * NJ);
  A(d.ckey, NC); A(d.cpair, NC); A(d.cman, 4 * (size_t)NC); A(d.cmat, NC); A(d.ccol, NC);
  A(d.nkey, NC); A(d.npair, NC); A(d.nman, 4 * (size_t)NC); A(d.nmat, NC); A(d.ncol, NC);
  A(d.touch, NC); A(d.tpos, NC); A(d.clist, NC); A(d.cbody, NC); A(d.ccolour, NC); A(d.csorted, NC);
  A(d.cisland, NC); A(d.cpend, NC); A(d.cleft, NC); A(d.events, (size_t)kEventWords * NC);
  for (int k = 0; k < 10; ++k) A(d.vc[k], NC);
  for (int k = 0; k < 5; ++k) A(d.pc[k], NC);
  A(d.col_count, kColourSlots + 1); A(d.col_cursor, kColourSlots + 1);
  A(d.bmask, NB); A(d.bslot, NB);
  A(d.btab, (size_t)kBodyTab * NB); A(d.parent, NB); A(d.sepbad, 3 * (size_t)NB); A(d.jok, 3 * (size_t)NB); A(d.solved, NB);
  A(d.mkey[0], NS); A(d.mkey[1], NS); A(d.skey, NS); A(d.xkey, NS); A(d.xval, NS); A(d.mval[0], NS); A(d.mval[1], NS);
  const size_t tiles = (size_t)(NS + kSortTile - 1) / kSortTile;
  A(d.sort_hist, kRadix * tiles);
  A(d.node_child, 2 * (size_t)NS); A(d.node_parent, NS); A(d.leaf_parent, NS); A(d.node_range, NS);
  A(d.node_rec, 3 * (size_t)NS); A(d.sfast, NS);
  {
    int levels = 1;
    while ((1 << levels) <= NS) ++levels;  // levels 0 .. floor(log2(NS))
    d.sparse_levels = levels;
    A(d.sparse, (size_t)levels * NS);
  }
  A(d.pslots, (size_t)kPairSlots * NS); A(d.heavy, NS);
  A(d.pcount, NS); A(d.pfound, NS); A(d.sinv, NS); A(d.poff, NS); A(d.poff_prev, NS);
  const size_t scan_n = std::max<size_t>(std::max<size_t>(NC, NS), kRadix * tiles);
  A(d.scan_tiles, scan_n / kScanTile + 2);
  // the counters and, behind them, the colour offsets: the step's one
  // read-back is a single copy
  A(d.cnt, kCntColOff + kColourSlots + 1);
  d.col_off = d.cnt + kCntColOff;
  // the cells solver
  A(d.bsflag, NB); A(d.bstime, NB); A(d.bwake, NB); A(d.bactive, NB); A(d.bncontact, NB); A(d.bjointed, NB);
  A(d.isl_awake, NB); A(d.isl_min, NB); A(d.isl_cons, NB);
  A(d.part_geo, 1); A(d.part_boff, 2 * kCells + 1); A(d.part_bcur, 2 * kCells); A(d.part_acc, 8); A(d.part_sync, 4);
  {  // the grid pair search: about four buckets per shape at most
    int t = 1024;
    while (t < 4 * NS) t <<= 1;
    d.grid_T = t;
    A(d.grid_acc, 1); A(d.gstatic, NS); A(d.gsoff, NW + 1); A(d.gcount, t + 1); A(d.goff, t + 1);
    A(d.gids, NS); A(d.gbox, NS); A(d.gfast, NS); A(d.gworld, NS);
    // the worlds' own tables: T + 1 words each, T a power of two of at least
    // 64 and below 4 x the world's shapes (or 64)
    A(d.gltab, NW + 1); A(d.gloff, 65 * (size_t)NW + 4 * (size_t)NS);
  } A(d.part_blist, 2 * (size_t)NB); A(d.part_brank, NB);
  A(d.part_ccnt, kPartGroups); A(d.part_bucket, (size_t)2 * kCells * kCellCons); A(d.part_cbucket, NC);
  A(d.part_coff, (size_t)kPartGroups * (kColourSlots + 1)); A(d.part_claim, NB); A(d.part_used, NB);
  A(d.part_xcol, NC); A(d.sloc, NC);
#undef A
  d.ccap = NC;
  d.ecap = NC;
  HIP_TRY(hipHostMalloc((void**)&pinned, 256 * sizeof(int), hipHostMallocDefault));
  HIP_TRY(hipHostMalloc((void**)&stage_wpar, (size_t)NW * sizeof(float4), hipHostMallocDefault));
  HIP_TRY(hipHostMalloc((void**)&stage_wflag, (size_t)NW * sizeof(int), hipHostMallocDefault));
  HIP_TRY(hipHostMalloc((void**)&stage_order, (size_t)NW * sizeof(int), hipHostMallocDefault));
  HIP_TRY(hipEventCreateWithFlags(&params_ev, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&order_ev, hipEventDisableTiming));
  HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  {
    // The side stream runs beside the main one (island labels beside the
    // contact set-up, the small worlds' solver beside the big worlds'). A
    // process gets a few hardware queues that its streams share, and two
    // streams on one queue run one after the other; a stream created with a
    // CU mask gets a queue of its own, and this mask names every CU.
    int dev = 0, cus = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    std::vector<uint32_t> all_cus((size_t)std::max(1, (cus + 31) / 32), 0xffffffffu);
    if (hipExtStreamCreateWithCUMask(&side, (uint32_t)all_cus.size(), all_cus.data()) != hipSuccess) {
      // no queue of its own to be had (many live batches): a shared one
      (void)hipGetLastError();
      side = nullptr;
      HIP_TRY(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    }
  }
  // fork and join order the side stream's launches against the main
  // stream's on this device only: no system-scope fence (by default an event
  // record writes back and invalidates the caches for the host's sake)
  HIP_TRY(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming | hipEventDisableSystemFence));
  HIP_TRY(hipEventCreateWithFlags(&join_ev, hipEventDisableTiming | hipEventDisableSystemFence));
  HIP_TRY(hipEventCreateWithFlags(&counts_ev, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&graph_ev, hipEventDisableTiming));
  for (auto& e : ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));  // timing only
  h_btype.assign(NB, 0);
  h_bworld.assign(NB, 0);
  h_shape_body.assign(NS, 0);
  h_wshape0.assign(NW, 0);
  HIP_TRY(hipMemcpy(d.wshape0, h_wshape0.data(), NW * sizeof(int), hipMemcpyHostToDevice));
  {  // the cells solver's bounds accumulator starts empty
    const unsigned acc[8] = {0xffffffffu, 0xffffffffu, 0u, 0u, 0u, 0u, 0u, 0u};
    HIP_TRY(hipMemcpy(d.part_acc, acc, sizeof(acc), hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(d.part_sync, 0, 4 * sizeof(unsigned)));  // each launch leaves its counter at 0
  }
  {  // colouring upkeep: nothing used yet, every world colours from scratch first
    std::vector<int> minus1(NW, -1), one(NW, 1);
    HIP_TRY(hipMemcpy(d.wcolmax, minus1.data(), NW * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.wrecolour, one.data(), NW * sizeof(int), hipMemcpyHostToDevice));
  }
  h_params.assign(NW, b2g_world_params{{0.0f, -10.0f}, 1.0f / 60.0f, 0.0f, 1, 1});
  std::vector<int> zero_off(std::max(NB, NW) + 1, 0);
  HIP_TRY(hipMemcpy(d.bj_off, zero_off.data(), (NB + 1) * sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d.wj_off, zero_off.data(), (NW + 1) * sizeof(int), hipMemcpyHostToDevice));
  return B2G_OK;
}

void b2g_batch::release() {
  graphs_clear();
  if (qbuf) (void)hipFree(qbuf);
  qbuf = nullptr;
  void* ptrs[] = {d.xf, d.pos, d.vel, d.prop, d.prop2, d.force, d.btype, d.bworld, d.wpar, d.wflag, d.wshape0, d.wshape1, d.wbody, d.worder, d.wact, d.wcol, d.wcolmax, d.wfresh,
                  d.wrecolour,
                  d.wbounds, d.wleaf, d.wroot, d.shdr, d.geo, d.smat, d.sfilt, d.aabb, d.joints, d.jsol,
                  d.jorder, d.wj_off, d.wj_list, d.jisland, d.bj_off, d.bj_other, d.ckey, d.cpair, d.cman, d.cmat, d.ccol,
                  d.nkey, d.npair, d.nman, d.nmat, d.ncol, d.touch, d.tpos, d.clist, d.cbody, d.ccolour,
                  d.csorted, d.cisland, d.cpend, d.cleft, d.events, d.col_count, d.col_cursor, d.bmask, d.bslot,
                  d.btab, d.parent, d.sepbad, d.jok, d.solved, d.mkey[0], d.mkey[1], d.skey, d.xkey, d.xval, d.wsort, d.mval[0], d.mval[1],
                  d.sort_hist, d.node_child, d.node_parent, d.leaf_parent, d.node_range, d.node_rec,
                  d.gltab, d.gloff, d.sfast, d.sparse, d.pslots, d.heavy, d.pcount, d.pfound, d.sinv, d.poff, d.poff_prev, d.scan_tiles,
                  d.cnt,
                  // sleep
                  d.bsflag, d.bstime, d.bwake, d.bactive, d.bncontact, d.bjointed, d.isl_awake, d.isl_min, d.isl_cons,
                  // the cells solver
                  d.part_geo, d.part_boff, d.part_bcur, d.part_acc, d.part_sync, d.part_blist, d.part_brank, d.part_ccnt,
                  d.part_bucket, d.part_cbucket, d.part_coff, d.part_claim, d.part_used, d.part_xcol, d.sloc,
                  // the grid pair search
                  d.grid_acc, d.gstatic, d.gsoff, d.gcount, d.goff, d.gids, d.gbox, d.gfast, d.gworld};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (int k = 0; k < 10; ++k)
    if (d.vc[k]) (void)hipFree(d.vc[k]);
  for (int k = 0; k < 5; ++k)
    if (d.pc[k]) (void)hipFree(d.pc[k]);
  if (pinned) (void)hipHostFree(pinned);
  if (stage_wpar) (void)hipHostFree(stage_wpar);
  if (stage_wflag) (void)hipHostFree(stage_wflag);
  if (stage_order) (void)hipHostFree(stage_order);
  if (params_ev) (void)hipEventDestroy(params_ev);
  if (order_ev) (void)hipEventDestroy(order_ev);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (fork_ev) (void)hipEventDestroy(fork_ev);
  if (join_ev) (void)hipEventDestroy(join_ev);
  if (counts_ev) (void)hipEventDestroy(counts_ev);
  if (graph_ev) (void)hipEventDestroy(graph_ev);
  if (side) (void)hipStreamDestroy(side);
  if (stream) (void)hipStreamDestroy(stream);
}

int b2g_batch::scan(const int* in, const int* n_dev, int n_host, int* out, int* total, long long n_bound) {
  if (n_bound <= 8192) {  // one launch; still correct if the device count turns out larger
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kScan1Threads), 0, stream, in, n_dev, n_host, out, total);
    return B2G_OK;
  }
  const int tiles = (int)std::max<long long>(1, (n_bound + kScanTile - 1) / kScanTile);
  const int g = std::min(tiles, 4096);
  hipLaunchKernelGGL(k_scan_tiles, dim3(g), dim3(kBlock), 0, stream, in, n_dev, n_host, d.scan_tiles);
  if (tiles <= kScanFoldTiles) {  // the tile offsets are folded into the apply pass
    hipLaunchKernelGGL(k_scan_apply_fold, dim3(g), dim3(kBlock), 0, stream, in, n_dev, n_host, d.scan_tiles, out,
                       total);
    return B2G_OK;
  }
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kBlock), 0, stream, d.scan_tiles, n_dev, n_host, total);
  hipLaunchKernelGGL(k_scan_apply, dim3(g), dim3(kBlock), 0, stream, in, n_dev, n_host, d.scan_tiles, out);
  return B2G_OK;
}

// LSD radix sort of n (key, value) pairs starting in (k0, v0), ping-ponging
// with (k1, v1)
int b2g_batch::radix_sort(int n, int bits, uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1) {
  uint64_t* keys[2] = {k0, k1};
  uint32_t* vals[2] = {v0, v1};
  const int tiles = (n + kSortTile - 1) / kSortTile;
  const int g = std::min(tiles, 4096);
  int src = 0;
  for (int shift = 0; shift < bits; shift += kRadixBits) {
    hipLaunchKernelGGL(k_radix_hist, dim3(g), dim3(kBlock), 0, stream, keys[src], n, shift, d.sort_hist);
    scan(d.sort_hist, nullptr, kRadix * tiles, d.sort_hist, nullptr, (long long)kRadix * tiles);
    hipLaunchKernelGGL(k_radix_scatter, dim3(g), dim3(kBlock), 0, stream, keys[src], vals[src], n, shift,
                       d.sort_hist, keys[1 - src], vals[1 - src]);
    src = 1 - src;
  }
  return src;  // buffer holding the result (0 for an even number of passes)
}

// The tree over the current AABBs (refreshed from the transforms unless the
// caller gave AABBs): world bounds, Morton keys, radix sort, Karras build
// (skipped by a refit), sparse table and node records.
constexpr int kRefitSteps = 3;

// the step's body finish, when no AABB launch is there to carry it
void b2g_batch::flush_finish() {
  if (!finish_pending) return;
  hipLaunchKernelGGL(k_finish_bodies, dim3(grid_for(d.nb)), dim3(kBlock), 0, stream, d);
  finish_pending = false;
}

int b2g_batch::build_tree(bool use_given_aabbs) {
  const int ns = d.ns;
  if (ns == 0 || use_given_aabbs) flush_finish();
  if (ns == 0) return B2G_OK;
  const int gs = grid_for(ns);
  // Refit: the last build's leaf order and topology with new boxes. Any tree
  // over the shapes answers the pair queries exactly; a refit one only gets
  // looser as bodies move, so it is rebuilt every kRefitSteps + 1 steps and
  // after any upload.
  const bool refit = !use_given_aabbs && !tree_stale && !capturing && refits_left > 0 && tree_ns == ns;
  if (refit) {
    --refits_left;
  } else {
    refits_left = use_given_aabbs ? 0 : kRefitSteps;
    tree_ns = use_given_aabbs ? -1 : ns;
  }
  if (!refit) hipLaunchKernelGGL(k_world_bounds_reset, dim3(grid_for(d.nw)), dim3(kBlock), 0, stream, d);
  if (!use_given_aabbs) {
    hipLaunchKernelGGL(k_aabb, dim3(std::max(gs, finish_pending ? grid_for(d.nb) : 1)), dim3(kBlock), 0, stream, d,
                       refit ? 0 : 1, finish_pending ? 1 : 0);
    finish_pending = false;
  } else {
    // only the world bounds of the caller's AABBs
    hipLaunchKernelGGL(k_world_bounds_from_aabbs, dim3(gs), dim3(kBlock), 0, stream, d);
  }
  if (!refit) {
    // With idle worlds kept (their leaf order stands), only the active
    // worlds' keys are sorted, then placed back at their worlds' leaf ranges.
    long long m = ns;
    if (d.idle && worlds_contiguous && idle_worlds() > 0) {
      m = 0;
      for (int w = 0; w < d.nw && w < (int)h_wshape1.size(); ++w)
        if (h_params[w].active) m += h_wshape1[w] - h_wshape0[w];
    }
    const bool active_only = m < ns;
    if (active_only) {
      hipLaunchKernelGGL(k_active_world_sizes, dim3(grid_for(d.nw)), dim3(kBlock), 0, stream, d);
      scan(d.wsort, nullptr, d.nw, d.wsort, nullptr, d.nw);
    }
    hipLaunchKernelGGL(k_morton, dim3(gs), dim3(kBlock), 0, stream, d, active_only ? 1 : 0);
    int bits = 32;
    int wbits = 0;
    while ((1 << wbits) < d.nw) ++wbits;
    bits += wbits;
    bits = ((bits + 2 * kRadixBits - 1) / (2 * kRadixBits)) * (2 * kRadixBits);  // even pass count
    if (active_only) {
      if (m > 0) {
        const int res = radix_sort((int)m, bits, d.mkey[1], d.mval[1], d.xkey, d.xval);
        if (res != 0) return fail(B2G_ERR_INVALID, "radix sort ended in the wrong buffer");
        hipLaunchKernelGGL(k_place_active_leaves, dim3(grid_for(m)), dim3(kBlock), 0, stream, d, (int)m);
      }
    } else {
      const int res = radix_sort(ns, bits, d.mkey[0], d.mval[0], d.mkey[1], d.mval[1]);
      if (res != 0) return fail(B2G_ERR_INVALID, "radix sort ended in the wrong buffer");
    }
    hipLaunchKernelGGL(k_world_leaves, dim3(gs), dim3(kBlock), 0, stream, d, d.mkey[0]);
    if (ns >= 2) hipLaunchKernelGGL(k_lbvh_build, dim3(grid_for(ns - 1)), dim3(kBlock), 0, stream, d, d.mkey[0]);
  }
  if (!refit) hipLaunchKernelGGL(k_leaf_aabbs, dim3(gs), dim3(kBlock), 0, stream, d, d.mval[0]);
  if (ns >= 2) {
    // range unions are only looked up inside one world (a range spanning
    // worlds gets an unbounded box first), so the table needs the levels of
    // the largest world only
    const int span = (!layout_dirty && max_world_shapes > 0) ? std::min(ns, std::max(max_world_shapes, 2)) : ns;
    int top = 0;
    while ((2 << top) <= span) ++top;  // floor(log2(span))
    Dev dj = d;
    dj.sparse_levels = top + 1;  // levels built this time
    for (int j = 1; j <= top; j += kSparseLevelsPerLaunch)
      hipLaunchKernelGGL(k_sparse_levels, dim3(gs), dim3(kBlock), 0, stream, dj, j);
    hipLaunchKernelGGL(k_node_records, dim3(grid_for(ns - 1)), dim3(kBlock), 0, stream, dj, d.mkey[0], d.mval[0]);
  }
  tree_stale = false;
  return B2G_OK;
}

// bodies, shapes or joints changed: packed pair-filter words, and the count
// of static shapes (which sizes the pair-segment sort)
void b2g_batch::refresh_filters() {
  hipLaunchKernelGGL(k_shape_filters, dim3(grid_for(d.ns)), dim3(kBlock), 0, stream, d);
  filters_dirty = false;
  static_shapes = 0;
  for (int i = 0; i < d.ns; ++i) static_shapes += h_btype[h_shape_body[i]] == B2G_STATIC ? 1 : 0;
  if (d.ns > 0) {  // the grid's cell size and static lists follow the shapes and body types
    hipLaunchKernelGGL(k_grid_reset, dim3(1), dim3(1), 0, stream, d);
    hipLaunchKernelGGL(k_grid_prepare, dim3(std::min(grid_for(d.ns), 64)), dim3(kBlock), 0, stream, d);
    const int nw = std::max(d.nw, 1);
    auto world_of = [&](int b) { return std::min(std::max(h_bworld[b], 0), nw - 1); };
    std::vector<int> off(nw + 1, 0);
    for (int i = 0; i < d.ns; ++i)
      if (h_btype[h_shape_body[i]] == B2G_STATIC) off[world_of(h_shape_body[i]) + 1]++;
    world_static_max = 0;
    for (int w = 0; w < nw; ++w) {
      world_static_max = std::max(world_static_max, off[w + 1]);
      off[w + 1] += off[w];
    }
    std::vector<int> list(std::max(off[nw], 1), 0), cur(off.begin(), off.end() - 1);
    for (int i = 0; i < d.ns; ++i)  // world after world, shape order inside each
      if (h_btype[h_shape_body[i]] == B2G_STATIC) list[cur[world_of(h_shape_body[i])]++] = i;
    (void)hipMemcpyAsync(d.gsoff, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice, stream);
    if (off[nw] > 0)
      (void)hipMemcpyAsync(d.gstatic, list.data(), off[nw] * sizeof(int), hipMemcpyHostToDevice, stream);
    // a bucket table per world (about two buckets per binned shape: four
    // measured slower on the 4096-world batch, whose worlds' tables are
    // cleared and scanned in LDS) when the worlds are laid out one after the
    // other and none bins more than kGridLocalShapes shapes
    grid_local_ok = false;
    std::vector<int> tab(nw + 1, 0);
    if (d.nw > 1 && worlds_contiguous && (int)h_wshape1.size() >= d.nw && (int)h_wshape0.size() >= d.nw) {
      grid_local_ok = true;
      for (int w = 0; w < d.nw && grid_local_ok; ++w) {
        const int n = (h_wshape1[w] - h_wshape0[w]) - (off[w + 1] - off[w]);
        int t = 64;
        while (t < 2 * n) t <<= 1;
        tab[w + 1] = tab[w] + t + 1;
        grid_local_ok = n <= kGridLocalShapes;
      }
      if (grid_local_ok)
        (void)hipMemcpyAsync(d.gltab, tab.data(), (d.nw + 1) * sizeof(int), hipMemcpyHostToDevice, stream);
    }
    (void)hipStreamSynchronize(stream);  // the host vectors end here
  }
}

// The broadphase's pair search on the grid (b2g_grid.h): AABBs, buckets,
// pairs into the slots, and the compacted pair list.
int b2g_batch::grid_pairs() {
  const int ns = d.ns;
  const int gs = grid_for(ns);
  const int gq = grid_for((long long)ns * kBlock / kQueryBlock);  // blocks of kQueryBlock threads
  // one-wave workgroups: a world of 20,000 shapes then spreads its queries
  // over every CU (157 workgroups of 128 lanes left a third of them idle)
  constexpr int kGridQueryThreads = 64;
  if (grid_local_ok && worlds_contiguous && !(d.idle && idle_worlds() > 0)) {
    // every world active and small: each bins into a table of its own
    Dev dg = d;
    dg.grid_T = 0;  // (the shared table's counts are not used)
    const LaunchShape ga = shape_for(std::max<long long>(ns, finish_pending ? d.nb : 1));
    hipLaunchKernelGGL(k_grid_aabb, ga.grid, ga.block, 0, stream, dg, finish_pending ? 1 : 0);
    finish_pending = false;
    hipLaunchKernelGGL(k_grid_bin_worlds, dim3(d.nw), dim3(kBlock), 0, stream, d);
    launch(B2G_KERNEL_PAIRS_WRITE, ns, k_grid_pairs<kPairSlots, kGridQueryThreads, kGridLocal>,
           dim3(grid_for((long long)ns * kBlock / kGridQueryThreads)), dim3(kGridQueryThreads), 0, d);
    int rc = scan(d.pcount, nullptr, ns, d.poff, d.cnt + C_PAIRS, ns);
    if (rc) return rc;
    hipLaunchKernelGGL((k_grid_pairs_compact<kPairSlots, kGridLocal>), dim3(gq), dim3(kQueryBlock), 0, stream, d);
    tree_stale = true;
    last_pair_search = 2;
    return B2G_OK;
  }
  // about four buckets per binned shape (two measured 40% slower queries on
  // the 20k pile: more tiles share a bucket): with idle worlds kept, only the
  // active worlds' shapes are binned (a serving tick has a few percent)
  long long binned = ns;
  if (d.idle && worlds_contiguous) {
    binned = 0;
    for (int w = 0; w < d.nw && w < (int)h_wshape1.size(); ++w)
      if (h_params[w].active) binned += h_wshape1[w] - h_wshape0[w];
  }
  Dev dg = d;
  dg.grid_T = 1024;
  while (dg.grid_T < 4 * binned && dg.grid_T < d.grid_T) dg.grid_T <<= 1;
  const LaunchShape ga = shape_for(std::max<long long>(ns, finish_pending ? d.nb : 1));
  hipLaunchKernelGGL(k_grid_aabb, ga.grid, ga.block, 0, stream, dg,
                     finish_pending ? 1 : 0);
  finish_pending = false;
  const LaunchShape gc = shape_for(ns);
  hipLaunchKernelGGL(k_grid_count, gc.grid, gc.block, 0, stream, dg);
  int rc = scan(dg.gcount, nullptr, dg.grid_T, dg.goff, dg.goff + dg.grid_T, dg.grid_T);
  if (rc) return rc;
  hipLaunchKernelGGL(k_grid_scatter, gc.grid, gc.block, 0, stream, dg);
  if (d.nw > 1)
    launch(B2G_KERNEL_PAIRS_WRITE, ns, k_grid_pairs<kPairSlots, kGridQueryThreads, kGridShared>,
           dim3(grid_for(binned * kBlock / kGridQueryThreads)), dim3(kGridQueryThreads), 0, dg);
  else
    launch(B2G_KERNEL_PAIRS_WRITE, ns, k_grid_pairs<kPairSlots, kGridQueryThreads, kGridOneWorld>,
           dim3(grid_for(binned * kBlock / kGridQueryThreads)), dim3(kGridQueryThreads), 0, dg);
  if ((rc = scan(d.pcount, nullptr, ns, d.poff, d.cnt + C_PAIRS, ns))) return rc;
  if (d.nw > 1)
    hipLaunchKernelGGL((k_grid_pairs_compact<kPairSlots, kGridShared>), dim3(gq), dim3(kQueryBlock), 0, stream, dg);
  else
    hipLaunchKernelGGL((k_grid_pairs_compact<kPairSlots, kGridOneWorld>), dim3(gq), dim3(kQueryBlock), 0, stream, dg);
  tree_stale = true;  // the LBVH is not kept: a user query builds it first
  last_pair_search = 2;
  return B2G_OK;
}

int b2g_batch::broadphase(bool use_given_aabbs) {
  const int ns = d.ns;
  if (ns == 0) {
    flush_finish();
    hipLaunchKernelGGL(k_clear_pairs, dim3(1), dim3(1), 0, stream, d);
    hipLaunchKernelGGL(k_finish_pairs, dim3(1), dim3(1), 0, stream, d);
    return B2G_OK;
  }
  // A world that is not active this step has not moved since the last
  // broadphase (nothing was uploaded since, and that broadphase ran): its
  // pairs are its segments of the current contact list, and on a refit its
  // boxes, leaves and nodes are those of the last tree update.
  const bool keep =
      keep_idle && pairs_current && worlds_contiguous && !filters_dirty && !capturing && !use_given_aabbs;
  if (filters_dirty) refresh_filters();
  if (!use_given_aabbs && use_grid()) {
    d.idle = keep ? 1 : 0;
    int rc = grid_pairs();
    if (rc) {
      d.idle = 0;
      return rc;
    }
    if (d.idle && idle_worlds() > 0)
      hipLaunchKernelGGL(k_keep_idle_pairs, dim3(std::min(d.nw, 4096)), dim3(kBlock), 0, stream, d);
    return finish_pair_list(use_given_aabbs);
  }
  last_pair_search = 1;
  d.idle = keep ? 1 : 0;
  int rc = build_tree(use_given_aabbs);
  if (rc) {
    d.idle = 0;
    return rc;
  }
  const int gq = grid_for((long long)ns * kBlock / kQueryBlock);  // blocks of kQueryBlock threads
  // The same slot count and stack for both kernels. Queries that fill the
  // GPU several times over are occupancy-bound: they keep 8 pair slots and,
  // while no world exceeds 65536 shapes, 16-bit stack entries (offsets inside
  // the world's leaf range): 24 KB of LDS per block instead of 48 KB. Fewer
  // queries keep 16 slots and 32-bit entries, which cost less per access.
  const bool dense = ns >= kDenseQueryShapes;
  const bool short_stack = dense && !layout_dirty && max_world_shapes > 0 && max_world_shapes <= 65536;
  if (short_stack)
    launch(B2G_KERNEL_PAIRS_WRITE, ns, k_lbvh_pairs<kPairSlotsDense, uint16_t>, dim3(gq), dim3(kQueryBlock), 0, d);
  else if (dense)
    launch(B2G_KERNEL_PAIRS_WRITE, ns, k_lbvh_pairs<kPairSlotsDense, int>, dim3(gq), dim3(kQueryBlock), 0, d);
  else
    launch(B2G_KERNEL_PAIRS_WRITE, ns, k_lbvh_pairs<kPairSlots, int>, dim3(gq), dim3(kQueryBlock), 0, d);
  scan(d.pcount, nullptr, ns, d.poff, d.cnt + C_PAIRS, ns);
  if (short_stack)
    hipLaunchKernelGGL((k_pairs_compact<kPairSlotsDense, uint16_t>), dim3(gq), dim3(kQueryBlock), 0, stream, d);
  else if (dense)
    hipLaunchKernelGGL((k_pairs_compact<kPairSlotsDense, int>), dim3(gq), dim3(kQueryBlock), 0, stream, d);
  else
    hipLaunchKernelGGL((k_pairs_compact<kPairSlots, int>), dim3(gq), dim3(kQueryBlock), 0, stream, d);
  if (d.idle && idle_worlds() > 0)
    hipLaunchKernelGGL(k_keep_idle_pairs, dim3(std::min(d.nw, 4096)), dim3(kBlock), 0, stream, d);
  return finish_pair_list(use_given_aabbs);
}

// The pair list's segments sorted, merged with the contact list, and the
// buffers swapped.
int b2g_batch::finish_pair_list(bool use_given_aabbs) {
  // one block per segment to sort: static shapes' segments, plus a few
  // over-full dynamic shapes
  hipLaunchKernelGGL(k_sort_pair_segments, dim3(std::min(std::max(static_shapes + 32, 64), 4096)), dim3(kBlock), 0,
                     stream, d);
  const LaunchShape mc = shape_for(std::max(est_contacts, d.ns) * 5 / 4);
  hipLaunchKernelGGL(k_merge_contacts, mc.grid, mc.block, 0, stream,
                     d);
  if (d.events_on || d.sleep_on)  // EndContact events; wake-ups of the destroyed touching contacts' bodies
    hipLaunchKernelGGL(k_dead_contact_events, dim3(grid_for(std::max(est_contacts, 256))), dim3(kBlock), 0, stream,
                       d);
  hipLaunchKernelGGL(k_finish_pairs, dim3(1), dim3(1), 0, stream, d);
  d.idle = 0;
  swap_contact_buffers();
  // a captured broadphase has not run yet: the replay's caller sets this
  if (!capturing) pairs_current = !use_given_aabbs;
  return B2G_OK;
}

// Host-side greedy colouring of the joint graph (topology changes only on
// upload), plus the CSR of joint partners that must not collide.
int b2g_batch::rebuild_joint_tables() {
  const int nj = d.nj, nb = d.nb;
  std::vector<unsigned long long> used(std::max(nb, 1), 0ull);
  std::vector<int> col(nj, 0);
  int ncol = 0;
  for (int q = 0; q < nj; ++q) {
    const b2g_revolute& j = h_joints[q];
    unsigned long long m = 0;
    const bool da = h_btype[j.body_a] == B2G_DYNAMIC, db = h_btype[j.body_b] == B2G_DYNAMIC;
    if (da) m |= used[j.body_a];
    if (db) m |= used[j.body_b];
    int c = 0;
    while (c < 64 && (m >> c) & 1ull) ++c;
    if (c >= 64) return fail(B2G_ERR_CAPACITY, "joint graph needs more than 64 colours");
    col[q] = c;
    if (da) used[j.body_a] |= 1ull << c;
    if (db) used[j.body_b] |= 1ull << c;
    ncol = std::max(ncol, c + 1);
  }
  std::vector<int> order(nj);
  h_jcol_off.assign(ncol + 1, 0);
  for (int q = 0; q < nj; ++q) h_jcol_off[col[q] + 1]++;
  for (int c = 0; c < ncol; ++c) h_jcol_off[c + 1] += h_jcol_off[c];
  std::vector<int> cur(h_jcol_off.begin(), h_jcol_off.end());
  for (int q = 0; q < nj; ++q) order[cur[col[q]]++] = q;
  // CSR of non-colliding joint partners
  std::vector<int> off(nb + 1, 0), other;
  for (int q = 0; q < nj; ++q) {
    if (h_joints[q].flags & 1) continue;
    off[h_joints[q].body_a + 1]++;
    off[h_joints[q].body_b + 1]++;
  }
  for (int b = 0; b < nb; ++b) off[b + 1] += off[b];
  other.assign(std::max(off[nb], 1), 0);
  std::vector<int> pos(off.begin(), off.end());
  for (int q = 0; q < nj; ++q) {
    if (h_joints[q].flags & 1) continue;
    other[pos[h_joints[q].body_a]++] = h_joints[q].body_b;
    other[pos[h_joints[q].body_b]++] = h_joints[q].body_a;
  }
  {  // bodies a joint attaches are no orphans (b2Body::GetJointList, b2_world.cpp:545)
    std::vector<int> jointed(std::max(nb, 1), 0);
    for (int q = 0; q < nj; ++q) jointed[h_joints[q].body_a] = jointed[h_joints[q].body_b] = 1;
    if (nb) HIP_TRY(hipMemcpy(d.bjointed, jointed.data(), nb * sizeof(int), hipMemcpyHostToDevice));
  }
  if (nj) HIP_TRY(hipMemcpy(d.jorder, order.data(), nj * sizeof(int), hipMemcpyHostToDevice));
  {  // joints grouped by world (of body A), for the per-world island pass
    const int nw = d.nw;
    std::vector<int> woff(nw + 1, 0), wlist(std::max(nj, 1));
    auto world_of = [&](int q) { return h_bworld[h_joints[q].body_a]; };
    for (int q = 0; q < nj; ++q)
      if (world_of(q) >= 0 && world_of(q) < nw) woff[world_of(q) + 1]++;
    for (int w = 0; w < nw; ++w) woff[w + 1] += woff[w];
    std::vector<int> wcur(woff.begin(), woff.end() - 1);
    for (int q = 0; q < nj; ++q)
      if (world_of(q) >= 0 && world_of(q) < nw) wlist[wcur[world_of(q)]++] = q;
    HIP_TRY(hipMemcpy(d.wj_off, woff.data(), (nw + 1) * sizeof(int), hipMemcpyHostToDevice));
    if (nj) HIP_TRY(hipMemcpy(d.wj_list, wlist.data(), nj * sizeof(int), hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMemcpy(d.bj_off, off.data(), (nb + 1) * sizeof(int), hipMemcpyHostToDevice));
  if (off[nb]) HIP_TRY(hipMemcpy(d.bj_other, other.data(), off[nb] * sizeof(int), hipMemcpyHostToDevice));
  joints_dirty = false;
  filters_dirty = true;
  pairs_current = false;
  return B2G_OK;
}

int b2g_batch::read_counters() {
  HIP_TRY(hipMemcpyAsync(pinned, d.cnt, (kCntColOff + kColourSlots + 1) * sizeof(int), hipMemcpyDeviceToHost,
                         stream));
  HIP_TRY(hipStreamSynchronize(stream));
  return B2G_OK;
}

// ---- the step -------------------------------------------------------------
// One launch sequence serves both modes. Eager mode reads the colour offsets
// back once per step (sizing the sweeps exactly, and running extra claim
// rounds when many contacts are new); graph mode replays a captured copy of
// the same sequence with no host round trip: `K` colour slots per sweep, a
// fixed number of claim rounds, and leftovers in the serial group.

void b2g_batch::swap_contact_buffers() {
  std::swap(d.ckey, d.nkey);
  std::swap(d.cpair, d.npair);
  std::swap(d.cman, d.nman);
  std::swap(d.cmat, d.nmat);
  std::swap(d.ccol, d.ncol);
  std::swap(d.poff, d.poff_prev);  // the new list's segments
  parity ^= 1;
}

void b2g_batch::graphs_clear() {
  for (auto& g : graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  graphs.clear();
}

// `collect_last`: the last round lists its losers for k_colour_finish
void b2g_batch::enqueue_collide_and_colour(int K, int rounds, bool collect_last) {
  enqueue_collide();
  const LaunchShape ck = shape_for(std::max(est_cons, 256) * 5 / 4);
  hipLaunchKernelGGL(k_colour_init, ck.grid, ck.block, 0, stream, d, K);
  for (int r = 0; r < rounds; ++r) enqueue_colour_round(K, collect_last && r == rounds - 1);
}

// An eager step's colouring: device-wide rounds; a single world finishes in
// one workgroup (batches keep their rounds device-wide: worlds that colour
// afresh would otherwise all land on one CU). The host still runs more
// rounds if the read-back reports leftovers.
void b2g_batch::enqueue_collide_and_colour_eager() {
  if (d.nw == 1) {
    enqueue_collide_and_colour(kMaxColours, kColourRoundsOneWorld, true);
    hipLaunchKernelGGL(k_colour_finish, dim3(1), dim3(kFinishThreads), 0, stream, d, rounds_pending,
                       kGraphFinishRounds, kMaxColours, kFinishMaxLeft);
    rounds_pending += kGraphFinishRounds;
  } else {
    enqueue_collide_and_colour(kMaxColours, kColourRoundsPerStep);
  }
}

// Collide (narrowphase), the bodies' velocity integration, and the compacted
// list of the contacts the step solves
void b2g_batch::enqueue_collide() {
  const int nb = d.nb;
  const int gc = grid_for(std::max(est_contacts, 256) * 5 / 4);
  const int gb = grid_for(nb);
  // narrowphase, and the bodies' velocity integration in the same launch
  launch(B2G_KERNEL_NARROWPHASE, est_contacts, k_narrowphase, dim3(std::max(gc, gb)), dim3(kBlock), 0, d, 1);
  // solved contacts: scan of the touching flags and the compacted list (the
  // bound only picks the launch shape; any device count is handled)
  const long long touch_bound = std::max(est_contacts, 1) * 5LL / 4;
  const int touch_tiles = (int)((touch_bound + kScanTile - 1) / kScanTile);
  if (touch_bound > 8192 && touch_tiles <= kScanFoldTiles) {
    const int g = std::min(touch_tiles, 4096);
    hipLaunchKernelGGL(k_scan_tiles, dim3(g), dim3(kBlock), 0, stream, d.touch, d.cnt + C_CONTACTS, 0, d.scan_tiles);
    hipLaunchKernelGGL(k_touch_scan_apply, dim3(g), dim3(kBlock), 0, stream, d, d.scan_tiles);
  } else {
    scan(d.touch, d.cnt + C_CONTACTS, 0, d.tpos, d.cnt + C_CONS, touch_bound);
    hipLaunchKernelGGL(k_compact_touching, dim3(gc), dim3(kBlock), 0, stream, d);
  }
}

// ---------------------------------------------------------------------------
// The cells solver (b2g_partition.h): a world's constraints split by the
// cells of two staggered grids, three launches per iteration, no host read.
// ---------------------------------------------------------------------------
// Only on request: on the 20,000-body pile the colour sweeps take 0.73 ms
// per step (steps 6-25) and 0.89 ms (steps 21-500) against the cells' 0.93
// and 1.06 ms, whose per-colour rounds inside a workgroup cost about as much
// as a launch per colour, so B2G_SOLVER_AUTO does not choose them.
bool b2g_batch::use_cells() const { return solver_mode == B2G_SOLVER_CELLS; }

void b2g_batch::enqueue_cells_solve(int vi, int pi, bool finish_in_broadphase) {
  const int gk = grid_for(std::max(est_cons, 256) * 5 / 4);
  const int gb = grid_for(d.nb);
  // the grid (bounds, then the cell size in the bounds launch's last
  // workgroup), the cells' body counts (scanned in the count launch's last
  // workgroup), the bodies and constraints into their cells, the colours
  // (the cross group's in the colour launch's last workgroup)
  hipLaunchKernelGGL(k_part_bounds, dim3(std::min(gb, kPartBoundsBlocks)), dim3(kBlock), 0, stream, d);
  hipLaunchKernelGGL(k_part_hist, dim3(std::min(gb, kPartBoundsBlocks)), dim3(kBlock), 0, stream, d);
  hipLaunchKernelGGL(k_part_classify, dim3(std::max(gk, gb)), dim3(kBlock), 0, stream, d);
  hipLaunchKernelGGL(k_part_colour_cells, dim3(2 * kCells + 1), dim3(kCellThreads), 0, stream, d);
  enqueue_islands_and_init();
  // the cells' share of the constraints (the cross group's count is the last
  // read-back's): the timed run is the two cell launches
  const long long cells_cons = std::max((long long)est_cons - (long long)pinned[C_PART_CROSS], 1LL);
  // one sweep: the A cells, the B cells, then the cross group
  auto sweep = [&](int id, auto kernel, auto cross_kernel, int it) {
    tbegin(id);
    hipLaunchKernelGGL(kernel, dim3(kCells), dim3(kCellThreads), 0, stream, d, 0, it);
    hipLaunchKernelGGL(kernel, dim3(kCells), dim3(kCellThreads), 0, stream, d, 1, it);
    tend(id, cells_cons, 2);
    hipLaunchKernelGGL(cross_kernel, dim3(1), dim3(kCrossThreads), 0, stream, d, 2, it);
  };
  auto joint_sweep = [&](auto kernel, auto... extra) {
    const int njc = (int)h_jcol_off.size() - 1;
    for (int c = 0; c < njc; ++c) {
      const int b = h_jcol_off[c], e = h_jcol_off[c + 1];
      if (e > b) hipLaunchKernelGGL(kernel, dim3(grid_for(e - b)), dim3(kBlock), 0, stream, d, b, e, extra...);
    }
  };
  sweep(B2G_KERNEL_WARM_START, k_part_sweep<kPartWarm, kCellThreads>, k_part_sweep<kPartWarm, kCrossThreads>, 0);
  joint_sweep(k_joint_init);
  for (int it = 0; it < vi; ++it) {
    joint_sweep(k_joint_velocity);
    sweep(B2G_KERNEL_CONTACT_VELOCITY, k_part_sweep<kPartVelocity, kCellThreads>,
          k_part_sweep<kPartVelocity, kCrossThreads>, 0);
  }
  (void)hipStreamWaitEvent(stream, join_ev, 0);
  {
    const LaunchShape si = shape_for(std::max(est_cons + d.nj, d.nb));
    hipLaunchKernelGGL(k_store_and_integrate, si.grid, si.block, 0, stream, d, 1);
  }
  for (int it = 0; it < pi; ++it) {
    sweep(B2G_KERNEL_CONTACT_POSITION, k_part_sweep<kPartPosition, kCellThreads>,
          k_part_sweep<kPartPosition, kCrossThreads>, it);
    joint_sweep(k_joint_position, it);
  }
  if (finish_in_broadphase)
    finish_pending = true;
  else
    hipLaunchKernelGGL(k_finish_bodies, dim3(grid_for(d.nb)), dim3(kBlock), 0, stream, d);
}

// an eager step by cells: the whole step is enqueued without a host read
int b2g_batch::cells_step(int vi, int pi, unsigned flags) {
  const bool prof_on = (flags & B2G_STEP_PROFILE) != 0;
  last_path = 3;
  if (prof_on) HIP_TRY(hipEventRecord(ev[0], stream));
  int rc;
  if (flags & B2G_STEP_FIND_NEW_CONTACTS) {
    if ((rc = broadphase(false))) return rc;
  }
  d.segs = pairs_current && worlds_contiguous ? 1 : 0;
  d.part_on = 1;
  enqueue_collide();
  if (prof_on) HIP_TRY(hipEventRecord(ev[1], stream));
  enqueue_cells_solve(vi, pi, !(flags & B2G_STEP_SKIP_BROADPHASE));
  d.part_on = 0;
  if (prof_on) HIP_TRY(hipEventRecord(ev[2], stream));
  if (!(flags & B2G_STEP_SKIP_BROADPHASE)) {
    if ((rc = broadphase(false))) return rc;
  }
  if (prof_on) HIP_TRY(hipEventRecord(ev[3], stream));
  if (tev_used > 64) tcollect();
  if (!(flags & B2G_STEP_ASYNC) || prof_on) {
    if ((rc = read_counters())) return rc;
    if ((rc = check_status())) return rc;
    est_cons = pinned[C_CONS];
  }
  if (prof_on) {
    float t01, t12, t23, t03;
    HIP_TRY(hipEventElapsedTime(&t01, ev[0], ev[1]));
    HIP_TRY(hipEventElapsedTime(&t12, ev[1], ev[2]));
    HIP_TRY(hipEventElapsedTime(&t23, ev[2], ev[3]));
    HIP_TRY(hipEventElapsedTime(&t03, ev[0], ev[3]));
    prof.collide = t01;
    prof.solve = t12 + t23;
    prof.broadphase = t23;
    prof.step = t03;
  }
  return B2G_OK;
}

void b2g_batch::enqueue_colour_round(int K, bool collect) {
  // the rounds walk the pending list: after the first steps it holds only the
  // newly touching constraints, so a modest grid covers it
  const LaunchShape ck = shape_for(std::max(est_cons, 256) * 5 / 4, 2048);
  hipLaunchKernelGGL(k_colour_claim, ck.grid, ck.block, 0, stream, d, rounds_pending);
  hipLaunchKernelGGL(k_colour_resolve, ck.grid, ck.block, 0, stream, d, rounds_pending, K, collect ? 1 : 0);
  ++rounds_pending;
}

void b2g_batch::enqueue_colour_groups() {
  const int gk = grid_for(std::max(est_cons, 256) * 5 / 4);
  hipLaunchKernelGGL(k_colour_count, dim3(gk), dim3(kBlock), 0, stream, d);
  hipLaunchKernelGGL(k_colour_scatter, dim3(gk), dim3(kBlock), 0, stream, d, rounds_pending);
  round_host += (unsigned)rounds_pending;
  rounds_pending = 0;
  hipLaunchKernelGGL(k_sort_serial_group, dim3(1), dim3(1), 0, stream, d);
}

// per-world island pass: LDS forest sizes (bodies); larger worlds use the
// device-wide union-find
constexpr int kIslandLdsSmall = 2048, kIslandLdsMid = 8192, kIslandLdsLarge = 32768;

// The island pass forks onto the side stream (it needs the compacted solved
// contacts, which exist by now, and nothing of the velocity phase); the
// position phase joins it (enqueue_islands_join).
void b2g_batch::enqueue_islands_and_init() {
  // With sleeping, the islands decide which bodies move at all, so they come
  // first, on the main stream: seeds, wake-ups and the velocity integration of
  // the bodies that move follow them.
  // A captured step keeps them on the main stream too: a graph's branch on
  // another stream costs a dependency across hardware queues at each end.
  const bool own_stream = d.sleep_on || capturing;
  hipStream_t is = own_stream ? stream : side;
  if (!own_stream) {
    (void)hipEventRecord(fork_ev, stream);
    (void)hipStreamWaitEvent(side, fork_ev, 0);
  }
  if (worlds_contiguous && max_world_bodies <= kIslandLdsLarge) {
    // one workgroup per world, the forest in LDS
    if (max_world_bodies <= kIslandLdsSmall)
      hipLaunchKernelGGL(k_world_islands<kIslandLdsSmall>, dim3(d.nw), dim3(256), 0, is, d);
    else if (max_world_bodies <= kIslandLdsMid)
      hipLaunchKernelGGL(k_world_islands<kIslandLdsMid>, dim3(d.nw), dim3(512), 0, is, d);
    else
      hipLaunchKernelGGL(k_world_islands<kIslandLdsLarge>, dim3(d.nw), dim3(1024), 0, is, d);
  } else {
    hipLaunchKernelGGL(k_island_link, dim3(grid_for(std::max(est_cons, 1) + d.nj)), dim3(kBlock), 0, is, d);
    hipLaunchKernelGGL(k_island_label, dim3(grid_for(d.nb)), dim3(kBlock), 0, is, d);
  }
  if (d.sleep_on) {
    hipLaunchKernelGGL(k_sleep_seed, dim3(grid_for(d.nb)), dim3(kBlock), 0, stream, d);
    hipLaunchKernelGGL(k_sleep_activate, dim3(grid_for(d.nb)), dim3(kBlock), 0, stream, d);
  }
  (void)hipEventRecord(join_ev, is);
  const LaunchShape ci = shape_for(std::max(est_cons, 256) * 5 / 4);
  hipLaunchKernelGGL(k_contact_init, ci.grid, ci.block, 0, stream, d);
  if (d.nj) hipLaunchKernelGGL(k_joint_prepare, dim3(grid_for(d.nj)), dim3(kBlock), 0, stream, d);
}

void b2g_batch::enqueue_islands_join() {
  (void)hipStreamWaitEvent(stream, join_ev, 0);
  hipLaunchKernelGGL(k_island_assign, dim3(grid_for(std::max(est_cons, 1) + d.nj)), dim3(kBlock), 0, stream, d);
}

// colours of at most this many slots (one pass of the tail launch's 1024
// lanes) can share the tail launch: on the 20k pile a 780-constraint colour
// there costs less than its own launch (0.71 against 0.72 ms per step); 2048
// (two passes) measured slower (0.77)
constexpr int kTailColourMax = 1024;
// Per-colour sweep launches run 64-lane workgroups: a colour's few thousand
// constraints then spread over as many CUs as possible, one wave each (the
// launch is latency-bound). On the 20k pile: 0.685 ms per step against 0.701
// with 128 lanes and 0.713 with 256 (steps 6-25), 0.858 against 0.880 and
// 0.898 (steps 21-500).
constexpr int kSweepThreads = 64;

// colours: the slots to sweep; sizes[c] (eager) or -1 (graph: unknown);
// begins[c]: first slot of each swept colour (eager; empty in graph mode)
void b2g_batch::enqueue_solve(int vi, int pi, const std::vector<int>& colours, const std::vector<int>& sizes,
                              const std::vector<int>& begins, int graph_grid, bool finish_in_broadphase) {
  // graph mode (sizes unknown): the shared tail launch also runs the serial
  // group, after the last colour
  const bool tail = !sizes.empty() && sizes[0] < 0;
  // the trailing colours [tq, nnorm) of `colours` share one launch: in eager
  // mode those of at most kTailColourMax slots; captured steps start the tail
  // where the last eager step's did (sizes are not known there)
  const int nq = (int)colours.size();
  const bool has_serial = nq > 0 && colours[nq - 1] == kSerialColour;
  const int nnorm = has_serial ? nq - 1 : nq;
  int tq = nnorm;
  if (!tail) {
    while (tq > 0 && sizes[tq - 1] <= kTailColourMax && nnorm - tq < kTailMax) --tq;
    if (nnorm - tq < 2) tq = nnorm;  // one small colour: its own launch is as cheap
    graph_tail_from = tq < nnorm ? colours[tq] : (nnorm ? colours[nnorm - 1] + 1 : 0);
    graph_big_items = 0;
    for (int q = 0; q < tq; ++q) graph_big_items = std::max(graph_big_items, sizes[q]);
  } else if (nnorm > 0) {
    // captured steps: the colours the last eager step launched one by one,
    // then one tail launch over every further colour slot (it reads their
    // offsets on the device and skips the empty ones) and the serial group
    tq = std::min(nnorm - 1, graph_tail_from);
  }
  TailSlots ts{};
  int tail_items = 0;
  if (tq < nnorm) {
    ts.c0 = colours[tq];
    ts.n = nnorm - tq;
    ts.known = tail ? 0 : 1;
    if (!tail) {
      for (int k = 0; k < ts.n; ++k) {
        ts.off[k] = begins[tq + k];
        tail_items += sizes[tq + k];
      }
      ts.off[ts.n] = begins[nnorm - 1] + sizes[nnorm - 1];
    }
  }
  auto contact_sweep = [&](int id, auto kernel, auto tail_kernel, auto... extra) {
    long long run_items = 0;
    if (tq > 0) tbegin(id);  // one timed run: the per-colour launches
    for (int q = 0; q < tq; ++q) {
      const int c = colours[q];
      const int n = sizes[q];
      const int2 slots = begins.empty() ? make_int2(-1, -1) : make_int2(begins[q], begins[q] + n);
      hipLaunchKernelGGL(kernel, dim3(n >= 0 ? grid_for((long long)n * kBlock / kSweepThreads) : graph_grid),
                         dim3(kSweepThreads), 0, stream, d, c, slots, extra...);
      run_items += n;
    }
    if (tq > 0) tend(id, run_items, tq);
    if (tq < nnorm) {  // graph mode: the serial group runs at the end of the tail launch
      // the tail is another kernel: not timed with the per-colour sweeps
      const int it_arg[] = {0, extra...};  // the position iteration, 0 for the velocity sweeps
      launch(kUntimed, tail ? -1 : tail_items, tail_kernel, dim3(1), dim3(kTailThreads), 0, d, ts,
             it_arg[sizeof...(extra)], tail ? 1 : 0);
    }
    if (has_serial) {
      const int q = nq - 1;
      const int2 slots = begins.empty() ? make_int2(-1, -1) : make_int2(begins[q], begins[q] + sizes[q]);
      launch(id, sizes[q], kernel, dim3(1), dim3(1), 0, d, kSerialColour, slots, extra...);
    }
  };
  auto joint_sweep = [&](auto kernel, auto... extra) {
    const int njc = (int)h_jcol_off.size() - 1;
    for (int c = 0; c < njc; ++c) {
      const int b = h_jcol_off[c], e = h_jcol_off[c + 1];
      if (e > b)
        hipLaunchKernelGGL(kernel, dim3(grid_for((long long)(e - b) * kBlock / kSweepThreads)), dim3(kSweepThreads), 0,
                           stream, d, b, e, extra...);
    }
  };
  // warm start: a gather per body when every colour has a body-table entry
  const int top_colour = nnorm ? colours[nnorm - 1] : -1;
  if (tail) {  // captured steps: the gather, then the colours past the body table and the serial group
    const LaunchShape wg = shape_for(d.nb);
    launch(B2G_KERNEL_WARM_START, -1, k_warm_gather, wg.grid, wg.block, 0, d);
    TailSlots wt{};
    wt.c0 = kBodyTab;
    wt.n = kMaxColours - kBodyTab;
    wt.known = 0;
    launch(kUntimed, -1, k_sweep_tail<kTailWarm>, dim3(1), dim3(kTailThreads), 0, d, wt, 0, 1);
  } else if (top_colour < kBodyTab) {
    int items = 0;
    for (int q = 0; q < nnorm; ++q) items += std::max(sizes[q], 0);
    const LaunchShape wg = shape_for(d.nb);
    launch(B2G_KERNEL_WARM_START, tail ? -1 : items, k_warm_gather, wg.grid, wg.block, 0, d);
    if (has_serial || tail) {  // captured steps: a serial group may appear
      const int2 slots = has_serial && !begins.empty() ? make_int2(begins[nq - 1], begins[nq - 1] + sizes[nq - 1])
                                                       : make_int2(-1, -1);
      hipLaunchKernelGGL(k_warm_start, dim3(1), dim3(1), 0, stream, d, kSerialColour, slots);
    }
  } else {
    contact_sweep(B2G_KERNEL_WARM_START, k_warm_start, k_sweep_tail<kTailWarm>);
  }
  joint_sweep(k_joint_init);
  for (int it = 0; it < vi; ++it) {
    joint_sweep(k_joint_velocity);
    contact_sweep(B2G_KERNEL_CONTACT_VELOCITY, k_contact_velocity, k_sweep_tail<kTailVelocity>);
  }
  // the island pass (side stream) ended long before the velocity phase did;
  // the same launch stores impulses, integrates positions and assigns islands
  (void)hipStreamWaitEvent(stream, join_ev, 0);
  {
    const LaunchShape si = shape_for(std::max(est_cons + d.nj, d.nb));
    hipLaunchKernelGGL(k_store_and_integrate, si.grid, si.block, 0, stream, d, 1);
  }
  for (int it = 0; it < pi; ++it) {  // each launch applies the island early exit itself
    contact_sweep(B2G_KERNEL_CONTACT_POSITION, k_contact_position, k_sweep_tail<kTailPosition>, it);
    joint_sweep(k_joint_position, it);
  }
  if (finish_in_broadphase)
    finish_pending = true;  // the broadphase's first launch (k_aabb) ends the step
  else
    hipLaunchKernelGGL(k_finish_bodies, dim3(grid_for(d.nb)), dim3(kBlock), 0, stream, d);
}

// the sticky error bits of the last counter read-back (pinned), as a status
int b2g_batch::status_error() const {
  const int s = pinned[C_STATUS];
  if (s & S_CONTACT_OVERFLOW)
    return fail(B2G_ERR_CAPACITY, "contact capacity exceeded (" + std::to_string(pinned[C_PAIRS]) + " pairs)");
  if (s & S_STACK_OVERFLOW) return fail(B2G_ERR_CAPACITY, "BVH traversal stack overflow");
  if (s & S_TREE_CORRUPT) return fail(B2G_ERR_INVALID, "BVH inconsistent (sort/build)");
  return B2G_OK;
}

int b2g_batch::check_status() {
  est_contacts = pinned[C_CONTACTS];
  return status_error();
}

// World ranges for the per-world solver: valid when every world's bodies and
// shapes are contiguous (the layout b2a_batch_from_scenes produces) and the
// batch has no joints.
// worlds above this many bodies get a solver launch of their own (see
// launch_solve_worlds)
constexpr int kBigWorldBodies = 1024;

// world-local shape numbering: first shape of every world; local indices
// must fit 24 bits (the colouring rank packs two of them)
int b2g_batch::update_first_shapes() {
  std::vector<int>& w0 = h_wshape0;
  std::fill(w0.begin(), w0.end(), INT_MAX);
  for (int i = 0; i < d.ns; ++i) {
    const int w = h_bworld[h_shape_body[i]];
    w0[w] = std::min(w0[w], i);
  }
  for (int i = 0; i < d.ns; ++i)
    if (i - w0[h_bworld[h_shape_body[i]]] >= (1 << 24))
      return fail(B2G_ERR_CAPACITY, "a world spans more than 2^24 shape indices");
  for (int& v : w0)
    if (v == INT_MAX) v = 0;
  HIP_TRY(hipMemcpy(d.wshape0, w0.data(), w0.size() * sizeof(int), hipMemcpyHostToDevice));
  return B2G_OK;
}

int b2g_batch::update_world_layout() {
  const int nw = std::max(d.nw, 1), nb = d.nb, ns = d.ns;
  // every body must belong to a declared world: the per-world tables below
  // (and the kernels' world records) have d.nw entries
  for (int b = 0; b < nb; ++b) {
    const int w = h_bworld[b];
    if (w < 0 || w >= d.nw)
      return fail(B2G_ERR_INVALID, "body " + std::to_string(b) + " belongs to world " + std::to_string(w) +
                                       ", but " + std::to_string(d.nw) +
                                       " worlds are declared (b2g_set_worlds / b2g_set_world_params)");
  }
  std::vector<int> b0(nw, INT_MAX), b1(nw, -1), bc(nw, 0), s0(nw, INT_MAX), s1(nw, -1), sc(nw, 0);
  for (int b = 0; b < nb; ++b) {
    const int w = h_bworld[b];
    b0[w] = std::min(b0[w], b);
    b1[w] = std::max(b1[w], b);
    ++bc[w];
  }
  for (int i = 0; i < ns; ++i) {
    if (h_shape_body[i] < 0 || h_shape_body[i] >= nb)
      return fail(B2G_ERR_INVALID, "shape " + std::to_string(i) + " refers to body " +
                                       std::to_string(h_shape_body[i]) + " of " + std::to_string(nb));
    const int w = h_bworld[h_shape_body[i]];
    s0[w] = std::min(s0[w], i);
    s1[w] = std::max(s1[w], i);
    ++sc[w];
  }
  bool ok = true;
  std::vector<int2> wb(nw);
  std::vector<int> ws1(nw), order(nw);
  max_world_bodies = 0;
  max_world_shapes = 0;
  for (int w = 0; w < nw; ++w) {
    if (bc[w] && b1[w] - b0[w] + 1 != bc[w]) ok = false;
    if (sc[w] && s1[w] - s0[w] + 1 != sc[w]) ok = false;
    if (sc[w] && (w >= (int)h_wshape0.size() || h_wshape0[w] != s0[w])) ok = false;  // as uploaded
    wb[w] = bc[w] ? make_int2(b0[w], b1[w] + 1) : make_int2(0, 0);
    ws1[w] = sc[w] ? s1[w] + 1 : h_wshape0[w];
    max_world_bodies = std::max(max_world_bodies, bc[w]);
    max_world_shapes = std::max(max_world_shapes, sc[w]);
    order[w] = w;
  }
  // largest worlds first: their workgroups start before the small ones fill in
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return bc[a] > bc[b]; });
  big_worlds = 0;
  max_small_world_bodies = 0;
  for (int w = 0; w < nw; ++w) {
    if (bc[w] > kBigWorldBodies) ++big_worlds;
    else max_small_world_bodies = std::max(max_small_world_bodies, bc[w]);
  }
  HIP_TRY(hipMemcpy(d.wbody, wb.data(), nw * sizeof(int2), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d.wshape1, ws1.data(), nw * sizeof(int), hipMemcpyHostToDevice));
  h_wshape1 = ws1;
  HIP_TRY(hipMemcpy(d.worder, order.data(), nw * sizeof(int), hipMemcpyHostToDevice));
  h_worder = order;
  worlds_contiguous = ok;
  per_world_ok = ok && d.nj == 0;
  layout_dirty = false;
  joints_dirty = true;  // joints are grouped by world
  return B2G_OK;
}

bool b2g_batch::use_per_world() const {
  if (!per_world_ok || solver_mode == B2G_SOLVER_COLOURS) return false;
  if (solver_mode == B2G_SOLVER_WORLDS) return true;
  return d.nw >= 2 && max_world_bodies <= 8192;
}

// autotuner candidates: workgroup size x bodies staged in LDS or not (one
// timed step each); a launch's stage holds up to 156 KB of bodies (a gfx950
// workgroup may take 160 KB of LDS, the whole CU's: every world the
// per-world solver takes, 8192 bodies at most, fits)
static const int kWorldBlocks[] = {64, 128, 256, 512, 1024};
constexpr int kWorldBlockCount = 5;
constexpr int kLdsBodyLimit = (156 * 1024) / 16;
constexpr int kSmallWorldLds = 512;

int b2g_batch::world_block_now() const {
  if (world_block > 0) return world_block;
  if (tuned_block > 0) return tuned_block;
  return kWorldBlocks[std::min(tune_step, 2 * kWorldBlockCount - 1) % kWorldBlockCount];
}

bool b2g_batch::world_lds_now() const {
  if (world_lds == 0) return false;  // (a world past a launch's stage reads its bodies from global memory)
  if (world_lds > 0) return true;
  if (world_block > 0 || tuned_block > 0) return tuned_lds;
  return std::min(tune_step, 2 * kWorldBlockCount - 1) >= kWorldBlockCount;
}

// Worlds of more than kBigWorldBodies bodies (the first ones in worder) get
// full 1024-thread workgroups in a launch of their own, beside the launch of
// the other worlds at the tuned workgroup size (on the side stream): one
// large world on a small workgroup would otherwise set the whole batch's
// solve time. Every world runs the same kernel either way.

void b2g_batch::launch_solve_worlds(int vi, int pi) {
  const int wblock = world_block_now();
  const bool lds = world_lds_now();
  // a mostly idle batch (serving) launches its active worlds only; a
  // captured step keeps every world's workgroup (its activity is not known)
  const int* order = d.worder;
  int n_all = d.nw, n_big = big_worlds;
  if (!capturing && (int)h_worder.size() >= d.nw && 2 * idle_worlds() > d.nw) {
    (void)hipEventSynchronize(order_ev);  // the last step's copy has left the staging
    n_all = n_big = 0;
    for (int q = 0; q < d.nw; ++q) {
      const int w = h_worder[q];
      if (!h_params[w].active) continue;
      stage_order[n_all++] = w;
      if (q < big_worlds) ++n_big;
    }
    if (n_all > 0) (void)hipMemcpyAsync(d.wact, stage_order, n_all * sizeof(int), hipMemcpyHostToDevice, stream);
    (void)hipEventRecord(order_ev, stream);
    order = d.wact;
  }
  // `bodies`: the LDS stage of each workgroup of the launch (its worlds of
  // more bodies read theirs from global memory)
  auto part = [&](hipStream_t st, int first, int count, int block, int bodies) {
    if (count <= 0) return;
    bodies = std::min(bodies, kLdsBodyLimit);
    if (lds && bodies > 0)
      hipLaunchKernelGGL(k_solve_worlds<true>, dim3(count), dim3(block), sizeof(float4) * (size_t)bodies, st, d, vi,
                         pi, first, order, bodies);
    else
      hipLaunchKernelGGL(k_solve_worlds<false>, dim3(count), dim3(block), 0, st, d, vi, pi, first, order, 0);
  };
  tbegin(B2G_KERNEL_SOLVER);
  if (n_big > 0 && n_big < n_all && wblock < 1024) {
    // the large worlds (the longest workgroups) are submitted first: where
    // both launches share a queue (a graph replay may), their workgroups are
    // then dispatched before the thousands of small ones
    (void)hipEventRecord(fork_ev, stream);
    (void)hipStreamWaitEvent(side, fork_ev, 0);
    part(stream, 0, n_big, 1024, max_world_bodies);
    // the small worlds' launch stages up to kSmallWorldLds bodies: a stage
    // sized for its largest world (1024 bodies, 16 KB) held 10 workgroups
    // per CU, and its thousands of worlds ran in two rounds
    part(side, n_big, n_all - n_big, wblock, std::min(max_small_world_bodies, kSmallWorldLds));
    (void)hipEventRecord(join_ev, side);
    (void)hipStreamWaitEvent(stream, join_ev, 0);
  } else {
    part(stream, 0, n_all, wblock, max_world_bodies);
  }
  tend(B2G_KERNEL_SOLVER, est_cons);
}

int b2g_batch::step(int vi, int pi, unsigned flags) {
  flush_finish();  // (a finish is only deferred to a broadphase of the same step)
  if (layout_dirty) {
    int rc = update_world_layout();
    if (rc) return rc;
  }
  if (joints_dirty) {
    int rc = rebuild_joint_tables();
    if (rc) return rc;
    graphs_clear();
  }
  if (round_host + 4096u > 0xffffu) {
    // the claim-round tag is 16 bits: clear the claim slots before it wraps
    hipLaunchKernelGGL(k_reset_colour_rounds, dim3(grid_for(d.nb)), dim3(kBlock), 0, stream, d);
    round_host = 0;
  }
  d.events_on = (flags & B2G_STEP_EVENTS) ? 1 : 0;
  d.sleep_on = sleep_now() ? 1 : 0;
  d.segs = 0;  // an eager step sets it once its contact list is known
  const bool collide_only = (flags & B2G_STEP_COLLIDE_ONLY) != 0;
  if ((flags & B2G_STEP_SKIP_BROADPHASE) && !collide_only) {  // bodies move without a rebuild or a pair query
    tree_stale = true;
    pairs_current = false;
  }
  if (d.events_on) hipLaunchKernelGGL(k_events_reset, dim3(1), dim3(1), 0, stream, d);
  if (collide_only) return collide_only_step(flags);
  const bool prof_on = (flags & B2G_STEP_PROFILE) != 0;
  const bool graph_ok = (flags & B2G_STEP_GRAPH) && !prof_on && timed_kernel < 0 && !d.sleep_on &&
                        !(flags & B2G_STEP_FIND_NEW_CONTACTS) &&
                        (graph_K > 0 || use_cells()) &&  // a cells step needs no colour count from an eager one
                        (!use_per_world() || world_block > 0 || tuned_block > 0);
  if (graph_ok) return graph_step(vi, pi, flags);
  return eager_step(vi, pi, flags);
}

// b2World::Step with dt <= 0 (src/dynamics/b2_world.cpp:1114-1138, 1164-1167):
// the broadphase when new fixtures wait for it, then Collide (manifolds,
// impulse carry-over, touching flags and their events) and ClearForces. Solve
// and its broadphase do not run, so bodies, joint impulses and the contacts'
// stored impulses stay as they are.
int b2g_batch::collide_only_step(unsigned flags) {
  if (flags & B2G_STEP_FIND_NEW_CONTACTS) {
    int rc = broadphase(false);
    if (rc) return rc;
  }
  flush_finish();
  const int gc = grid_for(std::max(est_contacts, 256) * 5 / 4);
  if (d.sleep_on) hipLaunchKernelGGL(k_sleep_reset, dim3(grid_for(d.nb)), dim3(kBlock), 0, stream, d);
  launch(B2G_KERNEL_NARROWPHASE, est_contacts, k_narrowphase, dim3(gc), dim3(kBlock), 0, d, 0);
  if (d.sleep_on) hipLaunchKernelGGL(k_sleep_wake, dim3(grid_for(d.nb)), dim3(kBlock), 0, stream, d);
  hipLaunchKernelGGL(k_clear_forces, dim3(grid_for(d.nb)), dim3(kBlock), 0, stream, d);
  if (!(flags & B2G_STEP_ASYNC)) {
    int rc = read_counters();
    if (rc) return rc;
    return check_status();
  }
  return B2G_OK;
}

// A step that applies the awake flags (a world allows sleeping, or a body
// sleeps): the colour sweeps, with the islands ahead of the velocity phase
// and the sleep timers after the position phase, before the broadphase's
// destroyed contacts wake their bodies (b2_world.cpp:526-666).
int b2g_batch::sleep_step(int vi, int pi, unsigned flags) {
  const bool prof_on = (flags & B2G_STEP_PROFILE) != 0;
  last_path = 1;
  if (prof_on) HIP_TRY(hipEventRecord(ev[0], stream));
  int rc;
  if (flags & B2G_STEP_FIND_NEW_CONTACTS) {
    if ((rc = broadphase(false))) return rc;
  }
  d.segs = pairs_current && worlds_contiguous ? 1 : 0;
  hipLaunchKernelGGL(k_sleep_reset, dim3(grid_for(d.nb)), dim3(kBlock), 0, stream, d);
  enqueue_collide_and_colour_eager();
  if (prof_on) HIP_TRY(hipEventRecord(ev[1], stream));
  enqueue_colour_groups();
  enqueue_islands_and_init();
  if ((rc = read_counters())) return rc;
  if ((rc = status_error())) return rc;
  est_cons = pinned[C_CONS];
  for (int extra = 0; pinned[C_REMAIN] > 0; extra += 2) {
    if (extra > 4096) return fail(B2G_ERR_CAPACITY, "colouring did not converge");
    enqueue_colour_round(kMaxColours);
    enqueue_colour_round(kMaxColours);
    if ((rc = read_counters())) return rc;
    if (pinned[C_REMAIN] == 0) {
      hipLaunchKernelGGL(k_colour_clear_counts, dim3(1), dim3(kBlock), 0, stream, d);
      enqueue_colour_groups();
      const LaunchShape ci = shape_for(std::max(est_cons, 256) * 5 / 4);
      hipLaunchKernelGGL(k_contact_init, ci.grid, ci.block, 0, stream, d);
      if ((rc = read_counters())) return rc;
    }
  }
  const int* off = pinned + kCntColOff;
  std::vector<int> colours, sizes, begins;
  for (int c = 0; c < kColourSlots; ++c) {
    if (off[c + 1] <= off[c]) continue;
    colours.push_back(c);
    sizes.push_back(off[c + 1] - off[c]);
    begins.push_back(off[c]);
  }
  last_ncolours = pinned[C_NCOLOURS];
  enqueue_solve(vi, pi, colours, sizes, begins, 0, !(flags & B2G_STEP_SKIP_BROADPHASE));
  hipLaunchKernelGGL(k_sleep_times, dim3(grid_for(d.nb)), dim3(kBlock), 0, stream, d);
  hipLaunchKernelGGL(k_sleep_apply, dim3(grid_for(d.nb)), dim3(kBlock), 0, stream, d, pi);
  if (prof_on) HIP_TRY(hipEventRecord(ev[2], stream));
  if (!(flags & B2G_STEP_SKIP_BROADPHASE)) {
    if ((rc = broadphase(false))) return rc;
  }
  if (prof_on) HIP_TRY(hipEventRecord(ev[3], stream));
  if (tev_used > 64) tcollect();
  if ((rc = read_counters())) return rc;
  if ((rc = check_status())) return rc;
  if (prof_on) {
    float t01, t12, t23, t03;
    HIP_TRY(hipEventElapsedTime(&t01, ev[0], ev[1]));
    HIP_TRY(hipEventElapsedTime(&t12, ev[1], ev[2]));
    HIP_TRY(hipEventElapsedTime(&t23, ev[2], ev[3]));
    HIP_TRY(hipEventElapsedTime(&t03, ev[0], ev[3]));
    prof.collide = t01;
    prof.solve = t12 + t23;
    prof.broadphase = t23;
    prof.step = t03;
  }
  return B2G_OK;
}

int b2g_batch::eager_step(int vi, int pi, unsigned flags) {
  if (d.sleep_on) return sleep_step(vi, pi, flags);
  if (use_cells()) return cells_step(vi, pi, flags);
  const bool prof_on = (flags & B2G_STEP_PROFILE) != 0;
  if (prof_on) HIP_TRY(hipEventRecord(ev[0], stream));
  if (flags & B2G_STEP_FIND_NEW_CONTACTS) {
    int rc = broadphase(false);
    if (rc) return rc;
  }
  // the contact list is the last broadphase's pair list: world contact
  // ranges are segment offsets
  d.segs = pairs_current && worlds_contiguous ? 1 : 0;
  enqueue_collide_and_colour_eager();
  if (prof_on) HIP_TRY(hipEventRecord(ev[1], stream));
  int rc;
  if (use_per_world()) {
    last_path = 2;
    if ((rc = read_counters())) return rc;
    if ((rc = status_error())) return rc;  // an earlier (asynchronous) step overflowed
    const int pw_colours = pinned[C_NCOLOURS];  // last step's (normal colours are what it reports)
    for (int extra = 0; pinned[C_REMAIN] > 0; extra += 2) {  // colour everything first
      if (extra > 4096) return fail(B2G_ERR_CAPACITY, "colouring did not converge");
      enqueue_colour_round(kMaxColours);
      enqueue_colour_round(kMaxColours);
      if ((rc = read_counters())) return rc;
    }
    est_cons = pinned[C_CONS];
    hipLaunchKernelGGL(k_world_colour_sort, dim3(d.nw), dim3(kBlock), 0, stream, d, rounds_pending);
    round_host += (unsigned)rounds_pending;
    rounds_pending = 0;
    enqueue_islands_and_init();
    enqueue_islands_join();
    const bool tuning = world_block == 0 && tuned_block == 0;
    if (tuning) HIP_TRY(hipEventRecord(ev[5], stream));
    launch_solve_worlds(vi, pi);
    if (tuning) {
      HIP_TRY(hipEventRecord(ev[6], stream));
      HIP_TRY(hipEventSynchronize(ev[6]));
      float ms = 0.0f;
      HIP_TRY(hipEventElapsedTime(&ms, ev[5], ev[6]));
      const bool lds_ok = world_lds != 0;
      const int ncand = lds_ok ? 2 * kWorldBlockCount : kWorldBlockCount;
      tune_ms[tune_step] = ms;
      if (++tune_step >= ncand) {
        int best = 0;
        for (int c = 1; c < ncand; ++c)
          if (tune_ms[c] < tune_ms[best]) best = c;
        tuned_block = kWorldBlocks[best % kWorldBlockCount];
        tuned_lds = best >= kWorldBlockCount;
      }
    }
    graph_K = kMaxColours;
    (void)pw_colours;
    if (prof_on) HIP_TRY(hipEventRecord(ev[2], stream));
    if (!(flags & B2G_STEP_SKIP_BROADPHASE)) {
      if ((rc = broadphase(false))) return rc;
    }
    if (prof_on) HIP_TRY(hipEventRecord(ev[3], stream));
    if (tev_used > 64) tcollect();
    if (!(flags & B2G_STEP_ASYNC) || prof_on) {
      if ((rc = read_counters())) return rc;
      if ((rc = check_status())) return rc;
    }
    if (prof_on) {
      float t01, t12, t23, t03;
      HIP_TRY(hipEventElapsedTime(&t01, ev[0], ev[1]));
      HIP_TRY(hipEventElapsedTime(&t12, ev[1], ev[2]));
      HIP_TRY(hipEventElapsedTime(&t23, ev[2], ev[3]));
      HIP_TRY(hipEventElapsedTime(&t03, ev[0], ev[3]));
      prof.collide = t01;
      prof.solve = t12 + t23;
      prof.broadphase = t23;
      prof.step = t03;
    }
    return B2G_OK;
  }
  last_path = 1;
  enqueue_colour_groups();
  // The counters and colour offsets are final here (the set-up kernels only
  // read them): the copy goes now, and the host waits for it alone while the
  // constraint set-up keeps the device busy.
  HIP_TRY(hipMemcpyAsync(pinned, d.cnt, (kCntColOff + kColourSlots + 1) * sizeof(int), hipMemcpyDeviceToHost,
                         stream));
  HIP_TRY(hipEventRecord(counts_ev, stream));
  enqueue_islands_and_init();
  HIP_TRY(hipEventSynchronize(counts_ev));
  if ((rc = status_error())) return rc;  // an earlier (asynchronous) step overflowed
  est_cons = pinned[C_CONS];
  for (int extra = 0; pinned[C_REMAIN] > 0; extra += 2) {  // first steps / bursts of new contacts
    if (extra > 4096) return fail(B2G_ERR_CAPACITY, "colouring did not converge");
    enqueue_colour_round(kMaxColours);
    enqueue_colour_round(kMaxColours);
    if ((rc = read_counters())) return rc;
    if (pinned[C_REMAIN] == 0) {
      hipLaunchKernelGGL(k_colour_clear_counts, dim3(1), dim3(kBlock), 0, stream, d);
      enqueue_colour_groups();
      const LaunchShape ci = shape_for(std::max(est_cons, 256) * 5 / 4);
      hipLaunchKernelGGL(k_contact_init, ci.grid, ci.block, 0, stream, d);
      if ((rc = read_counters())) return rc;
    }
  }
  const int* off = pinned + kCntColOff;
  std::vector<int> colours, sizes, begins;
  int normal_colours = 0;
  for (int c = 0; c < kColourSlots; ++c) {
    if (off[c + 1] <= off[c]) continue;
    colours.push_back(c);
    sizes.push_back(off[c + 1] - off[c]);
    begins.push_back(off[c]);
    if (c < kMaxColours) normal_colours = c + 1;
  }
  last_ncolours = pinned[C_NCOLOURS];
  // colour limit for captured steps: what this step needed plus one
  graph_K = std::min(kMaxColours, std::max(normal_colours, 1) + 1);
  enqueue_solve(vi, pi, colours, sizes, begins, 0, !(flags & B2G_STEP_SKIP_BROADPHASE));
  if (prof_on) HIP_TRY(hipEventRecord(ev[2], stream));
  if (!(flags & B2G_STEP_SKIP_BROADPHASE)) {
    if ((rc = broadphase(false))) return rc;
  }
  if (prof_on) HIP_TRY(hipEventRecord(ev[3], stream));
  if (tev_used > 64) tcollect();
  if (!(flags & B2G_STEP_ASYNC) || prof_on) {
    if ((rc = read_counters())) return rc;
    if ((rc = check_status())) return rc;
  }
  if (prof_on) {
    float t01, t12, t23, t03;
    HIP_TRY(hipEventElapsedTime(&t01, ev[0], ev[1]));
    HIP_TRY(hipEventElapsedTime(&t12, ev[1], ev[2]));
    HIP_TRY(hipEventElapsedTime(&t23, ev[2], ev[3]));
    HIP_TRY(hipEventElapsedTime(&t03, ev[0], ev[3]));
    prof.collide = t01;
    prof.solve = t12 + t23;  // b2Profile::solve includes the broadphase
    prof.broadphase = t23;
    prof.step = t03;
  }
  return B2G_OK;
}

// the whole step as one launch sequence with no host reads
int b2g_batch::enqueue_graph_body(int vi, int pi, unsigned flags, int K) {
  if (use_cells()) {  // the eager sequence as it is: nothing in it waits for the host
    d.segs = 0;  // a replay cannot know whether the list is the last broadphase's
    d.part_on = 1;
    enqueue_collide();
    enqueue_cells_solve(vi, pi, !(flags & B2G_STEP_SKIP_BROADPHASE));
    d.part_on = 0;
    if (!(flags & B2G_STEP_SKIP_BROADPHASE)) {
      int rc = broadphase(false);
      if (rc) return rc;
    }
    HIP_TRY(hipMemcpyAsync(pinned, d.cnt, C_COUNT * sizeof(int), hipMemcpyDeviceToHost, stream));
    return B2G_OK;
  }
  // per-world solver: colours below K; colour sweeps: any colour (those the
  // last eager step did not launch alone go to the tail launch)
  const int colour_limit = use_per_world() ? K : kMaxColours;
  enqueue_collide_and_colour(colour_limit, graph_colour_rounds(), true);
  hipLaunchKernelGGL(k_colour_finish, dim3(1), dim3(kFinishThreads), 0, stream, d, rounds_pending,
                     kGraphFinishRounds, colour_limit, INT_MAX);  // no host loop follows a replay
  rounds_pending += kGraphFinishRounds;
  hipLaunchKernelGGL(k_colour_leftover, dim3(grid_for(std::max(est_cons, 256) * 5 / 4)), dim3(kBlock), 0, stream,
                     d);
  if (use_per_world()) {
    hipLaunchKernelGGL(k_world_colour_sort, dim3(d.nw), dim3(kBlock), 0, stream, d, rounds_pending);
    round_host += (unsigned)rounds_pending;
    rounds_pending = 0;
    enqueue_islands_and_init();
    enqueue_islands_join();
    launch_solve_worlds(vi, pi);
    if (!(flags & B2G_STEP_SKIP_BROADPHASE)) {
      int rc = broadphase(false);
      if (rc) return rc;
    }
    HIP_TRY(hipMemcpyAsync(pinned, d.cnt, C_COUNT * sizeof(int), hipMemcpyDeviceToHost, stream));
    return B2G_OK;
  }
  enqueue_colour_groups();
  enqueue_islands_and_init();
  std::vector<int> colours, sizes;
  for (int c = 0; c < kMaxColours; ++c) {
    colours.push_back(c);
    sizes.push_back(-1);
  }
  // per-colour launches sized for the last eager step's largest colour and a
  // half (a larger colour is still covered: the launches stride over it)
  const int per_colour = std::max(256, graph_big_items + graph_big_items / 2);
  enqueue_solve(vi, pi, colours, sizes, {}, grid_for((long long)per_colour * kBlock / kSweepThreads),
                !(flags & B2G_STEP_SKIP_BROADPHASE));
  if (!(flags & B2G_STEP_SKIP_BROADPHASE)) {
    int rc = broadphase(false);
    if (rc) return rc;
  }
  HIP_TRY(hipMemcpyAsync(pinned, d.cnt, C_COUNT * sizeof(int), hipMemcpyDeviceToHost, stream));
  return B2G_OK;
}

int b2g_batch::graph_step(int vi, int pi, unsigned flags) {
  const unsigned key_flags = flags & (B2G_STEP_SKIP_BROADPHASE | B2G_STEP_EVENTS);
  const bool cells = use_cells();
  const int block = cells ? -1 : use_per_world() ? world_block_now() * 2 + (world_lds_now() ? 1 : 0) : 0;
  last_path = cells ? 3 : block ? 2 : 1;
  if (graph_ev_live && hipEventQuery(graph_ev) == hipSuccess) {
    // the last replay and its counter copy have finished: refuse to build on
    // a step that overflowed
    graph_ev_live = false;
    int rc = status_error();
    if (rc) return rc;
  }
  if (filters_dirty) refresh_filters();  // outside the captured sequence
  GraphEntry* ge = nullptr;
  for (auto& g : graphs)
    if (g.parity == parity && g.vi == vi && g.pi == pi && g.K == graph_K && g.flags == key_flags &&
        g.block == block)
      ge = &g;
  if (!ge) {
    if (graphs.size() >= 8) graphs_clear();
    const int parity0 = parity;
    const unsigned round0 = round_host;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    capturing = true;
    const int rc = enqueue_graph_body(vi, pi, flags, graph_K);
    capturing = false;
    hipError_t e = hipStreamEndCapture(stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(B2G_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(B2G_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    // capture only recorded the launches: restore the host view of the state
    if (parity != parity0) swap_contact_buffers();
    round_host = round0;
    graphs.push_back(GraphEntry{exec, parity0, vi, pi, graph_K, key_flags, block});
    ge = &graphs.back();
  }
  HIP_TRY(hipGraphLaunch(ge->exec, stream));
  HIP_TRY(hipEventRecord(graph_ev, stream));
  graph_ev_live = true;
  // mirror what the replay does to the host view
  if (!(flags & B2G_STEP_SKIP_BROADPHASE)) {
    swap_contact_buffers();
    pairs_current = true;
  }
  if (!cells) round_host += graph_colour_rounds() + kGraphFinishRounds;
  ++graph_steps;
  if (!(flags & B2G_STEP_ASYNC)) {
    HIP_TRY(hipStreamSynchronize(stream));
    int rc = check_status();
    if (rc) return rc;
    // constraints pushed to the serial group: raise the colour limit
    if (pinned[C_SERIAL] > 32 && graph_K < kMaxColours) graph_K = std::min(kMaxColours, graph_K + 1);
  }
  return B2G_OK;
}

// ---------------------------------------------------------------------------
// record conversion helpers
// ---------------------------------------------------------------------------
__host__ __device__ static inline void shape_to_device(const b2g_shape& s, int world, int4* hdr, float4* geo,
                                                       float4* mat, int4* filt) {
  *hdr = make_int4(s.type, s.body, s.count, world);
  geo[0] = make_float4(s.radius, s.centroid[0], s.centroid[1], 0.0f);
  for (int k = 0; k < 4; ++k) {
    geo[1 + k] = make_float4(s.vertices[4 * k], s.vertices[4 * k + 1], s.vertices[4 * k + 2], s.vertices[4 * k + 3]);
    geo[5 + k] = make_float4(s.normals[4 * k], s.normals[4 * k + 1], s.normals[4 * k + 2], s.normals[4 * k + 3]);
  }
  *mat = make_float4(s.friction, s.restitution, s.restitution_threshold, 0.0f);
  *filt = make_int4(s.category, s.mask, s.group, s.is_sensor);
}

__host__ __device__ static inline void manifold_to_record(const float4* m, b2g_contact* o) {
  const int packed = __builtin_bit_cast(int, m[3].z);
  o->type = packed & 0xff;
  o->point_count = (packed >> 8) & 0xff;
  o->touching = o->point_count > 0;
  o->local_normal[0] = m[2].x;
  o->local_normal[1] = m[2].y;
  o->local_point[0] = m[2].z;
  o->local_point[1] = m[2].w;
  o->point_local[0] = m[0].x;
  o->point_local[1] = m[0].y;
  o->point_local[2] = m[1].x;
  o->point_local[3] = m[1].y;
  o->normal_impulse[0] = m[0].z;
  o->tangent_impulse[0] = m[0].w;
  o->normal_impulse[1] = m[1].z;
  o->tangent_impulse[1] = m[1].w;
  o->id[0] = __builtin_bit_cast(uint32_t, m[3].x);
  o->id[1] = __builtin_bit_cast(uint32_t, m[3].y);
  if (o->point_count < 2) {
    o->point_local[2] = o->point_local[3] = 0.0f;
    o->normal_impulse[1] = o->tangent_impulse[1] = 0.0f;
    o->id[1] = 0;
  }
  if (o->point_count < 1) {
    o->point_local[0] = o->point_local[1] = 0.0f;
    o->normal_impulse[0] = o->tangent_impulse[0] = 0.0f;
    o->id[0] = 0;
  }
}

static void record_to_manifold(const b2g_contact& o, float4* m) {
  m[0] = make_float4(o.point_local[0], o.point_local[1], o.normal_impulse[0], o.tangent_impulse[0]);
  m[1] = make_float4(o.point_local[2], o.point_local[3], o.normal_impulse[1], o.tangent_impulse[1]);
  m[2] = make_float4(o.local_normal[0], o.local_normal[1], o.local_point[0], o.local_point[1]);
  const int packed = (o.type & 0xff) | ((o.point_count & 0xff) << 8);
  m[3] = make_float4(__builtin_bit_cast(float, o.id[0]), __builtin_bit_cast(float, o.id[1]),
                     __builtin_bit_cast(float, packed), 0.0f);
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* b2g_last_error(void) { return g_last_error.c_str(); }

int b2g_device_count(int32_t* count) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *count = n;
  return B2G_OK;
}

int b2g_batch_create(b2g_batch** out, const b2g_capacity* cap) {
  if (!out || !cap) return fail(B2G_ERR_INVALID, "null argument");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(B2G_ERR_NO_DEVICE, "no HIP device visible");
  if (cap->contacts >= (1 << 28)) return fail(B2G_ERR_INVALID, "contact capacity must stay below 2^28");
  b2g_batch* b = new b2g_batch();
  b->cap = *cap;
  HIP_TRY(hipGetDevice(&b->dev_id));
  int rc = b->alloc();
  if (rc) {
    b->release();
    delete b;
    return rc;
  }
  *out = b;
  return B2G_OK;
}

int b2g_batch_destroy(b2g_batch* b) {
  if (!b) return B2G_OK;
  (void)hipStreamSynchronize(b->stream);
  b->release();
  delete b;
  return B2G_OK;
}

int b2g_set_world_params(b2g_batch* b, int32_t first, int32_t count, const b2g_world_params* p) {
  if (!b || first < 0 || count < 0 || first + count > b->cap.worlds) return fail(B2G_ERR_INVALID, "world range");
  if (first + count > b->d.nw) {  // the world count is part of captured launches
    b->graphs_clear();
    b->graph_K = 0;
    b->filters_dirty = true;
    b->pairs_current = false;
    b->layout_dirty = true;
  }
  // the last call's copies have left the staging buffers (normally long ago)
  HIP_TRY(hipEventSynchronize(b->params_ev));
  float4* wp = b->stage_wpar + first;
  int* wf = b->stage_wflag + first;
  for (int i = 0; i < count; ++i) {
    b->h_params[first + i] = p[i];
    wp[i] = make_float4(p[i].gravity[0], p[i].gravity[1], p[i].dt, p[i].dt_ratio);
    wf[i] = (p[i].warm_starting ? 1 : 0) | (p[i].active ? 2 : 0) | (p[i].allow_sleep ? 4 : 0);
  }
  // in stream order: every later step sees the new parameters
  HIP_TRY(hipMemcpyAsync(b->d.wpar + first, wp, count * sizeof(float4), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(b->d.wflag + first, wf, count * sizeof(int), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipEventRecord(b->params_ev, b->stream));
  b->d.nw = std::max(b->d.nw, first + count);
  return B2G_OK;
}

int b2g_set_worlds(b2g_batch* b, int32_t count, const int32_t* body_begin, const int32_t* shape_begin,
                   const b2g_world_params* params) {
  (void)body_begin;
  (void)shape_begin;
  if (!b || count < 1 || count > b->cap.worlds) return fail(B2G_ERR_INVALID, "world count");
  b->graphs_clear();
  b->graph_K = 0;
  b->filters_dirty = true;
  b->pairs_current = false;
  b->layout_dirty = true;
  b->d.nw = count;
  return b2g_set_world_params(b, 0, count, params);
}

int b2g_set_counts(b2g_batch* b, int32_t bodies, int32_t shapes, int32_t joints) {
  if (!b || bodies < 0 || shapes < 0 || joints < 0 || bodies > b->cap.bodies || shapes > b->cap.shapes ||
      joints > b->cap.joints)
    return fail(B2G_ERR_INVALID, "record counts outside the capacity");
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->flush_finish();
  Dev& d = b->d;
  d.nb = bodies;
  d.ns = shapes;
  d.nj = joints;
  for (int i = 0; i < shapes; ++i)  // records above the old count are uploaded next; keep them in range
    if (b->h_shape_body[i] >= bodies) b->h_shape_body[i] = 0;
  if ((int)b->h_shape_rec.size() > shapes) b->h_shape_rec.resize(shapes);
  b->h_joints.resize(joints);
  for (b2g_revolute& j : b->h_joints) {
    if (j.body_a >= bodies) j.body_a = 0;
    if (j.body_b >= bodies) j.body_b = 0;
  }
  // the contact list named the old indices: the caller uploads the survivors
  const int zero = 0;
  HIP_TRY(hipMemcpy(d.cnt + C_CONTACTS, &zero, sizeof(int), hipMemcpyHostToDevice));
  b->est_contacts = 0;
  b->graphs_clear();
  b->graph_K = 0;
  b->filters_dirty = true;
  b->pairs_current = false;
  b->layout_dirty = true;
  b->joints_dirty = true;
  b->tree_stale = true;
  b->tree_ns = -1;
  return bodies > 0 && shapes > 0 ? b->update_first_shapes() : B2G_OK;
}

int b2g_upload_bodies(b2g_batch* b, int32_t first, int32_t count, const b2g_body* in) {
  if (!b || first < 0 || count < 0 || first + count > b->cap.bodies) return fail(B2G_ERR_INVALID, "body range");
  if (count == 0) return B2G_OK;
  b->filters_dirty = true;  // body types feed the pair filter
  b->pairs_current = false;
  b->tree_stale = true;
  if (first + count > b->d.nb) b->layout_dirty = true;
  std::vector<float4> xf(count), pos(count), vel(count), prop(count), prop2(count), force(count);
  std::vector<int> type(count), world(count), sflag(count);
  std::vector<float> stime(count);
  for (int i = 0; i < count; ++i) {
    const b2g_body& s = in[i];
    if (s.world < 0 || s.world >= b->cap.worlds) return fail(B2G_ERR_INVALID, "body world index");
    // a static body is never asleep (b2Body::SetAwake returns for it)
    sflag[i] = s.flags & (s.type == B2G_STATIC ? B2G_BODY_NO_AUTOSLEEP : (B2G_BODY_ASLEEP | B2G_BODY_NO_AUTOSLEEP));
    stime[i] = s.sleep_time;
    if (sflag[i] & B2G_BODY_ASLEEP) b->asleep_uploaded = true;
    xf[i] = make_float4(s.xf[0], s.xf[1], s.xf[2], s.xf[3]);
    pos[i] = make_float4(s.c[0], s.c[1], s.a, 0.0f);
    vel[i] = make_float4(s.v[0], s.v[1], s.w, 0.0f);
    prop[i] = make_float4(s.inv_mass, s.inv_i, s.local_center[0], s.local_center[1]);
    prop2[i] = make_float4(s.mass, s.gravity_scale, s.linear_damping, s.angular_damping);
    force[i] = make_float4(s.force[0], s.force[1], s.torque, 0.0f);
    type[i] = s.type;
    world[i] = s.world;
    if (b->h_bworld[first + i] != s.world) b->layout_dirty = true;
    b->h_btype[first + i] = s.type;
    b->h_bworld[first + i] = s.world;
  }
  hipStream_t st = b->stream;
  Dev& d = b->d;
  HIP_TRY(hipMemcpyAsync(d.xf + first, xf.data(), count * sizeof(float4), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d.pos + first, pos.data(), count * sizeof(float4), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d.vel + first, vel.data(), count * sizeof(float4), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d.prop + first, prop.data(), count * sizeof(float4), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d.prop2 + first, prop2.data(), count * sizeof(float4), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d.force + first, force.data(), count * sizeof(float4), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d.btype + first, type.data(), count * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d.bworld + first, world.data(), count * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d.bsflag + first, sflag.data(), count * sizeof(int), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d.bstime + first, stime.data(), count * sizeof(float), hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (first + count > d.nb) {
    b->graphs_clear();
    b->graph_K = 0;
  b->filters_dirty = true;
  b->pairs_current = false;
  b->layout_dirty = true;
  }
  if (first + count > d.nb) {
    // extend the joint-partner CSR with empty rows
    const int old_nb = d.nb;
    d.nb = first + count;
    if (b->h_joints.empty()) {
      std::vector<int> z(d.nb - old_nb, 0);
      HIP_TRY(hipMemcpy(d.bj_off + old_nb + 1, z.data(), z.size() * sizeof(int), hipMemcpyHostToDevice));
    } else {
      b->joints_dirty = true;
    }
  }
  return B2G_OK;
}

int b2g_upload_shapes(b2g_batch* b, int32_t first, int32_t count, const b2g_shape* in) {
  if (!b || first < 0 || count < 0 || first + count > b->cap.shapes) return fail(B2G_ERR_INVALID, "shape range");
  b->tree_stale = true;
  b->graphs_clear();  // captured launch sizes no longer match
  b->graph_K = 0;
  b->filters_dirty = true;
  b->pairs_current = false;
  b->layout_dirty = true;
  if (count == 0) return B2G_OK;
  std::vector<int4> hdr(count), filt(count);
  std::vector<float4> geo(9 * (size_t)count), mat(count);
  for (int i = 0; i < count; ++i) {
    const b2g_shape& s = in[i];
    if (s.body < 0 || s.body >= b->d.nb) return fail(B2G_ERR_INVALID, "shape body index (upload bodies first)");
    if (s.is_sensor) return fail(B2G_ERR_UNSUPPORTED, "sensor fixtures are not supported on the device path");
    if (s.type < 0 || s.type > 2 || (s.type == B2G_POLYGON && (s.count < 3 || s.count > 8)))
      return fail(B2G_ERR_INVALID, "shape type / vertex count");
    shape_to_device(s, b->h_bworld[s.body], &hdr[i], &geo[9 * (size_t)i], &mat[i], &filt[i]);
    b->h_shape_body[first + i] = s.body;
    if ((int)b->h_shape_rec.size() < first + count) b->h_shape_rec.resize(first + count);
    b->h_shape_rec[first + i] = s;
  }
  hipStream_t st = b->stream;
  Dev& d = b->d;
  HIP_TRY(hipMemcpyAsync(d.shdr + first, hdr.data(), count * sizeof(int4), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d.geo + 9 * (size_t)first, geo.data(), 9 * (size_t)count * sizeof(float4),
                         hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d.smat + first, mat.data(), count * sizeof(float4), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(d.sfilt + first, filt.data(), count * sizeof(int4), hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));
  d.ns = std::max(d.ns, first + count);
  return b->update_first_shapes();
}

int b2g_upload_joints(b2g_batch* b, int32_t first, int32_t count, const b2g_revolute* in) {
  if (!b || first < 0 || count < 0 || first + count > b->cap.joints) return fail(B2G_ERR_INVALID, "joint range");
  b->graphs_clear();  // captured launch sizes no longer match
  b->graph_K = 0;
  b->filters_dirty = true;
  b->pairs_current = false;
  b->layout_dirty = true;
  if (count == 0) return B2G_OK;
  if ((int)b->h_joints.size() < first + count) b->h_joints.resize(first + count);
  std::vector<Joint> dj(count);
  for (int i = 0; i < count; ++i) {
    const b2g_revolute& r = in[i];
    if (r.body_a < 0 || r.body_a >= b->d.nb || r.body_b < 0 || r.body_b >= b->d.nb)
      return fail(B2G_ERR_INVALID, "joint body index");
    b->h_joints[first + i] = r;
    Joint& j = dj[i];
    std::memset(&j, 0, sizeof(j));
    j.a = r.body_a;
    j.b = r.body_b;
    j.flags = r.flags;
    j.world = b->h_bworld[r.body_a];
    j.la_x = r.local_anchor_a[0]; j.la_y = r.local_anchor_a[1];
    j.lb_x = r.local_anchor_b[0]; j.lb_y = r.local_anchor_b[1];
    j.ref = r.reference_angle; j.lower = r.lower_angle; j.upper = r.upper_angle;
    j.speed = r.motor_speed; j.max_torque = r.max_motor_torque;
    j.imp_x = r.impulse[0]; j.imp_y = r.impulse[1];
    j.motor = r.motor_impulse; j.lower_imp = r.lower_impulse; j.upper_imp = r.upper_impulse;
  }
  HIP_TRY(hipMemcpy(b->d.joints + first, dj.data(), count * sizeof(Joint), hipMemcpyHostToDevice));
  b->d.nj = std::max(b->d.nj, first + count);
  b->joints_dirty = true;
  return B2G_OK;
}

int b2g_upload_contacts(b2g_batch* b, int32_t count, const b2g_contact* in) {
  if (!b || count < 0 || count > b->cap.contacts) return fail(B2G_ERR_INVALID, "contact count");
  b->graphs_clear();  // captured launch sizes no longer match
  b->graph_K = 0;
  b->filters_dirty = true;
  b->pairs_current = false;
  b->layout_dirty = true;
  std::vector<int> idx(count);
  std::vector<uint64_t> keys(count);
  for (int i = 0; i < count; ++i) {
    if (in[i].shape_a < 0 || in[i].shape_a >= b->d.ns || in[i].shape_b < 0 || in[i].shape_b >= b->d.ns ||
        in[i].shape_a == in[i].shape_b || in[i].point_count < 0 || in[i].point_count > 2)
      return fail(B2G_ERR_INVALID, "contact record out of range (upload shapes first)");
    const uint32_t a = (uint32_t)in[i].shape_a, c = (uint32_t)in[i].shape_b;
    keys[i] = ((uint64_t)std::min(a, c) << 32) | std::max(a, c);
    idx[i] = i;
  }
  std::sort(idx.begin(), idx.end(), [&](int x, int y) { return keys[x] < keys[y]; });
  std::vector<uint64_t> k(count);
  std::vector<int2> pr(count);
  std::vector<float4> man(4 * (size_t)count), mat(count);
  std::vector<int> col(count, -1);
  for (int q = 0; q < count; ++q) {
    const b2g_contact& c = in[idx[q]];
    k[q] = keys[idx[q]];
    pr[q] = make_int2(c.shape_a, c.shape_b);
    record_to_manifold(c, &man[4 * (size_t)q]);
    mat[q] = make_float4(c.friction, c.restitution, c.restitution_threshold, c.tangent_speed);
  }
  Dev& d = b->d;
  if (count) {
    HIP_TRY(hipMemcpy(d.ckey, k.data(), count * sizeof(uint64_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.cpair, pr.data(), count * sizeof(int2), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.cman, man.data(), man.size() * sizeof(float4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.cmat, mat.data(), count * sizeof(float4), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.ccol, col.data(), count * sizeof(int), hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMemcpy(d.cnt + C_CONTACTS, &count, sizeof(int), hipMemcpyHostToDevice));
  b->est_contacts = count;
  return B2G_OK;
}

int b2g_batch_step(b2g_batch* b, int32_t vi, int32_t pi, uint32_t flags) {
  if (!b) return fail(B2G_ERR_INVALID, "null batch");
  if (vi < 0 || pi < 0) return fail(B2G_ERR_INVALID, "iteration counts");
  return b->step(vi, pi, flags);
}

int b2g_batch_wait(b2g_batch* b) {
  if (!b) return fail(B2G_ERR_INVALID, "null batch");
  int rc = b->read_counters();
  if (rc) return rc;
  return b->check_status();
}

int b2g_get_counts(b2g_batch* b, b2g_counts* out) {
  if (!b || !out) return fail(B2G_ERR_INVALID, "null argument");
  int rc = b->read_counters();
  if (rc) return rc;
  out->worlds = b->d.nw;
  out->bodies = b->d.nb;
  out->shapes = b->d.ns;
  out->joints = b->d.nj;
  out->contacts = b->pinned[C_CONTACTS];
  out->touching = b->pinned[C_CONS];
  out->colours = b->pinned[C_NCOLOURS];
  out->islands = b->last_islands;
  out->pos_iters_run = b->last_pos_iters;
  out->status = b->pinned[C_STATUS];
  out->serial = b->pinned[C_SERIAL];
  out->graph_steps = (int32_t)std::min<long long>(b->graph_steps, INT32_MAX);
  return B2G_OK;
}

int b2g_set_option(b2g_batch* b, int32_t option, int32_t value) {
  if (!b) return fail(B2G_ERR_INVALID, "null batch");
  switch (option) {
    case B2G_OPT_SOLVER:
      if (value < B2G_SOLVER_AUTO || value > B2G_SOLVER_CELLS) return fail(B2G_ERR_INVALID, "solver mode");
      if (b->solver_mode != value) b->graphs_clear();  // the captured steps are the other solver's
      b->solver_mode = value;
      break;
    case B2G_OPT_WORLD_LDS:
      if (value < -1 || value > 1) return fail(B2G_ERR_INVALID, "world LDS staging: -1 (autotune), 0 or 1");
      b->world_lds = value;
      b->tuned_block = 0;
      b->tune_step = 0;
      break;
    case B2G_OPT_WORLD_BLOCK:
      if (value != 0 && (value < 64 || value > 1024 || (value & 63))) {
        return fail(B2G_ERR_INVALID, "world block: 0 (autotune) or a multiple of 64 in [64, 1024]");
      }
      b->world_block = value;
      b->tuned_block = 0;
      b->tune_step = 0;
      break;
    case B2G_OPT_KEEP_IDLE:
      if (value < 0 || value > 1) return fail(B2G_ERR_INVALID, "keep idle: 0 or 1");
      b->keep_idle = value != 0;
      break;
    case B2G_OPT_PAIR_SEARCH:
      if (value < 0 || value > 1) return fail(B2G_ERR_INVALID, "pair search: 0 (grid where it applies) or 1 (LBVH)");
      if (value != b->pair_search) {
        b->graphs_clear();  // the captured steps hold the other search
        b->tree_stale = true;
      }
      b->pair_search = value;
      break;
    default:
      return fail(B2G_ERR_INVALID, "unknown option");
  }
  b->graphs_clear();
  return B2G_OK;
}

int b2g_get_option(b2g_batch* b, int32_t option, int32_t* value) {
  if (!b || !value) return fail(B2G_ERR_INVALID, "null argument");
  if (b->layout_dirty) {
    int rc = b->update_world_layout();
    if (rc) return rc;
  }
  switch (option) {
    case B2G_OPT_SOLVER: *value = b->solver_mode; break;
    case B2G_OPT_WORLD_BLOCK: *value = b->world_block; break;
    case B2G_OPT_TUNED_WORLD_BLOCK: *value = b->world_block > 0 ? b->world_block : b->tuned_block; break;
    case B2G_OPT_LAST_SOLVER: *value = b->last_path; break;
    case B2G_OPT_PER_WORLD_ELIGIBLE: *value = b->per_world_ok ? 1 : 0; break;
    case B2G_OPT_WORLD_LDS: *value = b->world_lds; break;
    case B2G_OPT_TUNED_WORLD_LDS: *value = b->world_lds_now() ? 1 : 0; break;
    case B2G_OPT_KEEP_IDLE: *value = b->keep_idle ? 1 : 0; break;
    case B2G_OPT_PAIR_SEARCH: *value = b->pair_search; break;
    case B2G_OPT_LAST_PAIR_SEARCH: *value = b->last_pair_search; break;
    default: return fail(B2G_ERR_INVALID, "unknown option");
  }
  return B2G_OK;
}

int b2g_time_kernel(b2g_batch* b, int32_t kernel) {
  if (!b) return fail(B2G_ERR_INVALID, "null batch");
  b->tcollect();
  b->timed_kernel = kernel;
  b->t_ms = 0.0;
  b->t_launches = 0;
  b->t_total_items = 0;
  return B2G_OK;
}

int b2g_kernel_time(b2g_batch* b, double* total_ms, int64_t* launches, int64_t* items) {
  if (!b) return fail(B2G_ERR_INVALID, "null batch");
  b->tcollect();
  if (total_ms) *total_ms = b->t_ms;
  if (launches) *launches = b->t_launches;
  if (items) *items = b->t_total_items;
  return B2G_OK;
}

int b2g_get_profile(b2g_batch* b, b2g_profile* out) {
  if (!b || !out) return fail(B2G_ERR_INVALID, "null argument");
  *out = b->prof;
  return B2G_OK;
}

int b2g_download_bodies(b2g_batch* b, int32_t first, int32_t count, b2g_body* out) {
  if (!b || first < 0 || count < 0 || first + count > b->d.nb) return fail(B2G_ERR_INVALID, "body range");
  if (count == 0) return B2G_OK;
  std::vector<float4> xf(count), pos(count), vel(count), prop(count), prop2(count), force(count);
  Dev& d = b->d;
  hipStream_t st = b->stream;
  HIP_TRY(hipMemcpyAsync(xf.data(), d.xf + first, count * sizeof(float4), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(pos.data(), d.pos + first, count * sizeof(float4), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(vel.data(), d.vel + first, count * sizeof(float4), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(prop.data(), d.prop + first, count * sizeof(float4), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(prop2.data(), d.prop2 + first, count * sizeof(float4), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(force.data(), d.force + first, count * sizeof(float4), hipMemcpyDeviceToHost, st));
  std::vector<int> sflag(count);
  std::vector<float> stime(count);
  HIP_TRY(hipMemcpyAsync(sflag.data(), d.bsflag + first, count * sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(stime.data(), d.bstime + first, count * sizeof(float), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  for (int i = 0; i < count; ++i) {
    b2g_body& o = out[i];
    std::memset(&o, 0, sizeof(o));
    o.type = b->h_btype[first + i];
    o.world = b->h_bworld[first + i];
    o.xf[0] = xf[i].x; o.xf[1] = xf[i].y; o.xf[2] = xf[i].z; o.xf[3] = xf[i].w;
    o.c[0] = pos[i].x; o.c[1] = pos[i].y; o.a = pos[i].z;
    o.v[0] = vel[i].x; o.v[1] = vel[i].y; o.w = vel[i].z;
    o.inv_mass = prop[i].x; o.inv_i = prop[i].y;
    o.local_center[0] = prop[i].z; o.local_center[1] = prop[i].w;
    o.mass = prop2[i].x; o.gravity_scale = prop2[i].y;
    o.linear_damping = prop2[i].z; o.angular_damping = prop2[i].w;
    o.force[0] = force[i].x; o.force[1] = force[i].y; o.torque = force[i].z;
    o.flags = sflag[i];
    o.sleep_time = stime[i];
  }
  return B2G_OK;
}

int b2g_download_contacts(b2g_batch* b, int32_t first, int32_t count, b2g_contact* out) {
  if (!b) return fail(B2G_ERR_INVALID, "null batch");
  int rc = b->read_counters();
  if (rc) return rc;
  const int n = b->pinned[C_CONTACTS];
  if (!out) return n;  // count query
  if (first < 0 || count < 0 || first + count > n) return fail(B2G_ERR_INVALID, "contact range");
  if (count == 0) return B2G_OK;
  std::vector<int2> pr(count);
  std::vector<float4> man(4 * (size_t)count), mat(count);
  Dev& d = b->d;
  HIP_TRY(hipMemcpy(pr.data(), d.cpair + first, count * sizeof(int2), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(man.data(), d.cman + 4 * (size_t)first, man.size() * sizeof(float4), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(mat.data(), d.cmat + first, count * sizeof(float4), hipMemcpyDeviceToHost));
  for (int i = 0; i < count; ++i) {
    b2g_contact& o = out[i];
    std::memset(&o, 0, sizeof(o));
    o.shape_a = pr[i].x;
    o.shape_b = pr[i].y;
    manifold_to_record(&man[4 * (size_t)i], &o);
    o.friction = mat[i].x;
    o.restitution = mat[i].y;
    o.restitution_threshold = mat[i].z;
    o.tangent_speed = mat[i].w;
  }
  return B2G_OK;
}

int b2g_download_solver_points(b2g_batch* b, int32_t first, int32_t count, int32_t* out) {
  if (!b || (count && !out)) return fail(B2G_ERR_INVALID, "null argument");
  int rc = b->read_counters();
  if (rc) return rc;
  if (first < 0 || count < 0 || first + count > b->pinned[C_CONTACTS]) return fail(B2G_ERR_INVALID, "contact range");
  if (count == 0) return B2G_OK;
  // the fourth float4 of each manifold (ids, packed type / counts)
  std::vector<float4> m3(count);
  HIP_TRY(hipMemcpy2D(m3.data(), sizeof(float4), b->d.cman + 4 * (size_t)first + 3, 4 * sizeof(float4),
                      sizeof(float4), count, hipMemcpyDeviceToHost));
  for (int i = 0; i < count; ++i) out[i] = (__builtin_bit_cast(int, m3[i].z) >> 16) & 0xff;
  return B2G_OK;
}

int b2g_download_contact_colours(b2g_batch* b, int32_t first, int32_t count, int32_t* out) {
  if (!b || (count && !out)) return fail(B2G_ERR_INVALID, "null argument");
  int rc = b->read_counters();
  if (rc) return rc;
  if (first < 0 || count < 0 || first + count > b->pinned[C_CONTACTS]) return fail(B2G_ERR_INVALID, "contact range");
  if (count == 0) return B2G_OK;
  // colours are kept per contact in the (post-broadphase) contact buffer
  HIP_TRY(hipMemcpy(out, b->d.ccol + first, count * sizeof(int), hipMemcpyDeviceToHost));
  return B2G_OK;
}

}  // extern "C"

int b2g_batch::query_scratch(size_t bytes) {
  if (bytes <= qbuf_bytes) return B2G_OK;
  if (qbuf) (void)hipFree(qbuf);
  qbuf = nullptr;
  qbuf_bytes = 0;
  const size_t want = std::max<size_t>(bytes, 1 << 16);
  hipError_t e = hipMalloc(&qbuf, want);
  if (e != hipSuccess) return fail(B2G_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
  qbuf_bytes = want;
  return B2G_OK;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// shared by the AABB and ray queries: inputs (n float4 + n worlds) in, up to
// `cap` results of `rec` bytes per query + counts out
template <typename Launch>
static int run_queries(b2g_batch* b, int32_t n, const int32_t* worlds, const float* in4, int32_t cap, size_t rec,
                       void* out, int32_t* counts, Launch launch) {
  if (!b || n < 0 || cap < 0 || (n && (!worlds || !in4 || !counts)) || (n && cap && !out))
    return fail(B2G_ERR_INVALID, "null argument");
  for (int i = 0; i < n; ++i)
    if (worlds[i] < 0 || worlds[i] >= b->d.nw) return fail(B2G_ERR_INVALID, "query world index");
  if (n == 0) return B2G_OK;
  if (b->tree_stale) {  // bodies or shapes changed since the last broadphase
    int rc = b->build_tree(false);
    if (rc) return rc;
  }
  const size_t o_in = 0, o_w = align256(o_in + 16 * (size_t)n), o_out = align256(o_w + 4 * (size_t)n),
               o_cnt = align256(o_out + rec * (size_t)n * (size_t)cap), total = o_cnt + 4 * (size_t)n;
  int rc = b->query_scratch(total);
  if (rc) return rc;
  char* base = (char*)b->qbuf;
  HIP_TRY(hipMemcpyAsync(base + o_in, in4, 16 * (size_t)n, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(hipMemcpyAsync(base + o_w, worlds, 4 * (size_t)n, hipMemcpyHostToDevice, b->stream));
  launch((const float4*)(base + o_in), (const int*)(base + o_w), (void*)(base + o_out), (int*)(base + o_cnt));
  HIP_TRY(hipMemcpyAsync(counts, base + o_cnt, 4 * (size_t)n, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  if (cap) {
    // copy back each query's slots that hold results
    std::vector<char> tmp(rec * (size_t)n * (size_t)cap);
    HIP_TRY(hipMemcpy(tmp.data(), base + o_out, tmp.size(), hipMemcpyDeviceToHost));
    std::memcpy(out, tmp.data(), tmp.size());
  }
  rc = b->read_counters();
  if (rc) return rc;
  if (b->pinned[C_STATUS] & S_STACK_OVERFLOW) return fail(B2G_ERR_CAPACITY, "BVH traversal stack overflow");
  if (b->pinned[C_STATUS] & S_TREE_CORRUPT) return fail(B2G_ERR_INVALID, "BVH inconsistent (sort/build)");
  return B2G_OK;
}

extern "C" {

int b2g_query_aabbs(b2g_batch* b, int32_t n, const int32_t* worlds, const float* aabbs, int32_t capacity,
                    int32_t* out_shapes, int32_t* out_counts) {
  return run_queries(b, n, worlds, aabbs, capacity, sizeof(int32_t), out_shapes, out_counts,
                     [&](const float4* in, const int* w, void* out, int* cnt) {
                       const int g = grid_for((long long)n * kBlock / kQueryBlock);
                       hipLaunchKernelGGL(k_query_aabbs, dim3(g), dim3(kQueryBlock), 0, b->stream, b->d, n, in, w,
                                          capacity, (int*)out, cnt);
                     });
}

int b2g_raycast(b2g_batch* b, int32_t n, const int32_t* worlds, const float* segments, int32_t capacity,
                b2g_ray_hit* out_hits, int32_t* out_counts) {
  static_assert(sizeof(b2g_ray_hit) == sizeof(float4), "ray hit record is one float4");
  return run_queries(b, n, worlds, segments, capacity, sizeof(b2g_ray_hit), out_hits, out_counts,
                     [&](const float4* in, const int* w, void* out, int* cnt) {
                       const int g = grid_for((long long)n * kBlock / kQueryBlock);
                       hipLaunchKernelGGL(k_raycast, dim3(g), dim3(kQueryBlock), 0, b->stream, b->d, n, in, w,
                                          capacity, (float4*)out, cnt);
                     });
}

int b2g_download_events(b2g_batch* b, int32_t capacity, b2g_contact_event* out, int32_t* count) {
  if (!b || !count || (capacity && !out) || capacity < 0) return fail(B2G_ERR_INVALID, "null argument");
  int rc = b->read_counters();
  if (rc) return rc;
  const int n = std::min(b->pinned[C_EVENTS], b->d.ecap);
  *count = n;
  if (b->pinned[C_STATUS] & S_EVENT_OVERFLOW) return fail(B2G_ERR_CAPACITY, "contact event buffer overflow");
  if (capacity == 0 || n == 0) return B2G_OK;
  std::vector<float4> raw((size_t)kEventWords * n);
  HIP_TRY(hipMemcpy(raw.data(), b->d.events, raw.size() * sizeof(float4), hipMemcpyDeviceToHost));
  std::vector<b2g_contact_event> ev(n);
  for (int i = 0; i < n; ++i) {
    const float4* e = &raw[(size_t)kEventWords * i];
    b2g_contact_event& o = ev[i];
    std::memset(&o, 0, sizeof(o));
    o.kind = __builtin_bit_cast(int, e[0].z);
    // the step's stored point count travels in the packed manifold word
    o.solver_points = o.kind == B2G_EVENT_DESTROYED ? (__builtin_bit_cast(int, e[4].z) >> 16) & 0xff : 0;
    o.contact.shape_a = __builtin_bit_cast(int, e[0].x);
    o.contact.shape_b = __builtin_bit_cast(int, e[0].y);
    manifold_to_record(e + 1, &o.contact);
    o.contact.friction = e[5].x;
    o.contact.restitution = e[5].y;
    o.contact.restitution_threshold = e[5].z;
    o.contact.tangent_speed = e[5].w;
  }
  // Collide's events first, then the broadphase's; by shape pair inside each
  auto key = [](const b2g_contact_event& e) {
    const uint64_t a = (uint32_t)e.contact.shape_a, c = (uint32_t)e.contact.shape_b;
    return (std::min(a, c) << 32) | std::max(a, c);
  };
  std::sort(ev.begin(), ev.end(), [&](const b2g_contact_event& x, const b2g_contact_event& y) {
    const int px = x.kind == 3, py = y.kind == 3;
    return px != py ? px < py : key(x) < key(y);
  });
  const int m = std::min(n, (int)capacity);
  std::memcpy(out, ev.data(), m * sizeof(b2g_contact_event));
  *count = n;
  return B2G_OK;
}

int b2g_upload_contact_colours(b2g_batch* b, int32_t count, const int32_t* colours) {
  if (!b || (count && !colours)) return fail(B2G_ERR_INVALID, "null argument");
  int rc = b->read_counters();
  if (rc) return rc;
  if (count != b->pinned[C_CONTACTS]) return fail(B2G_ERR_INVALID, "colour count must equal the contact count");
  for (int i = 0; i < count; ++i)
    if (colours[i] < -1 || colours[i] > kSerialColour) return fail(B2G_ERR_INVALID, "colour out of range");
  if (count) HIP_TRY(hipMemcpy(b->d.ccol, colours, count * sizeof(int), hipMemcpyHostToDevice));
  b->graphs_clear();
  b->graph_K = 0;
  return B2G_OK;
}

// ---- snapshots ---------------------------------------------------------------
// File: "B2GSNAP1", then int32 version (2), capacity (5), world count, body,
// shape, joint and contact counts, then the world params, the per-world
// colouring state (fresh colour counts, then recolour flags), the body, shape,
// joint and contact records and the contact colours, all little-endian as
// laid out in include/b2g.h.
static const char kSnapMagic[8] = {'B', '2', 'G', 'S', 'N', 'A', 'P', '1'};

int b2g_snapshot_save(b2g_batch* b, const char* path) {
  if (!b || !path) return fail(B2G_ERR_INVALID, "null argument");
  HIP_TRY(hipStreamSynchronize(b->stream));
  int rc = b->read_counters();
  if (rc) return rc;
  const int nw = b->d.nw, nb = b->d.nb, ns = b->d.ns, nj = b->d.nj, nc = b->pinned[C_CONTACTS];
  std::vector<b2g_body> bodies(nb);
  std::vector<b2g_revolute> joints(nj);
  std::vector<b2g_contact> contacts(nc);
  std::vector<int32_t> colours(nc);
  if (nb && (rc = b2g_download_bodies(b, 0, nb, bodies.data()))) return rc;
  if (nj && (rc = b2g_download_joints(b, 0, nj, joints.data()))) return rc;
  if (nc && (rc = b2g_download_contacts(b, 0, nc, contacts.data()))) return rc;
  if (nc && (rc = b2g_download_contact_colours(b, 0, nc, colours.data()))) return rc;
  std::vector<int32_t> wstate(2 * (size_t)nw);  // per world: fresh colour count, recolour flag
  HIP_TRY(hipMemcpy(wstate.data(), b->d.wfresh, nw * sizeof(int), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(wstate.data() + nw, b->d.wrecolour, nw * sizeof(int), hipMemcpyDeviceToHost));
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(B2G_ERR_INVALID, std::string("cannot write ") + path);
  const int32_t head[11] = {2, b->cap.worlds, b->cap.bodies, b->cap.shapes, b->cap.joints, b->cap.contacts,
                            nw, nb, ns, nj, nc};
  bool ok = std::fwrite(kSnapMagic, 1, 8, f) == 8 && std::fwrite(head, sizeof(head), 1, f) == 1;
  ok = ok && (nw == 0 || std::fwrite(b->h_params.data(), sizeof(b2g_world_params), nw, f) == (size_t)nw);
  ok = ok && (nw == 0 || std::fwrite(wstate.data(), sizeof(int32_t), 2 * (size_t)nw, f) == 2 * (size_t)nw);
  ok = ok && (nb == 0 || std::fwrite(bodies.data(), sizeof(b2g_body), nb, f) == (size_t)nb);
  ok = ok && (ns == 0 || std::fwrite(b->h_shape_rec.data(), sizeof(b2g_shape), ns, f) == (size_t)ns);
  ok = ok && (nj == 0 || std::fwrite(joints.data(), sizeof(b2g_revolute), nj, f) == (size_t)nj);
  ok = ok && (nc == 0 || std::fwrite(contacts.data(), sizeof(b2g_contact), nc, f) == (size_t)nc);
  ok = ok && (nc == 0 || std::fwrite(colours.data(), sizeof(int32_t), nc, f) == (size_t)nc);
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(B2G_ERR_INVALID, std::string("short write to ") + path);
  return B2G_OK;
}

int b2g_snapshot_load(b2g_batch** out, const char* path) {
  if (!out || !path) return fail(B2G_ERR_INVALID, "null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(B2G_ERR_INVALID, std::string("cannot read ") + path);
  char magic[8];
  int32_t head[11];
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kSnapMagic, 8) == 0 &&
            std::fread(head, sizeof(head), 1, f) == 1 && head[0] == 2;
  const int nw = ok ? head[6] : 0, nb = ok ? head[7] : 0, ns = ok ? head[8] : 0, nj = ok ? head[9] : 0,
            nc = ok ? head[10] : 0;
  b2g_capacity cap{head[1], head[2], head[3], head[4], head[5]};
  ok = ok && nw >= 1 && nb >= 0 && ns >= 0 && nj >= 0 && nc >= 0 && nw <= cap.worlds && nb <= cap.bodies &&
       ns <= cap.shapes && nj <= cap.joints && nc <= cap.contacts;
  std::vector<b2g_world_params> params(ok ? nw : 0);
  std::vector<int32_t> wstate(ok ? 2 * (size_t)nw : 0);
  std::vector<b2g_body> bodies(ok ? nb : 0);
  std::vector<b2g_shape> shapes(ok ? ns : 0);
  std::vector<b2g_revolute> joints(ok ? nj : 0);
  std::vector<b2g_contact> contacts(ok ? nc : 0);
  std::vector<int32_t> colours(ok ? nc : 0);
  ok = ok && std::fread(params.data(), sizeof(b2g_world_params), nw, f) == (size_t)nw;
  ok = ok && std::fread(wstate.data(), sizeof(int32_t), 2 * (size_t)nw, f) == 2 * (size_t)nw;
  ok = ok && (nb == 0 || std::fread(bodies.data(), sizeof(b2g_body), nb, f) == (size_t)nb);
  ok = ok && (ns == 0 || std::fread(shapes.data(), sizeof(b2g_shape), ns, f) == (size_t)ns);
  ok = ok && (nj == 0 || std::fread(joints.data(), sizeof(b2g_revolute), nj, f) == (size_t)nj);
  ok = ok && (nc == 0 || std::fread(contacts.data(), sizeof(b2g_contact), nc, f) == (size_t)nc);
  ok = ok && (nc == 0 || std::fread(colours.data(), sizeof(int32_t), nc, f) == (size_t)nc);
  std::fclose(f);
  if (!ok) return fail(B2G_ERR_INVALID, std::string("not a readable snapshot: ") + path);
  b2g_batch* b = nullptr;
  int rc = b2g_batch_create(&b, &cap);
  if (rc) return rc;
  if (!rc) rc = b2g_set_worlds(b, nw, nullptr, nullptr, params.data());
  if (!rc && nb) rc = b2g_upload_bodies(b, 0, nb, bodies.data());
  if (!rc && ns) rc = b2g_upload_shapes(b, 0, ns, shapes.data());
  if (!rc && nj) rc = b2g_upload_joints(b, 0, nj, joints.data());
  if (!rc) rc = b2g_upload_contacts(b, nc, contacts.data());
  if (!rc) rc = b2g_upload_contact_colours(b, nc, colours.data());
  if (!rc && hipMemcpy(b->d.wfresh, wstate.data(), nw * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(B2G_ERR_HIP, "restoring the colouring state");
  if (!rc && hipMemcpy(b->d.wrecolour, wstate.data() + nw, nw * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(B2G_ERR_HIP, "restoring the colouring state");
  if (rc) {
    b2g_batch_destroy(b);
    return rc;
  }
  *out = b;
  return B2G_OK;
}

int b2g_download_joints(b2g_batch* b, int32_t first, int32_t count, b2g_revolute* out) {
  if (!b || first < 0 || count < 0 || first + count > b->d.nj) return fail(B2G_ERR_INVALID, "joint range");
  if (count == 0) return B2G_OK;
  std::vector<Joint> dj(count);
  HIP_TRY(hipMemcpy(dj.data(), b->d.joints + first, count * sizeof(Joint), hipMemcpyDeviceToHost));
  for (int i = 0; i < count; ++i) {
    b2g_revolute r = b->h_joints[first + i];
    r.impulse[0] = dj[i].imp_x;
    r.impulse[1] = dj[i].imp_y;
    r.motor_impulse = dj[i].motor;
    r.lower_impulse = dj[i].lower_imp;
    r.upper_impulse = dj[i].upper_imp;
    out[i] = r;
  }
  return B2G_OK;
}

int b2g_collide_pairs(int32_t n, const b2g_shape* sa, const float* xfa, const b2g_shape* sb, const float* xfb,
                      b2g_contact* out) {
  if (n < 0) return fail(B2G_ERR_INVALID, "negative count");
  if (n == 0) return B2G_OK;
  std::vector<int4> ha(n), hb(n), fa(n), fb(n);
  std::vector<float4> ga(9 * (size_t)n), gb(9 * (size_t)n), ma(n), mb(n);
  for (int i = 0; i < n; ++i) {
    shape_to_device(sa[i], 0, &ha[i], &ga[9 * (size_t)i], &ma[i], &fa[i]);
    shape_to_device(sb[i], 0, &hb[i], &gb[9 * (size_t)i], &mb[i], &fb[i]);
  }
  int4 *dha = nullptr, *dhb = nullptr;
  float4 *dga = nullptr, *dgb = nullptr, *dxa = nullptr, *dxb = nullptr, *dout = nullptr;
  int* dok = nullptr;
  int rc = B2G_OK;
  if ((rc = dalloc(&dha, n)) || (rc = dalloc(&dhb, n)) || (rc = dalloc(&dga, 9 * (size_t)n)) ||
      (rc = dalloc(&dgb, 9 * (size_t)n)) || (rc = dalloc(&dxa, n)) || (rc = dalloc(&dxb, n)) ||
      (rc = dalloc(&dout, 4 * (size_t)n)) || (rc = dalloc(&dok, n))) {
    // fall through to cleanup
  } else {
    std::vector<float4> out4(4 * (size_t)n);
    std::vector<int> ok(n);
    hipError_t e = hipMemcpy(dha, ha.data(), n * sizeof(int4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dhb, hb.data(), n * sizeof(int4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dga, ga.data(), ga.size() * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dgb, gb.data(), gb.size() * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dxa, xfa, n * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dxb, xfb, n * sizeof(float4), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_collide_pairs, dim3(grid_for(n)), dim3(kBlock), 0, 0, n, dha, dga, dxa, dhb, dgb, dxb,
                         dout, dok);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out4.data(), dout, out4.size() * sizeof(float4), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(ok.data(), dok, n * sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      rc = fail(B2G_ERR_HIP, std::string("b2g_collide_pairs: ") + hipGetErrorString(e));
    } else {
      for (int i = 0; i < n; ++i) {
        b2g_contact& o = out[i];
        std::memset(&o, 0, sizeof(o));
        o.shape_a = sa[i].body;
        o.shape_b = sb[i].body;
        manifold_to_record(&out4[4 * (size_t)i], &o);
        if (!ok[i]) o.point_count = -1;
      }
    }
  }
  void* ptrs[] = {dha, dhb, dga, dgb, dxa, dxb, dout, dok};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// b2g_collide_pairs without allocations: the same narrowphase kernel, with
// every staging buffer carved out of one file-scope device array. Each GPU
// holds its own copy of the array until the process ends; calls that overlap
// on one GPU share it and race.
// ---------------------------------------------------------------------------
namespace {

constexpr int kNoAllocMaxPairs = 4096;
constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// the array's device address per GPU: the array's name is a host-side
// stand-in, so host code only ever passes these addresses on
void* g_scratch_ptr[kNoAllocMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_for_current_device(unsigned char** out) {
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(B2G_ERR_INVALID, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(B2G_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  return B2G_OK;
}

// caller records -> the narrowphase's header and geometry arrays
__global__ void __launch_bounds__(kBlock) k_pair_shapes(int n, const b2g_shape* __restrict__ sa,
                                                        const b2g_shape* __restrict__ sb, int4* __restrict__ ha,
                                                        float4* __restrict__ ga, int4* __restrict__ hb,
                                                        float4* __restrict__ gb) {
  B2G_FOR(i, 0, n) {
    float4 mat;
    int4 filt;
    shape_to_device(sa[i], 0, &ha[i], ga + 9 * (size_t)i, &mat, &filt);
    shape_to_device(sb[i], 0, &hb[i], gb + 9 * (size_t)i, &mat, &filt);
  }
}

// the narrowphase's packed manifolds -> caller contact records, as
// b2g_collide_pairs fills them on the host
__global__ void __launch_bounds__(kBlock) k_pair_records(int n, const b2g_shape* __restrict__ sa,
                                                         const b2g_shape* __restrict__ sb,
                                                         const float4* __restrict__ man, const int* __restrict__ ok,
                                                         b2g_contact* __restrict__ out) {
  B2G_FOR(i, 0, n) {
    b2g_contact o{};
    o.shape_a = sa[i].body;
    o.shape_b = sb[i].body;
    manifold_to_record(man + 4 * (size_t)i, &o);
    if (!ok[i]) o.point_count = -1;
    out[i] = o;
  }
}

}  // namespace

extern "C" {

int b2g_collide_pairs_noalloc(int32_t n, const b2g_shape* sa, const float* xfa, const b2g_shape* sb,
                              const float* xfb, b2g_contact* out) {
  if (n < 0) return fail(B2G_ERR_INVALID, "negative count");
  if (n > kNoAllocMaxPairs)
    return fail(B2G_ERR_CAPACITY, "b2g_collide_pairs_noalloc takes at most " + std::to_string(kNoAllocMaxPairs) +
                                      " pairs per call (got " + std::to_string(n) + "); use b2g_collide_pairs");
  if (n == 0) return B2G_OK;
  if (!sa || !xfa || !sb || !xfb || !out) return fail(B2G_ERR_INVALID, "null argument");
  unsigned char* base = nullptr;
  int rc = scratch_for_current_device(&base);
  if (rc) return rc;
  const size_t N = (size_t)n;
  const size_t o_sa = 0, o_sb = align256(o_sa + N * sizeof(b2g_shape)), o_xa = align256(o_sb + N * sizeof(b2g_shape)),
               o_xb = align256(o_xa + N * sizeof(float4)), o_ha = align256(o_xb + N * sizeof(float4)),
               o_hb = align256(o_ha + N * sizeof(int4)), o_ga = align256(o_hb + N * sizeof(int4)),
               o_gb = align256(o_ga + 9 * N * sizeof(float4)), o_man = align256(o_gb + 9 * N * sizeof(float4)),
               o_ok = align256(o_man + 4 * N * sizeof(float4)), o_out = align256(o_ok + N * sizeof(int)),
               total = o_out + N * sizeof(b2g_contact);
  if (total > kNoAllocScratchBytes) return fail(B2G_ERR_CAPACITY, "scratch layout exceeds the scratch array");
  b2g_shape* dsa = reinterpret_cast<b2g_shape*>(base + o_sa);
  b2g_shape* dsb = reinterpret_cast<b2g_shape*>(base + o_sb);
  float4* dxa = reinterpret_cast<float4*>(base + o_xa);
  float4* dxb = reinterpret_cast<float4*>(base + o_xb);
  int4* dha = reinterpret_cast<int4*>(base + o_ha);
  int4* dhb = reinterpret_cast<int4*>(base + o_hb);
  float4* dga = reinterpret_cast<float4*>(base + o_ga);
  float4* dgb = reinterpret_cast<float4*>(base + o_gb);
  float4* dman = reinterpret_cast<float4*>(base + o_man);
  int* dok = reinterpret_cast<int*>(base + o_ok);
  b2g_contact* dout = reinterpret_cast<b2g_contact*>(base + o_out);
  HIP_TRY(hipMemcpy(dsa, sa, N * sizeof(b2g_shape), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dsb, sb, N * sizeof(b2g_shape), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dxa, xfa, N * sizeof(float4), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dxb, xfb, N * sizeof(float4), hipMemcpyHostToDevice));
  const dim3 grid(grid_for(n)), block(kBlock);
  hipLaunchKernelGGL(k_pair_shapes, grid, block, 0, 0, n, dsa, dsb, dha, dga, dhb, dgb);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_collide_pairs, grid, block, 0, 0, n, dha, dga, dxa, dhb, dgb, dxb, dman, dok);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_pair_records, grid, block, 0, 0, n, dsa, dsb, dman, dok, dout);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(out, dout, N * sizeof(b2g_contact), hipMemcpyDeviceToHost));
  return B2G_OK;
}

int b2g_find_pairs(int32_t n, const float* aabbs, const int32_t* shape_body, const int32_t* shape_type,
                   const int32_t* filter, int32_t body_count, const int32_t* body_type, int64_t max_pairs,
                   uint64_t* out_keys, int64_t* out_count) {
  if (n < 0 || body_count < 0 || !out_count) return fail(B2G_ERR_INVALID, "bad arguments");
  b2g_capacity cap{};
  cap.worlds = 1;
  cap.bodies = std::max(body_count, 1);
  cap.shapes = std::max(n, 1);
  cap.joints = 1;
  cap.contacts = (int)std::min<int64_t>(std::max<int64_t>(max_pairs, 16), (1 << 24) - 1);
  b2g_batch* b = nullptr;
  int rc = b2g_batch_create(&b, &cap);
  if (rc) return rc;
  b2g_world_params wp{{0.0f, 0.0f}, 1.0f / 60.0f, 0.0f, 1, 1};
  rc = b2g_set_worlds(b, 1, nullptr, nullptr, &wp);
  std::vector<b2g_body> bodies(body_count);
  for (int i = 0; i < body_count && !rc; ++i) {
    std::memset(&bodies[i], 0, sizeof(b2g_body));
    bodies[i].type = body_type[i];
    bodies[i].xf[3] = 1.0f;
  }
  if (!rc) rc = b2g_upload_bodies(b, 0, body_count, bodies.data());
  std::vector<b2g_shape> shapes(n);
  for (int i = 0; i < n && !rc; ++i) {
    b2g_shape& s = shapes[i];
    std::memset(&s, 0, sizeof(s));
    s.type = shape_type[i];
    s.body = shape_body[i];
    s.count = s.type == B2G_POLYGON ? 3 : 0;
    s.category = filter ? filter[3 * i] : 1;
    s.mask = filter ? filter[3 * i + 1] : 0xffff;
    s.group = filter ? filter[3 * i + 2] : 0;
  }
  if (!rc) rc = b2g_upload_shapes(b, 0, n, shapes.data());
  if (!rc && n) {
    std::vector<float4> ab(n);
    for (int i = 0; i < n; ++i) ab[i] = make_float4(aabbs[4 * i], aabbs[4 * i + 1], aabbs[4 * i + 2], aabbs[4 * i + 3]);
    hipError_t e = hipMemcpy(b->d.aabb, ab.data(), n * sizeof(float4), hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = fail(B2G_ERR_HIP, hipGetErrorString(e));
  }
  if (!rc) rc = b->broadphase(true);
  if (!rc) rc = b->read_counters();
  if (!rc) {
    const int64_t total = b->pinned[C_PAIRS];
    *out_count = total;
    const int64_t m = std::min<int64_t>(std::min<int64_t>(total, max_pairs), b->d.ccap);
    if (m > 0 && out_keys) {
      hipError_t e = hipMemcpy(out_keys, b->d.ckey, m * sizeof(uint64_t), hipMemcpyDeviceToHost);
      if (e != hipSuccess) rc = fail(B2G_ERR_HIP, hipGetErrorString(e));
    }
  }
  b2g_batch_destroy(b);
  return rc;
}

}  // extern "C"

