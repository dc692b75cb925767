// b2g_grid.h -- the broadphase's pair search on a hashed uniform grid
// instead of the LBVH, for any batch whose worlds have at most
// kGridStaticMax static shapes each.
//
// The pair set is the same: every pair of shapes whose AABBs overlap
// (b2TestOverlap, inclusive) and that pair_allowed admits, each found once
// under the LBVH's ownership rule (a shape reports partners with a larger
// index and static partners of any index; a static shape does not query),
// and written by the same slot and compaction scheme, so the pair list that
// the merge receives is the one the LBVH gives, bit for bit.
//
// Why it holds: the cell size G exceeds twice the largest distance from a
// non-static shape's body origin to any point of the shape, so a non-static
// shape's AABB is narrower than G on both axes. Two such AABBs that overlap
// have lower corners less than G apart on each axis, so their cells (taken
// at the lower corner, in double precision) are equal or adjacent: a shape
// finds every non-static partner in the 3 x 3 cells around its own. Cells
// (of each world: the world is part of the hash) are hashed onto a
// power-of-two table; cells that share a bucket only cost extra overlap
// tests (entries of another world in a shared bucket are skipped by their
// world index), and a bucket reached twice from one 3 x 3 block is scanned
// once.
// Static shapes (grounds, walls: any size) are not binned: a query tests its
// world's, at most kGridStaticMax of them. A world that is not active in the
// step keeps its pairs (its shapes are neither binned nor queried; the pair
// list's blocks of such worlds are copied whole, as with the LBVH).
//
// Steps of a broadphase (eager or captured):
//   k_grid_aabb     shape AABBs (and the step's finish, as k_aabb), zeroed
//                   pair counts and bucket counts;
//   k_grid_count    shapes per bucket;  scan -> bucket offsets;
//   k_grid_scatter  shape ids by bucket;
//   k_grid_pairs    each non-static shape's pairs into its slots;
//   scan of the pair counts; k_grid_pairs_compact (over-full shapes query
//   again straight into place). The segment sort and the merge follow as for
//   the LBVH. When shapes or body types change, k_grid_prepare recomputes G
//   and the host lists each world's static shapes.
#pragma once

#include "b2g_kernels.h"

namespace b2g {

constexpr int kGridStaticMax = 64;  // static shapes tested by every query; more: the LBVH is used

// Largest distance from the body origin to a point of the shape (the AABB of
// the shape at any pose lies within this radius of the body's position).
__device__ __forceinline__ float shape_reach(int type, int count, const float4* g) {
  const float4 h = g[0];
  const float r = h.x;
  if (type == kCircle) return sqrtf(h.y * h.y + h.z * h.z) + r;
  if (type == kEdge) {
    const float4 e = g[1], f = g[2];
    return fmaxf(sqrtf(e.z * e.z + e.w * e.w), sqrtf(f.x * f.x + f.y * f.y)) + r;
  }
  float m = 0.0f;
#pragma unroll
  for (int i = 0; i < kMaxVerts; ++i) {
    if (i < count) {
      const float4 a = g[1 + (i >> 1)];
      const float x = (i & 1) ? a.z : a.x, y = (i & 1) ? a.w : a.y;
      m = fmaxf(m, x * x + y * y);
    }
  }
  return sqrtf(m) + r;
}

__global__ void k_grid_reset(Dev d) {
  if (blockIdx.x == 0 && threadIdx.x == 0) d.grid_acc[0] = 0u;  // ordered-float maximum reach
}

// the maximum reach of the non-static shapes (one atomic per block)
__global__ void __launch_bounds__(kBlock) k_grid_prepare(Dev d) {
  __shared__ float s_m[kBlock / kWave];
  float m = 0.0f;
  B2G_FOR(i, 0, d.ns) {
    const int4 h = d.shdr[i];
    if (d.btype[h.y] != kStaticBody) m = fmaxf(m, shape_reach(h.x, h.z, d.geo + 9 * (size_t)i));
  }
  m = wave_max_all(m);
  if (lane_id() == 0) s_m[wave_id()] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)blockDim.x / kWave; ++w) m = fmaxf(m, s_m[w]);
    atomicMax(&d.grid_acc[0], float_to_ordered(m));
  }
}

struct GridGeom {
  double inv;  // 1 / cell size
  unsigned mask;
};
__device__ __forceinline__ GridGeom grid_geom(const Dev& d) {
  // twice the reach plus 1%, and a floor for point-like shapes
  const float reach = ordered_to_float(d.grid_acc[0]);
  const double g = fmax(2.0 * (double)reach * 1.01, 1e-3);
  return GridGeom{1.0 / g, (unsigned)d.grid_T - 1u};
}
__device__ __forceinline__ long long grid_coord(float v, double inv) {
  const double c = floor((double)v * inv);
  return c >= -4.0e15 && c <= 4.0e15 ? (long long)c : 0;  // NaN and far out land in cell 0
}
// An 8 x 8 tile of cells takes 64 consecutive buckets (Morton order inside
// the tile) and tiles are hashed: shapes sorted by bucket are sorted by place
// within a tile, so the lanes of a wave query neighbouring cells and share
// their cache lines.
__device__ __forceinline__ int grid_bucket(long long cx, long long cy, int w, unsigned mask) {
  const unsigned long long tx = (unsigned long long)(cx >> 3), ty = (unsigned long long)(cy >> 3);
  const unsigned long long h = tx * 0x9E3779B97F4A7C15ull ^ (ty + 0x632BE59BD9B4E019ull) * 0xC2B2AE3D27D4EB4Full ^
                               ((unsigned long long)(unsigned)w + 0x165667B19E3779F9ull) * 0xD6E8FEB86659FD93ull;
  const unsigned x = (unsigned)cx & 7u, y = (unsigned)cy & 7u;
  const unsigned m = (x & 1u) | ((y & 1u) << 1) | ((x & 2u) << 1) | ((y & 2u) << 2) | ((x & 4u) << 2) | ((y & 4u) << 3);
  return (int)((((h ^ (h >> 29)) << 6) | m) & mask);
}
// binned: a non-static shape of a world that is active in the step
__device__ __forceinline__ bool grid_binned(const Dev& d, int i) {
  return (d.sfast[i].y & 64) == 0 && !idle_world(d, d.shdr[i].w);
}

// AABBs, the step's finish (as k_aabb), and cleared pair and bucket counts;
// an idle world's shape keeps its box and takes its old segment's length as
// its pair count
__global__ void __launch_bounds__(kBlock) k_grid_aabb(Dev d, int finish) {
  if (blockIdx.x == 0 && threadIdx.x == 0) d.cnt[C_HEAVY] = 0;
  if (finish) finish_step(d);
  for (int c = (int)(blockIdx.x * blockDim.x + threadIdx.x); c <= d.grid_T; c += (int)(gridDim.x * blockDim.x))
    d.gcount[c] = 0;
  B2G_FOR(i, 0, d.ns) {
    const int4 h = d.shdr[i];
    if (idle_world(d, h.w)) {
      const int2 seg = prev_segment(d, i);
      d.pcount[i] = seg.y - seg.x;
      continue;
    }
    const float4 xf = finish && synced_body(d, h.y) ? synced_xf(d, h.y) : d.xf[h.y];
    d.aabb[i] = shape_aabb(h.x, h.z, d.geo + 9 * (size_t)i, xform_of(xf));
    d.pcount[i] = 0;
    d.pfound[i] = 0;  // a static shape's fill cursor; a queried shape's count is rewritten
  }
}

__global__ void __launch_bounds__(kBlock) k_grid_count(Dev d) {
  const GridGeom gg = grid_geom(d);
  B2G_FOR(i, 0, d.ns) {
    if (!grid_binned(d, i)) continue;
    const float4 b = d.aabb[i];
    atomicAdd(&d.gcount[grid_bucket(grid_coord(b.x, gg.inv), grid_coord(b.y, gg.inv), d.shdr[i].w, gg.mask)], 1);
  }
}

// bucket b's shapes go to gids[goff[b] .. goff[b + 1]), with a copy of their
// AABBs beside them (a query reads a bucket's boxes contiguously and touches
// a shape's index only on an overlap); the counts run back to 0
__global__ void __launch_bounds__(kBlock) k_grid_scatter(Dev d) {
  const GridGeom gg = grid_geom(d);
  B2G_FOR(i, 0, d.ns) {
    if (!grid_binned(d, i)) continue;
    const float4 b = d.aabb[i];
    const int h = grid_bucket(grid_coord(b.x, gg.inv), grid_coord(b.y, gg.inv), d.shdr[i].w, gg.mask);
    const int at = d.goff[h] + atomicSub(&d.gcount[h], 1) - 1;
    d.gids[at] = i;
    d.gbox[at] = b;
    d.gfast[at] = d.sfast[i];
    if (d.nw > 1) d.gworld[at] = d.shdr[i].w;
  }
}

// Batches of worlds of at most kGridLocalShapes binned shapes, every world
// active: each world bins its shapes into a table of its own (one workgroup
// per world, counts, offsets and placement in LDS), so no global atomic is
// needed and no bucket holds another world's shapes. World w's entries take
// the slots of its own shapes, from wshape0[w] on (static shapes are not
// binned, so they end before wshape1[w]).
constexpr int kGridLocalShapes = 4096;
constexpr int kGridLocalTable = 2 * kGridLocalShapes;  // buckets of the largest table
__global__ void __launch_bounds__(kBlock) k_grid_bin_worlds(Dev d) {
  __shared__ int cnt[kGridLocalTable];
  __shared__ int lds[kBlock / kWave];
  const GridGeom gg = grid_geom(d);
  const int tid = (int)threadIdx.x;
  for (int w = (int)blockIdx.x; w < d.nw; w += (int)gridDim.x) {
    const int s0 = d.wshape0[w], s1 = d.wshape1[w];
    const int tab = d.gltab[w];
    const int T = d.gltab[w + 1] - tab - 1;  // a power of two, 64 .. kGridLocalTable
    const unsigned mask = (unsigned)T - 1u;
    for (int h = tid; h < T; h += kBlock) cnt[h] = 0;
    __syncthreads();
    for (int i = s0 + tid; i < s1; i += kBlock) {
      if (d.sfast[i].y & 64) continue;
      const float4 b = d.aabb[i];
      atomicAdd(&cnt[grid_bucket(grid_coord(b.x, gg.inv), grid_coord(b.y, gg.inv), w, mask)], 1);
    }
    __syncthreads();
    // exclusive scan of the counts: thread t takes buckets t * per .. + per
    const int per = T / kBlock > 0 ? T / kBlock : 1;
    const int h0 = tid * per;
    int sum = 0;
    for (int k = 0; k < per; ++k)
      if (h0 + k < T) sum += cnt[h0 + k];
    int total;
    int ex = block_exclusive_scan(sum, lds, &total);
    for (int k = 0; k < per; ++k) {
      if (h0 + k < T) {
        const int c = cnt[h0 + k];
        cnt[h0 + k] = ex;  // the bucket's fill cursor
        d.gloff[tab + h0 + k] = s0 + ex;
        ex += c;
      }
    }
    if (tid == 0) d.gloff[tab + T] = s0 + total;
    __syncthreads();
    for (int i = s0 + tid; i < s1; i += kBlock) {
      const int4 f = d.sfast[i];
      if (f.y & 64) continue;
      const float4 b = d.aabb[i];
      const int at = s0 + atomicAdd(&cnt[grid_bucket(grid_coord(b.x, gg.inv), grid_coord(b.y, gg.inv), w, mask)], 1);
      d.gids[at] = i;
      d.gbox[at] = b;
      d.gfast[at] = f;
    }
    __syncthreads();  // cnt is the next world's
  }
}

// Every pair of shape i under the ownership rule: emit(key, s) as for
// lbvh_query (s >= 0: the static lower shape of a key that is not i's own).
// Where the buckets are: kGridOneWorld, the shared table of a one-world
// batch; kGridShared, the shared table of several worlds (an entry of another
// world that shares a bucket is skipped); kGridLocal, the table of i's own
// world (k_grid_bin_worlds), which holds no other world's shapes.
enum GridMode : int { kGridOneWorld = 0, kGridShared = 1, kGridLocal = 2 };
template <int kMode, class Emit>
__device__ __forceinline__ void grid_query(const Dev& d, const GridGeom& gg, int i, Emit emit) {
  constexpr bool kLocal = kMode == kGridLocal;
  const int4 fi = d.sfast[i];
  if (fi.y & 64) return;  // static: found by its partners
  const float4 q = d.aabb[i];
  // j's box overlaps q; fj is its filter word
  auto test = [&](int j, int4 fj) {
    if (j == i) return;
    const bool stat_j = (fj.y & 64) != 0;
    if (j < i && !stat_j) return;  // j's own query finds it
    if (!pair_allowed(d, fi, fj)) return;
    if (j < i)
      emit(((uint64_t)(uint32_t)j << 32) | (uint32_t)i, j);
    else
      emit(((uint64_t)(uint32_t)i << 32) | (uint32_t)j, -1);
  };
  const int w = d.shdr[i].w;
  const long long cx = grid_coord(q.x, gg.inv), cy = grid_coord(q.y, gg.inv);
  const int* off = d.goff;
  unsigned mask = gg.mask;
  if (kLocal) {
    const int tab = d.gltab[w];
    off = d.gloff + tab;
    mask = (unsigned)(d.gltab[w + 1] - tab - 2);
  }
  // the 3 x 3 block's buckets, each once; their ranges loaded together
  int hs[9], lo[9], hi[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) hs[c] = grid_bucket(cx + (c % 3) - 1, cy + (c / 3) - 1, w, mask);
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    bool again = false;
#pragma unroll
    for (int e = 0; e < c; ++e) again |= hs[e] == hs[c];
    lo[c] = again ? 0 : off[hs[c]];
    hi[c] = again ? 0 : off[hs[c] + 1];
  }
  // the candidates of the nine ranges as one sequence f = 0 .. total - 1,
  // kChunk boxes loaded before any is tested (position by compares, so the
  // arrays stay in registers)
  int pre[10];
  pre[0] = 0;
#pragma unroll
  for (int c = 0; c < 9; ++c) pre[c + 1] = pre[c] + (hi[c] - lo[c]);
  constexpr int kChunk = 8;
  for (int f0 = 0; f0 < pre[9]; f0 += kChunk) {
    int pos[kChunk];
    float4 bx[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {
      const int f = f0 + u;
      int p = -1;
#pragma unroll
      for (int c = 0; c < 9; ++c)
        if (f >= pre[c] && f < pre[c + 1]) p = lo[c] + (f - pre[c]);
      pos[u] = p;
    }
    int ids[kChunk], ws[kChunk];
    int4 fs[kChunk];
#pragma unroll
    for (int u = 0; u < kChunk; ++u) {  // the candidates' boxes, indices, filter words and worlds, all in flight
      const int at = pos[u] >= 0 ? pos[u] : 0;  // (in bounds; read but not used)
      bx[u] = d.gbox[at];
      ids[u] = d.gids[at];
      fs[u] = d.gfast[at];
      ws[u] = kMode == kGridShared ? d.gworld[at] : w;
    }
    // (a shared table's bucket can hold cells of other worlds: their shapes
    // are never partners)
#pragma unroll
    for (int u = 0; u < kChunk; ++u)
      if (pos[u] >= 0 && ws[u] == w && overlap(q, bx[u])) test(ids[u], fs[u]);
  }
  // the world's static shapes
  const int s0 = d.gsoff[w], s1 = d.gsoff[w + 1];
  for (int k0 = s0; k0 < s1; k0 += 4) {
    int js[4];
    float4 sb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) js[u] = k0 + u < s1 ? d.gstatic[k0 + u] : -1;
    int4 sf[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      sb[u] = d.aabb[js[u] >= 0 ? js[u] : i];
      sf[u] = d.sfast[js[u] >= 0 ? js[u] : i];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (js[u] >= 0 && overlap(q, sb[u])) test(js[u], sf[u]);
  }
}

// As k_lbvh_pairs: up to kSlots sorted keys per shape, the counts past them;
// lanes take the shapes in bucket order (neighbouring lanes, neighbouring
// cells). Shared table: entries 0 .. binned count. Tables of their own: world
// w's entries start at its first shape's index and end at its table's last
// word; the entry slots past that (its static shapes') are skipped.
template <int kSlots, int kThreads, int kMode>
__global__ void __launch_bounds__(kThreads) k_grid_pairs(Dev d) {
  constexpr bool kLocal = kMode == kGridLocal;
  __shared__ uint64_t slots[kSlots * kThreads];  // thread-minor
  const GridGeom gg = grid_geom(d);
  const int nq = kLocal ? d.ns : d.goff[d.grid_T];  // binned shapes
  B2G_FOR(t, 0, nq) {
    if (kLocal && t >= d.gloff[d.gltab[d.shdr[t].w + 1] - 1]) continue;
    const int i = d.gids[t];
    const int tid = (int)threadIdx.x;
    int found = 0, own = 0;
    grid_query<kMode>(d, gg, i, [&](uint64_t key, int s) {
      if (found < kSlots) {
        int b = found - 1;  // sorted insert
        while (b >= 0 && slots[b * kThreads + tid] > key) {
          slots[(b + 1) * kThreads + tid] = slots[b * kThreads + tid];
          --b;
        }
        slots[(b + 1) * kThreads + tid] = key;
      }
      ++found;
      if (s >= 0)
        atomicAdd(&d.pcount[s], 1);
      else
        ++own;
    });
    uint64_t* out = d.pslots + (size_t)i * kSlots;
    const int kept = found < kSlots ? found : kSlots;
    for (int k = 0; k < kept; ++k) out[k] = slots[k * kThreads + tid];
    d.pfound[i] = found;
    d.pcount[i] = own;
  }
}

// As k_pairs_compact, with the grid query for over-full shapes.
template <int kSlots, int kMode>
__global__ void __launch_bounds__(kQueryBlock) k_grid_pairs_compact(Dev d) {
  const GridGeom gg = grid_geom(d);
  B2G_FOR(i, 0, d.ns) {
    // an idle world's block of pairs is copied whole (k_keep_idle_pairs)
    if (idle_world(d, d.shdr[i].w)) continue;
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
      // over-full shape: in query order, sorted afterwards by k_sort_pair_segments
      int k = 0;
      grid_query<kMode>(d, gg, i, [&](uint64_t key, int s) {
        if (s >= 0)
          foreign(key, s);
        else if (k < own)
          d.nkey[base + k++] = key;
      });
      if (own > 1) d.heavy[atomicAdd(&d.cnt[C_HEAVY], 1)] = make_int2(base, own);
    }
  }
}

}  // namespace b2g
