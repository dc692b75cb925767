// b2g_partition.h -- the contact solver of one large world, by spatial cells.
//
// The colour sweeps of b2g_kernels.h need one launch per colour and
// iteration: a 20,000-body pile takes about nine per velocity iteration, and
// on MI355X a launch costs about as much as the work of a small colour. Here
// a large world's constraints are split three ways each step instead:
//   A  both dynamic bodies in one cell of a uniform grid (or one dynamic body
//      and a static or kinematic one): solved by that cell's workgroup;
//   B  otherwise, both in one cell of the same grid shifted by half a cell;
//   C  the few left over (bodies in different cells of both grids).
// A cell's workgroup stages its bodies in LDS, colours its constraints in LDS
// (Jones-Plassmann with fixed priorities) and sweeps the colours with a
// workgroup barrier between them. Cells of one grid share no dynamic body, so
// an iteration is three launches: A, B, C. The order is a Gauss-Seidel order
// like the colour sweeps' (every constraint once per iteration, sequential
// where bodies are shared); the colour reported per contact is
// phase * kPhaseStride + colour, which a restatement that sorts constraints by
// that value replays exactly. Joints keep their own sweeps around these
// (b2_island.cpp:329-338 velocity, 389-401 position).
#pragma once

#include "b2g_kernels.h"

namespace b2g {

constexpr int kCells = 1024;              // cells per grid (A, and the shifted B)
constexpr int kPartGroups = 2 * kCells + 1;  // A cells, B cells, the cross group C
constexpr int kCrossGroup = 2 * kCells;
constexpr int kCellBodies = 256;          // bodies of one cell (LDS); a fuller cell's constraints go to C
constexpr int kCellCons = 512;            // constraints of one cell; more go to C
constexpr int kCellTarget = 48;           // bodies per cell the grid aims at
constexpr int kPhaseStride = 1 << 16;     // reported colour: phase * kPhaseStride + colour
constexpr int kCellThreads = 128;         // workgroup of a cell (colouring and sweeps)
constexpr int kCrossThreads = 512;        // workgroup of the cross group's sweeps

// Grid geometry: part_geo[0] = (lo.x, lo.y, 1/L, L). Cell (i, j) of grid A
// holds the centres with floor((p - lo) / L) = (i, j); grid B is shifted by
// half a cell along both axes. The cells of each grid are hashed onto its
// kCells buckets: cells far apart may share a bucket (one workgroup then
// solves both, whose constraints are still internal to it), so the size of
// the region the bodies cover does not bound the cell size.
constexpr float kPartCoordMax = 1048576.0f;
__device__ __forceinline__ int part_coord(float v, float lo, float inv_l, float shift) {
  const float f = (v - lo) * inv_l + shift;
  return f >= 0.0f ? (int)fminf(f, kPartCoordMax) : 0;  // NaN lands in 0
}
__device__ __forceinline__ int part_bucket(int i, int j, unsigned seed) {
  unsigned h = (unsigned)i * 0x9E3779B1u ^ ((unsigned)j + seed) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  return (int)(h % (unsigned)kCells);
}
struct PartGrid {
  float lx, ly, inv_l;
  __device__ __forceinline__ int cell_a(float4 p) const {
    return part_bucket(part_coord(p.x, lx, inv_l, 0.0f), part_coord(p.y, ly, inv_l, 0.0f), 0u);
  }
  __device__ __forceinline__ int cell_b(float4 p) const {
    return part_bucket(part_coord(p.x, lx, inv_l, 0.5f), part_coord(p.y, ly, inv_l, 0.5f), 0x5BD1E995u);
  }
};
__device__ __forceinline__ PartGrid part_grid(const Dev& d) {
  const float4 g0 = d.part_geo[0];
  return PartGrid{g0.x, g0.y, g0.z};
}
__device__ __forceinline__ bool part_dynamic(const Dev& d, int b) { return d.btype[b] == kDynamicBody; }

// block-wide exclusive scan of one int per thread (any multiple of kWave
// threads up to 1024); lds: 16 ints
__device__ __forceinline__ int part_block_scan(int v, int* lds, int* total) {
  const int inc = wave_inclusive_scan(v);
  const int nw = (int)blockDim.x / kWave;
  if (lane_id() == kWave - 1) lds[wave_id()] = inc;
  __syncthreads();
  int before = 0, sum = 0;
  for (int w = 0; w < nw; ++w) {
    const int s = lds[w];
    if (w < wave_id()) before += s;
    sum += s;
  }
  __syncthreads();
  *total = sum;
  return before + inc - v;
}

// True in the workgroup of a launch that finishes last (every workgroup of
// the launch must call it once, uniformly); the other workgroups' global
// writes are visible to it. It puts the counter back to 0 for the next launch.
// One lane fences: after the barrier the workgroup's stores have reached its
// L2, and one agent-scope release writes that L2 back (a fence in every wave
// would write it back once per wave, which costs more than a launch).
__device__ __forceinline__ bool part_last_block(unsigned* counter) {
  __shared__ int last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();  // this workgroup's writes, device-wide
    const unsigned done = atomicAdd(counter, 1u);
    last = done == gridDim.x - 1 ? 1 : 0;
    if (last) {
      atomicExch(counter, 0u);
      __threadfence();  // and the others' writes before this workgroup reads them
    }
  }
  __syncthreads();
  return last != 0;
}

// The grid of a step, in three launches over the bodies and constraints:
//   k_part_bounds    bounds of the dynamic bodies' centres (ordered-float
//                    atomics, one set per workgroup), accumulated in
//                    part_acc; its last workgroup then sets the cell size for
//                    about kCellTarget bodies per cell, clears the
//                    histograms and constraint counts, and resets part_acc
//                    for the next step;
//   k_part_hist      bodies per cell of either grid; its last workgroup
//                    turns them into the offsets of the cells' body lists
//                    (A cells, then B cells);
//   k_part_classify  every dynamic body into its two cells' lists (its slot
//                    in each), and every solved constraint into its group.
// A cell of more than kCellBodies bodies (its LDS stage would not hold them)
// hands its constraints to the cross group.
__device__ __forceinline__ bool part_counts(const Dev& d, int b) {
  return part_dynamic(d, b) && world_active(d, d.bworld[b]);
}

// Bounds: each lane over a grid stride, then the workgroup (shuffles, then
// LDS across its waves), then one set of atomics per workgroup; the launch's
// last workgroup goes on to the grid (part_grid_from_bounds). Launched with
// at most kPartBoundsBlocks workgroups of kBlock threads. Beside the bounds
// it sums the bodies' footprints, 8 I / m (the square around a disc of the
// same radius of gyration), in fixed point so that the sum, and with it the
// grid, does not depend on the order of the additions.
constexpr int kPartBoundsBlocks = 64;
constexpr float kPartAreaScale = 256.0f;  // fixed point of the footprint sum (1/256 m^2)
__device__ __forceinline__ unsigned long long* part_area_acc(const Dev& d) {
  return reinterpret_cast<unsigned long long*>(d.part_acc + 6);
}
__device__ __forceinline__ void part_grid_from_bounds(const Dev& d);
__global__ void __launch_bounds__(kBlock) k_part_bounds(Dev d) {
  __shared__ float s_b[4][kBlock / kWave];
  __shared__ int s_n[kBlock / kWave];
  __shared__ unsigned long long s_area;
  if (threadIdx.x == 0) s_area = 0ull;
  __syncthreads();
  float lx = INFINITY, ly = INFINITY, hx = -INFINITY, hy = -INFINITY;
  int n = 0;
  unsigned long long area = 0ull;
  B2G_FOR(b, 0, d.nb) {
    if (!part_counts(d, b)) continue;
    const float4 p = d.pos[b];
    if (fabsf(p.x) <= 1e30f && fabsf(p.y) <= 1e30f) {  // NaN or far out: an edge cell
      lx = fminf(lx, p.x);
      ly = fminf(ly, p.y);
      hx = fmaxf(hx, p.x);
      hy = fmaxf(hy, p.y);
      ++n;
      const float4 pr = d.prop[b];  // inverse mass, inverse rotational inertia
      if (pr.y > 0.0f) area += (unsigned long long)(fminf(8.0f * pr.x / pr.y, 1e6f) * kPartAreaScale);
    }
  }
  if (area) atomicAdd(&s_area, area);
  for (int off = kWave / 2; off > 0; off >>= 1) {
    lx = fminf(lx, __shfl_xor(lx, off, kWave));
    ly = fminf(ly, __shfl_xor(ly, off, kWave));
    hx = fmaxf(hx, __shfl_xor(hx, off, kWave));
    hy = fmaxf(hy, __shfl_xor(hy, off, kWave));
    n += __shfl_xor(n, off, kWave);
  }
  if (lane_id() == 0) {
    s_b[0][wave_id()] = lx;
    s_b[1][wave_id()] = ly;
    s_b[2][wave_id()] = hx;
    s_b[3][wave_id()] = hy;
    s_n[wave_id()] = n;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)blockDim.x / kWave; ++w) {
      lx = fminf(lx, s_b[0][w]);
      ly = fminf(ly, s_b[1][w]);
      hx = fmaxf(hx, s_b[2][w]);
      hy = fmaxf(hy, s_b[3][w]);
      n += s_n[w];
    }
    if (n > 0) {
      atomicMin(&d.part_acc[0], float_to_ordered(lx));
      atomicMin(&d.part_acc[1], float_to_ordered(ly));
      atomicMax(&d.part_acc[2], float_to_ordered(hx));
      atomicMax(&d.part_acc[3], float_to_ordered(hy));
      atomicAdd(&d.part_acc[4], (unsigned)n);
      if (s_area) atomicAdd(part_area_acc(d), s_area);
    }
  }
  if (part_last_block(d.part_sync + 0)) part_grid_from_bounds(d);
}

// the empty accumulator of k_part_bounds
__device__ __forceinline__ void part_acc_reset(const Dev& d) {
  d.part_acc[0] = 0xffffffffu;
  d.part_acc[1] = 0xffffffffu;
  d.part_acc[2] = 0u;
  d.part_acc[3] = 0u;
  d.part_acc[4] = 0u;
  *part_area_acc(d) = 0ull;
}

__device__ __forceinline__ void part_grid_from_bounds(const Dev& d) {
  const int tid = (int)threadIdx.x;
  if (tid == 0) {
    const unsigned n = d.part_acc[4];
    float lx = 0.0f, ly = 0.0f, hx = 0.0f, hy = 0.0f;
    if (n > 0) {
      lx = ordered_to_float(d.part_acc[0]);
      ly = ordered_to_float(d.part_acc[1]);
      hx = ordered_to_float(d.part_acc[2]);
      hy = ordered_to_float(d.part_acc[3]);
    }
    const float w = fminf(fmaxf(hx - lx, 1e-3f), 1e30f), h = fminf(fmaxf(hy - ly, 1e-3f), 1e30f);
    const float per = 1.0f / (float)(n > 0 ? n : 1);
    // the cell for about kCellTarget bodies: of the mean footprint where the
    // bodies are packed, of the bounds where they are spread out
    const float l_bounds = sqrtf(w * h * (float)kCellTarget * per);
    const float l_packed = sqrtf((float)*part_area_acc(d) / kPartAreaScale * (float)kCellTarget * per);
    const float L = fmaxf(l_packed > 0.0f ? fminf(l_bounds, l_packed) : l_bounds, 1e-3f);
    d.part_geo[0] = make_float4(lx, ly, 1.0f / L, L);
    part_acc_reset(d);
  }
  for (int c = tid; c < 2 * kCells; c += blockDim.x) d.part_boff[c] = 0;  // the histogram, scanned in place
  for (int g = tid; g < kPartGroups; g += blockDim.x) d.part_ccnt[g] = 0;
}

// exclusive scan of the 2 * kCells body counts in place; the fill cursors
// start at the offsets (one workgroup of kBlock threads)
__device__ __forceinline__ void part_scan_counts(const Dev& d) {
  __shared__ int scan_lds[16];
  constexpr int kPer = 2 * kCells / kBlock;
  static_assert(kPer * kBlock == 2 * kCells, "whole cells per thread");
  const int tid = (int)threadIdx.x;
  int v[kPer], s = 0;
  for (int k = 0; k < kPer; ++k) {
    v[k] = d.part_boff[tid * kPer + k];
    s += v[k];
  }
  int total;
  int run = part_block_scan(s, scan_lds, &total);
  for (int k = 0; k < kPer; ++k) {
    d.part_boff[tid * kPer + k] = run;
    d.part_bcur[tid * kPer + k] = run;
    run += v[k];
  }
  if (tid == 0) d.part_boff[2 * kCells] = total;
}

// bodies per cell of either grid; the launch's last workgroup scans them
// (launched with at most kPartBoundsBlocks workgroups)
__global__ void __launch_bounds__(kBlock) k_part_hist(Dev d) {
  const PartGrid pg = part_grid(d);
  B2G_FOR(b, 0, d.nb) {
    if (!part_counts(d, b)) continue;
    const float4 p = d.pos[b];
    atomicAdd(&d.part_boff[pg.cell_a(p)], 1);
    atomicAdd(&d.part_boff[kCells + pg.cell_b(p)], 1);
  }
  if (part_last_block(d.part_sync + 1)) part_scan_counts(d);
}

__device__ __forceinline__ bool part_cell_fits(const Dev& d, int cell) {
  return d.part_boff[cell + 1] - d.part_boff[cell] <= kCellBodies;
}

// Bodies into their cells' lists; solved constraints (compacted slot k) into
// their groups' buckets.
__global__ void __launch_bounds__(kBlock) k_part_classify(Dev d) {
  const int n = d.cnt[C_CONS];
  const PartGrid pg = part_grid(d);
  const int m = max(n, d.nb);
  B2G_FOR(k, 0, m) {
    if (k < d.nb) {
      const int b = k;
      if (!part_counts(d, b)) {
        d.part_brank[b] = make_int2(-1, -1);
      } else {
        const float4 p = d.pos[b];
        const int ca = pg.cell_a(p), cb = kCells + pg.cell_b(p);
        const int ra = atomicAdd(&d.part_bcur[ca], 1), rb = atomicAdd(&d.part_bcur[cb], 1);
        d.part_blist[ra] = b;
        d.part_blist[rb] = b;
        d.part_brank[b] = make_int2(ra - d.part_boff[ca], rb - d.part_boff[cb]);
      }
    }
    if (k >= n) continue;
    const int2 bp = d.cbody[k];
    const int ia = bp.x >= 0 ? bp.x : ~bp.x, ib = bp.y >= 0 ? bp.y : ~bp.y;
    const bool da = part_dynamic(d, ia), db = part_dynamic(d, ib);
    int g = kCrossGroup;
    if (da && db) {
      const float4 pa = d.pos[ia], pb = d.pos[ib];
      const int a0 = pg.cell_a(pa), a1 = pg.cell_a(pb);
      if (a0 == a1) {
        g = a0;
      } else {
        const int b0 = pg.cell_b(pa), b1 = pg.cell_b(pb);
        if (b0 == b1) g = kCells + b0;
      }
    } else if (da || db) {
      g = pg.cell_a(d.pos[da ? ia : ib]);
    }
    if (g != kCrossGroup && part_cell_fits(d, g)) {
      const int r = atomicAdd(&d.part_ccnt[g], 1);
      if (r < kCellCons) {
        d.part_bucket[(size_t)g * kCellCons + r] = k;
        continue;
      }
    }
    const int r = atomicAdd(&d.part_ccnt[kCrossGroup], 1);
    d.part_cbucket[r] = k;
  }
}

__device__ __forceinline__ int part_group_size(const Dev& d, int g) {
  const int c = d.part_ccnt[g];
  return g == kCrossGroup ? c : min(c, kCellCons);
}

// Jones-Plassmann colouring priority of compacted constraint k: a bijection
// of k, so no two constraints tie; 0 means "no claim".
__device__ __forceinline__ unsigned long long part_priority(int k) {
  return (unsigned long long)hash32((uint32_t)k) + 1ull;
}

// Colours of one group's constraints, then their slots: the group's slots
// start after every earlier group's (A cells, B cells, C) and are sorted by
// colour; a group's constraints that find no colour below kMaxColours form
// its serial group, in constraint order. Writes the slot -> contact map, the
// LDS slots of each slot's bodies, the group's colour offsets and every
// contact's reported colour (in ccol, beyond kMaxColours, so the colour
// sweeps would treat it as no colour). `Claims` gives the claim and used-colour words
// of a body (LDS slots for a cell, global arrays for the cross group).
template <class Ids, class Claims>
__device__ void part_colour_group(const Dev& d, int g, int start, int n, Ids ids, Claims cl, int* s_col,
                                  int* hist, int* cursor) {
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  for (int i = tid; i < n; i += nt) {
    const int2 l = ids.bodies(i);
    if (l.x >= 0) {
      *cl.claim(l.x) = 0ull;
      *cl.used(l.x) = 0ull;
    }
    if (l.y >= 0) {
      *cl.claim(l.y) = 0ull;
      *cl.used(l.y) = 0ull;
    }
    s_col[i] = -1;
  }
  for (int c = tid; c <= kColourSlots; c += nt) hist[c] = 0;
  __syncthreads();
  for (;;) {
    int left = 0;
    for (int i = tid; i < n; i += nt) {
      if (s_col[i] >= 0) continue;
      left = 1;
      const int2 l = ids.bodies(i);
      const unsigned long long p = part_priority(ids.k(i));
      if (l.x >= 0) atomicMax(cl.claim(l.x), p);
      if (l.y >= 0) atomicMax(cl.claim(l.y), p);
    }
    if (!__syncthreads_or(left)) break;
    for (int i = tid; i < n; i += nt) {
      if (s_col[i] >= 0) continue;
      const int2 l = ids.bodies(i);
      const unsigned long long p = part_priority(ids.k(i));
      const bool win = (l.x < 0 || *cl.claim(l.x) == p) && (l.y < 0 || *cl.claim(l.y) == p);
      if (!win) continue;
      // winners share no body, so their used-colour words are theirs alone
      const unsigned long long m = (l.x >= 0 ? *cl.used(l.x) : 0ull) | (l.y >= 0 ? *cl.used(l.y) : 0ull);
      const int c = ~m ? __ffsll((long long)~m) - 1 : kSerialColour;
      s_col[i] = c;
      if (c < kMaxColours) {
        if (l.x >= 0) *cl.used(l.x) |= 1ull << c;
        if (l.y >= 0) *cl.used(l.y) |= 1ull << c;
      }
    }
    __syncthreads();
    for (int i = tid; i < n; i += nt) {
      const int2 l = ids.bodies(i);
      if (l.x >= 0) *cl.claim(l.x) = 0ull;
      if (l.y >= 0) *cl.claim(l.y) = 0ull;
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += nt) atomicAdd(&hist[s_col[i]], 1);
  __syncthreads();
  if (tid == 0) {
    int run = start;
    for (int c = 0; c < kColourSlots; ++c) {
      const int h = hist[c];
      hist[c] = run;
      cursor[c] = run;
      run += h;
    }
    hist[kColourSlots] = run;
  }
  __syncthreads();
  int* coff = d.part_coff + (size_t)g * (kColourSlots + 1);
  for (int c = tid; c <= kColourSlots; c += nt) coff[c] = hist[c];
  const int phase = g < kCells ? 0 : g < kCrossGroup ? 1 : 2;
  for (int i = tid; i < n; i += nt) {
    const int col = s_col[i];
    const int s = atomicAdd(&cursor[col], 1);
    const int k = ids.k(i);
    const int c = d.clist[k];
    d.csorted[s] = c;
    d.sloc[s] = ids.local(i);
    d.ccol[c] = phase * kPhaseStride + col;  // out of the colour sweeps' range: they recolour it
  }
  __syncthreads();
  if (tid == 0 && hist[kSerialColour + 1] > hist[kSerialColour]) {
    // the serial group in contact order (contact indices grow with the keys)
    const int b = hist[kSerialColour], e = hist[kSerialColour + 1];
    for (int a = b + 1; a < e; ++a) {
      const int v = d.csorted[a];
      const int2 lv = d.sloc[a];
      int q = a - 1;
      while (q >= b && d.csorted[q] > v) {
        d.csorted[q + 1] = d.csorted[q];
        d.sloc[q + 1] = d.sloc[q];
        --q;
      }
      d.csorted[q + 1] = v;
      d.sloc[q + 1] = lv;
    }
  }
}

// cell groups: constraint ids and body slots in LDS
struct CellIds {
  const int* k_;
  const int2* l_;
  __device__ __forceinline__ int k(int i) const { return k_[i]; }
  __device__ __forceinline__ int2 bodies(int i) const { return l_[i]; }
  __device__ __forceinline__ int2 local(int i) const { return l_[i]; }
};
struct LdsClaims {
  unsigned long long* c;
  unsigned long long* u;
  __device__ __forceinline__ unsigned long long* claim(int r) const { return c + r; }
  __device__ __forceinline__ unsigned long long* used(int r) const { return u + r; }
};
// the cross group: constraint ids from its bucket, dynamic bodies by index
struct CrossIds {
  const Dev& d;
  __device__ __forceinline__ int k(int i) const { return d.part_cbucket[i]; }
  __device__ __forceinline__ int2 bodies(int i) const {
    const int2 bp = d.cbody[d.part_cbucket[i]];
    const int ia = bp.x >= 0 ? bp.x : ~bp.x, ib = bp.y >= 0 ? bp.y : ~bp.y;
    return make_int2(part_dynamic(d, ia) ? ia : -1, part_dynamic(d, ib) ? ib : -1);
  }
  __device__ __forceinline__ int2 local(int) const { return make_int2(-1, -1); }
};
struct GlobalClaims {
  unsigned long long* c;
  unsigned long long* u;
  __device__ __forceinline__ unsigned long long* claim(int b) const { return c + b; }
  __device__ __forceinline__ unsigned long long* used(int b) const { return u + b; }
};

__device__ __forceinline__ int part_group_start(const Dev& d, int g, int* scan_lds) {
  int s = 0;
  for (int q = (int)threadIdx.x; q < g; q += blockDim.x) s += part_group_size(d, q);
  int total;
  (void)part_block_scan(s, scan_lds, &total);
  return total;
}

__device__ __forceinline__ void part_colour_cross(const Dev& d);

// one workgroup per cell of either grid, then one for the cross group
__global__ void __launch_bounds__(kCellThreads) k_part_colour_cells(Dev d) {
  if ((int)blockIdx.x == 2 * kCells) {
    part_colour_cross(d);
    return;
  }
  __shared__ int s_k[kCellCons];
  __shared__ int2 s_l[kCellCons];
  __shared__ int s_col[kCellCons];
  __shared__ unsigned long long s_claim[kCellBodies];
  __shared__ unsigned long long s_used[kCellBodies];
  __shared__ int hist[kColourSlots + 1];
  __shared__ int cursor[kColourSlots];
  __shared__ int scan_lds[16];
  const int g = (int)blockIdx.x;
  const int start = part_group_start(d, g, scan_lds);
  const int n = part_group_size(d, g);
  const int phase = g < kCells ? 0 : 1;
  for (int i = (int)threadIdx.x; i < n; i += blockDim.x) {
    const int k = d.part_bucket[(size_t)g * kCellCons + i];
    const int2 bp = d.cbody[k];
    const int ia = bp.x >= 0 ? bp.x : ~bp.x, ib = bp.y >= 0 ? bp.y : ~bp.y;
    const int2 ra = d.part_brank[ia], rb = d.part_brank[ib];
    s_k[i] = k;
    s_l[i] = make_int2(part_dynamic(d, ia) ? (phase ? ra.y : ra.x) : -1, part_dynamic(d, ib) ? (phase ? rb.y : rb.x) : -1);
  }
  __syncthreads();
  part_colour_group(d, g, start, n, CellIds{s_k, s_l}, LdsClaims{s_claim, s_used}, s_col, hist, cursor);
}

// the cross group, one workgroup; its per-constraint colours in global memory
__device__ __forceinline__ void part_colour_cross(const Dev& d) {
  __shared__ int hist[kColourSlots + 1];
  __shared__ int cursor[kColourSlots];
  __shared__ int scan_lds[16];
  const int start = part_group_start(d, kCrossGroup, scan_lds);
  const int n = part_group_size(d, kCrossGroup);
  if ((int)threadIdx.x == 0) d.cnt[C_PART_CROSS] = n;
  part_colour_group(d, kCrossGroup, start, n, CrossIds{d}, GlobalClaims{d.part_claim, d.part_used}, d.part_xcol,
                    hist, cursor);
}

// Body access of a cell's sweep: LDS slots [0, kCellBodies) hold the cell's
// dynamic bodies (written back at the end), the slots after them a copy of
// each staged constraint's static or kinematic body (read only); a body
// without a slot (loc < 0) is read and written in global memory.
struct CellBodies {
  float4* lds;
  float4* glob;
  int ia;
  int2 loc;
  __device__ __forceinline__ float4* operator()(int i) const {
    const int l = i == ia ? loc.x : loc.y;
    return l >= 0 ? lds + l : glob + i;
  }
  __device__ __forceinline__ bool writes(const Dev& d, int i) const {
    const int l = i == ia ? loc.x : loc.y;
    return l >= 0 ? l < kCellBodies : writes_body(d, i);
  }
};

enum PartOp : int { kPartWarm = 0, kPartVelocity = 1, kPartPosition = 2 };

// A group's constraint records staged in LDS for one sweep: the group's slots
// are contiguous, so one coalesced pass loads them all at the start and every
// colour round then reads LDS. Warm start and velocity: the ten velocity
// words; position: word 2 of the velocity record, the five position words
// and the island. Slots past kStageCons are read from global memory.
constexpr int kStageCons = 96;
template <int kOp>
struct PartStage {
  static constexpr int kWords = kOp == kPartPosition ? 6 : 10;
};
// A staged constraint's solved impulses stay in LDS until the sweep's end, and
// its island's position test is made once at the start of the sweep (it only
// reads the previous iteration's marks): the colour rounds then touch no
// global memory, so their barriers wait for LDS alone.
struct StagedVelocity {  // velocity word k of staged constraint i
  float4* base;
  int i;
  __device__ __forceinline__ float4 vc(int k) const { return base[k * kStageCons + i]; }
  __device__ __forceinline__ void put_impulses(const Dev&, int, float4 q0, float4 q1, int count) const {
    base[4 * kStageCons + i] = q0;
    if (count > 1) base[6 * kStageCons + i] = q1;
  }
};
struct StagedPosition {  // velocity word 2, then position words 0..4
  const float4* base;
  const int* skip;
  int i;
  __device__ __forceinline__ float4 vc(int) const { return base[i]; }
  __device__ __forceinline__ float4 pc(int k) const { return base[(1 + k) * kStageCons + i]; }
  __device__ __forceinline__ bool island_skip(const Dev&, int, int) const { return skip[i] != 0; }
};

// LDS of one group's sweep
template <int kOp>
struct PartLds {
  int off[kColourSlots + 1];
  float4 body[kCellBodies + 2 * kStageCons];
  float4 con[PartStage<kOp>::kWords * kStageCons];
  int isl[kOp == kPartPosition ? kStageCons : 1];
  int skip[kOp == kPartPosition ? kStageCons : 1];  // the island's iterations are over
  float sep[kOp == kPartPosition ? kStageCons : 1];  // minimum separation of the solve
  int2 loc[kStageCons];
};

// One group's share of a sweep: a cell (its bodies and constraints staged in
// LDS) or the cross group (global memory).
template <int kOp>
__device__ __forceinline__ void part_sweep_group(const Dev& d, int g, int it, PartLds<kOp>& L) {
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  const bool staged = g != kCrossGroup;
  float4* body = kOp == kPartPosition ? d.pos : d.vel;
  const int* coff = d.part_coff + (size_t)g * (kColourSlots + 1);
  int b0 = 0, b1 = 0;
  if (staged) {  // (issued together with the colour offsets)
    b0 = d.part_boff[g];
    b1 = d.part_boff[g + 1];
  }
  for (int c = tid; c <= kColourSlots; c += nt) L.off[c] = coff[c];
  __syncthreads();
  const int s0 = L.off[0];
  if (L.off[kColourSlots] <= s0) return;  // the whole workgroup
  const int staged_n = min(L.off[kColourSlots] - s0, kStageCons);
  for (int r = tid; r < b1 - b0; r += nt) L.body[r] = body[d.part_blist[b0 + r]];
  for (int i = tid; i < staged_n; i += nt) {
    float4 v2;
    if (kOp == kPartPosition) {
      v2 = d.vc[2][s0 + i];
      L.con[i] = v2;
#pragma unroll
      for (int k = 0; k < 5; ++k) L.con[(1 + k) * kStageCons + i] = d.pc[k][s0 + i];
      const int isl = d.cisland[s0 + i];
      L.isl[i] = isl;
      L.skip[i] = as_int(L.con[5 * kStageCons + i].w) != 0 && island_skip_global(d, isl, it) ? 1 : 0;
      L.sep[i] = 0.0f;
    } else {
#pragma unroll
      for (int k = 0; k < 10; ++k) L.con[k * kStageCons + i] = d.vc[k][s0 + i];
      v2 = L.con[2 * kStageCons + i];
    }
    int2 loc = make_int2(-1, -1);
    if (staged) {  // a cell's constraint: its static or kinematic body gets a read-only slot
      loc = d.sloc[s0 + i];
      if (loc.x < 0) {
        loc.x = kCellBodies + 2 * i;
        L.body[loc.x] = body[as_int(v2.x)];
      }
      if (loc.y < 0) {
        loc.y = kCellBodies + 2 * i + 1;
        L.body[loc.y] = body[as_int(v2.y)];
      }
    }
    L.loc[i] = loc;
  }
  __syncthreads();
  // slot s (valid) with its bodies in the cell's stage or in global memory
  auto solve = [&](int s, bool valid) {
    const int i = s - s0;
    const bool in_lds = valid && i < staged_n;
    const int2 loc = !valid || !staged ? make_int2(-1, -1) : in_lds ? L.loc[i] : d.sloc[s];
    const float4 v2 = !valid ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : in_lds ? L.con[(kOp == kPartPosition ? 0 : 2) * kStageCons + i] : d.vc[2][s];
    const CellBodies bv{L.body, body, as_int(v2.x), loc};
    if (kOp == kPartWarm) {
      if (in_lds)
        warm_start_src(d, bv, StagedVelocity{L.con, i});
      else if (valid)
        warm_start_src(d, bv, SlotFields{d, s});
    } else if (kOp == kPartVelocity) {
      if (in_lds)
        contact_velocity_src(d, s, bv, StagedVelocity{L.con, i});
      else if (valid)
        contact_velocity_src(d, s, bv, SlotFields{d, s});
    } else {
      const int isl = !valid ? 0 : in_lds ? L.isl[i] : d.cisland[s];
      float m = 0.0f;
      if (in_lds)
        m = contact_position_src(d, isl, it, bv, StagedPosition{L.con, L.skip, i});
      else if (valid)
        m = contact_position_src(d, isl, it, bv, SlotFields{d, s});
      return m;
    }
    return 0.0f;
  };
  for (int c = 0; c < kColourSlots; ++c) {
    const int lo = L.off[c], hi = L.off[c + 1];
    if (hi <= lo) continue;  // uniform across the workgroup
    if (c == kSerialColour) {
      if (tid == 0) {
        for (int s = lo; s < hi; ++s) {
          const float m = solve(s, true);
          if (kOp == kPartPosition && separation_bad(m)) {
            const int isl = s - s0 < staged_n ? L.isl[s - s0] : d.cisland[s];
            d.sepbad[island_slot(d, isl, it)] = 1u;
          }
        }
      }
    } else {
      for (int base = lo; base < hi; base += nt) {  // every lane of a wave takes part in the island marks
        const int s = base + tid;
        const bool valid = s < hi;
        const float m = solve(s, valid);
        if (kOp == kPartPosition) {
          const bool in_lds = valid && s - s0 < staged_n;
          if (in_lds) L.sep[s - s0] = m;  // marked at the end of the sweep
          mark_island_sep(d, it, in_lds || !valid ? 0 : d.cisland[s], valid && !in_lds, m);
        }
      }
    }
    __syncthreads();
  }
  if (kOp == kPartVelocity) {  // the staged constraints' impulses
    for (int i = tid; i < staged_n; i += nt) {
      d.vc[4][s0 + i] = L.con[4 * kStageCons + i];
      if ((as_int(L.con[2 * kStageCons + i].z) & 0xff) > 1) d.vc[6][s0 + i] = L.con[6 * kStageCons + i];
    }
  }
  if (kOp == kPartPosition) {  // the staged constraints' island marks
    for (int base = 0; base < staged_n; base += nt) {
      const int i = base + tid;
      const bool valid = i < staged_n;
      mark_island_sep(d, it, valid ? L.isl[i] : 0, valid, valid ? L.sep[i] : 0.0f);
    }
  }
  for (int r = tid; r < b1 - b0; r += nt) body[d.part_blist[b0 + r]] = L.body[r];
}

// One phase of a sweep: the cells of grid A (phase 0) or of grid B (phase
// 1), a workgroup each (one past the grid's cells has nothing to do), or the
// cross group (phase 2, one workgroup), whose constraints touch bodies of
// both grids' cells.
template <int kOp, int kThreads>
__global__ void __launch_bounds__(kThreads) k_part_sweep(Dev d, int phase, int it) {
  __shared__ PartLds<kOp> L;
  part_sweep_group<kOp>(d, phase == 2 ? kCrossGroup : phase * kCells + (int)blockIdx.x, it, L);
}

}  // namespace b2g
