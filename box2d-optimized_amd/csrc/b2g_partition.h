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
constexpr int kCellBodies = 512;          // bodies of one cell (LDS); larger cells make the grid finer
constexpr int kCellCons = 512;            // constraints of one cell; more go to C
constexpr int kCellTarget = 48;           // bodies per cell the grid aims at
constexpr int kPhaseStride = 1 << 16;     // reported colour: phase * kPhaseStride + colour
constexpr int kPartGridThreads = 1024;
constexpr int kCellThreads = 128;         // workgroup of a cell (colouring and sweeps)
constexpr int kCrossThreads = 512;        // workgroup of the cross group

// grid geometry: part_geo[0] = (lo.x, lo.y, 1/L, L), part_geo[1] = (Gx, Gy,
// all-cross, -) with the ints as bits
__device__ __forceinline__ int part_coord(float v, float lo, float inv_l, float shift, int n) {
  const float f = (v - lo) * inv_l + shift;
  return f >= 0.0f ? (int)fminf(f, (float)(n - 1)) : 0;  // NaN lands in 0
}
struct PartGrid {
  float lx, ly, inv_l;
  int gx, gy, all_cross;
  __device__ __forceinline__ int cell_a(float4 p) const {
    return part_coord(p.x, lx, inv_l, 0.0f, gx) + gx * part_coord(p.y, ly, inv_l, 0.0f, gy);
  }
  __device__ __forceinline__ int cell_b(float4 p) const {
    return part_coord(p.x, lx, inv_l, 0.5f, gx + 1) + (gx + 1) * part_coord(p.y, ly, inv_l, 0.5f, gy + 1);
  }
};
__device__ __forceinline__ PartGrid part_grid(const Dev& d) {
  const float4 g0 = d.part_geo[0], g1 = d.part_geo[1];
  return PartGrid{g0.x, g0.y, g0.z, as_int(g1.x), as_int(g1.y), as_int(g1.z)};
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

// One workgroup: the grid for this step and every dynamic body's cells. The
// bounds of the dynamic bodies' centres give a cell size for about
// kCellTarget bodies per cell; it grows until no cell of either grid holds
// more than kCellBodies bodies and both grids fit kCells cells. Then the
// bodies are listed cell by cell (A cells, then B cells) with their slot in
// each. Also clears the constraint counts of the step.
__global__ void __launch_bounds__(kPartGridThreads) k_part_bodies(Dev d) {
  __shared__ int hist[2 * kCells];
  __shared__ float red[4][kPartGridThreads / kWave];
  __shared__ int redn[kPartGridThreads / kWave];
  __shared__ float s_geo[4];
  __shared__ int s_int[4];
  __shared__ int scan_lds[16];
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  for (int g = tid; g < kPartGroups; g += nt) d.part_ccnt[g] = 0;
  float lx = INFINITY, ly = INFINITY, hx = -INFINITY, hy = -INFINITY;
  int ndyn = 0;
  for (int b = tid; b < d.nb; b += nt) {
    if (!part_dynamic(d, b) || !world_active(d, d.bworld[b])) continue;
    const float4 p = d.pos[b];
    if (!(fabsf(p.x) <= 1e30f) || !(fabsf(p.y) <= 1e30f)) continue;  // NaN or far out: an edge cell
    lx = fminf(lx, p.x);
    ly = fminf(ly, p.y);
    hx = fmaxf(hx, p.x);
    hy = fmaxf(hy, p.y);
    ++ndyn;
  }
  for (int off = kWave / 2; off > 0; off >>= 1) {
    lx = fminf(lx, __shfl_xor(lx, off, kWave));
    ly = fminf(ly, __shfl_xor(ly, off, kWave));
    hx = fmaxf(hx, __shfl_xor(hx, off, kWave));
    hy = fmaxf(hy, __shfl_xor(hy, off, kWave));
    ndyn += __shfl_xor(ndyn, off, kWave);
  }
  if (lane_id() == 0) {
    red[0][wave_id()] = lx;
    red[1][wave_id()] = ly;
    red[2][wave_id()] = hx;
    red[3][wave_id()] = hy;
    redn[wave_id()] = ndyn;
  }
  __syncthreads();
  if (tid == 0) {
    int n = 0;
    for (int w = 0; w < nt / kWave; ++w) {
      lx = fminf(lx, red[0][w]);
      ly = fminf(ly, red[1][w]);
      hx = fmaxf(hx, red[2][w]);
      hy = fmaxf(hy, red[3][w]);
      n += redn[w];
    }
    if (n == 0) lx = ly = hx = hy = 0.0f;
    const float w = fminf(fmaxf(hx - lx, 1e-3f), 1e30f), h = fminf(fmaxf(hy - ly, 1e-3f), 1e30f);
    float L = sqrtf(w * h * (float)kCellTarget / (float)(n > 0 ? n : 1));
    L = fmaxf(L, 1e-3f);
    s_geo[0] = lx;
    s_geo[1] = ly;
    s_geo[2] = w;
    s_geo[3] = h;
    s_int[0] = __float_as_int(L);
  }
  __syncthreads();
  // grow the cell until both grids fit kCells cells and no cell is over full
  int all_cross = 1;
  for (int attempt = 0; attempt < 24; ++attempt) {
    float L = __int_as_float(s_int[0]);
    // (in floats: the cell counts are converted once they are known to be small)
    while ((s_geo[2] / L + 2.0f) * (s_geo[3] / L + 2.0f) > (float)kCells) L *= 1.25f;
    const int gx = (int)(s_geo[2] / L) + 1, gy = (int)(s_geo[3] / L) + 1;
    const PartGrid pg{s_geo[0], s_geo[1], 1.0f / L, gx, gy, 0};
    for (int c = tid; c < 2 * kCells; c += nt) hist[c] = 0;
    __syncthreads();
    for (int b = tid; b < d.nb; b += nt) {
      if (!part_dynamic(d, b) || !world_active(d, d.bworld[b])) continue;
      const float4 p = d.pos[b];
      atomicAdd(&hist[pg.cell_a(p)], 1);
      atomicAdd(&hist[kCells + pg.cell_b(p)], 1);
    }
    __syncthreads();
    int over = 0;
    for (int c = tid; c < 2 * kCells; c += nt) over |= hist[c] > kCellBodies;
    over = __syncthreads_or(over);
    if (!over) {
      all_cross = 0;
      if (tid == 0) {
        d.part_geo[0] = make_float4(pg.lx, pg.ly, pg.inv_l, L);
        d.part_geo[1] = make_float4(as_float(gx), as_float(gy), as_float(0), 0.0f);
      }
      break;
    }
    if (tid == 0) s_int[0] = __float_as_int(L * 1.25f);
    __syncthreads();
  }
  if (all_cross) {  // no grid keeps the cells small enough: every constraint goes to C
    if (tid == 0) {
      d.part_geo[0] = make_float4(0.0f, 0.0f, 1.0f, 1.0f);
      d.part_geo[1] = make_float4(as_float(1), as_float(1), as_float(1), 0.0f);
    }
    for (int c = tid; c < 2 * kCells; c += nt) hist[c] = 0;
    __syncthreads();
  }
  // offsets of the cells' body lists (A cells, then B cells)
  constexpr int kPer = 2 * kCells / kPartGridThreads;
  static_assert(kPer * kPartGridThreads == 2 * kCells, "two cells per thread");
  int v[kPer], s = 0;
  for (int k = 0; k < kPer; ++k) {
    v[k] = hist[tid * kPer + k];
    s += v[k];
  }
  int total;
  int run = part_block_scan(s, scan_lds, &total);
  for (int k = 0; k < kPer; ++k) {
    d.part_boff[tid * kPer + k] = run;
    hist[tid * kPer + k] = run;  // becomes the fill cursor
    run += v[k];
  }
  if (tid == 0) d.part_boff[2 * kCells] = total;
  __syncthreads();
  if (all_cross) return;
  const PartGrid pg = part_grid(d);  // written by thread 0 above, visible after the barrier
  for (int b = tid; b < d.nb; b += nt) {
    if (!part_dynamic(d, b) || !world_active(d, d.bworld[b])) {
      d.part_brank[b] = make_int2(-1, -1);
      continue;
    }
    const float4 p = d.pos[b];
    const int ca = pg.cell_a(p), cb = kCells + pg.cell_b(p);
    const int ra = atomicAdd(&hist[ca], 1), rb = atomicAdd(&hist[cb], 1);
    d.part_blist[ra] = b;
    d.part_blist[rb] = b;
    d.part_brank[b] = make_int2(ra - d.part_boff[ca], rb - d.part_boff[cb]);
  }
}

// Group of each solved constraint (compacted slot k), into its group's bucket.
__global__ void __launch_bounds__(kBlock) k_part_classify(Dev d) {
  const int n = d.cnt[C_CONS];
  const PartGrid pg = part_grid(d);
  B2G_FOR(k, 0, n) {
    const int2 bp = d.cbody[k];
    const int ia = bp.x >= 0 ? bp.x : ~bp.x, ib = bp.y >= 0 ? bp.y : ~bp.y;
    const bool da = part_dynamic(d, ia), db = part_dynamic(d, ib);
    int g = kCrossGroup;
    if (!pg.all_cross) {
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
    }
    if (g != kCrossGroup) {
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

// one workgroup per cell of either grid
__global__ void __launch_bounds__(kCellThreads) k_part_colour_cells(Dev d) {
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
__global__ void __launch_bounds__(kCrossThreads) k_part_colour_cross(Dev d) {
  __shared__ int hist[kColourSlots + 1];
  __shared__ int cursor[kColourSlots];
  __shared__ int scan_lds[16];
  const int start = part_group_start(d, kCrossGroup, scan_lds);
  const int n = part_group_size(d, kCrossGroup);
  if ((int)threadIdx.x == 0) d.cnt[C_PART_CROSS] = n;
  part_colour_group(d, kCrossGroup, start, n, CrossIds{d}, GlobalClaims{d.part_claim, d.part_used}, d.part_xcol,
                    hist, cursor);
}

// Body access of a cell's sweep: the constraint's dynamic bodies in the
// cell's LDS stage, a static or kinematic one (read only) in global memory.
struct CellBodies {
  float4* lds;
  float4* glob;
  int ia;
  int2 loc;
  __device__ __forceinline__ float4* operator()(int i) const {
    const int l = i == ia ? loc.x : loc.y;
    return l >= 0 ? lds + l : glob + i;
  }
  __device__ __forceinline__ bool writes(const Dev& d, int i) const { return writes_body(d, i); }
};

enum PartOp : int { kPartWarm = 0, kPartVelocity = 1, kPartPosition = 2 };

// One phase of a sweep: every cell of the phase's grid (a workgroup each; a
// workgroup past the grid's cells has nothing to do), or the cross group.
template <int kOp, int kThreads>
__global__ void __launch_bounds__(kThreads) k_part_sweep(Dev d, int phase, int it) {
  __shared__ int off[kColourSlots + 1];
  __shared__ float4 s_body[kCellBodies];
  const int g = phase == 2 ? kCrossGroup : phase * kCells + (int)blockIdx.x;
  const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
  const int* coff = d.part_coff + (size_t)g * (kColourSlots + 1);
  for (int c = tid; c <= kColourSlots; c += nt) off[c] = coff[c];
  __syncthreads();
  if (off[kColourSlots] <= off[0]) return;  // the whole workgroup
  float4* body = kOp == kPartPosition ? d.pos : d.vel;
  const bool staged = phase < 2;
  int b0 = 0, b1 = 0;
  if (staged) {
    b0 = d.part_boff[g];
    b1 = d.part_boff[g + 1];
    for (int r = tid; r < b1 - b0; r += nt) s_body[r] = body[d.part_blist[b0 + r]];
    __syncthreads();
  }
  auto solve = [&](int s, bool valid) {
    const int2 loc = valid ? (staged ? d.sloc[s] : make_int2(-1, -1)) : make_int2(-1, -1);
    const int ia = valid ? as_int(d.vc[2][s].x) : 0;
    const CellBodies bv{s_body, body, ia, loc};
    if (kOp == kPartWarm) {
      if (valid) warm_start_t(d, s, bv);
    } else if (kOp == kPartVelocity) {
      if (valid) contact_velocity_t(d, s, bv);
    } else {
      const int isl = valid ? d.cisland[s] : 0;
      const float m = valid ? contact_position_t(d, s, isl, it, bv) : 0.0f;
      mark_island_sep(d, it, isl, valid, m);
    }
  };
  for (int c = 0; c < kColourSlots; ++c) {
    const int lo = off[c], hi = off[c + 1];
    if (hi <= lo) continue;  // uniform across the workgroup
    if (c == kSerialColour) {
      if (tid == 0) {
        for (int s = lo; s < hi; ++s) {
          const int2 loc = staged ? d.sloc[s] : make_int2(-1, -1);
          const CellBodies bv{s_body, body, as_int(d.vc[2][s].x), loc};
          if (kOp == kPartWarm) {
            warm_start_t(d, s, bv);
          } else if (kOp == kPartVelocity) {
            contact_velocity_t(d, s, bv);
          } else {
            const int isl = d.cisland[s];
            const float m = contact_position_t(d, s, isl, it, bv);
            if (separation_bad(m)) d.sepbad[island_slot(d, isl, it)] = 1u;
          }
        }
      }
    } else {
      for (int base = lo; base < hi; base += nt) solve(base + tid, base + tid < hi);
    }
    __syncthreads();
  }
  if (staged)
    for (int r = tid; r < b1 - b0; r += nt) body[d.part_blist[b0 + r]] = s_body[r];
}

}  // namespace b2g
