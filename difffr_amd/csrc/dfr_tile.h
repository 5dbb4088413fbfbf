// Brick tiles shared by the forward and reverse gather passes.
//
// A workgroup owns one brick of 4x4x4 cells.  After the counting sort the
// brick's particles are contiguous and all their neighbours lie in the brick's
// 6x6x6-cell halo; tile slots number the halo's particles (halo cells in
// (z, y, x) order, particles of a cell in sorted order).  Fluid neighbour
// lists hold 16-bit tile slots, eight per 16-byte pack, packs stored
// pack-major (nbr8[p * N + i]).
#pragma once
#include "dfr_kernels.h"

namespace dfr {

constexpr int HALO = 216;

template <int NT> struct TileHdr {
  uint32_t cstart[HALO];      // first particle (global sorted index) of each halo cell
  uint32_t coff[HALO + 1];    // tile slot of each halo cell's first particle
  uint32_t scan[NT / 64 + 1];
  uint32_t own_begin, own_end, total;
  int bx, by, bz;
};

template <int NT>
__device__ inline void tile_setup(const GridDesc& g, const uint32_t* __restrict__ fstart, uint32_t brick,
                                  TileHdr<NT>& T) {
  const int bx = (int)(brick % (uint32_t)g.bn[0]);
  const int by = (int)((brick / (uint32_t)g.bn[0]) % (uint32_t)g.bn[1]);
  const int bz = (int)(brick / ((uint32_t)g.bn[0] * (uint32_t)g.bn[1]));
  const int t = threadIdx.x;
  uint32_t cnt = 0;
  if (t < HALO) {
    const int cx = 4 * bx - 1 + t % 6, cy = 4 * by - 1 + (t / 6) % 6, cz = 4 * bz - 1 + t / 36;
    uint32_t st = 0;
    if (cx >= 0 && cy >= 0 && cz >= 0 && cx < g.n[0] && cy < g.n[1] && cz < g.n[2]) {
      const uint32_t key = cell_key(g, cx, cy, cz);
      st = fstart[key];
      cnt = fstart[key + 1] - st;
    }
    T.cstart[t] = st;
  }
  uint32_t total;
  const uint32_t ex = block_exclusive_scan<NT>(cnt, T.scan, total);
  if (t < HALO) T.coff[t] = ex;
  if (t == 0) {
    T.coff[HALO] = total;
    T.total = total;
    T.own_begin = fstart[brick * 64u];
    T.own_end = fstart[brick * 64u + 64u];
    T.bx = bx; T.by = by; T.bz = bz;
  }
  __syncthreads();
}

// Global sorted index of a tile slot (slot < total).
template <int NT> __device__ inline uint32_t tile_global(const TileHdr<NT>& T, uint32_t slot) {
  int lo = 0, hi = HALO - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (T.coff[mid] <= slot) lo = mid; else hi = mid - 1;
  }
  return T.cstart[lo] + (slot - T.coff[lo]);
}

// Halo cell index of the cell (cx, cy, cz) + (dx, dy, dz) for a particle of the brick.
template <int NT>
__device__ inline int halo_cell(const TileHdr<NT>& T, int cx, int cy, int cz, int dx, int dy, int dz) {
  return (cz + dz - 4 * T.bz + 1) * 36 + (cy + dy - 4 * T.by + 1) * 6 + (cx + dx - 4 * T.bx + 1);
}

// Two-phase halo staging: every thread first resolves its slots to particle
// indices (LDS-only binary searches, independent), then issues all its global
// loads before writing LDS, so the loads are in flight together.  CAP is the
// largest capacity (registers per thread); slots >= total are skipped.
template <int NT, int CAP = TILE_CAP> constexpr int per_thread() { return (CAP + NT - 1) / NT; }

template <int NT, int CAP = TILE_CAP>
__device__ inline void tile_indices(const TileHdr<NT>& T, uint32_t total, uint32_t* js) {
#pragma unroll
  for (int r = 0; r < per_thread<NT, CAP>(); ++r) {
    const uint32_t s = threadIdx.x + r * NT;
    js[r] = (s < total) ? tile_global<NT>(T, s) : 0u;
  }
}

// Uniform early exit for bricks that own no particle (about half of a tank's
// bricks lie above the free surface): two scalar loads, no halo setup.
__device__ inline bool brick_empty(const uint32_t* __restrict__ fstart, uint32_t brick) {
  return fstart[brick * 64u] == fstart[brick * 64u + 64u];
}

__device__ inline uint32_t pack_get(const uint4& v, int u) {
  const uint32_t w = (u < 2) ? v.x : (u < 4) ? v.y : (u < 6) ? v.z : v.w;
  return (u & 1) ? (w >> 16) : (w & 0xffffu);
}

struct PassArgs {
  StepConst c;
  GridDesc g;
  const uint32_t* fstart;
  const float4* pos;
  float4* vel;
  const float4* bpos;
  const float4* bvel;
  const float4* bx0;
  const uint4* nbr8;
  const uint4* thdr;      // brick headers and halo slot maps written by the build
  const uint32_t* tmap;
  const uint32_t* nbrB;
  const uint32_t* cnt;
  int nb_static;
  int cap;  // halo slots staged in LDS (larger halos read global memory)
};

inline unsigned nbricks(const GridDesc& g) { return (unsigned)(g.bn[0] * g.bn[1] * g.bn[2]); }

inline PassArgs make_args(const StepConst& c, const GridDesc& g, GridBufs& gb, FluidBufs& f, BoundaryBufs& b,
                          NbrBufs& nb) {
  PassArgs a;
  a.c = c; a.g = g; a.fstart = gb.fstart; a.pos = f.pos; a.vel = f.vel; a.bpos = b.pos; a.bvel = b.vel;
  a.bx0 = b.x0; a.nbr8 = nb.nbr8; a.nbrB = nb.nbrB; a.cnt = nb.cnt; a.nb_static = b.n - b.nd;
  a.thdr = nb.thdr; a.tmap = nb.tmap;
  a.cap = nb.cap;
  return a;
}

// A pass's view of its brick.  The build leaves, per brick, a header (own
// particle range, halo total) and the halo's slot -> particle map, so a pass
// starts with one header load and coalesced map loads; only a halo beyond the
// LDS capacity sets up the cell table for the binary-search fallback.
struct BrickView {
  uint32_t own_begin, own_end, total;
  bool in_lds;
};
template <int NT>
__device__ inline bool brick_open(const GridDesc& g, const uint32_t* __restrict__ fstart, const uint4* __restrict__ thdr,
                                  uint32_t brick, int cap, TileHdr<NT>& T, BrickView& B) {
  const uint4 h = thdr[brick];
  if (h.x == h.y) return false;  // no own particle (uniform exit)
  B.own_begin = h.x;
  B.own_end = h.y;
  B.total = h.z;
  B.in_lds = h.z <= (uint32_t)cap;
  if (!B.in_lds) tile_setup<NT>(g, fstart, brick, T);
  return true;
}
template <int NT, int CAP = TILE_CAP>
__device__ inline void map_indices(const uint32_t* __restrict__ tmap, uint32_t brick, uint32_t total, uint32_t* js) {
#pragma unroll
  for (int r = 0; r < per_thread<NT, CAP>(); ++r) {
    const uint32_t s = threadIdx.x + r * NT;
    js[r] = (s < total) ? tmap[(size_t)brick * TILE_CAP + s] : 0u;
  }
}

// Visits the nF fluid neighbour slots of particle i in list order: a loop over
// the packs with the eight slots of a pack unrolled, the pack two ahead in
// flight (a compact loop body keeps the instruction stream small).
template <class Body>
__device__ inline void for_fluid_slots(const uint4* __restrict__ nbr8, size_t stride, uint32_t i, uint32_t nF,
                                       Body&& body) {
  // rows past the last used one are read as that row again (in bounds, in
  // cache) and never used
  const uint32_t np = (nF + PACK - 1) / PACK;
  const uint32_t last = np > 0 ? np - 1 : 0;
  uint4 cur = nbr8[i];
  uint4 nxt = nbr8[(size_t)min(1u, last) * stride + i];
#pragma unroll 1
  for (uint32_t p = 0; p < np; ++p) {
    const uint4 pk = cur;
    cur = nxt;
    nxt = nbr8[(size_t)min(p + 2, last) * stride + i];
    const uint32_t base = p * PACK;
#pragma unroll
    for (int u = 0; u < PACK; ++u)
      if (base + (uint32_t)u < nF) body(pack_get(pk, u));
  }
}

// Branch-free visit of the same slots for LDS-staged halos.  Within a pack the
// slots go in groups of G: `load(s)` of every slot of a group is issued before
// the first `body`, so the group's LDS reads are in flight together, then
// `body(v, live)` runs in slot order.  `live` is false for the slots past nF
// in the last pack; those hold slot 0 (a staged slot, so the read stays in
// bounds), and the body must select their contribution to zero.  A group
// wholly past nF is skipped.
template <int G, class Load, class Body>
__device__ inline void for_fluid_groups(const uint4* __restrict__ nbr8, size_t stride, uint32_t i, uint32_t nF,
                                        Load&& load, Body&& body) {
  static_assert(PACK % G == 0, "the group size divides the pack");
  const uint32_t np = (nF + PACK - 1) / PACK;
  const uint32_t last = np > 0 ? np - 1 : 0;
  uint4 cur = nbr8[i];
  uint4 nxt = nbr8[(size_t)min(1u, last) * stride + i];
#pragma unroll 1
  for (uint32_t p = 0; p < np; ++p) {
    const uint4 pk = cur;
    cur = nxt;
    nxt = nbr8[(size_t)min(p + 2, last) * stride + i];
    const uint32_t rem = nF - p * PACK;  // >= 1
#pragma unroll
    for (int g0 = 0; g0 < PACK; g0 += G) {
      if (g0 > 0 && rem <= (uint32_t)g0) break;
      decltype(load(0u)) v[G];
#pragma unroll
      for (int u = 0; u < G; ++u) v[u] = load(pack_get(pk, g0 + u));
#pragma unroll
      for (int u = 0; u < G; ++u) body(v[u], (uint32_t)(g0 + u) < rem);
    }
  }
}

// Visits the nB boundary neighbours of particle i in list order, G at a time:
// the group's sample indices are loaded first, then the group's sample data
// (`load(j)`), so a group costs two dependent trips to memory instead of two
// per neighbour; `body(v, j)` then runs for the group's neighbours in order.
template <uint32_t G = 4, class Load, class Body>
__device__ inline void for_boundary_groups(const uint32_t* __restrict__ nbrB, size_t stride, uint32_t i,
                                           uint32_t nB, Load&& load, Body&& body) {
#pragma unroll 1
  for (uint32_t k0 = 0; k0 < nB; k0 += G) {
    uint32_t jj[G];
#pragma unroll
    for (uint32_t u = 0; u < G; ++u) jj[u] = nbrB[(size_t)min(k0 + u, nB - 1) * stride + i];
    decltype(load(0u)) v[G];
#pragma unroll
    for (uint32_t u = 0; u < G; ++u) v[u] = load(jj[u]);
#pragma unroll
    for (uint32_t u = 0; u < G; ++u)
      if (k0 + u < nB) body(v[u], jj[u]);
  }
}

// Halo slot records in LDS: the fields a pass gathers per neighbour,
// interleaved so one slot is one or two 16-byte reads.
struct Slot6 {
  float4 a;
  float2 b;
};
struct Slot8 {
  float4 a, b;
};

// Same visit order without unrolling, for the global-memory fallback (a halo
// beyond the LDS capacity): keeps the rare path out of the unrolled hot loop.
template <class Body>
__device__ inline void for_fluid_slots_rolled(const uint4* __restrict__ nbr8, size_t stride, uint32_t i,
                                              uint32_t nF, Body&& body) {
#pragma unroll 1
  for (uint32_t k = 0; k < nF; ++k) {
    const uint4 v = nbr8[(size_t)(k >> 3) * stride + i];
    body(pack_get(v, (int)(k & 7u)));
  }
}

}  // namespace dfr
