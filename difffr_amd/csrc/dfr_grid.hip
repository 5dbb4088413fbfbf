// Counting-sort cell grid and compact (transposed ELL) neighbour lists.
//
// Replaces CompactNSearch::find_neighbors / z_sort as used by
// Simulation::performNeighborhoodSearch(Sort) (SPlisHSPlasH/Simulation.cpp:743-777)
// and TimeStepDiffDFSPH::performNeighborhoodSearch (DiffDFSPH/TimeStepDiffDFSPH.cpp:2044-2056).
//
// Every step the fluid is counting-sorted by cell key (brick-Morton order) and
// physically permuted, so that the gather passes read neighbours from nearby
// addresses.  Inside a cell particles are ordered by particle id, which makes
// the order (and hence every FP32 sum) independent of atomic scheduling.
// Boundary samples are sorted as an index list only (their data never moves).
#include "dfr_kernels.h"

namespace dfr {

// Cell key and a slot within the cell.  The fluid arrives in the previous
// step's cell order, so a wavefront's keys form runs of equal values; each run
// takes its slots with one atomic (the leader's), the other lanes add their
// rank in the run.  Slot order within a cell is then fixed by k_cell_sort.
// The fluid and the boundary samples are sorted by the same launches: the
// blocks of a grid over particles split at S.split between the two sets, the
// per-cell kernels take the set from blockIdx.y.
__device__ inline const SortSet& set_of_block(const SortSets& S, uint32_t b, uint32_t& lb) {
  const int k = b >= S.split ? 1 : 0;
  lb = b - (k ? S.split : 0u);
  return S.s[k];
}

__global__ void k_keys(SortSets S, GridDesc g) {
  uint32_t lb;
  const SortSet& T = set_of_block(S, xcd_block(), lb);
  const uint32_t i = lb * blockDim.x + threadIdx.x;
  const bool valid = i < T.n;
  uint32_t k = 0xffffffffu;  // lanes past the end form their own run and take no slot
  if (valid) {
    const float4 p = T.pos[i];
    int cx, cy, cz;
    cell_coords(g, p.x, p.y, p.z, cx, cy, cz);
    k = cell_key(g, cx, cy, cz);
    T.key[i] = k;
  }
  const int lane = (int)(threadIdx.x & 63u);
  const uint32_t kprev = __shfl_up(k, 1, 64);
  const unsigned long long heads = __ballot(lane == 0 || kprev != k);
  const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane
  const int first = 63 - __clzll(heads & upto);
  const unsigned long long after = heads & ~upto;
  const int end = after ? __ffsll((long long)after) - 1 : 64;
  uint32_t base = 0;
  if (valid && lane == first) base = atomicAdd(&T.count[k], (uint32_t)(end - first));
  base = __shfl(base, first, 64);
  if (valid) T.slot[i] = base + (uint32_t)(lane - first);
}

__global__ void k_scatter(SortSets S) {
  uint32_t lb;
  const SortSet& T = set_of_block(S, xcd_block(), lb);
  const uint32_t i = lb * blockDim.x + threadIdx.x;
  if (i >= T.n) return;
  T.perm[T.start[T.key[i]] + T.slot[i]] = i;
}

// Sort of each cell's entries by an order key (particle id for the fluid,
// sample index for the boundary; keys are unique, so the result does not
// depend on the slot order k_keys produced).  Cells hold about ten particles:
// up to 16 entries are loaded together and sorted in registers by a bitonic
// network; larger cells use an insertion sort.
constexpr int CELL_NET = 16;
__global__ void k_cell_sort(SortSets S, uint32_t nkeys) {
  const SortSet& T = S.s[blockIdx.y];
  const uint32_t* __restrict__ start = T.start;
  uint32_t* __restrict__ perm = T.perm;
  const uint32_t* __restrict__ order = T.order;
  const uint32_t c = gid();
  if (c >= nkeys) return;
  const uint32_t s = start[c], e = start[c + 1];
  const uint32_t m = e - s;
  if (m < 2) return;
  if (m <= (uint32_t)CELL_NET) {
    uint32_t v[CELL_NET], kv[CELL_NET];
#pragma unroll
    for (int a = 0; a < CELL_NET; ++a) v[a] = ((uint32_t)a < m) ? perm[s + a] : 0u;
#pragma unroll
    for (int a = 0; a < CELL_NET; ++a) kv[a] = ((uint32_t)a < m) ? (order ? order[v[a]] : v[a]) : 0xffffffffu;
#pragma unroll
    for (int kk = 2; kk <= CELL_NET; kk <<= 1)
#pragma unroll
      for (int j = kk >> 1; j > 0; j >>= 1)
#pragma unroll
        for (int a = 0; a < CELL_NET; ++a) {
          const int b = a ^ j;
          if (b > a) {
            const bool asc = (a & kk) == 0;
            const bool sw = asc ? (kv[a] > kv[b]) : (kv[a] < kv[b]);
            const uint32_t ka = kv[a], kb = kv[b], va = v[a], vb = v[b];
            kv[a] = sw ? kb : ka; kv[b] = sw ? ka : kb;
            v[a] = sw ? vb : va; v[b] = sw ? va : vb;
          }
        }
#pragma unroll
    for (int a = 0; a < CELL_NET; ++a)
      if ((uint32_t)a < m) perm[s + a] = v[a];
    return;
  }
  for (uint32_t a = s + 1; a < e; ++a) {
    const uint32_t v = perm[a];
    const uint32_t kv = order ? order[v] : v;
    uint32_t b = a;
    while (b > s) {
      const uint32_t u = perm[b - 1];
      const uint32_t ku = order ? order[u] : u;
      if (ku <= kv) break;
      perm[b] = u;
      --b;
    }
    perm[b] = v;
  }
}

// Permute the per-particle fluid state into sorted order.
__global__ void k_gather_fluid(int n, const uint32_t* __restrict__ perm, const float4* __restrict__ pos,
                               const float4* __restrict__ vel, const float* __restrict__ kappa,
                               const float* __restrict__ kappav, const uint32_t* __restrict__ id,
                               float4* __restrict__ pos2, float4* __restrict__ vel2, float* __restrict__ kappa2,
                               float* __restrict__ kappav2, uint32_t* __restrict__ id2) {
  const uint32_t k = gid();
  if (k >= (uint32_t)n) return;
  const uint32_t i = perm[k];
  pos2[k] = pos[i];
  vel2[k] = vel[i];
  kappa2[k] = kappa[i];
  kappav2[k] = kappav[i];
  id2[k] = id[i];
}

// ---------------------------------------------------------------------------
// Exclusive scan of the per-cell counts of each set (blockIdx.y, or blockIdx.x
// for the tile sums): three passes, 2048 keys per block.  The last pass also
// clears the counts for the next sort.
// ---------------------------------------------------------------------------
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = DFR_BLOCK * SCAN_ITEMS;

__global__ void k_scan_tiles(SortSets S, uint32_t n) {
  const SortSet& T = S.s[blockIdx.y];
  __shared__ uint32_t lds[DFR_BLOCK / 64 + 1];
  const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = (base + k < n) ? T.count[base + k] : 0u;
    sum += v[k];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(sum, lds, total);
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    if (base + k < n) T.start[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) T.tile_sums[blockIdx.x] = total;
}

__global__ void k_scan_sums(SortSets S, uint32_t m, uint32_t n) {
  const SortSet& T = S.s[blockIdx.x];
  uint32_t* __restrict__ sums = T.tile_sums;
  __shared__ uint32_t lds[DFR_BLOCK / 64 + 1];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < m; base += DFR_BLOCK) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = (i < m) ? sums[i] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(v, lds, total);
    if (i < m) sums[i] = ex + carry;
    carry += total;
  }
  if (threadIdx.x == 0) T.start[n] = carry;
}

__global__ void k_scan_add(SortSets S, uint32_t n) {
  const SortSet& T = S.s[blockIdx.y];
  const uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  const uint32_t add = T.tile_sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k)
    if (base + k < n) {
      T.start[base + k] += add;
      T.count[base + k] = 0u;
    }
}

// Fluid neighbours of the samples of dynamic bodies (the B->F lists the
// gradient loop of computeRigidBodyGradient walks, TimeStepDiffDFSPH.cpp:1246-1273).
// One wavefront per dynamic sample.  The candidates of its 27 cells (cells
// in (dz, dy, dx) order, particles of a cell in sorted order) are numbered
// consecutively; each round tests 64 of them, one per lane, and the hits are
// placed by their rank in a ballot, so the list comes out in candidate order,
// as a sequential scan would write it, while the loads of a round are in
// flight together.
constexpr int BF_WAVES = 4;
__global__ void __launch_bounds__(64 * BF_WAVES)
k_build_bf_nbrs(const float4* __restrict__ bpos, const uint32_t* __restrict__ dyn_idx, int nd,
                const float4* __restrict__ pos, GridDesc g, const uint32_t* __restrict__ fstart, float h2, int maxbf,
                uint32_t* __restrict__ bf, uint32_t* __restrict__ bfcnt, uint32_t* __restrict__ overflow) {
  __shared__ uint32_t cst[BF_WAVES][27], cpre[BF_WAVES][28];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t k = blockIdx.x * BF_WAVES + w;
  const bool valid = k < (uint32_t)nd;  // uniform over the wavefront
  float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
  if (valid) p = bpos[dyn_idx[k]];
  if (lane < 27) {
    uint32_t st = 0, cnt = 0;
    if (valid) {
      int cx, cy, cz;
      cell_coords(g, p.x, p.y, p.z, cx, cy, cz);
      const int x = cx + lane % 3 - 1, y = cy + (lane / 3) % 3 - 1, z = cz + lane / 9 - 1;
      if (x >= 0 && y >= 0 && z >= 0 && x < g.n[0] && y < g.n[1] && z < g.n[2]) {
        const uint32_t c = cell_key(g, x, y, z);
        st = fstart[c];
        cnt = fstart[c + 1] - st;
      }
    }
    cst[w][lane] = st;
    cpre[w][lane + 1] = cnt;
  }
  __syncthreads();
  if (lane == 0) {
    cpre[w][0] = 0;
    for (int c = 1; c <= 27; ++c) cpre[w][c] += cpre[w][c - 1];
  }
  __syncthreads();
  if (!valid) return;
  const uint32_t total = cpre[w][27];
  uint32_t n = 0;
  for (uint32_t base = 0; base < total; base += 64) {
    const uint32_t t = base + (uint32_t)lane;
    bool hit = false;
    uint32_t j = 0;
    if (t < total) {
      int lo = 0, hi = 26;  // the cell holding candidate t: cpre[c] <= t < cpre[c + 1]
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cpre[w][mid] <= t) lo = mid; else hi = mid - 1;
      }
      j = cst[w][lo] + (t - cpre[w][lo]);
      const float4 q = pos[j];
      hit = dist2_exact(q.x, q.y, q.z, p.x, p.y, p.z) < h2;
    }
    const unsigned long long m = __ballot(hit);
    const uint32_t at = n + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (hit && at < (uint32_t)maxbf) bf[(size_t)at * nd + k] = j;
    n += (uint32_t)__popcll(m);
  }
  if (lane == 0) {
    if (n > (uint32_t)maxbf) { atomicMax(overflow, n); n = maxbf; }
    bfcnt[k] = n;
  }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(size_t n, unsigned b = DFR_BLOCK) { return (unsigned)((n + b - 1) / b); }

size_t scan_tile_count(uint32_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

// Sorts the sets of S (nsets = 1 or 2) by cell.  The counts must be zero on
// entry (they are left zero).
static void launch_sort_sets(const SortSets& S, int nsets, const GridDesc& g, hipStream_t s) {
  const uint32_t nk = g.nkeys;
  const unsigned particle_blocks = S.split + (nsets > 1 ? blocks_for(S.s[1].n) : 0u);
  const unsigned tiles = (unsigned)scan_tile_count(nk);
  if (particle_blocks) hipLaunchKernelGGL(k_keys, dim3(particle_blocks), dim3(DFR_BLOCK), 0, s, S, g);
  hipLaunchKernelGGL(k_scan_tiles, dim3(tiles, nsets), dim3(DFR_BLOCK), 0, s, S, nk);
  hipLaunchKernelGGL(k_scan_sums, dim3(nsets), dim3(DFR_BLOCK), 0, s, S, tiles, nk);
  hipLaunchKernelGGL(k_scan_add, dim3(tiles, nsets), dim3(DFR_BLOCK), 0, s, S, nk);
  if (particle_blocks) {
    hipLaunchKernelGGL(k_scatter, dim3(particle_blocks), dim3(DFR_BLOCK), 0, s, S);
    hipLaunchKernelGGL(k_cell_sort, dim3(blocks_for(nk), nsets), dim3(DFR_BLOCK), 0, s, S, nk);
  }
}

static SortSet fluid_set(FluidBufs& f, GridBufs& gb) {
  return SortSet{f.pos, (uint32_t)f.n, gb.fcount, gb.key, gb.slot, gb.fstart, gb.perm, f.id, gb.tile_sums};
}
static SortSet boundary_set(BoundaryBufs& b, GridBufs& gb, uint32_t nk) {
  return SortSet{b.pos, (uint32_t)b.n, gb.bcount, gb.bkey, gb.bslot, gb.bstart, gb.bperm, nullptr,
                 gb.tile_sums + scan_tile_count(nk) + 1};
}

void launch_sort_fluid_and_boundary(FluidBufs& f, BoundaryBufs& b, const GridDesc& g, GridBufs& gb,
                                    hipStream_t s) {
  SortSets S{};
  S.s[0] = fluid_set(f, gb);
  S.s[1] = boundary_set(b, gb, g.nkeys);
  S.split = blocks_for(S.s[0].n);
  launch_sort_sets(S, 2, g, s);
  if (f.n > 0) {
    hipLaunchKernelGGL(k_gather_fluid, dim3(blocks_for(f.n)), dim3(DFR_BLOCK), 0, s, f.n, gb.perm, f.pos, f.vel,
                       f.kappa, f.kappav, f.id, f.pos2, f.vel2, f.kappa2, f.kappav2, f.id2);
    f.swap_sorted();
  }
}

void launch_sort_boundary(BoundaryBufs& b, const GridDesc& g, GridBufs& gb, hipStream_t s) {
  SortSets S{};
  S.s[0] = boundary_set(b, gb, g.nkeys);
  S.split = blocks_for(S.s[0].n);
  launch_sort_sets(S, 1, g, s);
}

void launch_build_bf(FluidBufs& f, BoundaryBufs& b, const GridDesc& g, GridBufs& gb, NbrBufs& nb, float h2,
                     hipStream_t s) {
  if (b.nd > 0) {
    hipLaunchKernelGGL(k_build_bf_nbrs, dim3(blocks_for(b.nd, BF_WAVES)), dim3(64 * BF_WAVES), 0, s, b.pos,
                       b.dyn_idx, b.nd, f.pos, g, gb.fstart, h2, nb.maxbf, nb.bf, nb.bfcnt, nb.overflow + 1);
  }
}

}  // namespace dfr
