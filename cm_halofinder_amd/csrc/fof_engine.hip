// fof_engine.hip -- Friends-of-Friends linking pipeline for gfx950.
//
// Replaces the reference's cell lists + serial sweep + ring merge
// (src_v2/FOF.c:44-161, Befriend :341-376) and computes what the host's
// retrieval / Checkhalo / Output_Halos loops (src_v2/FOF.c:188-203,
// :378-430, :797-847) derive from the ring.
//
// Pipeline (one launch per phase for all volumes of a batch):
//   1 k_pa_count     reference cell of every particle (src_v2/FOF.c:79-83),
//                    per-tile histograms of coarse cell buckets
//   2 scan           (bucket, tile) offsets
//   3 k_pa_scatter, k_pb_scatter, k_pc_sort   two LDS-staged partitions
//                    (coarse, then 256 fine buckets of 2^14 cells), then a
//                    counting sort per fine bucket -> cell starts, dense-cell
//                    lists and float4 {x,y,z,index} in cell order
//                    (sparse addressing: composite keys, radix sort, hash)
//   4 k_subsort      cells with >= max(T, 256) particles: counting sort by
//                    sub-cell; sub-cells have side <= l/sqrt(3), so each is a
//                    clique and its particles start in one union-find set
//                    ("runs")
//   5 k_link_sparse  one thread per particle of a sparse cell: own cell and
//                    the 13 forward cells (src_v2/FOF.c:112-153)
//   6 k_link_dense   dense cells of 256 particles or more, each linked by
//                    one workgroup that took it from a queue: run pairs with
//                    early exit
//     k_link_small   dense cells below 256 particles, one wave each: every
//                    pair tested from LDS broadcasts, components by BFS over
//                    a bit matrix in registers
//   7 union-find finish: flatten, group minimum index, sizes
//   8 groups of two or more ranked by label through a label bitmap, each
//     given a segment of the member list; labels (written by index,
//     straight or through scatter_u32) while every member goes to its
//     group's segment; each segment sorted by index (a wave in registers,
//     or a workgroup)
//   9 ring in ascending index order (written the same way), halo list,
//     FP64 halo sums
//
// Pair rule (src_v2/FOF.c:351-357): dx = fl32(xi - xj) widened to double,
// distsq = (dx^2 + dy^2) + dz^2 in FP64, linked iff !(distsq > linksq).
// dx^2 is exact in FP64, so only the two additions round.  The FP32 test
// d2 = fma(dz,dz,fma(dy,dy,dx*dx)) has relative error < 2e-7; outside the
// band |d2/l2 - 1| >= band (>= 1e-6) it decides exactly as the FP64
// expression would, inside it the FP64 expression is evaluated as written.
#include <math.h>
#include <string.h>

#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <mutex>
#include <type_traits>

#include "fof_internal.h"

namespace fof {

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int vol_of(const uint32_t *__restrict__ vbeg, int nvol,
                                      uint32_t p) {
  int lo = 0, hi = nvol - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (vbeg[mid] <= p) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Volume lookups for waves whose active lanes hold ascending elements (a
// grid-stride loop over 64-aligned runs, a tile, a label-sorted list): the
// volumes of the wave's first and last element are found with scalar loads,
// starting from the previous iteration's volume (hint, wave-uniform), and the
// lanes only search between them -- in a batch nearly every wave lies in one
// volume, and a cold binary search is six dependent loads.
__device__ __forceinline__ int vol_from(const uint32_t *__restrict__ vbeg, int nvol, uint32_t a,
                                        int hint) {
  if (vbeg[hint] <= a && a < vbeg[hint + 1]) return hint;
  return vol_of(vbeg, nvol, a);
}

__device__ __forceinline__ int vol_between(const uint32_t *__restrict__ vbeg, int nvol,
                                           uint32_t p, uint32_t a, uint32_t b, int &hint) {
  const int lo0 = vol_from(vbeg, nvol, a, hint);
  const int hi0 = (b < vbeg[lo0 + 1]) ? lo0 : vol_of(vbeg, nvol, b);
  hint = hi0;
  int lo = lo0, hi = hi0;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (vbeg[mid] <= p) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// lanes hold consecutive elements p (n elements in all)
__device__ __forceinline__ int vol_of_run(const uint32_t *__restrict__ vbeg, int nvol,
                                          uint32_t p, uint32_t n, int &hint) {
  if (nvol == 1) return 0;
  const uint32_t a = __builtin_amdgcn_readfirstlane(p);
  return vol_between(vbeg, nvol, p, a, min(a + 63u, n - 1u), hint);
}

// lanes hold ascending elements p
__device__ __forceinline__ int vol_of_sorted(const uint32_t *__restrict__ vbeg, int nvol,
                                             uint32_t p, int &hint) {
  if (nvol == 1) return 0;
  const uint32_t a = __builtin_amdgcn_readfirstlane(p);
  const int last = 63 - __clzll(__ballot(1));
  return vol_between(vbeg, nvol, p, a, __builtin_amdgcn_readlane(p, last), hint);
}

// Reference cell coordinate along one axis (src_v2/FOF.c:79-81), clamped.
__device__ __forceinline__ int cell1(double x, double rmin_buff, double lcell, int n,
                                     double *u) {
  double t = (x - rmin_buff) / lcell;
  *u = t;
  int i = (int)floor(t);
  return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

struct CellPos {
  int ix, iy, iz;   // reference cell
  int fx, fy, fz;   // fine (sub-cell) coordinate = i*m + s
};

__device__ __forceinline__ int sub1(double u, int i, int m) {
  int s = (int)floor((u - (double)i) * (double)m);
  return s < 0 ? 0 : (s >= m ? m - 1 : s);
}

__device__ __forceinline__ CellPos cell_pos(const VolDev &V, float x, float y, float z) {
  CellPos c;
  double ux, uy, uz;
  c.ix = cell1((double)x, V.rmin_buff[0], V.lcell[0], V.nx, &ux);
  c.iy = cell1((double)y, V.rmin_buff[1], V.lcell[1], V.ny, &uy);
  c.iz = cell1((double)z, V.rmin_buff[2], V.lcell[2], V.nz, &uz);
  if (V.m[0] > 0) {
    c.fx = c.ix * V.m[0] + sub1(ux, c.ix, V.m[0]);
    c.fy = c.iy * V.m[1] + sub1(uy, c.iy, V.m[1]);
    c.fz = c.iz * V.m[2] + sub1(uz, c.iz, V.m[2]);
  } else {
    c.fx = c.ix, c.fy = c.iy, c.fz = c.iz;
  }
  return c;
}

// A float4 load rebuilt from its components: copies of the vector type
// itself into a private array keep the array in scratch memory.
__device__ __forceinline__ float4 ld4(const float4 *p) {
  const float4 v = *p;
  return make_float4(v.x, v.y, v.z, v.w);
}

// Fine (sub-cell) coordinates of a particle of a dense cell, 21 bits per axis,
// recorded once by k_subsort and read by the link kernels.
__device__ __forceinline__ unsigned long long pack_fine(int fx, int fy, int fz) {
  return (unsigned long long)fx | ((unsigned long long)fy << 21) | ((unsigned long long)fz << 42);
}
__device__ __forceinline__ int fine_x(unsigned long long f) { return (int)(f & 0x1FFFFFull); }
__device__ __forceinline__ int fine_y(unsigned long long f) { return (int)((f >> 21) & 0x1FFFFFull); }
__device__ __forceinline__ int fine_z(unsigned long long f) { return (int)((f >> 42) & 0x1FFFFFull); }

__device__ __forceinline__ uint32_t cell_id(const VolDev &V, int ix, int iy, int iz) {
  return V.cbeg + ((uint32_t)iz * (uint32_t)V.ny + (uint32_t)iy) * (uint32_t)V.nx +
         (uint32_t)ix;
}

// Cell addressing.  Dense: a table of cell starts over every cell of every
// volume (cell id = cbeg + linear index).  Sparse: only occupied cells, in
// (volume, linear index) order, found through an open-addressing hash table
// keyed by the composite key (volume << volshift) | linear index.
// A hash slot: the composite key (empty = ~0) and the cell index side by side,
// so that a probe is one 16-byte load (the table is gigabytes: every probe
// misses the caches).
struct __align__(16) HEntry {
  unsigned long long k;
  unsigned long long v;
};

struct Grid {
  const uint32_t *start;           // cell start offsets (+1 entry)
  const struct HEntry *htab;       // sparse: hash table (key, cell index)
  const uint32_t *cell_of;         // sparse: cell of each sorted particle
  unsigned long long hmask;
  int sparse;
  int volshift;
};

__device__ __forceinline__ unsigned long long mix64d(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ unsigned long long comp_key(const Grid &G, const VolDev &V, int ix,
                                                       int iy, int iz) {
  return ((unsigned long long)V.vid << G.volshift) |
         (((unsigned long long)iz * (unsigned)V.ny + (unsigned)iy) * (unsigned)V.nx +
          (unsigned)ix);
}

// Home slot of a key: four consecutive x-cells (keys k with one k >> 2) get
// four consecutive slots, so that the three cells of a neighbour row mostly
// share one 64-byte line of the table.
__device__ __forceinline__ unsigned long long hslot(const Grid &G, unsigned long long k) {
  return ((mix64d(k >> 2) << 2) | (k & 3ull)) & G.hmask;
}

__device__ __forceinline__ uint32_t hlookup(const Grid &G, unsigned long long k) {
  unsigned long long i = hslot(G, k);
  for (;;) {
    const HEntry h = G.htab[i];
    const unsigned long long kk = h.k;
    if (kk == k) return (uint32_t)h.v;
    if (kk == ~0ull) return kNone;
    i = (i + 1) & G.hmask;
  }
}

// Cell index of (ix,iy,iz) (in range), kNone when a sparse cell is empty.
__device__ __forceinline__ uint32_t cell_index(const Grid &G, const VolDev &V, int ix, int iy,
                                               int iz) {
  if (!G.sparse) return cell_id(V, ix, iy, iz);
  return hlookup(G, comp_key(G, V, ix, iy, iz));
}

// Particle range covering cells (x0..x1, iy, iz).  Sparse: the occupied
// cells among consecutive keys are consecutive in the cell list.
__device__ __forceinline__ void row_range(const Grid &G, const VolDev &V, int x0, int x1, int iy,
                                          int iz, uint32_t &r0, uint32_t &r1) {
  if (!G.sparse) {
    const uint32_t c0 = cell_id(V, x0, iy, iz);
    r0 = G.start[c0];
    r1 = G.start[c0 + (uint32_t)(x1 - x0) + 1];
    return;
  }
  uint32_t lo = kNone, hi = kNone;
  const unsigned long long k0 = comp_key(G, V, x0, iy, iz);
  for (int x = 0; x <= x1 - x0; x++) {
    const uint32_t c = hlookup(G, k0 + (unsigned long long)x);
    if (c != kNone) {
      if (lo == kNone) lo = c;
      hi = c;
    }
  }
  if (lo == kNone) {
    r0 = r1 = 0;
    return;
  }
  r0 = G.start[lo];
  r1 = G.start[hi + 1];
}

// Squared linking length of a home cell: constant, or the variable rule
// (same operation order as the oracle; built with -ffp-contract=off).
__device__ __forceinline__ double cell_linksq(const VolDev &V, int ix, int iy, int iz) {
  if (V.vntab == 0) return V.linksq;
  double c0 = V.rmin_buff[0] + ((double)ix + 0.5) * V.lcell[0];
  double c1 = V.rmin_buff[1] + ((double)iy + 0.5) * V.lcell[1];
  double c2 = V.rmin_buff[2] + ((double)iz + 0.5) * V.lcell[2];
  double e0 = c0 - V.vorigin[0], e1 = c1 - V.vorigin[1], e2 = c2 - V.vorigin[2];
  double r2 = e0 * e0 + e1 * e1 + e2 * e2;
  double top = (double)(V.vntab - 1);
  double t = r2 / V.vr2max * top;
  if (!(t < top)) return V.vtab[V.vntab - 1];
  int k = (int)floor(t);
  double f = t - (double)k;
  return V.vtab[k] + f * (V.vtab[k + 1] - V.vtab[k]);
}

struct Rule {
  double linksq;  // FP64 threshold (src_v2/FOF.c:357)
  double prune2;  // squared sub-cell gap beyond which no pair can link
  float lo, hi;   // FP32 decision thresholds
};

__device__ __forceinline__ Rule make_rule(const VolDev &V, int ix, int iy, int iz) {
  Rule r;
  if (V.vntab == 0) {
    r.linksq = V.linksq;
    r.lo = V.lo;
    r.hi = V.hi;
  } else {
    r.linksq = cell_linksq(V, ix, iy, iz);
    r.lo = (float)(r.linksq * (1.0 - (double)V.band));
    r.hi = (float)(r.linksq * (1.0 + (double)V.band));
  }
  r.prune2 = r.linksq * (1.0 + 1e-4);
  return r;
}

// The reference's test in FP64 (src_v2/FOF.c:345-357): dx = fl32(xi - xj),
// distsq summed in double, linked unless distsq > linksq.
__device__ __forceinline__ bool pair_exact(float ax, float ay, float az, float bx, float by,
                                           float bz, double linksq) {
  const float dx = ax - bx, dy = ay - by, dz = az - bz;
  const double ex = dx, ey = dy, ez = dz;
  double ds = ex * ex + ey * ey;
  ds = ds + ez * ez;
  return !(ds > linksq);
}

// The reference predicate with the FP32 fast path.
__device__ __forceinline__ bool pair_linked(const float4 a, const float4 b, const Rule &r,
                                            uint32_t &nre) {
  const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  if (d2 < r.lo) return true;
  if (d2 > r.hi) return false;
  nre++;
  return pair_exact(a.x, a.y, a.z, b.x, b.y, b.z, r.linksq);
}

// Union-find on cell-sorted indices; roots are hooked larger-under-smaller,
// so every root is the smallest sorted index of its set and parent[x] <= x.
// Loads are agent-coherent (skip the non-coherent per-CU cache); a stale
// value can only be an older (larger) ancestor, so walks still terminate.
__device__ __forceinline__ uint32_t ld_par(uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_par(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t uf_find(uint32_t *par, uint32_t x) {
  uint32_t cur = ld_par(par + x);
  if (cur == x) return x;
  uint32_t prev = x, nxt;
  while (cur > (nxt = ld_par(par + cur))) {
    st_par(par + prev, nxt);  // path halving
    prev = cur;
    cur = nxt;
  }
  return cur;
}

// The roots of a and b, the two walks advancing together so that their
// loads overlap (path halving on both).
__device__ __forceinline__ void uf_find2(uint32_t *par, uint32_t &a, uint32_t &b) {
  uint32_t ca = ld_par(par + a), cb = ld_par(par + b);
  uint32_t qa = a, qb = b;
  bool fa = (ca == a), fb = (cb == b);
  while (!fa || !fb) {
    const uint32_t na = fa ? ca : ld_par(par + ca);
    const uint32_t nb = fb ? cb : ld_par(par + cb);
    if (!fa) {
      if (ca > na) st_par(par + qa, na), qa = ca, ca = na;
      else fa = true;
    }
    if (!fb) {
      if (cb > nb) st_par(par + qb, nb), qb = cb, cb = nb;
      else fb = true;
    }
  }
  a = ca, b = cb;
}

__device__ __forceinline__ void uf_union(uint32_t *par, uint32_t a, uint32_t b) {
  for (;;) {
    uf_find2(par, a, b);
    if (a == b) return;
    if (a > b) {
      uint32_t t = a;
      a = b;
      b = t;
    }
    uint32_t old = atomicCAS(par + b, b, a);
    if (old == b) return;
    b = old;
  }
}

// Join j to the set whose root (or an ancestor of it) is rx; returns the
// current root.  Skips the hook when j is already in the set.
__device__ __forceinline__ uint32_t uf_join(uint32_t *par, uint32_t rx, uint32_t j) {
  uint32_t b = uf_find(par, j);
  for (;;) {
    uint32_t a = uf_find(par, rx);
    if (a == b) return a;
    if (a > b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    const uint32_t old = atomicCAS(par + b, b, a);
    if (old == b) return a;
    rx = a;
    b = uf_find(par, old);
  }
}

// Statistics counters are spread over kCtrSlots copies, one 128-byte line
// each, picked by workgroup: a million waves adding to one address serialise
// at one L2 channel.  The host sums the copies.  Counter k of a slot:
constexpr int kCtrTestsSparse = 0;  // pair tests in k_link_sparse
constexpr int kCtrRechecks = 1;     // tests re-decided in FP64 (all kernels)
constexpr int kCtrGroups = 2;       // groups of any size
constexpr int kCtrOwned = 3;        // owned halos
constexpr int kCtrTestsBig = 4;     // pair tests in k_link_dense, 256 threads
constexpr int kCtrTestsSmall = 5;   // pair tests in k_link_small
constexpr int kCtrJoined = 6;       // candidate run pairs found already joined
constexpr int kCtrSlots = 64;
constexpr int kCtrStride = 16;  // u64 per slot
__device__ __forceinline__ void ctr_add(unsigned long long *ctr, int k, unsigned long long v) {
  atomicAdd(ctr + (blockIdx.x & (kCtrSlots - 1)) * kCtrStride + k, v);
}

__device__ __forceinline__ void add_counters(unsigned long long *ctr, int ktests, uint32_t tests,
                                             uint32_t rechecks, uint32_t joined = 0) {
  uint64_t t = tests, r = rechecks, j = joined;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    t += __shfl_down(t, o, 64);
    r += __shfl_down(r, o, 64);
    j += __shfl_down(j, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (t) ctr_add(ctr, ktests, t);
    if (r) ctr_add(ctr, kCtrRechecks, r);
    if (j) ctr_add(ctr, kCtrJoined, j);
  }
}

// Device-side counts and flags of a call (u32 words of the `small` block);
// the counters follow at word kSmCtr.
constexpr int kSmErr = 0;       // bit 0: particle outside [rmin_buff, rmax_buff]
constexpr int kSmNDense = 1;    // dense cells below kBigCell particles
constexpr int kSmM = 2;         // members of groups of two or more (member list)
constexpr int kSmH = 3;         // halos
constexpr int kSmOverflow = 4;  // run-table overflow
constexpr int kSmU = 5;         // occupied cells (sparse addressing)
constexpr int kSmNBig = 7;      // dense cells from kBigCell particles
constexpr int kSmR = 8;         // groups of two or more (their roots)
constexpr int kSmNtB = 10;      // tiles of the fine partition pass
constexpr int kSmDq = 11;       // next cell of k_link_dense's queue
constexpr int kSmCtr = 64;

// Dense cells with at least this many particles are linked by a 256-thread
// workgroup each, smaller ones by a single wave each.
constexpr uint32_t kBigCell = 256;

// ---------------------------------------------------------------------------
// 1-3 (dense grid): counting sort by cell without global atomics.
// Cells are grouped into fine buckets of 2^14 consecutive cells and those into
// coarse buckets of 256 fine buckets.
//   A: per tile of kPartTile particles, cell keys and an LDS histogram of
//      coarse buckets; scan; scatter {key}, {x,y,z,index} by coarse bucket
//   B: the same inside each coarse bucket, by fine bucket (256 ways)
//   C: one workgroup per fine bucket: LDS counters for its 2^14 cells,
//      scan -> cell starts and dense-cell lists, LDS cursors -> cell order
// 256-way partitions keep a tile's open output lines few enough to be
// combined in L2; all counting is in LDS.
// ---------------------------------------------------------------------------
constexpr uint32_t kPartTile = 32768;
constexpr int kFineBits = 14;
constexpr uint32_t kFineCells = 1u << kFineBits;
constexpr int kMaxCoarse = 512;  // C < 2^31 cells

__device__ __forceinline__ uint32_t cell_of_xyz(float x, float y, float z, const VolDev &V,
                                                uint32_t *__restrict__ err) {
  // src_v2/read_data.c:810-812 bounds; non-finite values are rejected too.
  const bool ok = ((double)x >= V.rmin_buff[0]) && ((double)x <= V.rmax_buff[0]) &&
                  ((double)y >= V.rmin_buff[1]) && ((double)y <= V.rmax_buff[1]) &&
                  ((double)z >= V.rmin_buff[2]) && ((double)z <= V.rmax_buff[2]);
  if (!ok) atomicOr(err, 1u);
  double u;
  const int ix = ok ? cell1((double)x, V.rmin_buff[0], V.lcell[0], V.nx, &u) : 0;
  const int iy = ok ? cell1((double)y, V.rmin_buff[1], V.lcell[1], V.ny, &u) : 0;
  const int iz = ok ? cell1((double)z, V.rmin_buff[2], V.lcell[2], V.nz, &u) : 0;
  return cell_id(V, ix, iy, iz);
}

// A, count: keys, validation, hist[b * ntiles + tile] (bucket-major, so one
// exclusive scan gives each (bucket, tile) its first output slot).
__global__ __launch_bounds__(256) void k_pa_count(const uint8_t *__restrict__ base,
                                                  uint32_t stride, uint32_t offpos, uint32_t N,
                                                  const VolDev *__restrict__ vols,
                                                  const uint32_t *__restrict__ vbeg, int nvol,
                                                  uint32_t Bc, uint32_t ntiles,
                                                  uint32_t *__restrict__ key,
                                                  uint32_t *__restrict__ hist,
                                                  uint32_t *__restrict__ err) {
  __shared__ uint32_t h[kMaxCoarse];
  const uint32_t tile = blockIdx.x;
  for (uint32_t b = threadIdx.x; b < Bc; b += blockDim.x) h[b] = 0;
  __syncthreads();
  const uint32_t p0 = tile * kPartTile, p1 = min(N, p0 + kPartTile);
  int hint = 0;
  // kCnt records per thread loaded before any is used; the volume's
  // constants are read once per batch when the batch lies in one volume
  // (read per element, their loads were the pass's critical path: 11 ms
  // against 7.5 at 1024^3)
  constexpr int kCnt = 8;
  for (uint32_t q0 = p0; q0 < p1; q0 += kCnt * 256) {
    float x[kCnt], y[kCnt], z[kCnt];
#pragma unroll
    for (int k = 0; k < kCnt; k++) {
      const uint32_t p = min(q0 + threadIdx.x + k * 256, p1 - 1);
      const float *pp = reinterpret_cast<const float *>(base + (uint64_t)p * stride + offpos);
      x[k] = pp[0], y[k] = pp[1], z[k] = pp[2];
    }
    const uint32_t qe = min(q0 + kCnt * 256, p1) - 1;
    int va = 0;
    bool one = true;
    if (nvol > 1) {
      va = vol_from(vbeg, nvol, q0, hint);
      hint = va;
      one = qe < vbeg[va + 1];
    }
    uint32_t c[kCnt];
    if (one) {
      const VolDev V = vols[va];
#pragma unroll
      for (int k = 0; k < kCnt; k++)
        c[k] = q0 + threadIdx.x + k * 256 < p1 ? cell_of_xyz(x[k], y[k], z[k], V, err) : 0u;
    } else {
#pragma unroll
      for (int k = 0; k < kCnt; k++) {
        const uint32_t p = q0 + threadIdx.x + k * 256;
        c[k] = p < p1 ? cell_of_xyz(x[k], y[k], z[k], vols[vol_of_run(vbeg, nvol, p, N, hint)], err)
                      : 0u;
      }
    }
#pragma unroll
    for (int k = 0; k < kCnt; k++) {
      const uint32_t p = q0 + threadIdx.x + k * 256;
      if (p < p1) {
        key[p] = c[k];
        atomicAdd(&h[c[k] >> (kFineBits + 8)], 1u);
      }
    }
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < Bc; b += blockDim.x) hist[(uint64_t)b * ntiles + tile] = h[b];
}

// Partition scatter through LDS: a sub-tile of kSub elements is laid out in
// bucket order in LDS, then each bucket's elements leave as one contiguous
// run at the bucket's cursor.  cur[] holds the block's global cursors.
// 2048 elements (8 per thread) give each of 256 buckets about 8 per flush,
// so that most output lines are complete before L2 evicts them (1024 was
// 10% slower on the 1024^3 box: partial lines written twice).
constexpr int kSub = 2048;
constexpr int kPartNT = 256;             // threads of the partition scatter kernels
constexpr int kPer = kSub / kPartNT;     // elements a thread holds per sub-tile

template <int NB>
__device__ __forceinline__ void staged_part(uint32_t nsub, const uint32_t (&kk)[kPer],
                                            const float4 (&pp)[kPer],
                                            const uint32_t (&bk)[kPer], uint32_t nbk,
                                            uint32_t *cur, uint32_t *lcnt, uint32_t *loff,
                                            uint32_t *skey, float4 *spos, uint32_t *sbk,
                                            uint32_t *ws, uint32_t *__restrict__ okey,
                                            float4 *__restrict__ opos) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t rank[kPer];
#pragma unroll
  for (int k = 0; k < kPer; k++)
    if (threadIdx.x + k * kPartNT < nsub) rank[k] = atomicAdd(&lcnt[bk[k]], 1u);
  __syncthreads();
  // exclusive scan of lcnt over the buckets (kB consecutive per thread)
  constexpr int kB = (NB + kPartNT - 1) / kPartNT;
  uint32_t c[kB], loc = 0;
#pragma unroll
  for (int q = 0; q < kB; q++) {
    const uint32_t b = threadIdx.x * kB + q;
    c[q] = b < nbk ? lcnt[b] : 0u;
    loc += c[q];
  }
  uint32_t inc = loc;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) ws[w] = inc;
  __syncthreads();
  uint32_t pre = inc - loc;
  for (int q = 0; q < w; q++) pre += ws[q];
#pragma unroll
  for (int q = 0; q < kB; q++) {
    const uint32_t b = threadIdx.x * kB + q;
    if (b < nbk) loff[b] = pre;
    pre += c[q];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kPer; k++)
    if (threadIdx.x + k * kPartNT < nsub) {
      const uint32_t j = loff[bk[k]] + rank[k];
      skey[j] = kk[k];
      spos[j] = pp[k];
      sbk[j] = bk[k];
    }
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < nsub; j += kPartNT) {
    const uint32_t b = sbk[j];
    const uint32_t g = cur[b] + (j - loff[b]);
    okey[g] = skey[j];
    opos[g] = spos[j];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kB; q++) {
    const uint32_t b = threadIdx.x * kB + q;
    if (b < nbk) cur[b] += lcnt[b], lcnt[b] = 0;
  }
  __syncthreads();
}

// A, scatter: keys and {x,y,z,index} to the tile's slots of each coarse bucket.
__global__ __launch_bounds__(kPartNT) void k_pa_scatter(const uint8_t *__restrict__ base,
                                                    uint32_t stride, uint32_t offpos, uint32_t N,
                                                    uint32_t Bc, uint32_t ntiles,
                                                    const uint32_t *__restrict__ key,
                                                    const uint32_t *__restrict__ off,
                                                    uint32_t *__restrict__ okey,
                                                    float4 *__restrict__ opos) {
  __shared__ uint32_t cur[kMaxCoarse], lcnt[kMaxCoarse], loff[kMaxCoarse];
  __shared__ uint32_t skey[kSub], sbk[kSub];
  __shared__ float4 spos[kSub];
  __shared__ uint32_t ws[kPartNT / 64];
  const uint32_t tile = blockIdx.x;
  for (uint32_t b = threadIdx.x; b < kMaxCoarse; b += blockDim.x) {
    cur[b] = b < Bc ? off[(uint64_t)b * ntiles + tile] : 0u;
    lcnt[b] = 0;
  }
  __syncthreads();
  const uint32_t p0 = tile * kPartTile, p1 = min(N, p0 + kPartTile);
  for (uint32_t q0 = p0; q0 < p1; q0 += kSub) {
    const uint32_t nsub = min((uint32_t)kSub, p1 - q0);
    uint32_t kk[kPer], bk[kPer];
    float4 pp[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint32_t i = threadIdx.x + k * kPartNT;
      if (i < nsub) {
        const uint32_t p = q0 + i;
        const uint32_t c = key[p];
        const float *fp = reinterpret_cast<const float *>(base + (uint64_t)p * stride + offpos);
        kk[k] = c;
        bk[k] = c >> (kFineBits + 8);
        pp[k] = make_float4(fp[0], fp[1], fp[2], __uint_as_float(p));
      }
    }
    staged_part<kMaxCoarse>(nsub, kk, pp, bk, Bc, cur, lcnt, loff, skey, spos, sbk, ws, okey,
                            opos);
  }
}

// Coarse bucket starts (Bc + 1 entries) out of the scanned A histogram.
__global__ void k_pa_starts(uint32_t Bc, uint32_t ntiles, const uint32_t *__restrict__ off,
                            uint32_t *__restrict__ cstart) {
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b <= Bc; b += gridDim.x * blockDim.x)
    cstart[b] = off[(uint64_t)b * ntiles];
}

// B: tiles inside coarse buckets.  tinfo[t] = coarse bucket of tile t;
// t0[c] = first tile of bucket c (t0[Bc] = number of tiles, also left in
// *ntb).  Built on the device from the coarse bucket starts; the pass-B
// kernels are launched for the most tiles the buckets can need
// (ntiles_A + Bc) and the blocks past t0[Bc] do no work.
__global__ __launch_bounds__(kMaxCoarse) void k_pb_tiles(uint32_t Bc, uint32_t ntmax,
                                                         const uint32_t *__restrict__ cstart,
                                                         uint32_t *__restrict__ t0,
                                                         uint32_t *__restrict__ tinfo,
                                                         uint32_t *__restrict__ ntb) {
  __shared__ uint32_t ws[kMaxCoarse / 64];
  const uint32_t c = threadIdx.x, lane = c & 63, w = c >> 6;
  const uint32_t nt = c < Bc ? (cstart[c + 1] - cstart[c] + kPartTile - 1) / kPartTile : 0u;
  uint32_t inc = nt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) ws[w] = inc;
  __syncthreads();
  uint32_t pre = inc - nt, tot = 0;
  for (uint32_t q = 0; q < kMaxCoarse / 64; q++) {
    pre += q < w ? ws[q] : 0u;
    tot += ws[q];
  }
  if (c < Bc) {
    t0[c] = pre;
    for (uint32_t k = 0; k < nt; k++) tinfo[pre + k] = c;
  }
  if (c == 0) {
    t0[Bc] = tot;
    *ntb = tot;
  }
  for (uint32_t t = tot + c; t < ntmax; t += blockDim.x) tinfo[t] = kNone;
}
__device__ __forceinline__ void pb_tile(uint32_t t, const uint32_t *__restrict__ tinfo,
                                        const uint32_t *__restrict__ t0,
                                        const uint32_t *__restrict__ cstart, uint32_t &c,
                                        uint32_t &i0, uint32_t &i1, uint32_t &nt, uint32_t &k) {
  c = tinfo[t];
  k = t - t0[c];
  nt = t0[c + 1] - t0[c];
  i0 = cstart[c] + k * kPartTile;
  i1 = min(cstart[c + 1], i0 + kPartTile);
}

__global__ __launch_bounds__(256) void k_pb_count(const uint32_t *__restrict__ tinfo,
                                                  const uint32_t *__restrict__ t0,
                                                  const uint32_t *__restrict__ cstart,
                                                  const uint32_t *__restrict__ ikey,
                                                  uint32_t *__restrict__ hist) {
  __shared__ uint32_t h[256];
  if (tinfo[blockIdx.x] == kNone) {  // past the last tile: zero its histogram slots
    hist[256ull * blockIdx.x + threadIdx.x] = 0;
    return;
  }
  uint32_t c, i0, i1, nt, k;
  pb_tile(blockIdx.x, tinfo, t0, cstart, c, i0, i1, nt, k);
  h[threadIdx.x] = 0;
  __syncthreads();
  for (uint32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x)
    atomicAdd(&h[(ikey[i] >> kFineBits) & 255u], 1u);
  __syncthreads();
  hist[256ull * t0[c] + (uint64_t)threadIdx.x * nt + k] = h[threadIdx.x];
}

__global__ __launch_bounds__(kPartNT) void k_pb_scatter(const uint32_t *__restrict__ tinfo,
                                                    const uint32_t *__restrict__ t0,
                                                    const uint32_t *__restrict__ cstart,
                                                    const uint32_t *__restrict__ off,
                                                    const uint32_t *__restrict__ ikey,
                                                    const float4 *__restrict__ ipos,
                                                    uint32_t *__restrict__ okey,
                                                    float4 *__restrict__ opos) {
  __shared__ uint32_t cur[256], lcnt[256], loff[256];
  __shared__ uint32_t skey[kSub], sbk[kSub];
  __shared__ float4 spos[kSub];
  __shared__ uint32_t ws[kPartNT / 64];
  if (tinfo[blockIdx.x] == kNone) return;  // past the last tile
  uint32_t c, i0, i1, nt, k;
  pb_tile(blockIdx.x, tinfo, t0, cstart, c, i0, i1, nt, k);
  for (uint32_t b = threadIdx.x; b < 256; b += kPartNT) {
    cur[b] = off[256ull * t0[c] + (uint64_t)b * nt + k];
    lcnt[b] = 0;
  }
  __syncthreads();
  for (uint32_t q0 = i0; q0 < i1; q0 += kSub) {
    const uint32_t nsub = min((uint32_t)kSub, i1 - q0);
    uint32_t kk[kPer], bk[kPer];
    float4 pp[kPer];
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      // loads clamped to the sub-tile (the extra elements are masked in
      // staged_part), so the arrays live in registers, not scratch
      const uint32_t i = min(threadIdx.x + q * kPartNT, nsub - 1);
      kk[q] = ikey[q0 + i];
      bk[q] = (kk[q] >> kFineBits) & 255u;
      const float4 v = ipos[q0 + i];
      pp[q] = make_float4(v.x, v.y, v.z, v.w);
    }
    staged_part<256>(nsub, kk, pp, bk, 256, cur, lcnt, loff, skey, spos, sbk, ws, okey, opos);
  }
}

// Fine bucket starts (F + 1 entries).
__global__ void k_pb_starts(uint32_t F, const uint32_t *__restrict__ t0,
                            const uint32_t *__restrict__ cstart, const uint32_t *__restrict__ off,
                            uint32_t N, uint32_t *__restrict__ fstart) {
  for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f <= F; f += gridDim.x * blockDim.x) {
    if (f == F) {
      fstart[f] = N;
      continue;
    }
    const uint32_t c = f >> 8, sub = f & 255u, nt = t0[c + 1] - t0[c];
    fstart[f] = nt ? off[256ull * t0[c] + (uint64_t)sub * nt] : cstart[c];
  }
}

__device__ __forceinline__ int vol_of_cell(const VolDev *__restrict__ vols, int nvol, uint32_t c) {
  int lo = 0, hi = nvol - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (vols[mid].cbeg <= c) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// C: one workgroup per fine bucket, all counting in LDS.  The dense-cell
// lists are gathered per workgroup in LDS and appended with one atomic per
// list (per-wave appends to the two global counters were a third of the
// kernel's time); a workgroup whose lists outgrow LDS appends the rest per
// wave.
constexpr int kPcThreads = 1024;
constexpr uint32_t kPcListS = 1536, kPcListB = 512;

__global__ __launch_bounds__(kPcThreads) void k_pc_sort(uint64_t C, const uint32_t *__restrict__ fstart,
                                                 const uint32_t *__restrict__ ikey,
                                                 const float4 *__restrict__ ipos,
                                                 uint32_t *__restrict__ start,
                                                 float4 *__restrict__ pos4,
                                                 const VolDev *__restrict__ vols, int nvol,
                                                 uint32_t T, uint32_t *__restrict__ dense,
                                                 uint32_t *__restrict__ ndense,
                                                 uint32_t *__restrict__ dbig,
                                                 uint32_t *__restrict__ nbig) {
  // counters padded by one word per 64 so that the per-thread segments of
  // the scan fall in different LDS banks
  __shared__ uint32_t lc[kFineCells + kFineCells / 64 + 1];
  auto P = [](uint32_t e) { return e + (e >> 6); };
  __shared__ uint32_t ws[kPcThreads / 64];
  __shared__ uint32_t lsml[kPcListS], lbig[kPcListB];
  // entries reserved, first reservation that did not fit, global bases
  __shared__ uint32_t nres[2], nfit[2], gbase[2];
  const uint32_t f = blockIdx.x;
  const uint32_t q0 = fstart[f], q1 = fstart[f + 1];
  const uint64_t c0 = (uint64_t)f << kFineBits;
  const uint32_t nc = (uint32_t)min((uint64_t)kFineCells, C - c0);
  for (uint32_t e = threadIdx.x; e < kFineCells + kFineCells / 64 + 1; e += blockDim.x) lc[e] = 0;
  if (threadIdx.x < 2) nres[threadIdx.x] = 0, nfit[threadIdx.x] = threadIdx.x ? kPcListB : kPcListS;
  __syncthreads();
  for (uint32_t i = q0 + threadIdx.x; i < q1; i += blockDim.x)
    atomicAdd(&lc[P(ikey[i] & (kFineCells - 1))], 1u);
  __syncthreads();
  // exclusive scan of the 2^14 counters: kPer consecutive per thread
  constexpr uint32_t kPer = kFineCells / kPcThreads;
  const uint32_t e0 = threadIdx.x * kPer;
  uint32_t loc = 0;
  for (uint32_t q = 0; q < kPer; q++) loc += lc[P(e0 + q)];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t inc = loc;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) ws[w] = inc;
  __syncthreads();
  uint32_t pre = inc - loc;
  for (int q = 0; q < w; q++) pre += ws[q];
  for (uint32_t q = 0; q < kPer; q++) {
    const uint32_t v = lc[P(e0 + q)];
    lc[P(e0 + q)] = pre;
    pre += v;
  }
  if (threadIdx.x == kPcThreads - 1) lc[P(kFineCells)] = pre;  // bucket total
  __syncthreads();
  // one wave's dense cells of one kind: into the LDS list while it fits
  // (the fitting reservations are a prefix of the list), else appended
  auto append = [&](int kind, bool mine, uint32_t cell) {
    const unsigned long long bal = __ballot(mine);
    if (!bal) return;
    const uint32_t cnt = (uint32_t)__popcll(bal), below = (uint32_t)__popcll(bal & ((1ull << lane) - 1));
    const uint32_t cap = kind ? kPcListB : kPcListS;
    uint32_t base = 0;
    if (lane == 0) {
      base = atomicAdd(&nres[kind], cnt);
      if (base + cnt > cap) {
        atomicMin(&nfit[kind], base);
        base = 0x80000000u | atomicAdd(kind ? nbig : ndense, cnt);
      }
    }
    base = __shfl(base, 0, 64);
    if (!mine) return;
    if (base & 0x80000000u) (kind ? dbig : dense)[(base & 0x7FFFFFFFu) + below] = cell;
    else (kind ? lbig : lsml)[base + below] = cell;
  };
  // cell starts (coalesced) and the dense-cell lists
  for (uint32_t e0 = threadIdx.x & ~63u; e0 < nc; e0 += blockDim.x) {
    const uint32_t e = e0 + lane;
    bool sml = false, big = false;
    if (e < nc) {
      const uint32_t s = lc[P(e)], n = lc[P(e + 1)] - s;
      start[c0 + e] = q0 + s;
      if (n >= T) {  // dense: below kBigCell always, from it only with sub-cells
        sml = n < kBigCell;
        big = !sml && vols[vol_of_cell(vols, nvol, (uint32_t)(c0 + e))].m[0] > 0;
      }
    }
    append(0, sml, (uint32_t)(c0 + e));
    append(1, big, (uint32_t)(c0 + e));
  }
  if (c0 + nc == C && threadIdx.x == 0) start[C] = q1;
  __syncthreads();
  if (threadIdx.x < 2) {
    const uint32_t k = threadIdx.x, m = min(nres[k], nfit[k]);
    nfit[k] = m;
    gbase[k] = m ? atomicAdd(k ? nbig : ndense, m) : 0u;
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < nfit[0]; k += blockDim.x) dense[gbase[0] + k] = lsml[k];
  for (uint32_t k = threadIdx.x; k < nfit[1]; k += blockDim.x) dbig[gbase[1] + k] = lbig[k];
  // scatter with LDS cursors
  for (uint32_t i = q0 + threadIdx.x; i < q1; i += blockDim.x) {
    const uint32_t slot = q0 + atomicAdd(&lc[P(ikey[i] & (kFineCells - 1))], 1u);
    pos4[slot] = ipos[i];
  }
}

// 1-3 (sparse addressing): composite keys, sorted, occupied cells.
__global__ __launch_bounds__(256) void k_keys_sparse(const uint8_t *__restrict__ base,
                                                     uint32_t stride, uint32_t offpos, uint32_t N,
                                                     const VolDev *__restrict__ vols,
                                                     const uint32_t *__restrict__ vbeg, int nvol,
                                                     Grid G, uint32_t *__restrict__ klo,
                                                     uint32_t *__restrict__ khi,
                                                     uint32_t *__restrict__ val,
                                                     uint32_t *__restrict__ err) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < N;
       p += gridDim.x * blockDim.x) {
    const VolDev &V = vols[vol_of(vbeg, nvol, p)];
    const float *pp = reinterpret_cast<const float *>(base + (uint64_t)p * stride + offpos);
    const float x = pp[0], y = pp[1], z = pp[2];
    const bool ok = ((double)x >= V.rmin_buff[0]) && ((double)x <= V.rmax_buff[0]) &&
                    ((double)y >= V.rmin_buff[1]) && ((double)y <= V.rmax_buff[1]) &&
                    ((double)z >= V.rmin_buff[2]) && ((double)z <= V.rmax_buff[2]);
    if (!ok) atomicOr(err, 1u);
    double u;
    const int ix = ok ? cell1((double)x, V.rmin_buff[0], V.lcell[0], V.nx, &u) : 0;
    const int iy = ok ? cell1((double)y, V.rmin_buff[1], V.lcell[1], V.ny, &u) : 0;
    const int iz = ok ? cell1((double)z, V.rmin_buff[2], V.lcell[2], V.nz, &u) : 0;
    const unsigned long long k = comp_key(G, V, ix, iy, iz);
    klo[p] = (uint32_t)k;
    khi[p] = (uint32_t)(k >> 32);
    val[p] = p;
  }
}

__global__ __launch_bounds__(256) void k_gather_u32(uint32_t N, const uint32_t *__restrict__ src,
                                                    const uint32_t *__restrict__ idx,
                                                    uint32_t *__restrict__ dst) {
  for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < N; x += gridDim.x * blockDim.x)
    dst[x] = src[idx[x]];
}

__global__ __launch_bounds__(256) void k_scatter_sparse(const uint8_t *__restrict__ base,
                                                        uint32_t stride, uint32_t offpos,
                                                        uint32_t N, const uint32_t *__restrict__ perm,
                                                        const VolDev *__restrict__ vols,
                                                        const uint32_t *__restrict__ vbeg,
                                                        int nvol, Grid G,
                                                        float4 *__restrict__ pos4,
                                                        uint32_t *__restrict__ sklo,
                                                        uint32_t *__restrict__ skhi) {
  for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < N;
       x += gridDim.x * blockDim.x) {
    const uint32_t p = perm[x];
    const float *pp = reinterpret_cast<const float *>(base + (uint64_t)p * stride + offpos);
    const float4 q = make_float4(pp[0], pp[1], pp[2], __uint_as_float(p));
    pos4[x] = q;
    const VolDev &V = vols[vol_of(vbeg, nvol, p)];
    const CellPos c = cell_pos(V, q.x, q.y, q.z);
    const unsigned long long k = comp_key(G, V, c.ix, c.iy, c.iz);
    sklo[x] = (uint32_t)k;
    skhi[x] = (uint32_t)(k >> 32);
  }
}

__global__ __launch_bounds__(256) void k_cell_flags(uint32_t N, const uint32_t *__restrict__ sklo,
                                                    const uint32_t *__restrict__ skhi,
                                                    uint32_t *__restrict__ flags) {
  for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < N;
       x += gridDim.x * blockDim.x)
    flags[x] = (x == 0 || sklo[x] != sklo[x - 1] || skhi[x] != skhi[x - 1]) ? kFlag : 0u;
}

// Per occupied cell: close the start list, record each particle's cell,
// insert the key into the hash table, list dense cells.
__global__ __launch_bounds__(256) void k_cells_finish(
    const uint32_t *__restrict__ dU, uint32_t N, uint32_t *__restrict__ cstart,
    const uint32_t *__restrict__ sklo, const uint32_t *__restrict__ skhi, Grid G,
    HEntry *__restrict__ htab,
    uint32_t *__restrict__ cell_of, const VolDev *__restrict__ vols,
    const uint32_t *__restrict__ vbeg, int nvol, uint32_t T, uint32_t *__restrict__ dense,
    uint32_t *__restrict__ ndense, uint32_t *__restrict__ dbig, uint32_t *__restrict__ nbig) {
  const uint32_t U = *dU;
  const uint32_t lane = threadIdx.x & 63;
  int hint = 0;
  // wave-uniform trip count: the dense-cell lists are appended per wave
  for (uint32_t cb = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); cb < U;
       cb += gridDim.x * blockDim.x) {
    const uint32_t c = cb + lane;
    bool sml = false, big = false;
    if (c < U) {
      const uint32_t s = cstart[c], e = (c + 1 < U) ? cstart[c + 1] : N;
      if (c + 1 == U) cstart[U] = N;
      for (uint32_t x = s; x < e; x++) cell_of[x] = c;
      const unsigned long long k = (unsigned long long)sklo[s] | ((unsigned long long)skhi[s] << 32);
      unsigned long long i = hslot(G, k);
      for (;;) {
        if (atomicCAS(&htab[i].k, ~0ull, k) == ~0ull) {
          htab[i].v = c;
          break;
        }
        i = (i + 1) & G.hmask;
      }
      const VolDev &V = vols[vol_of_sorted(vbeg, nvol, s, hint)];  // s ascends with c
      if (e - s >= T) {  // dense: below kBigCell always, from it only with sub-cells
        sml = e - s < kBigCell;
        big = !sml && V.m[0] > 0;
      }
    }
    const unsigned long long bs = __ballot(sml), bb = __ballot(big);
    const unsigned long long below = (1ull << lane) - 1;
    if (bs) {
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(ndense, (uint32_t)__popcll(bs));
      base = __shfl(base, 0, 64);
      if (sml) dense[base + __popcll(bs & below)] = c;
    }
    if (bb) {
      uint32_t base = 0;
      if (lane == 0) base = atomicAdd(nbig, (uint32_t)__popcll(bb));
      base = __shfl(base, 0, 64);
      if (big) dbig[base + __popcll(bb & below)] = c;
    }
  }
}

__global__ __launch_bounds__(256) void k_init_uf(uint32_t N, uint32_t *__restrict__ par,
                                                 uint32_t *__restrict__ run_end) {
  for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < N;
       x += gridDim.x * blockDim.x) {
    par[x] = x;
    run_end[x] = x + 1;
  }
}

// ---------------------------------------------------------------------------
// 4: sub-cell counting sort of dense cells; runs become cliques.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int sub_bin(const VolDev &V, const float4 p) {
  double ux, uy, uz;
  const int ix = cell1((double)p.x, V.rmin_buff[0], V.lcell[0], V.nx, &ux);
  const int iy = cell1((double)p.y, V.rmin_buff[1], V.lcell[1], V.ny, &uy);
  const int iz = cell1((double)p.z, V.rmin_buff[2], V.lcell[2], V.nz, &uz);
  return (sub1(uz, iz, V.m[2]) * V.m[1] + sub1(uy, iy, V.m[1])) * V.m[0] + sub1(ux, ix, V.m[0]);
}

// Sub-cell counting sort, one workgroup of NT threads per dense cell of the
// list (cells of 256 particles or more; smaller dense cells stay unsorted:
// k_link_small tests all their pairs).
constexpr int kMaxCellBins = 12 * 12 * 12;  // m <= 12 per axis

template <int NT>
__global__ __launch_bounds__(NT) void k_subsort(const uint32_t *__restrict__ dense,
                                                const uint32_t *__restrict__ ndense,
                                                const uint32_t *__restrict__ start,
                                                const VolDev *__restrict__ vols,
                                                const uint32_t *__restrict__ vbeg, int nvol,
                                                float4 *__restrict__ pos4,
                                                float4 *__restrict__ tmp4,
                                                uint32_t *__restrict__ par,
                                                uint32_t *__restrict__ run_end,
                                                unsigned long long *__restrict__ fine,
                                                uint32_t *__restrict__ rtoff,
                                                uint32_t *__restrict__ rtab) {
  constexpr int kWaves = NT / 64;
  constexpr int kPer = (kMaxCellBins + NT - 1) / NT;  // bins per thread in the scans
  __shared__ uint32_t bcur[kMaxCellBins];    // cursors
  __shared__ uint32_t bstart[kMaxCellBins];  // run starts
  __shared__ uint32_t ws[kWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // block-wide exclusive scan of one value per thread; returns (prefix, total)
  auto block_scan = [&](uint32_t loc, uint32_t &tot) -> uint32_t {
    uint32_t inc = loc;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(inc, o, 64);
      if (lane >= o) inc += t;
    }
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    uint32_t off = 0;
    tot = 0;
#pragma unroll
    for (int q = 0; q < kWaves; q++) {
      off += (q < w) ? ws[q] : 0u;
      tot += ws[q];
    }
    return off + inc - loc;
  };
  const uint32_t nd = *ndense;
  for (uint32_t d = blockIdx.x; d < nd; d += gridDim.x) {
    const uint32_t c = dense[d];
    const uint32_t s = start[c], e = start[c + 1], n = e - s;
    const VolDev &V = vols[vol_of(vbeg, nvol, s)];
    if (V.m[0] == 0) continue;  // no sub-cells in this volume (uniform per block)
    const int nbins = V.m[0] * V.m[1] * V.m[2];
    for (int b = threadIdx.x; b < kMaxCellBins; b += NT) bcur[b] = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += NT) atomicAdd(&bcur[sub_bin(V, pos4[s + i])], 1u);
    __syncthreads();
    // exclusive scan of the counts -> starts and cursors (kPer bins a thread)
    {
      const int b0 = threadIdx.x * kPer;
      uint32_t loc = 0, tot;
      for (int q = 0; q < kPer; q++)
        if (b0 + q < nbins) loc += bcur[b0 + q];
      uint32_t pre = block_scan(loc, tot);
      for (int q = 0; q < kPer; q++)
        if (b0 + q < nbins) {
          const uint32_t cc = bcur[b0 + q];
          bstart[b0 + q] = pre;
          bcur[b0 + q] = pre;
          pre += cc;
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
      const float4 p = pos4[s + i];
      tmp4[s + atomicAdd(&bcur[sub_bin(V, p)], 1u)] = p;
    }
    __syncthreads();
    // run table of the cell: heads in sub-cell order at rtab[s ..), their
    // number at rtoff[s] (a cell has at most as many runs as particles)
    {
      const int b0 = threadIdx.x * kPer;
      uint32_t loc = 0, tot;
      for (int q = 0; q < kPer; q++)
        if (b0 + q < nbins && bcur[b0 + q] != bstart[b0 + q]) loc++;
      const uint32_t pre = block_scan(loc, tot);
      if (threadIdx.x == 0) rtoff[s] = tot;
      uint32_t r = s + pre;
      for (int q = 0; q < kPer; q++)
        if (b0 + q < nbins && bcur[b0 + q] != bstart[b0 + q]) rtab[r++] = s + bstart[b0 + q];
    }
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
      const float4 p = tmp4[s + i];
      const CellPos cp = cell_pos(V, p.x, p.y, p.z);
      const int b = ((cp.fz - cp.iz * V.m[2]) * V.m[1] + (cp.fy - cp.iy * V.m[1])) * V.m[0] +
                    (cp.fx - cp.ix * V.m[0]);
      pos4[s + i] = p;
      par[s + i] = s + bstart[b];
      run_end[s + i] = s + bcur[b];
      fine[s + i] = pack_fine(cp.fx, cp.fy, cp.fz);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// 5: sparse cells -- one thread per particle.
// ---------------------------------------------------------------------------
// FAST: every volume of the call has a constant l, a dense grid and
// sub-cells, so only the first branch below is compiled (it then fits 8
// waves per SIMD without spilling); otherwise both branches.  Both run at 8
// waves per SIMD (the full kernel spills a little: at 6 waves without spills
// it was 15% slower on a lightcone).
template <bool FAST>
__global__ __launch_bounds__(256, 8) void k_link_sparse(uint32_t N, const float4 *__restrict__ pos4,
                                                     const Grid G,
                                                     const uint32_t *__restrict__ run_end,
                                                     const unsigned long long *__restrict__ fine,
                                                     uint32_t *__restrict__ par,
                                                     const VolDev *__restrict__ vols,
                                                     const uint32_t *__restrict__ vbeg,
                                                     int nvol, uint32_t T,
                                                     unsigned long long *__restrict__ ctr) {
  uint32_t tests = 0, rechecks = 0;
  int hint = 0;
  for (uint32_t x = blockIdx.x * blockDim.x + threadIdx.x; x < N;
       x += gridDim.x * blockDim.x) {
    const float4 px = pos4[x];
    const VolDev &V = vols[vol_of_run(vbeg, nvol, x, N, hint)];
    const CellPos cp = cell_pos(V, px.x, px.y, px.z);
    const uint32_t c = G.sparse ? G.cell_of[x] : cell_id(V, cp.ix, cp.iy, cp.iz);
    const uint32_t s = G.start[c], e = G.start[c + 1];
    // dense cells are linked by k_link_small (below kBigCell particles) and
    // k_link_dense (from it, with sub-cells)
    if (e - s >= T && (e - s < kBigCell || V.m[0] > 0)) continue;
    const Rule R = make_rule(V, cp.ix, cp.iy, cp.iz);
    uint32_t rx = x;  // x's root, refreshed by every join
    const int nx = V.nx, ny = V.ny, nz = V.nz;
    const int xl = (cp.ix > 0) ? 1 : 0, xr = (cp.ix + 1 < nx) ? 1 : 0;
    if (FAST || (!G.sparse && V.vntab == 0 && V.m[0] > 0)) {
      // Constant l on a dense grid: k_link_dense links every pair that has a
      // particle in a dense cell (its forward cells, and its backward cells
      // that are sparse).  Here: the sparse cells among the own cell (after
      // x) and the 13 forward cells, partners loaded four at a time.
      // Row segments: own row (ix, ix+1); (iy+1, iz) and (iy-1..iy+1, iz+1)
      // over ix-1..ix+1.  b[r][0..3] are the segment's cell boundaries.
      uint32_t b[5][4];
      bool ok[5];
      b[0][0] = x + 1, b[0][1] = e, b[0][2] = xr ? G.start[c + 2] : e, b[0][3] = b[0][2];
      ok[0] = true;
#pragma unroll
      for (int r = 1; r < 5; r++) {
        const int ry = (r == 1) ? cp.iy + 1 : cp.iy + (r - 3);
        const int rz = (r == 1) ? cp.iz : cp.iz + 1;
        ok[r] = ry >= 0 && ry < ny && rz < nz;
        const uint32_t c0 = ok[r] ? cell_id(V, cp.ix - xl, ry, rz) : 0u;
        const int ncell = xl + 1 + xr;
        // the four cell starts from c0 in one load (the table has slack
        // behind its last entry)
        uint32_t q[4];
        __builtin_memcpy(q, G.start + c0, sizeof(q));
        b[r][0] = ok[r] ? q[0] : 0u;
        b[r][1] = ok[r] ? q[1] : 0u;
        b[r][2] = ok[r] ? (ncell >= 2 ? q[2] : q[1]) : 0u;
        b[r][3] = ok[r] ? (ncell >= 3 ? q[3] : b[r][2]) : 0u;
      }
#pragma unroll
      for (int r = 0; r < 5; r++) {
        if (!ok[r]) continue;
        // a dense cell (own row: only ix+1 can be) is skipped
        const bool d0 = r > 0 && b[r][1] - b[r][0] >= T;
        const bool d1 = b[r][2] - b[r][1] >= T;
        const bool d2 = r > 0 && b[r][3] - b[r][2] >= T;
        uint32_t j = b[r][0];
        for (;;) {
          if (d0 && j < b[r][1]) j = b[r][1];
          if (d1 && j >= b[r][1] && j < b[r][2]) j = b[r][2];
          if (d2 && j >= b[r][2]) j = b[r][3];
          if (j >= b[r][3]) break;
          uint32_t je = min(j + 4u, b[r][3]);
          if (d1 && j < b[r][1]) je = min(je, b[r][1]);
          if (d2 && j < b[r][2]) je = min(je, b[r][2]);
          const uint32_t n = je - j;
          float4 q[4];
#pragma unroll
          for (int k = 0; k < 4; k++) q[k] = ((uint32_t)k < n) ? ld4(pos4 + j + k) : px;
          uint32_t hit = 0;
#pragma unroll
          for (int k = 0; k < 4; k++)
            hit |= ((uint32_t)k < n && pair_linked(px, q[k], R, rechecks)) ? (1u << k) : 0u;
          tests += n;
          while (hit) {
            const int k = __builtin_ctz(hit);
            hit &= hit - 1;
            rx = uf_join(par, rx, j + k);
          }
          j = je;
        }
      }
      continue;
    }
    if (FAST) continue;  // not reached: the branch above always continues
    // Up to five contiguous ranges cover the own cell (after x) and the 13
    // forward cells: (ix+1,iy,iz); row (iy+1,iz); rows (iy-1..iy+1, iz+1).
    uint32_t r0[5], r1[5];
    int nr = 0;
    r0[nr] = x + 1;
    if (!G.sparse) {
      r1[nr] = G.start[c + 1 + xr];
    } else {
      r1[nr] = e;
      if (xr) {
        const uint32_t c1 = hlookup(G, comp_key(G, V, cp.ix + 1, cp.iy, cp.iz));
        if (c1 != kNone) r1[nr] = G.start[c1 + 1];
      }
    }
    nr++;
    if (cp.iy + 1 < ny) {
      row_range(G, V, cp.ix - xl, cp.ix + xr, cp.iy + 1, cp.iz, r0[nr], r1[nr]);
      nr++;
    }
    if (cp.iz + 1 < nz) {
      for (int dy = -1; dy <= 1; dy++) {
        if (cp.iy + dy < 0 || cp.iy + dy >= ny) continue;
        row_range(G, V, cp.ix - xl, cp.ix + xr, cp.iy + dy, cp.iz + 1, r0[nr], r1[nr]);
        nr++;
      }
    }
    const bool subcells = V.m[0] > 0;
    for (int r = 0; r < nr; r++) {
      uint32_t j = r0[r];
      const uint32_t je = r1[r];
      if (!subcells || je - j < T) {
        // fewer than T particles: no dense cell here, every particle is a run
        for (; j + 4 <= je; j += 4) {
          float4 q[4];
#pragma unroll
          for (int k = 0; k < 4; k++) q[k] = ld4(pos4 + j + k);
          uint32_t hit = 0;
#pragma unroll
          for (int k = 0; k < 4; k++) hit |= pair_linked(px, q[k], R, rechecks) ? (1u << k) : 0u;
          tests += 4;
          while (hit) {
            const int k = __builtin_ctz(hit);
            hit &= hit - 1;
            rx = uf_join(par, rx, j + k);
          }
        }
        for (; j < je; j++) {
          tests++;
          if (pair_linked(px, pos4[j], R, rechecks)) rx = uf_join(par, rx, j);
        }
        continue;
      }
      while (j < je) {
        const uint32_t re = run_end[j];
        if (re == j + 1) {
          tests++;
          if (pair_linked(px, pos4[j], R, rechecks)) rx = uf_join(par, rx, j);
          j++;
          continue;
        }
        // a clique run of a dense cell: skip it when out of reach or joined
        const unsigned long long fj = fine[j];
        const int gx = max(abs(fine_x(fj) - cp.fx) - 1, 0);
        const int gy = max(abs(fine_y(fj) - cp.fy) - 1, 0);
        const int gz = max(abs(fine_z(fj) - cp.fz) - 1, 0);
        const double g2 = (double)(gx * gx) * V.s2[0] + (double)(gy * gy) * V.s2[1] +
                          (double)(gz * gz) * V.s2[2];
        if (!(g2 > R.prune2) && uf_find(par, rx) != uf_find(par, j)) {
          for (uint32_t k = j; k < re; k++) {
            tests++;
            if (pair_linked(px, pos4[k], R, rechecks)) {
              rx = uf_join(par, rx, k);
              break;
            }
          }
        }
        j = re;
      }
    }
  }
  add_counters(ctr, kCtrTestsSparse, tests, rechecks);
}

// ---------------------------------------------------------------------------
// 6: dense cells -- one workgroup per cell, run pairs with early exit.
// ---------------------------------------------------------------------------
// Run-table capacities.  The home cell has one run per occupied sub-cell,
// at most 12^3 (and at most one per particle); neighbour runs (dense
// neighbours) or particles (sparse neighbours) are taken a chunk at a time.
constexpr int kMaxRuns = 1728;
constexpr uint32_t kMaxT = 512;
// Default dense threshold: on clumpy 1024^3 boxes (b = 0.2, about one
// particle per cell) the link and sub-sort phases take 148-149 ms for T = 16,
// 20 and 24, 153 ms at 32 and 155 ms at 40.
constexpr uint32_t kDefaultT = 24;

__device__ __forceinline__ uint32_t pack_rel(int a, int b, int c) {
  return ((uint32_t)(a & 255)) | ((uint32_t)(b & 255) << 8) | ((uint32_t)(c & 255) << 16);
}
__device__ __forceinline__ int rel_x(uint32_t r) { return (int)(int8_t)(r & 255); }
__device__ __forceinline__ int rel_y(uint32_t r) { return (int)(int8_t)((r >> 8) & 255); }
__device__ __forceinline__ int rel_z(uint32_t r) { return (int)(int8_t)((r >> 16) & 255); }

// Union-find over run slots in LDS (A's runs 0..na-1, the neighbour's runs
// after them).  Hooking larger-under-smaller keeps A's roots inside A, so the
// neighbour's slots can be reset for the next neighbour.
__device__ __forceinline__ uint32_t lfind(volatile uint32_t *lp, uint32_t x) {
  for (;;) {
    const uint32_t p = lp[x];
    if (p == x) return x;
    const uint32_t gp = lp[p];
    if (gp == p) return p;
    lp[x] = gp;
    x = gp;
  }
}

__device__ __forceinline__ void lunion(volatile uint32_t *lp, uint32_t a, uint32_t b) {
  for (;;) {
    a = lfind(lp, a);
    b = lfind(lp, b);
    if (a == b) return;
    if (a > b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    if (atomicCAS((uint32_t *)&lp[b], b, a) == b) return;
  }
}

// Any linking pair between two runs (both cliques)?  Early exit.  Partner
// positions are loaded two at a time: the dense kernels run at 6-8 waves per
// SIMD, where wider batches cost more in registers than they hide latency.
__device__ __forceinline__ bool runs_link(const float4 *__restrict__ pos4, uint32_t ha,
                                          uint32_t ea, uint32_t hb, uint32_t eb, const Rule &R,
                                          uint32_t &tests, uint32_t &rechecks) {
  if (ea == ha + 1 && eb == hb + 1) {  // two single particles: both loads at once
    const float4 a = pos4[ha], b = pos4[hb];
    tests++;
    return pair_linked(a, b, R, rechecks);
  }
  for (uint32_t i = ha; i < ea; i++) {
    const float4 pi = pos4[i];
    uint32_t j = hb;
    for (; j + 2 <= eb; j += 2) {
      float4 q[2];
#pragma unroll
      for (int k = 0; k < 2; k++) q[k] = ld4(pos4 + j + k);
      bool any = false;
#pragma unroll
      for (int k = 0; k < 2; k++) any |= pair_linked(pi, q[k], R, rechecks);
      tests += 2;
      if (any) return true;
    }
    for (; j < eb; j++) {
      tests++;
      if (pair_linked(pi, pos4[j], R, rechecks)) return true;
    }
  }
  return false;
}

// Run pairs with more than BIGPAIR particle pairs (a template parameter of
// k_link_dense) are queued in LDS and tested by 16-lane groups: 16 particle
// pairs per step, ballot exit.  Smaller pairs are tested in place by the
// thread that found them.

// Home-cell sub-cell coordinates (0..11) packed 4 bits per axis.
__device__ __forceinline__ uint16_t pack_home(int x, int y, int z) {
  return (uint16_t)(x | (y << 4) | (z << 8));
}

// Cells with more runs than this enumerate candidates through a bitmap of
// A's occupied sub-cells, one x-row of the window at a time; cells with fewer
// test every run pair.  The 256-thread kernel walks the window twice: first
// the sub-cells that touch the run's own (|offset| <= 1 on every axis), whose
// pairs link most often, so that their merges prune the second walk over the
// rest.
constexpr int kMapRuns = 24;

// 0, 1, -1, 2, -2, ...: window rows visited from the centre out.
__device__ __forceinline__ int zig(int k) { return (k & 1) ? (k + 1) >> 1 : -(k >> 1); }

// Cell offset of neighbour k of the dense kernels' table: k < 13 the 13
// forward cells of src_v2/FOF.c:122-153 ((1,0,0); (-1..1,1,0); (-1..1,-1..1,1)),
// k >= 13 the same offsets negated.
__device__ __forceinline__ void nb_offset(int k, int dd[3]) {
  const int kf = k % 13, sg = k < 13 ? 1 : -1;
  dd[0] = dd[1] = dd[2] = 0;
  if (kf == 0) dd[0] = 1;
  else if (kf < 4) dd[1] = 1, dd[0] = kf - 2;
  else dd[2] = 1, dd[1] = (kf - 4) / 3 - 1, dd[0] = (kf - 4) % 3 - 1;
  dd[0] *= sg, dd[1] *= sg, dd[2] *= sg;
}

// Bytes of dynamic LDS for ra home-run slots and MAXRUNSB neighbour slots:
// ahead, agr, lp (ra + MAXRUNSB), arel (u16).
constexpr size_t dense_lds(uint32_t ra, uint32_t maxrunsb) {
  return 4 * ((size_t)3 * ra + maxrunsb) + 2 * (size_t)ra + 8;
}

// WPE: waves per SIMD to fit (the register budget; LDS per workgroup is
// sized at launch to the runs a cell of the batch can have).
template <int NT, int WPE, int MAXRUNS, int MAXRUNSB, int BIGQ, int ROUND, uint32_t BIGPAIR>
__global__ __launch_bounds__(NT, WPE) void k_link_dense(const uint32_t *__restrict__ dense,
                                                    const uint32_t *__restrict__ ndense,
                                                    const float4 *__restrict__ pos4,
                                                    const Grid G,
                                                    const uint32_t *__restrict__ run_end,
                                                    const unsigned long long *__restrict__ fine,
                                                    const uint32_t *__restrict__ rtoff,
                                                    const uint32_t *__restrict__ rtab,
                                                    uint32_t T,
                                                    uint32_t *__restrict__ par,
                                                    const VolDev *__restrict__ vols,
                                                    const uint32_t *__restrict__ vbeg,
                                                    int nvol,
                                                    unsigned long long *__restrict__ ctr,
                                                    uint32_t *__restrict__ overflow,
                                                    uint32_t ra, uint32_t *__restrict__ dq) {
  // per-run slots of the home cell (ra <= MAXRUNS of them: the most
  // sub-cells of a cell in the batch), in dynamic LDS
  extern __shared__ uint32_t dyn[];
  uint32_t *const ahead = dyn, *const agr = dyn + ra, *const lp = dyn + 2 * ra;
  uint16_t *const arel = reinterpret_cast<uint16_t *>(dyn + 3 * ra + MAXRUNSB);
  __shared__ uint32_t bhead[MAXRUNSB], brel[MAXRUNSB], bgr[MAXRUNSB];
  __shared__ uint32_t bigq[BIGQ];
  // occupied sub-cells of A (<= 12^3) and the occupied count before each
  // word: A's runs are listed in sub-cell order, so the run slot of an
  // occupied sub-cell is its rank in the bitmap
  constexpr int kOccW = kMaxRuns / 64 + 1;
  __shared__ unsigned long long occ[kOccW];
  __shared__ uint16_t wpre[kOccW];
  // after the touching pairs: the dominant LDS set (mainr, kNone if none)
  // and a bitmap of the occupied sub-cells whose runs were outside it
  // ("stragglers"); a pair of two runs of the dominant set is joined, so a
  // run or neighbour item in it needs only the stragglers of a window row
  __shared__ unsigned long long smap[kOccW];
  __shared__ uint32_t mainr;
  __shared__ uint32_t nbS[26], nbOff[27];
  __shared__ uint8_t nbDense[26];
  __shared__ uint32_t cnt, qn, dnext;
  uint32_t tests = 0, rechecks = 0, joined = 0;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t nd = *ndense;
  // cells are taken from a queue (*dq, zero at launch) one at a time: their
  // costs differ by orders of magnitude, and a fixed share per workgroup
  // left a tail (51.2 against 49.4 ms on the clumpy 1024^3 box)
  if (threadIdx.x == 0) dnext = atomicAdd(dq, 1u);
  __syncthreads();
  for (uint32_t d = dnext; d < nd; d = dnext) {
    const uint32_t c = dense[d];
    const uint32_t sA = G.start[c];
    const VolDev &V = vols[vol_of(vbeg, nvol, sA)];
    const unsigned long long f0 = fine[sA];
    const int mm[3] = {V.m[0], V.m[1], V.m[2]};
    const int ix = fine_x(f0) / mm[0], iy = fine_y(f0) / mm[1], iz = fine_z(f0) / mm[2];
    const Rule R = make_rule(V, ix, iy, iz);
    const int ox = ix * mm[0], oy = iy * mm[1], oz = iz * mm[2];
    int reach[3];
    for (int a = 0; a < 3; a++)  // largest sub-cell offset whose gap can be within reach
      reach[a] = min((int)floor(sqrt(R.prune2 / V.s2[a])) + 1, mm[a]);
    // Sub-cell gap pruning in FP32: the prune radius carries a 1e-4 margin,
    // far above the FP32 rounding of the gap.
    const float gs0 = (float)V.s2[0], gs1 = (float)V.s2[1], gs2 = (float)V.s2[2];
    const float gp2 = (float)R.prune2;
    auto gap_ok = [&](int dx, int dy, int dz) {
      const int gx = max(abs(dx) - 1, 0), gy = max(abs(dy) - 1, 0), gz = max(abs(dz) - 1, 0);
      return (float)(gx * gx) * gs0 + (float)(gy * gy) * gs1 + (float)(gz * gz) * gs2 <= gp2;
    };
    // neighbour table: the 13 forward cells of src_v2/FOF.c:122-153 and,
    // for a constant l on a dense grid, the 13 backward cells when sparse
    // (k_link_sparse leaves their pairs with dense cells to this kernel);
    // their items (runs of dense neighbours, particles of sparse ones) and
    // the window of sub-cells that can reach A
    if (threadIdx.x < 64) {  // wave 0: one neighbour a lane, then a scan of the counts
      const int k = threadIdx.x;
      uint32_t items = 0;
      if (k < 26) {
        int dd[3];
        nb_offset(k, dd);
        const int bx = ix + dd[0], by = iy + dd[1], bz = iz + dd[2];
        uint32_t sB = 0, eB = 0;
        const bool want = k < 13 || (!G.sparse && V.vntab == 0);
        if (want && !(bx < 0 || bx >= V.nx || by < 0 || by >= V.ny || bz < 0 || bz >= V.nz)) {
          const uint32_t cb = cell_index(G, V, bx, by, bz);
          if (cb != kNone) sB = G.start[cb], eB = G.start[cb + 1];
        }
        if (k >= 13 && eB - sB >= T) eB = sB;  // a dense backward cell links this pair itself
        // runs of a cell sub-sorted into runs, particles of any other: dense
        // cells of kBigCell particles or more always are; the smaller ones
        // only where k_link_sparse scans runs (sparse addressing or a
        // variable l)
        const bool dn = (eB - sB >= T) && (eB - sB >= kBigCell || G.sparse || V.vntab != 0);
        nbS[k] = sB;
        nbDense[k] = dn ? 1 : 0;
        items = dn ? rtoff[sB] : eB - sB;  // runs of a dense cell, else particles
      }
#pragma unroll
      for (int o = 1; o < 32; o <<= 1) {
        const uint32_t t = __shfl_up(items, o, 64);
        if (k >= o) items += t;
      }
      if (k < 26) nbOff[k + 1] = items;
      if (k == 0) nbOff[0] = 0;
    }
    if (threadIdx.x == 0) qn = 0, mainr = kNone;
    if (threadIdx.x < kOccW) occ[threadIdx.x] = 0, smap[threadIdx.x] = 0;
    __syncthreads();
    const int wy = 2 * reach[1] + 1;
    // the home cell's runs (one per occupied sub-cell, in sub-cell order)
    // from its run table; runs <= min(particles, sub-cells) <= ra
    const int na = (int)rtoff[sA];
    for (uint32_t k = threadIdx.x; k < (uint32_t)na; k += blockDim.x) {
      const uint32_t j = rtab[sA + k];
      const unsigned long long f = fine[j];
      const int rx = fine_x(f) - ox, ry = fine_y(f) - oy, rz = fine_z(f) - oz;
      ahead[k] = j;
      arel[k] = pack_home(rx, ry, rz);
      agr[k] = uf_find(par, j);  // snapshot of the global root
      lp[k] = k;
      const int bin = (rz * mm[1] + ry) * mm[0] + rx;
      atomicOr(&occ[bin >> 6], 1ull << (bin & 63));
    }
    __syncthreads();
    if (threadIdx.x < kOccW) {
      uint32_t n = 0;
      for (uint32_t w = 0; w < threadIdx.x; w++) n += __popcll(occ[w]);
      wpre[threadIdx.x] = (uint16_t)n;
    }
    __syncthreads();
    // run slot of the first occupied sub-cell at or after bin
    auto rank = [&](int bin) -> uint32_t {
      const int w = bin >> 6;
      return wpre[w] + (uint32_t)__popcll(occ[w] & ((1ull << (bin & 63)) - 1));
    };
    // bits of the occupied sub-cells [bin0, bin0 + len) of A (len <= 25)
    auto rowbits = [&](int bin0, int len) -> uint32_t {
      const int w = bin0 >> 6, sh = bin0 & 63;
      unsigned long long v = occ[w] >> sh;
      if (sh + len > 64) v |= occ[w + 1] << (64 - sh);
      return (uint32_t)(v & ((1ull << len) - 1));
    };
    // the same for the stragglers
    auto srowbits = [&](int bin0, int len) -> uint32_t {
      const int w = bin0 >> 6, sh = bin0 & 63;
      unsigned long long v = smap[w] >> sh;
      if (sh + len > 64) v |= smap[w + 1] << (64 - sh);
      return (uint32_t)(v & ((1ull << len) - 1));
    };
    const bool bmap = na > kMapRuns;
    // bits of x in [c - 1, c + 1] for a row starting at x0, given s = c - 1 - x0
    auto touchbits = [](int s) -> uint32_t { return s >= 0 ? 7u << s : 7u >> -s; };
    // a run pair (LDS slots sa, sb with heads ha, hb) not yet joined: small
    // pairs are tested here, big ones go to the wave queue
    auto test_pair = [&](uint32_t sa, uint32_t sb, uint32_t ha, uint32_t hb) {
      const uint32_t ea = run_end[ha], eb = run_end[hb];
      if ((ea - ha) * (eb - hb) > BIGPAIR) {
        const uint32_t q = atomicAdd(&qn, 1u);
        if (q < (uint32_t)BIGQ) {
          bigq[q] = (sa << 16) | sb;
          return;
        }
      }
      if (runs_link(pos4, ha, ea, hb, eb, R, tests, rechecks)) lunion(lp, sa, sb);
    };
    // the queued pairs, four per wave (16-lane groups): skip when joined in
    // LDS, join when under the same global root now, else a group-wide test
    auto drain = [&]() {
      __syncthreads();
      const uint32_t nq = min(qn, (uint32_t)BIGQ);
      const uint32_t grp = lane >> 4, gl = lane & 15;
      for (uint32_t k0 = (threadIdx.x >> 6) * 4; k0 < nq; k0 += (blockDim.x >> 6) * 4) {
        const uint32_t k = k0 + grp;
        uint32_t sa = 0, sb = 0, ha = 0, hb = 0;
        if (k < nq) {
          const uint32_t e = bigq[k];
          sa = e >> 16, sb = e & 0xFFFFu;
          ha = ahead[sa], hb = sb < (uint32_t)na ? ahead[sb] : bhead[sb - na];
        }
        uint32_t v = 2u;  // 2: nothing to do, 1: join, 0: test
        if (k < nq && gl == 0) {
          if (lfind(lp, sa) != lfind(lp, sb)) v = (uf_find(par, ha) == uf_find(par, hb)) ? 1u : 0u;
          joined += (v != 0u) ? 1u : 0u;
        }
        v = __shfl(v, lane & ~15u, 64);
        const uint32_t sza = v == 0u ? run_end[ha] - ha : 0u;
        const uint32_t szb = v == 0u ? run_end[hb] - hb : 1u;
        const uint32_t total = sza * szb;
        bool linked = (v == 1u), done = (v != 0u);
        for (uint32_t base = 0;; base += 16) {
          bool l = false;
          const uint32_t idx = base + gl;
          if (!done && idx < total) {
            const uint32_t i = idx / szb, j = idx - i * szb;
            l = pair_linked(pos4[ha + i], pos4[hb + j], R, rechecks);
            tests++;
          }
          const uint32_t gbits = (uint32_t)(__ballot(l) >> (lane & ~15u)) & 0xFFFFu;
          if (!done && gbits) linked = true, done = true;
          if (base + 16 >= total) done = true;
          if (!__ballot(!done)) break;
        }
        if (linked && gl == 0) lunion(lp, sa, sb);
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        qn = 0;
        if (mainr != kNone) mainr = lfind(lp, mainr);  // merges may hook it
      }
      __syncthreads();
    };
    // a candidate pair of slots: skip if joined, join on equal snapshots,
    // else test
    auto consider = [&](uint32_t sa, uint32_t sb, uint32_t ga, uint32_t gb, uint32_t ha,
                        uint32_t hb) {
      if (lfind(lp, sa) == lfind(lp, sb)) {
        joined++;
        return;
      }
      if (ga == gb) {
        joined++;
        lunion(lp, sa, sb);
      } else {
        test_pair(sa, sb, ha, hb);
      }
    };
    // (A, A), in rounds each followed by a drain, so that merges found early
    // prune the later rounds (a full queue falls back to testing in place)
    if (bmap) {
      // items (row of the forward window, run), row-major, rows centre-out;
      // the first walk (touching sub-cells: the 5 forward rows at |dy|, dz
      // <= 1) before the second (every forward row)
      const int ry = reach[1], rz = reach[2];
      const uint32_t nrow1 = 5, nrow2 = (uint32_t)(ry + 1 + rz * wy);
      auto aa_items = [&](bool first) {
      const uint32_t to = (first ? nrow1 : nrow2) * (uint32_t)na;
      for (uint32_t base = 0; base < to; base += ROUND) {
        const uint32_t pe = min(to, base + (uint32_t)ROUND);
        const bool masked = mainr != kNone;
        for (uint32_t p = base + threadIdx.x; p < pe; p += blockDim.x) {
          const uint32_t r2 = p / (uint32_t)na, ia = p - r2 * (uint32_t)na;
          int dz, dy;
          if (first) {
            dz = r2 < 2 ? 0 : 1;
            dy = zig((int)(r2 < 2 ? r2 : r2 - 2));
          } else if (r2 <= (uint32_t)ry) {
            dz = 0;
            dy = (int)r2;
          } else {
            const uint32_t t = r2 - (uint32_t)(ry + 1);
            dz = 1 + (int)(t / (uint32_t)wy);
            dy = zig((int)(t % (uint32_t)wy));
          }
          const bool trow = dz <= 1 && dy >= -1 && dy <= 1;
          const uint32_t ra = arel[ia];
          const int ax = ra & 15, ay = (ra >> 4) & 15, az = ra >> 8;
          const int by = ay + dy, bz = az + dz;
          if (bz >= mm[2] || by < 0 || by >= mm[1] || !gap_ok(0, dy, dz)) continue;
          const int x0 = (dz == 0 && dy == 0) ? ax + 1 : max(ax - reach[0], 0);
          const int x1 = min(ax + reach[0], mm[0] - 1);
          if (x0 > x1) continue;
          const int bin0 = (bz * mm[1] + by) * mm[0] + x0;
          const uint32_t all = rowbits(bin0, x1 - x0 + 1), r0 = rank(bin0);
          uint32_t bits = all;
          if (trow) bits &= first ? touchbits(ax - 1 - x0) : ~touchbits(ax - 1 - x0);
          if (masked) {  // a run of the dominant set: only the stragglers
            const int abin = (az * mm[1] + ay) * mm[0] + ax;
            if (!((smap[abin >> 6] >> (abin & 63)) & 1ull)) bits &= srowbits(bin0, x1 - x0 + 1);
          }
          while (bits) {
            const int b = __builtin_ctz(bits);
            bits &= bits - 1;
            if (!gap_ok(x0 + b - ax, dy, dz)) continue;
            const uint32_t ib = r0 + __popc(all & ((1u << b) - 1));
            const uint32_t s0 = min(ia, ib), s1 = max(ia, ib);
            consider(s0, s1, agr[s0], agr[s1], ahead[s0], ahead[s1]);
          }
        }
        drain();
      }
      };
      aa_items(true);
      // the dominant set: a vote of 64 runs spread over the cell (wave 0);
      // kept when it holds at least a quarter of them
      if (threadIdx.x < 64) {
        const uint32_t r = lfind(lp, (uint32_t)(((uint64_t)threadIdx.x * (uint32_t)na) >> 6));
        uint32_t best = kNone, bestc = 0;
        unsigned long long pend = __ballot(true);
        while (pend) {
          const uint32_t rl = __shfl(r, __ffsll((long long)pend) - 1, 64);
          const unsigned long long m = __ballot(r == rl);
          const uint32_t cnt1 = (uint32_t)__popcll(m);
          if (cnt1 > bestc) bestc = cnt1, best = rl;
          pend &= ~m;
        }
        if (threadIdx.x == 0) mainr = bestc >= 16 ? best : kNone;
      }
      __syncthreads();
      if (mainr != kNone) {
        for (uint32_t k = threadIdx.x; k < (uint32_t)na; k += blockDim.x)
          if (lfind(lp, k) != mainr) {
            const uint32_t rk = arel[k];
            const int bin = ((rk >> 8) * mm[1] + ((rk >> 4) & 15)) * mm[0] + (rk & 15);
            atomicOr(&smap[bin >> 6], 1ull << (bin & 63));
          }
      }
      __syncthreads();
      aa_items(false);
    } else {
      const uint32_t total = (uint32_t)na * (uint32_t)na;  // na <= kMapRuns
      for (uint32_t base = 0; base < total; base += ROUND) {
        const uint32_t pe = min(total, base + (uint32_t)ROUND);
        for (uint32_t p = base + threadIdx.x; p < pe; p += blockDim.x) {
          const uint32_t ia = p / (uint32_t)na, ib = p - ia * (uint32_t)na;
          if (ib <= ia) continue;
          const uint32_t ra = arel[ia], rb = arel[ib];
          if (!gap_ok((int)(rb & 15) - (int)(ra & 15),
                      (int)((rb >> 4) & 15) - (int)((ra >> 4) & 15),
                      (int)(rb >> 8) - (int)(ra >> 8)))
            continue;
          consider(ia, ib, agr[ia], agr[ib], ahead[ia], ahead[ib]);
        }
        drain();
      }
    }
    // (A, B): the neighbours' items in chunks of MAXRUNSB
    const uint32_t ntot = nbOff[26];
    for (uint32_t c0 = 0; c0 < ntot; c0 += MAXRUNSB) {
      const uint32_t c1 = min(ntot, c0 + (uint32_t)MAXRUNSB);
      if (threadIdx.x == 0) cnt = 0;
      __syncthreads();
      for (uint32_t g = c0 + threadIdx.x; g < c1; g += blockDim.x) {
        int q = 0, qh = 25;  // the last neighbour whose items start at or before g
        while (q < qh) {
          const int mid = (q + qh + 1) >> 1;
          if (nbOff[mid] <= g) q = mid;
          else qh = mid - 1;
        }
        const uint32_t item = g - nbOff[q];
        const uint32_t j = nbDense[q] ? rtab[nbS[q] + item] : nbS[q] + item;
        int fx, fy, fz;
        if (nbDense[q]) {
          const unsigned long long f = fine[j];
          fx = fine_x(f), fy = fine_y(f), fz = fine_z(f);
        } else {
          const float4 pj = pos4[j];
          const CellPos cj = cell_pos(V, pj.x, pj.y, pj.z);
          fx = cj.fx, fy = cj.fy, fz = cj.fz;
        }
        // within reach of A: at most reach sub-cells beyond A's face
        const int rr[3] = {fx - ox, fy - oy, fz - oz};
        int dd[3];
        nb_offset(q, dd);
        bool far = false;
        for (int a = 0; a < 3; a++)
          far |= (dd[a] == 1 && rr[a] > mm[a] - 1 + reach[a]) || (dd[a] == -1 && rr[a] < -reach[a]);
        if (far) continue;
        const int rx = rr[0], ry = rr[1], rz = rr[2];
        const uint32_t k = atomicAdd(&cnt, 1u);  // < MAXRUNSB: one per item at most
        bhead[k] = j;
        brel[k] = pack_rel(rx, ry, rz);
        bgr[k] = uf_find(par, j);
        lp[na + k] = na + k;
      }
      __syncthreads();
      const uint32_t nb = cnt;
      // every (home run, neighbour item) pair within reach, in rounds
      if (bmap) {
        // items (row of the window, item), row-major, rows centre-out: the
        // 9 touching rows (touching sub-cells) first, then every row
        const uint32_t nrow = (uint32_t)((2 * reach[2] + 1) * wy);
        const uint32_t total = (9 + nrow) * nb;
        for (uint32_t base = 0; base < total; base += ROUND) {
          const uint32_t pe = min(total, base + (uint32_t)ROUND);
          const uint32_t mr = mainr;
          for (uint32_t p = base + threadIdx.x; p < pe; p += blockDim.x) {
            const uint32_t r2 = p / nb, jb = p - r2 * nb;
            const bool first = r2 < 9;
            const uint32_t ri = first ? r2 : r2 - 9;
            const uint32_t rw = first ? 3u : (uint32_t)wy;
            const int dz = zig((int)(ri / rw)), dy = zig((int)(ri % rw));
            const bool trow = dz >= -1 && dz <= 1 && dy >= -1 && dy <= 1;
            const uint32_t rb = brel[jb];
            const int bx = rel_x(rb), ay = rel_y(rb) + dy, az = rel_z(rb) + dz;
            if (az < 0 || az >= mm[2] || ay < 0 || ay >= mm[1] || !gap_ok(0, dy, dz)) continue;
            const int x0 = max(bx - reach[0], 0), x1 = min(bx + reach[0], mm[0] - 1);
            if (x0 > x1) continue;
            const int bin0 = (az * mm[1] + ay) * mm[0] + x0;
            const uint32_t all = rowbits(bin0, x1 - x0 + 1), r0 = rank(bin0);
            uint32_t bits = all;
            if (trow) bits &= first ? touchbits(bx - 1 - x0) : ~touchbits(bx - 1 - x0);
            // an item already in the dominant set: only the stragglers
            if (mr != kNone && lfind(lp, na + jb) == mr) bits &= srowbits(bin0, x1 - x0 + 1);
            while (bits) {
              const int b = __builtin_ctz(bits);
              bits &= bits - 1;
              if (!gap_ok(x0 + b - bx, dy, dz)) continue;
              const uint32_t ia = r0 + __popc(all & ((1u << b) - 1));
              consider(ia, na + jb, agr[ia], bgr[jb], ahead[ia], bhead[jb]);
            }
          }
          drain();
        }
      } else {
        const uint32_t total = (uint32_t)na * nb;
        for (uint32_t base = 0; base < total; base += ROUND) {
          const uint32_t pe = min(total, base + (uint32_t)ROUND);
          for (uint32_t p = base + threadIdx.x; p < pe; p += blockDim.x) {
            const uint32_t ib = p / (uint32_t)na, ia = p - ib * (uint32_t)na;
            const uint32_t ra = arel[ia], rb = brel[ib];
            if (!gap_ok(rel_x(rb) - (int)(ra & 15), rel_y(rb) - (int)((ra >> 4) & 15),
                        rel_z(rb) - (int)(ra >> 8)))
              continue;
            consider(ia, na + ib, agr[ia], bgr[ib], ahead[ia], bhead[ib]);
          }
          drain();
        }
      }
      // commit: a neighbour run joined in LDS joins its (home-cell) root's run
      for (uint32_t ib = threadIdx.x; ib < nb; ib += blockDim.x) {
        const uint32_t r = lfind(lp, na + ib);
        if (r != na + ib) uf_union(par, bhead[ib], ahead[r]);
      }
      if (c1 < ntot) __syncthreads();  // the slots are reused by the next chunk
    }
    // commit the home cell's runs (all their merges)
    for (int ia = threadIdx.x; ia < na; ia += blockDim.x) {
      const uint32_t r = lfind(lp, ia);
      if (r != (uint32_t)ia) uf_union(par, ahead[ia], ahead[r]);
    }
    if (threadIdx.x == 0) dnext = atomicAdd(dq, 1u);
    __syncthreads();
  }
  add_counters(ctr, kCtrTestsBig, tests, rechecks, joined);
}

// ---------------------------------------------------------------------------
// 6b: small dense cells (T <= n < kBigCell particles) -- one wave per cell,
// every pair tested in registers, components by bit-matrix closure.
// ---------------------------------------------------------------------------
// At about one particle per sub-cell (the usual case for cells below 256
// particles) the sub-cell cliques of k_link_dense save few tests, while their
// run-pair enumeration costs each lane a divergent walk per candidate.  Here
// a wave tests 64 pairs per step with no branch and no memory access: each
// lane holds one particle, the partners are held one per lane and broadcast
// with readlane.  Pairs the FP32 test leaves undecided (within the band
// around l^2) are re-decided in FP64 afterwards.  Links inside the cell form a
// symmetric bit matrix in registers (lane i: rows i, i+64, ...; a ballot gives
// the transposed columns), closed transitively (Warshall: for every k, the
// rows holding k absorb row k), so each particle knows its component without
// one union per link.  Neighbour particles within reach of the cell's box are
// then tested 64 at a time against the cell; each joins the first particle it
// links to, and one that links two components merges them in an LDS
// union-find seeded with the components (the four largest components' masks
// are kept in scalar registers, so a candidate skips whole components).
// Which pairs belong to this kernel is the rule of k_link_dense: the cell's
// own pairs, its 13 forward neighbours (src_v2/FOF.c:122-153) and, for a
// constant l on a dense grid, its backward neighbours that are sparse.
//
// Each wave of a workgroup works on its own cells (no workgroup barriers);
// the neighbours' particles are read as one index space over the 26 ranges,
// two 64-particle chunks in flight, and the next cell's bounds are fetched
// while the current one is linked.
constexpr int kSmallQ = 128;    // neighbour candidates queued in LDS (< 64 + 64)
constexpr int kSmallWaves = 4;  // waves per workgroup
constexpr float kFar = 1e18f;   // position of an absent particle: links nothing

constexpr int kBlk = 4;  // partners per block of a wave's pair tests
struct SmallLds {
  // the cell's particles, x, y and z apart (the w field is not needed; the
  // float4 table cost 6 KB more per workgroup and 6% of the kernel's time),
  // then kBlk absent ones
  float ax[kBigCell + kBlk], ay[kBigCell + kBlk], az[kBigCell + kBlk];
  unsigned long long cm[4][4];       // the four largest components (bit masks)
  uint32_t lp[kBigCell];             // union-find over the cell's components
  float qx[kSmallQ], qy[kSmallQ], qz[kSmallQ];  // queued neighbour candidates
  uint32_t qidx[kSmallQ];
};

// LDS written by some lanes of a wave and read by others afterwards: the
// wave's LDS operations execute in order; this keeps the compiler from
// moving them across.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The lane's particle (xi, yi, zi) against the kBlk partners p[0..kBlk-1]
// in LDS, the same for every lane (broadcast reads, all issued before the
// tests).  Past the cell's last particle the table holds kBlk absent ones at
// -kFar (absent lanes sit at +kFar), so a short last block needs no masking:
// nothing links them.  Returns bit k = linked under the reference's test: the FP32
// decision outside the band; when any lane has a pair inside the band, the
// block is re-decided pair by pair with pair_linked.  With COLS, cols[k] is
// the lane mask of partner k's links (the transposed block).
template <bool COLS>
__device__ __forceinline__ uint32_t testblk(float xi, float yi, float zi,
                                           const float *__restrict__ px,
                                           const float *__restrict__ py,
                                           const float *__restrict__ pz, const Rule &R,
                                           uint32_t &rechecks,
                                           unsigned long long (&cols)[kBlk]) {
  uint32_t lk = 0;
  unsigned long long band = 0;  // lanes with an undecided pair (uniform)
  float qx[kBlk], qy[kBlk], qz[kBlk];
#pragma unroll
  for (int k = 0; k < kBlk; k++) qx[k] = px[k], qy[k] = py[k], qz[k] = pz[k];
#pragma unroll
  for (int k = 0; k < kBlk; k++) {
    const float dx = xi - qx[k], dy = yi - qy[k], dz = zi - qz[k];
    const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const bool l = d2 < R.lo;
    const unsigned long long c = __ballot(l);
    if (COLS) cols[k] = c;
    lk |= l ? (1u << k) : 0u;
    band |= __ballot(d2 <= R.hi) & ~c;
  }
  if (band) {  // rare: re-decide the block exactly
    lk = 0;
#pragma unroll
    for (int k = 0; k < kBlk; k++) {
      const float4 q = make_float4(qx[k], qy[k], qz[k], 0.f);
      const bool l = pair_linked(make_float4(xi, yi, zi, 0.f), q, R, rechecks);
      if (COLS) cols[k] = __ballot(l);
      lk |= l ? (1u << k) : 0u;
    }
  }
  return lk;
}

// The cell's own pairs (NW 64-particle tiles, 64 (NW - 1) < nA <= 64 NW;
// the particles in S.ax, S.ay, S.az): the link matrix, its components (breadth-first,
// by pulling: a particle joins when its row meets the frontier), each
// particle's component root (the lowest member) into lp, and the four
// largest components' masks into S.cm.
template <int NW>
__device__ __forceinline__ void small_cell_own(SmallLds &S, uint32_t nA, const Rule &R,
                                               uint32_t &rechecks) {
  const uint32_t lane = threadIdx.x & 63;
  float xr[NW], yr[NW], zr[NW];
#pragma unroll
  for (int r = 0; r < NW; r++) {
    const uint32_t i = 64 * r + lane;
    const uint32_t ic = i < nA ? i : 0;
    xr[r] = i < nA ? S.ax[ic] : kFar, yr[r] = i < nA ? S.ay[ic] : kFar;
    zr[r] = i < nA ? S.az[ic] : kFar;
  }
  unsigned long long rows[NW][NW];  // rows[r][w]: links of 64 r + lane to 64 w ..
#pragma unroll
  for (int r = 0; r < NW; r++)
#pragma unroll
    for (int w = 0; w < NW; w++) rows[r][w] = 0;
#pragma unroll
  for (int ta = 0; ta < NW; ta++) {
#pragma unroll
    for (int tb = ta; tb < NW; tb++) {
      const uint32_t jn = min(64u, nA - 64 * tb);
      for (uint32_t j0 = 0; j0 < jn; j0 += kBlk) {
        unsigned long long cols[kBlk];
        const uint32_t o = 64 * tb + j0;
        const uint32_t h =
            testblk<true>(xr[ta], yr[ta], zr[ta], S.ax + o, S.ay + o, S.az + o, R, rechecks, cols);
        rows[ta][tb] |= (unsigned long long)h << j0;
        if (tb != ta) {  // the transposed block: rows of 64 tb + j0 + k, word ta
#pragma unroll
          for (int k = 0; k < kBlk; k++)
            if (lane == j0 + k) rows[tb][ta] = cols[k];
        }
      }
    }
  }
  // particles present, and those linked to none (their own component)
  unsigned long long unv[NW];
#pragma unroll
  for (int r = 0; r < NW; r++) {
    const uint32_t n = min(64u, nA - 64 * r);
    const unsigned long long present = n >= 64 ? ~0ull : ((1ull << n) - 1);
    unsigned long long other = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) other |= (w == r) ? (rows[r][w] & ~(1ull << lane)) : rows[r][w];
    const unsigned long long iso = __ballot(other == 0) & present;
    if ((iso >> lane) & 1ull) S.lp[64 * r + lane] = 64 * r + lane;
    unv[r] = present & ~iso;
  }
  uint32_t csz[4] = {0, 0, 0, 0};
  unsigned long long cmask[4][NW];
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int w = 0; w < NW; w++) cmask[c][w] = 0;
  for (;;) {
    int sw = -1;
#pragma unroll
    for (int w = NW - 1; w >= 0; w--)
      if (unv[w]) sw = w;
    if (sw < 0) break;
    unsigned long long comp[NW], F[NW];
#pragma unroll
    for (int w = 0; w < NW; w++) {
      comp[w] = (w == sw) ? (unv[w] & (~unv[w] + 1)) : 0ull;  // the seed: lowest unvisited
      F[w] = comp[w];
      unv[w] &= ~comp[w];
    }
    uint32_t seed = 0;
#pragma unroll
    for (int w = 0; w < NW; w++)
      if (comp[w]) seed = 64 * w + (uint32_t)__builtin_ctzll(comp[w]);
    for (;;) {
      unsigned long long nx[NW], any = 0;
#pragma unroll
      for (int r = 0; r < NW; r++) {
        unsigned long long meet = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) meet |= rows[r][w] & F[w];
        nx[r] = __ballot(meet != 0) & unv[r];
        any |= nx[r];
      }
      if (!any) break;
#pragma unroll
      for (int r = 0; r < NW; r++) comp[r] |= nx[r], unv[r] &= ~nx[r], F[r] = nx[r];
    }
    uint32_t size = 0;
#pragma unroll
    for (int r = 0; r < NW; r++) {
      if ((comp[r] >> lane) & 1ull) S.lp[64 * r + lane] = seed;
      size += (uint32_t)__popcll(comp[r]);
    }
    // keep the four largest (descending)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      if (size > csz[c]) {
        const uint32_t ts = csz[c];
        csz[c] = size;
        size = ts;
#pragma unroll
        for (int w = 0; w < NW; w++) {
          const unsigned long long tm = cmask[c][w];
          cmask[c][w] = comp[w];
          comp[w] = tm;
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int w = 0; w < NW; w++)
      if (lane == 0) S.cm[c][w] = cmask[c][w];
}

// One cell of NW 64-particle tiles (64 (NW - 1) < nA <= 64 NW): its own
// pairs, its neighbours' particles, the commit.  nxt: the wave's next cell
// (kNone: none), whose bounds are fetched into nS, nE on the way.
template <int NW>
__device__ __forceinline__ void small_cell(SmallLds &S, uint32_t sA, uint32_t nA,
                                           const VolDev &V, const CellPos &cp, const Rule &R,
                                           uint32_t nbS, uint32_t nbN,
                                           const float4 *__restrict__ pos4, const Grid &G,
                                           uint32_t *__restrict__ par, uint32_t nxt,
                                           uint32_t &nS, uint32_t &nE, uint32_t &tests,
                                           uint32_t &rechecks) {
  const uint32_t lane = threadIdx.x & 63;
  volatile uint32_t *const lp = S.lp;
  // (A, A)
  small_cell_own<NW>(S, nA, R, rechecks);
  tests += (lane == 0) ? nA * (nA - 1) / 2 : 0u;
  wave_sync();
  // the next cell's bounds (its index has arrived by now)
  if (nxt != kNone) nS = G.start[nxt], nE = G.start[nxt + 1];
  // (A, B): the 26 ranges as one index space: off = exclusive prefix
  uint32_t off = nbN;
#pragma unroll
  for (int o = 1; o < 32; o <<= 1) {
    const uint32_t t = __shfl_up(off, o, 64);
    if (lane >= (uint32_t)o) off += t;
  }
  const uint32_t total = __shfl(off, 25, 64);
  off -= nbN;
  // the bounding box of the cell's particles: a neighbour particle farther
  // than l from it links to none of them (prune2 carries a 1e-4 margin on
  // l^2; the gap is evaluated in FP64 from the float coordinates)
  double lo[3], hi[3];
  {
    float mn[3] = {kFar, kFar, kFar}, mx[3] = {-kFar, -kFar, -kFar};
#pragma unroll
    for (int r = 0; r < NW; r++) {
      const uint32_t i = 64 * r + lane;
      if (i < nA) {
        const float qx = S.ax[i], qy = S.ay[i], qz = S.az[i];
        mn[0] = fminf(mn[0], qx), mn[1] = fminf(mn[1], qy), mn[2] = fminf(mn[2], qz);
        mx[0] = fmaxf(mx[0], qx), mx[1] = fmaxf(mx[1], qy), mx[2] = fmaxf(mx[2], qz);
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
      for (int a = 0; a < 3; a++) {
        mn[a] = fminf(mn[a], __shfl_xor(mn[a], o, 64));
        mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], o, 64));
      }
#pragma unroll
    for (int a = 0; a < 3; a++) lo[a] = mn[a], hi[a] = mx[a];
  }
  // index g of the space -> particle, for the 64 consecutive indices g0 +
  // lane: only the ranges from the one holding g0 to the one holding the
  // chunk's last index are scanned
  auto locate = [&](uint32_t g0) -> uint32_t {
    const uint32_t g = g0 + lane, gl = min(g0 + 63u, total - 1u);
    const unsigned long long la = __ballot(lane < 26 && off <= g0);
    const unsigned long long lb = __ballot(lane < 26 && off <= gl);
    const int ka = 63 - __clzll(la), kb = 63 - __clzll(lb);
    uint32_t base = (uint32_t)__builtin_amdgcn_readlane((int)nbS, ka) -
                    (uint32_t)__builtin_amdgcn_readlane((int)off, ka);
    for (int k = ka + 1; k <= kb; k++) {
      const uint32_t ok = (uint32_t)__builtin_amdgcn_readlane((int)off, k);
      const uint32_t sk = (uint32_t)__builtin_amdgcn_readlane((int)nbS, k);
      if (g >= ok) base = sk - ok;
    }
    return base + g;
  };
  uint32_t qn = 0;  // uniform
  auto flush = [&](uint32_t cnt) {
    wave_sync();
    const bool vb = lane < cnt;
    const uint32_t lq = vb ? lane : 0;
    const uint32_t jb = S.qidx[lq];
    const float bx = vb ? S.qx[lq] : kFar, by = vb ? S.qy[lq] : kFar, bz = vb ? S.qz[lq] : kFar;
    tests += vb ? nA : 0u;
    unsigned long long m[NW];
#pragma unroll
    for (int w = 0; w < NW; w++) {
      const uint32_t jn = min(64u, nA - 64 * w);
      unsigned long long mw = 0;
      for (uint32_t j0 = 0; j0 < jn; j0 += kBlk) {
        unsigned long long cols[kBlk];
        const uint32_t o = 64 * w + j0;
        mw |= (unsigned long long)testblk<false>(bx, by, bz, S.ax + o, S.ay + o, S.az + o, R,
                                                 rechecks, cols)
              << j0;
      }
      m[w] = vb ? mw : 0ull;
    }
    // join the first linked particle; every further component linked
    // merges with its component
    bool any = false;
#pragma unroll
    for (int q = 0; q < NW; q++) any |= m[q] != 0;
    if (any) {
      int i0 = -1;
#pragma unroll
      for (int q = NW - 1; q >= 0; q--)
        if (m[q]) i0 = 64 * q + __builtin_ctzll(m[q]);
      uf_union(par, jb, sA + (uint32_t)i0);
#pragma unroll
      for (int c = 0; c < 4; c++) {
        bool hit = false;
#pragma unroll
        for (int q = 0; q < NW; q++) hit |= (m[q] & S.cm[c][q]) != 0;
        if (!hit) continue;
        int ic0 = -1;
#pragma unroll
        for (int q = NW - 1; q >= 0; q--)
          if (m[q] & S.cm[c][q]) ic0 = 64 * q + __builtin_ctzll(m[q] & S.cm[c][q]);
        lunion(lp, (uint32_t)i0, (uint32_t)ic0);
#pragma unroll
        for (int q = 0; q < NW; q++) m[q] &= ~S.cm[c][q];
      }
      // members of smaller components, one at a time
      for (;;) {
        int i = -1;
#pragma unroll
        for (int q = NW - 1; q >= 0; q--)
          if (m[q]) i = 64 * q + __builtin_ctzll(m[q]);
        if (i < 0) break;
        lunion(lp, (uint32_t)i0, (uint32_t)i);
#pragma unroll
        for (int q = 0; q < NW; q++)
          if (q == (i >> 6)) m[q] &= m[q] - 1;
      }
    }
    wave_sync();
  };
  auto push = [&](float4 pb, uint32_t j, bool valid) {
    bool keep = false;
    if (valid) {
      const double x = pb.x, y = pb.y, z = pb.z;
      const double gx = fmax(fmax(lo[0] - x, x - hi[0]), 0.0);
      const double gy = fmax(fmax(lo[1] - y, y - hi[1]), 0.0);
      const double gz = fmax(fmax(lo[2] - z, z - hi[2]), 0.0);
      keep = !(gx * gx + gy * gy + gz * gz > R.prune2);
    }
    const unsigned long long bal = __ballot(keep);
    if (keep) {
      const uint32_t at = qn + (uint32_t)__popcll(bal & ((1ull << lane) - 1));
      S.qx[at] = pb.x, S.qy[at] = pb.y, S.qz[at] = pb.z;
      S.qidx[at] = j;
    }
    qn += (uint32_t)__popcll(bal);
    if (qn >= 64) {
      flush(64);
      // move the overflow down (qn - 64 < 64 entries)
      const bool mv = lane < qn - 64;
      const uint32_t lm = 64 + (mv ? lane : 0);
      const float tx = S.qx[lm], ty = S.qy[lm], tz = S.qz[lm];
      const uint32_t ti = S.qidx[lm];
      wave_sync();
      if (mv) S.qx[lane] = tx, S.qy[lane] = ty, S.qz[lane] = tz, S.qidx[lane] = ti;
      qn -= 64;
    }
  };
  for (uint32_t g0 = 0; g0 < total; g0 += 128) {
    const uint32_t ga = g0 + lane, gb = g0 + 64 + lane;
    const bool va = ga < total, vb = gb < total;
    const uint32_t ja = locate(g0);
    const uint32_t jb = g0 + 64 < total ? locate(g0 + 64) : 0u;
    const float4 pa = va ? ld4(pos4 + ja) : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 pb = vb ? ld4(pos4 + jb) : make_float4(0.f, 0.f, 0.f, 0.f);
    push(pa, ja, va);
    if (g0 + 64 < total) push(pb, jb, vb);
  }
  if (qn) flush(qn);
  wave_sync();
  // commit the cell's components (and the merges made through
  // neighbours): a particle still its own global root is hooked under its
  // component's root with one CAS (the root is lower: larger-under-
  // smaller), any other joins by union
  for (uint32_t i = lane; i < nA; i += 64) {
    const uint32_t r = lfind(lp, i);
    if (r != i && atomicCAS(par + sA + i, sA + i, sA + r) != sA + i) uf_union(par, sA + i, sA + r);
  }
  wave_sync();
}

// Five waves per SIMD (at most 96 VGPRs, a few spilled).  On the clumpy
// 1024^3 box: 33.4 ms with blocks of 8 partners and no cap (134 VGPRs, three
// waves), 28.3 ms at four waves, 25.0 ms with blocks of 4 at five, 23.5 ms
// with the x, y, z tables; six waves (80 VGPRs) spill enough to take 31.4 ms.
__global__ __launch_bounds__(64 * kSmallWaves) __attribute__((amdgpu_waves_per_eu(5))) void k_link_small(
    const uint32_t *__restrict__ dense, const uint32_t *__restrict__ ndense,
    const float4 *__restrict__ pos4, const Grid G, uint32_t T, uint32_t *__restrict__ par,
    const VolDev *__restrict__ vols, const uint32_t *__restrict__ vbeg, int nvol,
    unsigned long long *__restrict__ ctr) {
  // the wave's index, made visibly uniform so that the cell's bounds, its
  // particles as partners and the scan state live in scalar registers
  const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  __shared__ SmallLds Sall[kSmallWaves];
  SmallLds &S = Sall[wv];
  uint32_t tests = 0, rechecks = 0;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t nd = *ndense;
  const uint32_t w0 = blockIdx.x * kSmallWaves + wv;
  const uint32_t ws = gridDim.x * kSmallWaves;
  // bounds of the wave's next cell, fetched ahead
  uint32_t nS = 0, nE = 0;
  if (w0 < nd) {
    const uint32_t c = dense[w0];
    nS = G.start[c], nE = G.start[c + 1];
  }
  for (uint32_t d = w0; d < nd; d += ws) {
    const uint32_t sA = nS, nA = nE - nS;  // T <= nA < kBigCell
    const uint32_t cN = (d + ws < nd) ? dense[d + ws] : 0u;
    const VolDev &V = vols[vol_of(vbeg, nvol, sA)];
    const float4 p0 = pos4[sA];
    for (uint32_t i = lane; i < nA; i += 64) {
      const float4 q = ld4(pos4 + sA + i);
      S.ax[i] = q.x, S.ay[i] = q.y, S.az[i] = q.z;
    }
    if (lane < (uint32_t)kBlk) S.ax[nA + lane] = -kFar, S.ay[nA + lane] = -kFar, S.az[nA + lane] = -kFar;
    const CellPos cp = cell_pos(V, p0.x, p0.y, p0.z);
    const Rule R = make_rule(V, cp.ix, cp.iy, cp.iz);
    // neighbour k of the table (lane k): its particle range
    uint32_t nbS = 0, nbN = 0;
    if (lane < 26) {
      int dd[3];
      nb_offset((int)lane, dd);
      const int bx = cp.ix + dd[0], by = cp.iy + dd[1], bz = cp.iz + dd[2];
      const bool want = lane < 13 || (!G.sparse && V.vntab == 0 && V.m[0] > 0);
      if (want && !(bx < 0 || bx >= V.nx || by < 0 || by >= V.ny || bz < 0 || bz >= V.nz)) {
        const uint32_t cb = cell_index(G, V, bx, by, bz);
        if (cb != kNone) nbS = G.start[cb], nbN = G.start[cb + 1] - nbS;
      }
      if (lane >= 13 && nbN >= T) nbN = 0;  // a dense backward cell links this pair itself
    }
    wave_sync();
    const uint32_t nw = (nA + 63) >> 6;
    const uint32_t nxt = (d + ws < nd) ? cN : kNone;
    if (nw <= 1) small_cell<1>(S, sA, nA, V, cp, R, nbS, nbN, pos4, G, par, nxt, nS, nE, tests, rechecks);
    else if (nw == 2) small_cell<2>(S, sA, nA, V, cp, R, nbS, nbN, pos4, G, par, nxt, nS, nE, tests, rechecks);
    else if (nw == 3) small_cell<3>(S, sA, nA, V, cp, R, nbS, nbN, pos4, G, par, nxt, nS, nE, tests, rechecks);
    else small_cell<4>(S, sA, nA, V, cp, R, nbS, nbN, pos4, G, par, nxt, nS, nE, tests, rechecks);
  }
  add_counters(ctr, kCtrTestsSmall, tests, rechecks);
}

// ---------------------------------------------------------------------------
// 7: union-find finish
// ---------------------------------------------------------------------------
// Particle ID of input record p (the reference's P[p].ID, src_v2/vars.h:36-47).
__device__ __forceinline__ unsigned long long rec_id(const uint8_t *__restrict__ base,
                                                     uint32_t stride, uint32_t offid, uint32_t p) {
  const uint32_t *q = reinterpret_cast<const uint32_t *>(base + (uint64_t)p * stride + offid);
  return (unsigned long long)q[0] | ((unsigned long long)q[1] << 32);
}

// Flatten every particle to its root; the other members of a group
// accumulate their minimum input index and their count at the root (minorig
// and gsize start at ~0 and 0; the root itself is added in k_label).  With
// idroot, the members' minimum particle ID is accumulated the same way.
__global__ __launch_bounds__(256) void k_flatten(uint32_t N, uint32_t *__restrict__ par,
                                                 const float4 *__restrict__ pos4,
                                                 uint32_t *__restrict__ minorig,
                                                 uint32_t *__restrict__ gsize,
                                                 unsigned long long *__restrict__ ctr,
                                                 const uint8_t *__restrict__ base, uint32_t stride,
                                                 uint32_t offid,
                                                 unsigned long long *__restrict__ idroot) {
  uint32_t roots = 0;
  const uint32_t lane = threadIdx.x & 63;
  // wave-uniform loop: lanes of a wave reduce together
  for (uint32_t xb = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); xb < N;
       xb += gridDim.x * blockDim.x) {
    const uint32_t x = xb + lane;
    uint32_t r = x;
    bool member = false;
    if (x < N) {
      uint32_t p = par[x];
      while (p != r) {
        r = p;
        p = par[r];
      }
      member = (r != x);
      if (member) par[x] = r;
      else roots++;
    }
    const uint32_t o = member ? __float_as_uint(pos4[x].w) : 0xFFFFFFFFu;
    // members of one group are mostly neighbours in the cell order: one
    // atomicMin / atomicAdd per distinct root in the wave
    unsigned long long pend = __ballot(member);
    while (pend) {
      const int leader = __ffsll((long long)pend) - 1;
      const uint32_t rl = __shfl(r, leader, 64);
      const bool mine = member && r == rl;
      const unsigned long long m = __ballot(mine);
      uint32_t v = mine ? o : 0xFFFFFFFFu;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor(v, off, 64));
      if ((int)lane == leader) {
        atomicMin(minorig + rl, v);
        atomicAdd(gsize + rl, (uint32_t)__popcll(m));
      }
      if (idroot) {
        unsigned long long id = mine ? rec_id(base, stride, offid, o) : ~0ull;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          const unsigned long long t = __shfl_xor(id, off, 64);
          id = t < id ? t : id;
        }
        if ((int)lane == leader) atomicMin(idroot + rl, id);
      }
      pend &= ~m;
    }
  }
  uint64_t t = roots;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
  if ((threadIdx.x & 63) == 0 && t) ctr_add(ctr, kCtrGroups, t);
}

// Members of groups of two or more: label[orig] = group minimum index |
// kFlag, and each member's (label, orig) at the next free slot of its group
// in the member list (the group's cursor, indexed by its label, starts
// at the group's first slot; lanes of one group take consecutive slots with
// one atomic).  The head's entry of hsize receives the group size.  Two ways
// to write the index-ordered label array from the cell-ordered particles:
//  DIRECT  random 4-byte writes straight into label -- when every volume's
//          label range is small enough to stay in the last-level cache
//          while its cells are processed (the 4^3-batched 1024^3 box: 11 ms
//          less than the partitioned writes); with label_out given (no
//          minimum-ID output wanted) straight into the caller's output, as
//          volume-local labels over the defaults k_label_init wrote;
//  else    scatter_u32 from the member list afterwards, partitioned by index
//          (one 1.09e9-particle volume: 4 ms less than direct writes spread
//          over 4.4 GB).
constexpr uint32_t kLabTile = 4096;
// Largest volume (particles) written directly: 128 MB of labels.
constexpr uint64_t kDirectMax = 1ull << 25;

template <bool DIRECT>
__global__ __launch_bounds__(256) void k_label(uint32_t N, const uint32_t *__restrict__ par,
                                               const float4 *__restrict__ pos4,
                                               const uint32_t *__restrict__ minorig,
                                               const uint32_t *__restrict__ gsize,
                                               uint32_t *__restrict__ out,
                                               uint32_t *__restrict__ hsize,
                                               uint32_t *__restrict__ cursor,
                                               uint32_t *__restrict__ mkey,
                                               uint32_t *__restrict__ mval,
                                               const uint8_t *__restrict__ base, uint32_t stride,
                                               uint32_t offid,
                                               const unsigned long long *__restrict__ idroot,
                                               unsigned long long *__restrict__ idhead, int single,
                                               const VolDev *__restrict__ vols,
                                               const uint32_t *__restrict__ vbeg, int nvol,
                                               uint32_t *__restrict__ label_out) {
  const uint32_t lane = threadIdx.x & 63;
  const unsigned long long lt = (1ull << lane) - 1;
  const uint64_t x0 = (uint64_t)blockIdx.x * kLabTile;
  // the tile's volume base when the tile lies in one volume (volumes are
  // contiguous in the cell order too), else looked up per particle
  uint32_t btile = kNone;
  if (DIRECT && label_out && x0 < N) {
    const int va = vol_of(vbeg, nvol, (uint32_t)x0);
    if (min((uint64_t)N, x0 + kLabTile) <= vbeg[va + 1]) btile = vbeg[va];
  }
  for (uint32_t q = 0; q < kLabTile; q += 256) {
    const uint64_t xx = x0 + q + threadIdx.x;
    bool mem = false;
    uint32_t o = 0, lab = 0, r = 0;
    if (xx < N) {
      const uint32_t x = (uint32_t)xx;
      r = par[x];
      const uint32_t g = gsize[r];
      // singletons: k_label_init wrote them, unless their volume keeps
      // groups of one (those join the member list)
      if (!(r == x && g == 0) || (single && !(1.0 < vols[vol_of(vbeg, nvol, x)].nphalomin))) {
        mem = true;
        o = __float_as_uint(pos4[x].w);
        const uint32_t orr = (r == x) ? o : __float_as_uint(pos4[r].w);
        lab = min(minorig[r], orr);
        if (DIRECT && label_out)  // straight to the output, volume-local
          label_out[o] = lab - (btile != kNone ? btile : vbeg[vol_of(vbeg, nvol, x)]);
        else if (DIRECT)
          out[o] = lab | kFlag;
        if (o == lab) {
          hsize[lab] = g + 1u;
          if (idhead) {  // the root's own ID joins the members' minimum
            const unsigned long long a = idroot[r], b = rec_id(base, stride, offid, orr);
            idhead[lab] = a < b ? a : b;
          }
        }
      }
    }
    // member-list slots: the wave's first few groups one atomic each (a
    // wave's particles are mostly in one or two groups), any other lane its own
    unsigned long long todo = __ballot(mem);
    for (int it = 0; it < 4 && todo; it++) {
      const int first = __builtin_ctzll(todo);
      const uint32_t l0 = (uint32_t)__builtin_amdgcn_readlane((int)lab, first);
      const bool mine = mem && lab == l0;
      const unsigned long long same = __ballot(mine);
      uint32_t b = 0;
      if ((int)lane == first) b = atomicAdd(cursor + l0, (uint32_t)__popcll(same));
      b = (uint32_t)__builtin_amdgcn_readlane((int)b, first);
      if (mine) {
        const uint32_t slot = b + (uint32_t)__popcll(same & lt);
        mkey[slot] = lab;
        mval[slot] = o;
      }
      todo &= ~same;
    }
    if ((todo >> lane) & 1ull) {
      const uint32_t slot = atomicAdd(cursor + lab, 1u);
      mkey[slot] = lab;
      mval[slot] = o;
    }
  }
}

// Streaming kernels whose elements' volumes are looked up per wave take one
// contiguous tile of kTileN elements per block, so that a wave's next
// elements usually lie in the volume of its last ones (the hint of
// vol_of_run); a grid-stride loop would jump across volumes every step.
constexpr uint32_t kTileN = 4096;
static inline uint32_t tiles_for(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, (n + kTileN - 1) / kTileN); }

// Every particle as a singleton (label = own index, group size 1, kept only
// when Nphalomin <= 1); k_label overwrites the members of larger groups.
__global__ __launch_bounds__(256) void k_label_init(uint32_t N, const VolDev *__restrict__ vols,
                                                    const uint32_t *__restrict__ vbeg, int nvol,
                                                    uint32_t *__restrict__ label,
                                                    uint32_t *__restrict__ hsize,
                                                    uint32_t *__restrict__ label_out,
                                                    uint32_t *__restrict__ next_out) {
  const uint64_t i0 = (uint64_t)blockIdx.x * kTileN, i1 = min((uint64_t)N, i0 + kTileN);
  if (i0 >= i1) return;
  // internal defaults (label, hsize) or, when k_label writes label_out
  // itself, the outputs' defaults: own label and own ring entry, 1-based
  auto body = [&](uint32_t i, double nphalomin, uint32_t b) {
    if (label) label[i] = i | (!(1.0 < nphalomin) ? kFlag : 0u);
    if (hsize) hsize[i] = 1u;
    if (label_out) label_out[i] = i - b;
    if (next_out) next_out[i] = i - b + 1u;
  };
  const int va = vol_of(vbeg, nvol, (uint32_t)i0);
  if (i1 <= vbeg[va + 1]) {  // the tile lies in one volume
    const double nph = vols[va].nphalomin;
    const uint32_t b = vbeg[va];
    for (uint32_t i = (uint32_t)i0 + threadIdx.x; i < i1; i += blockDim.x) body(i, nph, b);
    return;
  }
  int hint = va;
  for (uint32_t i = (uint32_t)i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const int v = vol_of_run(vbeg, nvol, i, N, hint);
    body(i, vols[v].nphalomin, vbeg[v]);
  }
}

// Outputs for every particle: label_out (volume-local), for particles not in
// the sorted list the trivial ring entry next = self (1-based), and the
// group's minimum particle ID (own ID for a singleton).
__global__ __launch_bounds__(256) void k_out_init(uint32_t N, const uint32_t *__restrict__ label,
                                                  const uint32_t *__restrict__ vbeg, int nvol,
                                                  uint32_t *__restrict__ label_out,
                                                  uint32_t *__restrict__ next_out,
                                                  const uint32_t *__restrict__ hsize,
                                                  const uint8_t *__restrict__ base, uint32_t stride,
                                                  uint32_t offid,
                                                  const unsigned long long *__restrict__ idhead,
                                                  unsigned long long *__restrict__ minid_out) {
  const uint64_t i0 = (uint64_t)blockIdx.x * kTileN, i1 = min((uint64_t)N, i0 + kTileN);
  if (i0 >= i1) return;
  auto body = [&](uint32_t i, uint32_t b) {
    const uint32_t l = label[i];
    const uint32_t head = l & ~kFlag;
    if (label_out) label_out[i] = head - b;
    if (next_out && !(l & kFlag)) next_out[i] = i - b + 1u;
    if (minid_out)
      minid_out[i] = (head == i && hsize[i] == 1u) ? rec_id(base, stride, offid, i) : idhead[head];
  };
  const int va = vol_of(vbeg, nvol, (uint32_t)i0);
  if (i1 <= vbeg[va + 1]) {  // the tile lies in one volume
    const uint32_t b = vbeg[va];
    for (uint32_t i = (uint32_t)i0 + threadIdx.x; i < i1; i += blockDim.x) body(i, b);
    return;
  }
  int hint = va;
  for (uint32_t i = (uint32_t)i0 + threadIdx.x; i < i1; i += blockDim.x)
    body(i, vbeg[vol_of_run(vbeg, nvol, i, N, hint)]);
}

// Ring in ascending index order over the label-sorted list (1-based,
// volume-local successors), and halo starts.  The entries go straight into
// next_out, or, with pidx, leave as (index, successor) pairs in list order
// for scatter_u32 (the choice of k_label).
// ---------------------------------------------------------------------------
// The member list: every group of two or more as one segment, segments in
// ascending label (head index) order, members inside a segment in ascending
// input index -- the order of the reference's rings (src_v2/FOF.c:359-373
// joins rings so that each runs up the input order from its head) and of its
// halo loops (src_v2/FOF.c:797-847).  Built without sorting the members by
// label: the groups' roots are sorted by label (one entry per group), a scan
// of their sizes gives each segment's first slot, k_label puts every member
// in its segment, and each segment is then sorted by index on its own.
// ---------------------------------------------------------------------------
// The group list in label order without a sort: each group of two or more
// (and, where a volume keeps groups of one -- Nphalomin <= 1 -- each such
// particle) sets the bit of its label (the group's lowest input index) in a
// bitmap over the input indices and puts its size at lsize[label]; a scan of
// the words' bit counts ranks the labels.
__global__ __launch_bounds__(256) void k_group_bits(uint32_t N, const uint32_t *__restrict__ par,
                                                    const uint32_t *__restrict__ gsize,
                                                    const uint32_t *__restrict__ minorig,
                                                    const float4 *__restrict__ pos4,
                                                    uint32_t *__restrict__ bits,
                                                    uint32_t *__restrict__ lsize, int single,
                                                    const VolDev *__restrict__ vols,
                                                    const uint32_t *__restrict__ vbeg, int nvol) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < N; j += gridDim.x * blockDim.x) {
    if (par[j] != j) continue;
    const uint32_t g = gsize[j];
    if (g == 0 && (!single || 1.0 < vols[vol_of(vbeg, nvol, j)].nphalomin)) continue;
    const uint32_t lab = min(minorig[j], __float_as_uint(pos4[j].w));  // minorig: ~0 alone
    atomicOr(bits + (lab >> 5), 1u << (lab & 31));
    lsize[lab] = g + 1u;
  }
}

__global__ __launch_bounds__(256) void k_word_counts(uint64_t nw, const uint32_t *__restrict__ bits,
                                                     uint32_t *__restrict__ cnt) {
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nw;
       w += (uint64_t)gridDim.x * blockDim.x)
    cnt[w] = (uint32_t)__popc(bits[w]);
}

// Labels in ascending order and their groups' sizes (scanned in place into
// first slots); the group count to the device counts.
__global__ __launch_bounds__(256) void k_group_list(uint64_t nw, const uint32_t *__restrict__ bits,
                                                    const uint32_t *__restrict__ wpre,
                                                    const uint32_t *__restrict__ lsize,
                                                    uint32_t *__restrict__ glab,
                                                    uint32_t *__restrict__ offs,
                                                    uint32_t *__restrict__ dR) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) *dR = wpre[nw];
  for (uint64_t w = t; w < nw; w += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t m = bits[w], at = wpre[w];
    while (m) {
      const uint32_t lab = (uint32_t)(w * 32) + (uint32_t)__builtin_ctz(m);
      m &= m - 1;
      glab[at] = lab;
      offs[at] = lsize[lab];
      at++;
    }
  }
}

// Each group's cursor (indexed by its label) at its first slot; the member
// count to the device counts.
__global__ __launch_bounds__(256) void k_group_cursors(const uint32_t *__restrict__ dR,
                                                       const uint32_t *__restrict__ glab,
                                                       const uint32_t *__restrict__ offs,
                                                       uint32_t *__restrict__ cursor,
                                                       uint32_t *__restrict__ dM) {
  const uint32_t R = *dR;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) *dM = offs[R];
  for (uint32_t i = t; i < R; i += gridDim.x * blockDim.x) cursor[glab[i]] = offs[i];
}

// Sorting each segment of the member list by index.  Groups of at most
// kSegWave members are sorted by one wave each in registers (a bitonic
// network; 64 E elements for E per lane, lane-major, so that the stages
// below distance E stay inside a lane and the others are lane shuffles);
// larger ones are listed for k_seg_sort_large: a workgroup each, a bitonic
// sort in LDS up to kSegLds members, beyond that an LSD radix sort (four
// 8-bit digits, stable scatter per 256-element tile) through scratch.
constexpr uint32_t kSegWave = 1024;
constexpr uint32_t kSegLds = 4096;
constexpr int kSmLarge = 9;  // groups listed for k_seg_sort_large

// n <= 64: one element a lane, the network only as wide as n needs
__device__ __forceinline__ void seg_sort_lane(uint32_t *__restrict__ p, uint32_t n, uint32_t lane) {
  uint32_t v = lane < n ? p[lane] : 0xFFFFFFFFu;
  uint32_t P = 2;
  while (P < n) P <<= 1;
  for (uint32_t k = 2; k <= P; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      const uint32_t u = (uint32_t)__shfl_xor((int)v, (int)j, 64);
      const bool up = (lane & k) == 0, lower = (lane & j) == 0;
      v = (lower == up) ? min(v, u) : max(v, u);
    }
  if (lane < n) p[lane] = v;
}

// 32 E < n <= 64 E: element a = E lane + e in register e of the lane
template <int E>
__device__ __forceinline__ void seg_sort_regs(uint32_t *__restrict__ p, uint32_t n, uint32_t lane) {
  constexpr uint32_t P = 64u * E;
  uint32_t v[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const uint32_t a = lane * E + e;
    v[e] = a < n ? p[a] : 0xFFFFFFFFu;
  }
#pragma unroll
  for (uint32_t k = 2; k <= P; k <<= 1)
#pragma unroll
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      if (j < (uint32_t)E) {
#pragma unroll
        for (int e = 0; e < E; e++) {
          if (e & j) continue;
          const bool up = ((lane * E + e) & k) == 0;
          const uint32_t x = v[e], y = v[e | j];
          v[e] = up ? min(x, y) : max(x, y);
          v[e | j] = up ? max(x, y) : min(x, y);
        }
      } else {
        const uint32_t lj = j / E;
        const bool lower = (lane & lj) == 0;
#pragma unroll
        for (int e = 0; e < E; e++) {
          const uint32_t u = (uint32_t)__shfl_xor((int)v[e], (int)lj, 64);
          const bool up = ((lane * E + e) & k) == 0;
          v[e] = (lower == up) ? min(v[e], u) : max(v[e], u);
        }
      }
    }
#pragma unroll
  for (int e = 0; e < E; e++) {
    const uint32_t a = lane * E + e;
    if (a < n) p[a] = v[e];
  }
}

// A wave takes 64 consecutive groups at a time: groups of at most four
// members are sorted by their lane alone (most groups are pairs), the others
// one at a time by the whole wave; groups above kSegWave members are
// appended to the list for k_seg_sort_large.
__global__ __launch_bounds__(256) void k_seg_sort(const uint32_t *__restrict__ dR,
                                                  const uint32_t *__restrict__ offs,
                                                  uint32_t *__restrict__ mval,
                                                  uint32_t *__restrict__ large,
                                                  uint32_t *__restrict__ dL) {
  const uint32_t R = *dR;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t w0 = blockIdx.x * 4 + (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  for (uint32_t c0 = w0 * 64; c0 < R; c0 += gridDim.x * 4 * 64) {
    const uint32_t g = c0 + lane;
    uint32_t s = 0, n = 0;
    if (g < R) s = offs[g], n = offs[g + 1] - s;
    if (n <= 4) {  // a comparator network on the lane's own group
      uint32_t v[4];
#pragma unroll
      for (int e = 0; e < 4; e++) v[e] = (uint32_t)e < n ? mval[s + e] : 0xFFFFFFFFu;
      auto cas = [&](int a, int b) {
        const uint32_t x = v[a], y = v[b];
        v[a] = min(x, y), v[b] = max(x, y);
      };
      cas(0, 1), cas(2, 3), cas(0, 2), cas(1, 3), cas(1, 2);
#pragma unroll
      for (int e = 0; e < 4; e++)
        if ((uint32_t)e < n) mval[s + e] = v[e];
    }
    unsigned long long big = __ballot(n > 4);
    while (big) {
      const int l = __builtin_ctzll(big);
      big &= big - 1;
      const uint32_t sb = (uint32_t)__builtin_amdgcn_readlane((int)s, l);
      const uint32_t nb = (uint32_t)__builtin_amdgcn_readlane((int)n, l);
      uint32_t *p = mval + sb;
      if (nb <= 64) seg_sort_lane(p, nb, lane);
      else if (nb <= 128) seg_sort_regs<2>(p, nb, lane);
      else if (nb <= 256) seg_sort_regs<4>(p, nb, lane);
      else if (nb <= 512) seg_sort_regs<8>(p, nb, lane);
      else if (nb <= kSegWave) seg_sort_regs<16>(p, nb, lane);
      else if (lane == 0) large[atomicAdd(dL, 1u)] = c0 + (uint32_t)l;
    }
  }
}

__global__ __launch_bounds__(256) void k_seg_sort_large(const uint32_t *__restrict__ dL,
                                                        const uint32_t *__restrict__ large,
                                                        const uint32_t *__restrict__ offs,
                                                        uint32_t *__restrict__ mval,
                                                        uint32_t *__restrict__ scratch) {
  __shared__ uint32_t sk[kSegLds];
  __shared__ uint32_t hist[256], wc[4][256], wsum[4];
  const uint32_t nl = *dL;
  const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
  const unsigned long long lt = (1ull << lane) - 1;
  for (uint32_t q = blockIdx.x; q < nl; q += gridDim.x) {
    const uint32_t g = large[q];
    const uint32_t s = offs[g], n = offs[g + 1] - s;
    if (n <= kSegLds) {
      uint32_t P = 128;
      while (P < n) P <<= 1;
      __syncthreads();
      for (uint32_t i = t; i < P; i += 256) sk[i] = i < n ? mval[s + i] : 0xFFFFFFFFu;
      __syncthreads();
      for (uint32_t k = 2; k <= P; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
          for (uint32_t i = t; i < P / 2; i += 256) {
            const uint32_t a = ((i & ~(j - 1)) << 1) | (i & (j - 1)), b = a + j;
            const uint32_t x = sk[a], y = sk[b];
            if ((x > y) == ((a & k) == 0)) sk[a] = y, sk[b] = x;
          }
          __syncthreads();
        }
      for (uint32_t i = t; i < n; i += 256) mval[s + i] = sk[i];
      continue;
    }
    uint32_t *src = mval + s, *dst = scratch + s;
    for (int shift = 0; shift < 32; shift += 8) {
      __syncthreads();
      hist[t] = 0;
      __syncthreads();
      for (uint32_t i = t; i < n; i += 256) atomicAdd(&hist[(src[i] >> shift) & 255u], 1u);
      __syncthreads();
      {  // exclusive scan of the 256 digit counts: the digits' cursors
        const uint32_t c = hist[t];
        uint32_t inc = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t u = __shfl_up(inc, o, 64);
          if (lane >= (uint32_t)o) inc += u;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        uint32_t pre = inc - c;
        for (uint32_t q2 = 0; q2 < w; q2++) pre += wsum[q2];
        hist[t] = pre;
      }
      for (uint32_t b0 = 0; b0 < n; b0 += 256) {
        const uint32_t i = b0 + t;
        const bool valid = i < n;
        const uint32_t key = valid ? src[i] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
          const bool on = (d >> bit) & 1u;
          const unsigned long long bb = __ballot(on);
          peers &= on ? bb : ~bb;
        }
        __syncthreads();  // the previous tile's reads of wc and hist are done
        wc[w][lane] = 0, wc[w][lane + 64] = 0, wc[w][lane + 128] = 0, wc[w][lane + 192] = 0;
        __syncthreads();
        if (valid && (peers & lt) == 0) wc[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
          uint32_t at = hist[d] + (uint32_t)__popcll(peers & lt);
          for (uint32_t q2 = 0; q2 < w; q2++) at += wc[q2][d];
          dst[at] = key;
        }
        __syncthreads();
        hist[t] += wc[0][t] + wc[1][t] + wc[2][t] + wc[3][t];
      }
      uint32_t *tmp = src;
      src = dst, dst = tmp;
    }
  }
}

// Halo flags of the groups in label order: the first slot | kFlag for a
// group passing its volume's size cut (src_v2/FOF.c:290-298), else 0.
__global__ __launch_bounds__(256) void k_halo_flags(const uint32_t *__restrict__ dR,
                                                    const uint32_t *__restrict__ rkey,
                                                    const uint32_t *__restrict__ offs,
                                                    const VolDev *__restrict__ vols,
                                                    const uint32_t *__restrict__ vbeg, int nvol,
                                                    uint32_t *__restrict__ hsrc) {
  const uint32_t R = *dR;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < R; i += gridDim.x * blockDim.x) {
    const uint32_t n = offs[i + 1] - offs[i];
    const double nph = vols[vol_of(vbeg, nvol, rkey[i])].nphalomin;
    hsrc[i] = !((double)n < nph) ? (offs[i] | kFlag) : 0u;
  }
}

__global__ __launch_bounds__(256) void k_ring(const uint32_t *__restrict__ dM,
                                              const uint32_t *__restrict__ skey,
                                              const uint32_t *__restrict__ sval,
                                              const uint32_t *__restrict__ hsize,
                                              const VolDev *__restrict__ vols,
                                              const uint32_t *__restrict__ vbeg, int nvol,
                                              uint32_t *__restrict__ next_out,
                                              uint32_t *__restrict__ pidx,
                                              uint32_t *__restrict__ pval,
                                              uint32_t *__restrict__ hflag) {
  const uint32_t M = *dM;
  const uint64_t k0 = (uint64_t)blockIdx.x * kTileN, k1 = min((uint64_t)M, k0 + kTileN);
  if (k0 >= k1) return;
  auto body = [&](uint32_t k, uint32_t lab, uint32_t b, double nphalomin) {
    const uint32_t i = sval[k];
    const uint32_t nx = (k + 1 < M && skey[k + 1] == lab) ? sval[k + 1] : lab;
    if (pidx) pidx[k] = i, pval[k] = nx - b + 1u;
    else if (next_out) next_out[i] = nx - b + 1u;
    if (hflag) {
      const bool gstart = (k == 0) || (skey[k - 1] != lab);
      hflag[k] = (gstart && !((double)hsize[lab] < nphalomin)) ? kFlag : 0u;
    }
  };
  // the list is sorted by label, so by volume; a member shares its head's
  // volume, and a tile whose first and last labels share one has no other
  const int va = vol_of(vbeg, nvol, skey[k0]);
  if (skey[k1 - 1] < vbeg[va + 1]) {
    const uint32_t b = vbeg[va];
    const double nph = vols[va].nphalomin;
    for (uint32_t k = (uint32_t)k0 + threadIdx.x; k < k1; k += blockDim.x) body(k, skey[k], b, nph);
    return;
  }
  int hint = va;
  for (uint32_t k = (uint32_t)k0 + threadIdx.x; k < k1; k += blockDim.x) {
    const uint32_t lab = skey[k];
    const int v = vol_of_sorted(vbeg, nvol, lab, hint);
    body(k, lab, vbeg[v], vols[v].nphalomin);
  }
}

// Halo sums: FP64 accumulation of the float Pos / Vel of the members in
// ascending index order from the head (the order the host walks the ring
// returned in next_out), so the host's Checkhalo / Output_Halos loops give
// the same bits.  One wavefront per halo: the 64 lanes gather 64 members'
// records at a time into LDS and then each of the first NQ lanes adds one
// quantity over them, in order.  Quantities (q = lane):
//   0-5    Pos x,y,z, Vel x,y,z                          (src_v2/FOF.c:805-810)
//   with MOM, the second moments (src_v2/FOF.c:811-824):
//   6-11   Pos xx,yy,zz,xy,xz,yz   12-17  Vel xx,...     raw products, ring order
//   18-20  Pos x,y,z - head        21-23  Vel - head     deviations from the head
//   24-29  Pos products of those   30-35  Vel products   (for centred moments)
// A float difference widened to double is exact here (same-halo members), and
// products of two such numbers (< 53 significant bits) are exact, so only the
// additions round.
__device__ __forceinline__ double halo_quantity(int q, const float (&v)[6], const double (&ref)[6]) {
  if (q < 6) return (double)v[q];
  // component pairs of the products: xx yy zz xy xz yz
  const int r = q < 18 ? q - 6 : (q < 24 ? q - 18 : q - 24);
  const int blk = r / 6, k = r % 6, o = blk * 3;  // blk 0: Pos, 1: Vel
  if (q >= 18 && q < 24) {
    const int c = r;  // 0-5: Pos xyz, Vel xyz
    return (double)v[c] - ref[c];
  }
  const int ca = o + (k < 3 ? k : (k < 5 ? 0 : 1));
  const int cb = o + (k < 3 ? k : (k == 3 ? 1 : 2));
  if (q < 18) return (double)v[ca] * (double)v[cb];
  return ((double)v[ca] - ref[ca]) * ((double)v[cb] - ref[cb]);
}

template <bool MOM>
__global__ __launch_bounds__(256) void k_halo_sums(const uint32_t *__restrict__ dH,
                                                   const uint32_t *__restrict__ hk,
                                                   const uint32_t *__restrict__ skey,
                                                   const uint32_t *__restrict__ sval,
                                                   const uint32_t *__restrict__ hsize,
                                                   const uint8_t *__restrict__ base,
                                                   uint32_t stride, uint32_t offpos,
                                                   uint32_t offvel,
                                                   const VolDev *__restrict__ vols,
                                                   const uint32_t *__restrict__ vbeg, int nvol,
                                                   fof_halo_v1 *__restrict__ out,
                                                   fof_moments_v1 *__restrict__ mout,
                                                   uint64_t cap,
                                                   uint32_t *__restrict__ hhead,
                                                   unsigned long long *__restrict__ ctr) {
  constexpr int NQ = MOM ? 36 : 6;
  // two stages per wave: the records of the next 64 members are loaded while
  // the first NQ lanes add up the current ones
  __shared__ float stage[4][2][6][64];
  const uint32_t H = *dH;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t nw = gridDim.x * (blockDim.x >> 6);
  uint32_t owned_cnt = 0;
  int hint = 0;
  for (uint32_t h = blockIdx.x * (blockDim.x >> 6) + w; h < H; h += nw) {
    const uint32_t k0 = hk[h];
    const uint32_t lab = skey[k0];
    const uint32_t n = hsize[lab];
    hint = vol_from(vbeg, nvol, lab, hint);  // halos ascend by head, per wave too
    const int v = hint;
    const VolDev &V = vols[v];
    double acc = 0;
    uint32_t warn = 0;
    float r[6];
    double ref[6];
    auto fetch = [&](uint32_t k) {
      if (lane < min(64u, n - k)) {
        const uint8_t *rec = base + (uint64_t)sval[k0 + k + lane] * stride;
        const float *pp = reinterpret_cast<const float *>(rec + offpos);
        const float *vv = reinterpret_cast<const float *>(rec + offvel);
        r[0] = pp[0], r[1] = pp[1], r[2] = pp[2], r[3] = vv[0], r[4] = vv[1], r[5] = vv[2];
      }
    };
    fetch(0);
    if (MOM) {  // the head (first member) is the reference point of the deviations
#pragma unroll
      for (int c = 0; c < 6; c++) ref[c] = (double)__shfl(r[c], 0, 64);
    }
    int b = 0;
    for (uint32_t k = 0; k < n; k += 64, b ^= 1) {
      const uint32_t m = min(64u, n - k);
      if ((uint32_t)lane < m) {
#pragma unroll
        for (int c = 0; c < 6; c++) stage[w][b][c][lane] = r[c];
#pragma unroll
        for (int d = 0; d < 3; d++) {
          const double q = r[d];
          if ((q - V.rmin_buff[d] < V.linklength) || (V.rmax_buff[d] - q < V.linklength))
            warn |= 1u << d;
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (k + 64 < n) fetch(k + 64);
      if (lane < NQ) {
        if (lane < 6) {
          for (uint32_t q = 0; q < m; q++) acc += (double)stage[w][b][lane][q];
        } else {
          for (uint32_t q = 0; q < m; q++) {
            float vv[6];
#pragma unroll
            for (int c = 0; c < 6; c++) vv[c] = stage[w][b][c][q];
            acc += halo_quantity(lane, vv, ref);
          }
        }
      }
    }
    const double s0 = __shfl(acc, 0, 64), s1 = __shfl(acc, 1, 64), s2 = __shfl(acc, 2, 64);
    const double s3 = __shfl(acc, 3, 64), s4 = __shfl(acc, 4, 64), s5 = __shfl(acc, 5, 64);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) warn |= __shfl_xor(warn, o, 64);
    // src_v2/FOF.c:394-401
    const double cx = s0 / (double)n, cy = s1 / (double)n, cz = s2 / (double)n;
    const bool own = !((cx < V.rmin[0]) || (cx >= V.rmax[0]) || (cy < V.rmin[1]) ||
                       (cy >= V.rmax[1]) || (cz < V.rmin[2]) || (cz >= V.rmax[2]));
    if (!own) warn = 0;
    if (MOM && mout && h < cap) {
      // raw products (lanes 6-17), centred products from the deviations:
      // sum (a - ma)(b - mb) = S_ab - S_a S_b / n with a, b taken about the head
      double dev[6];
#pragma unroll
      for (int c = 0; c < 6; c++) dev[c] = __shfl(acc, 18 + c, 64);
      double prod[12];
#pragma unroll
      for (int c = 0; c < 12; c++) prod[c] = __shfl(acc, 24 + c, 64);
      const double rawq = __shfl(acc, 6 + lane % 12, 64);  // every lane takes part
      if (lane < 12) {
        const int blk = lane / 6, k = lane % 6, o = blk * 3;
        const int ca = o + (k < 3 ? k : (k < 5 ? 0 : 1));
        const int cb = o + (k < 3 ? k : (k == 3 ? 1 : 2));
        double pc = 0, da = 0, db = 0;
#pragma unroll
        for (int c = 0; c < 12; c++) pc = (c == lane) ? prod[c] : pc;
#pragma unroll
        for (int c = 0; c < 6; c++) {
          da = (c == ca) ? dev[c] : da;
          db = (c == cb) ? dev[c] : db;
        }
        double *mo = reinterpret_cast<double *>(mout + h);
        mo[lane] = rawq;                            // pos2[6], vel2[6]
        mo[12 + lane] = pc - da * db / (double)n;   // cpos2[6], cvel2[6]
      }
    }
    if (lane == 0) {
      if (h < cap) {
        fof_halo_v1 H1;
        H1.head = lab - vbeg[v];
        H1.n = n;
        H1.owned = own ? 1u : 0u;
        H1.edge_warn = warn;
        H1.sum_pos[0] = s0, H1.sum_pos[1] = s1, H1.sum_pos[2] = s2;
        H1.sum_vel[0] = s3, H1.sum_vel[1] = s4, H1.sum_vel[2] = s5;
        out[h] = H1;
      }
      hhead[h] = lab;
      owned_cnt += own ? 1u : 0u;
    }
  }
  uint64_t t = owned_cnt;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
  if ((threadIdx.x & 63) == 0 && t) ctr_add(ctr, kCtrOwned, t);
}

// halo_offsets[v] = first halo whose global head >= vbeg[v].
__global__ void k_halo_offsets(const uint32_t *__restrict__ dH,
                               const uint32_t *__restrict__ hhead,
                               const uint32_t *__restrict__ vbeg, int nvol,
                               uint64_t *__restrict__ off) {
  const uint32_t H = *dH;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v <= nvol; v += gridDim.x * blockDim.x) {
    const uint32_t key = vbeg[v];
    uint32_t lo = 0, hi = H;
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      if (hhead[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    off[v] = (v == nvol) ? H : lo;
  }
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
static inline uint32_t grid_for(uint64_t n, uint32_t cap = 65535u * 4u) {
  uint64_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (uint32_t)g;
}

static int bits_for(uint64_t n) {
  int b = 1;
  while (b < 32 && (1ull << b) < n) b++;
  return b;
}

// Host-side volume preparation: grid and thresholds, validated.
struct HostVol {
  VolDev d;
  uint64_t ncell;
  double ncell_d;
};

// Smallest l^2 of the cells holding particles, per volume with a variable
// l (as the bits of a non-negative double: their order is the doubles').
// Each wave takes kVlRun consecutive particles and keeps a running minimum
// per lane, flushed when the lane's volume changes and once at the end (an
// atomic per wave and step on one address per volume serialises).
constexpr uint32_t kVlRun = 2048;

__global__ __launch_bounds__(256) void k_varlink_min(const uint8_t *__restrict__ base,
                                                     uint32_t stride, uint32_t offpos, uint32_t N,
                                                     const VolDev *__restrict__ vols,
                                                     const uint32_t *__restrict__ vbeg, int nvol,
                                                     unsigned long long *__restrict__ vmin) {
  const int lane = threadIdx.x & 63;
  const uint64_t w0 = ((uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * kVlRun;
  int cv = -1, hint = 0;
  unsigned long long cm = ~0ull;
  for (uint32_t r = 0; r < kVlRun; r += 64) {
    const uint64_t pp64 = w0 + r + lane;
    if (pp64 >= N) break;
    const uint32_t p = (uint32_t)pp64;
    const int v = vol_of_run(vbeg, nvol, p, N, hint);
    const VolDev &V = vols[v];
    const float *pp = reinterpret_cast<const float *>(base + (uint64_t)p * stride + offpos);
    const double x = pp[0], y = pp[1], z = pp[2];
    const bool ok = V.vntab > 0 && x >= V.rmin_buff[0] && x <= V.rmax_buff[0] &&
                    y >= V.rmin_buff[1] && y <= V.rmax_buff[1] && z >= V.rmin_buff[2] &&
                    z <= V.rmax_buff[2];
    unsigned long long bits = ~0ull;
    if (ok) {
      double u;
      const int ix = cell1(x, V.rmin_buff[0], V.lcell[0], V.nx, &u);
      const int iy = cell1(y, V.rmin_buff[1], V.lcell[1], V.ny, &u);
      const int iz = cell1(z, V.rmin_buff[2], V.lcell[2], V.nz, &u);
      const double l2 = cell_linksq(V, ix, iy, iz);
      memcpy(&bits, &l2, 8);
    }
    if (v != cv) {
      if (cv >= 0 && cm != ~0ull) atomicMin(vmin + cv, cm);
      cv = v, cm = bits;
    } else if (bits < cm) {
      cm = bits;
    }
  }
  const int v0 = __shfl(cv, 0, 64);
  if (__ballot(cv != v0) == 0ull) {  // one volume: reduce over the wave first
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned long long t = __shfl_xor(cm, o, 64);
      cm = t < cm ? t : cm;
    }
    if (lane == 0 && v0 >= 0 && cm != ~0ull) atomicMin(vmin + v0, cm);
  } else if (cv >= 0 && cm != ~0ull) {
    atomicMin(vmin + cv, cm);
  }
}

// Sub-cells: side <= l_min (1 - 1e-6) / sqrt(3) so that any two particles
// of one sub-cell link (see file header).  Disabled (m = 0) when more than
// 12 per axis would be needed, or l_min = 0.
static void set_subcells(VolDev &V, double lmin2) {
  V.m[0] = V.m[1] = V.m[2] = 0;
  if (!(lmin2 > 0)) return;
  const double s = sqrt(lmin2) * (1.0 - 1e-6) / sqrt(3.0);
  int m[3];
  bool ok = true;
  for (int k = 0; k < 3; k++) {
    double q = ceil(V.lcell[k] / s);
    if (q < 1) q = 1;
    if (q > 12) ok = false;
    m[k] = (int)q;
  }
  const int nn[3] = {V.nx, V.ny, V.nz};
  for (int k = 0; k < 3; k++)
    if ((double)nn[k] * m[k] >= 2097152.0) ok = false;  // 21-bit fine coordinates
  if (ok)
    for (int k = 0; k < 3; k++) {
      V.m[k] = m[k];
      const double side = V.lcell[k] / m[k];
      V.s2[k] = side * side;
    }
}

int prepare_volume(const fof_volume_v1 &in, double band, HostVol &hv, std::vector<const fof_varlink_v1 *> &vl) {
  const fof_domain_v1 &D = in.domain;
  VolDev &V = hv.d;
  memset(&V, 0, sizeof(V));
  for (int k = 0; k < 3; k++) {
    V.rmin_buff[k] = D.rmin_buff[k];
    V.rmax_buff[k] = D.rmax_buff[k];
    V.rmin[k] = D.rmin[k];
    V.rmax[k] = D.rmax[k];
    V.lcell[k] = D.lcell[k];
    if (!(D.lcell[k] > 0) || !std::isfinite(D.lcell[k]) || !(D.rmax_buff[k] >= D.rmin_buff[k]))
      return FOF_E_ARG;
  }
  if (!(D.linklength >= 0) || !std::isfinite(D.linklength)) return FOF_E_ARG;
  V.linklength = D.linklength;
  V.linksq = D.linklength * D.linklength;  // src_v2/FOF.c:60
  V.nphalomin = D.nphalomin;
  V.band = (float)band;
  V.lo = (float)(V.linksq * (1.0 - band));
  V.hi = (float)(V.linksq * (1.0 + band));
  // An internal grid finer than the reference's, for a constant l well
  // inside the reference cells.  The reference predicate links a pair only
  // if its true distance is below l(1 + 1e-7) (dx = fl32(xi - xj) errs by
  // < 6e-8).  On an axis where lcell > l(1 + 1e-6), every such pair is at
  // most one cell apart on the reference grid and on any grid whose cells
  // are longer still, so both sweeps find the same linked pairs and the
  // groups do not depend on the choice.  Such an axis is cut into
  // r = ceil(lcell / 6.9 l) cells (so sub-cells, m <= 12, stay on: the
  // all-pairs scan of coarse cells is what a coarse grid otherwise costs).
  // A variable l keeps the reference grid: its cell's l^2 is part of the
  // rule (src_v2/FOF.c:102-109).
  if (!D.varlink && D.linklength > 0) {
    for (int k = 0; k < 3; k++) {
      const double L = D.lcell[k], l = D.linklength;
      if (!(L > l * (1.0 + 1e-6))) continue;
      double r = ceil(L / (6.9 * l));
      const double ext = D.rmax_buff[k] - D.rmin_buff[k];
      while (r > 1 && ceil(ext / (L / r)) + 1.0 > 1048576.0) r -= 1;  // at most 2^20 cells
      if (r > 1) V.lcell[k] = L / r;
    }
  }
  // src_v2/FOF.c:63-65 (on the grid in use)
  double e[3];
  for (int k = 0; k < 3; k++) e[k] = ceil((D.rmax_buff[k] - D.rmin_buff[k]) / V.lcell[k]) + 1.0;
  if (!(e[0] < 2e9 && e[1] < 2e9 && e[2] < 2e9)) return FOF_E_RANGE;
  V.nx = (int)e[0], V.ny = (int)e[1], V.nz = (int)e[2];
  hv.ncell_d = e[0] * e[1] * e[2];
  hv.ncell = (hv.ncell_d < 1.8e19) ? (uint64_t)V.nx * (uint64_t)V.ny * (uint64_t)V.nz : ~0ull;
  // Smallest linking length that any home cell can use.
  double lmin2 = V.linksq;
  if (D.varlink) {
    const fof_varlink_v1 *L = D.varlink;
    if (L->ntab < 2 || !L->l2tab || !(L->r2max > 0)) return FOF_E_ARG;
    vl.push_back(L);
    V.vntab = L->ntab;
    V.vr2max = L->r2max;
    for (int k = 0; k < 3; k++) V.vorigin[k] = L->origin[k];
    lmin2 = -1;  // set by the caller from the occupied cells
  }
  set_subcells(V, lmin2);
  return 0;
}

// Bump allocator over a workspace: run once with base == nullptr to size it,
// then over the allocated arena to carve it.  Every buffer is 256-byte
// aligned.
struct Carve {
  uint8_t *base = nullptr;
  size_t off = 0;
  template <class T>
  T *take(size_t bytes) {
    const size_t o = (off + 255) & ~(size_t)255;
    off = o + bytes;
    return base ? reinterpret_cast<T *>(base + o) : nullptr;
  }
};

// Host-side description of a call: volumes, grids, sizes of every buffer.
struct Plan {
  uint64_t N = 0, C = 0, Csz = 0;
  double Cd = 0;
  uint32_t nvol = 0;
  bool sparse = false;
  int key_bits = 0, volshift = 0;
  uint32_t T = kDefaultT;
  uint32_t big_grid = 8192;  // workgroups of the 256-thread dense kernels
  double band = 1e-5;
  std::vector<HostVol> hv;
  std::vector<uint32_t> vbeg;
  std::vector<double> vt_host;
  std::vector<uint32_t> vt_off;  // per volume: first table entry (or ~0)
  size_t vt_entries = 0;
  bool varlink = false;
  // dense grid partitions
  uint32_t F = 0, Bc = 0, ntA = 0, ntBmax = 0;
  uint64_t nA = 0, nB = 0;
  uint64_t ndense_cap = 0, Hbound = 0, Hstage = 0;
  bool single_halos = false;  // some volume keeps groups of one (Nphalomin <= 1)
  size_t tmpb = 0;
  // what the call produces
  uint32_t flags = 0;
  bool host_input = false, host_out = false;
  bool want_ring = false, want_label = false, want_minid = false, want_halos = false,
       want_mom = false;
  bool stage_halos = false;  // halos (and moments) into the workspace first
  bool direct_writes = true;  // labels and ring straight by index (k_label)
  uint32_t stride = 32, offpos = 8, offvel = 20;
  int32_t offid = 0;
};

// Device buffers of a call, carved from the arena.
struct Bufs {
  float4 *pos4;
  uint32_t *key, *start, *par, *runend, *rtab, *gsize, *label, *sortbuf, *dense, *hhead, *cnt,
      *part, *vbeg, *small;
  unsigned long long *fine, *idhead;
  void *tmp;
  VolDev *vols;
  double *vtab;
  uint64_t *hoff;
  uint8_t *input;
  uint32_t *st_next, *st_label;
  unsigned long long *st_minid;
  fof_halo_v1 *st_halo;
  fof_moments_v1 *st_mom;
};

constexpr size_t kSmallBytes = 256 + (size_t)kCtrSlots * kCtrStride * 8;

static size_t layout(const Plan &pl, uint8_t *base, Bufs &b) {
  Carve c;
  c.base = base;
  const uint64_t N = pl.N, q4 = (N + 63) / 64 * 64;
  b.pos4 = c.take<float4>(N * 16);
  b.key = c.take<uint32_t>(N * 4 + 64);
  b.start = c.take<uint32_t>((pl.Csz + 2) * 4 + 64);
  b.par = c.take<uint32_t>(N * 4 + 64);
  b.runend = c.take<uint32_t>(N * 4 + 64);
  b.fine = c.take<unsigned long long>(N * 8 + 64);
  b.rtab = c.take<uint32_t>(N * 4 + 64);
  b.gsize = c.take<uint32_t>(N * 4 + 64);
  b.label = c.take<uint32_t>(N * 4 + 64);
  b.sortbuf = c.take<uint32_t>(16 * q4 + 256);
  b.dense = c.take<uint32_t>(2 * pl.ndense_cap * 4 + 64);
  b.tmp = c.take<void>(pl.tmpb);
  b.vols = c.take<VolDev>(pl.nvol * sizeof(VolDev));
  b.vtab = c.take<double>(pl.vt_entries * sizeof(double) + 8);
  b.vbeg = c.take<uint32_t>((pl.nvol + 1) * 4 + 64);
  b.small = c.take<uint32_t>(kSmallBytes);
  b.hhead = c.take<uint32_t>(pl.Hbound * 4 + 64);
  b.hoff = c.take<uint64_t>((pl.nvol + 1) * 8 + 64);
  b.cnt = pl.sparse ? c.take<uint32_t>(N * 4 + 64) : nullptr;
  b.part = pl.sparse ? nullptr
                     : c.take<uint32_t>((2 * pl.nA + 2 * pl.nB + 2 * (uint64_t)pl.Bc + pl.ntBmax +
                                         pl.F + 512) * 4);
  b.input = pl.host_input ? c.take<uint8_t>(N * pl.stride) : nullptr;
  b.idhead = pl.want_minid ? c.take<unsigned long long>(N * 8 + 64) : nullptr;
  const bool ho = pl.host_out;
  b.st_next = ho && pl.want_ring ? c.take<uint32_t>(N * 4) : nullptr;
  b.st_label = ho && pl.want_label ? c.take<uint32_t>(N * 4) : nullptr;
  b.st_minid = ho && pl.want_minid ? c.take<unsigned long long>(N * 8) : nullptr;
  const bool hs = ho || pl.stage_halos;
  b.st_halo = hs && pl.want_halos ? c.take<fof_halo_v1>(pl.Hstage * sizeof(fof_halo_v1)) : nullptr;
  b.st_mom = hs && pl.want_mom ? c.take<fof_moments_v1>(pl.Hstage * sizeof(fof_moments_v1)) : nullptr;
  return c.off + 256;
}

static int make_plan(const fof_particles_v1 *P, const fof_volume_v1 *vols, uint32_t nvol,
                     const fof_opts_v1 *opts, const fof_outputs_v1 *out, Plan &pl) {
  pl.flags = opts ? opts->flags : 0u;
  pl.band = (opts && opts->band > 0) ? opts->band : 1e-5;
  if (!(pl.band >= 1e-6 && pl.band <= 0.1)) return FOF_E_ARG;
  pl.T = std::min(kMaxT, (opts && opts->dense_threshold) ? opts->dense_threshold : kDefaultT);
  pl.big_grid = (opts && opts->big_grid) ? std::min(opts->big_grid, 65536u) : 8192u;
  pl.nvol = nvol;
  pl.stride = P->stride, pl.offpos = P->off_pos, pl.offvel = P->off_vel, pl.offid = P->off_id;
  pl.host_input = P->location != FOF_MEM_DEVICE;
  pl.hv.resize(nvol);
  pl.vbeg.resize(nvol + 1);
  pl.vt_off.assign(nvol, ~0u);
  std::vector<const fof_varlink_v1 *> vls;
  double nphmin = 1e300;
  for (uint32_t v = 0; v < nvol; v++) {
    if (vols[v].offset != pl.N) return FOF_E_ARG;  // volumes must be back to back
    int rc = prepare_volume(vols[v], pl.band, pl.hv[v], vls);
    if (rc) return rc;
    pl.vbeg[v] = (uint32_t)pl.N;
    pl.hv[v].d.pbeg = (uint32_t)pl.N;
    pl.hv[v].d.vid = v;
    pl.N += vols[v].n;
    if (pl.N >= 0x7FFFFFFFull) return FOF_E_RANGE;
    pl.hv[v].d.pend = (uint32_t)pl.N;
    pl.Cd += pl.hv[v].ncell_d;
    nphmin = std::min(nphmin, vols[v].domain.nphalomin);
  }
  pl.vbeg[nvol] = (uint32_t)pl.N;
  // labels and ring: straight random writes by index while every volume's
  // range stays in the last-level cache, else through an index partition
  uint64_t maxvol = 0;
  for (uint32_t v = 0; v < nvol; v++) maxvol = std::max<uint64_t>(maxvol, pl.vbeg[v + 1] - pl.vbeg[v]);
  pl.direct_writes = (pl.flags & FOF_OPT_WRITE_PAIRS) ? false
                     : (pl.flags & FOF_OPT_WRITE_DIRECT) ? true
                                                          : maxvol <= kDirectMax;
  if (pl.N != P->n) return FOF_E_ARG;
  // Dense cell tables unless they would dwarf the particle arrays or reach
  // the 2^31 cells a u32 cell index (with room for the table's end) allows.
  pl.sparse = (pl.flags & FOF_OPT_SPARSE) || pl.Cd > 3.0 * (double)pl.N + 4194304.0 ||
              pl.Cd >= (double)0x7FFFFFF0u;
  if (pl.sparse) {
    int lb = 1, vb = 0;
    for (uint32_t v = 0; v < nvol; v++) {
      int b = 1;
      while (b < 64 && (double)(1ull << b) < pl.hv[v].ncell_d) b++;
      lb = std::max(lb, b);
    }
    while ((1u << vb) < nvol) vb++;
    for (uint32_t v = 0; v < nvol; v++)
      if (pl.hv[v].ncell_d > 1.8e19) return FOF_E_RANGE;
    if (lb + vb > 64) return FOF_E_RANGE;
    pl.volshift = lb;
    pl.key_bits = lb + vb;
  } else {
    for (uint32_t v = 0; v < nvol; v++) {
      pl.hv[v].d.cbeg = (uint32_t)pl.C;
      pl.hv[v].d.ncell = (uint32_t)pl.hv[v].ncell;
      pl.C += pl.hv[v].ncell;
    }
  }
  pl.Csz = pl.sparse ? pl.N : pl.C;
  for (uint32_t v = 0; v < nvol; v++)
    if (vols[v].domain.varlink) {
      pl.vt_off[v] = (uint32_t)pl.vt_entries;
      pl.vt_entries += vols[v].domain.varlink->ntab;
    }
  pl.varlink = pl.vt_entries > 0;
  // outputs
  const bool od = out->location == FOF_MEM_DEVICE;
  pl.host_out = !od;
  pl.want_ring = out->next && !(pl.flags & FOF_OPT_NO_RING);
  pl.want_label = out->label != nullptr;
  pl.want_minid = out->minid != nullptr;
  if (pl.want_minid && P->off_id < 0) return FOF_E_ARG;
  pl.want_halos = !(pl.flags & FOF_OPT_NO_HALOS);
  pl.want_mom = pl.want_halos && out->moments != nullptr;
  // sizes
  const uint64_t N = pl.N;
  pl.ndense_cap = N / (pl.T ? pl.T : 1) + 2;
  // kept groups have at least max(1, ceil(nphalomin)) members and are disjoint
  const double kmin = std::max(1.0, std::ceil(std::min(nphmin, 4e9)));
  pl.Hbound = std::min<uint64_t>(N, (uint64_t)((double)N / kmin) + 1);
  pl.single_halos = !(1.0 < nphmin);
  pl.Hstage = std::min<uint64_t>(pl.Hbound, out->halo_cap);
  pl.tmpb = std::max(sort_tmp_bytes(N), compact_tmp_bytes(N));
  if (!pl.sparse) {
    pl.F = (uint32_t)((pl.C + kFineCells - 1) >> kFineBits);
    pl.Bc = (pl.F + 255) >> 8;
    if (pl.Bc > (uint32_t)kMaxCoarse) return FOF_E_RANGE;
    pl.ntA = (uint32_t)((N + kPartTile - 1) / kPartTile);
    pl.ntBmax = pl.ntA + pl.Bc;
    pl.nA = (uint64_t)pl.Bc * pl.ntA;
    pl.nB = 256ull * pl.ntBmax;
    pl.tmpb = std::max(pl.tmpb, scan_tmp_bytes(std::max(pl.nA, pl.nB) + 1));
  }
  // varlink tables (checked on the host)
  pl.vt_host.resize(pl.vt_entries);
  for (uint32_t v = 0; v < nvol; v++) {
    const fof_varlink_v1 *L = vols[v].domain.varlink;
    if (!L) continue;
    double *dst = pl.vt_host.data() + pl.vt_off[v];
    if (L->location == FOF_MEM_DEVICE)
      FOF_HIP(hipMemcpy(dst, L->l2tab, L->ntab * sizeof(double), hipMemcpyDeviceToHost));
    else
      memcpy(dst, L->l2tab, L->ntab * sizeof(double));
    for (uint32_t k = 0; k < L->ntab; k++)
      if (!(dst[k] >= 0) || !std::isfinite(dst[k])) return FOF_E_ARG;
  }
  return 0;
}

// The volume table (VolDev[nvol], varlink tables, volume starts) as uploaded.
static void build_vblob(const Plan &pl, const Bufs &b, std::vector<uint8_t> &blob) {
  const size_t nv = pl.nvol * sizeof(VolDev);
  blob.resize(nv + pl.vt_entries * sizeof(double) + (pl.nvol + 1) * 4);
  for (uint32_t v = 0; v < pl.nvol; v++) {
    VolDev V = pl.hv[v].d;
    if (pl.vt_off[v] != ~0u) V.vtab = b.vtab + pl.vt_off[v];
    memcpy(&blob[v * sizeof(VolDev)], &V, sizeof(VolDev));
  }
  if (pl.vt_entries) memcpy(&blob[nv], pl.vt_host.data(), pl.vt_entries * sizeof(double));
  memcpy(&blob[nv + pl.vt_entries * sizeof(double)], pl.vbeg.data(), (pl.nvol + 1) * 4);
}

static int upload_vblob(const Plan &pl, const Bufs &b, const std::vector<uint8_t> &blob,
                        hipStream_t st) {
  const size_t nv = pl.nvol * sizeof(VolDev), nt = pl.vt_entries * sizeof(double);
  FOF_HIP(hipMemcpyAsync(b.vols, blob.data(), nv, hipMemcpyHostToDevice, st));
  if (nt) FOF_HIP(hipMemcpyAsync(b.vtab, blob.data() + nv, nt, hipMemcpyHostToDevice, st));
  FOF_HIP(hipMemcpyAsync(b.vbeg, blob.data() + nv + nt, (pl.nvol + 1) * 4, hipMemcpyHostToDevice,
                         st));
  return 0;
}

// ---------------------------------------------------------------------------
// One pipeline's share of a call: plan, carved workspace, output pointers.
// ---------------------------------------------------------------------------
struct Run {
  Plan pl;
  Bufs B;
  Pipe *pp = nullptr;
  const uint8_t *dbase = nullptr;  // particle records on the device
  const void *host_src = nullptr;  // host records to upload first (host input)
  uint32_t *d_next = nullptr, *d_label = nullptr;
  unsigned long long *d_minid = nullptr;
  fof_halo_v1 *d_halo = nullptr;
  fof_moments_v1 *d_mom = nullptr;
  uint64_t hcap = 0;
  uint64_t poff = 0;    // first particle of this pipeline in the call
  uint32_t v0 = 0;      // first volume of this pipeline in the call
  bool timing = false;  // per-phase events on this pipeline's stream
  uint32_t spans = 0;
  int first_span = -1, last_span = -1;
};

// Plan a pipeline over volumes [v0, v0 + nv) of the call (rebased so that
// the pipeline sees particles from 0) and carve its workspace.
static int prepare_run(Run &r, Pipe &pp, const fof_particles_v1 *P, const fof_volume_v1 *vols,
                       uint32_t v0, uint32_t nv, const fof_opts_v1 *opts,
                       const fof_outputs_v1 *out, bool stage_halos, uint8_t *fixed,
                       size_t fixed_bytes) {
  r.pp = &pp;
  r.v0 = v0;
  r.poff = vols[v0].offset;
  fof_particles_v1 Pg = *P;
  std::vector<fof_volume_v1> vg(vols + v0, vols + v0 + nv);
  uint64_t n = 0;
  for (auto &v : vg) {
    v.offset -= r.poff;
    n += v.n;
  }
  Pg.n = n;
  Pg.base = static_cast<const uint8_t *>(P->base) + r.poff * P->stride;
  int rc = make_plan(&Pg, vg.data(), nv, opts, out, r.pl);
  if (rc) return rc;
  Plan &pl = r.pl;
  pl.stage_halos = stage_halos && pl.want_halos;
  if (pl.N == 0) return 0;
  const size_t need = layout(pl, nullptr, r.B);
  if (fixed) {
    // the caller's fixed workspace: nothing is allocated
    if (need > fixed_bytes || pl.sparse) return FOF_E_RANGE;
    layout(pl, fixed, r.B);
  } else {
    // a workspace far larger than this call needs (a previous call with
    // fewer pipelines) is given back rather than kept beside the others
    if (pp.arena.bytes > 2 * need && pp.arena.bytes - need > (1ull << 30)) release(pp.arena);
    if ((rc = ensure(pp.arena, need))) return rc;
    layout(pl, static_cast<uint8_t *>(pp.arena.ptr), r.B);
  }
  if (pl.host_input) {
    r.host_src = Pg.base;
    r.dbase = r.B.input;
  } else {
    r.dbase = static_cast<const uint8_t *>(Pg.base);
  }
  const bool od = !pl.host_out;
  r.d_next = pl.want_ring ? (od ? out->next + r.poff : r.B.st_next) : nullptr;
  r.d_label = pl.want_label ? (od ? out->label + r.poff : r.B.st_label) : nullptr;
  r.d_minid = pl.want_minid ? (od ? reinterpret_cast<unsigned long long *>(out->minid) + r.poff
                                  : r.B.st_minid)
                            : nullptr;
  const bool direct = od && !pl.stage_halos;
  r.d_halo = pl.want_halos ? (direct ? out->halos : r.B.st_halo) : nullptr;
  r.d_mom = pl.want_mom ? (direct ? out->moments : r.B.st_mom) : nullptr;
  r.hcap = r.d_halo ? (direct ? out->halo_cap : pl.Hstage) : 0;
  build_vblob(pl, r.B, pp.vblob);
  pp.hres.resize(kSmallBytes + (nv + 1) * 8);
  return 0;
}

// Uploads that precede a pipeline's kernels: the volume table and, for
// host-resident input, the particle records.
static int upload_run(Run &r, hipStream_t st) {
  if (r.pl.N == 0) return 0;
  if (r.host_src)
    FOF_HIP(hipMemcpyAsync(r.B.input, r.host_src, r.pl.N * r.pl.stride, hipMemcpyHostToDevice, st));
  return upload_vblob(r.pl, r.B, r.pp->vblob, st);
}

// Every kernel of a pipeline, in stream order.  Without a variable linking
// length or occupied-cell addressing nothing here waits for the host, so the
// sequence can be captured in a graph.
static int enqueue_run(Run &r, hipStream_t st) {
  Plan &pl = r.pl;
  Bufs &B = r.B;
  Pipe &pp = *r.pp;
  const uint64_t N = pl.N;
  if (N == 0) return 0;
  const uint32_t nvol = pl.nvol;
  int rc = 0;
  const uint8_t *dbase = r.dbase;
  uint32_t *small = B.small;
  unsigned long long *ctr = reinterpret_cast<unsigned long long *>(small + kSmCtr);
  uint32_t *hsm = reinterpret_cast<uint32_t *>(pp.hres.data());
  const bool timing = r.timing;
  if (timing && !pp.events) {
    for (auto &e : pp.ev) FOF_HIP(hipEventCreate(&e));
    pp.events = true;
  }
  r.spans = 0;
  r.first_span = r.last_span = -1;
  auto tb = [&](int k) -> int {
    if (timing) {
      FOF_HIP(hipEventRecord(pp.ev[2 * k], st));
      if (r.first_span < 0) r.first_span = k;
    }
    return 0;
  };
  auto te = [&](int k) -> int {
    if (timing) {
      FOF_HIP(hipEventRecord(pp.ev[2 * k + 1], st));
      r.spans |= 1u << k;
      r.last_span = k;
    }
    return 0;
  };
  if ((rc = fill_u32(small, kSmallBytes / 4, 0u, st))) return rc;

  if (pl.varlink) {
    // A variable l: the sub-cell size follows the smallest l of any occupied
    // cell (only occupied cells are home cells of a pair), not of the table.
    unsigned long long *dmin = reinterpret_cast<unsigned long long *>(B.hoff);
    if ((rc = fill_u32(reinterpret_cast<uint32_t *>(dmin), 2ull * nvol, ~0u, st))) return rc;
    k_varlink_min<<<(uint32_t)((N + 4 * kVlRun - 1) / (4 * kVlRun)), 256, 0, st>>>(
        dbase, pl.stride, pl.offpos, (uint32_t)N, B.vols, B.vbeg, (int)nvol, dmin);
    FOF_HIP(hipGetLastError());
    std::vector<unsigned long long> hmin(nvol);
    FOF_HIP(hipMemcpyAsync(hmin.data(), dmin, nvol * 8, hipMemcpyDeviceToHost, st));
    FOF_HIP(hipStreamSynchronize(st));
    for (uint32_t v = 0; v < nvol; v++) {
      if (pl.vt_off[v] == ~0u || hmin[v] == ~0ull) continue;
      double l2;
      memcpy(&l2, &hmin[v], 8);
      set_subcells(pl.hv[v].d, l2);
    }
    build_vblob(pl, B, pp.vblob);
    if ((rc = upload_vblob(pl, B, pp.vblob, st))) return rc;
  }

  const VolDev *dV = B.vols;
  const uint32_t *dVB = B.vbeg;
  float4 *pos4 = B.pos4;
  uint32_t *key = B.key, *start = B.start, *par = B.par, *runend = B.runend, *rtab = B.rtab;
  unsigned long long *fine = B.fine;
  uint32_t *rtoff = key;  // free between the cell sort and k_label (head sizes)
  uint32_t *gsize = B.gsize, *label = B.label;
  const uint64_t q4 = (N + 63) / 64 * 64;
  // four u32 arrays; viewed as float4[N] they are the sub-cell sort buffer
  uint32_t *sk0 = B.sortbuf, *sv0 = sk0 + q4, *sk1 = sk0 + 2 * q4, *sv1 = sk0 + 3 * q4;
  uint32_t *dense = B.dense, *dbig = dense + pl.ndense_cap;
  const uint32_t NN = (uint32_t)N;
  const uint32_t stride = pl.stride, offpos = pl.offpos, offvel = pl.offvel;
  const uint32_t offid = pl.offid >= 0 ? (uint32_t)pl.offid : 0u;
  const uint32_t gN = grid_for(N);
  const uint32_t T = pl.T;

  Grid G;
  G.start = start;
  G.htab = nullptr;
  G.cell_of = label;  // sparse: the label buffer is free until k_label
  G.hmask = 0;
  G.sparse = pl.sparse ? 1 : 0;
  G.volshift = pl.volshift;
  if (!pl.sparse) {
    // 1-3: counting sort by cell: partitions A (coarse) and B (fine), then C
    const uint32_t F = pl.F, Bc = pl.Bc, ntA = pl.ntA, ntBmax = pl.ntBmax;
    const uint64_t nA = pl.nA, nB = pl.nB;
    uint32_t *histA = B.part;
    uint32_t *offA = histA + nA + 32;      // nA + 1
    uint32_t *histB = offA + nA + 32;      // nB
    uint32_t *offB = histB + nB + 32;      // nB + 1
    uint32_t *cst = offB + nB + 32;        // Bc + 1
    uint32_t *t0 = cst + Bc + 32;          // Bc + 1
    uint32_t *tinfo = t0 + Bc + 32;        // ntBmax
    uint32_t *fst = tinfo + ntBmax + 32;   // F + 1
    uint32_t *k1 = gsize, *k2 = label;     // free until the union-find finish / k_label
    float4 *p1 = pos4;                     // pass A output, consumed before pass C writes
    float4 *p2 = reinterpret_cast<float4 *>(sk0);
    if ((rc = tb(0))) return rc;
    k_pa_count<<<ntA, 256, 0, st>>>(dbase, stride, offpos, NN, dV, dVB, (int)nvol, Bc, ntA, key,
                                    histA, small + kSmErr);
    FOF_HIP(hipGetLastError());
    if ((rc = te(0))) return rc;
    if ((rc = tb(1))) return rc;
    if ((rc = scan_u32(histA, offA, nA, B.tmp, st))) return rc;
    k_pa_starts<<<grid_for(Bc + 1), 256, 0, st>>>(Bc, ntA, offA, cst);
    FOF_HIP(hipGetLastError());
    k_pb_tiles<<<1, kMaxCoarse, 0, st>>>(Bc, ntBmax, cst, t0, tinfo, small + kSmNtB);
    FOF_HIP(hipGetLastError());
    if ((rc = te(1))) return rc;
    if ((rc = tb(2))) return rc;
    k_pa_scatter<<<ntA, kPartNT, 0, st>>>(dbase, stride, offpos, NN, Bc, ntA, key, offA, k1, p1);
    FOF_HIP(hipGetLastError());
    k_pb_count<<<ntBmax, 256, 0, st>>>(tinfo, t0, cst, k1, histB);
    FOF_HIP(hipGetLastError());
    if ((rc = scan_u32(histB, offB, nB, B.tmp, st))) return rc;
    k_pb_scatter<<<ntBmax, kPartNT, 0, st>>>(tinfo, t0, cst, offB, k1, p1, k2, p2);
    FOF_HIP(hipGetLastError());
    k_pb_starts<<<grid_for(F + 1), 256, 0, st>>>(F, t0, cst, offB, NN, fst);
    FOF_HIP(hipGetLastError());
    k_pc_sort<<<F, kPcThreads, 0, st>>>((uint64_t)pl.C, fst, k2, p2, start, pos4, dV, (int)nvol, T,
                                        dense, small + kSmNDense, dbig, small + kSmNBig);
    FOF_HIP(hipGetLastError());
    if ((rc = te(2))) return rc;
  } else {
    // 1-3: composite keys, LSD sort by (volume, cell), occupied cells, hash
    uint32_t *cnt = B.cnt;
    if ((rc = tb(0))) return rc;
    k_keys_sparse<<<gN, 256, 0, st>>>(dbase, stride, offpos, NN, dV, dVB, (int)nvol, G, sk0, key,
                                      sv0, small + kSmErr);
    FOF_HIP(hipGetLastError());
    if ((rc = te(0))) return rc;
    if ((rc = tb(1))) return rc;
    if ((rc = radix_sort_u32(sk0, sv0, sk1, sv1, N, nullptr, std::min(pl.key_bits, 32), B.tmp, st)))
      return rc;
    if (pl.key_bits > 32) {
      k_gather_u32<<<gN, 256, 0, st>>>(NN, key, sv0, gsize);
      FOF_HIP(hipGetLastError());
      if ((rc = radix_sort_u32(gsize, sv0, sk1, sv1, N, nullptr, pl.key_bits - 32, B.tmp, st)))
        return rc;
    }
    if ((rc = te(1))) return rc;
    if ((rc = tb(2))) return rc;
    k_scatter_sparse<<<gN, 256, 0, st>>>(dbase, stride, offpos, NN, sv0, dV, dVB, (int)nvol, G,
                                         pos4, par, runend);
    FOF_HIP(hipGetLastError());
    if ((rc = te(2))) return rc;
    k_cell_flags<<<gN, 256, 0, st>>>(NN, par, runend, cnt);
    FOF_HIP(hipGetLastError());
    if ((rc = compact_flagged(cnt, N, nullptr, kFlag, nullptr, start, small + kSmU, B.tmp, st)))
      return rc;
    // the hash table is sized from the occupied-cell count
    FOF_HIP(hipMemcpyAsync(hsm, small, 64, hipMemcpyDeviceToHost, st));
    FOF_HIP(hipStreamSynchronize(st));
    const uint32_t U = hsm[kSmU];
    uint64_t hsz = 1024;
    while (hsz < 2ull * U) hsz <<= 1;
    if ((rc = ensure(pp.hash, hsz * sizeof(HEntry) + 256))) return rc;
    HEntry *htab = static_cast<HEntry *>(pp.hash.ptr);
    if ((rc = fill_u32(reinterpret_cast<uint32_t *>(htab), hsz * 4, ~0u, st))) return rc;
    G.htab = htab;
    G.hmask = hsz - 1;
    k_cells_finish<<<grid_for(U), 256, 0, st>>>(small + kSmU, NN, start, par, runend, G, htab,
                                                label, dV, dVB, (int)nvol, T, dense,
                                                small + kSmNDense, dbig, small + kSmNBig);
    FOF_HIP(hipGetLastError());
  }
  if ((rc = tb(3))) return rc;
  k_init_uf<<<gN, 256, 0, st>>>(NN, par, runend);
  FOF_HIP(hipGetLastError());
  if ((rc = te(3))) return rc;
  // 4: sub-cell sort of dense cells (tmp4 aliases the sort buffers s0..s3);
  // grids for the most dense cells there can be, the lists' lengths are read
  // on the device
  const uint32_t gsml = (uint32_t)std::min<uint64_t>(pl.ndense_cap, 65536);
  const uint32_t gbig = (uint32_t)std::min<uint64_t>(N / kBigCell + 1, pl.big_grid);
  bool any_sub = false;
  for (uint32_t v = 0; v < nvol; v++) any_sub = any_sub || pl.hv[v].d.m[0] > 0;
  // constant l on a dense grid: k_link_sparse skips every dense cell, so only
  // the cells k_link_dense links need runs; otherwise it scans dense forward
  // cells run by run, and the smaller dense cells are sub-sorted too
  bool fast = !pl.sparse;
  for (uint32_t v = 0; v < nvol; v++) fast = fast && pl.hv[v].d.vntab == 0 && pl.hv[v].d.m[0] > 0;
  if (any_sub) {
    if ((rc = tb(4))) return rc;
    if (!fast) {
      k_subsort<256><<<gsml, 256, 0, st>>>(dense, small + kSmNDense, start, dV, dVB, (int)nvol,
                                           pos4, reinterpret_cast<float4 *>(sk0), par, runend,
                                           fine, rtoff, rtab);
      FOF_HIP(hipGetLastError());
    }
    k_subsort<256><<<gbig, 256, 0, st>>>(dbig, small + kSmNBig, start, dV, dVB, (int)nvol,
                                                pos4, reinterpret_cast<float4 *>(sk0), par, runend,
                                                fine, rtoff, rtab);
    FOF_HIP(hipGetLastError());
    if ((rc = te(4))) return rc;
  }
  // 5-6: linking
  if ((rc = tb(5))) return rc;
  if (fast)
    k_link_sparse<true><<<gN, 256, 0, st>>>(NN, pos4, G, runend, fine, par, dV, dVB, (int)nvol, T, ctr);
  else
    k_link_sparse<false><<<gN, 256, 0, st>>>(NN, pos4, G, runend, fine, par, dV, dVB, (int)nvol, T, ctr);
  FOF_HIP(hipGetLastError());
  if ((rc = te(5))) return rc;
  // the most sub-cells (so runs) a dense cell of the batch can have
  uint32_t maxbins = 1;
  for (uint32_t v = 0; v < nvol; v++)
    if (pl.hv[v].d.m[0] > 0)
      maxbins = std::max(maxbins, (uint32_t)(pl.hv[v].d.m[0] * pl.hv[v].d.m[1] * pl.hv[v].d.m[2]));
  if (any_sub) {
    if ((rc = tb(6))) return rc;
    constexpr int kRB = 448;
    const uint32_t ra = std::min(maxbins, (uint32_t)kMaxRuns);
    // rounds of 768 candidates: 256 and 384 gave 51.8 and 51.1 ms at 1024^3,
    // 512 and 768 49.5-50.6 and 49.2-49.4, 1024 50.5-50.7, 2048 59.5 (the
    // test queue overflows into in-place tests); a queue of 384 entries
    // 48.6-48.8 against 48.9-49.1 for 256 and 320, 49.1-49.2 for 512 and
    // 53-54 for 768 or 1024
    k_link_dense<256, 8, kMaxRuns, kRB, 384, 768, 1>
        <<<gbig, 256, dense_lds(ra, kRB), st>>>(
            dbig, small + kSmNBig, pos4, G, runend, fine, rtoff, rtab, T, par, dV, dVB, (int)nvol,
            ctr, small + kSmOverflow, ra, small + kSmDq);
    FOF_HIP(hipGetLastError());
    if ((rc = te(6))) return rc;
  }
  // dense cells below kBigCell particles, with or without sub-cells
  if ((rc = tb(15))) return rc;
  k_link_small<<<(gsml + kSmallWaves - 1) / kSmallWaves, 64 * kSmallWaves, 0, st>>>(
      dense, small + kSmNDense, pos4, G, T, par, dV, dVB, (int)nvol, ctr);
  FOF_HIP(hipGetLastError());
  if ((rc = te(15))) return rc;
  // 7: union-find finish (minorig reuses runend, hsize reuses key, the
  // members' minimum ID reuses fine)
  uint32_t *minorig = runend, *hsize = key;
  unsigned long long *idroot = pl.want_minid ? fine : nullptr;
  if ((rc = tb(7))) return rc;
  if ((rc = fill_u32(minorig, N, ~0u, st))) return rc;
  if ((rc = fill_u32(gsize, N, 0u, st))) return rc;
  if (idroot && (rc = fill_u32(reinterpret_cast<uint32_t *>(idroot), 2 * N, ~0u, st))) return rc;
  k_flatten<<<gN, 256, 0, st>>>(NN, par, pos4, minorig, gsize, ctr, dbase, stride, offid, idroot);
  FOF_HIP(hipGetLastError());
  if ((rc = te(7))) return rc;
  // 8: the group list -- groups of two or more (and kept groups of one) in
  // label order, each group's first slot in the member list -- then the
  // labels, which also fill the list.  The label bitmap and its word prefix
  // counts live where the cursors will (rtab), the sizes by label in the
  // label array (before its defaults), the first slots in sk1 and the labels
  // in order in sv1 (behind 64 words: R + 1 first slots may reach N + 1);
  // large groups, halo flags and the halo list go to the cell-sorted
  // positions once k_label has read them.
  const int single = pl.single_halos ? 1 : 0;
  const uint64_t rcap = std::max<uint64_t>(single ? N : N / 2, 1);
  const uint64_t nw = (N + 31) / 32;
  uint32_t *gbits = rtab, *wpre = rtab + nw, *lsize = label;
  uint32_t *offs = sk1, *glist = sv1 + 64;
  uint32_t *cursor = rtab;             // per label, from its group's first slot
  uint32_t *p32 = reinterpret_cast<uint32_t *>(pos4);
  uint32_t *large = p32 + 3 * N, *hflags = p32 + 2 * N;
  if ((rc = tb(9))) return rc;
  if ((rc = fill_u32(gbits, nw, 0u, st))) return rc;
  k_group_bits<<<gN, 256, 0, st>>>(NN, par, gsize, minorig, pos4, gbits, lsize, single, dV, dVB,
                                   (int)nvol);
  FOF_HIP(hipGetLastError());
  k_word_counts<<<grid_for(nw), 256, 0, st>>>(nw, gbits, wpre);
  FOF_HIP(hipGetLastError());
  if ((rc = scan_u32(wpre, wpre, nw, B.tmp, st))) return rc;
  k_group_list<<<grid_for(nw), 256, 0, st>>>(nw, gbits, wpre, lsize, glist, offs, small + kSmR);
  FOF_HIP(hipGetLastError());
  if ((rc = scan_u32_count(offs, offs, rcap, small + kSmR, B.tmp, st))) return rc;
  k_group_cursors<<<grid_for(rcap), 256, 0, st>>>(small + kSmR, glist, offs, cursor, small + kSmM);
  FOF_HIP(hipGetLastError());
  const bool direct = pl.direct_writes;
  // direct writes and no minimum IDs: k_label writes label_out itself and
  // k_out_init is not needed
  const bool fused = direct && !r.d_minid;
  k_label_init<<<tiles_for(N), 256, 0, st>>>(NN, dV, dVB, (int)nvol, fused ? nullptr : label,
                                             fused ? nullptr : hsize, fused ? r.d_label : nullptr,
                                             fused ? r.d_next : nullptr);
  FOF_HIP(hipGetLastError());
  if ((rc = te(9))) return rc;
  if ((rc = tb(8))) return rc;
  const uint32_t glab = (uint32_t)((N + kLabTile - 1) / kLabTile);
  if (direct)
    k_label<true><<<glab, 256, 0, st>>>(NN, par, pos4, minorig, gsize, label, hsize, cursor, sk0,
                                        sv0, dbase, stride, offid, idroot, B.idhead, single, dV,
                                        dVB, (int)nvol, fused ? r.d_label : nullptr);
  else
    k_label<false><<<glab, 256, 0, st>>>(NN, par, pos4, minorig, gsize, label, hsize, cursor, sk0,
                                         sv0, dbase, stride, offid, idroot, B.idhead, single, dV,
                                         dVB, (int)nvol, nullptr);
  FOF_HIP(hipGetLastError());
  if (!direct && (rc = scatter_u32(sv0, sk0, p32, p32 + N, N, small + kSmM, N, label, kFlag,
                                   B.tmp, st)))
    return rc;
  if (!fused) {
    k_out_init<<<tiles_for(N), 256, 0, st>>>(NN, label, dVB, (int)nvol, r.d_label, r.d_next,
                                             hsize, dbase, stride, offid, B.idhead, r.d_minid);
    FOF_HIP(hipGetLastError());
  }
  if ((rc = te(8))) return rc;
  // members in index order inside each group (large groups through the
  // free cell-sorted positions)
  if ((rc = tb(10))) return rc;
  k_seg_sort<<<(uint32_t)std::min<uint64_t>((rcap + 255) / 256, 8192), 256, 0, st>>>(
      small + kSmR, offs, sv0, large, small + kSmLarge);
  FOF_HIP(hipGetLastError());
  k_seg_sort_large<<<(uint32_t)std::min<uint64_t>(N / kSegWave + 1, 8192), 256, 0, st>>>(
      small + kSmLarge, large, offs, sv0, p32);
  FOF_HIP(hipGetLastError());
  if ((rc = te(10))) return rc;
  // 9: the rings
  if ((rc = tb(11))) return rc;
  {
    // ring pairs (the partitioned way) in the free cell-sorted positions
    uint32_t *pidx = (r.d_next && !direct) ? p32 : nullptr;
    k_ring<<<tiles_for(N), 256, 0, st>>>(small + kSmM, sk0, sv0, hsize, dV, dVB, (int)nvol,
                                        r.d_next, pidx, p32 + N, nullptr);
    FOF_HIP(hipGetLastError());
    if (pidx && (rc = scatter_u32(p32, p32 + N, p32 + 2 * N, p32 + 3 * N, N, small + kSmM, N,
                                  r.d_next, 0u, B.tmp, st)))
      return rc;
  }
  if ((rc = te(11))) return rc;
  if (pl.want_halos) {
    // halo starts: first slots of the groups passing the size cut, in label
    // order
    if ((rc = tb(12))) return rc;
    k_halo_flags<<<grid_for(rcap), 256, 0, st>>>(small + kSmR, glist, offs, dV, dVB, (int)nvol,
                                                  hflags);
    FOF_HIP(hipGetLastError());
    if ((rc = compact_flagged(hflags, rcap, small + kSmR, kFlag, p32, p32 + N, small + kSmH, B.tmp,
                              st)))
      return rc;
    if ((rc = te(12))) return rc;
    if ((rc = tb(13))) return rc;
    const uint32_t gh = (uint32_t)std::min<uint64_t>((pl.Hbound * 64 + 255) / 256, 8192);
    if (r.d_mom)
      k_halo_sums<true><<<gh, 256, 0, st>>>(small + kSmH, p32, sk0, sv0, hsize, dbase, stride,
                                            offpos, offvel, dV, dVB, (int)nvol, r.d_halo, r.d_mom,
                                            r.hcap, B.hhead, ctr);
    else
      k_halo_sums<false><<<gh, 256, 0, st>>>(small + kSmH, p32, sk0, sv0, hsize, dbase, stride,
                                             offpos, offvel, dV, dVB, (int)nvol, r.d_halo, nullptr,
                                             r.hcap, B.hhead, ctr);
    FOF_HIP(hipGetLastError());
    if ((rc = te(13))) return rc;
    k_halo_offsets<<<grid_for(nvol + 1), 256, 0, st>>>(small + kSmH, B.hhead, dVB, (int)nvol, B.hoff);
    FOF_HIP(hipGetLastError());
  }
  return 0;
}

// Counts, flags, counters and halo offsets back to the host (after the
// pipeline's kernels, on the same stream).
static int readback_run(Run &r, hipStream_t st) {
  if (r.pl.N == 0) return 0;
  FOF_HIP(hipMemcpyAsync(r.pp->hres.data(), r.B.small, kSmallBytes, hipMemcpyDeviceToHost, st));
  if (r.pl.want_halos)
    FOF_HIP(hipMemcpyAsync(r.pp->hres.data() + kSmallBytes, r.B.hoff, (r.pl.nvol + 1) * 8,
                           hipMemcpyDeviceToHost, st));
  return 0;
}

static uint32_t run_nhalos(const Run &r) {
  if (r.pl.N == 0 || !r.pl.want_halos) return 0;
  return reinterpret_cast<const uint32_t *>(r.pp->hres.data())[kSmH];
}

// Staged halos (and moments) of the pipelines of a call into the output, at
// each pipeline's prefix of the halo counts (read on the device).
struct HaloParts {
  const fof_halo_v1 *src[kMaxPipes];
  const fof_moments_v1 *msrc[kMaxPipes];
  const uint32_t *cnt[kMaxPipes];
  uint64_t stage[kMaxPipes];
  int k;
};

__global__ __launch_bounds__(256) void k_gather_halos(HaloParts hp, fof_halo_v1 *__restrict__ out,
                                                      fof_moments_v1 *__restrict__ mout,
                                                      uint64_t cap) {
  uint64_t base = 0;
  for (int g = 0; g < hp.k; g++) {
    const uint64_t h = hp.cnt[g] ? min((uint64_t)*hp.cnt[g], hp.stage[g]) : 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < h; i += (uint64_t)gridDim.x * 256) {
      if (base + i >= cap) break;
      if (out) out[base + i] = hp.src[g][i];
      if (mout) mout[base + i] = hp.msrc[g][i];
    }
    base += hp.cnt[g] ? *hp.cnt[g] : 0u;
  }
}

// ---------------------------------------------------------------------------
// Autotuning: dense-cell threshold, concurrent pipelines and the grid of the
// 256-thread dense kernels, per workload shape.
// ---------------------------------------------------------------------------
static uint32_t ilog2(uint64_t x) {
  uint32_t b = 0;
  while (x > 1) x >>= 1, b++;
  return b;
}

// Shape of a call: log2 of particles and volumes, addressing, variable l,
// particles per cell (x16, log2) and the linking length over the cell size
// (x64).  Calls with one key behave alike for the tuned parameters.
static TuneKey tune_key(const Plan &pl) {
  TuneKey k;
  k.lg_n = ilog2(pl.N);
  k.lg_nvol = ilog2(pl.nvol);
  k.sparse = pl.sparse ? 1 : 0;
  k.varlink = pl.varlink ? 1 : 0;
  k.ppc16 = ilog2((uint64_t)(16.0 * (double)pl.N / std::max(1.0, pl.Cd)) + 1);
  double b = 0;
  for (const HostVol &h : pl.hv) b += h.d.linklength / h.d.lcell[0];
  k.b64 = (uint32_t)std::lround(64.0 * b / std::max<size_t>(1, pl.hv.size()));
  return k;
}

static std::mutex g_tune_mu;

static const char *tune_path() {
  const char *p = getenv("FOF_TUNE_CACHE");
  return (p && *p) ? p : nullptr;
}

static bool tune_lookup(EngineState &S, const TuneKey &k, TuneVal &v) {
  for (auto &e : S.tuned)
    if (e.first == k) {
      v = e.second;
      return true;
    }
  const char *path = tune_path();
  if (!path) return false;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  FILE *f = fopen(path, "r");
  if (!f) return false;
  char line[256];
  bool found = false;
  while (fgets(line, sizeof line, f)) {
    TuneKey q;
    TuneVal w;
    if (sscanf(line, "%u %u %u %u %u %u -> %u %u %u", &q.lg_n, &q.lg_nvol, &q.sparse, &q.varlink,
               &q.ppc16, &q.b64, &w.T, &w.pipes, &w.gbig) == 9 &&
        q == k) {
      v = w;  // the last entry for a key wins
      found = true;
    }
  }
  fclose(f);
  if (found) S.tuned.push_back({k, v});
  return found;
}

static void tune_store(EngineState &S, const TuneKey &k, const TuneVal &v) {
  S.tuned.push_back({k, v});
  const char *path = tune_path();
  if (!path) return;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  FILE *f = fopen(path, "a");
  if (!f) return;
  fprintf(f, "%u %u %u %u %u %u -> %u %u %u\n", k.lg_n, k.lg_nvol, k.sparse, k.varlink, k.ppc16,
          k.b64, v.T, v.pipes, v.gbig);
  fclose(f);
}

}  // namespace fof

using namespace fof;

extern "C" int fof_engine_plan(const fof_particles_v1 *P, const fof_volume_v1 *vols,
                               uint32_t nvol, const fof_opts_v1 *opts, fof_plan_v1 *out) {
  Plan pl;
  fof_outputs_v1 o;
  memset(&o, 0, sizeof(o));
  o.location = FOF_MEM_DEVICE;
  o.label = reinterpret_cast<uint32_t *>(1);  // a typical call: ring, labels, halos
  o.next = reinterpret_cast<uint32_t *>(1);
  int rc = make_plan(P, vols, nvol, opts, &o, pl);
  if (rc) return rc;
  Bufs b;
  memset(out, 0, sizeof(*out));
  out->version = FOF_ABI_VERSION;
  out->sparse = pl.sparse ? 1u : 0u;
  out->n = pl.N;
  out->cells = pl.sparse ? 0 : pl.C;
  out->grid_cells = pl.Cd;
  out->workspace_bytes = layout(pl, nullptr, b);
  out->dense_threshold = pl.T;
  return 0;
}

namespace fof {

// FNV-1a over the bytes that fix a call's launch sequence.
struct Sig {
  uint64_t h = 1469598103934665603ull;
  void add(const void *p, size_t n) {
    const uint8_t *b = static_cast<const uint8_t *>(p);
    for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
  }
  template <class T>
  void add(const T &x) {
    add(&x, sizeof(T));
  }
};

// Volumes split into k contiguous groups of about equal particle counts.
static std::vector<uint32_t> split_volumes(const fof_volume_v1 *vols, uint32_t nvol, uint32_t k) {
  uint64_t tot = 0;
  for (uint32_t v = 0; v < nvol; v++) tot += vols[v].n;
  std::vector<uint32_t> cut{0};
  uint64_t acc = 0;
  for (uint32_t v = 0; v < nvol; v++) {
    acc += vols[v].n;
    const uint32_t g = (uint32_t)cut.size();
    // cut after v when this group has its share, or when every later group
    // needs one of the remaining volumes
    if (g < k && v + 1 < nvol && (acc * k >= tot * g || nvol - (v + 1) <= k - g)) cut.push_back(v + 1);
  }
  cut.push_back(nvol);
  return cut;
}

// One call with a fixed configuration: pipelines `pipes`, graph replay when
// asked for and possible.
static int run_call(fof_ctx *ctx, const fof_particles_v1 *P, const fof_volume_v1 *vols,
                    uint32_t nvol, const fof_opts_v1 *opts, fof_outputs_v1 *out, uint32_t pipes,
                    double *wall_ms, uint8_t *fixed = nullptr, size_t fixed_bytes = 0) {
  EngineState &S = ctx->st;
  const uint32_t flags = opts ? opts->flags : 0u;
  fof_stats_v1 *stats = out->stats;
  out->nhalos = 0;
  if (out->halo_offsets)
    for (uint32_t v = 0; v <= nvol; v++) out->halo_offsets[v] = 0;
  pipes = std::max(1u, std::min({pipes, nvol, (uint32_t)kMaxPipes}));
  const std::vector<uint32_t> cut = split_volumes(vols, nvol, pipes);
  const uint32_t K = (uint32_t)cut.size() - 1;
  if (S.pipes.empty()) S.pipes.push_back(nullptr);  // pipeline 0 is S.main
  while (S.pipes.size() < K) {
    S.pipes.push_back(new Pipe);
    FOF_HIP(hipStreamCreate(&S.pipes.back()->stream));
  }
  if (!S.sync_events) {  // created at context set-up; kept for safety
    FOF_HIP(hipEventCreateWithFlags(&S.fork, hipEventDisableTiming));
    for (auto &e : S.join) FOF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    FOF_HIP(hipEventCreate(&S.t0));
    FOF_HIP(hipEventCreate(&S.t1));
    S.sync_events = true;
  }
  std::vector<Run> runs(K);
  int rc = 0;
  const bool timing = (flags & FOF_OPT_TIMING) != 0;
  for (uint32_t g = 0; g < K; g++) {
    Pipe &pp = g == 0 ? S.main : *S.pipes[g];
    if ((rc = prepare_run(runs[g], pp, P, vols, cut[g], cut[g + 1] - cut[g], opts, out, K > 1,
                          fixed, fixed_bytes)))
      return rc;
  }
  const Plan &pl0 = runs[0].pl;
  uint64_t Ntot = 0;
  bool varlink = false, sparse = false, host_io = false;
  for (auto &r : runs) {
    Ntot += r.pl.N;
    varlink = varlink || r.pl.varlink;
    sparse = sparse || r.pl.sparse;
    host_io = host_io || r.pl.host_input || r.pl.host_out;
  }
  if (stats) {
    memset(stats, 0, sizeof(*stats));
    stats->version = FOF_ABI_VERSION;
    stats->nvol = nvol;
    stats->n = Ntot;
    stats->dense_threshold = pl0.T;
    stats->big_grid = pl0.big_grid;
    stats->sparse = sparse ? 1u : 0u;
    stats->pipelines = K;
    for (auto &r : runs) stats->ncells += r.pl.sparse ? 0 : r.pl.C;
  }
  if (Ntot == 0) return 0;
  hipStream_t st = S.main.stream;
  const bool want_halos = pl0.want_halos;
  const bool gather = K > 1 && want_halos && !pl0.host_out && (out->halos || out->moments);
  const bool graph = (flags & FOF_OPT_GRAPH) && !varlink && !sparse && !host_io;
  // the pipelines' kernels, forked from and joined back to the main stream
  auto enqueue_all = [&](bool with_timing) -> int {
    for (uint32_t g = 0; g < K; g++) runs[g].timing = with_timing && g == 0;
    if (K > 1) FOF_HIP(hipEventRecord(S.fork, st));
    for (uint32_t g = 0; g < K; g++) {
      hipStream_t sg = runs[g].pp->stream;
      if (g > 0) FOF_HIP(hipStreamWaitEvent(sg, S.fork, 0));
      int e = enqueue_run(runs[g], sg);
      if (e) return e;
      if (g > 0) {
        FOF_HIP(hipEventRecord(S.join[g], sg));
        FOF_HIP(hipStreamWaitEvent(st, S.join[g], 0));
      }
    }
    if (gather) {
      HaloParts hp;
      memset(&hp, 0, sizeof(hp));
      hp.k = (int)K;
      for (uint32_t g = 0; g < K; g++) {
        hp.src[g] = runs[g].B.st_halo;
        hp.msrc[g] = runs[g].B.st_mom;
        hp.cnt[g] = runs[g].pl.N ? runs[g].B.small + kSmH : nullptr;
        hp.stage[g] = runs[g].pl.Hstage;
      }
      k_gather_halos<<<1024, 256, 0, st>>>(hp, out->halos, pl0.want_mom ? out->moments : nullptr,
                                           out->halo_cap);
      FOF_HIP(hipGetLastError());
    }
    return 0;
  };
  // uploads, on the main stream ahead of everything
  for (auto &r : runs)
    if ((rc = upload_run(r, st))) return rc;
  FOF_HIP(hipEventRecord(S.t0, st));
  uint32_t graph_state = 0;
  if (graph) {
    Sig sig;
    sig.add(*P);
    sig.add(nvol);
    for (uint32_t v = 0; v < nvol; v++) sig.add(vols[v]);
    if (opts) sig.add(*opts);
    sig.add(out->location);
    sig.add(out->next), sig.add(out->label), sig.add(out->minid), sig.add(out->halos);
    sig.add(out->moments), sig.add(out->halo_cap);
    sig.add(K);
    for (auto &r : runs) sig.add(r.pp->arena.ptr);
    if (S.graph_exec && S.graph_sig == sig.h) {
      FOF_HIP(hipGraphLaunch(S.graph_exec, st));
      graph_state = 1;
    } else {
      if (S.graph_exec) (void)hipGraphExecDestroy(S.graph_exec);
      S.graph_exec = nullptr;
      FOF_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      rc = enqueue_all(false);
      hipGraph_t gr = nullptr;
      hipError_t e = hipStreamEndCapture(st, &gr);
      if (rc) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
      }
      if (e != hipSuccess) return fof_set_hip_error(e, "hipStreamEndCapture", __FILE__, __LINE__);
      e = hipGraphInstantiate(&S.graph_exec, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      if (e != hipSuccess) {
        S.graph_exec = nullptr;
        return fof_set_hip_error(e, "hipGraphInstantiate", __FILE__, __LINE__);
      }
      S.graph_sig = sig.h;
      FOF_HIP(hipGraphLaunch(S.graph_exec, st));
      graph_state = 2;
    }
  } else {
    if ((rc = enqueue_all(timing))) return rc;
  }
  FOF_HIP(hipEventRecord(S.t1, st));
  for (auto &r : runs)
    if ((rc = readback_run(r, st))) return rc;
  FOF_HIP(hipStreamSynchronize(st));
  float tot_ms = 0;
  FOF_HIP(hipEventElapsedTime(&tot_ms, S.t0, S.t1));
  if (wall_ms) *wall_ms = tot_ms;

  // ---- results ----
  uint64_t H = 0;
  bool dom_err = false, ovf = false;
  for (auto &r : runs) {
    if (r.pl.N == 0) continue;
    const uint32_t *hsm = reinterpret_cast<const uint32_t *>(r.pp->hres.data());
    dom_err = dom_err || hsm[kSmErr];
    ovf = ovf || hsm[kSmOverflow];
    H += run_nhalos(r);
  }
  out->nhalos = H;
  if (dom_err) return FOF_E_DOMAIN;
  if (ovf) return FOF_E_RANGE;  // run table overflow (cannot happen with m <= 12)
  if (want_halos && H > out->halo_cap && (out->halos || out->moments)) {
    if (out->halo_offsets) out->halo_offsets[nvol] = H;  // the capacity needed
    if (stats) stats->nhalos = H;
    return FOF_E_RANGE;
  }
  if (out->halo_offsets && want_halos) {
    uint64_t base = 0;
    for (auto &r : runs) {
      const uint64_t *ho = reinterpret_cast<const uint64_t *>(r.pp->hres.data() + kSmallBytes);
      for (uint32_t v = 0; v < r.pl.nvol; v++) out->halo_offsets[r.v0 + v] = base + (r.pl.N ? ho[v] : 0);
      base += run_nhalos(r);
    }
    out->halo_offsets[nvol] = H;
  }
  // host outputs
  if (pl0.host_out) {
    uint64_t hb = 0;
    for (auto &r : runs) {
      const uint64_t n = r.pl.N, h = run_nhalos(r);
      if (n) {
        if (r.pl.want_ring)
          FOF_HIP(hipMemcpyAsync(out->next + r.poff, r.B.st_next, n * 4, hipMemcpyDeviceToHost, st));
        if (r.pl.want_label)
          FOF_HIP(hipMemcpyAsync(out->label + r.poff, r.B.st_label, n * 4, hipMemcpyDeviceToHost, st));
        if (r.pl.want_minid)
          FOF_HIP(hipMemcpyAsync(out->minid + r.poff, r.B.st_minid, n * 8, hipMemcpyDeviceToHost, st));
        if (out->halos && h)
          FOF_HIP(hipMemcpyAsync(out->halos + hb, r.B.st_halo, h * sizeof(fof_halo_v1),
                                 hipMemcpyDeviceToHost, st));
        if (out->moments && h)
          FOF_HIP(hipMemcpyAsync(out->moments + hb, r.B.st_mom, h * sizeof(fof_moments_v1),
                                 hipMemcpyDeviceToHost, st));
      }
      hb += h;
    }
    FOF_HIP(hipStreamSynchronize(st));
  }
  if (stats) {
    unsigned long long hc[kCtrStride] = {};
    for (auto &r : runs) {
      if (r.pl.N == 0) continue;
      const uint32_t *hsm = reinterpret_cast<const uint32_t *>(r.pp->hres.data());
      const unsigned long long *hcs = reinterpret_cast<const unsigned long long *>(hsm + kSmCtr);
      for (int q = 0; q < kCtrSlots; q++)
        for (int k = 0; k < kCtrStride; k++) hc[k] += hcs[q * kCtrStride + k];
      stats->ndense += (uint64_t)hsm[kSmNDense] + hsm[kSmNBig];
      stats->nmembers += hsm[kSmM];
      if (r.pl.sparse) stats->ncells += hsm[kSmU];
    }
    stats->pair_tests_kernel[0] = hc[kCtrTestsSparse];
    stats->pair_tests_kernel[1] = hc[kCtrTestsBig];
    stats->pair_tests_kernel[2] = hc[kCtrTestsSmall];
    stats->pair_tests = hc[kCtrTestsSparse] + hc[kCtrTestsBig] + hc[kCtrTestsSmall];
    stats->pairs_joined = hc[kCtrJoined];
    stats->rechecks = hc[kCtrRechecks];
    stats->ngroups = hc[kCtrGroups];
    stats->nhalos = H;
    stats->nhalos_owned = hc[kCtrOwned];
    stats->graph = graph_state;
    if (timing) {
      const Run &r = runs[0];
      Pipe &pp = *r.pp;
      float ms;
      if (!graph_state)
        for (int k = 0; k < FOF_NPHASE - 1; k++) {
          if (!(r.spans & (1u << k))) continue;
          FOF_HIP(hipEventElapsedTime(&ms, pp.ev[2 * k], pp.ev[2 * k + 1]));
          stats->phase_ms[k] = ms;
        }
      stats->phase_ms[FOF_NPHASE - 1] = tot_ms;
      stats->link_kernel_ms = stats->phase_ms[5] + stats->phase_ms[6] + stats->phase_ms[15];
    }
  }
  return 0;
}

}  // namespace fof

extern "C" int fof_engine_run(fof_ctx *ctx, const fof_particles_v1 *P, const fof_volume_v1 *vols,
                              uint32_t nvol, const fof_opts_v1 *opts, fof_outputs_v1 *out) {
  EngineState &S = ctx->st;
  fof_opts_v1 o;
  memset(&o, 0, sizeof(o));
  o.version = FOF_ABI_VERSION;
  if (opts) o = *opts;
  uint32_t pipes = o.pipelines ? o.pipelines : 1u;
  if (!(o.flags & FOF_OPT_AUTOTUNE)) return run_call(ctx, P, vols, nvol, &o, out, pipes, nullptr);
  // Autotuned: the shape's cached choice, or a search on this call's own
  // input (every candidate links the same groups, so the last run's outputs
  // are the call's outputs).
  Plan pl;
  fof_outputs_v1 probe = *out;
  int rc = make_plan(P, vols, nvol, &o, &probe, pl);
  if (rc) return rc;
  const TuneKey key = tune_key(pl);
  TuneVal best;
  if (!tune_lookup(S, key, best)) {
    fof_stats_v1 *user_stats = out->stats;
    out->stats = nullptr;
    auto trial = [&](uint32_t T, uint32_t K, uint32_t gb, double &ms) -> int {
      fof_opts_v1 q = o;
      q.flags &= ~(FOF_OPT_TIMING | FOF_OPT_GRAPH);
      q.dense_threshold = T;
      q.big_grid = gb;
      double best_ms = 1e300;
      for (int rep = 0; rep < 2; rep++) {  // the first run also warms the workspace
        double w = 0;
        int e = run_call(ctx, P, vols, nvol, &q, out, K, &w);
        if (e) return e;
        best_ms = std::min(best_ms, w);
      }
      ms = best_ms;
      return 0;
    };
    best = TuneVal{kDefaultT, 1, 0};
    double bms = 1e300, ms;
    for (uint32_t T : {16u, 24u, 32u}) {
      if ((rc = trial(T, 1, 0, ms))) break;
      if (ms < bms) bms = ms, best.T = T;
    }
    for (uint32_t gb : {2048u, 32768u}) {
      if (rc || (rc = trial(best.T, 1, gb, ms))) break;
      if (ms < bms * 0.99) bms = ms, best.gbig = gb;
    }
    for (uint32_t K : {2u, 3u, 4u}) {
      if (rc || K > nvol) break;
      if ((rc = trial(best.T, K, best.gbig, ms))) break;
      if (ms < bms * 0.99) bms = ms, best.pipes = K;
    }
    out->stats = user_stats;
    if (rc) return rc;
    tune_store(S, key, best);
  }
  o.dense_threshold = best.T;
  o.big_grid = best.gbig;
  return run_call(ctx, P, vols, nvol, &o, out, best.pipes, nullptr);
}

// ---------------------------------------------------------------------------
// Fixed workspace for callers that cannot allocate (fof_link_gpu_noalloc).
// Each device holds its own copy of the array from the time the library's
// code is loaded on it until the process ends; host code must use the
// address hipGetSymbolAddress returns for the device, never the array's name.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocBytes = (size_t)256 << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {
std::mutex g_scratch_mu;
uint8_t *g_scratch[64];  // per device, from hipGetSymbolAddress
}  // namespace

extern "C" int fof_engine_init(fof_ctx *ctx) {
  EngineState &S = ctx->st;
  FOF_HIP(hipStreamCreate(&S.main.stream));
  // everything a later call may need besides workspace: the fork / join and
  // timing events, and the fixed workspace's address on this device
  FOF_HIP(hipEventCreateWithFlags(&S.fork, hipEventDisableTiming));
  for (auto &e : S.join) FOF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  FOF_HIP(hipEventCreate(&S.t0));
  FOF_HIP(hipEventCreate(&S.t1));
  S.sync_events = true;
  if (ctx->device >= 0 && ctx->device < 64) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[ctx->device]) {
      void *ptr = nullptr;
      FOF_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
      g_scratch[ctx->device] = static_cast<uint8_t *>(ptr);
    }
  }
  return 0;
}

extern "C" int fof_engine_run_noalloc(fof_ctx *ctx, const fof_particles_v1 *P,
                                      const fof_volume_v1 *vol, const fof_opts_v1 *opts,
                                      fof_outputs_v1 *out) {
  if (ctx->device < 0 || ctx->device >= 64 || !g_scratch[ctx->device]) return FOF_E_NODEV;
  if (P->location != FOF_MEM_DEVICE || out->location != FOF_MEM_DEVICE) return FOF_E_ARG;
  if (vol->domain.varlink) return FOF_E_ARG;
  fof_opts_v1 o;
  memset(&o, 0, sizeof(o));
  o.version = FOF_ABI_VERSION;
  if (opts) o = *opts;
  // no graph (instantiation allocates), no events, one pipeline, no tuning
  o.flags &= ~(FOF_OPT_TIMING | FOF_OPT_GRAPH | FOF_OPT_AUTOTUNE);
  o.pipelines = 1;
  return run_call(ctx, P, vol, 1, &o, out, 1, nullptr, g_scratch[ctx->device], kNoAllocBytes);
}

extern "C" uint64_t fof_engine_noalloc_bytes(void) { return kNoAllocBytes; }

static void free_pipe(Pipe &pp) {
  release(pp.arena);
  release(pp.hash);
  if (pp.events)
    for (auto &e : pp.ev) (void)hipEventDestroy(e);
  pp.events = false;
  if (pp.stream) (void)hipStreamDestroy(pp.stream);
  pp.stream = nullptr;
}

extern "C" void fof_engine_free(fof_ctx *ctx) {
  EngineState &S = ctx->st;
  if (S.graph_exec) (void)hipGraphExecDestroy(S.graph_exec);
  S.graph_exec = nullptr;
  free_pipe(S.main);
  for (auto &p : S.pipes) {
    if (!p) continue;
    free_pipe(*p);
    delete p;
  }
  S.pipes.clear();
  if (S.sync_events) {
    (void)hipEventDestroy(S.fork);
    for (auto &e : S.join) (void)hipEventDestroy(e);
    (void)hipEventDestroy(S.t0);
    (void)hipEventDestroy(S.t1);
  }
  S.sync_events = false;
}
