// fof_prims.hip -- device-wide primitives for the FoF engine: exclusive scan,
// stable compaction and stable LSD radix sort of u32 (key, value) pairs.
// All kernels use 256-thread blocks (four wave64s) and 4096-element tiles.
#include "fof_internal.h"

namespace fof {

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;  // 4096

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// Exclusive block scan over 256 threads; *total receives the block sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *ws,
                                                    uint32_t *total) {
  const uint32_t inc = wave_incl_scan(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) ws[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kBlock / 64; i++) {
    uint32_t s = ws[i];
    off += (i < w) ? s : 0u;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

__device__ __forceinline__ uint32_t block_reduce(uint32_t v, uint32_t *ws) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
  __syncthreads();
  uint32_t t = 0;
#pragma unroll
  for (int i = 0; i < kBlock / 64; i++) t += ws[i];
  __syncthreads();
  return t;
}

// Element count of a pass: the host-known capacity, or the smaller count
// that an earlier kernel left in device memory (dn != nullptr), so that a
// sequence of launches needs no host round trip between them.
__device__ __forceinline__ uint64_t dev_count(const uint32_t *dn, uint64_t cap) {
  return dn ? min((uint64_t)*dn, cap) : cap;
}

// ---- scan ------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t *__restrict__ in,
                                                        uint64_t ncap, const uint32_t *dn,
                                                        uint32_t *__restrict__ bsum) {
  __shared__ uint32_t ws[kBlock / 64];
  const uint64_t n = dev_count(dn, ncap);
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
    if (idx < n) s += in[idx];
  }
  s = block_reduce(s, ws);
  if (threadIdx.x == 0) bsum[blockIdx.x] = s;
}

// in may alias out.  boff[b] is the exclusive prefix of tile b; boff[nb] is
// the grand total, copied to out[n] by the last block.
__global__ __launch_bounds__(kBlock) void k_scan_down(const uint32_t *in, uint32_t *out,
                                                      uint64_t ncap, const uint32_t *dn,
                                                      const uint32_t *__restrict__ boff) {
  __shared__ uint32_t ws[kBlock / 64];
  const uint64_t n = dev_count(dn, ncap);
  if ((uint64_t)blockIdx.x * kTile > n) return;  // past the count (block sums are 0)
  const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kItems;
  uint32_t v[kItems];
  if (base + kItems <= n) {
    const uint4 *p = reinterpret_cast<const uint4 *>(in + base);
#pragma unroll
    for (int q = 0; q < kItems / 4; q++) {
      uint4 t = p[q];
      v[4 * q] = t.x, v[4 * q + 1] = t.y, v[4 * q + 2] = t.z, v[4 * q + 3] = t.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < kItems; i++) v[i] = (base + i < n) ? in[base + i] : 0u;
  }
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kItems; i++) s += v[i];
  uint32_t tot;
  uint32_t pre = block_excl_scan(s, ws, &tot) + boff[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    uint32_t x = v[i];
    v[i] = pre;
    pre += x;
  }
  if (base + kItems <= n) {
    uint4 *p = reinterpret_cast<uint4 *>(out + base);
#pragma unroll
    for (int q = 0; q < kItems / 4; q++)
      p[q] = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
  } else {
#pragma unroll
    for (int i = 0; i < kItems; i++)
      if (base + i < n) out[base + i] = v[i];
  }
  if (threadIdx.x == 0 && (uint64_t)blockIdx.x == min(n / kTile, (uint64_t)gridDim.x - 1))
    out[n] = boff[gridDim.x];
}

__global__ void k_scan_single(uint32_t *bsum, uint32_t nb) {
  // nb <= kTile: one block scans in place and appends the total.
  __shared__ uint32_t ws[kBlock / 64];
  const uint32_t base = threadIdx.x * kItems;
  uint32_t v[kItems], s = 0;
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    v[i] = (base + i < nb) ? bsum[base + i] : 0u;
    s += v[i];
  }
  uint32_t tot;
  uint32_t pre = block_excl_scan(s, ws, &tot);
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    if (base + i < nb) bsum[base + i] = pre;
    pre += v[i];
  }
  if (threadIdx.x == 0) bsum[nb] = tot;
}

// ---- compaction -------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_compact_count(const uint32_t *__restrict__ src,
                                                          uint64_t ncap, const uint32_t *dn,
                                                          uint32_t flag,
                                                          uint32_t *__restrict__ bsum) {
  __shared__ uint32_t ws[kBlock / 64];
  const uint64_t n = dev_count(dn, ncap);
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
    if (idx < n) s += (src[idx] & flag) ? 1u : 0u;
  }
  s = block_reduce(s, ws);
  if (threadIdx.x == 0) bsum[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_compact_emit(const uint32_t *__restrict__ src,
                                                         uint64_t ncap, const uint32_t *dn,
                                                         uint32_t flag,
                                                         const uint32_t *__restrict__ boff,
                                                         uint32_t *__restrict__ okey,
                                                         uint32_t *__restrict__ oval,
                                                         uint32_t *__restrict__ dcount) {
  __shared__ uint32_t ws[kBlock / 64];
  const uint64_t n = dev_count(dn, ncap);
  const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kItems;
  uint32_t v[kItems], s = 0;
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    v[i] = (base + i < n) ? src[base + i] : 0u;
    s += (v[i] & flag) ? 1u : 0u;
  }
  uint32_t tot;
  uint32_t pre = block_excl_scan(s, ws, &tot) + boff[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    if (v[i] & flag) {
      if (okey) okey[pre] = v[i] & ~flag;
      oval[pre] = (uint32_t)(base + i);
      pre++;
    }
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *dcount = boff[gridDim.x];
}

// ---- radix sort -------------------------------------------------------------
// Histograms are digit-major, hist[d * nb + b], with nb the tiles of the
// count (device-side when dn is given); *hn receives 256 nb, the length of
// the scan that follows.  Blocks past the count write nothing.
__global__ __launch_bounds__(kBlock) void k_rs_hist(const uint32_t *__restrict__ keys,
                                                    uint64_t ncap, const uint32_t *dn, int shift,
                                                    uint32_t *__restrict__ hist,
                                                    uint32_t *__restrict__ hn) {
  __shared__ uint32_t h[4][256];
  const uint64_t n = dev_count(dn, ncap);
  const uint32_t nb = (uint32_t)((n + kTile - 1) / kTile);
  if (blockIdx.x == 0 && threadIdx.x == 0) *hn = 256u * nb;
  if (blockIdx.x >= nb) return;
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 4; q++) h[q][threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
#pragma unroll
  for (int i = 0; i < kItems; i++) {
    uint64_t idx = base + (uint64_t)i * kBlock + threadIdx.x;
    if (idx < n) atomicAdd(&h[w][(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * nb + blockIdx.x] =
      h[0][threadIdx.x] + h[1][threadIdx.x] + h[2][threadIdx.x] + h[3][threadIdx.x];
}

// Stable scatter.  Wave w of the block owns elements [w*1024, w*1024+1024) of
// the tile and walks them 64 at a time; lanes holding the same digit find one
// another with eight ballots, so ranks follow element order.  The tile is
// then laid out in LDS in digit order and written out digit run by digit
// run, so that consecutive threads write consecutive addresses.
__global__ __launch_bounds__(kBlock) void k_rs_scatter(const uint32_t *__restrict__ kin,
                                                       const uint32_t *__restrict__ vin,
                                                       uint32_t *__restrict__ kout,
                                                       uint32_t *__restrict__ vout,
                                                       uint64_t ncap, const uint32_t *dn,
                                                       int shift,
                                                       const uint32_t *__restrict__ hoff) {
  const uint64_t n = dev_count(dn, ncap);
  const uint32_t nb = (uint32_t)((n + kTile - 1) / kTile);
  if (blockIdx.x >= nb) return;
  __shared__ uint32_t wcnt[4][256];
  __shared__ uint32_t boff[256];   // global start of each digit for this tile
  __shared__ uint32_t toff[257];   // start of each digit inside the tile
  __shared__ uint32_t lk[kTile], lv[kTile];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < 4; q++) wcnt[q][threadIdx.x] = 0;
  boff[threadIdx.x] = hoff[(uint64_t)threadIdx.x * nb + blockIdx.x];
  __syncthreads();
  const uint64_t tbase = (uint64_t)blockIdx.x * kTile;
  const uint64_t chunk = tbase + (uint64_t)w * (kTile / 4);
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  uint32_t kk[kItems], vv[kItems], rk[kItems];
#pragma unroll
  for (int r = 0; r < kItems; r++) {
    const uint64_t idx = chunk + (uint64_t)r * 64 + lane;
    const bool valid = idx < n;
    const uint32_t k = valid ? kin[idx] : 0u;
    kk[r] = k;
    vv[r] = valid ? vin[idx] : 0u;
    const uint32_t d = (k >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t rin = (uint32_t)__popcll(peers & lt);
    const uint32_t cnt = (uint32_t)__popcll(peers);
    uint32_t old = 0;
    if (valid) old = wcnt[w][d];
    if (valid && rin == 0) wcnt[w][d] = old + cnt;
    rk[r] = valid ? old + rin : kNone;
  }
  __syncthreads();
  // per digit: offsets of the waves inside the digit, and the digit's total
  uint32_t dtot;
  {
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint32_t c = wcnt[q][threadIdx.x];
      wcnt[q][threadIdx.x] = s;
      s += c;
    }
    dtot = s;
  }
  // exclusive scan of the digit totals over the block -> toff
  {
    uint32_t inc = dtot;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t t = __shfl_up(inc, o, 64);
      if (lane >= o) inc += t;
    }
    __shared__ uint32_t wsum[4];
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t off = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) off += (q < w) ? wsum[q] : 0u;
    toff[threadIdx.x] = off + inc - dtot;
    if (threadIdx.x == 255) toff[256] = off + inc;
  }
  __syncthreads();
  // tile in digit order in LDS
#pragma unroll
  for (int r = 0; r < kItems; r++) {
    if (rk[r] != kNone) {
      const uint32_t d = (kk[r] >> shift) & 255u;
      const uint32_t j = toff[d] + wcnt[w][d] + rk[r];
      lk[j] = kk[r];
      lv[j] = vv[r];
    }
  }
  __syncthreads();
  // write out: element j of the tile goes to its digit's run
  const uint32_t valid = toff[256];
  for (uint32_t j = threadIdx.x; j < valid; j += kBlock) {
    const uint32_t k = lk[j];
    const uint32_t d = (k >> shift) & 255u;
    const uint32_t pos = boff[d] + (j - toff[d]);
    kout[pos] = k;
    vout[pos] = lv[j];
  }
}

uint64_t scan_levels_bytes(uint64_t n) {
  uint64_t b = 0;
  for (;;) {
    uint64_t nb = (n + kTile - 1) / kTile;
    b += (nb + 1 + 63) / 64 * 64 * sizeof(uint32_t);
    if (nb <= (uint64_t)kTile) break;
    n = nb;
  }
  return b;
}

// Scan of the per-tile sums held in tmp[0..nb) in place (tmp[nb] = total);
// deeper levels use the space behind.
int scan_bsums(uint32_t *bs, uint64_t nb, hipStream_t st) {
  if (nb <= (uint64_t)kTile) {
    k_scan_single<<<1, kBlock, 0, st>>>(bs, (uint32_t)nb);
    FOF_HIP(hipGetLastError());
    return 0;
  }
  uint32_t *next = bs + (nb + 1 + 63) / 64 * 64;
  uint64_t nb2 = (nb + kTile - 1) / kTile;
  k_scan_reduce<<<(uint32_t)nb2, kBlock, 0, st>>>(bs, nb, nullptr, next);
  FOF_HIP(hipGetLastError());
  int rc = scan_bsums(next, nb2, st);
  if (rc) return rc;
  k_scan_down<<<(uint32_t)nb2, kBlock, 0, st>>>(bs, bs, nb, nullptr, next);
  FOF_HIP(hipGetLastError());
  return 0;
}

// Exclusive scan of min(*dn, n) values (dn may be null); out[count] = total.
int scan_u32_dev(const uint32_t *in, uint32_t *out, uint64_t n, const uint32_t *dn, void *tmp,
                 hipStream_t st) {
  if (n == 0) return fill_u32(out, 1, 0u, st);
  uint32_t *bs = static_cast<uint32_t *>(tmp);
  uint64_t nb = (n + kTile - 1) / kTile;
  k_scan_reduce<<<(uint32_t)nb, kBlock, 0, st>>>(in, n, dn, bs);
  FOF_HIP(hipGetLastError());
  int rc = scan_bsums(bs, nb, st);
  if (rc) return rc;
  k_scan_down<<<(uint32_t)nb, kBlock, 0, st>>>(in, out, n, dn, bs);
  FOF_HIP(hipGetLastError());
  return 0;
}

}  // namespace

size_t scan_tmp_bytes(uint64_t n) { return scan_levels_bytes(n) + 256; }

int scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, void *tmp, hipStream_t st) {
  return scan_u32_dev(in, out, n, nullptr, tmp, st);
}

int scan_u32_count(const uint32_t *in, uint32_t *out, uint64_t n, const uint32_t *dn, void *tmp,
                   hipStream_t st) {
  return scan_u32_dev(in, out, n, dn, tmp, st);
}

size_t compact_tmp_bytes(uint64_t n) { return scan_levels_bytes(n) + 256; }

int compact_flagged(const uint32_t *src, uint64_t n, const uint32_t *dn, uint32_t flag,
                    uint32_t *okey, uint32_t *oval, uint32_t *dcount, void *tmp,
                    hipStream_t st) {
  if (n == 0) return fill_u32(dcount, 1, 0u, st);
  uint32_t *bs = static_cast<uint32_t *>(tmp);
  uint64_t nb = (n + kTile - 1) / kTile;
  k_compact_count<<<(uint32_t)nb, kBlock, 0, st>>>(src, n, dn, flag, bs);
  FOF_HIP(hipGetLastError());
  int rc = scan_bsums(bs, nb, st);
  if (rc) return rc;
  k_compact_emit<<<(uint32_t)nb, kBlock, 0, st>>>(src, n, dn, flag, bs, okey, oval, dcount);
  FOF_HIP(hipGetLastError());
  return 0;
}

size_t sort_tmp_bytes(uint64_t n) {
  uint64_t nb = (n + kTile - 1) / kTile;
  uint64_t h = 256 * nb;
  return (h + 64 + 64) * sizeof(uint32_t) + scan_tmp_bytes(h) + 256;
}

__global__ __launch_bounds__(kBlock) void k_copy_pairs(uint32_t *__restrict__ k0,
                                                       uint32_t *__restrict__ v0,
                                                       const uint32_t *__restrict__ k1,
                                                       const uint32_t *__restrict__ v1,
                                                       uint64_t ncap, const uint32_t *dn) {
  const uint64_t n = dev_count(dn, ncap);
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBlock) {
    k0[i] = k1[i];
    v0[i] = v1[i];
  }
}

int radix_sort_u32(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, uint64_t n,
                   const uint32_t *dn, int bits, void *tmp, hipStream_t st) {
  if (n <= 1 || bits <= 0) return 0;
  const uint64_t nb = (n + kTile - 1) / kTile;
  uint32_t *hist = static_cast<uint32_t *>(tmp);
  uint32_t *hn = hist + 256 * nb + 64;  // histogram length, set on the device
  void *stmp = hn + 64;
  uint32_t *ka = k0, *va = v0, *kb = k1, *vb = v1;
  int passes = 0;
  for (int shift = 0; shift < bits; shift += 8, passes++) {
    k_rs_hist<<<(uint32_t)nb, kBlock, 0, st>>>(ka, n, dn, shift, hist, hn);
    FOF_HIP(hipGetLastError());
    int rc = scan_u32_dev(hist, hist, 256 * nb, hn, stmp, st);
    if (rc) return rc;
    k_rs_scatter<<<(uint32_t)nb, kBlock, 0, st>>>(ka, va, kb, vb, n, dn, shift, hist);
    FOF_HIP(hipGetLastError());
    uint32_t *t = ka;
    ka = kb, kb = t;
    t = va, va = vb, vb = t;
  }
  if (passes & 1) {
    k_copy_pairs<<<(uint32_t)std::min<uint64_t>(nb * 4, 65536), kBlock, 0, st>>>(k0, v0, ka, va, n,
                                                                                 dn);
    FOF_HIP(hipGetLastError());
  }
  return 0;
}

// ---- scatter through one partition pass -------------------------------------
// One tile of kTile consecutive pairs per block: the blocks in flight then
// cover one stretch of the partitioned list (a grid-stride loop would spread
// them over many partitions).
__global__ __launch_bounds__(kBlock) void k_put(const uint32_t *__restrict__ idx,
                                                const uint32_t *__restrict__ val, uint64_t ncap,
                                                const uint32_t *dn, uint32_t orv,
                                                uint32_t *__restrict__ out) {
  const uint64_t n = dev_count(dn, ncap);
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
#pragma unroll
  for (int r = 0; r < kItems; r++) {
    const uint64_t i = base + (uint64_t)r * kBlock + threadIdx.x;
    if (i < n) out[idx[i]] = val[i] | orv;
  }
}

int scatter_u32(const uint32_t *idx, const uint32_t *val, uint32_t *idx2, uint32_t *val2,
                uint64_t n, const uint32_t *dn, uint64_t nout, uint32_t *out, uint32_t orv,
                void *tmp, hipStream_t st) {
  if (n == 0) return 0;
  int bits = 0;
  while (bits < 32 && (nout - 1) >> bits) bits++;
  const int shift = bits > 8 ? bits - 8 : 0;
  const uint64_t nb = (n + kTile - 1) / kTile;
  uint32_t *hist = static_cast<uint32_t *>(tmp);
  uint32_t *hn = hist + 256 * nb + 64;
  void *stmp = hn + 64;
  k_rs_hist<<<(uint32_t)nb, kBlock, 0, st>>>(idx, n, dn, shift, hist, hn);
  FOF_HIP(hipGetLastError());
  int rc = scan_u32_dev(hist, hist, 256 * nb, hn, stmp, st);
  if (rc) return rc;
  k_rs_scatter<<<(uint32_t)nb, kBlock, 0, st>>>(idx, val, idx2, val2, n, dn, shift, hist);
  FOF_HIP(hipGetLastError());
  k_put<<<(uint32_t)nb, kBlock, 0, st>>>(idx2, val2, n, dn, orv, out);
  FOF_HIP(hipGetLastError());
  return 0;
}

// ---- fills ------------------------------------------------------------------
// Kernels rather than hipMemsetAsync, so that a captured call sequence holds
// kernel nodes only.
__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t *__restrict__ p, uint64_t n,
                                                     uint32_t v, int vec) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  const uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  uint64_t done = 0;
  if (vec) {  // p is 16-byte aligned
    const uint64_t n4 = n / 4;
    uint4 *q = reinterpret_cast<uint4 *>(p);
    const uint4 w = make_uint4(v, v, v, v);
    for (uint64_t i = t; i < n4; i += stride) q[i] = w;
    done = n4 * 4;
  }
  for (uint64_t i = done + t; i < n; i += stride) p[i] = v;
}

int fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st) {
  if (n == 0) return 0;
  const int vec = (reinterpret_cast<uintptr_t>(p) & 15u) == 0 ? 1 : 0;
  const uint64_t per = vec ? 4 : 1;
  const uint64_t blocks = std::min<uint64_t>((n / per + kBlock - 1) / kBlock + 1, 16384);
  k_fill_u32<<<(uint32_t)blocks, kBlock, 0, st>>>(p, n, v, vec);
  FOF_HIP(hipGetLastError());
  return 0;
}

int ensure(Scratch &s, size_t bytes) {
  if (bytes <= s.bytes) return 0;
  if (s.ptr) FOF_HIP(hipFree(s.ptr));
  s.ptr = nullptr;
  s.bytes = 0;
  size_t b = (bytes + 4095) / 4096 * 4096;
  hipError_t e = hipMalloc(&s.ptr, b);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    s.ptr = nullptr;
    return FOF_E_NOMEM;
  }
  s.bytes = b;
  return 0;
}

void release(Scratch &s) {
  if (s.ptr) (void)hipFree(s.ptr);
  s.ptr = nullptr;
  s.bytes = 0;
}

}  // namespace fof
