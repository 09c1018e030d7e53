// fof_internal.h -- declarations shared by the HIP translation units of the
// MI355X Friends-of-Friends engine.  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/fof_gpu.h"

#define FOF_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) return fof_set_hip_error(e_, #call, __FILE__, __LINE__); \
  } while (0)

int fof_set_hip_error(hipError_t e, const char *what, const char *file, int line);

namespace fof {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kFlag = 0x80000000u;  // "member of a group that is kept"
constexpr int kMaxSub = 16;              // sub-cells per axis inside a cell

// Per-volume constants, read by every kernel (global memory, scalar loads).
struct alignas(16) VolDev {
  double rmin_buff[3], rmax_buff[3], rmin[3], rmax[3], lcell[3];
  double linklength, linksq, nphalomin;
  double s2[3];        // squared sub-cell side per axis
  double vorigin[3];   // variable linking length: origin, r2max, table
  double vr2max;
  const double *vtab;
  uint32_t vntab;      // 0 = constant linking length
  float lo, hi;        // FP32 decision thresholds around linksq (constant l)
  float band;
  int nx, ny, nz;
  int m[3];            // sub-cells per axis; m[0] == 0 disables sub-cells
  uint32_t pbeg, pend; // particle range in the batch (input order and
                       // cell-sorted order coincide per volume)
  uint32_t cbeg, ncell; // dense cell addressing: first cell, cell count
  uint32_t vid;         // volume index (sparse cell keys)
};

// Device-wide primitives (fof_prims.hip).
struct Scratch {
  // grow-only device buffer
  void *ptr = nullptr;
  size_t bytes = 0;
};
int ensure(Scratch &s, size_t bytes);
void release(Scratch &s);

// Exclusive scan of n u32 values in[0..n) into out[0..n]; out[n] = total.
// tmp must hold scan_tmp_bytes(n).
size_t scan_tmp_bytes(uint64_t n);
int scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, void *tmp,
             hipStream_t st);

// Stable LSD radix sort of (key,val) u32 pairs on bits [0,bits).  Result is
// left in (k0,v0); (k1,v1) are ping-pong buffers of the same size.
size_t sort_tmp_bytes(uint64_t n);
int radix_sort_u32(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1,
                   uint64_t n, int bits, void *tmp, hipStream_t st);

// out[idx[i]] = val[i] for i in [0,n), every idx[i] < nout: the pairs are
// first partitioned (one radix pass) by the top 8 bits of idx, so that the
// writes of a stretch of the list fall in one 1/256 of out and merge in the
// caches (random 4-byte writes cost one memory request each).  (idx2, val2)
// receive the partitioned pairs; tmp: sort_tmp_bytes(n).
int scatter_u32(uint32_t *idx, uint32_t *val, uint32_t *idx2, uint32_t *val2, uint64_t n,
                uint64_t nout, uint32_t *out, void *tmp, hipStream_t st);

// Stable compaction: for every i in [0,n) with (src[i] & flag) != 0, in
// ascending i, append key = src[i] & ~flag and value = i.  The count goes to
// *dcount (device).  tmp: compact_tmp_bytes(n).
size_t compact_tmp_bytes(uint64_t n);
int compact_flagged(const uint32_t *src, uint64_t n, uint32_t flag, uint32_t *out_key,
                    uint32_t *out_val, uint32_t *dcount, void *tmp, hipStream_t st);

// Device workspace owned by a context (grow-only).
struct Work {
  Scratch pos4, key, cnt, start, par, runend, gsize, label, sortbuf, dense, tmp, vols, vbeg,
      small, input, out_next, out_label, out_halo, hhead, hoff, hash, fine, rtab, part;
};

struct EngineState {
  Work w;
  hipStream_t stream = nullptr;
  hipEvent_t ev[2 * FOF_NPHASE] = {};
  bool events = false;
};

}  // namespace fof

struct fof_ctx {
  int device = 0;
  fof::EngineState st;
};
