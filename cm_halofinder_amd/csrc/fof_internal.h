// fof_internal.h -- declarations shared by the HIP translation units of the
// MI355X Friends-of-Friends engine.  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>
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
// The same over min(*dn, n) values (the count left on the device).
int scan_u32_count(const uint32_t *in, uint32_t *out, uint64_t n, const uint32_t *dn, void *tmp,
                   hipStream_t st);

// Passes over a list whose length an earlier kernel left in device memory
// take the host-known capacity n and dn (the device count, n when dn is
// null): grids are sized for the capacity and blocks past the count do no
// work, so a call sequence needs no host round trip.

// Stable LSD radix sort of (key,val) u32 pairs on bits [0,bits).  Result is
// left in (k0,v0); (k1,v1) are ping-pong buffers of the same size.
size_t sort_tmp_bytes(uint64_t n);
int radix_sort_u32(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1,
                   uint64_t n, const uint32_t *dn, int bits, void *tmp, hipStream_t st);

// out[idx[i]] = val[i] | orv for i in [0,n), every idx[i] < nout: the pairs
// are first partitioned (one radix pass) by the top 8 bits of idx, so that
// the writes of a stretch of the list fall in one 1/256 of out and merge in
// the caches (random 4-byte writes cost one memory request each).  (idx2,
// val2) receive the partitioned pairs; tmp: sort_tmp_bytes(n).
int scatter_u32(const uint32_t *idx, const uint32_t *val, uint32_t *idx2, uint32_t *val2,
                uint64_t n, const uint32_t *dn, uint64_t nout, uint32_t *out, uint32_t orv,
                void *tmp, hipStream_t st);

// Stable compaction: for every i in [0,n) with (src[i] & flag) != 0, in
// ascending i, append key = src[i] & ~flag and value = i.  The count goes to
// *dcount (device; must not alias dn).  tmp: compact_tmp_bytes(n).
size_t compact_tmp_bytes(uint64_t n);
int compact_flagged(const uint32_t *src, uint64_t n, const uint32_t *dn, uint32_t flag,
                    uint32_t *out_key, uint32_t *out_val, uint32_t *dcount, void *tmp,
                    hipStream_t st);

// p[0..n) = v (a kernel, not hipMemsetAsync: captured sequences hold kernel
// nodes only).
int fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st);

// One pipeline: a stream, grow-only device workspace (one arena carved into
// the call's buffers) and the host copies of what the call uploads and reads
// back.  Kept by the context between calls.
struct Pipe {
  hipStream_t stream = nullptr;
  Scratch arena;  // every buffer of a call
  Scratch hash;   // occupied-cell hash table (sized once the cell count is known)
  std::vector<uint8_t> vblob;  // volume table, varlink tables, volume starts
  std::vector<uint8_t> hres;   // counts, flags and counters read back
  hipEvent_t ev[2 * FOF_NPHASE] = {};
  bool events = false;
};

constexpr int kMaxPipes = 8;  // concurrent pipelines of one call

// Autotuning: a workload shape and the parameters chosen for it.
struct TuneKey {
  uint32_t lg_n, lg_nvol, sparse, varlink, ppc16, b64;
  bool operator==(const TuneKey &o) const {
    return lg_n == o.lg_n && lg_nvol == o.lg_nvol && sparse == o.sparse && varlink == o.varlink &&
           ppc16 == o.ppc16 && b64 == o.b64;
  }
};
struct TuneVal {
  uint32_t T, pipes, gbig;  // dense threshold, pipelines, dense-kernel grid (0: default)
};

struct EngineState {
  Pipe main;                       // pipeline 0, on the context's stream
  std::vector<Pipe *> pipes;       // pipelines 1.. (entry 0 unused)
  hipEvent_t fork = nullptr, join[kMaxPipes] = {}, t0 = nullptr, t1 = nullptr;
  bool sync_events = false;
  hipGraphExec_t graph_exec = nullptr;  // the last captured call
  uint64_t graph_sig = 0;
  std::vector<std::pair<TuneKey, TuneVal>> tuned;
};

}  // namespace fof

struct fof_ctx {
  int device = 0;
  fof::EngineState st;
};

// Engine entry (fof_engine.hip); arguments checked by the caller.
extern "C" int fof_engine_run(fof_ctx *ctx, const fof_particles_v1 *P, const fof_volume_v1 *vols,
                              uint32_t nvol, const fof_opts_v1 *opts, fof_outputs_v1 *out);
extern "C" int fof_engine_plan(const fof_particles_v1 *P, const fof_volume_v1 *vols,
                               uint32_t nvol, const fof_opts_v1 *opts, fof_plan_v1 *out);
extern "C" int fof_engine_run_noalloc(fof_ctx *ctx, const fof_particles_v1 *P,
                                      const fof_volume_v1 *vol, const fof_opts_v1 *opts,
                                      fof_outputs_v1 *out);
extern "C" uint64_t fof_engine_noalloc_bytes(void);
extern "C" int fof_engine_init(fof_ctx *ctx);
extern "C" void fof_engine_free(fof_ctx *ctx);
