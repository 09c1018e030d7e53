// fof_api.hip -- C ABI of the FoF engine: contexts, argument checks, errors.
// Argument checks run on the host before any device work, so a bad call
// fails the same way with or without a GPU.  The domain check (every
// particle inside [rmin_buff, rmax_buff]) runs on the host for host-resident
// particles, before any device work, and inside the first kernel for
// device-resident ones (reported when the call ends).
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "fof_internal.h"

static thread_local char g_last_error[256];

int fof_set_hip_error(hipError_t e, const char *what, const char *file, int line) {
  snprintf(g_last_error, sizeof g_last_error, "%s failed: %s (%s:%d)", what,
           hipGetErrorString(e), file, line);
  return FOF_E_HIP;
}

namespace {

int check_device(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    (void)hipGetLastError();
    return FOF_E_NODEV;
  }
  if (device < 0 || device >= n) return FOF_E_NODEV;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return FOF_E_NODEV;
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return FOF_E_NODEV;
  return 0;
}

int check_particles(const fof_particles_v1 *P) {
  if (!P || P->version != FOF_ABI_VERSION) return FOF_E_ARG;
  if (P->location != FOF_MEM_HOST && P->location != FOF_MEM_DEVICE) return FOF_E_ARG;
  if (P->n >= 0x7FFFFFFFull) return FOF_E_RANGE;
  if (P->n && !P->base) return FOF_E_ARG;
  if (P->stride % 4 || P->off_pos % 4 || P->off_vel % 4) return FOF_E_ARG;
  if (P->off_pos + 12 > P->stride || P->off_vel + 12 > P->stride) return FOF_E_ARG;
  if (P->off_id >= 0 && (uint32_t)P->off_id + 8 > P->stride) return FOF_E_ARG;
  return 0;
}

int check_opts(const fof_opts_v1 *o) {
  if (!o) return 0;
  if (o->version != FOF_ABI_VERSION) return FOF_E_ARG;
  if (o->band != 0 && !(o->band >= 1e-6 && o->band <= 0.1)) return FOF_E_ARG;
  return 0;
}

// Domain check of host-resident particles, before any device work: every
// position inside its volume's [rmin_buff, rmax_buff] (the bounds of
// src_v2/read_data.c:810-812; NaN fails the comparisons).  Split over up to
// 16 host threads.
int check_domain_host(const fof_particles_v1 *P, const fof_volume_v1 *vols, uint32_t nvol) {
  if (P->location != FOF_MEM_HOST || P->n == 0) return 0;
  const uint8_t *base = static_cast<const uint8_t *>(P->base);
  const uint64_t N = P->n;
  const unsigned nth =
      (unsigned)std::max<uint64_t>(1, std::min<uint64_t>({16u, std::thread::hardware_concurrency(), N / 65536 + 1}));
  std::vector<int> bad(nth, 0);
  auto work = [&](unsigned t) {
    const uint64_t a = N * t / nth, e = N * (t + 1) / nth;
    uint32_t v = 0;
    while (v + 1 < nvol && vols[v].offset + vols[v].n <= a) v++;
    for (uint64_t p = a; p < e; p++) {
      while (p >= vols[v].offset + vols[v].n) v++;
      const fof_domain_v1 &D = vols[v].domain;
      const float *q = reinterpret_cast<const float *>(base + p * P->stride + P->off_pos);
      for (int d = 0; d < 3; d++) {
        const double x = q[d];
        if (!(x >= D.rmin_buff[d] && x <= D.rmax_buff[d])) {
          bad[t] = 1;
          return;
        }
      }
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 1; t < nth; t++) th.emplace_back(work, t);
  work(0);
  for (auto &x : th) x.join();
  for (int x : bad)
    if (x) return FOF_E_DOMAIN;
  return 0;
}

std::mutex g_default_mu;
fof_ctx *g_default[64];

}  // namespace

extern "C" {

const char *fof_strerror(int code) {
  switch (code) {
    case FOF_OK: return "success";
    case FOF_E_ARG: return "invalid argument";
    case FOF_E_DOMAIN: return "particle outside [rmin_buff, rmax_buff] or non-finite";
    case FOF_E_RANGE: return "size out of range (particles, cells or output capacity)";
    case FOF_E_NOMEM: return "device memory allocation failed";
    case FOF_E_HIP: return g_last_error[0] ? g_last_error : "HIP runtime error";
    case FOF_E_NODEV: return "no gfx950 device available";
    default: return "unknown error";
  }
}

int fof_ctx_create(int device, fof_ctx **out) {
  if (!out) return FOF_E_ARG;
  *out = nullptr;
  int rc = check_device(device);
  if (rc) return rc;
  FOF_HIP(hipSetDevice(device));
  fof_ctx *c = new (std::nothrow) fof_ctx;
  if (!c) return FOF_E_NOMEM;
  c->device = device;
  rc = fof_engine_init(c);
  if (rc) {
    delete c;
    return rc;
  }
  *out = c;
  return 0;
}

void fof_ctx_destroy(fof_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  fof_engine_free(ctx);
  delete ctx;
}

static int resolve_ctx(fof_ctx **ctx) {
  if (*ctx) {
    FOF_HIP(hipSetDevice((*ctx)->device));
    return 0;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return FOF_E_NODEV;
  }
  if (dev < 0 || dev >= 64) return FOF_E_NODEV;
  std::lock_guard<std::mutex> lk(g_default_mu);
  if (!g_default[dev]) {
    int rc = fof_ctx_create(dev, &g_default[dev]);
    if (rc) return rc;
  }
  *ctx = g_default[dev];
  return 0;
}

static int check_batch(const fof_particles_v1 *particles, const fof_volume_v1 *vols,
                       uint32_t nvol, const fof_opts_v1 *opts) {
  int rc = check_particles(particles);
  if (rc) return rc;
  if ((rc = check_opts(opts))) return rc;
  if (nvol == 0 || !vols) return FOF_E_ARG;
  uint64_t total = 0;
  for (uint32_t v = 0; v < nvol; v++) {
    if (vols[v].offset != total) return FOF_E_ARG;
    total += vols[v].n;
  }
  if (total != particles->n) return FOF_E_ARG;
  return 0;
}

int fof_link_gpu_ex(fof_ctx *ctx, const fof_particles_v1 *particles, const fof_volume_v1 *vols,
                    uint32_t nvol, const fof_opts_v1 *opts, fof_outputs_v1 *out) {
  if (!out || out->version != FOF_ABI_VERSION) return FOF_E_ARG;
  if (out->location != FOF_MEM_HOST && out->location != FOF_MEM_DEVICE) return FOF_E_ARG;
  out->nhalos = 0;
  int rc = check_batch(particles, vols, nvol, opts);
  if (rc) return rc;
  if (out->minid && particles->off_id < 0) return FOF_E_ARG;
  if ((rc = check_domain_host(particles, vols, nvol))) return rc;
  if ((rc = resolve_ctx(&ctx))) return rc;
  return fof_engine_run(ctx, particles, vols, nvol, opts, out);
}

static fof_outputs_v1 legacy_outputs(const fof_opts_v1 *opts, uint32_t *next_out,
                                     uint32_t *label_out, fof_halo_v1 *halos, uint64_t halo_cap,
                                     uint64_t *halo_offsets, fof_stats_v1 *stats) {
  fof_outputs_v1 o;
  memset(&o, 0, sizeof(o));
  o.version = FOF_ABI_VERSION;
  o.location = (opts && (opts->flags & FOF_OPT_OUT_DEVICE)) ? FOF_MEM_DEVICE : FOF_MEM_HOST;
  o.next = next_out;
  o.label = label_out;
  o.halos = halos;
  o.halo_cap = halo_cap;
  o.halo_offsets = halo_offsets;
  o.stats = stats;
  return o;
}

int fof_link_gpu_batch(fof_ctx *ctx, const fof_particles_v1 *particles,
                       const fof_volume_v1 *vols, uint32_t nvol, const fof_opts_v1 *opts,
                       uint32_t *next_out, uint32_t *label_out, fof_halo_v1 *halos,
                       uint64_t halo_cap, uint64_t *halo_offsets, fof_stats_v1 *stats) {
  fof_outputs_v1 o = legacy_outputs(opts, next_out, label_out, halos, halo_cap, halo_offsets, stats);
  return fof_link_gpu_ex(ctx, particles, vols, nvol, opts, &o);
}

int fof_link_gpu(fof_ctx *ctx, const fof_particles_v1 *particles, const fof_domain_v1 *domain,
                 const fof_opts_v1 *opts, uint32_t *next_out, uint32_t *label_out,
                 fof_halo_v1 *halos, uint64_t halo_cap, uint64_t *nhalos, fof_stats_v1 *stats) {
  if (nhalos) *nhalos = 0;
  int rc = check_particles(particles);
  if (rc) return rc;
  if (!domain) return FOF_E_ARG;
  fof_volume_v1 v;
  v.offset = 0;
  v.n = particles->n;
  v.domain = *domain;
  fof_outputs_v1 o = legacy_outputs(opts, next_out, label_out, halos, halo_cap, nullptr, stats);
  rc = fof_link_gpu_ex(ctx, particles, &v, 1, opts, &o);
  if (nhalos) *nhalos = o.nhalos;
  return rc;
}

int fof_link_gpu_noalloc(fof_ctx *ctx, const fof_particles_v1 *particles,
                         const fof_domain_v1 *domain, const fof_opts_v1 *opts,
                         fof_outputs_v1 *out) {
  if (!out || out->version != FOF_ABI_VERSION) return FOF_E_ARG;
  out->nhalos = 0;
  int rc = check_particles(particles);
  if (rc) return rc;
  if ((rc = check_opts(opts))) return rc;
  if (!domain) return FOF_E_ARG;
  if (particles->location != FOF_MEM_DEVICE || out->location != FOF_MEM_DEVICE) return FOF_E_ARG;
  if (out->minid && particles->off_id < 0) return FOF_E_ARG;
  if ((rc = resolve_ctx(&ctx))) return rc;
  fof_volume_v1 v;
  v.offset = 0;
  v.n = particles->n;
  v.domain = *domain;
  return fof_engine_run_noalloc(ctx, particles, &v, opts, out);
}

int fof_plan(const fof_particles_v1 *particles, const fof_volume_v1 *vols, uint32_t nvol,
             const fof_opts_v1 *opts, fof_plan_v1 *out) {
  if (!out) return FOF_E_ARG;
  int rc = check_batch(particles, vols, nvol, opts);
  if (rc) return rc;
  return fof_engine_plan(particles, vols, nvol, opts, out);
}

int fof_device_info(int device, char *buf, size_t len) {
  if (!buf || !len) return FOF_E_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    (void)hipGetLastError();
    snprintf(buf, len, "no device %d", device);
    return FOF_E_NODEV;
  }
  hipDeviceProp_t p;
  FOF_HIP(hipGetDeviceProperties(&p, device));
  snprintf(buf, len, "%s %s CUs=%d mem=%.1fGiB lds/block=%zu clock=%dkHz", p.name,
           p.gcnArchName, p.multiProcessorCount, p.totalGlobalMem / 1073741824.0,
           (size_t)p.sharedMemPerBlock, p.clockRate);
  return 0;
}

}  // extern "C"
