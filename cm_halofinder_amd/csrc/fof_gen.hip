// fof_gen.hip -- synthetic particle sets and periodic ghost shells.
//
// These build the input that the reference's Read_Data hands to FOF()
// (src_v2/read_data.c:5-856): 32-byte records {u64 ID, float Pos[3],
// float Vel[3]} (src_v2/vars.h:36-47 with PARTICLE_ID + MEMORY_MODE).
// Every quantity is a pure function of (seed, index), computed by the same
// __host__ __device__ code on both sides; only libm / device-libm rounding of
// log, sqrt, sin, cos and pow can differ in the last bit between the two.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "fof_internal.h"

namespace fof {
namespace gen {

struct Rec {
  uint64_t id;
  float pos[3];
  float vel[3];
};
static_assert(sizeof(Rec) == 32, "record layout");

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t rnd(uint64_t seed, uint64_t stream, uint64_t i) {
  return mix64(mix64(seed * 0x2545F4914F6CDD1Dull + stream) ^ (i * 0x9E3779B97F4A7C15ull));
}

// uniform in (0,1)
__host__ __device__ inline double u01(uint64_t r) {
  return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Keyed bijection of [0,n): 4-round Feistel network on 2h bits with
// cycle walking.
struct Perm {
  uint64_t n, key;
  int h;
  uint64_t mask;
};

inline Perm make_perm(uint64_t n, uint64_t seed) {
  Perm p;
  p.n = n;
  p.key = mix64(seed ^ 0x5DEECE66Dull);
  int bits = 2;
  while ((1ull << bits) < n) bits++;
  if (bits & 1) bits++;
  p.h = bits / 2;
  p.mask = (1ull << p.h) - 1;
  return p;
}

__host__ __device__ inline uint64_t perm_apply(const Perm &p, uint64_t x) {
  do {
    uint64_t l = x >> p.h, r = x & p.mask;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      uint64_t f = mix64(r ^ (p.key + (uint64_t)k * 0x632BE59BD9B4E019ull)) & p.mask;
      uint64_t t = l ^ f;
      l = r;
      r = t;
    }
    x = (l << p.h) | r;
  } while (x >= p.n);
  return x;
}

struct Tables {
  uint64_t nclumped;         // particles in clumps
  uint64_t nclumps;
  const uint64_t *off;       // [nclumps+1] prefix of clump sizes
  const double *cen;         // [3*nclumps]
  const double *rad;         // [nclumps] Plummer scale radius
  uint64_t npairs;           // borderline pairs
  uint64_t pair_base;        // first generation index of the pairs
};

__host__ __device__ inline double wrap(double x, double L) {
  if (x < 0) x += L;
  if (x >= L) x -= L;
  if (x < 0 || x >= L) x = x - floor(x / L) * L;
  return x;
}

// Reference pair arithmetic (src_v2/FOF.c:351-355) on two float points.
__host__ __device__ inline double ref_d2(const float *a, const float *b) {
  float fx = a[0] - b[0], fy = a[1] - b[1], fz = a[2] - b[2];
  double dx = fx, dy = fy, dz = fz;
  double s = dx * dx + dy * dy;
  return s + dz * dz;
}

// One borderline pair: A uniform in [0.5, L-0.5)^3, B = A + l*u nudged by
// whole float ulps until |d^2/l^2 - 1| < rel under the reference arithmetic.
__host__ __device__ inline void borderline_pair(const fof_gen_v1 &g, uint64_t q, float *A,
                                                float *B) {
  const double L = g.box, l = g.pair_l, l2 = l * l;
  for (int k = 0; k < 3; k++)
    A[k] = (float)(0.5 + (L - 1.0) * u01(rnd(g.seed, 20 + k, q)));
  for (int attempt = 0; attempt < 64; attempt++) {
    const double ct = 2.0 * u01(rnd(g.seed, 30 + attempt, q)) - 1.0;
    const double ph = 6.283185307179586 * u01(rnd(g.seed, 100 + attempt, q));
    const double st = sqrt(fmax(0.0, 1.0 - ct * ct));
    const double dir[3] = {st * cos(ph), st * sin(ph), ct};
    float b0[3];
    for (int k = 0; k < 3; k++) b0[k] = (float)((double)A[k] + l * dir[k]);
    for (int dz = -6; dz <= 6; dz++)
      for (int dy = -6; dy <= 6; dy++)
        for (int dx = -6; dx <= 6; dx++) {
          float c[3] = {b0[0], b0[1], b0[2]};
          const int dd[3] = {dx, dy, dz};
          for (int k = 0; k < 3; k++) {
            for (int s = 0; s < (dd[k] > 0 ? dd[k] : -dd[k]); s++)
              c[k] = nextafterf(c[k], dd[k] > 0 ? 3.0e38f : -3.0e38f);
          }
          const double r = ref_d2(A, c) / l2 - 1.0;
          if (fabs(r) < g.pair_rel) {
            B[0] = c[0], B[1] = c[1], B[2] = c[2];
            return;
          }
        }
  }
  // no lattice point found (not expected): place B on top of A
  B[0] = A[0], B[1] = A[1], B[2] = A[2];
}

// Generation index -> record.
__host__ __device__ inline void make_particle(const fof_gen_v1 &g, const Tables &t, uint64_t gi,
                                              Rec &r) {
  const double L = g.box;
  double p[3];
  if (gi < t.nclumped) {
    // clump k with off[k] <= gi < off[k+1]
    uint64_t lo = 0, hi = t.nclumps - 1;
    while (lo < hi) {
      uint64_t mid = (lo + hi + 1) >> 1;
      if (t.off[mid] <= gi) lo = mid;
      else hi = mid - 1;
    }
    const double a = t.rad[lo];
    const double u = g.umax * u01(rnd(g.seed, 1, gi));
    const double rr = a / sqrt(pow(u, -2.0 / 3.0) - 1.0);
    const double ct = 2.0 * u01(rnd(g.seed, 2, gi)) - 1.0;
    const double ph = 6.283185307179586 * u01(rnd(g.seed, 3, gi));
    const double st = sqrt(fmax(0.0, 1.0 - ct * ct));
    p[0] = t.cen[3 * lo] + rr * st * cos(ph);
    p[1] = t.cen[3 * lo + 1] + rr * st * sin(ph);
    p[2] = t.cen[3 * lo + 2] + rr * ct;
    for (int k = 0; k < 3; k++) r.pos[k] = (float)wrap(p[k], L);
  } else if (t.npairs && gi >= t.pair_base && gi < t.pair_base + 2 * t.npairs) {
    float A[3], B[3];
    borderline_pair(g, (gi - t.pair_base) >> 1, A, B);
    const float *s = ((gi - t.pair_base) & 1) ? B : A;
    for (int k = 0; k < 3; k++) r.pos[k] = s[k];
  } else {
    for (int k = 0; k < 3; k++) r.pos[k] = (float)(L * u01(rnd(g.seed, 4 + k, gi)));
  }
  for (int k = 0; k < 3; k++)
    if (!(r.pos[k] < (float)L)) r.pos[k] = 0.0f;  // float rounding reached L
  // velocities: Box-Muller
  const double u1 = u01(rnd(g.seed, 8, gi)), u2 = u01(rnd(g.seed, 9, gi));
  const double u3 = u01(rnd(g.seed, 10, gi)), u4 = u01(rnd(g.seed, 11, gi));
  const double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
  r.vel[0] = (float)(g.sigma_v * r1 * cos(6.283185307179586 * u2));
  r.vel[1] = (float)(g.sigma_v * r1 * sin(6.283185307179586 * u2));
  r.vel[2] = (float)(g.sigma_v * r2 * cos(6.283185307179586 * u4));
  r.id = gi;
}

// Clump table: quantile-sampled sizes from dN/dM ~ M^-alpha on [mmin,mmax],
// rescaled so that they sum to the clumped count exactly; centres uniform;
// scale radius a0 * (size/100)^(1/3).
struct HostTables {
  std::vector<uint64_t> off;
  std::vector<double> cen, rad;
  Tables t;
};

inline int build_tables(const fof_gen_v1 &g, HostTables &H) {
  const uint64_t n = g.n;
  uint64_t npairs = 0;
  if (g.pair_frac > 0) npairs = (uint64_t)(g.pair_frac * (double)n / 2.0);
  uint64_t nclumped = (g.nclumps > 0) ? (uint64_t)(g.clump_frac * (double)n) : 0;
  if (nclumped + 2 * npairs > n) return FOF_E_ARG;
  const uint64_t K = nclumped ? g.nclumps : 0;
  if (K > nclumped) return FOF_E_ARG;
  H.off.assign(K + 1, 0);
  H.cen.assign(3 * K + 3, 0.0);
  H.rad.assign(K + 1, 0.0);
  if (K) {
    const double a1 = 1.0 - g.alpha;
    const double lo = pow(g.mmin, a1), hi = pow(g.mmax, a1);
    std::vector<double> m(K);
    double tot = 0;
    for (uint64_t k = 0; k < K; k++) {
      const double u = ((double)k + 0.5) / (double)K;
      m[k] = pow(lo + u * (hi - lo), 1.0 / a1);
      tot += m[k];
    }
    std::vector<uint64_t> sz(K);
    uint64_t used = 0;
    for (uint64_t k = 0; k < K; k++) {
      sz[k] = std::max<uint64_t>(1, (uint64_t)floor(m[k] * (double)nclumped / tot));
      used += sz[k];
    }
    // settle the rounding remainder one particle at a time: add from the
    // smallest clumps up, remove from the largest down (sizes stay >= 1)
    for (uint64_t k = 0; used < nclumped; k = (k + 1) % K) sz[k]++, used++;
    for (uint64_t k = K - 1; used > nclumped; k = (k == 0) ? K - 1 : k - 1)
      if (sz[k] > 1) sz[k]--, used--;
    for (uint64_t k = 0; k < K; k++) {
      H.off[k + 1] = H.off[k] + sz[k];
      for (int d = 0; d < 3; d++) H.cen[3 * k + d] = g.box * u01(rnd(g.seed, 12 + d, k));
      H.rad[k] = g.a0 * cbrt((double)sz[k] / 100.0);
    }
  }
  H.t.nclumped = nclumped;
  H.t.nclumps = K;
  H.t.off = H.off.data();
  H.t.cen = H.cen.data();
  H.t.rad = H.rad.data();
  H.t.npairs = npairs;
  H.t.pair_base = n - 2 * npairs;
  return 0;
}

inline int check_gen(const fof_gen_v1 *g) {
  if (!g || g->n == 0 || !(g->box > 1.0) || g->n >= (1ull << 40)) return FOF_E_ARG;
  if (g->nclumps && (!(g->mmin > 0) || !(g->mmax >= g->mmin) || g->alpha == 1.0))
    return FOF_E_ARG;
  if (!(g->umax > 0 && g->umax < 1)) return FOF_E_ARG;
  if (g->pair_frac > 0 && (!(g->pair_l > 0) || !(g->pair_rel > 0))) return FOF_E_ARG;
  return 0;
}

__global__ __launch_bounds__(256) void k_gen(fof_gen_v1 g, Tables t, Perm P, Rec *out) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < g.n;
       j += (uint64_t)gridDim.x * blockDim.x) {
    Rec r;
    make_particle(g, t, perm_apply(P, j), r);
    out[j] = r;
  }
}

// ---------------------------------------------------------------------------
// Ghost shells.  Subvolume (a,b,c) of a periodic box cut gx*gy*gz ways owns
// [rmin, rmax) = [i*box/g, (i+1)*box/g) per axis (src_v2/startup.c:36-47) and
// receives every image x + s*box (s in {-1,0,1}) that lands in
// [rmin - shell, rmax + shell] -- the owner routing of
// src_v2/read_data.c:425-437 plus the shifted copies of :418-522, :785-799,
// also when an axis is cut only once.
// ---------------------------------------------------------------------------
struct Cut {
  double box, shell;
  int g[3];
};

// Per axis: up to 3 (volume index, image coordinate) candidates.
__host__ __device__ inline int axis_cands(const Cut &c, int d, float x, int *vi, float *xi) {
  const int G = c.g[d];
  const double Ls = c.box / (double)G;
  int own = (int)floor((double)G * ((double)x / c.box));
  if (own >= G) own = G - 1;
  if (own < 0) own = 0;
  int n = 0;
  for (int t = -1; t <= 1; t++) {
    int v = own + t;
    double shift = 0;
    if (v < 0) v += G, shift = c.box;
    if (v >= G) v -= G, shift = -c.box;
    const float img = (float)((double)x + shift);
    const double lo = (double)v * Ls - c.shell, hi = (double)v * Ls + Ls + c.shell;
    if ((double)img >= lo && (double)img <= hi) {
      // an identical (volume, image) pair can arise only when G == 1 and
      // shift == 0 twice, which t in {-1,0,1} never produces
      vi[n] = v;
      xi[n] = img;
      n++;
    }
  }
  return n;
}

__host__ __device__ inline int count_images(const Cut &c, const float *p) {
  int vi[3];
  float xi[3];
  int n = 1;
  for (int d = 0; d < 3; d++) n *= axis_cands(c, d, p[d], vi, xi);
  return n;
}

// k-th image (in z-major, x-fastest candidate order) of particle p.
__host__ __device__ inline int image_k(const Cut &c, const float *p, int k, float *q) {
  int vi[3][3], nc[3];
  float xi[3][3];
  for (int d = 0; d < 3; d++) nc[d] = axis_cands(c, d, p[d], vi[d], xi[d]);
  const int a = k % nc[0], b = (k / nc[0]) % nc[1], e = k / (nc[0] * nc[1]);
  q[0] = xi[0][a], q[1] = xi[1][b], q[2] = xi[2][e];
  return (vi[2][e] * c.g[1] + vi[1][b]) * c.g[0] + vi[0][a];
}

inline void fill_vols(const fof_ghost_v1 &g, const uint64_t *counts, fof_volume_v1 *vols) {
  const int G[3] = {(int)g.gx, (int)g.gy, (int)g.gz};
  uint64_t off = 0;
  for (int c = 0; c < G[2]; c++)
    for (int b = 0; b < G[1]; b++)
      for (int a = 0; a < G[0]; a++) {
        const int v = (c * G[1] + b) * G[0] + a;
        fof_volume_v1 &V = vols[v];
        memset(&V, 0, sizeof(V));
        V.offset = off;
        V.n = counts[v];
        off += counts[v];
        const int idx[3] = {a, b, c};
        for (int d = 0; d < 3; d++) {
          const double Ls = g.box / (double)G[d];
          V.domain.rmin[d] = idx[d] * Ls;  // src_v2/startup.c:36-39
          V.domain.rmax[d] = V.domain.rmin[d] + Ls;
          V.domain.rmin_buff[d] = V.domain.rmin[d] - g.shell;
          V.domain.rmax_buff[d] = V.domain.rmax[d] + g.shell;
          V.domain.lcell[d] = g.box / (double)g.pside;  // src_v2/FOF.c:46-48
        }
        V.domain.linklength = g.linklength;
        V.domain.nphalomin = g.nphalomin;
        V.domain.varlink = nullptr;
      }
}

inline int check_ghost(const fof_ghost_v1 *g) {
  if (!g || !(g->box > 0) || !(g->shell >= 0) || g->gx == 0 || g->gy == 0 || g->gz == 0 ||
      g->pside == 0)
    return FOF_E_ARG;
  if ((uint64_t)g->gx * g->gy * g->gz > 65536) return FOF_E_ARG;
  for (uint32_t G : {g->gx, g->gy, g->gz})
    if (!(g->shell < g->box / (double)G / 2.0)) return FOF_E_ARG;
  return 0;
}

__global__ __launch_bounds__(256) void k_img_count(Cut c, const Rec *in, uint64_t n,
                                                   uint32_t *cnt) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (uint64_t)gridDim.x * blockDim.x)
    cnt[j] = (uint32_t)count_images(c, in[j].pos);
}

// Expanded item list: (volume, particle, image) in input order.
__global__ __launch_bounds__(256) void k_img_emit(Cut c, const Rec *in, uint64_t n,
                                                  const uint32_t *off, uint32_t *vkey,
                                                  uint32_t *item, uint32_t *ipart,
                                                  uint8_t *icode) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t o = off[j], m = off[j + 1] - o;
    for (uint32_t k = 0; k < m; k++) {
      float q[3];
      vkey[o + k] = (uint32_t)image_k(c, in[j].pos, (int)k, q);
      item[o + k] = o + k;
      ipart[o + k] = (uint32_t)j;
      icode[o + k] = (uint8_t)k;
    }
  }
}

__global__ __launch_bounds__(256) void k_img_gather(Cut c, const Rec *in, uint64_t m,
                                                    const uint32_t *item, const uint32_t *ipart,
                                                    const uint8_t *icode, Rec *out) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t it = item[t];
    const Rec r = in[ipart[it]];
    Rec o = r;
    image_k(c, r.pos, icode[it], o.pos);
    out[t] = o;
  }
}

__global__ void k_vol_count(const uint32_t *vkey, uint64_t m, int nvol, uint64_t *counts) {
  // vkey sorted: counts[v] = upper_bound(v) - lower_bound(v)
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nvol; v += gridDim.x * blockDim.x) {
    uint64_t lo = 0, hi = m;
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if (vkey[mid] < (uint32_t)v) lo = mid + 1;
      else hi = mid;
    }
    uint64_t a = lo;
    lo = a, hi = m;
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if (vkey[mid] <= (uint32_t)v) lo = mid + 1;
      else hi = mid;
    }
    counts[v] = lo - a;
  }
}

}  // namespace gen
}  // namespace fof

using namespace fof;
using namespace fof::gen;

extern "C" int fof_gen_host(const fof_gen_v1 *g, void *out) {
  int rc = check_gen(g);
  if (rc) return rc;
  if (!out) return FOF_E_ARG;
  HostTables H;
  if ((rc = build_tables(*g, H))) return rc;
  const Perm P = make_perm(g->n, g->seed);
  Rec *o = static_cast<Rec *>(out);
  for (uint64_t j = 0; j < g->n; j++) make_particle(*g, H.t, perm_apply(P, j), o[j]);
  return 0;
}

extern "C" int fof_gen_device(const fof_gen_v1 *g, void *out) {
  int rc = check_gen(g);
  if (rc) return rc;
  if (!out) return FOF_E_ARG;
  HostTables H;
  if ((rc = build_tables(*g, H))) return rc;
  const Perm P = make_perm(g->n, g->seed);
  const uint64_t K = H.t.nclumps;
  void *d = nullptr;
  const size_t bytes = (K + 1) * 8 + (3 * K + 3) * 8 + (K + 1) * 8;
  FOF_HIP(hipMalloc(&d, bytes));
  uint8_t *b = static_cast<uint8_t *>(d);
  Tables t = H.t;
  t.off = reinterpret_cast<const uint64_t *>(b);
  t.cen = reinterpret_cast<const double *>(b + (K + 1) * 8);
  t.rad = reinterpret_cast<const double *>(b + (K + 1) * 8 + (3 * K + 3) * 8);
  hipError_t e = hipMemcpy(b, H.off.data(), (K + 1) * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b + (K + 1) * 8, H.cen.data(), (3 * K + 3) * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b + (K + 1) * 8 + (3 * K + 3) * 8, H.rad.data(), (K + 1) * 8,
                  hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    uint64_t blocks = std::min<uint64_t>((g->n + 255) / 256, 1u << 20);
    k_gen<<<(uint32_t)blocks, 256>>>(*g, t, P, static_cast<Rec *>(out));
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  (void)hipFree(d);
  if (e != hipSuccess) return fof_set_hip_error(e, "fof_gen_device", __FILE__, __LINE__);
  return 0;
}

extern "C" int fof_ghosts_host(const fof_ghost_v1 *g, const void *in, uint64_t n, void *out,
                               uint64_t out_cap, uint64_t *counts, fof_volume_v1 *vols) {
  int rc = check_ghost(g);
  if (rc) return rc;
  if ((!in && n) || !counts) return FOF_E_ARG;
  const Cut c = {g->box, g->shell, {(int)g->gx, (int)g->gy, (int)g->gz}};
  const int nvol = c.g[0] * c.g[1] * c.g[2];
  const Rec *r = static_cast<const Rec *>(in);
  for (int v = 0; v < nvol; v++) counts[v] = 0;
  for (uint64_t j = 0; j < n; j++) {
    const int m = count_images(c, r[j].pos);
    for (int k = 0; k < m; k++) {
      float q[3];
      counts[image_k(c, r[j].pos, k, q)]++;
    }
  }
  if (vols) fill_vols(*g, counts, vols);
  if (!out) return 0;
  uint64_t total = 0;
  std::vector<uint64_t> cur(nvol);
  for (int v = 0; v < nvol; v++) cur[v] = total, total += counts[v];
  if (total > out_cap) return FOF_E_RANGE;
  Rec *o = static_cast<Rec *>(out);
  for (uint64_t j = 0; j < n; j++) {
    const int m = count_images(c, r[j].pos);
    for (int k = 0; k < m; k++) {
      Rec x = r[j];
      const int v = image_k(c, r[j].pos, k, x.pos);
      o[cur[v]++] = x;
    }
  }
  return 0;
}

extern "C" int fof_ghosts_device(const fof_ghost_v1 *g, const void *in, uint64_t n, void *out,
                                 uint64_t out_cap, uint64_t *counts, fof_volume_v1 *vols) {
  int rc = check_ghost(g);
  if (rc) return rc;
  if ((!in && n) || !counts) return FOF_E_ARG;
  const Cut c = {g->box, g->shell, {(int)g->gx, (int)g->gy, (int)g->gz}};
  const int nvol = c.g[0] * c.g[1] * c.g[2];
  if (n == 0) {
    for (int v = 0; v < nvol; v++) counts[v] = 0;
    if (vols) fill_vols(*g, counts, vols);
    return 0;
  }
  if (n >= 0x7FFFFFFFull) return FOF_E_RANGE;
  const Rec *din = static_cast<const Rec *>(in);
  // per-particle image counts and their prefix
  Scratch cnt, tmp;
  if ((rc = ensure(cnt, (n + 2) * 4))) return rc;
  uint32_t *dc = static_cast<uint32_t *>(cnt.ptr);
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 1u << 20);
  k_img_count<<<blocks, 256>>>(c, din, n, dc);
  FOF_HIP(hipGetLastError());
  if ((rc = ensure(tmp, scan_tmp_bytes(n)))) return rc;
  if ((rc = scan_u32(dc, dc, n, tmp.ptr, 0))) return rc;
  uint32_t m32 = 0;
  FOF_HIP(hipMemcpy(&m32, dc + n, 4, hipMemcpyDeviceToHost));
  const uint64_t m = m32;
  // items: key (volume), value (item index) + ping-pong, particle, code
  Scratch items;
  const uint64_t q = (m + 63) / 64 * 64;
  if ((rc = ensure(items, q * 4 * 5 + q + 256))) {
    release(cnt), release(tmp);
    return rc;
  }
  uint32_t *vk = static_cast<uint32_t *>(items.ptr), *it = vk + q, *vk1 = vk + 2 * q,
           *it1 = vk + 3 * q, *ip = vk + 4 * q;
  uint8_t *ic = reinterpret_cast<uint8_t *>(vk + 5 * q);
  k_img_emit<<<blocks, 256>>>(c, din, n, dc, vk, it, ip, ic);
  FOF_HIP(hipGetLastError());
  int bits = 1;
  while ((1 << bits) < nvol) bits++;
  if (nvol > 1) {
    if ((rc = ensure(tmp, sort_tmp_bytes(m)))) return rc;
    if ((rc = radix_sort_u32(vk, it, vk1, it1, m, nullptr, bits, tmp.ptr, 0))) return rc;
  }
  Scratch dcount;
  if ((rc = ensure(dcount, (nvol + 1) * 8))) return rc;
  k_vol_count<<<(nvol + 255) / 256, 256>>>(vk, m, nvol, static_cast<uint64_t *>(dcount.ptr));
  FOF_HIP(hipGetLastError());
  FOF_HIP(hipMemcpy(counts, dcount.ptr, nvol * 8, hipMemcpyDeviceToHost));
  if (vols) fill_vols(*g, counts, vols);
  if (out) {
    if (m > out_cap) rc = FOF_E_RANGE;
    else {
      const uint32_t gb = (uint32_t)std::min<uint64_t>((m + 255) / 256, 1u << 20);
      k_img_gather<<<gb, 256>>>(c, din, m, it, ip, ic, static_cast<Rec *>(out));
      FOF_HIP(hipGetLastError());
      FOF_HIP(hipDeviceSynchronize());
    }
  }
  release(cnt), release(tmp), release(items), release(dcount);
  return rc;
}
