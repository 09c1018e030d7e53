// fof_io.cpp -- host readers of the reference's particle formats and writers
// of its halo catalogue (declared in include/fof_io.h).  Plain C++ on the
// host; compiled with -ffp-contract=off so that the catalogue values round as
// the reference's x86-64 build rounds them.
#include "../../include/fof_io.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

namespace {

// The reference's 32-byte particle record (src_v2/vars.h:36-47).
struct Rec {
  unsigned long long id;
  float pos[3];
  float vel[3];
};
static_assert(sizeof(Rec) == 32, "record layout");

struct File {
  FILE *f = nullptr;
  explicit File(const char *path, const char *mode) : f(path ? fopen(path, mode) : nullptr) {}
  ~File() {
    if (f) fclose(f);
  }
};

bool read_exact(FILE *f, void *p, size_t n) { return fread(p, 1, n, f) == n; }

// A Fortran record marker that must equal `expect`.
bool marker(FILE *f, uint32_t expect) {
  uint32_t m;
  return read_exact(f, &m, 4) && m == expect;
}

bool blank(const char *s) {
  for (; *s; s++)
    if (*s != ' ' && *s != '\t' && *s != '\r' && *s != '\n') return false;
  return true;
}

}  // namespace

extern "C" {

// src_v2/read_data.c:286-312: fscanf "%llu %f %f %f %f %f %f" per line (MEMORY_MODE +
// PARTICLE_ID); strtoull / strtof are the conversions scanf's %llu and %f use.
int fof_read_ascii(const char *path, int with_id, void *out, uint64_t cap, uint64_t *n) {
  if (!path || !n) return FOF_E_ARG;
  *n = 0;
  File F(path, "r");
  if (!F.f) return FOF_E_ARG;
  Rec *o = static_cast<Rec *>(out);
  std::vector<char> line(4096);
  uint64_t k = 0;
  while (fgets(line.data(), (int)line.size(), F.f)) {
    // lines longer than the buffer: grow and continue reading the line
    while (strlen(line.data()) == line.size() - 1 && line[line.size() - 2] != '\n') {
      const size_t old = line.size();
      line.resize(old * 2);
      if (!fgets(line.data() + old - 1, (int)(line.size() - old + 1), F.f)) break;
    }
    if (blank(line.data())) continue;
    if (out) {
      if (k >= cap) return FOF_E_RANGE;
      char *p = line.data(), *e;
      Rec r;
      if (with_id) {
        r.id = strtoull(p, &e, 10);
        if (e == p) return FOF_E_ARG;
        p = e;
      } else {
        r.id = k;
      }
      float v[6];
      for (int q = 0; q < 6; q++) {
        v[q] = strtof(p, &e);
        if (e == p) return FOF_E_ARG;
        p = e;
      }
      memcpy(r.pos, v, 12);
      memcpy(r.vel, v + 3, 12);
      o[k] = r;
    }
    k++;
  }
  *n = k;
  return 0;
}

// src_v2/read_data.c:321-365: [256][header][256] [12n][Pos][12n] [12n][Vel][12n]
// [8n][ID][8n]; the particle count is header.npart[1].
int fof_read_gadget1(const char *path, int with_id, void *out, uint64_t cap, uint64_t *n,
                     fof_gadget_header_v1 *hdr) {
  static_assert(sizeof(fof_gadget_header_v1) == 256, "Gadget-1 header is 256 bytes");
  if (!path || !n) return FOF_E_ARG;
  *n = 0;
  File F(path, "rb");
  if (!F.f) return FOF_E_ARG;
  fof_gadget_header_v1 h;
  if (!marker(F.f, 256) || !read_exact(F.f, &h, 256) || !marker(F.f, 256)) return FOF_E_ARG;
  if (hdr) *hdr = h;
  const uint64_t np = h.npart[1];
  *n = np;
  if (!out) return 0;
  if (np > cap) return FOF_E_RANGE;
  Rec *o = static_cast<Rec *>(out);
  std::vector<float> buf(3 * np);
  const uint32_t b3 = (uint32_t)(12 * np);
  if (!marker(F.f, b3) || !read_exact(F.f, buf.data(), 12 * np) || !marker(F.f, b3)) return FOF_E_ARG;
  for (uint64_t j = 0; j < np; j++) memcpy(o[j].pos, &buf[3 * j], 12);
  if (!marker(F.f, b3) || !read_exact(F.f, buf.data(), 12 * np) || !marker(F.f, b3)) return FOF_E_ARG;
  for (uint64_t j = 0; j < np; j++) memcpy(o[j].vel, &buf[3 * j], 12);
  if (with_id) {
    std::vector<unsigned long long> ids(np);
    const uint32_t b8 = (uint32_t)(8 * np);
    if (!marker(F.f, b8) || !read_exact(F.f, ids.data(), 8 * np) || !marker(F.f, b8)) return FOF_E_ARG;
    for (uint64_t j = 0; j < np; j++) o[j].id = ids[j];
  } else {
    for (uint64_t j = 0; j < np; j++) o[j].id = j;
  }
  return 0;
}

// src_v2/read_data.c:248-277: records [4][npart][4] then [len][npart particles][len]
// until the end of the file; a particle is {u32 ID (PARTICLE_ID), Pos, Vel}.
int fof_read_picola(const char *path, int with_id, void *out, uint64_t cap, uint64_t *n) {
  if (!path || !n) return FOF_E_ARG;
  *n = 0;
  File F(path, "rb");
  if (!F.f) return FOF_E_ARG;
  Rec *o = static_cast<Rec *>(out);
  const uint32_t psz = with_id ? 28u : 24u;
  uint64_t k = 0;
  std::vector<uint8_t> buf;
  for (;;) {
    uint32_t m;
    if (fread(&m, 1, 4, F.f) != 4) break;  // end of file
    uint32_t np;
    if (m != 4 || !read_exact(F.f, &np, 4) || !marker(F.f, 4)) return FOF_E_ARG;
    const uint64_t len = (uint64_t)np * psz;
    if (!marker(F.f, (uint32_t)len)) return FOF_E_ARG;
    if (out && k + np > cap) return FOF_E_RANGE;
    buf.resize(len);
    if (!read_exact(F.f, buf.data(), len) || !marker(F.f, (uint32_t)len)) return FOF_E_ARG;
    if (out)
      for (uint32_t j = 0; j < np; j++) {
        const uint8_t *p = buf.data() + (uint64_t)j * psz;
        Rec r;
        if (with_id) {
          uint32_t id;
          memcpy(&id, p, 4);
          r.id = id;
          p += 4;
        } else {
          r.id = k + j;
        }
        memcpy(r.pos, p, 12);
        memcpy(r.vel, p + 12, 12);
        o[k + j] = r;
      }
    k += np;
  }
  *n = k;
  return 0;
}

}  // extern "C"

namespace {

// The values Output_Halos writes for one halo after N (src_v2/FOF.c:848-860):
// CoM and mean velocity, then the inertia and dispersion tensors by the
// reference's single-pass formulas from the raw FP64 moments.
int halo_values(const fof_halo_v1 &h, const fof_moments_v1 *m, uint32_t flags, float *v) {
  const double inv = 1.0 / (double)h.n;
  const double xh = h.sum_pos[0], yh = h.sum_pos[1], zh = h.sum_pos[2];
  const double vxh = h.sum_vel[0], vyh = h.sum_vel[1], vzh = h.sum_vel[2];
  int k = 0;
  v[k++] = (float)(xh * inv);
  v[k++] = (float)(yh * inv);
  v[k++] = (float)(zh * inv);
  v[k++] = (float)(vxh * inv);
  v[k++] = (float)(vyh * inv);
  v[k++] = (float)(vzh * inv);
  if (flags & FOF_OUT_INERTIA) {
    const double xh2 = m->pos2[0], yh2 = m->pos2[1], zh2 = m->pos2[2];
    const double xhyh = m->pos2[3], xhzh = m->pos2[4], yhzh = m->pos2[5];
    v[k++] = (float)(yh2 + zh2 - inv * (yh * yh + zh * zh));
    v[k++] = (float)(xh2 + zh2 - inv * (xh * xh + zh * zh));
    v[k++] = (float)(xh2 + yh2 - inv * (xh * xh + yh * yh));
    v[k++] = (float)(inv * xh * yh - xhyh);
    v[k++] = (float)(inv * xh * zh - xhzh);
    v[k++] = (float)(inv * yh * zh - yhzh);
  }
  if (flags & FOF_OUT_DISPERSION) {
    const double vxh2 = m->vel2[0], vyh2 = m->vel2[1], vzh2 = m->vel2[2];
    const double vxhvyh = m->vel2[3], vxhvzh = m->vel2[4], vyhvzh = m->vel2[5];
    v[k++] = (float)((vxh2 - vxh * vxh * inv) * inv);
    v[k++] = (float)((vyh2 - vyh * vyh * inv) * inv);
    v[k++] = (float)((vzh2 - vzh * vzh * inv) * inv);
    v[k++] = (float)((vxhvyh - vxh * vyh * inv) * inv);
    v[k++] = (float)((vxhvzh - vxh * vzh * inv) * inv);
    v[k++] = (float)((vyhvzh - vyh * vzh * inv) * inv);
  }
  return k;
}

void put_rec(FILE *f, uint32_t len) { fwrite(&len, 4, 1, f); }

}  // namespace

extern "C" {

int fof_write_halos(const char *path, uint32_t flags, const fof_halo_v1 *halos,
                    const fof_moments_v1 *moments, uint64_t nhalos,
                    const fof_particles_v1 *P, const uint32_t *next) {
  if (!path || (nhalos && !halos)) return FOF_E_ARG;
  if ((flags & (FOF_OUT_INERTIA | FOF_OUT_DISPERSION)) && nhalos && !moments) return FOF_E_ARG;
  const bool parts = (flags & FOF_OUT_PARTICLES) != 0, unf = (flags & FOF_OUT_UNFORMATTED) != 0;
  const bool withid = (flags & FOF_OUT_PARTICLE_ID) != 0;
  if (parts && (!P || !next || P->location != FOF_MEM_HOST || !P->base)) return FOF_E_ARG;
  if (parts && withid && P->off_id < 0) return FOF_E_ARG;
  // owned halos only (Checkhalo, src_v2/FOF.c:394-401)
  std::vector<uint64_t> own;
  for (uint64_t k = 0; k < nhalos; k++)
    if (halos[k].owned) own.push_back(k);
  const int32_t nh = (int32_t)own.size();
  File F(path, unf ? "wb" : "w");
  if (!F.f) return FOF_E_ARG;
  FILE *f = F.f;
  const int nval = 6 + ((flags & FOF_OUT_INERTIA) ? 6 : 0) + ((flags & FOF_OUT_DISPERSION) ? 6 : 0);
  const int halosize = 1 + nval;
  const uint8_t *base = parts ? static_cast<const uint8_t *>(P->base) : nullptr;
  auto member = [&](uint32_t j0, float *o, unsigned long long *id) {
    const uint8_t *r = base + (uint64_t)j0 * P->stride;
    memcpy(o, r + P->off_pos, 12);
    memcpy(o + 3, r + P->off_vel, 12);
    if (withid) memcpy(id, r + P->off_id, 8);
  };
  auto write_members = [&](const fof_halo_v1 &h) {
    const uint32_t head1 = h.head + 1;  // 1-based, as ihalo[]
    uint32_t j = head1;
    do {
      float o[6];
      unsigned long long id = 0;
      member(j - 1, o, &id);
      if (unf) {
        const uint32_t len = withid ? 32u : 24u;
        put_rec(f, len);
        if (withid) fwrite(&id, 8, 1, f);
        fwrite(o, 4, 6, f);
        put_rec(f, len);
      } else if (withid) {
        fprintf(f, "%15llu %15.6f %15.6f %15.6f %15.6f %15.6f %15.6f\n", id, o[0], o[1], o[2], o[3],
                o[4], o[5]);
      } else {
        fprintf(f, "%15.6f %15.6f %15.6f %15.6f %15.6f %15.6f\n", o[0], o[1], o[2], o[3], o[4], o[5]);
      }
      j = next[j - 1];
    } while (j != head1);
  };
  float v[18];
  if (unf) {
    put_rec(f, 4);
    fwrite(&nh, 4, 1, f);
    put_rec(f, 4);
    if (!parts) put_rec(f, (uint32_t)((uint64_t)nh * halosize * 4));
  } else if (parts) {
    fprintf(f, "%15d\n", nh);
  }
  for (uint64_t k : own) {
    const fof_halo_v1 &h = halos[k];
    halo_values(h, moments ? moments + k : nullptr, flags, v);
    if (unf) {
      const float nf = (float)h.n;
      if (parts) put_rec(f, (uint32_t)(halosize * 4));
      fwrite(&nf, 4, 1, f);
      fwrite(v, 4, nval, f);
      if (parts) put_rec(f, (uint32_t)(halosize * 4));
    } else {
      fprintf(f, "%15d", (int)h.n);
      for (int q = 0; q < nval; q++) fprintf(f, " %15.6f", v[q]);
      fprintf(f, "\n");
    }
    if (parts) write_members(h);
  }
  if (unf && !parts) put_rec(f, (uint32_t)((uint64_t)nh * halosize * 4));
  return ferror(f) ? FOF_E_ARG : 0;
}

// src_v2/FOF.c:172-284.  The visited flags play the part of the reference's
// head[] (every member but the first-found one is marked by the counting
// walk, :192-200), so groups are taken in ascending order of their lowest
// index.
int fof_subsample(const fof_particles_v1 *P, const uint32_t *next, const fof_subsample_v1 *s,
                  fof_uniform_fn uniform, void *state, uint32_t *inhalo, uint64_t in_cap,
                  uint64_t *n_in, uint32_t *outhalo, uint64_t out_cap, uint64_t *n_out) {
  if (!P || !s || !n_in || !n_out) return FOF_E_ARG;
  *n_in = 0, *n_out = 0;
  if (P->version != FOF_ABI_VERSION || P->location != FOF_MEM_HOST) return FOF_E_ARG;
  if (P->n > 0xFFFFFFFEull) return FOF_E_RANGE;
  const uint32_t n = (uint32_t)P->n;
  if (n && (!P->base || !next)) return FOF_E_ARG;
  const bool sin = s->samp_in > 0, sout = s->samp_out > 0;
  if ((sin || sout) && !uniform) return FOF_E_ARG;
  if ((sin && !inhalo && in_cap) || (sout && !outhalo && out_cap)) return FOF_E_ARG;
  for (uint32_t i = 0; i < n; i++)
    if (next[i] < 1 || next[i] > n) return FOF_E_ARG;
  const uint8_t *base = static_cast<const uint8_t *>(P->base);
  // Pos is float, the box double: the comparisons promote (src_v2/FOF.c:209-231)
  auto outside = [&](uint32_t ip) {
    const uint8_t *r = base + (uint64_t)(ip - 1) * P->stride + P->off_pos;
    for (int a = 0; a < 3; a++) {
      float x;
      memcpy(&x, r + 4 * a, 4);
      if ((double)x < s->rmin[a]) return true;
      if (((s->closed_upper >> a) & 1u) ? ((double)x > s->rmax[a]) : ((double)x >= s->rmax[a]))
        return true;
    }
    return false;
  };
  std::vector<uint8_t> seen((size_t)n + 1, 0);
  uint64_t ni = 0, no = 0;
  int rc = 0;
  for (uint32_t i = 1; i <= n && !rc; i++) {
    if (seen[i]) continue;
    uint32_t ip = i;
    uint64_t np = 0;
    for (;;) {  // the group's size
      ip = next[ip - 1];
      np++;
      if (ip == i) break;
      if (seen[ip] || np >= n) return FOF_E_ARG;  // not a ring through i
      seen[ip] = 1;
    }
    const bool halo = !((double)np < s->nphalomin);  // src_v2/FOF.c:202
    const double rate = halo ? s->samp_in : s->samp_out;
    if (!(rate > 0)) continue;
    uint32_t *list = halo ? inhalo : outhalo;
    uint64_t &cnt = halo ? ni : no;
    const uint64_t cap = halo ? in_cap : out_cap;
    ip = i;
    for (;;) {
      ip = next[ip - 1];
      if (uniform(state) < rate) {
        if (outside(ip)) break;
        if (cnt >= cap) {
          rc = FOF_E_RANGE;
          break;
        }
        list[cnt++] = ip;
      }
      if (ip == i) break;
    }
  }
  *n_in = ni, *n_out = no;
  return rc;
}

// src_v2/FOF.c:432-544
int fof_write_subsample(const char *path, uint32_t flags, const fof_particles_v1 *P,
                        const uint32_t *idx, uint64_t n) {
  if (!path || !P || (n && !idx)) return FOF_E_ARG;
  if (P->version != FOF_ABI_VERSION || P->location != FOF_MEM_HOST || (P->n && !P->base))
    return FOF_E_ARG;
  const bool unf = (flags & FOF_OUT_UNFORMATTED) != 0, withid = (flags & FOF_OUT_PARTICLE_ID) != 0;
  if (withid && P->off_id < 0) return FOF_E_ARG;
  for (uint64_t k = 0; k < n; k++)
    if (idx[k] < 1 || idx[k] > P->n) return FOF_E_ARG;
  if (n == 0) return 0;  // nsubsamp > 0 guards the file (src_v2/FOF.c:457)
  File F(path, unf ? "wb" : "w");
  if (!F.f) return FOF_E_ARG;
  FILE *f = F.f;
  const uint8_t *base = static_cast<const uint8_t *>(P->base);
  auto rec = [&](uint64_t k) { return base + (uint64_t)(idx[k] - 1) * P->stride; };
  if (unf) {
    // the reference's record lengths are int products (truncated like them)
    const uint32_t cnt = (uint32_t)n;
    put_rec(f, 4);
    fwrite(&cnt, 4, 1, f);
    put_rec(f, 4);
    for (int part = 0; part < 2; part++) {  // all positions, then all velocities
      const uint32_t off = part ? P->off_vel : P->off_pos;
      put_rec(f, (uint32_t)(12u * n));
      for (uint64_t k = 0; k < n; k++) fwrite(rec(k) + off, 4, 3, f);
      put_rec(f, (uint32_t)(12u * n));
    }
    if (withid) {
      put_rec(f, (uint32_t)(8u * n));
      for (uint64_t k = 0; k < n; k++) fwrite(rec(k) + P->off_id, 8, 1, f);
      put_rec(f, (uint32_t)(8u * n));
    }
  } else {
    for (uint64_t k = 0; k < n; k++) {
      float v[6];
      memcpy(v, rec(k) + P->off_pos, 12);
      memcpy(v + 3, rec(k) + P->off_vel, 12);
      if (withid) {
        unsigned long long id;
        memcpy(&id, rec(k) + P->off_id, 8);
        fprintf(f, "%15llu %15.6f %15.6f %15.6f %15.6f %15.6f %15.6f\n", id, v[0], v[1], v[2], v[3],
                v[4], v[5]);
      } else {
        fprintf(f, "%15.6f %15.6f %15.6f %15.6f %15.6f %15.6f\n", v[0], v[1], v[2], v[3], v[4], v[5]);
      }
    }
  }
  return ferror(f) ? FOF_E_ARG : 0;
}

// src_v2/FOF.c:917-957
int fof_write_info(const char *path, const double lmin[3], const double lmax[3], int nx, int ny,
                   int nz, const int32_t *nhalos, const uint32_t *ninhalo,
                   const uint32_t *nouthalo) {
  if (!path || !lmin || !lmax || nx < 1 || ny < 1 || nz < 1 || !nhalos) return FOF_E_ARG;
  File F(path, "w");
  if (!F.f) return FOF_E_ARG;
  fprintf(F.f, "#    FILENUM      XMIN         YMIN        ZMIN         XMAX         YMAX         "
               "ZMAX         NHALOS     NSUBSAMPIN     NSUBSAMPOUT    \n");
  const int ntask = nx * ny * nz;
  for (int i = 0; i < ntask; i++) {
    const int tz = i / (nx * ny);
    const int ty = (i - nx * ny * tz) / nx;
    const int tx = i - nx * ty - nx * ny * tz;
    const double x0 = tx * ((lmax[0] - lmin[0]) / (double)nx) + lmin[0];
    const double y0 = ty * ((lmax[1] - lmin[1]) / (double)ny) + lmin[1];
    const double z0 = tz * ((lmax[2] - lmin[2]) / (double)nz) + lmin[2];
    const double x1 = x0 + ((lmax[0] - lmin[0]) / (double)nx);
    const double y1 = y0 + ((lmax[1] - lmin[1]) / (double)ny);
    const double z1 = z0 + ((lmax[2] - lmin[2]) / (double)nz);
    fprintf(F.f, "%12d %12.6lf %12.6lf %12.6lf %12.6lf %12.6lf %12.6lf %12d %12u %12u\n", i, x0, y0,
            z0, x1, y1, z1, nhalos[i], ninhalo ? ninhalo[i] : 0u, nouthalo ? nouthalo[i] : 0u);
  }
  return ferror(F.f) ? FOF_E_ARG : 0;
}

}  // extern "C"
