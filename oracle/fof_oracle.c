/*
 * fof_oracle.c -- CPU oracle for the CM_Halofinder Friends-of-Friends path.
 *
 * TEST INFRASTRUCTURE ONLY (see fof_oracle.h).  Plain C, single thread.
 * Built with -ffp-contract=off so that every FP64 expression rounds exactly
 * as written, as in the reference's x86-64 SSE2 build (no FMA, float
 * expressions evaluated in float, FLT_EVAL_METHOD == 0).
 */
#include "fof_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* src_v2/FOF.c:351-357: dx = P[ip-1].Pos[0]-P[ip2-1].Pos[0] is a float
 * subtraction (both operands float) widened to double; distsq is summed left
 * to right in double; the pair links unless distsq > linksq. */
int or_pair_linked(const float *pi, const float *pj, double linksq) {
  float fx = pi[0] - pj[0];
  float fy = pi[1] - pj[1];
  float fz = pi[2] - pj[2];
  double dx = fx, dy = fy, dz = fz;
  double distsq = dx * dx + dy * dy + dz * dz;
  return !(distsq > linksq);
}

double or_cell_linksq(const or_domain *d, const or_varlink *vl, int ix, int iy,
                      int iz) {
  if (!vl) return d->linklength * d->linklength; /* src_v2/FOF.c:60 */
  /* Intent of src_v2/FOF.c:105-108: squared distance of the cell centre to
   * the origin, looked up in a table.  Centre uses +0.5 on every axis. */
  double c0 = d->rmin_buff[0] + ((double)ix + 0.5) * d->lcell[0];
  double c1 = d->rmin_buff[1] + ((double)iy + 0.5) * d->lcell[1];
  double c2 = d->rmin_buff[2] + ((double)iz + 0.5) * d->lcell[2];
  double e0 = c0 - vl->origin[0], e1 = c1 - vl->origin[1],
         e2 = c2 - vl->origin[2];
  double r2 = e0 * e0 + e1 * e1 + e2 * e2;
  double top = (double)(vl->ntab - 1);
  double t = r2 / vl->r2max * top;
  if (!(t < top)) return vl->l2tab[vl->ntab - 1];
  int k = (int)floor(t);
  double f = t - (double)k;
  return vl->l2tab[k] + f * (vl->l2tab[k + 1] - vl->l2tab[k]);
}

typedef struct grid {
  int nx, ny, nz;
  uint64_t ncells;
} grid;

/* src_v2/FOF.c:63-65. */
static int grid_setup(const or_domain *d, grid *g) {
  double ex = ceil((d->rmax_buff[0] - d->rmin_buff[0]) / d->lcell[0]) + 1.0;
  double ey = ceil((d->rmax_buff[1] - d->rmin_buff[1]) / d->lcell[1]) + 1.0;
  double ez = ceil((d->rmax_buff[2] - d->rmin_buff[2]) / d->lcell[2]) + 1.0;
  if (!(ex >= 1 && ey >= 1 && ez >= 1) || ex * ey * ez > 4294967294.0)
    return -3;
  g->nx = (int)ex;
  g->ny = (int)ey;
  g->nz = (int)ez;
  g->ncells = (uint64_t)g->nx * g->ny * g->nz;
  return 0;
}

/* src_v2/read_data.c:810-812 bounds, plus finiteness. */
static int check_domain(uint32_t n, const float *pos, const or_domain *d) {
  for (uint32_t i = 0; i < n; i++)
    for (int k = 0; k < 3; k++) {
      double x = pos[3 * i + k];
      if (!(x >= d->rmin_buff[k] && x <= d->rmax_buff[k])) return -2;
    }
  return 0;
}

/* src_v2/FOF.c:79-83. */
static void cell_of(const float *p, const or_domain *d, int *ix, int *iy,
                    int *iz) {
  *ix = (int)floor(((double)p[0] - d->rmin_buff[0]) / d->lcell[0]);
  *iy = (int)floor(((double)p[1] - d->rmin_buff[1]) / d->lcell[1]);
  *iz = (int)floor(((double)p[2] - d->rmin_buff[2]) / d->lcell[2]);
}

/* ---------------------------------------------------------------------------
 * Ring linker: src_v2/FOF.c:63-161 with Befriend, src_v2/FOF.c:341-376.
 * next/chain/first are 1-based as in the reference (index 0 = "none").
 * ------------------------------------------------------------------------- */
static void befriend(uint32_t ip, uint32_t ip2, const float *pos,
                     double linksq, uint32_t *next, uint64_t *pairs) {
  (*pairs)++;
  if (!or_pair_linked(&pos[3 * (ip - 1)], &pos[3 * (ip2 - 1)], linksq)) return;
  uint32_t ip3 = ip;
  for (;;) { /* src_v2/FOF.c:359-367 */
    ip3 = next[ip3];
    if (ip3 == ip) break;
    if (ip3 == ip2) return;
  }
  ip3 = next[ip]; /* src_v2/FOF.c:369-373 */
  next[ip] = next[ip2];
  next[ip2] = ip3;
}

int or_link_ring(uint32_t n, const float *pos, const or_domain *d,
                 const or_varlink *vl, uint32_t *next0, uint64_t *pairs_out) {
  grid g;
  int rc = grid_setup(d, &g);
  if (rc) return rc;
  if ((rc = check_domain(n, pos, d))) return rc;
  uint32_t *next = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
  uint32_t *chain = (uint32_t *)calloc((size_t)n + 1, sizeof(uint32_t));
  uint32_t *first = (uint32_t *)calloc(g.ncells, sizeof(uint32_t));
  if (!next || !chain || !first) {
    free(next), free(chain), free(first);
    return -4;
  }
  const uint64_t nx = g.nx, ny = g.ny;
  for (uint32_t i = 1; i <= n; i++) { /* src_v2/FOF.c:76-92 */
    next[i] = i;
    int ix, iy, iz;
    cell_of(&pos[3 * (i - 1)], d, &ix, &iy, &iz);
    uint64_t ind = (uint64_t)iz * ny * nx + (uint64_t)iy * nx + (uint64_t)ix;
    uint32_t ip = first[ind];
    if (ip != 0) chain[i] = ip;
    first[ind] = i;
  }
  uint64_t pairs = 0;
  double linksq = d->linklength * d->linklength;
  for (int iz = 0; iz < g.nz; iz++)
    for (int iy = 0; iy < g.ny; iy++)
      for (int ix = 0; ix < g.nx; ix++) { /* src_v2/FOF.c:96-161 */
        uint32_t ip =
            first[(uint64_t)iz * ny * nx + (uint64_t)iy * nx + (uint64_t)ix];
        if (ip == 0) continue;
        if (vl) linksq = or_cell_linksq(d, vl, ix, iy, iz);
        do {
          uint32_t ip2 = ip;
          for (;;) {
            ip2 = chain[ip2];
            if (ip2 == 0) break;
            befriend(ip, ip2, pos, linksq, next, &pairs);
          }
          for (int iz2 = iz; iz2 <= iz + 1; iz2++) {
            if (iz2 >= g.nz) continue;
            int iy1 = (iz2 == iz) ? iy : iy - 1;
            for (int iy2 = iy1; iy2 <= iy + 1; iy2++) {
              if (iy2 >= g.ny || iy2 < 0) continue;
              int ix1 = (iz2 == iz && iy2 == iy) ? ix : ix - 1;
              for (int ix2 = ix1; ix2 <= ix + 1; ix2++) {
                if (ix2 >= g.nx || ix2 < 0) continue;
                if (ix2 == ix && iy2 == iy && iz2 == iz) continue;
                ip2 = first[(uint64_t)iz2 * ny * nx + (uint64_t)iy2 * nx +
                            (uint64_t)ix2];
                if (ip2 == 0) continue;
                befriend(ip, ip2, pos, linksq, next, &pairs);
                for (;;) {
                  ip2 = chain[ip2];
                  if (ip2 == 0) break;
                  befriend(ip, ip2, pos, linksq, next, &pairs);
                }
              }
            }
          }
          ip = chain[ip];
        } while (ip != 0);
      }
  for (uint32_t i = 1; i <= n; i++) next0[i - 1] = next[i] - 1;
  if (pairs_out) *pairs_out = pairs;
  free(next), free(chain), free(first);
  return 0;
}

/* ---------------------------------------------------------------------------
 * Union-find linker over the same candidate set (same cell + 13 forward
 * cells, src_v2/FOF.c:112-153) and the same predicate.  Roots are hooked
 * larger-under-smaller, so each root is its component's minimum index.
 * ------------------------------------------------------------------------- */
static uint32_t uf_find(uint32_t *par, uint32_t x) {
  while (par[x] != x) {
    par[x] = par[par[x]];
    x = par[x];
  }
  return x;
}
static void uf_union(uint32_t *par, uint32_t a, uint32_t b) {
  a = uf_find(par, a);
  b = uf_find(par, b);
  if (a == b) return;
  if (a < b) par[b] = a;
  else par[a] = b;
}

int or_link_uf(uint32_t n, const float *pos, const or_domain *d,
               const or_varlink *vl, uint32_t *label0, uint64_t *pairs_out) {
  grid g;
  int rc = grid_setup(d, &g);
  if (rc) return rc;
  if ((rc = check_domain(n, pos, d))) return rc;
  uint64_t *start = (uint64_t *)calloc(g.ncells + 1, sizeof(uint64_t));
  uint32_t *order = (uint32_t *)malloc((size_t)n * sizeof(uint32_t) + 1);
  uint64_t *cell = (uint64_t *)malloc((size_t)n * sizeof(uint64_t) + 1);
  if (!start || !order || !cell) {
    free(start), free(order), free(cell);
    return -4;
  }
  const uint64_t nx = g.nx, ny = g.ny;
  for (uint32_t i = 0; i < n; i++) {
    int ix, iy, iz;
    cell_of(&pos[3 * i], d, &ix, &iy, &iz);
    cell[i] = (uint64_t)iz * ny * nx + (uint64_t)iy * nx + (uint64_t)ix;
    start[cell[i] + 1]++;
  }
  for (uint64_t c = 0; c < g.ncells; c++) start[c + 1] += start[c];
  for (uint32_t i = 0; i < n; i++) order[start[cell[i]]++] = i;
  for (uint64_t c = g.ncells; c > 0; c--) start[c] = start[c - 1];
  start[0] = 0;
  for (uint32_t i = 0; i < n; i++) label0[i] = i;
  uint64_t pairs = 0;
  double linksq = d->linklength * d->linklength;
  for (int iz = 0; iz < g.nz; iz++)
    for (int iy = 0; iy < g.ny; iy++)
      for (int ix = 0; ix < g.nx; ix++) {
        uint64_t c = (uint64_t)iz * ny * nx + (uint64_t)iy * nx + (uint64_t)ix;
        uint64_t a0 = start[c], a1 = start[c + 1];
        if (a0 == a1) continue;
        if (vl) linksq = or_cell_linksq(d, vl, ix, iy, iz);
        for (uint64_t a = a0; a < a1; a++) {
          uint32_t i = order[a];
          for (uint64_t b = a + 1; b < a1; b++) {
            pairs++;
            if (or_pair_linked(&pos[3 * i], &pos[3 * order[b]], linksq))
              uf_union(label0, i, order[b]);
          }
        }
        for (int iz2 = iz; iz2 <= iz + 1; iz2++) {
          if (iz2 >= g.nz) continue;
          int iy1 = (iz2 == iz) ? iy : iy - 1;
          for (int iy2 = iy1; iy2 <= iy + 1; iy2++) {
            if (iy2 >= g.ny || iy2 < 0) continue;
            int ix1 = (iz2 == iz && iy2 == iy) ? ix : ix - 1;
            for (int ix2 = ix1; ix2 <= ix + 1; ix2++) {
              if (ix2 >= g.nx || ix2 < 0) continue;
              if (ix2 == ix && iy2 == iy && iz2 == iz) continue;
              uint64_t c2 =
                  (uint64_t)iz2 * ny * nx + (uint64_t)iy2 * nx + (uint64_t)ix2;
              for (uint64_t a = a0; a < a1; a++)
                for (uint64_t b = start[c2]; b < start[c2 + 1]; b++) {
                  pairs++;
                  if (or_pair_linked(&pos[3 * order[a]], &pos[3 * order[b]],
                                     linksq))
                    uf_union(label0, order[a], order[b]);
                }
            }
          }
        }
      }
  for (uint32_t i = 0; i < n; i++) label0[i] = uf_find(label0, i);
  if (pairs_out) *pairs_out = pairs;
  free(start), free(order), free(cell);
  return 0;
}

/* ---------------------------------------------------------------------------
 * Brute force: every pair, the reference predicate, and the reference's
 * candidate rule (cells at Chebyshev distance <= 1; the home cell is the one
 * with the lower linear index).  No cell lists.
 * ------------------------------------------------------------------------- */
int or_link_brute(uint32_t n, const float *pos, const or_domain *d,
                  const or_varlink *vl, uint32_t *label0) {
  grid g;
  int rc = grid_setup(d, &g);
  if (rc) return rc;
  if ((rc = check_domain(n, pos, d))) return rc;
  int *ci = (int *)malloc(3 * (size_t)n * sizeof(int) + 1);
  if (!ci) return -4;
  for (uint32_t i = 0; i < n; i++)
    cell_of(&pos[3 * i], d, &ci[3 * i], &ci[3 * i + 1], &ci[3 * i + 2]);
  for (uint32_t i = 0; i < n; i++) label0[i] = i;
  const double linksq0 = d->linklength * d->linklength;
  for (uint32_t i = 0; i < n; i++)
    for (uint32_t j = i + 1; j < n; j++) {
      int ddx = abs(ci[3 * i] - ci[3 * j]), ddy = abs(ci[3 * i + 1] - ci[3 * j + 1]),
          ddz = abs(ci[3 * i + 2] - ci[3 * j + 2]);
      if (ddx > 1 || ddy > 1 || ddz > 1) continue;
      double linksq = linksq0;
      if (vl) {
        uint64_t li = ((uint64_t)ci[3 * i + 2] * g.ny + ci[3 * i + 1]) * g.nx +
                      ci[3 * i];
        uint64_t lj = ((uint64_t)ci[3 * j + 2] * g.ny + ci[3 * j + 1]) * g.nx +
                      ci[3 * j];
        const int *h = (li <= lj) ? &ci[3 * i] : &ci[3 * j];
        linksq = or_cell_linksq(d, vl, h[0], h[1], h[2]);
      }
      if (or_pair_linked(&pos[3 * i], &pos[3 * j], linksq))
        uf_union(label0, i, j);
    }
  for (uint32_t i = 0; i < n; i++) label0[i] = uf_find(label0, i);
  free(ci);
  return 0;
}

void or_labels_from_ring(uint32_t n, const uint32_t *next0, uint32_t *label0) {
  const uint32_t none = 0xFFFFFFFFu;
  for (uint32_t i = 0; i < n; i++) label0[i] = none;
  for (uint32_t i = 0; i < n; i++) {
    if (label0[i] != none) continue;
    uint32_t j = i;
    do {
      label0[j] = i;
      j = next0[j];
    } while (j != i && j < n);
  }
}

void or_ring_from_labels(uint32_t n, const uint32_t *label0, uint32_t *next0) {
  /* last[h] = last member of group h seen so far in ascending order. */
  uint32_t *last = (uint32_t *)malloc((size_t)n * sizeof(uint32_t) + 1);
  for (uint32_t i = 0; i < n; i++) last[i] = 0xFFFFFFFFu;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t h = label0[i];
    if (last[h] != 0xFFFFFFFFu) next0[last[h]] = i;
    last[h] = i;
  }
  for (uint32_t i = 0; i < n; i++)
    if (last[i] != 0xFFFFFFFFu) next0[last[i]] = i; /* close: last -> head */
  free(last);
}

/* src_v2/FOF.c:188-203 (heads, size cut), :295-298 + Checkhalo :378-430
 * (ownership, boundary warnings) and Output_Halos :797-847 (sums). */
int64_t or_retrieve(uint32_t n, const uint32_t *next0, const float *pos,
                    const float *vel, const or_domain *d, or_halo *out,
                    uint64_t cap) {
  unsigned char *head = (unsigned char *)calloc((size_t)n + 1, 1);
  if (!head) return -4;
  for (uint32_t i = 0; i < n; i++) {
    if (head[i]) continue;
    uint32_t ip = i;
    int nphalo = 0;
    for (;;) {
      ip = next0[ip];
      nphalo++;
      if (ip == i) break;
      head[ip] = 1;
    }
    if ((double)nphalo < d->nphalomin) head[i] = 1;
  }
  int64_t nh = 0;
  for (uint32_t i = 0; i < n; i++) {
    if (head[i]) continue;
    or_halo h;
    memset(&h, 0, sizeof(h));
    h.head = i;
    double xh = 0, yh = 0, zh = 0, vx = 0, vy = 0, vz = 0;
    uint32_t j = i;
    int np = 0;
    do {
      np++;
      xh += pos[3 * j];
      yh += pos[3 * j + 1];
      zh += pos[3 * j + 2];
      vx += vel[3 * j];
      vy += vel[3 * j + 1];
      vz += vel[3 * j + 2];
      j = next0[j];
    } while (j != i);
    h.n = (uint32_t)np;
    h.sum_pos[0] = xh, h.sum_pos[1] = yh, h.sum_pos[2] = zh;
    h.sum_vel[0] = vx, h.sum_vel[1] = vy, h.sum_vel[2] = vz;
    double cx = xh / (double)np, cy = yh / (double)np, cz = zh / (double)np;
    h.owned = !((cx < d->rmin[0]) || (cx >= d->rmax[0]) || (cy < d->rmin[1]) ||
                (cy >= d->rmax[1]) || (cz < d->rmin[2]) || (cz >= d->rmax[2]));
    if (h.owned) {
      j = i;
      do {
        for (int k = 0; k < 3; k++) {
          double p = pos[3 * j + k];
          if ((p - d->rmin_buff[k] < d->linklength) ||
              (d->rmax_buff[k] - p < d->linklength))
            h.edge_warn |= 1u << k;
        }
        j = next0[j];
      } while (j != i);
    }
    double inv = 1.0 / (double)np;
    h.out[0] = (float)(xh * inv), h.out[1] = (float)(yh * inv);
    h.out[2] = (float)(zh * inv), h.out[3] = (float)(vx * inv);
    h.out[4] = (float)(vy * inv), h.out[5] = (float)(vz * inv);
    if ((uint64_t)nh < cap) out[nh] = h;
    nh++;
  }
  free(head);
  return nh;
}

/* src_v2/FOF.c:172-284, the PERIODIC and the non-PERIODIC bounds tests
 * (:209-231, :251-273) selected per axis by closed_upper. */
static int or_sub_outside(const float *p, const or_subsamp *s) {
  for (int a = 0; a < 3; a++) {
    if (p[a] < s->rmin[a]) return 1;
    if ((s->closed_upper >> a) & 1u) {
      if (p[a] > s->rmax[a]) return 1;
    } else {
      if (p[a] >= s->rmax[a]) return 1;
    }
  }
  return 0;
}

int or_subsample(uint32_t n, const uint32_t *next0, const float *pos,
                 const or_subsamp *s, const double *u, uint64_t nu,
                 uint64_t *used, uint32_t *inhalo, uint64_t *ninhalo,
                 uint32_t *outhalo, uint64_t *nouthalo) {
  /* 1-based next[] and head[] as in the reference */
  uint32_t *next = (uint32_t *)malloc(((size_t)n + 1) * sizeof(uint32_t));
  unsigned char *head = (unsigned char *)calloc((size_t)n + 1, 1);
  uint64_t k = 0, ni = 0, no = 0;
  int rc = 0;
  for (uint32_t i = 0; i < n; i++) next[i + 1] = next0[i] + 1;
  for (uint32_t i = 1; i <= n && !rc; i++) {
    if (head[i] == 0) {
      uint32_t ip = i;
      int nphalo = 0;
      do {
        ip = next[ip];
        nphalo++;
        if (ip == i) break;
        head[ip] = 1;
      } while (1);
      if (nphalo < s->nphalomin) {
        head[ip] = 1;
        if (s->samp_out > 0) {
          do {
            ip = next[ip];
            if (k >= nu) { rc = -1; break; }
            if (u[k++] < s->samp_out) {
              if (or_sub_outside(pos + 3 * (size_t)(ip - 1), s)) break;
              outhalo[no++] = ip;
            }
            if (ip == i) break;
          } while (1);
        }
      } else {
        if (s->samp_in > 0) {
          do {
            ip = next[ip];
            if (k >= nu) { rc = -1; break; }
            if (u[k++] < s->samp_in) {
              if (or_sub_outside(pos + 3 * (size_t)(ip - 1), s)) break;
              inhalo[ni++] = ip;
            }
            if (ip == i) break;
          } while (1);
        }
      }
    }
  }
  free(next);
  free(head);
  *used = k, *ninhalo = ni, *nouthalo = no;
  return rc;
}

/* Neumaier summation: s + c carries ~106 bits. */
static void neu_add(double *s, double *c, double x) {
  double t = *s + x;
  if (fabs(*s) >= fabs(x)) *c += (*s - t) + x;
  else *c += (x - t) + *s;
  *s = t;
}

void or_props_comp(uint32_t n, const uint32_t *label0, const float *pos,
                   const float *vel, uint64_t nh, const uint32_t *heads,
                   double *sum_pos, double *sum_vel) {
  uint32_t *slot = (uint32_t *)malloc((size_t)n * sizeof(uint32_t) + 1);
  double *acc = (double *)calloc(12 * nh + 1, sizeof(double));
  for (uint32_t i = 0; i < n; i++) slot[i] = 0xFFFFFFFFu;
  for (uint64_t k = 0; k < nh; k++) slot[heads[k]] = (uint32_t)k;
  for (uint32_t i = 0; i < n; i++) {
    uint32_t k = slot[label0[i]];
    if (k == 0xFFFFFFFFu) continue;
    for (int q = 0; q < 3; q++) {
      neu_add(&acc[12 * k + q], &acc[12 * k + 6 + q], pos[3 * i + q]);
      neu_add(&acc[12 * k + 3 + q], &acc[12 * k + 9 + q], vel[3 * i + q]);
    }
  }
  for (uint64_t k = 0; k < nh; k++)
    for (int q = 0; q < 3; q++) {
      sum_pos[3 * k + q] = acc[12 * k + q] + acc[12 * k + 6 + q];
      sum_vel[3 * k + q] = acc[12 * k + 3 + q] + acc[12 * k + 9 + q];
    }
  free(slot), free(acc);
}

/* Second moments of halos (src_v2/FOF.c:811-824: INERTIA / DISPERSION
 * accumulators xh2, yh2, zh2, xhyh, xhzh, yhzh and vxh2 ... vyhvzh).
 * raw[12k + q]: FP64 sums in ring order from heads[k] of the products
 * (double)a * (double)b of the float members, q = 0-5 Pos {xx,yy,zz,xy,xz,yz},
 * 6-11 Vel.  cen[12k + q]: the centred sums sum (a - <a>)(b - <b>), from a
 * compensated mean and a compensated second pass. */
static const int mom_a[6] = {0, 1, 2, 0, 0, 1}, mom_b[6] = {0, 1, 2, 1, 2, 2};

void or_moments(uint32_t n, const uint32_t *next0, const float *pos,
                const float *vel, uint64_t nh, const uint32_t *heads,
                double *raw, double *cen) {
  for (uint64_t k = 0; k < nh; k++) {
    const uint32_t i = heads[k];
    double r[12] = {0};
    double s[6] = {0}, c[6] = {0};
    uint32_t j = i;
    uint32_t np = 0;
    do {
      const float *p = &pos[3 * j], *v = &vel[3 * j];
      for (int q = 0; q < 6; q++) {
        r[q] += (double)p[mom_a[q]] * (double)p[mom_b[q]];
        r[6 + q] += (double)v[mom_a[q]] * (double)v[mom_b[q]];
      }
      for (int q = 0; q < 3; q++) {
        neu_add(&s[q], &c[q], p[q]);
        neu_add(&s[3 + q], &c[3 + q], v[q]);
      }
      np++;
      j = next0[j];
    } while (j != i && j < n);
    double m[6];
    for (int q = 0; q < 6; q++) m[q] = (s[q] + c[q]) / (double)np;
    double t[12] = {0}, tc[12] = {0};
    j = i;
    do {
      const float *p = &pos[3 * j], *v = &vel[3 * j];
      for (int q = 0; q < 6; q++) {
        neu_add(&t[q], &tc[q], ((double)p[mom_a[q]] - m[mom_a[q]]) * ((double)p[mom_b[q]] - m[mom_b[q]]));
        neu_add(&t[6 + q], &tc[6 + q],
                ((double)v[mom_a[q]] - m[3 + mom_a[q]]) * ((double)v[mom_b[q]] - m[3 + mom_b[q]]));
      }
      j = next0[j];
    } while (j != i && j < n);
    for (int q = 0; q < 12; q++) {
      raw[12 * k + q] = r[q];
      cen[12 * k + q] = t[q] + tc[q];
    }
  }
}

/* The values the reference writes for a halo with INERTIA and DISPERSION
 * (src_v2/FOF.c:848-860, stored as float): inertia {Ixx, Iyy, Izz, Ixy, Ixz,
 * Iyz} from the raw sums with its single-pass formulas, and the velocity
 * dispersion tensor {xx, yy, zz, xy, xz, yz}. */
void or_moments_reference_out(uint32_t np, const double *sum_pos,
                              const double *sum_vel, const double *raw,
                              float *inertia, float *disp) {
  const double inv = 1.0 / (double)np;
  const double xh = sum_pos[0], yh = sum_pos[1], zh = sum_pos[2];
  const double vxh = sum_vel[0], vyh = sum_vel[1], vzh = sum_vel[2];
  const double xh2 = raw[0], yh2 = raw[1], zh2 = raw[2], xhyh = raw[3], xhzh = raw[4],
               yhzh = raw[5];
  const double vxh2 = raw[6], vyh2 = raw[7], vzh2 = raw[8], vxhvyh = raw[9],
               vxhvzh = raw[10], vyhvzh = raw[11];
  inertia[0] = (float)(yh2 + zh2 - inv * (yh * yh + zh * zh));
  inertia[1] = (float)(xh2 + zh2 - inv * (xh * xh + zh * zh));
  inertia[2] = (float)(xh2 + yh2 - inv * (xh * xh + yh * yh));
  inertia[3] = (float)(inv * xh * yh - xhyh);
  inertia[4] = (float)(inv * xh * zh - xhzh);
  inertia[5] = (float)(inv * yh * zh - yhzh);
  disp[0] = (float)((vxh2 - vxh * vxh * inv) * inv);
  disp[1] = (float)((vyh2 - vyh * vyh * inv) * inv);
  disp[2] = (float)((vzh2 - vzh * vzh * inv) * inv);
  disp[3] = (float)((vxhvyh - vxh * vyh * inv) * inv);
  disp[4] = (float)((vxhvzh - vxh * vzh * inv) * inv);
  disp[5] = (float)((vyhvzh - vyh * vzh * inv) * inv);
}
