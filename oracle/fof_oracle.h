/*
 * fof_oracle.h -- CPU oracle for the CM_Halofinder Friends-of-Friends path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference's linking and catalogue arithmetic, used as the checker by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The
 * product (cm_halofinder_amd) never links or calls it.
 *
 * Positions and velocities are passed as interleaved float xyz triples
 * (pos[3*i+d]), i.e. the Pos/Vel members of the reference's particle record
 * (src_v2/vars.h:36-47, MEMORY_MODE).  Indices are 0-based; the reference's
 * 1-based next[i] is next0[i-1]+1.
 *
 * Pinning (what checks each function independently of this restatement):
 *   or_pair_linked     exact ties d^2 == l^2 (Pythagorean quadruples at
 *                      l = 5/16, 25/128), one-ulp offsets, NaN/duplicate rules
 *   or_link_brute      O(n^2) over all pairs; no grid
 *   or_link_uf         equals or_link_brute when l < lcell; invariants
 *                      (permutation, lattice a<=l / a>l, chains)
 *   or_link_ring       same partition as or_link_brute / or_link_uf
 *   or_retrieve        closed forms (lattice clumps, symmetric groups),
 *                      heads = minimum index, size cut, ownership boundaries
 *   or_props_comp      compensated sums vs exact rational sums in tests
 *   or_moments         closed forms (axis-symmetric clusters), exact
 *                      rational sums (raw in ring order to FP64 rounding,
 *                      centred to a few ulps)
 *   or_moments_reference_out  the reference's formulas on hand-checked
 *                      small cases (exact in binary fractions)
 *   or_subsample       hand-built rings with scripted draws: every member
 *                      at rate 1, alternate draws, a member outside the box
 *                      ending its group's walk, the closed upper face, the
 *                      size cut's < comparison, the number of draws taken
 * The reference itself cannot be compiled here (vars.h includes GSL headers
 * the image lacks), so no golden output of the reference exists.
 */
#ifndef CM_FOF_ORACLE_H
#define CM_FOF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_domain {
  double rmin[3], rmax[3], rmin_buff[3], rmax_buff[3];
  double lcell[3];
  double linklength;
  double nphalomin;
} or_domain;

/* Variable linking length: l^2 of a cell = linear interpolation on
 * uniform r2 samples of the squared distance of the cell centre to origin. */
typedef struct or_varlink {
  double origin[3];
  double r2max;
  uint32_t ntab;
  const double *l2tab;
} or_varlink;

typedef struct or_halo {
  uint32_t head;      /* 0-based minimum index                              */
  uint32_t n;
  uint32_t owned;
  uint32_t edge_warn;
  double sum_pos[3];  /* FP64, in ring order from the head                  */
  double sum_vel[3];
  float out[6];       /* Output_Halos values: (float)(sum * (1.0/n))        */
} or_halo;

/* Reference predicate, src_v2/FOF.c:351-357. */
int or_pair_linked(const float *pi, const float *pj, double linksq);

/* Squared linking length of reference cell (ix,iy,iz). */
double or_cell_linksq(const or_domain *d, const or_varlink *vl, int ix, int iy,
                      int iz);

/* Linkers.  Return 0 or a negative error (-2: particle outside the buffered
 * box or non-finite; -3: grid too large; -4: out of memory).  *pairs (may be
 * NULL) receives the number of predicate evaluations. */
int or_link_ring(uint32_t n, const float *pos, const or_domain *d,
                 const or_varlink *vl, uint32_t *next0, uint64_t *pairs);
int or_link_uf(uint32_t n, const float *pos, const or_domain *d,
               const or_varlink *vl, uint32_t *label0, uint64_t *pairs);
int or_link_brute(uint32_t n, const float *pos, const or_domain *d,
                  const or_varlink *vl, uint32_t *label0);

/* Ring <-> labels. */
void or_labels_from_ring(uint32_t n, const uint32_t *next0, uint32_t *label0);
void or_ring_from_labels(uint32_t n, const uint32_t *label0, uint32_t *next0);

/* Retrieval + Checkhalo + Output_Halos sums over a ring.  Writes every group
 * passing the size cut (owned or not) in ascending head order, up to cap;
 * returns the number of such groups. */
int64_t or_retrieve(uint32_t n, const uint32_t *next0, const float *pos,
                    const float *vel, const or_domain *d, or_halo *out,
                    uint64_t cap);

/* Compensated (Neumaier) FP64 sums of pos/vel per group head: for each halo
 * head h[k] (label value), sums over all i with label0[i]==h[k]. */
void or_props_comp(uint32_t n, const uint32_t *label0, const float *pos,
                   const float *vel, uint64_t nh, const uint32_t *heads,
                   double *sum_pos, double *sum_vel);

/* Second moments over a ring (src_v2/FOF.c:811-824), raw[12*nh] in ring
 * order and centred cen[12*nh] (compensated two-pass); q = 0-5 Pos
 * {xx,yy,zz,xy,xz,yz}, 6-11 Vel. */
void or_moments(uint32_t n, const uint32_t *next0, const float *pos,
                const float *vel, uint64_t nh, const uint32_t *heads,
                double *raw, double *cen);

/* The reference's written INERTIA / DISPERSION values (src_v2/FOF.c:848-860)
 * from the FP64 sums and raw second moments of one halo. */
void or_moments_reference_out(uint32_t np, const double *sum_pos,
                              const double *sum_vel, const double *raw,
                              float *inertia, float *disp);

/* Subsampling (src_v2/FOF.c:172-284) over a ring, the draws taken in order
 * from u[0..nu) (the reference's gsl_rng_uniform calls).  Kept particles go
 * 1-based, as the reference's ip, to inhalo / outhalo (room for n each);
 * *used receives the number of draws taken.  Returns -1 when the draws run
 * out, else 0. */
typedef struct or_subsamp {
  double samp_in, samp_out;
  double rmin[3], rmax[3];
  double nphalomin;
  uint32_t closed_upper; /* bit a: the last rank along axis a without PERIODIC */
} or_subsamp;

int or_subsample(uint32_t n, const uint32_t *next0, const float *pos,
                 const or_subsamp *s, const double *u, uint64_t nu,
                 uint64_t *used, uint32_t *inhalo, uint64_t *ninhalo,
                 uint32_t *outhalo, uint64_t *nouthalo);

#ifdef __cplusplus
}
#endif
#endif
