/*
 * fof_io.h -- host-side readers of CM_Halofinder's input formats and writers
 * of its halo catalogue, around the GPU linking engine of fof_gpu.h.
 *
 * Readers produce the reference's particle record with PARTICLE_ID and
 * MEMORY_MODE (src_v2/vars.h:36-47): 32 bytes {u64 ID @0, float Pos[3] @8,
 * float Vel[3] @20}, ready for fof_ghosts_host / fof_ghosts_device or the
 * linking calls.  Writers take the linking engine's halo table (and its
 * second moments and ring) and write the layouts of Output_Halos
 * (src_v2/FOF.c:546-913) and Output_Info (src_v2/FOF.c:917-957).
 *
 * All functions return 0 or a negative FOF_E_* code (FOF_E_ARG also for an
 * unreadable or malformed file, FOF_E_RANGE for a capacity too small).
 */
#ifndef CM_FOF_IO_H
#define CM_FOF_IO_H

#include <stdint.h>

#include "fof_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The 256-byte Gadget-1 header of a snapshot file (src_v2/vars.h:67-87). */
typedef struct fof_gadget_header_v1 {
  uint32_t npart[6];       /* npart[1]: particles of this file             */
  double mass[6];
  double time;
  double redshift;
  int32_t flag_sfr;
  int32_t flag_feedback;
  uint32_t npart_total[6];
  int32_t flag_cooling;
  int32_t num_files;
  double box_size;
  double omega0;
  double omega_lambda;
  double hubble_param;
  int32_t flag_stellarage;
  int32_t flag_metals;
  int32_t hashtabsize;
  char fill[84];
} fof_gadget_header_v1;

/* ASCII particles, one per line: "%llu %f %f %f %f %f %f" (ID, Pos, Vel)
 * with with_id, else "%f %f %f %f %f %f" and ID = the line's index
 * (src_v2/read_data.c:286-312 with MEMORY_MODE).  With out == NULL only the
 * particle count is returned in *n; otherwise up to cap records are written
 * and a count above cap is FOF_E_RANGE. */
int fof_read_ascii(const char *path, int with_id, void *out, uint64_t cap, uint64_t *n);

/* Gadget-1 snapshot (GADGET_STYLE, src_v2/read_data.c:321-365): Fortran
 * records of the header, npart[1] float positions, float velocities and,
 * with with_id, u64 IDs (else ID = index).  hdr may be NULL.  Record markers
 * are checked against the block sizes. */
int fof_read_gadget1(const char *path, int with_id, void *out, uint64_t cap, uint64_t *n,
                     fof_gadget_header_v1 *hdr);

/* PICOLA lightcone, unformatted (LIGHTCONE + UNFORMATTED,
 * src_v2/read_data.c:207-277): pairs of Fortran records, a u32 particle
 * count then that many particles of {u32 ID (with with_id), float Pos[3],
 * float Vel[3]}, until the end of the file.  The u32 ID is widened to u64
 * (the reference reads it into the low half of an uninitialised u64). */
int fof_read_picola(const char *path, int with_id, void *out, uint64_t cap, uint64_t *n);

/* Catalogue layouts (Output_Halos compile-time options of the reference). */
#define FOF_OUT_UNFORMATTED 0x1u /* Fortran-record binary, else ASCII          */
#define FOF_OUT_PARTICLES 0x2u   /* every member after its halo               */
#define FOF_OUT_INERTIA 0x4u     /* the inertia tensor (6 values)             */
#define FOF_OUT_DISPERSION 0x8u  /* the velocity dispersion tensor (6 values) */
#define FOF_OUT_PARTICLE_ID 0x10u /* member IDs with FOF_OUT_PARTICLES        */

/* One rank's halo file (src_v2/FOF.c:546-913): the owned halos of the table
 * (halos[k].owned != 0), in table order (ascending head = the reference's
 * ihalo[] order).  Per halo the reference's values: N, then
 * (float)(sum * (1.0 / N)) for Pos and Vel, then with FOF_OUT_INERTIA the
 * inertia tensor and with FOF_OUT_DISPERSION the dispersion tensor, both by
 * the reference's single-pass formulas from the raw moments (moments must
 * then be given).  With FOF_OUT_PARTICLES the members follow each halo in
 * ring order from the head: P (host records, stride/offsets as in
 * fof_particles_v1) and next (the 1-based ring of fof_link_gpu) are needed.
 * ASCII with FOF_OUT_DISPERSION but not FOF_OUT_INERTIA prints the
 * dispersion values; the reference prints array entries 12-17 there, beyond
 * its 12-value buffer (src_v2/FOF.c:708-710, :894-897). */
int fof_write_halos(const char *path, uint32_t flags, const fof_halo_v1 *halos,
                    const fof_moments_v1 *moments, uint64_t nhalos,
                    const fof_particles_v1 *P, const uint32_t *next);

/* The .info file (src_v2/FOF.c:917-957): one line per rank of an
 * nx x ny x nz decomposition of [lmin, lmax) (x-fastest), with its region,
 * halo count and the two subsample counts (NULL: zeros). */
int fof_write_info(const char *path, const double lmin[3], const double lmax[3], int nx, int ny,
                   int nz, const int32_t *nhalos, const uint32_t *ninhalo,
                   const uint32_t *nouthalo);

/* Subsampling (src_v2/FOF.c:172-284), on the host after the link: the
 * groups of a successor ring are visited as the reference's retrieval loop
 * visits them -- heads in ascending index order, a group of fewer than
 * nphalomin members "outside halos", any other "inside" -- and when the
 * matching rate is above zero the group's ring is walked from the head's
 * successor back to the head with one draw from the caller's generator per
 * particle visited.  A particle whose draw is below the rate is kept, unless
 * it lies outside the rank box: that ends the walk of its group (the
 * reference's `break`, :209-231 and :251-273).  The reference draws from
 * gsl_rng_uniform on a ranlxd1 generator seeded with Seed (:41-42); passing
 * that generator here gives the reference's lists for the ring given.
 * next is the 1-based ring of fof_link_gpu (next[i-1] == the reference's
 * next[i]) and P the host records; inhalo / outhalo receive the kept
 * particles' 1-based indices in visit order (each may be NULL when its rate
 * is not above zero).  FOF_E_RANGE when a list would pass its capacity: the
 * reference sizes them rint(1.5 n rate), at most n, and stops with a fatal
 * error once a count reaches that size (:234-238, :276-280).  FOF_E_ARG for
 * a next[] that is not a set of rings over 1..n. */
typedef double (*fof_uniform_fn)(void *state);

typedef struct fof_subsample_v1 {
  double samp_in;        /* SampInHalos  (src_v2/vars.h:106); <= 0: off     */
  double samp_out;       /* SampOutHalos (src_v2/vars.h:107); <= 0: off     */
  double rmin[3], rmax[3]; /* the rank box                                   */
  double nphalomin;      /* src_v2/vars.h:111                               */
  uint32_t closed_upper; /* bit a: Pos[a] == rmax[a] counts as inside: built
                            without PERIODIC, on the last rank along axis a
                            (src_v2/FOF.c:216-229, :258-271)                 */
  uint32_t reserved;
} fof_subsample_v1;

int fof_subsample(const fof_particles_v1 *P, const uint32_t *next, const fof_subsample_v1 *s,
                  fof_uniform_fn uniform, void *state, uint32_t *inhalo, uint64_t in_cap,
                  uint64_t *n_in, uint32_t *outhalo, uint64_t out_cap, uint64_t *n_out);

/* A subsample file (Subsample, src_v2/FOF.c:432-544; the reference names it
 * <OutputFileBase>_inhalo.<task> or _outhalo.<task>): the particles idx[0..n)
 * (1-based, as fof_subsample returns them) of the host records P.  ASCII: one
 * line per particle, the ID first with FOF_OUT_PARTICLE_ID.  With
 * FOF_OUT_UNFORMATTED, Fortran records: the count, all positions, all
 * velocities, then with FOF_OUT_PARTICLE_ID all IDs.  No file is created
 * when n is 0, as in the reference. */
int fof_write_subsample(const char *path, uint32_t flags, const fof_particles_v1 *P,
                        const uint32_t *idx, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif /* CM_FOF_IO_H */
