/*
 * fof_gpu.h -- C ABI of the MI355X Friends-of-Friends linking engine.
 *
 * Drop-in for the linking routine of CM_Halofinder.  The reference's FOF()
 * (src_v2/FOF.c:23-339, identical in src/FOF.c) builds a cell grid and
 * linked cell lists (src_v2/FOF.c:44-92), sweeps the 13 forward neighbour
 * cells (src_v2/FOF.c:94-161) and merges ring-linked lists in Befriend
 * (src_v2/FOF.c:341-376).  Its only product is the successor ring next[1..n]
 * that the rest of FOF() walks (retrieval src_v2/FOF.c:188-203, Checkhalo
 * src_v2/FOF.c:378-430, Output_Halos src_v2/FOF.c:797-847).  fof_link_gpu()
 * replaces src_v2/FOF.c:44-161 and returns that ring (plus labels and the
 * halo table those later loops compute), so the host code after line 163
 * keeps working unchanged.
 *
 * All entry points are extern "C", take plain pointers and sizes, return 0
 * on success and a negative FOF_E_* code on failure.  A context may be used
 * by one host thread at a time; separate contexts are independent.
 */
#ifndef CM_FOF_GPU_H
#define CM_FOF_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FOF_ABI_VERSION 1u

/* Error codes.  The host maps any non-zero return to
 * FatalError("FOF.c", line) exactly as src_v2/main.c:107-112 does. */
#define FOF_OK 0
#define FOF_E_ARG (-1)    /* bad pointer, version, stride or parameter        */
#define FOF_E_DOMAIN (-2) /* particle outside [rmin_buff,rmax_buff] or non-finite:
                             the reference would index first[] out of bounds
                             (src_v2/FOF.c:79-85)                              */
#define FOF_E_RANGE (-3)  /* n >= 2^31-1 particles in one call (labels carry a
                             flag in bit 31; the reference's u32 indices,
                             src_v2/FOF.c:29, allow 2^32-2), a grid too large
                             for 64-bit cell keys, or an output capacity too
                             small                                             */
#define FOF_E_NOMEM (-4)  /* device allocation failed                          */
#define FOF_E_HIP (-5)    /* HIP runtime error                                 */
#define FOF_E_NODEV (-6)  /* no usable gfx950 device                           */

/* Where a buffer lives. */
#define FOF_MEM_HOST 0u
#define FOF_MEM_DEVICE 1u

/* Particle array: the reference's `struct part_data *P` (src_v2/vars.h:36-47),
 * AoS with an optional u64 ID (PARTICLE_ID) followed by float Pos[3], Vel[3]
 * (MEMORY_MODE).  With PARTICLE_ID + MEMORY_MODE the record is 32 bytes:
 * ID @0, Pos @8, Vel @20.  Without PARTICLE_ID it is 24 bytes: Pos @0,
 * Vel @12.  Only float (MEMORY_MODE) records are accepted. */
typedef struct fof_particles_v1 {
  uint32_t version;  /* FOF_ABI_VERSION                                      */
  uint32_t location; /* FOF_MEM_HOST or FOF_MEM_DEVICE                       */
  const void *base;  /* P                                                    */
  uint64_t n;        /* nparticles_tot (src_v2/vars.h:58)                    */
  uint32_t stride;   /* sizeof(struct part_data), multiple of 4              */
  int32_t off_id;    /* byte offset of ID, or -1 without PARTICLE_ID         */
  uint32_t off_pos;  /* byte offset of Pos[0]                                */
  uint32_t off_vel;  /* byte offset of Vel[0]                                */
} fof_particles_v1;

/* Variable linking length (the intent of VARLINK, src_v2/FOF.c:55-61,
 * :102-109 and src_v2/link.c:17-95, which is dead code in the reference).
 * The squared linking length of a pair is the value of its home cell: the
 * reference-grid cell with the lower linear index of the two particles'
 * cells (the cell being swept, since the forward stencil only visits higher
 * indices).  A cell's value is the piecewise-linear interpolation, in the
 * squared distance r2 of the cell centre to `origin`, of l2tab[0..ntab-1]
 * sampled uniformly on r2 in [0, r2max]; beyond r2max the last entry holds.
 * Cell centre = rmin_buff + (i + 0.5) * lcell on every axis. */
typedef struct fof_varlink_v1 {
  double origin[3];
  double r2max;
  uint32_t ntab; /* >= 2 */
  uint32_t location; /* where l2tab lives                                    */
  const double *l2tab;
} fof_varlink_v1;

/* One rank's domain (src_v2/startup.c:33-47) plus the FoF parameters. */
typedef struct fof_domain_v1 {
  double rmin[3], rmax[3];           /* owned box; Checkhalo keeps halos whose
                                        CoM lies in [rmin,rmax)              */
  double rmin_buff[3], rmax_buff[3]; /* owned box + ghost shell              */
  double lcell[3];      /* reference cell size, src_v2/FOF.c:45-53           */
  double linklength;    /* linksq = linklength^2, src_v2/FOF.c:60            */
  double nphalomin;     /* size cut, src_v2/FOF.c:202                        */
  const fof_varlink_v1 *varlink; /* NULL: constant linking length           */
} fof_domain_v1;

/* Options. */
#define FOF_OPT_TIMING 0x1u     /* record per-phase device times in stats     */
#define FOF_OPT_OUT_DEVICE 0x2u /* next_out / label_out / halos are device
                                   pointers                                   */
#define FOF_OPT_NO_RING 0x4u    /* do not write next_out                      */
#define FOF_OPT_NO_HALOS 0x8u   /* skip the halo table                        */
#define FOF_OPT_GRAPH 0x10u    /* capture the call's launch sequence in a HIP
                                   graph and replay it on later calls with
                                   identical arguments (device-resident input
                                   and outputs, constant linking length, dense
                                   cell addressing; other calls run eagerly)  */
#define FOF_OPT_SPARSE 0x20u    /* address only occupied cells (sorted keys +
                                   hash table); chosen automatically when a
                                   dense cell table would be much larger than
                                   the particle count or reach 2^31 cells     */
#define FOF_OPT_AUTOTUNE 0x40u  /* pick the dense-cell threshold and the number
                                   of concurrent pipelines for this workload
                                   shape by timing candidates on the call's own
                                   input (once per shape; results are cached
                                   per context and in the file named by the
                                   environment variable FOF_TUNE_CACHE)       */
#define FOF_OPT_WRITE_PAIRS 0x80u   /* write labels and the ring through a
                                       partition of the entries by index (the
                                       default for a volume of more than 2^25
                                       particles, whose random writes would
                                       spread beyond the last-level cache)    */
#define FOF_OPT_WRITE_DIRECT 0x100u /* write them straight by index (the
                                       default up to 2^25 particles a volume);
                                       with both flags, WRITE_PAIRS holds      */

typedef struct fof_opts_v1 {
  uint32_t version;
  uint32_t flags;
  double band;              /* relative FP32 band around l^2 that is re-decided
                               in FP64; 0 selects 1e-5; must lie in [1e-6,0.1] */
  uint32_t dense_threshold; /* cells with at least this many particles are
                               linked at sub-cell level; 0 = default (24),
                               values above 512 act as 512                   */
  uint32_t pipelines;       /* a batch is split over this many concurrent
                               pipelines (volumes in contiguous groups of
                               about equal particle counts, each on its own
                               stream and workspace); 0 or 1: one, at most 8 */
  uint32_t big_grid;        /* workgroups of the 256-thread dense-cell kernels;
                               0 = default (8192)                           */
  uint32_t reserved;
} fof_opts_v1;

/* One halo: a group whose size passes the cut (src_v2/FOF.c:202).  `head` is
 * the group's minimum local index, 0-based (the reference's ihalo[] holds
 * head + 1).  Sums are FP64 accumulations of the float Pos / Vel of the
 * members in ascending index order starting at the head -- the order in which
 * the host walks the ring returned in next_out -- so the host's own Checkhalo
 * and Output_Halos loops reproduce them bit for bit. */
typedef struct fof_halo_v1 {
  uint32_t head;
  uint32_t n;
  uint32_t owned;     /* 1 if CoM = sum_pos / n lies in [rmin, rmax)
                         (src_v2/FOF.c:394-401)                              */
  uint32_t edge_warn; /* bit d set if a member lies within linklength of the
                         buffered box in axis d (src_v2/FOF.c:403-422), only
                         evaluated for owned halos                          */
  double sum_pos[3];
  double sum_vel[3];
} fof_halo_v1;

/* Counters and timings.  With FOF_OPT_TIMING, phase_ms[k] is the device time
 * (HIP events on the library's stream) of span k of the last call, 0 when the
 * span did not run:
 *   0 cell keys                     1 bucket offsets (scan) or key sort (sparse)
 *   2 partitions + bucket sort      3 union-find init
 *   4 sub-cell sort of dense cells  5 pair tests, sparse cells
 *     of 256 particles or more
 *   6 pair tests, dense cells of    7 union-find flatten + group minima
 *     256 particles or more
 *   8 labels, member list filled    9 label defaults + group list
 *  10 members sorted in each group 11 ring
 *  12 halo list                    13 halo sums
 *  14 output copies                15 pair tests, dense cells below 256
 *  16 total, first span to last                                             */
#define FOF_NPHASE 17
typedef struct fof_stats_v1 {
  uint32_t version;
  uint32_t nvol;
  uint64_t n;             /* particles linked                                */
  uint64_t ncells;        /* cells in the grids used (occupied cells when
                             sparse)                                         */
  uint64_t ndense;        /* cells linked at sub-cell level                  */
  uint64_t pair_tests;    /* particle-pair distance tests                    */
  uint64_t rechecks;      /* tests re-decided in FP64                        */
  uint64_t ngroups;       /* groups of any size                              */
  uint64_t nmembers;      /* particles in groups of two or more (the ring)  */
  uint64_t nhalos;        /* groups passing the size cut                     */
  uint64_t nhalos_owned;  /* of those, owned (Checkhalo)                     */
  uint32_t sparse;        /* 1 if occupied-cell addressing was used          */
  uint32_t dense_threshold; /* the dense-cell threshold in effect            */
  double phase_ms[FOF_NPHASE];
  double link_kernel_ms;  /* spans 5 + 6 + 15: the pair-test kernels         */
  uint64_t pair_tests_kernel[3]; /* pair tests by kernel: sparse cells, dense
                                    cells of 256 particles or more, smaller
                                    dense cells (sum = pair_tests)            */
  uint64_t pairs_joined;  /* candidate run pairs (cells of 256 particles or
                             more) not tested because both ends were already
                             in one set                                      */
  uint32_t pipelines;     /* concurrent pipelines the volumes were split over */
  uint32_t graph;         /* 1: replayed a captured graph, 2: captured one    */
  uint32_t big_grid;      /* workgroups of the 256-thread dense kernels       */
  uint32_t reserved;
} fof_stats_v1;

typedef struct fof_ctx fof_ctx;

/* Context: owns device workspace that is reused across calls. */
int fof_ctx_create(int device, fof_ctx **out);
void fof_ctx_destroy(fof_ctx *ctx);

/* Link one domain.  Replaces src_v2/FOF.c:44-161.
 *  next_out[n]  : the successor ring, 1-based like the reference's next[1..n]
 *                 (next_out[i-1] == next[i]); each group's ring visits its
 *                 members in ascending index order.  May be NULL.
 *  label_out[n] : 0-based minimum index of each particle's group.  May be NULL.
 *  halos        : groups passing the size cut in ascending head order
 *                 (the order of ihalo[]); capacity halo_cap.  May be NULL.
 *  nhalos       : number of halos written (or needed, with FOF_E_RANGE).
 *  stats        : may be NULL.
 * ctx may be NULL: a per-device default context is used. */
int fof_link_gpu(fof_ctx *ctx, const fof_particles_v1 *particles,
                 const fof_domain_v1 *domain, const fof_opts_v1 *opts,
                 uint32_t *next_out, uint32_t *label_out, fof_halo_v1 *halos,
                 uint64_t halo_cap, uint64_t *nhalos, fof_stats_v1 *stats);

/* Batched linking: many independent domains (subvolumes, snapshots,
 * lightcone shells) stored back to back in one particle array, linked by
 * one sequence of launches.  Volume v owns particles
 * [vols[v].offset, vols[v].offset + vols[v].n).  next_out / label_out are
 * indexed like the particle array and hold volume-local values.  Halos are
 * written volume after volume; halo_offsets[nvol+1] receives the prefix. */
typedef struct fof_volume_v1 {
  uint64_t offset;
  uint64_t n;
  fof_domain_v1 domain;
} fof_volume_v1;

int fof_link_gpu_batch(fof_ctx *ctx, const fof_particles_v1 *particles,
                       const fof_volume_v1 *vols, uint32_t nvol,
                       const fof_opts_v1 *opts, uint32_t *next_out,
                       uint32_t *label_out, fof_halo_v1 *halos,
                       uint64_t halo_cap, uint64_t *halo_offsets,
                       fof_stats_v1 *stats);

/* Second moments of a halo (INERTIA / DISPERSION, src_v2/FOF.c:848-872).
 * pos2 / vel2: FP64 sums over the members, in the ring order of fof_halo_v1's
 * sums, of the products {xx, yy, zz, xy, xz, yz} of the float Pos / Vel
 * widened to double: the reference's accumulators xh2, yh2, zh2, xhyh, xhzh,
 * yhzh (and vxh2 ... vyhvzh) bit for bit, from which its single-pass
 * formulas follow.  cpos2 / cvel2: the same sums taken about the mean,
 * sum (a - <a>)(b - <b>), accumulated about the head particle and recentred,
 * so without the cancellation of the single-pass form.  The reference's
 * inertia tensor is {cyy + czz, cxx + czz, cxx + cyy, -cxy, -cxz, -cyz} and
 * its dispersion tensor cvel2 / n. */
typedef struct fof_moments_v1 {
  double pos2[6], vel2[6];
  double cpos2[6], cvel2[6];
} fof_moments_v1;

/* Outputs of fof_link_gpu_ex.  Every array lives where `location` says;
 * NULL skips an output. */
typedef struct fof_outputs_v1 {
  uint32_t version;        /* FOF_ABI_VERSION                                 */
  uint32_t location;       /* FOF_MEM_HOST or FOF_MEM_DEVICE                  */
  uint32_t *next;          /* [n] successor ring, as next_out above          */
  uint32_t *label;         /* [n] group minimum index (volume-local)         */
  uint64_t *minid;         /* [n] minimum particle ID of the group: the
                              canonical label, independent of input order
                              (needs off_id >= 0)                            */
  fof_halo_v1 *halos;      /* [halo_cap] halos, volume after volume          */
  fof_moments_v1 *moments; /* [halo_cap] second moments of the same halos    */
  uint64_t halo_cap;
  uint64_t *halo_offsets;  /* host [nvol + 1]: first halo of each volume     */
  uint64_t nhalos;         /* out: halos found (also with FOF_E_RANGE, when
                              halo_cap was too small)                        */
  fof_stats_v1 *stats;     /* host, may be NULL                              */
} fof_outputs_v1;

/* Link a batch of volumes (one volume: nvol = 1) with every output the
 * engine computes.  fof_link_gpu and fof_link_gpu_batch are this call with
 * FOF_OPT_OUT_DEVICE selecting the outputs' location. */
int fof_link_gpu_ex(fof_ctx *ctx, const fof_particles_v1 *particles,
                    const fof_volume_v1 *vols, uint32_t nvol,
                    const fof_opts_v1 *opts, fof_outputs_v1 *out);

/* fof_link_gpu_ex for one domain without allocating device memory.  This
 * entry point has no counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The call's workspace is a fixed array
 * of FOF_NOALLOC_BYTES in the library, one per device, held from the time
 * the library's code is loaded on the device until the process ends.
 * Particles and every output must be device-resident; constant linking
 * length only; timing, graphs, autotuning and pipelines are not used.  A
 * domain whose workspace (fof_plan's workspace_bytes, plus 8 bytes per
 * particle for minid) exceeds the array, or that needs occupied-cell
 * addressing, returns FOF_E_RANGE.  Every call on a device shares that
 * device's array: calls that overlap on one device race. */
#define FOF_NOALLOC_BYTES (256ull << 20)
int fof_link_gpu_noalloc(fof_ctx *ctx, const fof_particles_v1 *particles,
                         const fof_domain_v1 *domain, const fof_opts_v1 *opts,
                         fof_outputs_v1 *out);

/* What a call would do, computed on the host (no device needed): the cell
 * addressing chosen, the grid sizes and the device workspace it needs. */
typedef struct fof_plan_v1 {
  uint32_t version;
  uint32_t sparse;          /* 1: occupied-cell addressing                   */
  uint64_t n;               /* particles                                     */
  uint64_t cells;           /* dense cell table entries (0 when sparse)      */
  double grid_cells;        /* cells of the grids linked on, all volumes: the
                               reference grid, or a finer one on the axes
                               where a constant l is below lcell (1 - 1e-6)
                               (the same linked pairs; see fof_link_gpu)    */
  uint64_t workspace_bytes; /* device workspace of the call                  */
  uint32_t dense_threshold; /* in effect                                     */
  uint32_t reserved;
} fof_plan_v1;

int fof_plan(const fof_particles_v1 *particles, const fof_volume_v1 *vols,
             uint32_t nvol, const fof_opts_v1 *opts, fof_plan_v1 *out);

/* Text for an error code. */
const char *fof_strerror(int code);

/* Library / device information: writes a NUL-terminated description. */
int fof_device_info(int device, char *buf, size_t len);

/* ---- Workload tools (synthetic input, ghost shells).  Not part of the
 * reference interface: they build the inputs the reference's Read_Data
 * (src_v2/read_data.c:5-856) would hand to FOF(). ---- */

/* Clump + background generator.  The n particles are written as 32-byte
 * records {u64 ID, float Pos[3], float Vel[3]} in a shuffled order (ID is
 * the generation index).  A fraction clump_frac lies in nclumps Plummer
 * spheres whose sizes follow dN/dM ~ M^-alpha on [mmin,mmax] (quantile
 * sampled and rescaled to the clumped count), scale radius
 * a = a0 * (size/100)^(1/3); the rest is uniform.  Positions are periodic in
 * [0, L)^3; velocities are Gaussian with dispersion sigma_v per axis. */
typedef struct fof_gen_v1 {
  uint64_t seed;
  uint64_t n;
  double box;
  uint64_t nclumps;
  double clump_frac;
  double mmin, mmax, alpha;
  double a0;
  double sigma_v;
  double umax;        /* Plummer mass fraction at which radii are truncated  */
  double pair_frac;   /* fraction of particles re-placed as borderline pairs */
  double pair_rel;    /* |d^2/l^2 - 1| bound for borderline pairs            */
  double pair_l;      /* linking length the borderline pairs straddle        */
} fof_gen_v1;

/* Host implementation (any machine).  out: n records. */
int fof_gen_host(const fof_gen_v1 *g, void *out);
/* Device implementation: out is a device pointer to n records. */
int fof_gen_device(const fof_gen_v1 *g, void *out);

/* Periodic ghost-shell builder: cut the periodic box [0,box)^3 into
 * gx*gy*gz subvolumes and give each its owned particles plus every periodic
 * image within `shell` of it (the shifted-copy semantics of
 * src_v2/read_data.c:418-522 and :785-799, image = (float)(x +/- box)).
 * Records are {u64 ID, float Pos[3], float Vel[3]}.  With out == NULL only
 * the per-volume counts are computed.  Volume order is x-fastest
 * (src_v2/startup.c:33-35); within a volume, particles keep input order,
 * owned copies and images interleaved.  counts[g] receives each volume's
 * size; vols (may be NULL) receives offsets and domains for fof_link_gpu_batch
 * with the given linklength, nphalomin and reference grid particles/side. */
typedef struct fof_ghost_v1 {
  double box;
  double shell;
  uint32_t gx, gy, gz;
  uint32_t pside;   /* particles per side of the whole box (Px=Py=Pz) used
                       for the reference cell size lcell = box/pside         */
  double linklength;
  double nphalomin;
} fof_ghost_v1;

int fof_ghosts_host(const fof_ghost_v1 *g, const void *in, uint64_t n,
                    void *out, uint64_t out_cap, uint64_t *counts,
                    fof_volume_v1 *vols);
int fof_ghosts_device(const fof_ghost_v1 *g, const void *in, uint64_t n,
                      void *out, uint64_t out_cap, uint64_t *counts,
                      fof_volume_v1 *vols);

#ifdef __cplusplus
}
#endif
#endif /* CM_FOF_GPU_H */
