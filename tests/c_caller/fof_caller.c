/*
 * A C caller written the way INTEGRATION.md patches src_v2/FOF.c: the
 * reference's particle record and globals, fof_link_gpu() in place of
 * src_v2/FOF.c:44-161, then the reference's own retrieval loop
 * (src_v2/FOF.c:188-203: ring walks from ascending i, head marks, size cut)
 * and Checkhalo's centre of mass (src_v2/FOF.c:384-401) over the returned
 * next[1..n].  Input: a Gadget-1 file read with the library's reader.
 * Output, one line per owned halo: head (1-based, as ihalo[]), N, CoM.
 * Then the subsampling walk (src_v2/FOF.c:172-284) through fof_subsample
 * with the rates given and a 64-bit LCG in place of gsl_rng_uniform: two
 * lines, "S in <count> <indices>" and "S out <count> <indices>".
 *
 *   fof_caller <gadget1 file> <box> <pside> <linklength> <nphalomin>
 *              [<SampInHalos> <SampOutHalos> <seed>]
 */
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "fof_gpu.h"
#include "fof_io.h"

/* src_v2/vars.h:36-47 with PARTICLE_ID and MEMORY_MODE */
struct part_data {
  unsigned long long ID;
  float Pos[3];
  float Vel[3];
};

static struct part_data *P;
static unsigned int nparticles_tot;
static unsigned int *next;
static double rmin[3], rmax[3], rmin_buff[3], rmax_buff[3];
static double linklength, nphalomin;

/* the caller's generator: x <- a x + c (mod 2^64), u = top 53 bits / 2^53 */
static double lcg_uniform(void *state) {
  unsigned long long *x = (unsigned long long *)state;
  *x = *x * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(*x >> 11) * (1.0 / 9007199254740992.0);
}

static void FatalError(const char *file, int line) {
  printf("FatalError %s:%d\n", file, line);
  exit(1);
}

int main(int argc, char **argv) {
  if (argc < 6) return 2;
  const double box = atof(argv[2]);
  const int pside = atoi(argv[3]);
  linklength = atof(argv[4]);
  nphalomin = atof(argv[5]);
  uint64_t n = 0;
  if (fof_read_gadget1(argv[1], 1, NULL, 0, &n, NULL)) FatalError("fof_caller.c", __LINE__);
  P = (struct part_data *)malloc(n * sizeof(struct part_data));
  if (fof_read_gadget1(argv[1], 1, P, n, &n, NULL)) FatalError("fof_caller.c", __LINE__);
  nparticles_tot = (unsigned int)n;
  for (int k = 0; k < 3; k++) rmin[k] = rmin_buff[k] = 0.0, rmax[k] = rmax_buff[k] = box;

  /* ---- the INTEGRATION.md patch in place of src_v2/FOF.c:44-161 ---- */
  double lcell_[3] = {box / pside, box / pside, box / pside};
  fof_particles_v1 fp = {FOF_ABI_VERSION, FOF_MEM_HOST, P, nparticles_tot, sizeof(struct part_data),
                         (int32_t)offsetof(struct part_data, ID),
                         (uint32_t)offsetof(struct part_data, Pos),
                         (uint32_t)offsetof(struct part_data, Vel)};
  fof_domain_v1 fd;
  for (int k = 0; k < 3; k++) {
    fd.rmin[k] = rmin[k];
    fd.rmax[k] = rmax[k];
    fd.rmin_buff[k] = rmin_buff[k];
    fd.rmax_buff[k] = rmax_buff[k];
    fd.lcell[k] = lcell_[k];
  }
  fd.linklength = linklength;
  fd.nphalomin = nphalomin;
  fd.varlink = NULL;
  unsigned int *next0 = (unsigned int *)malloc(nparticles_tot * sizeof(unsigned int));
  int rc = fof_link_gpu(NULL, &fp, &fd, NULL, next0, NULL, NULL, 0, NULL, NULL);
  if (rc != 0) {
    printf("\nERROR: fof_link_gpu: %s\n", fof_strerror(rc));
    FatalError("FOF.c", 161);
  }
  next = next0 - 1; /* the reference indexes next[1..n] */

  /* ---- src_v2/FOF.c:180-203, unchanged ---- */
  unsigned int *head0 = (unsigned int *)calloc(nparticles_tot, sizeof(unsigned int));
  unsigned int *head = head0 - 1;
  unsigned int i, ip;
  int nphalo;
  for (i = 1; i <= nparticles_tot; i++) {
    if (head[i] == 0) {
      ip = i;
      nphalo = 0;
      do {
        ip = next[ip];
        nphalo++;
        if (ip == i) break;
        head[ip] = 1;
      } while (1);
      if (nphalo < nphalomin) head[ip] = 1;
    }
  }
  /* ---- Checkhalo (src_v2/FOF.c:384-401) for every remaining head ---- */
  for (i = 1; i <= nparticles_tot; i++) {
    if (head[i] != 0) continue;
    double xh = 0, yh = 0, zh = 0;
    int np = 0;
    unsigned int j = i;
    do {
      xh += (double)P[j - 1].Pos[0];
      yh += (double)P[j - 1].Pos[1];
      zh += (double)P[j - 1].Pos[2];
      np++;
      j = next[j];
    } while (j != i);
    xh /= (double)np, yh /= (double)np, zh /= (double)np;
    if (xh < rmin[0] || xh >= rmax[0] || yh < rmin[1] || yh >= rmax[1] || zh < rmin[2] ||
        zh >= rmax[2])
      continue;
    printf("%u %d %.17g %.17g %.17g\n", i, np, xh, yh, zh);
  }
  if (argc >= 9) {
    /* ---- subsampling, as INTEGRATION.md binds it ---- */
    fof_subsample_v1 s = {atof(argv[6]), atof(argv[7]), {rmin[0], rmin[1], rmin[2]},
                          {rmax[0], rmax[1], rmax[2]}, nphalomin, 0u, 0u};
    unsigned long long state = strtoull(argv[8], NULL, 10);
    unsigned int *inhalo = (unsigned int *)malloc(nparticles_tot * sizeof(unsigned int));
    unsigned int *outhalo = (unsigned int *)malloc(nparticles_tot * sizeof(unsigned int));
    uint64_t nin = 0, nout = 0;
    rc = fof_subsample(&fp, next + 1, &s, lcg_uniform, &state, inhalo, nparticles_tot, &nin, outhalo,
                       nparticles_tot, &nout);
    if (rc != 0) {
      printf("\nERROR: fof_subsample: %s\n", fof_strerror(rc));
      FatalError("FOF.c", 284);
    }
    printf("S in %llu", (unsigned long long)nin);
    for (uint64_t k = 0; k < nin; k++) printf(" %u", inhalo[k]);
    printf("\nS out %llu", (unsigned long long)nout);
    for (uint64_t k = 0; k < nout; k++) printf(" %u", outhalo[k]);
    printf("\n");
    free(inhalo);
    free(outhalo);
  }
  free(head0);
  free(next0);
  free(P);
  return 0;
}
