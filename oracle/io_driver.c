/*
 * io_driver.c -- drives the reference's own serial grid output
 * (dumpalloct_serial, io.c:1233-1363, with assign_grid_field, io.c:32-200) on
 * a tree described by flat arrays, so that tests can check the library's
 * emma_rt_dump_grid against it byte for byte.
 * TEST INFRASTRUCTURE ONLY (checker); oracle/Makefile builds it with the
 * reference's io.c, oct.c and friedmann.c into oracle/_ref/.
 *
 * Tree: levels lmin..lmax, noct[k] octs on level lmin+k, child[] the
 * levels' child tables concatenated (8 ints per oct: the child oct's index on
 * the next level, or < 0 for a leaf), opos[] the octs' lower corners (3
 * doubles per oct).  F[k*16 + j] points to level k's field j, ncell (or
 * ngrp*ncell, group-major) doubles, or is NULL: j = 0 e, 1 fx, 2 fy, 3 fz,
 * 4 src, 5 nhplus, 6 nh, 7 temp, 8 t_first_xion, 9 t_last_xion, 10 d, 11 u,
 * 12 v, 13 w, 14 p, 15 dX.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "prototypes.h"

void dumpalloct_serial(char folder[], REAL tsim, struct RUNPARAMS *param, struct CPUINFO *cpu);

/* the output field names of readOutputParam_grid (parameters.c:132-177), by field id */
static char *kFieldName[] = {
    "x", "y", "z", "l", "cpu", "gdata_d", "density", "gdata.p", "res", "f0", "f1", "f2", "marked",
    "field_d", "field_u", "field_v", "field_w", "field_p", "rfield_e0", "rfield_fx0", "rfield_fy0",
    "rfield_fz0", "rfield_e1", "rfield_fx1", "rfield_fy1", "rfield_fz1", "rfield_snfb", "rfield_e2",
    "rfield_fx2", "rfield_fy2", "rfield_fz2", "rfield_src", "xion", "field_dX", "field_dXHE", "field_dXXHE",
    "field_xHE", "field_xxHE", "rfield_temp", "t_first_xion", "t_last_xion", "rho_first_xion"};
#define NFIELD_TOT 42
#define NF 16

static double fv(const double *const *F, int k, int j, size_t i) {
  const double *p = F[k * NF + j];
  return p ? p[i] : 0.;
}

int ref_dump_grid(const char *folder, int ndump, int rank, double tsim, int lmin, int lmax, int ngrp,
                  const int *noct, const int *child, const double *opos, const double *const *F, int nfield,
                  const int *ids) {
  const int nl = lmax - lmin + 1;
  if (nl < 1 || ngrp != NGRP) return -1;
  struct OCT **octs = (struct OCT **)calloc(nl, sizeof(struct OCT *));
  for (int k = 0; k < nl; k++) octs[k] = (struct OCT *)calloc(noct[k] > 0 ? noct[k] : 1, sizeof(struct OCT));
  size_t oc = 0, op = 0;
  for (int k = 0; k < nl; k++) {
    const size_t n = (size_t)noct[k] * 8;
    for (int o = 0; o < noct[k]; o++) {
      struct OCT *oct = &octs[k][o];
      oct->x = opos[(op + o) * 3 + 0];
      oct->y = opos[(op + o) * 3 + 1];
      oct->z = opos[(op + o) * 3 + 2];
      oct->level = lmin + k;
      oct->cpu = rank;
      for (int c = 0; c < 8; c++) {
        struct CELL *cell = &oct->cell[c];
        const size_t i = (size_t)o * 8 + c;
        const int ch = child[oc + i];
        cell->idx = c;
        cell->child = (ch >= 0 && k + 1 < nl) ? &octs[k + 1][ch] : NULL;
        for (int g = 0; g < NGRP; g++) {
          cell->rfield.e[g] = fv(F, k, 0, g * n + i);
          cell->rfield.fx[g] = fv(F, k, 1, g * n + i);
          cell->rfield.fy[g] = fv(F, k, 2, g * n + i);
          cell->rfield.fz[g] = fv(F, k, 3, g * n + i);
          cell->rfield.src[g] = fv(F, k, 4, g * n + i);
        }
        cell->rfield.nhplus = fv(F, k, 5, i);
        cell->rfield.nh = fv(F, k, 6, i);
        cell->rfield.temp = fv(F, k, 7, i);
        cell->t_first_xion = fv(F, k, 8, i);
        cell->t_last_xion = fv(F, k, 9, i);
        cell->field.d = fv(F, k, 10, i);
        cell->field.u = fv(F, k, 11, i);
        cell->field.v = fv(F, k, 12, i);
        cell->field.w = fv(F, k, 13, i);
        cell->field.p = fv(F, k, 14, i);
        cell->field.dX = fv(F, k, 15, i);
      }
    }
    oc += (size_t)noct[k] * 8;
    op += noct[k];
  }
  struct CPUINFO cpu;
  memset(&cpu, 0, sizeof(cpu));
  cpu.rank = rank;
  cpu.ndumps = &ndump;
  cpu.locNoct = (int *)calloc(lmax, sizeof(int));
  cpu.octList = (struct OCT ***)calloc(lmax, sizeof(struct OCT **));
  for (int k = 0; k < nl; k++) {
    cpu.locNoct[lmin + k - 1] = noct[k];
    cpu.octList[lmin + k - 1] = (struct OCT **)calloc(noct[k] > 0 ? noct[k] : 1, sizeof(struct OCT *));
    for (int o = 0; o < noct[k]; o++) cpu.octList[lmin + k - 1][o] = &octs[k][o];
  }
  struct OUTPUTPARAM out;
  memset(&out, 0, sizeof(out));
  out.n_field_tot = NFIELD_TOT;
  for (int i = 0; i < NFIELD_TOT; i++) out.field_name[i] = kFieldName[i];
  for (int f = 0; f < nfield; f++) {
    if (ids[f] < 0 || ids[f] >= NFIELD_TOT) return -1;
    out.field_id[ids[f]] = 1;
  }
  for (int i = 0; i < NFIELD_TOT; i++) out.n_field += out.field_id[i];
  struct RUNPARAMS param;
  memset(&param, 0, sizeof(param));
  param.lcoarse = lmin;
  param.lmax = lmax;
  param.out_grid = &out;
  char dir[512];
  snprintf(dir, sizeof(dir), "%s", folder);
  dumpalloct_serial(dir, (REAL)tsim, &param, &cpu);
  for (int k = 0; k < nl; k++) {
    free(cpu.octList[lmin + k - 1]);
    free(octs[k]);
  }
  free(cpu.octList);
  free(cpu.locNoct);
  free(octs);
  return 0;
}
