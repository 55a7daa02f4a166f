/*
 * ref_driver.c -- links the reference's own CPU functions (rad_utils.c,
 * chem_utils.c, oct.c compiled from /root/reference/src by oracle/Makefile)
 * and exposes them through a flat C interface so that tests can pin the
 * oracle restatement (emma_oracle.c) and write golden vectors.
 * TEST INFRASTRUCTURE ONLY (checker); outputs go to oracle/_ref/.
 *
 * The driver builds a real struct OCT tree (prototypes.h:1028-1050) from the
 * flat description shared with the oracle:
 *   per level l = lmin..lmax: noct, parent cell (level l-1 cell index),
 *   face-neighbour cells (level l-1 cell indices, 6 per oct),
 *   child oct (level l+1 oct index or -1, 8 per oct).
 * The level lmin-1 is a structural layer of fully split cells; structural
 * cell s is the parent of level-lmin oct s.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stddef.h>
#include <math.h>

#include "prototypes.h"
#include "oct.h"
#include "rad_utils.h"
#include "chem_utils.h"

int rad_sweepX(struct RGRID *stencil, int level, int curcpu, int nread, int stride, REAL dx, REAL dt, REAL c);
int rad_sweepY(struct RGRID *stencil, int level, int curcpu, int nread, int stride, REAL dx, REAL dt, REAL c);
int rad_sweepZ(struct RGRID *stencil, int level, int curcpu, int nread, int stride, REAL dx, REAL dt, REAL c);
void updatefieldrad(struct OCT *octstart, struct RGRID *stencil, int nread, int stride, struct CPUINFO *cpu,
                    REAL dxcur, REAL dtnew, REAL cloc);
REAL Eddington(REAL fx, REAL fy, REAL fz, REAL ee, REAL c, int i, int j);
int advancerad(struct OCT **firstoct, int level, struct CPUINFO *cpu, struct RGRID *stencil, int stride,
               REAL dxcur, REAL dtnew, REAL aexp, struct RUNPARAMS *param, int chemonly);

REAL cucompute_alpha_b(REAL temp, REAL unit_number, REAL aexp);
REAL cucompute_alpha_a(REAL temp, REAL unit_number, REAL aexp);
REAL cucompute_beta(REAL temp, REAL unit_number, REAL aexp);
void cuCompCooling(REAL temp, REAL x, REAL nH, REAL *lambda, REAL *tcool, REAL aexp, REAL CLUMPF, int src);

#ifdef COARSERAD
int putE2coarse(struct CELL *cell, struct RUNPARAMS *param, int level, REAL aexp, REAL tcur);
#endif
#if defined(WRADHYD) || defined(STARS)
int FillRad(int level, struct RUNPARAMS *param, struct OCT **firstoct, struct CPUINFO *cpu, int init, REAL aexp,
            REAL tloc);
#endif

#define MAXLEV 24

static struct RUNPARAMS g_param;
static struct COSMOPARAM g_cosmo;
static struct CPUINFO g_cpu;
static int g_noct[MAXLEV + 1];
static struct OCT *g_lev[MAXLEV + 1];   /* g_lev[l] = octs of level l */
static struct OCT *g_struct = NULL;      /* level lmin-1 structural layer */
static struct OCT *g_firstoct[MAXLEV + 1];
static int g_lmin = 0, g_lmax = -1;
static REAL g_hnu[NGRP], g_ae[NGRP], g_ai[NGRP], g_fg[NGRP];
static REAL g_time_bound[NGRP_TIME + 1];

/* rad_utils.c times its phases with MPI_Wtime, which MPICH only serves after
   MPI_Init: initialise a singleton MPI process once */
static void ensure_mpi(void) {
  int ok = 0;
  MPI_Initialized(&ok);
  if (!ok) MPI_Init(NULL, NULL);
}

int refdrv_ngrp(void) { return NGRP; }

/* 1 when built with COARSERAD (transport on the coarse level only) */
int refdrv_coarserad(void) {
#ifdef COARSERAD
  return 1;
#else
  return 0;
#endif
}
int refdrv_real_bytes(void) { return (int)sizeof(REAL); }

/* layout table of the reference structs for the library's layout self-check */
void refdrv_layout(long *out) {
  out[0] = sizeof(struct Rtype);
  out[1] = offsetof(struct Rtype, e);
  out[2] = offsetof(struct Rtype, fx);
  out[3] = offsetof(struct Rtype, fy);
  out[4] = offsetof(struct Rtype, fz);
  out[5] = offsetof(struct Rtype, src);
  out[6] = offsetof(struct Rtype, nhplus);
  out[7] = offsetof(struct Rtype, eint);
  out[8] = offsetof(struct Rtype, nh);
  out[9] = offsetof(struct Rtype, temp);
  out[10] = sizeof(struct CELLLIGHT_R);
  out[11] = offsetof(struct CELLLIGHT_R, split);
  out[12] = sizeof(struct CELLFLUX_R);
  out[13] = offsetof(struct CELLFLUX_R, deltaR);
  out[14] = offsetof(struct CELLFLUX_R, rfield);
  out[15] = offsetof(struct CELLFLUX_R, rfieldnew);
  out[16] = offsetof(struct CELLFLUX_R, rflux);
  out[17] = sizeof(struct OCTLIGHT_R);
  out[18] = sizeof(struct OCTFLUX_R);
  out[19] = sizeof(struct RGRID);
  out[20] = offsetof(struct RGRID, oct);
  out[21] = offsetof(struct RGRID, New);
}

/* time_bound: NGRP_TIME + 1 bounds in years as readAtomic leaves them
   (parameters.c:455-468, the last one INFINITY) */
void refdrv_set_params(const double *units7, double clight, double fudgecool, int ncvgcool, const double *hnu,
                       const double *alphae, const double *alphai, const double *factgrp, double H0, double om,
                       double ov, double tphy, const double *time_bound) {
  memset(&g_param, 0, sizeof(g_param));
  memset(&g_cosmo, 0, sizeof(g_cosmo));
  g_param.unit.unit_l = units7[0];
  g_param.unit.unit_v = units7[1];
  g_param.unit.unit_t = units7[2];
  g_param.unit.unit_n = units7[3];
  g_param.unit.unit_mass = units7[4];
  g_param.unit.unit_d = units7[5];
  g_param.unit.unit_N = units7[6];
  g_param.clight = clight;
  g_param.clightorg = clight;
  g_param.fudgecool = fudgecool;
  g_param.ncvgcool = ncvgcool;
  for (int g = 0; g < NGRP; g++) {
    g_hnu[g] = hnu[g];
    g_ae[g] = alphae[g];
    g_ai[g] = alphai[g];
    g_fg[g] = factgrp[g];
  }
  g_param.atomic.n = NGRP;
  g_param.atomic.ngrp_space = NGRP_SPACE;
  g_param.atomic.ngrp_time = NGRP_TIME;
  for (int i = 0; i <= NGRP_TIME; i++) g_time_bound[i] = time_bound[i];
  g_param.atomic.time_bound = g_time_bound;
  g_param.atomic.hnu = g_hnu;
  g_param.atomic.alphae = g_ae;
  g_param.atomic.alphai = g_ai;
  g_param.atomic.factgrp = g_fg;
  g_cosmo.H0 = H0;
  g_cosmo.om = om;
  g_cosmo.ov = ov;
  g_cosmo.ob = om;
  g_cosmo.aexp = 1.;
  g_cosmo.tphy = tphy;
  g_param.cosmo = &g_cosmo;
  g_param.lcoarse = g_lmin;
  g_param.lmax = g_lmax;
}

void refdrv_set_tphy(double tphy) { g_cosmo.tphy = tphy; }

static void free_tree(void) {
  for (int l = 0; l <= MAXLEV; l++) {
    free(g_lev[l]);
    g_lev[l] = NULL;
    g_noct[l] = 0;
    g_firstoct[l] = NULL;
  }
  free(g_struct);
  g_struct = NULL;
  free(g_cpu.noct);
  g_cpu.noct = NULL;
}

static struct CELL *cell_at(int level, int idx) {
  if (level == g_lmin - 1) return &g_struct[idx / 8].cell[idx % 8];
  return &g_lev[level][idx / 8].cell[idx % 8];
}

/* flat arrays are the levels' arrays concatenated in level order */
int refdrv_build(int lmin, int lmax, const int *noct, const int *parent, const int *neicell, const int *child) {
  if (lmin < 1 || lmax < lmin || lmax > MAXLEV) return -1;
  ensure_mpi();
  free_tree();
  g_lmin = lmin;
  g_lmax = lmax;
  int nstruct = (noct[0] + 7) / 8;
  g_struct = (struct OCT *)calloc(nstruct, sizeof(struct OCT));
  for (int l = lmin; l <= lmax; l++) {
    g_noct[l] = noct[l - lmin];
    g_lev[l] = (struct OCT *)calloc(g_noct[l] > 0 ? g_noct[l] : 1, sizeof(struct OCT));
  }
  for (int j = 0; j < nstruct; j++) {
    g_struct[j].level = lmin - 1;
    for (int k = 0; k < 8; k++) {
      int s = j * 8 + k;
      g_struct[j].cell[k].idx = k;
      g_struct[j].cell[k].child = (s < g_noct[lmin]) ? &g_lev[lmin][s] : NULL;
    }
  }
  size_t op = 0, on = 0, oc = 0;
  for (int l = lmin; l <= lmax; l++) {
    struct OCT *L = g_lev[l];
    for (int i = 0; i < g_noct[l]; i++) {
      struct OCT *o = &L[i];
      o->level = l;
      o->cpu = 0;
      o->parent = (l == lmin) ? cell_at(l - 1, i) : cell_at(l - 1, parent[op + i]);
      for (int f = 0; f < 6; f++) o->nei[f] = cell_at(l - 1, neicell[on + (size_t)i * 6 + f]);
      for (int c = 0; c < 8; c++) {
        int ch = child[oc + (size_t)i * 8 + c];
        o->cell[c].idx = c;
        o->cell[c].child = (ch >= 0) ? &g_lev[l + 1][ch] : NULL;
        o->cell[c].t_first_xion = -1;
        o->cell[c].t_last_xion = -1;
      }
      o->next = (i + 1 < g_noct[l]) ? &L[i + 1] : NULL;
      o->prev = (i > 0) ? &L[i - 1] : NULL;
    }
    g_firstoct[l - 1] = g_noct[l] > 0 ? &L[0] : NULL;
    op += g_noct[l];
    on += (size_t)g_noct[l] * 6;
    oc += (size_t)g_noct[l] * 8;
  }
  memset(&g_cpu, 0, sizeof(g_cpu));
  g_cpu.rank = 0;
  g_cpu.nproc = 1;
  g_cpu.noct = (int *)calloc(MAXLEV + 1, sizeof(int));
  for (int l = lmin; l <= lmax; l++) g_cpu.noct[l - 1] = g_noct[l];
  g_cpu.levelcoarse = lmin;
  g_param.lcoarse = lmin;
  g_param.lmax = lmax;
  return 0;
}

/* which: 0 = rfield, 1 = rfieldnew. Arrays are SoA, group-major. */
static struct Rtype *rt_of(int level, int i, int which) {
  struct CELL *c = cell_at(level, i);
  return which ? &c->rfieldnew : &c->rfield;
}

int refdrv_set_level(int level, int which, const double *e, const double *fx, const double *fy, const double *fz,
                     const double *src, const double *nhplus, const double *eint, const double *nh,
                     const double *temp) {
  if (level < g_lmin || level > g_lmax) return -1;
  int n = g_noct[level] * 8;
  for (int i = 0; i < n; i++) {
    struct Rtype *R = rt_of(level, i, which);
    for (int g = 0; g < NGRP; g++) {
      R->e[g] = e[g * n + i];
      R->fx[g] = fx[g * n + i];
      R->fy[g] = fy[g * n + i];
      R->fz[g] = fz[g * n + i];
      R->src[g] = src[g * n + i];
    }
    R->nhplus = nhplus[i];
    R->eint = eint[i];
    R->nh = nh[i];
    R->temp = temp[i];
  }
  return 0;
}

int refdrv_get_level(int level, int which, double *e, double *fx, double *fy, double *fz, double *src,
                     double *nhplus, double *eint, double *nh, double *temp) {
  if (level < g_lmin || level > g_lmax) return -1;
  int n = g_noct[level] * 8;
  for (int i = 0; i < n; i++) {
    struct Rtype *R = rt_of(level, i, which);
    for (int g = 0; g < NGRP; g++) {
      e[g * n + i] = R->e[g];
      fx[g * n + i] = R->fx[g];
      fy[g * n + i] = R->fy[g];
      fz[g * n + i] = R->fz[g];
      src[g * n + i] = R->src[g];
    }
    nhplus[i] = R->nhplus;
    eint[i] = R->eint;
    nh[i] = R->nh;
    temp[i] = R->temp;
  }
  return 0;
}

int refdrv_get_txion(int level, double *t_first, double *t_last) {
  if (level < g_lmin || level > g_lmax) return -1;
  int n = g_noct[level] * 8;
  for (int i = 0; i < n; i++) {
    struct CELL *c = cell_at(level, i);
    t_first[i] = c->t_first_xion;
    t_last[i] = c->t_last_xion;
  }
  return 0;
}

void refdrv_clean_new_rad(int level, double aexp) { clean_new_rad(level, &g_param, g_firstoct, &g_cpu, aexp); }

#ifdef GPUAXL
/* ---- GPUAXL build: RadSolver's GPU branch (advanceradGPU) and the device
   set-up of emma.c come from the drop-in file emma_amd/dropin/rad_utils_hip.c
   over the HIP library ---- */
#include "rad_utils_gpu.h"
#include "interface.h"

static struct RGRID *g_rstencil = NULL;
static int g_rstride = 0;

int refdrv_gpuaxl(void) { return 1; }

/* emma.c:568-614, 1536-1540 in order: devices, the page-locked batch, the
   device batch, the parameter copy (again after every parameter change) */
int refdrv_gpu_init(int stride) {
  if (countdevices(0) < 1) return -1;
  initlocaldevice(0, 1);
  checkdevice(0);
  if (g_rstencil == NULL) {
    create_pinned_stencil_rad(&g_rstencil, stride);
    g_rstride = stride;
    create_radstencil_GPU(&g_cpu, stride);
  }
  create_param_GPU(&g_param, &g_cpu);
  return 0;
}

/* emma.c:2105-2122 */
void refdrv_gpu_finish(void) {
  if (g_rstencil != NULL) {
    destroy_radstencil_GPU(&g_cpu, g_rstride);
    destroy_pinned_stencil_rad(&g_rstencil, g_rstride);
  }
}
#else
int refdrv_gpuaxl(void) { return 0; }
#endif

/* RadSolver (rad_utils.c:1861) with a host stencil of `stride` RGRIDs (the
   page-locked batch of refdrv_gpu_init in a GPUAXL build) */
double refdrv_radsolver(int level, double dtnew, double aexp, int chemonly, int stride) {
#ifdef GPUAXL
  if (g_rstencil != NULL) return RadSolver(level, &g_param, g_firstoct, &g_cpu, g_rstencil, g_rstride, dtnew, aexp,
                                           chemonly);
#endif
  struct RGRID *st = (struct RGRID *)calloc(stride, sizeof(struct RGRID));
  REAL t = RadSolver(level, &g_param, g_firstoct, &g_cpu, st, stride, dtnew, aexp, chemonly);
  free(st);
  return t;
}

/* advancerad only (no final update) */
int refdrv_advancerad(int level, double dtnew, double aexp, int chemonly, int stride) {
  struct RGRID *st = (struct RGRID *)calloc(stride, sizeof(struct RGRID));
  int n = advancerad(g_firstoct, level, &g_cpu, st, stride, pow(0.5, level), dtnew, aexp, &g_param, chemonly);
  free(st);
  return n;
}

/* ---- stencil-level entry points on a caller-owned struct RGRID array ---- */
int refdrv_sizeof_rgrid(void) { return (int)sizeof(struct RGRID); }

void refdrv_stencil_sweeps(void *stencil, int nread, double dx, double dt, double c) {
  struct RGRID *st = (struct RGRID *)stencil;
  rad_sweepX(st, 0, 0, nread, nread, dx, dt, c);
  rad_sweepY(st, 0, 0, nread, nread, dx, dt, c);
  rad_sweepZ(st, 0, 0, nread, nread, dx, dt, c);
}

void refdrv_stencil_update(void *stencil, int nread, double dx, double dt, double c) {
  updatefieldrad(NULL, (struct RGRID *)stencil, nread, nread, &g_cpu, dx, dt, c);
}

void refdrv_stencil_chem(void *stencil, int nread, double dx, double dt, double aexp, int chemonly) {
  chemrad((struct RGRID *)stencil, nread, nread, &g_cpu, dx, dt, &g_param, aexp, chemonly);
}

/* ---- scalar functions ---- */
void refdrv_eddington(int n, const double *fx, const double *fy, const double *fz, const double *e, double c,
                      const int *i, const int *j, double *out) {
  for (int k = 0; k < n; k++) out[k] = Eddington(fx[k], fy[k], fz[k], e[k], c, i[k], j[k]);
}

void refdrv_getcellnei(int icell, int *vnei, int *vcell) { getcellnei(icell, vnei, vcell); }

void refdrv_e2t(int n, const double *nhplus, const double *eint, const double *nh, double aexp, double *temp) {
  for (int k = 0; k < n; k++) {
    struct Rtype R;
    memset(&R, 0, sizeof(R));
    R.nhplus = nhplus[k];
    R.eint = eint[k];
    R.nh = nh[k];
    E2T(&R, aexp, &g_param);
    temp[k] = R.temp;
  }
}

/* rate fits and cooling function of chem_utils.c:37-187, as chemrad calls them */
void refdrv_rates_n(int n, const double *temp, const double *x, const double *nH, double *alpha_a, double *alpha_b,
                    double *beta, double *lambda, double *tcool) {
  for (int k = 0; k < n; k++) {
    alpha_a[k] = cucompute_alpha_a(temp[k], 1., 1.);
    alpha_b[k] = cucompute_alpha_b(temp[k], 1., 1.);
    beta[k] = cucompute_beta(temp[k], 1., 1.);
    REAL l, t;
    cuCompCooling(temp[k], x[k], nH[k], &l, &t, 1., 1., 0);
    lambda[k] = l;
    tcool[k] = t;
  }
}

/* FillRad's COARSERAD block (src_utils.c:717-733): on the coarse level, every
   split cell gathers the volume-weighted E and F of the leaves below it */
int refdrv_fillrad_coarse(int level, double aexp) {
#ifdef COARSERAD
  if (level != g_param.lcoarse) return 0;
  int ns = 0;
  for (int i = 0; i < g_noct[level]; i++)
    for (int c = 0; c < 8; c++)
      if (g_lev[level][i].cell[c].child != NULL) ns += putE2coarse(&g_lev[level][i].cell[c], &g_param, level, aexp, 0.);
  return ns;
#else
  (void)level;
  (void)aexp;
  return -1;
#endif
}

/* oct lower corners (struct OCT x, y, z), 3 per oct: read by stromgren_source
   and by the TRANS* domain faces of the WRADTEST gather */
int refdrv_set_oct_pos(int level, const double *xyz) {
  if (level < g_lmin || level > g_lmax) return -1;
  for (int i = 0; i < g_noct[level]; i++) {
    g_lev[level][i].x = xyz[3 * i + 0];
    g_lev[level][i].y = xyz[3 * i + 1];
    g_lev[level][i].z = xyz[3 * i + 2];
  }
  return 0;
}

#ifdef WRADHYD
/* ---- hydro coupling (WRADHYD) and the WRADTEST Strömgren source ---- */
int refdrv_wradhyd(void) { return 1; }

static struct UNITARY_STARS_TEST g_star;

/* per-cell hydro state of a level (struct Wtype, prototypes.h:611-627) */
int refdrv_set_hydro(int level, const double *d, const double *u, const double *v, const double *w, const double *p,
                     const double *dX) {
  if (level < g_lmin || level > g_lmax) return -1;
  int n = g_noct[level] * 8;
  for (int i = 0; i < n; i++) {
    struct Wtype *W = &cell_at(level, i)->field;
    W->d = d[i];
    W->u = u[i];
    W->v = v[i];
    W->w = w[i];
    W->p = p[i];
    W->dX = dX[i];
    W->E = 0.;
  }
  return 0;
}

int refdrv_get_hydro(int level, double *p, double *dX, double *E) {
  if (level < g_lmin || level > g_lmax) return -1;
  int n = g_noct[level] * 8;
  for (int i = 0; i < n; i++) {
    const struct Wtype *W = &cell_at(level, i)->field;
    p[i] = W->p;
    dX[i] = W->dX;
    E[i] = W->E;
  }
  return 0;
}

/* stromgren_source parameters (src_utils.c:339-412): source position, star
   mass [solar masses], lifetime [yr], srcint [photons/s/kg]; the time groups
   come from refdrv_set_params */
void refdrv_set_stromgren(double x, double y, double z, double mass, double lifetime, double srcint) {
  g_star.src_pos_x = x;
  g_star.src_pos_y = y;
  g_star.src_pos_z = z;
  g_star.mass = mass;
  g_star.lifetime = lifetime;
  g_param.unitary_stars_test = &g_star;
  g_param.srcint = srcint;
}

#else
int refdrv_wradhyd(void) { return 0; }
#endif

#if defined(WRADHYD) || defined(STARS)
/* FillRad (src_utils.c:692-849) on one level, init = 0, tcur = cosmo->tphy */
int refdrv_fillrad(int level, double aexp) {
  g_param.lcoarse = g_lmin;
  return FillRad(level, &g_param, g_firstoct, &g_cpu, 0, aexp, 0.);
}
#endif

#ifdef STARS
/* ---- STARS build: star particles linked from the cells (CELL.phead,
   prototypes.h:832; struct PART, prototypes.h:722-759) and the source
   parameters stars_sources reads (src_utils.c:209-335) ---- */
int refdrv_stars(void) { return 1; }

static struct STARSPARAM g_starsparam;
static struct PART *g_parts[MAXLEV + 1];

/* srcint [photons/s/kg] and the stars' lifetime (STARSPARAM.tlife) */
void refdrv_set_star_params(double srcint, double tlife) {
  g_param.srcint = srcint;
  memset(&g_starsparam, 0, sizeof(g_starsparam));
  g_starsparam.tlife = tlife;
  g_param.stars = &g_starsparam;
}

/* the particles of one level, sorted by cell; within a cell the listed order
   becomes the cell's particle list order (phead -> next -> ...) */
int refdrv_set_stars(int level, int n, const int *cell, const double *mass, const double *age,
                     const int *radiative_state, const int *is_star) {
  if (level < g_lmin || level > g_lmax) return -1;
  free(g_parts[level]);
  g_parts[level] = (struct PART *)calloc(n > 0 ? n : 1, sizeof(struct PART));
  for (int i = 0; i < g_noct[level] * 8; i++) cell_at(level, i)->phead = NULL;
  for (int k = 0; k < n; k++) {
    struct PART *p = &g_parts[level][k];
    p->mass = mass[k];
    p->age = age[k];
    p->radiative_state = radiative_state[k];
    p->isStar = is_star[k];
    p->level = level;
    struct CELL *c = cell_at(level, cell[k]);
    if (c->phead == NULL) {
      c->phead = p;
    } else {
      struct PART *q = c->phead;
      while (q->next != NULL) q = q->next;
      q->next = p;
      p->prev = q;
    }
  }
  return 0;
}
#else
int refdrv_stars(void) { return 0; }
#endif
