/*
 * rad_utils_hip.c -- EMMA's GPU radiation entry points over libemma_rt.so.
 *
 * This file takes the place of the reference's CUDA module for the radiation
 * step (src/rad_utils_gpu.cu, and the device management of src/interface.cu)
 * in an EMMA build with GPUAXL: it defines the symbols RadSolver and emma.c
 * call, with the signatures of src/rad_utils_gpu.h:1-6 and
 * src/interface.h:16-18, and reaches the GPU only through the C-ABI of
 * include/emma_rt.h.  It is C, compiled with EMMA's own headers and flags
 * (REAL, NGRP, COARSERAD, ... as the rest of the build), e.g.
 *
 *   gcc -DGPUAXL <EMMA flags> -I<emma src> -I<this repo>/include -c rad_utils_hip.c
 *   ... link with -lemma_rt
 *
 * The host keeps gathering struct RGRID batches with gatherstencilrad and
 * scattering them back with scatterstencilrad (rad_utils.c:1482, 1658), as
 * advanceradGPU does (rad_utils_gpu.cu:987-1122); the sweeps, the update and
 * the chemistry of each batch run on the GPU (emma_rt_advance_stencil), in
 * RADSTEP and in COARSERAD builds.
 *
 * Errors: the reference's void entry points end the run on a device error
 * (WCUDA_ERR, gpu_type.h:4-15: message and exit(EXIT_FAILURE)); these do the
 * same with the library's status text.  The N<0 / NaN and EMIN conditions the
 * CPU path aborts on (rad_utils.c:1611-1650, chem_utils.c:382-385) end the run
 * the same way.
 */
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "prototypes.h"
#include "rad_utils_gpu.h"
#include "interface.h"
#include "emma_rt.h"

/* the host's gather and scatter (rad_utils.c:1482, 1658), which
   rad_utils_gpu.cu:18-19 declares the same way */
struct OCT *gatherstencilrad(struct OCT *octstart, struct RGRID *stencil, int stride, struct CPUINFO *cpu,
                             int *nread, REAL cloc);
struct OCT *scatterstencilrad(struct OCT *octstart, struct RGRID *stencil, int stride, struct CPUINFO *cpu,
                              REAL dxcur, REAL dtnew, REAL cloc);

/* one library context per process (one GPU per MPI rank, initlocaldevice) */
static emma_rt_ctx *g_ctx = NULL;
static int g_stride = 0;
static emma_rt_params g_p;

static void fatal(const char *where, int st) {
  fprintf(stderr, "emma_rt: %s failed: %s (status %d)\n", where, emma_rt_strerror(st), st);
  exit(EXIT_FAILURE);
}

/* The RGRID byte layout this file is compiled with (prototypes.h:939-1168),
   in the order emma_rt_check_layout reads it. */
static void host_layout(long *v) {
  v[0] = sizeof(struct Rtype);
  v[1] = offsetof(struct Rtype, e);
  v[2] = offsetof(struct Rtype, fx);
  v[3] = offsetof(struct Rtype, fy);
  v[4] = offsetof(struct Rtype, fz);
  v[5] = offsetof(struct Rtype, src);
  v[6] = offsetof(struct Rtype, nhplus);
  v[7] = offsetof(struct Rtype, eint);
  v[8] = offsetof(struct Rtype, nh);
  v[9] = offsetof(struct Rtype, temp);
  v[10] = sizeof(struct CELLLIGHT_R);
  v[11] = offsetof(struct CELLLIGHT_R, split);
  v[12] = sizeof(struct CELLFLUX_R);
  v[13] = offsetof(struct CELLFLUX_R, deltaR);
  v[14] = offsetof(struct CELLFLUX_R, rfield);
  v[15] = offsetof(struct CELLFLUX_R, rfieldnew);
  v[16] = offsetof(struct CELLFLUX_R, rflux);
  v[17] = sizeof(struct OCTLIGHT_R);
  v[18] = sizeof(struct OCTFLUX_R);
  v[19] = sizeof(struct RGRID);
  v[20] = offsetof(struct RGRID, oct);
  v[21] = offsetof(struct RGRID, New);
}

/* the parts of RUNPARAMS the radiation step reads (prototypes.h:299-413) */
static void params_of(const struct RUNPARAMS *param, emma_rt_params *p) {
  int g;
  memset(p, 0, sizeof(*p));
  p->ngrp = NGRP;
  p->ncvgcool = param->ncvgcool;
  p->fudgecool = param->fudgecool;
  p->clight = param->clight;
  p->unit_l = param->unit.unit_l;
  p->unit_v = param->unit.unit_v;
  p->unit_t = param->unit.unit_t;
  p->unit_n = param->unit.unit_n;
  p->unit_mass = param->unit.unit_mass;
  p->unit_d = param->unit.unit_d;
  p->unit_N = param->unit.unit_N;
  for (g = 0; g < NGRP; g++) {
    p->hnu[g] = param->atomic.hnu[g];
    p->alphae[g] = param->atomic.alphae[g];
    p->alphai[g] = param->atomic.alphai[g];
    p->factgrp[g] = param->atomic.factgrp[g];
  }
  if (param->cosmo != NULL) { /* create_param_GPU copies the cosmology too (rad_utils_gpu.cu:62-67) */
    p->H0 = param->cosmo->H0;
    p->om = param->cosmo->om;
    p->ov = param->cosmo->ov;
  }
  p->lcoarse = param->lcoarse;
  p->ngrp_time = NGRP_TIME;
  for (g = 0; g <= NGRP_TIME; g++) /* readAtomic's bounds, one group from 0 yr if absent */
    p->time_bound[g] = param->atomic.time_bound != NULL ? param->atomic.time_bound[g] : (g < NGRP_TIME ? 0. : INFINITY);
}

/* ------------------------------------------------------------------ interface.cu:61-84 */
int countdevices(int rank) {
  int n = 0;
  emma_rt_countdevices(rank, &n);
  return n;
}

void initlocaldevice(int rank, int count) {
  int st = emma_rt_initlocaldevice(rank, count);
  if (st) fatal("initlocaldevice", st);
}

void checkdevice(int rank) {
  int d = -1;
  int st = emma_rt_checkdevice(rank, &d);
  if (st) fatal("checkdevice", st);
}

/* ------------------------------------------------------------------ rad_utils_gpu.cu:30-106 */
/* create_radstencil_GPU comes before create_param_GPU in emma.c (605-614):
   the device buffer of `stride` RGRIDs is made as soon as the library
   context exists.  cpu->rad_stencil stays NULL: the library owns the buffer. */
void create_radstencil_GPU(struct CPUINFO *cpu, int stride) {
  (void)cpu;
  g_stride = stride;
  if (g_ctx != NULL) {
    int st = emma_rt_create_radstencil(g_ctx, stride, (int)sizeof(REAL));
    if (st) fatal("create_radstencil_GPU", st);
  }
}

void create_pinned_stencil_rad(struct RGRID **stencil, int stride) {
  void *p = NULL;
  int st = emma_rt_alloc_pinned((size_t)stride * sizeof(struct RGRID), &p);
  if (st) fatal("create_pinned_stencil_rad", st);
  *stencil = (struct RGRID *)p;
}

void destroy_radstencil_GPU(struct CPUINFO *cpu, int stride) {
  (void)cpu;
  (void)stride;
  if (g_ctx != NULL) emma_rt_destroy_radstencil(g_ctx);
  g_stride = 0;
}

/* frees the page-locked batch and clears the caller's pointer (the
   reference's version frees the RGRID** itself, rad_utils_gpu.cu:51-53) */
void destroy_pinned_stencil_rad(struct RGRID **stencil, int stride) {
  (void)stride;
  if (stencil != NULL && *stencil != NULL) {
    emma_rt_destroy_pinned_stencil(*stencil);
    *stencil = NULL;
  }
}

/* the one-time parameter copy (rad_utils_gpu.cu:57-106): the library
   context takes the radiation parameters, checks the RGRID layout this file
   was compiled with and, in a COARSERAD build, switches to that mode */
void create_param_GPU(struct RUNPARAMS *param, struct CPUINFO *cpu) {
  long lay[22];
  int st;
  (void)cpu;
  params_of(param, &g_p);
  if (g_ctx == NULL) {
    st = emma_rt_create(&g_p, &g_ctx);
    if (st) fatal("create_param_GPU", st);
  } else {
    st = emma_rt_set_params(g_ctx, &g_p);
    if (st) fatal("create_param_GPU", st);
  }
  host_layout(lay);
  st = emma_rt_check_layout(NGRP, (int)sizeof(REAL), lay);
  if (st) fatal("create_param_GPU (struct RGRID layout)", st);
#ifdef COARSERAD
  st = emma_rt_set_mode(g_ctx, EMMA_RT_COARSERAD);
#else
  st = emma_rt_set_mode(g_ctx, EMMA_RT_RADSTEP);
#endif
  if (st) fatal("create_param_GPU", st);
  if (g_stride > 0) {
    st = emma_rt_create_radstencil(g_ctx, g_stride, (int)sizeof(REAL));
    if (st) fatal("create_param_GPU (stencil buffer)", st);
  }
}

/* ------------------------------------------------------------------ rad_utils_gpu.cu:987-1122 */
/* RadSolver's GPU branch (rad_utils.c:1884-1888): batches of at most
   `stride` octs of the level, gathered and scattered by the host, advanced
   on the GPU.  Returns the number of octs processed. */
int advanceradGPU(struct OCT **firstoct, int level, struct CPUINFO *cpu, struct RGRID *stencil, int stride,
                  REAL dxcur, REAL dtnew, REAL aexp, struct RUNPARAMS *param, int chemonly) {
  struct OCT *nextoct, *curoct;
  int nreadtot = 0, nread = 0, st;
  const REAL cloc = aexp * param->clight * LIGHT_SPEED_IN_M_PER_S / param->unit.unit_v;
  if (g_ctx == NULL) fatal("advanceradGPU (create_param_GPU not called)", EMMA_RT_ERR_STATE);
  if (stride > g_stride) fatal("advanceradGPU (stride larger than the device stencil)", EMMA_RT_ERR_ARG);
  if (param->clight != g_p.clight) { /* the light speed may change between calls (advanceamr.c:198-215) */
    g_p.clight = param->clight;
    st = emma_rt_set_params(g_ctx, &g_p);
    if (st) fatal("advanceradGPU", st);
  }
  nextoct = firstoct[level - 1];
  if ((nextoct != NULL) && (cpu->noct[level - 1] != 0)) {
    do {
      curoct = nextoct;
      nextoct = gatherstencilrad(curoct, stencil, stride, cpu, &nread, cloc);
      if (nread > 0) {
        st = emma_rt_advance_stencil(g_ctx, stencil, nread, level, dxcur, dtnew, aexp, chemonly);
        if (st) fatal("advanceradGPU", st);
        nextoct = scatterstencilrad(curoct, stencil, nread, cpu, dxcur, dtnew, cloc);
        nreadtot += nread;
      }
    } while ((nextoct != NULL) && (nread > 0));
  }
  return nreadtot;
}
