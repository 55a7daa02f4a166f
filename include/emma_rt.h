/*
 * emma_rt.h -- C-ABI of the MI355X (gfx950) M1 radiative-transfer + hydrogen
 * chemistry library (libemma_rt.so).
 *
 * This is the boundary through which EMMA's C host reaches the GPU for the
 * radiation step.  It replaces the reference's CUDA module
 * (src/rad_utils_gpu.cu, src/chem_utils_gpu.cu, src/interface.cu, declared in
 * src/rad_utils_gpu.h:1-6 and src/interface.h:1-13).  Signatures take plain
 * pointers and sizes only.  Every entry point returns a status code
 * (EMMA_RT_OK or a negative EMMA_RT_ERR_*); the library never aborts or exits
 * where the reference calls abort()/exit().
 *
 * Two ways in:
 *  1. Stencil path (drop-in for advanceradGPU's kernel sequence,
 *     rad_utils_gpu.cu:1054-1086): the host keeps gathering struct RGRID
 *     batches with gatherstencilrad and scattering them back with
 *     scatterstencilrad; emma_rt_advance_stencil runs the three sweeps, the
 *     update and the chemistry on the batch in the reference's own byte layout.
 *  2. Resident-level path (fast path): the levels' radiation state lives on the
 *     device as structure-of-arrays with index tables for the oct neighbours;
 *     emma_rt_radsolver is RadSolver (rad_utils.c:1861-2022) for one level,
 *     gather and scatter included, and emma_rt_coarse_step runs the whole
 *     RADSTEP level recursion of one or more coarse steps as one HIP graph.
 *
 * Layout of the resident path (shared with the checker and the host mirror):
 *   cell index = oct*8 + icell, icell = ix + 2 iy + 4 iz (oct.c:116-185);
 *   group fields are group-major, e[g*ncell + cell];
 *   nei_oct[oct*6 + face] >= 0 : same-level neighbour oct on that face,
 *                          < 0 : the level-1 neighbour cell is a leaf whose
 *                                cell index is -1 - nei_oct (it is injected
 *                                into the stencil, rad_utils.c:1462-1476);
 *   child[cell] >= 0 : the cell is split, index of its child oct at level+1.
 */
#ifndef EMMA_RT_H
#define EMMA_RT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMMA_RT_MAX_GRP 9

/* status codes */
#define EMMA_RT_OK 0
#define EMMA_RT_ERR_ARG (-1)        /* bad argument or size */
#define EMMA_RT_ERR_HIP (-2)        /* HIP runtime error (WCUDA_ERR exit, gpu_type.h:4-15) */
#define EMMA_RT_ERR_NEG_PHOTON (-3) /* N<0 or NaN after transport (abort, rad_utils.c:1611-1650) */
#define EMMA_RT_ERR_EMIN (-4)       /* photon density below EMIN (abort, chem_utils.c:382-385) */
#define EMMA_RT_ERR_NOCONV (-5)     /* chemistry sub-step underflow (abort, chem_utils.c:453-466) */
#define EMMA_RT_ERR_NEIGHBOR (-6)   /* missing neighbour (abort, rad_utils.c:1441-1444) */
#define EMMA_RT_ERR_LAYOUT (-7)     /* host struct layout differs from the library's */
#define EMMA_RT_ERR_NODEVICE (-8)   /* no GPU visible */
#define EMMA_RT_ERR_STATE (-9)      /* call out of order (no tree / no stencil buffer) */
#define EMMA_RT_ERR_SCHED (-10)     /* persistent coarse step: a dependency wait passed its time bound */

/* level-stepping mode of the resident path (compile-time COARSERAD flag of
   the reference, prototypes.h:81-83) */
#define EMMA_RT_RADSTEP 0   /* every level transports with its own dt (advanceamr.c:243) */
#define EMMA_RT_COARSERAD 1 /* transport on the coarse level only (advanceamr.c:1309) */

/* precision of the resident path */
#define EMMA_RT_FP64 0  /* everything FP64 */
#define EMMA_RT_MIXED 1 /* photon density, flux and source FP32; x, eint, T, n_H FP64 */

/* Run parameters the path reads from struct RUNPARAMS (prototypes.h:299-413):
   units (struct UNITS, prototypes.h:171-179), clight/fudgecool/ncvgcool
   (prototypes.h:366-371), the atomic group table as readAtomic leaves it
   (parameters.c:490-497: hnu in J, alphae/alphai in m^3/s at the reduced
   light speed), the COSMOPARAM entries chemrad uses (chem_utils.c:738), the
   coarse level (RUNPARAMS.lcoarse: advancerad's COARSERAD test,
   rad_utils.c:1808-1812, and the test source's radius, src_utils.c:342), and
   the atomic table's time groups (readAtomic, parameters.c:455-468):
   ngrp = NGRP_SPACE x ngrp_time groups, time group i covering
   [time_bound[i], time_bound[i+1]) years, time_bound[ngrp_time] = INFINITY. */
typedef struct emma_rt_params {
  int ngrp;
  int ncvgcool;
  double fudgecool;
  double clight;
  double unit_l, unit_v, unit_t, unit_n, unit_mass, unit_d, unit_N;
  double hnu[EMMA_RT_MAX_GRP], alphae[EMMA_RT_MAX_GRP], alphai[EMMA_RT_MAX_GRP], factgrp[EMMA_RT_MAX_GRP];
  double H0, om, ov;
  int lcoarse;
  int ngrp_time;
  double time_bound[EMMA_RT_MAX_GRP + 1];
} emma_rt_params;

typedef struct emma_rt_ctx emma_rt_ctx;

/* Counters of the last radsolver / coarse-step call. */
typedef struct emma_rt_stats {
  long long cell_updates;   /* leaf cells advanced (one per leaf per level step); -1 with the
                               counters off (emma_rt_enable_counters), as chem_substeps and
                               deferred_cells */
  long long chem_substeps;  /* accepted chemistry sub-steps, summed over cells */
  long long level_steps;    /* level-step kernel launches */
  double kernel_ms;         /* device time (HIP events), see emma_rt_enable_timing */
  double photons;           /* FP64 sum of N*dx^3 over leaf cells after the call (photon budget) */
  int status;               /* first error raised on the device, EMMA_RT_OK if none */
  int bad_level;            /* level and cell of that error */
  int bad_cell;
  int timed_spans;          /* event-bracketed spans summed into kernel_ms */
  long long deferred_cells; /* leaves whose chemistry the straggler pass ran (emma_rt_set_chem_budget) */
} emma_rt_stats;

/* ---- device management: interface.cu:61-84 (countdevices, initlocaldevice,
   checkdevice).  initlocaldevice binds rank % count, which the reference
   meant to do but does not (interface.cu:73 always selects device 0). ---- */
int emma_rt_countdevices(int rank, int *count);
int emma_rt_initlocaldevice(int rank, int count);
int emma_rt_checkdevice(int rank, int *device);
const char *emma_rt_strerror(int status);

/* ---- parameters: create_param_GPU (rad_utils_gpu.cu:57-106) makes a
   one-time device copy of RUNPARAMS; here the context owns the copy and
   emma_rt_set_params refreshes it (clight may change, advanceamr.c:198-215). */
int emma_rt_create(const emma_rt_params *p, emma_rt_ctx **out);
int emma_rt_set_params(emma_rt_ctx *ctx, const emma_rt_params *p);
int emma_rt_destroy(emma_rt_ctx *ctx);

/* ---- stencil path (drop-in for the kernels of advanceradGPU) ----
   The caller's struct RGRID layout (prototypes.h:1165-1168) is described by
   22 numbers: sizeof/offsetof of Rtype (e,fx,fy,fz,src,nhplus,eint,nh,temp),
   CELLLIGHT_R (size, split), CELLFLUX_R (size, deltaR, rfield, rfieldnew,
   rflux), OCTLIGHT_R, OCTFLUX_R, RGRID (size, oct, New).  The library checks
   it against the layout it computes for (ngrp, real_bytes) and refuses a
   mismatch with EMMA_RT_ERR_LAYOUT. */
int emma_rt_rgrid_layout(int ngrp, int real_bytes, long *layout22);
int emma_rt_check_layout(int ngrp, int real_bytes, const long *layout22);
/* create_radstencil_GPU (rad_utils_gpu.cu:30-36): device buffer of `stride`
   RGRIDs of REAL = double (real_bytes 8) or float (4). */
int emma_rt_create_radstencil(emma_rt_ctx *ctx, int stride, int real_bytes);
int emma_rt_destroy_radstencil(emma_rt_ctx *ctx);
/* create_pinned_stencil_rad / destroy_pinned_stencil_rad (rad_utils_gpu.cu:
   39-43, 51-53): page-locked host memory for `stride` RGRIDs of the context's
   ngrp (zeroed), so the copies of emma_rt_advance_stencil run at DMA speed. */
int emma_rt_create_pinned_stencil(emma_rt_ctx *ctx, int stride, int real_bytes, void **out);
int emma_rt_destroy_pinned_stencil(void *stencil);
/* `bytes` of page-locked host memory, no context needed (the allocation of
   create_pinned_stencil_rad, rad_utils_gpu.cu:39-43, which emma.c:605-606
   makes before the parameters exist); free with
   emma_rt_destroy_pinned_stencil */
int emma_rt_alloc_pinned(size_t bytes, void **out);
/* rad_utils_gpu.cu:1054-1086: copy `nread` RGRIDs of `level` host->device,
   run rad_sweepX/Y/Z + updatefieldrad + chemrad semantics (CPU semantics of
   rad_utils.c/chem_utils.c), copy back.  Only the parts the kernel reads (oct
   slots, New.cell[*].rfieldnew) go to the device and only New.cell[*].rfieldnew
   and rflux, the parts that change, come back.  `stencil` is host memory in
   the reference's RGRID layout (page-locked memory from
   emma_rt_create_pinned_stencil is fastest).  In EMMA_RT_COARSERAD mode
   (emma_rt_set_mode) the batch follows the COARSERAD build: the sweeps and
   the update run only when level == params.lcoarse and !chemonly (condadvec,
   rad_utils.c:1808-1812), fluxes of split neighbours are kept and split
   cells are transport-updated (rad_utils.c:518-526, 1548-1552); without the
   sweeps only rfieldnew comes back. */
int emma_rt_advance_stencil(emma_rt_ctx *ctx, void *stencil, int nread, int level, double dxcur, double dtnew,
                            double aexp, int chemonly);

/* emma_rt_advance_stencil without the device buffer of
   emma_rt_create_radstencil: for callers that cannot allocate device memory
   (code under HIP graph capture, serving loops, harnesses that forbid
   hipMalloc).  No counterpart in the reference.  The batch is staged in a
   static device area of EMMA_RT_NOALLOC_STENCIL_BYTES per GPU (allocated with
   the library's code object, held until the process ends); the call allocates
   nothing.  nread * sizeof(struct RGRID) beyond that size (e.g. 4128 records
   at NGRP=1, 1826 at NGRP=3, REAL=double) returns EMMA_RT_ERR_ARG.  real_bytes
   is sizeof(REAL).  Results equal emma_rt_advance_stencil's.  All calls on one
   GPU share its area: calls that overlap in time on one GPU (other host
   threads, other contexts) race and must be serialised by the caller. */
#define EMMA_RT_NOALLOC_STENCIL_BYTES (32u << 20)
int emma_rt_advance_stencil_noalloc(emma_rt_ctx *ctx, void *stencil, int nread, int level, double dxcur,
                                    double dtnew, double aexp, int chemonly, int real_bytes);

/* ---- resident-level path ---- */
/* nei_oct entries of faces on a non-periodic domain face, the WRADTEST
   TRANSXM..TRANSZP and REFXM/REFYM/REFZM build flags of the reference
   (rad_utils.c:1065-1409): the cells of the boundary oct see a copy of the
   oct's own cells across the face (each cell itself, except on +y where the
   reference's face table makes cells 6 and 7 see each other), with the flux
   component normal to the face negated for REFLECTIVE; nothing is refluxed
   through the face (rad_utils.c:1700-1734).  The reference decides from the
   octs' positions (x == 0 for the minus x face, x + 2 dx == 1 for the plus
   x face, ...); the caller marks those faces. */
#define EMMA_RT_NEI_TRANSMISSIVE (-2147483647 - 1)
#define EMMA_RT_NEI_REFLECTIVE (-2147483647)

/* Device tree of levels lmin..lmax: noct[lmax-lmin+1]; nei_oct and child are
   the levels' arrays concatenated in level order (6 and 8 ints per oct).
   nei_oct: the same-level oct across the face, -1-c for the coarser level's
   leaf cell c, or one of the EMMA_RT_NEI_* domain-face codes.
   A level holds at most min(2^29, (2^31-1)/ngrp) cells, and fewer than 2^28
   (leaf, face) pairs with a finer oct behind the face (EMMA_RT_ERR_ARG
   beyond: cell indices and per-group byte offsets are 32-bit, and a leaf's
   child entry encodes its first pair).  Each such face has its own reflux
   accumulators, written by one thread: AMR results are the same bit for bit
   from run to run. */
int emma_rt_tree_create(emma_rt_ctx *ctx, int lmin, int lmax, const int *noct, const int *nei_oct,
                        const int *child, int precision);
/* host SoA (double) <-> device state of one level (rfield, prototypes.h:849) */
int emma_rt_set_level(emma_rt_ctx *ctx, int level, const double *e, const double *fx, const double *fy,
                      const double *fz, const double *src, const double *nhplus, const double *eint,
                      const double *nh, const double *temp);
int emma_rt_get_level(emma_rt_ctx *ctx, int level, double *e, double *fx, double *fy, double *fz, double *src,
                      double *nhplus, double *eint, double *nh, double *temp);
/* reionisation times (rad_utils.c:1901-1913): enable per-cell t_first/t_last
   bookkeeping (initialised to -1) and read them back */
int emma_rt_enable_xion_times(emma_rt_ctx *ctx, int enable);
int emma_rt_get_xion_times(emma_rt_ctx *ctx, int level, double *t_first, double *t_last);
int emma_rt_set_tphy(emma_rt_ctx *ctx, double tphy);

/* RadSolver (rad_utils.c:1861-2022) on one level, RADSTEP mode: transport
   from the level's current state (coarse leaves injected), update, chemistry,
   1/8 refluxing of coarse-fine faces into the coarser level, final update
   (rfield := rfieldnew; split cells := mean of their 8 children).
   Synchronous; returns the first device error as a status. */
int emma_rt_radsolver(emma_rt_ctx *ctx, int level, double dtnew, double aexp, int chemonly);

/* `ncoarse` coarse steps of the RADSTEP recursion (advanceamr.c:243-1300):
   level l takes 2^(l-lmin) steps of dt_rad(l) (advanceamr.c:226) per coarse
   step, finer levels first.  Captured once as a HIP graph and replayed. */
int emma_rt_coarse_step(emma_rt_ctx *ctx, double aexp, int ncoarse);

/* COARSERAD (EMMA_RT_COARSERAD): the RadSolver step of the coarse level
   transports every cell (split cells included, fluxes of split neighbours
   kept, no refluxing; rad_utils.c:518-526, 1548-1552), finer levels run the
   chemistry only, and after the coarse step the coarse E, F are pushed down to
   the leaves (distribE, rad_utils.c:26-45, 2009-2013).  emma_rt_fillrad is
   FillRad's COARSERAD block (src_utils.c:717-733): before the coarse step,
   every split coarse cell gathers the volume-weighted E, F of its leaves
   (putE2coarse/collectE, src_utils.c:94-150); a no-op in RADSTEP mode and on
   finer levels.  emma_rt_coarse_step follows Advance_level_RAD
   (advanceamr.c:1309-1537) in this mode: one step of the coarse dt per level. */
int emma_rt_set_mode(emma_rt_ctx *ctx, int mode);
/* FillRad (src_utils.c:692-849) on one level: the COARSERAD block above, then
   the per-cell coupling / source block described below (when enabled) */
int emma_rt_fillrad(emma_rt_ctx *ctx, int level, double aexp);

/* ---- coupling to the hydro state and the test source (WRADHYD / WRADTEST).
   FillRad's per-cell block (src_utils.c:738-838) and RadSolver's injection
   into the hydro fields (rad_utils.c:1941-1951, 1993-2004) on the device.
   emma_rt_set_hydro stores a level's struct Wtype fields (prototypes.h:611-627;
   d, u, v, w, p, dX, host arrays of ncell doubles).  With the coupling on
   (emma_rt_set_coupling(ctx, 1)), emma_rt_fillrad copies n_H = d (1-YHE),
   eint = p/(GAMMA-1), x n_H = dX into the radiation state of every cell, zeroes
   the source of split cells and applies E2T to unsplit ones; emma_rt_radsolver
   and emma_rt_coarse_step then write p = (GAMMA-1) eint, dX = x d (1-YHE) and
   E = p/(GAMMA-1) + d |u|^2/2 (getE, hydro_utils.c:29-31) for every cell after
   the level's final update, read back with emma_rt_get_hydro (any pointer may
   be NULL).  emma_rt_coarse_step calls FillRad before every level step, as
   Advance_level does (advanceamr.c:1035, 1459).
   emma_rt_set_stromgren_source enables stromgren_source (src_utils.c:339-412)
   in FillRad: unsplit cells whose centre lies within 2^-lcoarse of (x, y, z)
   and in the positive octant get mass*SOLAR_MASS*srcint/8 photons/s spread over
   (2^-lcoarse unit_l)^3, in code units, in the NGRP_SPACE groups of the time
   group tphy (years) lies in (emma_rt_params.time_bound), while
   fmod(tphy, lifetime) < lifetime; every other group and cell gets 0.  As
   under the reference's TESTCOSMO build (FillRad reads cosmo->tphy =
   a2t(aexp), src_utils.c:706, advanceamr.c:279, 860), the source's time is
   the context's tphy, one value per call (aexp is one value per call too).
   It needs the octs' lower corners (struct OCT x, y, z; 3 doubles per oct)
   from emma_rt_set_oct_pos.
   Calls return EMMA_RT_ERR_STATE when a level lacks what the coupling or the
   source reads. */
int emma_rt_set_hydro(emma_rt_ctx *ctx, int level, const double *d, const double *u, const double *v,
                      const double *w, const double *p, const double *dX);
int emma_rt_get_hydro(emma_rt_ctx *ctx, int level, double *p, double *dX, double *E);
int emma_rt_set_coupling(emma_rt_ctx *ctx, int hydro);
int emma_rt_set_oct_pos(emma_rt_ctx *ctx, int level, const double *xyz);
int emma_rt_set_stromgren_source(emma_rt_ctx *ctx, int enable, double x, double y, double z, double mass,
                                 double lifetime, double srcint);

/* Star particles as sources: putsource -> stars_sources (src_utils.c:209-335,
   446-520), the source of the reference's STARS builds, in FillRad on the
   device.  emma_rt_set_stars hands over the star particles of one level's
   leaf cells (struct PART, prototypes.h:722-759, linked from CELL.phead):
   listed cell by cell in ascending cell order and, within a cell, in the
   order of the cell's particle list (the order the reference adds them in);
   cell = the level's cell index, mass (code units), age (the unit of
   tphy), radiative_state (the time group the particle radiates in, <
   ngrp_time), is_star (isStar).  nstar = 0 removes them.
   emma_rt_set_star_sources switches FillRad to these sources (and the
   Stromgren source off): every unsplit cell's src is cleared, then each of
   its particles with isStar >= 6, other than 100 (the AGN case, not built),
   adds to the NGRP_SPACE groups of its radiative_state
   ((mass unit_d) srcint / dx^3) unit_t / unit_N aexp^2 while tphy > age;
   srcint is RUNPARAMS.srcint [photons/s/kg], tlife STARSPARAM.tlife;
   decrease_after_tlife is the reference's DECREASE_EMMISIVITY_AFTER_TLIFE
   flag (stars older than tlife radiate that value times
   ((tphy - age)/tlife)^-4). */
int emma_rt_set_stars(emma_rt_ctx *ctx, int level, int nstar, const int *cell, const double *mass, const double *age,
                      const int *radiative_state, const int *is_star);
int emma_rt_set_star_sources(emma_rt_ctx *ctx, int enable, double srcint, double tlife, int decrease_after_tlife);

/* A restart point on the device: emma_rt_save_state copies every level's
   current state (rfield: e, fx, fy, fz, src, nhplus, eint, nh, temp), the
   reflux accumulators, the reionisation times and the hydro fields the
   injection writes into a device copy (allocated by the first save),
   emma_rt_restore_state copies them back.  The reference restarts from its
   dump of the whole tree (io.c); this keeps a run's state on the GPU, e.g. to
   repeat a coarse step from the same state. */
int emma_rt_save_state(emma_rt_ctx *ctx);
int emma_rt_restore_state(emma_rt_ctx *ctx);

/* Chemistry stragglers (mixed precision, 1 or 3 groups).  The reference's
   chemrad sub-steps each cell until its dt is covered (chem_utils.c:453-767),
   so a few cells near ionisation fronts take many more sub-steps than the rest;
   on the GPU a wave then runs as long as its slowest lane.  With a budget
   b > 0 for a level, a leaf still sub-stepping after b tries keeps its
   transported N, F and leaves its chemistry to a second pass over the list of
   such leaves, which restarts it from the same inputs, one thread per leaf.
   The results are those of the undivided loop (the same sub-steps, rounded
   alike up to FP64 contraction); the choice only changes the speed.  level -1
   sets every level; budget 0 turns the pass off (the default).  No counterpart
   in the reference (its GPU chemistry, chem_utils_gpu.cu:155-582, runs one
   thread per cell to the end).  EMMA_RT_ERR_ARG on the FP64 and generic-group
   paths. */
int emma_rt_set_chem_budget(emma_rt_ctx *ctx, int level, int budget);
int emma_rt_get_chem_budget(emma_rt_ctx *ctx, int level, int *budget);
/* With a budget set, a leaf whose chain looks long when the budget runs out
   can stay in the first pass instead: no sub-step accepted yet, or more than
   `keep` sub-steps to go at the pace of its accepted ones (its remaining time
   over their mean length), so that long chains overlap the rest of the level
   step.  keep 0 (the default) defers every leaf that is still sub-stepping;
   on the benchmark workloads that is the faster choice (a long chain kept in
   the first pass holds its whole work-group's registers and LDS). */
int emma_rt_set_chem_keep(emma_rt_ctx *ctx, int level, double keep);
/* Halo variant of the mixed-precision level step (1 or 3 groups) for one
   level or every level (-1).  The results are the same; the variants differ
   in how a tile reaches its neighbours outside the tile (the halo):
   LISTED  each wave lists its halo faces, one round (at most 64 per wave:
           complete levels; EMMA_RT_ERR_ARG on a level with more);
   PERDIR  each lane's own halo face per direction pass, chemistry inputs
           loaded late (fewest live registers; large sparse levels);
   LATENCY all halo faces and inputs loaded up front (one memory round trip;
           small levels whose step is one work-group's latency);
   AUTO    the library's choice at tree creation (LISTED where possible,
           LATENCY on sparse levels of at most 256 tiles of 32 octs, PERDIR
           otherwise).  The reference's analogue is the CUDA launch shape it
           takes from the run parameters (nthread, nstream, rad_utils_gpu.cu:
           1025-1043).  emma_rt_get_level_tuning reads a level's variant,
           chemistry budget, whether it has more halo faces per wave than
           LISTED holds, and the chemistry keep factor (any pointer may be
           NULL). */
#define EMMA_RT_VARIANT_AUTO (-1)
#define EMMA_RT_VARIANT_LISTED 0
#define EMMA_RT_VARIANT_PERDIR 1
#define EMMA_RT_VARIANT_LATENCY 2
int emma_rt_set_level_variant(emma_rt_ctx *ctx, int level, int variant);
int emma_rt_get_level_tuning(emma_rt_ctx *ctx, int level, int *variant, int *chem_budget, int *halo_per_dir,
                             double *chem_keep);

/* Grid output: the serial STDIO dump of dumpalloct_serial (io.c:1233-1363)
   for the fields this library holds, from the device state.  For each
   requested field id (the ids of assign_grid_field, io.c:32-200, and
   readOutputParam_grid's names, parameters.c:132-177) one file
   <folder>grid_<name>/<name>.<ndump %05d>.p<rank %05d> holding int n,
   the first sizeof(float) bytes of the double tsim (as the reference writes
   its REAL tsim), float xmin, xmax, ymin, ymax, zmin, zmax over the dumped
   cells' octs, then n floats: every leaf cell (and every cell of level lmax),
   levels lcoarse..lmax, octs in the level's order, cells 0..7.  Held ids:
   0-2 cell corner x, y, z, 3 level, 4 cpu (= rank), 13-17 d, u, v, w, p
   (with the hydro coupling), 18-25 and 27-30 E, Fx, Fy, Fz of groups 0-2
   (groups < ngrp), 31 sum of the sources, 32 x_HII, 33 dX (hydro), 38 T,
   39-40 reionisation times (with emma_rt_enable_xion_times); any other id
   returns EMMA_RT_ERR_ARG.  Needs the octs' corners (emma_rt_set_oct_pos) on
   every level.  folder ends with '/'; the grid_<name>/ folders are created. */
int emma_rt_dump_grid(emma_rt_ctx *ctx, const char *folder, int ndump, int rank, double tsim, int nfield,
                      const int *field_ids);
/* The values of one field in that order (the data of its file): *n_out =
   the number of dumped cells; out == NULL asks for the count only, else out
   holds capacity floats (EMMA_RT_ERR_ARG when fewer than the count).  For a
   host writing its own files, e.g. dump_HDF5_grid's per-field dataset "data"
   (io.c:1495-1611) with its own HDF5 library. */
int emma_rt_grid_field(emma_rt_ctx *ctx, int field_id, int rank, float *out, long long capacity, long long *n_out);

/* ---- coarse-step schedule.  No counterpart in the reference, whose
   Advance_level recursion (advanceamr.c:243-1300, 1309-1537) launches
   advanceradGPU level by level from the host.  EMMA_RT_SCHED_GRAPH (the
   default): emma_rt_coarse_step replays a captured HIP graph of one launch per
   level step.  EMMA_RT_SCHED_PERSISTENT: every run of consecutive level steps
   (all of a RADSTEP coarse step; in COARSERAD mode the finer levels' steps,
   then the coarse level's, with the transfers between them) is ONE launch of
   persistent work-groups fed by a device work queue.  An item is one tile
   (32 octs) of one level step, with FillRad's block before and the hydro
   injection after it (several consecutive tiles per item on the largest
   levels); items wait only for the tiles whose results they read
   or whose inputs they overwrite, so tiles of different levels and steps run
   at once (a small level's latency chain beside the large levels' tiles).
   The queue's order comes from a list schedule the host computes once per
   tree; results are identical bit for bit to the graph schedule's.  The
   chemistry straggler pass (emma_rt_set_chem_budget) is not used inside the
   persistent kernel, and the latency halo variant runs as per-direction.
   Per-level device times are not available in this schedule: with
   EMMA_RT_TIMING_LAUNCH each persistent launch is one timed span. */
/* Tile order of the tiled mixed kernel on a complete periodic cube of 2^m
   octs a side stored in Morton order (the library checks the level's
   nei_oct against that layout at emma_rt_tree_create): EMMA_RT_ORDER_MORTON
   (the default) keeps the tiles' Morton order; EMMA_RT_ORDER_RASTER walks
   each XCD's octant of 4 x 4 x 2-oct tiles z fastest, then y, then x, so
   that the neighbours across a tile's larger faces are processed next to it
   (the autotuner picks per cube; Morton order read less from HBM on the
   128^3 and 256^3 cubes measured).  Speed only, results unchanged.
   level < 0: every level that can take it.
   EMMA_RT_ERR_ARG for RASTER on a level that is not such a cube. */
#define EMMA_RT_ORDER_MORTON 0
#define EMMA_RT_ORDER_RASTER 1
int emma_rt_set_level_order(emma_rt_ctx *ctx, int level, int order);
int emma_rt_get_level_order(emma_rt_ctx *ctx, int level, int *order, int *cube_side_log2);

#define EMMA_RT_SCHED_GRAPH 0
#define EMMA_RT_SCHED_PERSISTENT 1
typedef struct emma_rt_sched_info {
  long long items;   /* tile items of the last planned coarse-step call */
  long long deps;    /* dependency edges among them */
  int blocks;        /* persistent work-groups per launch */
  int launches;      /* persistent launches per call */
  double est_span;   /* list-schedule makespan, in tile steps */
  double est_chain;  /* longest dependency chain, in tile steps */
  double est_work;   /* total work, in tile steps */
} emma_rt_sched_info;
int emma_rt_set_schedule(emma_rt_ctx *ctx, int sched);
/* the schedule and (sched_info of the last coarse-step call; zero for the
   graph schedule) */
int emma_rt_get_schedule(emma_rt_ctx *ctx, int *sched, emma_rt_sched_info *info);
/* The planner alone (host code, no device needed): the tree as for
   emma_rt_tree_create (nlev levels, coarsest first), the level index of every
   level step in sequential order, a relative cost per tile step of each level
   (NULL: 1), the tiles per item of each level (NULL: the library's rule, one
   tile below 16 x slots tiles per level, else about 8 x slots items per level
   step), the work-group count and the number of queues (the library uses one
   per XCD, 8; each level step's items split into that many contiguous
   ranges).  Writes nitems / ndeps; with cap_items == 0 returns after that
   (sizes only), else fills qoff[nqueues + 1] (queue q's tickets are qoff[q]
   .. qoff[q+1]-1) and, by ticket, the level step, first tile and tile count
   of each item, its flags (1 raised priority, 2 a dependency in another
   queue, 4 a dependent in another queue), its place in the simulated start
   order, and the tickets it waits for (deps[dep_off[t] .. dep_off[t+1])),
   and est[3] = makespan, longest chain, total work (may be NULL).
   EMMA_RT_ERR_ARG on bad input or short buffers. */
int emma_rt_plan_schedule(int nlev, const int *noct, const int *nei_oct, const int *child, int nsteps,
                          const int *step_level, const double *tile_cost, const int *chunk, int slots, int nqueues,
                          long long cap_items, long long cap_deps, long long *nitems, long long *ndeps, int *qoff,
                          int *item_step, int *item_tile, int *item_ntile, unsigned char *item_flags,
                          int *item_rank, int *dep_off, int *deps, double *est);

int emma_rt_get_stats(emma_rt_ctx *ctx, emma_rt_stats *out);
/* device timing into emma_rt_stats.kernel_ms: EMMA_RT_TIMING_LAUNCH puts
   HIP events around every level-step kernel (coarse steps then launch
   eagerly, not from the captured graph); EMMA_RT_TIMING_GRAPH does the same
   for single level steps, and coarse steps replay their graph between one
   pair of events (the replay's device time, launch gaps inside the graph
   and the clearing of the call's status words at its head included) */
#define EMMA_RT_TIMING_OFF 0
#define EMMA_RT_TIMING_LAUNCH 1
#define EMMA_RT_TIMING_GRAPH 2
int emma_rt_enable_timing(emma_rt_ctx *ctx, int enable);
/* the kernels' device counters behind emma_rt_stats.cell_updates,
   chem_substeps and deferred_cells (on by default: enable = 1).  Off (0),
   the level steps skip their per-tile reduction (a barrier and a few atomics
   per 256 cells) and those three fields read -1; the results are the same
   bit for bit.  No counterpart in the reference, whose RadSolver counts
   nothing on the device. */
int emma_rt_enable_counters(emma_rt_ctx *ctx, int enable);
/* per-level device time of the last call's event-timed level-step launches
   (EMMA_RT_TIMING_LAUNCH, or single level steps): ms[k] and launches[k] for
   level lmin + k, k < nlev (either pointer may be NULL).  The reference times
   the same phases on the host (advancerad's tg/th/tu/ts, rad_utils.c:
   1796-1845; RadSolver's return value, rad_utils.c:2020-2021). */
int emma_rt_get_level_times(emma_rt_ctx *ctx, int nlev, double *ms, int *launches);

#ifdef __cplusplus
}
#endif
#endif
