/*
 * emma_oracle.c -- CPU restatement (FP64, scalar) of EMMA's M1 transport +
 * hydrogen chemistry hot path.  TEST INFRASTRUCTURE ONLY: the checker the HIP
 * library is compared against; never linked into or called by the product.
 *
 * Each function names the reference file:line it restates.  Floating-point
 * expressions keep the reference's operation order so that, compiled with
 * -O2 -ffp-contract=off, results are bit-identical to the reference CPU path
 * (checked against oracle/_ref and the committed golden vectors).
 */
#include "emma_oracle.h"

#include <math.h>
#include <string.h>

/* param.h constants (src/param.h:4-37,52-56) */
#define O_GAMMA (5. / 3.)
#define O_CFL (0.85)
#define O_FRAC_VAR (0.25)
#define O_PMIN (1e-10)
#define O_EMIN (1e-8)
#define O_LIGHT (299792458.)
#define O_KBOLTZ (1.3806e-23)
#define O_PARSEC (3.085677e16)
#define O_YHE (0.)
#define O_yHE (O_YHE / (1. - O_YHE) / 4.002)

/* face neighbour rule of oct.c:116-185: crossing face f=2d+s of cell c flips
   bit d; the neighbour stays in the oct (slot 6) when the move is inward */
void orc_getcellnei(int icell, int *vnei, int *vcell) {
  for (int f = 0; f < 6; f++) {
    int d = f >> 1, s = f & 1;
    int b = (icell >> d) & 1;
    vcell[f] = icell ^ (1 << d);
    vnei[f] = (b != s) ? 6 : f;
  }
}

/* rad_utils.c:448-475 */
double orc_eddington(double fx, double fy, double fz, double ee, double c, int i, int j) {
  double c2e = ee * c * c;
  double ff = 0.;
  double arg, chi, res = 0.;
  double n[3] = {0., 0., 0.};
  if (ee > 0) {
    ff = sqrt(fx * fx + fy * fy + fz * fz);
    if (ff > 0) {
      n[0] = fx / ff;
      n[1] = fy / ff;
      n[2] = fz / ff;
    }
    ff = ff / (c * ee);
  }
  arg = fmax(4. - 3. * ff * ff, 0.);
  chi = (3. + 4. * ff * ff) / (5. + 2. * sqrt(arg));
  if (i == j) res = (1. - chi) / 2. * c2e;
  arg = (3. * chi - 1.) / 2. * c2e;
  res += arg * n[i] * n[j];
  return res;
}

/* advanceamr.c:226 (L_comptstep_rad) */
double orc_dtrad(int level, double aexp, const orc_params *P) {
  return O_CFL / (aexp * P->clight * O_LIGHT / P->unit_v) / 3.0 / (double)(1 << level);
}

/* rad_utils.c:1784 */
double orc_cloc(double aexp, const orc_params *P) { return aexp * P->clight * O_LIGHT / P->unit_v; }

/* chem_utils.c:21-32 */
void orc_e2t(orc_cell *R, double aexp, const orc_params *P) {
  double eint = R->eint;
  double nH = R->nh;
  double x = R->nhplus / R->nh;
  double pstar = P->unit_n * P->unit_d * pow(P->unit_v, 2);
  nH = nH / pow(aexp, 3) * P->unit_N;
  eint = eint / pow(aexp, 5) * pstar;
  R->temp = eint / (1.5 * nH * O_KBOLTZ * (1. + x) * (1. + O_yHE));
}

/* chem_utils.c:37-51 (unit_number = 1 as called from chemrad) */
double orc_alpha_b(double temp) {
  double lambda = 2e0 * 157807e0 / temp;
  double a = 2.753e-14 * pow(lambda, 1.5) / pow(1e0 + pow(lambda / 2.740, 0.407), 2.242);
  return a * 1e-6 * 1.;
}

/* chem_utils.c:56-70 */
double orc_alpha_a(double temp) {
  double lambda = 2e0 * 157807e0 / temp;
  double a = 1.269e-13 * pow(lambda, 1.503) / pow(1e0 + pow(lambda / 0.522, 0.470), 1.923);
  return a * 1e-6 * 1.;
}

/* chem_utils.c:75-88 */
double orc_beta(double temp) {
  double T5 = temp / 1e5;
  double b = 1.17e-10 * sqrt(temp) / (1 + sqrt(T5)) * exp(-(157809e0 / temp));
  return b * 1e-6 * 1.;
}

/* chem_utils.c:92-187 with OTSA, WRADTEST (no Compton), CLUMPF=1, yHE=0 */
void orc_cooling(double temp, double x, double nH, double *lambda, double *tcool) {
  double c1, c2, c3, c4, c5, c6, unsurtc, nh2;
  const double CLUMPF = 1.;
  nh2 = nH * 1e-6;
  c1 = exp(-157809.1e0 / temp) * 2.54e-21 * sqrt(temp) / (1. + sqrt(temp / 1e5)) * x * (1. - x) * nh2 * nh2 *
       CLUMPF * (1 + O_yHE);
  c2 = 0.;
  c6 = 3.435e-30 * temp * pow(2e0 * 157807e0 / temp, 1.970e0) /
       pow(1. + (pow(2e0 * 157807e0 / temp / 2.250e0, 0.376e0)), 3.720e0) * x * x * nh2 * nh2 * CLUMPF;
  c3 = exp(-118348e0 / temp) * 7.5e-19 / (1 + sqrt(temp / 1e5)) * x * (1. - x) * nh2 * nh2 * CLUMPF * (1 + O_yHE);
  c4 = 1.42e-27 * 1.5e0 * sqrt(temp) * x * (1. + O_yHE) * nh2 * x * nh2 * (1. + O_yHE) * CLUMPF;
  c5 = 0;
  *lambda = c1 + c2 + c3 + c4 + c5 + c6;
  *lambda = (*lambda) * 1e-7 * 1e6;
  unsurtc = fmax(c1, c2);
  unsurtc = fmax(unsurtc, c3);
  unsurtc = fmax(unsurtc, c4);
  unsurtc = fmax(unsurtc, fabs(c5));
  unsurtc = fmax(unsurtc, c6) * 1e-7;
  *tcool = 1.5e0 * nh2 * (1. + x) * (1 + O_yHE) * O_KBOLTZ * temp / unsurtc;
}

/* chemrad body for one unsplit cell: chem_utils.c:266-791 */
int orc_chem_cell(const orc_params *P, orc_cell *R, double dtnew, double aexporg, int chemonly, int *nsub) {
  const int G = P->ngrp;
  const double hnu0 = 13.6 * 1.6022e-19;
  double egyloc[ORC_MAXGRP], floc[3][ORC_MAXGRP], srcloc[ORC_MAXGRP], et[ORC_MAXGRP], p[ORC_MAXGRP];
  double x0, nH, eint, emin, Cool, tcool, tcool1, dtcool, alpha = 0., alphab, beta, ai_tmp1, xt, eintt;
  double aexp, fudgecool, currentcool_t;
  int nitcool, g;
  const double unit_l = P->unit_l, unit_t = P->unit_t, unit_N = P->unit_N;
  const double unit_n = P->unit_n, unit_d = P->unit_d, unit_v = P->unit_v;

  double dt = dtnew * unit_t * pow(aexporg, 2); /* 273 */

  for (g = 0; g < G; g++) { /* 317-323 */
    egyloc[g] = R->e[g] / (aexporg * aexporg * aexporg) * unit_N;
    floc[0][g] = R->fx[g] / pow(aexporg, 4) * unit_l / unit_t * unit_N;
    floc[1][g] = R->fy[g] / pow(aexporg, 4) * unit_l / unit_t * unit_N;
    floc[2][g] = R->fz[g] / pow(aexporg, 4) * unit_l / unit_t * unit_N;
  }
  x0 = R->nhplus / R->nh;                                                   /* 325 */
  nH = R->nh / (aexporg * aexporg * aexporg) * unit_N;                      /* 328 */
  eint = R->eint / pow(aexporg, 5) * unit_n * unit_d * pow(unit_v, 2);      /* 330 */
  emin = O_PMIN / (O_GAMMA - 1.) / pow(aexporg, 5) * unit_n * unit_d * pow(unit_v, 2); /* 331 */
  for (g = 0; g < G; g++) srcloc[g] = (R->src[g] * unit_N / unit_t / (aexporg * aexporg)) / pow(aexporg, 3);

  if (egyloc[0] < O_EMIN) return ORC_ERR_EMIN; /* 382-385 */

  aexp = aexporg;
  fudgecool = P->fudgecool;
  currentcool_t = 0.;
  nitcool = 0;
  while (currentcool_t < dt) { /* 431 */
    double Nfree = (1. + x0) * nH;
    double tloc = eint / (1.5 * Nfree * O_KBOLTZ);
    int test = 0;
    orc_cooling(tloc, x0, nH, &Cool, &tcool1);
    if (fudgecool < 1e-20) { /* 453-466 */
      if (nsub) *nsub = nitcool;
      return ORC_ERR_NOCONV;
    }
    ai_tmp1 = 0.;
    for (g = 0; g < G; g++) ai_tmp1 += (P->alphae[g] * P->hnu[g] - P->alphai[g] * hnu0) * egyloc[g] * nH * (1.0 - x0);
    tcool = fabs(eint / (ai_tmp1 * (!chemonly) - Cool)); /* 477 */
    ai_tmp1 = 0.;
    dtcool = fmin(fudgecool * tcool, dt - currentcool_t);
    alpha = orc_alpha_a(tloc) * 1.; /* overwritten by alphab under OTSA (500) */
    alphab = orc_alpha_b(tloc) * 1.;
    beta = orc_beta(tloc) * 1.;
    for (g = 0; g < G; g++) { /* 491-555 */
      double factotsa = 0;
      alpha = alphab;
      ai_tmp1 = P->alphai[g] * (1. - x0);
      if (chemonly) {
        et[g] = egyloc[g];
      } else {
        et[g] = ((alpha - alphab) * x0 * x0 * nH * nH * dtcool * factotsa + egyloc[g] + srcloc[g] * dtcool * P->factgrp[g]) /
                (1. + dtcool * (ai_tmp1 * nH));
      }
      if ((et[g] < 0) || (isnan(et[g]))) test = 1;
      p[g] = (1. + (ai_tmp1 * nH) * dtcool);
    }
    if (test) {
      fudgecool = fudgecool / 10.;
      continue;
    }
    ai_tmp1 = 0.; /* 565-570 SEMI_IMPLICIT */
    for (g = 0; g < G; g++) ai_tmp1 += P->alphai[g] * et[g] * (!chemonly);
    {
      double RECT = alpha * x0 * x0 * nH * dtcool; /* 588 */
      xt = 1. - (RECT + (1.0 - x0)) / (1. + dtcool * (beta * x0 * nH + ai_tmp1));
    }
    ai_tmp1 = 0.;
    if (((xt > 1.) || (xt < 0.)) || (isnan(xt))) { /* 599-604 */
      fudgecool /= 10.;
      continue;
    }
    orc_cooling(tloc, xt, nH, &Cool, &tcool1); /* 607 */
    ai_tmp1 = 0.;                              /* 650-659 */
    for (g = 0; g < G; g++) ai_tmp1 += (P->alphae[g] * P->hnu[g] - P->alphai[g] * hnu0) * et[g] * nH * (1.0 - xt);
    eintt = (eint + dtcool * ((ai_tmp1 * (!chemonly)) - Cool));
    if (eintt < 0.) { /* 703-708 */
      fudgecool = fudgecool / 10.;
      continue;
    }
    {
      double deltaE = fabs(eintt / eint - 1.); /* 710-719 */
      if (deltaE > O_FRAC_VAR) {
        fudgecool = fudgecool / 10.;
        continue;
      } else {
        fudgecool = fmin(fudgecool * 1.5, P->fudgecool);
      }
    }
    eintt = fmax(emin, eintt); /* 724 */
    {
      double aold = aexp; /* 737-765 */
      double hubblet = P->H0 * sqrt(P->om / aexp + P->ov * (aexp * aexp)) / aexp * (1e3 / (1e6 * O_PARSEC));
      double da = hubblet * dtcool * aexp;
      aexp += da;
      for (g = 0; g < G; g++) {
        egyloc[g] = et[g] * pow(aold / aexp, 3);
        if (!chemonly) {
          floc[0][g] = floc[0][g] / p[g] * pow(aold / aexp, 4);
          floc[1][g] = floc[1][g] / p[g] * pow(aold / aexp, 4);
          floc[2][g] = floc[2][g] / p[g] * pow(aold / aexp, 4);
        }
      }
      x0 = xt;
      eint = eintt * pow(aold / aexp, 5);
    }
    currentcool_t += dtcool;
    fudgecool = P->fudgecool;
    nitcool++;
    if ((nitcool == P->ncvgcool) && (P->ncvgcool != 0)) break;
  }
  (void)alpha;
  (void)tcool1;

  if (!chemonly) { /* 775-783 */
    for (g = 0; g < G; g++) {
      R->e[g] = fmax(egyloc[g] * pow(aexp, 3) / unit_N, O_EMIN * P->factgrp[g] / unit_N);
      R->fx[g] = floc[0][g] * pow(aexp, 4) / unit_l * unit_t / unit_N;
      R->fy[g] = floc[1][g] * pow(aexp, 4) / unit_l * unit_t / unit_N;
      R->fz[g] = floc[2][g] * pow(aexp, 4) / unit_l * unit_t / unit_N;
    }
  }
  R->nhplus = x0 * R->nh; /* 788-790 */
  R->eint = eint * pow(aexp, 5) / unit_n / unit_d / pow(unit_v, 2);
  orc_e2t(R, aexp, P);
  if (nsub) *nsub = nitcool;
  return ORC_OK;
}

static void load_cell(int G, const orc_fields *f, int i, orc_cell *R) {
  const int n = f->ncell;
  for (int g = 0; g < G; g++) {
    R->e[g] = f->e[g * n + i];
    R->fx[g] = f->fx[g * n + i];
    R->fy[g] = f->fy[g * n + i];
    R->fz[g] = f->fz[g * n + i];
    R->src[g] = f->src[g * n + i];
  }
  R->nhplus = f->nhplus[i];
  R->eint = f->eint[i];
  R->nh = f->nh[i];
  R->temp = f->temp[i];
}

static void store_cell(int G, orc_fields *f, int i, const orc_cell *R) {
  const int n = f->ncell;
  for (int g = 0; g < G; g++) {
    f->e[g * n + i] = R->e[g];
    f->fx[g * n + i] = R->fx[g];
    f->fy[g * n + i] = R->fy[g];
    f->fz[g * n + i] = R->fz[g];
    f->src[g * n + i] = R->src[g];
  }
  f->nhplus[i] = R->nhplus;
  f->eint[i] = R->eint;
  f->nh[i] = R->nh;
  f->temp[i] = R->temp;
}

int orc_chem_cells(const orc_params *P, orc_fields *f, double dtnew, double aexp, int chemonly, int *nsub,
                   int *bad_cell) {
  for (int i = 0; i < f->ncell; i++) {
    orc_cell R;
    int ns = 0;
    load_cell(P->ngrp, f, i, &R);
    int st = orc_chem_cell(P, &R, dtnew, aexp, chemonly, &ns);
    if (nsub) nsub[i] = ns;
    if (st != ORC_OK) {
      if (bad_cell) *bad_cell = i;
      return st;
    }
    store_cell(P->ngrp, f, i, &R);
  }
  return ORC_OK;
}

/* clean_new_rad without MPI (rad_utils.c:2050-2070) */
void orc_clean_new(int G, const orc_fields *old, orc_fields *newf) {
  size_t n = (size_t)G * old->ncell * sizeof(double), m = (size_t)old->ncell * sizeof(double);
  memcpy(newf->e, old->e, n);
  memcpy(newf->fx, old->fx, n);
  memcpy(newf->fy, old->fy, n);
  memcpy(newf->fz, old->fz, n);
  memcpy(newf->src, old->src, n);
  memcpy(newf->nhplus, old->nhplus, m);
  memcpy(newf->eint, old->eint, m);
  memcpy(newf->nh, old->nh, m);
  memcpy(newf->temp, old->temp, m);
}

/* One stencil slot as gathered by gatherstencilrad and
   recursive_neighbor_gather_oct_rad (rad_utils.c:1413-1477, 1482-1534): the
   cell itself, a cell of the same oct, a cell of the same-level neighbour oct
   (with that cell's split flag), or the coarse leaf cell injected into every
   slot with split=0. */
typedef struct {
  double e[ORC_MAXGRP], fx[ORC_MAXGRP], fy[ORC_MAXGRP], fz[ORC_MAXGRP];
  int split;
} slot_state;

static void read_slot(int G, const orc_fields *f, int idx, int split, slot_state *S) {
  const int n = f->ncell;
  for (int g = 0; g < G; g++) {
    S->e[g] = f->e[g * n + idx];
    S->fx[g] = f->fx[g * n + idx];
    S->fy[g] = f->fy[g * n + idx];
    S->fz[g] = f->fz[g * n + idx];
  }
  S->split = split;
}

/* the cell of the boundary oct whose copy the cell `icell` sees across the
   domain face f (the reference's TRANS* face tables, rad_utils.c:1106-1213) */
int orc_trans_source(int f, int icell) { return (f == 3 && icell >= 6) ? icell ^ 1 : icell; }

static int cell_gather(int G, const int *nei_oct, const int *child, const orc_fields *old, const orc_fields *cold,
                       int oct, int icell, slot_state *C, slot_state N[6]) {
  int vnei[6], vcell[6];
  int i = oct * 8 + icell;
  read_slot(G, old, i, child[i] >= 0, C);
  orc_getcellnei(icell, vnei, vcell);
  for (int f = 0; f < 6; f++) {
    if (vnei[f] == 6) {
      int j = oct * 8 + vcell[f];
      read_slot(G, old, j, child[j] >= 0, &N[f]);
    } else {
      int no = nei_oct[oct * 6 + f];
      if (no >= 0) {
        int j = no * 8 + vcell[f];
        read_slot(G, old, j, child[j] >= 0, &N[f]);
      } else if (no <= ORC_NEI_REFLECTIVE) {
        /* domain boundary (rad_utils.c:1106-1213, 1376-1407): the ghost oct
           copies the boundary oct through the reference's face tables, which
           give each cell its own copy except on +y, where cells 6 and 7 see
           each other (face[4] = 7, face[5] = 6, rad_utils.c:1124-1139); the
           normal flux is negated for REF* */
        const int m = orc_trans_source(f, icell);
        read_slot(G, old, oct * 8 + m, child[oct * 8 + m] >= 0, &N[f]);
        if (no == ORC_NEI_REFLECTIVE) {
          for (int g = 0; g < G; g++) {
            if (f >> 1 == 0) N[f].fx[g] *= -1.;
            if (f >> 1 == 1) N[f].fy[g] *= -1.;
            if (f >> 1 == 2) N[f].fz[g] *= -1.;
          }
        }
      } else {
        if (cold == NULL) return ORC_ERR_NEIGHBOR;
        read_slot(G, cold, -1 - no, 0, &N[f]);
      }
    }
  }
  return ORC_OK;
}

static double comp(const slot_state *S, int i, int g) {
  return i == 0 ? S->fx[g] : (i == 1 ? S->fy[g] : S->fz[g]);
}

/* rad_sweepX/Y/Z (rad_utils.c:478-1060) for one cell: 6 faces x G x 5.
   On face f = 2d+s the minus-side state is the neighbour (s=0) or the cell
   (s=1); the Lax-Friedrichs flux in +d is 0.5(fp+fm) + 0.5 c (um-up) with
   p = plus side, m = minus side, zeroed when the neighbour is split. */
static void cell_fluxes(int G, const slot_state *C, const slot_state *N, double c, double *F) {
  for (int f = 0; f < 6; f++) {
    int d = f >> 1;
    const slot_state *M = (f & 1) ? C : &N[f];
    const slot_state *Pp = (f & 1) ? &N[f] : C;
    double ffact = (double)(!N[f].split);
    double *out = F + f * 5 * G;
    for (int g = 0; g < G; g++) {
      double up = Pp->e[g], um = M->e[g];
      double fp = comp(Pp, d, g), fm = comp(M, d, g);
      out[0 + g * 5] = 0.5 * (fp + fm) + 0.5 * c * (um - up);
      for (int i = 0; i < 3; i++) {
        up = comp(Pp, i, g);
        um = comp(M, i, g);
        fp = orc_eddington(Pp->fx[g], Pp->fy[g], Pp->fz[g], Pp->e[g], c, i, d);
        fm = orc_eddington(M->fx[g], M->fy[g], M->fz[g], M->e[g], c, i, d);
        out[1 + i + g * 5] = 0.5 * (fp + fm) + 0.5 * c * (um - up);
      }
      out[4 + g * 5] = 0.;
    }
    for (int k = 0; k < 5 * G; k++) out[k] *= ffact;
  }
}

int orc_level_fluxes(int G, int noct, const int *nei_oct, const int *child, const orc_fields *old,
                     const orc_fields *cold, double c, double *rflux) {
  for (int oct = 0; oct < noct; oct++)
    for (int icell = 0; icell < 8; icell++) {
      slot_state C, N[6];
      int st = cell_gather(G, nei_oct, child, old, cold, oct, icell, &C, N);
      if (st != ORC_OK) return st;
      cell_fluxes(G, &C, N, c, rflux + (size_t)(oct * 8 + icell) * 30 * G);
    }
  return ORC_OK;
}

/* advancerad for one level in RADSTEP mode (rad_utils.c:1775-1857): fluxes
   from the old state, update of rfieldnew (updatefieldrad 1538-1655), chemrad
   on unsplit cells, then scatterstencilrad's 1/8 refluxing into coarse leaf
   neighbours (1738-1762) in oct, cell, face order. */
int orc_level_advance(const orc_params *P, int noct, const int *nei_oct, const int *child,
                      const orc_fields *old, orc_fields *newf, const orc_fields *cold, orc_fields *cnew,
                      double dx, double dt, double aexp, int chemonly, int *nsub, int *bad_cell) {
  return orc_level_advance_mode(P, 0, 1, noct, nei_oct, child, old, newf, cold, cnew, dx, dt, aexp, chemonly, nsub,
                                bad_cell);
}

/* mode 0 = RADSTEP; mode 1 = COARSERAD, where transport runs only on the
   coarse level (is_coarse) and not for chemonly calls (rad_utils.c:1808-1812),
   fluxes of split neighbours are kept (condsplit=1, 518-526), split cells are
   updated too (1548-1552) and nothing is refluxed (1740) */
int orc_level_advance_mode(const orc_params *P, int mode, int is_coarse, int noct, const int *nei_oct,
                           const int *child, const orc_fields *old, orc_fields *newf, const orc_fields *cold,
                           orc_fields *cnew, double dx, double dt, double aexp, int chemonly, int *nsub,
                           int *bad_cell) {
  const int G = P->ngrp;
  const int condadvec = (mode == 0) ? 1 : (is_coarse && !chemonly);
  const double c = orc_cloc(aexp, P);
  const double dtsurdx = dt / dx;
  double F[30 * ORC_MAXGRP];
  for (int oct = 0; oct < noct; oct++) {
    for (int icell = 0; icell < 8; icell++) {
      int i = oct * 8 + icell;
      slot_state C, N[6];
      if (condadvec) {
        int st = cell_gather(G, nei_oct, child, old, cold, oct, icell, &C, N);
        if (st != ORC_OK) {
          if (bad_cell) *bad_cell = i;
          return st;
        }
        if (mode == 1)
          for (int f = 0; f < 6; f++) N[f].split = 0;
        cell_fluxes(G, &C, N, c, F);
      }
      orc_cell R;
      int updated = 0;
      if (condadvec && (mode == 1 || child[i] < 0)) { /* updatefieldrad */
        double Re[ORC_MAXGRP] = {0}, Rx[ORC_MAXGRP] = {0}, Ry[ORC_MAXGRP] = {0}, Rz[ORC_MAXGRP] = {0};
        double one = 1.;
        for (int flx = 0; flx < 6; flx++) { /* 1566-1574 */
          for (int g = 0; g < G; g++) {
            Re[g] += F[0 + g * 5 + flx * 5 * G] * dtsurdx * one;
            Rx[g] += F[1 + g * 5 + flx * 5 * G] * dtsurdx * one;
            Ry[g] += F[2 + g * 5 + flx * 5 * G] * dtsurdx * one;
            Rz[g] += F[3 + g * 5 + flx * 5 * G] * dtsurdx * one;
          }
          one *= -1.;
        }
        load_cell(G, newf, i, &R);
        for (int g = 0; g < G; g++) {
          R.e[g] += Re[g];
          R.fx[g] += Rx[g];
          R.fy[g] += Ry[g];
          R.fz[g] += Rz[g];
        }
        /* 1611-1650: negative or NaN photon density aborts the reference
           (its second test reads e[1], which is fx[0] when NGRP=1) */
        if (R.e[0] < 0 || isnan(R.e[0]) || isnan(G > 1 ? R.e[1] : R.fx[0])) {
          if (bad_cell) *bad_cell = i;
          return ORC_ERR_NEG_PHOTON;
        }
        updated = 1;
      } else {
        load_cell(G, newf, i, &R);
      }
      if (child[i] >= 0) { /* chemrad skips split cells (chem_utils.c:302) */
        if (nsub) nsub[i] = 0;
        if (updated) store_cell(G, newf, i, &R);
      } else {
        int ns = 0;
        int st = orc_chem_cell(P, &R, dt, aexp, chemonly, &ns);
        if (nsub) nsub[i] = ns;
        if (st != ORC_OK) {
          if (bad_cell) *bad_cell = i;
          return st;
        }
        store_cell(G, newf, i, &R);
      }
      if (mode == 0 && cnew != NULL) { /* refluxing, 1738-1762 */
        int vnei[6], vcell[6];
        orc_getcellnei(icell, vnei, vcell);
        for (int inei = 0; inei < 6; inei++) {
          if (vnei[inei] == 6) continue;
          int no = nei_oct[oct * 6 + vnei[inei]];
          if (no >= 0 || no <= ORC_NEI_REFLECTIVE) continue; /* TRANS* faces: rad_utils.c:1700-1734 */
          int ci = -1 - no, nc = cnew->ncell;
          double one = pow(-1., inei + 1);
          const double *Ff = F + inei * 5 * G;
          for (int g = 0; g < G; g++) {
            cnew->e[g * nc + ci] += Ff[0 + g * 5] * dtsurdx * one * 0.125;
            cnew->fx[g * nc + ci] += Ff[1 + g * 5] * dtsurdx * one * 0.125;
            cnew->fy[g * nc + ci] += Ff[2 + g * 5] * dtsurdx * one * 0.125;
            cnew->fz[g * nc + ci] += Ff[3 + g * 5] * dtsurdx * one * 0.125;
          }
        }
      }
    }
  }
  return ORC_OK;
}

/* RadSolver final update (rad_utils.c:1891-2006), RADSTEP mode */
void orc_level_final(int G, int noct, const int *child, orc_fields *old, const orc_fields *newf,
                     const orc_fields *fine_old, double tphy, double *t_first, double *t_last,
                     const double *fine_t_first, const double *fine_t_last) {
  orc_level_final_mode(G, 0, 0, 0, noct, child, old, newf, fine_old, tphy, t_first, t_last, fine_t_first,
                       fine_t_last);
}

/* COARSERAD (mode 1): on the coarse level, unless chemonly, every cell takes
   its rfieldnew (rad_utils.c:1917-1929); the push of the coarse values down
   to the leaves (distribE) is orc_distribE */
void orc_level_final_mode(int G, int mode, int is_coarse, int chemonly, int noct, const int *child, orc_fields *old,
                          const orc_fields *newf, const orc_fields *fine_old, double tphy, double *t_first,
                          double *t_last, const double *fine_t_first, const double *fine_t_last) {
  const int ncell = noct * 8;
  for (int i = 0; i < ncell; i++) {
    if (t_first && t_last) { /* 1901-1913 */
      double xion = old->nhplus[i] / old->nh[i];
      if (xion < 0.5) t_last[i] = tphy;
      if ((xion >= 0.5) && (t_first[i] == -1)) t_first[i] = tphy;
    }
    orc_cell R;
    const int unsplit = (mode == 0 || chemonly) ? (child[i] < 0) : ((child[i] < 0) || is_coarse);
    if (unsplit) { /* 1934-1940 */
      load_cell(G, newf, i, &R);
      for (int g = 0; g < G; g++) R.src[g] = old->src[g * ncell + i];
      R.nh = old->nh[i];
    } else { /* 1955-2006 */
      const int co = child[i], nf = fine_old->ncell;
      memset(&R, 0, sizeof(R));
      if (t_first && t_last) {
        t_first[i] = 0;
        t_last[i] = 0;
      }
      for (int k = 0; k < 8; k++) {
        int j = co * 8 + k;
        for (int g = 0; g < G; g++) {
          R.e[g] += fine_old->e[g * nf + j] * 0.125;
          R.fx[g] += fine_old->fx[g * nf + j] * 0.125;
          R.fy[g] += fine_old->fy[g * nf + j] * 0.125;
          R.fz[g] += fine_old->fz[g * nf + j] * 0.125;
          R.src[g] += fine_old->src[g * nf + j] * 0.125;
        }
        if (t_first && t_last && fine_t_first && fine_t_last) {
          t_first[i] += fine_t_first[j] * 0.125;
          t_last[i] += fine_t_last[j] * 0.125;
        }
        R.nhplus += fine_old->nhplus[j] * 0.125;
        R.eint += fine_old->eint[j] * 0.125;
        R.temp += fine_old->temp[j] * 0.125;
        R.nh += fine_old->nh[j] * 0.125;
      }
    }
    store_cell(G, old, i, &R);
  }
}

/* vectorised helpers for the tests */
void orc_eddington_n(int n, const double *fx, const double *fy, const double *fz, const double *e, double c,
                     const int *i, const int *j, double *out) {
  for (int k = 0; k < n; k++) out[k] = orc_eddington(fx[k], fy[k], fz[k], e[k], c, i[k], j[k]);
}

void orc_rates_n(int n, const double *temp, const double *x, const double *nH, double *alpha_a, double *alpha_b,
                 double *beta, double *lambda, double *tcool) {
  for (int k = 0; k < n; k++) {
    alpha_a[k] = orc_alpha_a(temp[k]);
    alpha_b[k] = orc_alpha_b(temp[k]);
    beta[k] = orc_beta(temp[k]);
    orc_cooling(temp[k], x[k], nH[k], &lambda[k], &tcool[k]);
  }
}

void orc_e2t_n(const orc_params *P, int n, const double *nhplus, const double *eint, const double *nh, double aexp,
               double *temp) {
  for (int k = 0; k < n; k++) {
    orc_cell R;
    memset(&R, 0, sizeof(R));
    R.nhplus = nhplus[k];
    R.eint = eint[k];
    R.nh = nh[k];
    orc_e2t(&R, aexp, P);
    temp[k] = R.temp;
  }
}

/* ---- COARSERAD transfers between the coarse level and the leaves ----
   Levels are passed as arrays indexed by level - lcoarse. */

/* collectE (src_utils.c:94-123): depth-first over the subtree, leaves add
   their E, F (and nhplus, eint) times the volume fraction 2^(3(lcoarse-l)) */
static void collect_rec(int G, int lcoarse, int l, int nlev, int cell, const int *const *child,
                        const orc_fields *const *F, orc_fields *C, int ci) {
  const int k = l - lcoarse;
  const int ch = child[k][cell];
  if (ch < 0) {
    const double fact = pow(2., 3 * (lcoarse - l));
    const orc_fields *L = F[k];
    const int n = L->ncell, nc = C->ncell;
    for (int g = 0; g < G; g++) {
      C->e[g * nc + ci] += L->e[g * n + cell] * fact;
      C->fx[g * nc + ci] += L->fx[g * n + cell] * fact;
      C->fy[g * nc + ci] += L->fy[g * n + cell] * fact;
      C->fz[g * nc + ci] += L->fz[g * n + cell] * fact;
    }
    C->nhplus[ci] += L->nhplus[cell] * fact;
    C->eint[ci] += L->eint[cell] * fact;
  } else if (k + 1 < nlev) {
    for (int c = 0; c < 8; c++) collect_rec(G, lcoarse, l + 1, nlev, ch * 8 + c, child, F, C, ci);
  }
}

/* putE2coarse (src_utils.c:127-150) for every split cell of the coarse level,
   as FillRad does under COARSERAD (src_utils.c:717-733) */
void orc_putE2coarse(int G, int lcoarse, int nlev, const int *const *child, orc_fields *const *old,
                     orc_fields *newc) {
  orc_fields *C = old[0];
  const int nc = C->ncell;
  for (int i = 0; i < nc; i++) {
    if (child[0][i] < 0) continue;
    for (int g = 0; g < G; g++) {
      C->e[g * nc + i] = 0.;
      C->fx[g * nc + i] = 0.;
      C->fy[g * nc + i] = 0.;
      C->fz[g * nc + i] = 0.;
    }
    for (int c = 0; c < 8; c++)
      if (nlev > 1) collect_rec(G, lcoarse, lcoarse + 1, nlev, child[0][i] * 8 + c, child, (const orc_fields *const *)old, C, i);
    for (int g = 0; g < G; g++) {
      newc->e[g * nc + i] = C->e[g * nc + i];
      newc->fx[g * nc + i] = C->fx[g * nc + i];
      newc->fy[g * nc + i] = C->fy[g * nc + i];
      newc->fz[g * nc + i] = C->fz[g * nc + i];
    }
  }
}

/* distribE (rad_utils.c:26-45): every leaf below a split coarse cell takes the
   coarse cell's E and F */
static void distrib_rec(int G, int lcoarse, int l, int nlev, int cell, const int *const *child, orc_fields *const *F,
                        const orc_fields *C, int ci) {
  const int k = l - lcoarse;
  const int ch = child[k][cell];
  if (ch < 0) {
    orc_fields *L = F[k];
    const int n = L->ncell, nc = C->ncell;
    for (int g = 0; g < G; g++) {
      L->e[g * n + cell] = C->e[g * nc + ci];
      L->fx[g * n + cell] = C->fx[g * nc + ci];
      L->fy[g * n + cell] = C->fy[g * nc + ci];
      L->fz[g * n + cell] = C->fz[g * nc + ci];
    }
  } else if (k + 1 < nlev) {
    for (int c = 0; c < 8; c++) distrib_rec(G, lcoarse, l + 1, nlev, ch * 8 + c, child, F, C, ci);
  }
}

void orc_distribE(int G, int lcoarse, int nlev, const int *const *child, orc_fields *const *old) {
  const orc_fields *C = old[0];
  for (int i = 0; i < C->ncell; i++) {
    if (child[0][i] < 0) continue;
    for (int c = 0; c < 8; c++)
      if (nlev > 1) distrib_rec(G, lcoarse, lcoarse + 1, nlev, child[0][i] * 8 + c, child, old, C, i);
  }
}

/* ---------------------------------------------------------------- FillRad
   src_utils.c:692-849 with init = 0 (the COARSERAD block is orc_putE2coarse).
   Constants: param.h:31 (MYR), 37 (SOLAR_MASS). */
#define O_MYR (3.1536e13)
#define O_SOLAR_MASS (1.989e30)

/* stromgren_source (src_utils.c:339-412) for one cell; src already cleaned
   to 0 by putsource (src_utils.c:450-455) */
static void orc_stromgren(const orc_params *P, const orc_star *S, int level, int lcoarse, const double *octpos,
                          int icell, double tcur, double aexp, double *src, int ncell, int cell) {
  const double X0 = 1. / (double)(1ULL << lcoarse);
  const double tcur_in_yrs = tcur * P->unit_t / O_MYR * 1e6;
  const double dxcur = pow(0.5, level);
  const double xc = octpos[0] + (icell & 1) * dxcur + dxcur * 0.5;
  const double yc = octpos[1] + ((icell >> 1) & 1) * dxcur + dxcur * 0.5;
  const double zc = octpos[2] + ((icell >> 2)) * dxcur + dxcur * 0.5;
  const int nsub = 1;
  const double R = sqrt(pow(xc - S->x, 2) + pow(yc - S->y, 2) + pow(zc - S->z, 2));
  if ((R <= X0) && (fmod(tcur_in_yrs, S->lifetime) < S->lifetime / nsub)) {
    if ((xc > 0.) && (yc > 0.) && (zc > 0.)) {
      /* the NGRP_SPACE groups of the time group tcur lies in (src_utils.c:
         366-392); the bounds as readAtomic leaves them (parameters.c:455-468) */
      const int nspace = P->ngrp / P->ngrp_time;
      for (int it = 0; it < P->ngrp_time; it++) {
        if ((tcur_in_yrs >= P->time_bound[it]) && (tcur_in_yrs < P->time_bound[it + 1])) {
          for (int is = 0; is < nspace; is++) {
            const int g = it * nspace + is;
            const double srcint = S->mass * O_SOLAR_MASS * S->srcint * (tcur > 0 ? 1. : 0) * nsub / 8;
            src[g * ncell + cell] = srcint / pow(X0 * P->unit_l, 3) * P->unit_t / P->unit_N * pow(aexp, 2);
          }
        }
      }
    }
  } else {
    for (int g = 0; g < P->ngrp; g++) src[g * ncell + cell] = 0.;
  }
}

/* stars_sources (src_utils.c:209-335) for one unsplit cell: clear every
   group, then each particle of the cell with isStar >= 6 adds to the
   NGRP_SPACE groups of its radiative_state (isStar == 100 is the AGN case,
   empty without AGN); *k walks the cell-sorted particle list */
static void orc_stars_cell(const orc_params *P, const orc_stars *S, int level, double tcur, double aexp, double *src,
                           int ncell, int cell, int *k) {
  const double dxcur = pow(0.5, level);
  const int nspace = P->ngrp / P->ngrp_time;
  for (int g = 0; g < P->ngrp; g++) src[g * ncell + cell] = 0.;
  while (*k < S->nstar && S->cell[*k] < cell) (*k)++;
  for (; *k < S->nstar && S->cell[*k] == cell; (*k)++) {
    const int i = *k;
    if (!(S->is_star[i] >= 6) || S->is_star[i] == 100) continue;
    if (!S->decrease || (tcur - S->age[i]) < S->tlife) {
      const double t = (tcur - S->age[i]) / S->tlife;
      for (int ig = 0; ig < nspace; ig++)
        src[(S->radiative_state[i] * nspace + ig) * ncell + cell] += (S->mass[i] * P->unit_d) * S->srcint /
                                                                      pow(dxcur, 3) * P->unit_t / P->unit_N *
                                                                      pow(aexp, 2) * (t > 0 ? 1. : 0.);
    } else { /* DECREASE_EMMISIVITY_AFTER_TLIFE, src_utils.c:301-323 */
      const double slope = -4.;
      const double s = (S->mass[i] * P->unit_d) * S->srcint / pow(dxcur, 3) * P->unit_t / P->unit_N * pow(aexp, 2);
      const double t = (tcur - S->age[i]) / S->tlife;
      for (int ig = 0; ig < nspace; ig++)
        src[(S->radiative_state[i] * nspace + ig) * ncell + cell] += s * (t < 1. ? 1. : pow(t, slope)) * (t > 0 ? 1 : 0);
    }
  }
}

int orc_fillrad_cells(const orc_params *P, int level, int lcoarse, orc_fields *F, const int *child,
                      const double *opos, const orc_hydro *H, const orc_star *star, const orc_stars *stars,
                      double tcur, double aexp) {
  const int n = F->ncell;
  int kstar = 0;
  for (int cell = 0; cell < n; cell++) {
    if (H != NULL) {
      F->nh[cell] = H->d[cell] * (1. - O_YHE);
      F->eint[cell] = H->p[cell] / (O_GAMMA - 1.);
      F->nhplus[cell] = H->dX[cell];
    }
    if (child[cell] >= 0) {
      for (int g = 0; g < P->ngrp; g++) F->src[g * n + cell] = 0.;
      continue;
    }
    if (star != NULL && star->on) {
      for (int g = 0; g < P->ngrp; g++) F->src[g * n + cell] = 0.;
      orc_stromgren(P, star, level, lcoarse, opos + 3 * (cell / 8), cell % 8, tcur, aexp, F->src, n, cell);
    } else if (stars != NULL && stars->on) {
      orc_stars_cell(P, stars, level, tcur, aexp, F->src, n, cell, &kstar);
    }
    if (H != NULL) {
      orc_cell R;
      R.nhplus = F->nhplus[cell];
      R.eint = F->eint[cell];
      R.nh = F->nh[cell];
      orc_e2t(&R, aexp, P);
      F->temp[cell] = R.temp;
    }
  }
  return ORC_OK;
}

/* rad_utils.c:1941-1951 (unsplit), 1993-2004 (split, after the restriction) */
void orc_rad_to_hydro(const orc_fields *F, orc_hydro *H) {
  for (int i = 0; i < F->ncell; i++) {
    H->p[i] = (O_GAMMA - 1.) * F->eint[i];
    H->dX[i] = F->nhplus[i] / F->nh[i] * H->d[i] * (1. - O_YHE);
    /* getE, hydro_utils.c:29-31 */
    H->E[i] = H->p[i] / (O_GAMMA - 1.) + 0.5 * H->d[i] * (H->u[i] * H->u[i] + H->v[i] * H->v[i] + H->w[i] * H->w[i]);
  }
}

