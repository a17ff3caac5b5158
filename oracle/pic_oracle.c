/*
 * pic_oracle.c -- CPU FP64 restatement of ALaDyn's LP PIC step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never part of
 * the product path.  Each routine cites the reference routine it restates and
 * keeps the reference's operation order, so that compiled without FMA
 * contraction (-ffp-contract=off) it reproduces the reference's rounding.
 * Array layouts and the packed weight slot are described in pic_oracle.h.
 */
#include "pic_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Fortran ebf(i,j,k,c) / jc(i,j,k,c) with 1-based indices. */
#define F4(a, i, j, k, c)                                                    \
  (a)[((((int64_t)(c)-1) * g->n3p + ((k)-1)) * (int64_t)g->n2p + ((j)-1)) * \
          (int64_t)g->n1p +                                                  \
      ((i)-1)]
/* Fortran part(n,c) with 0-based n and 1-based column c. */
#define PC(a, n, c) (a)[((int64_t)(c)-1) * ld + (n)]

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

void orc_grid_setup(orc_grid *g, double cfl) {
  /* set_grid (set_grid_param.f90:52-225): x(i) = dx*(i-1),
   * y(i) = dy*(i-1-ny/2), z likewise; interior index 3..n+2 (code_util.f90:37,
   * set_fyzxgrid p_ind = min(sh, n_loc) .. n_loc+p_ind-1). */
  if (g->ndim == 2) g->nz = 1;
  g->nfield = g->ndim == 3 ? 6 : 3;
  g->curr_ndim = g->ndim == 3 ? 3 : 2;
  g->n1p = g->nx + 5;
  g->n2p = g->ny + 5;
  g->n3p = g->ndim == 3 ? g->nz + 5 : 1;
  g->ix1 = 3;
  g->ix2 = g->nx + 2;
  g->jy1 = 3;
  g->jy2 = g->ny + 2;
  if (g->ndim == 3) {
    g->kz1 = 3;
    g->kz2 = g->nz + 2;
  } else {
    g->kz1 = 1;
    g->kz2 = 1;
  }
  g->dx_inv = 1.0 / g->dx;
  g->dy_inv = 1.0 / g->dy;
  g->dz_inv = g->ndim == 3 ? 1.0 / g->dz : 1.0;
  if (g->ndim == 2) g->dz = 1.0;
  g->xmn = 0.0;
  g->xmx = g->dx * (double)g->nx;
  g->ymn = g->dy * (double)(-(g->ny / 2));
  g->ymx = g->dy * (double)(g->ny - g->ny / 2);
  if (g->ndim == 3) {
    g->zmn = g->dz * (double)(-(g->nz / 2));
    g->zmx = g->dz * (double)(g->nz - g->nz / 2);
  } else {
    g->zmn = 0.0;
    g->zmx = 0.0;
  }
  /* djc (set_init_param.f90:92-102) */
  g->djc[0] = g->dx;
  g->djc[1] = 0.0;
  g->djc[2] = 0.0;
  if (g->ndim == 2) {
    g->djc[0] = 0.5 * g->dx;
    g->djc[1] = 0.5 * g->dy;
  } else if (g->ndim == 3) {
    g->djc[0] = g->dx / 3.;
    g->djc[1] = g->dy / 3.;
    g->djc[2] = g->dz / 3.;
  }
  /* se_coeff (grid_field_param.f90:36-99) */
  double yx_rat = g->dy / g->dx;
  double nu = cfl;
  if (g->ndim > 1)
    nu = cfl * yx_rat / sqrt(yx_rat * yx_rat + (double)g->ndim - 1.);
  switch (g->der_ord) {
    case 3:
      g->se_coeff[0] = 1. + 0.125 * (1. - nu * nu);
      g->se_coeff[1] = (1. - g->se_coeff[0]) / 3.;
      break;
    case 4:
      g->se_coeff[0] = 1.125;
      g->se_coeff[1] = (1. - g->se_coeff[0]) / 3.;
      break;
    default:
      g->der_ord = 2;
      g->se_coeff[0] = 1.;
      g->se_coeff[1] = 0.;
      break;
  }
}

double orc_pack_weight(int charge, int part_ind, float wgh) {
  unsigned char b[8];
  int16_t c = (int16_t)charge, p = (int16_t)part_ind;
  memcpy(b, &c, 2);
  memcpy(b + 2, &p, 2);
  memcpy(b + 4, &wgh, 4);
  double w;
  memcpy(&w, b, 8);
  return w;
}

void orc_unpack_weight(double w, int *charge, int *part_ind, float *wgh) {
  unsigned char b[8];
  memcpy(b, &w, 8);
  int16_t c, p;
  float f;
  memcpy(&c, b, 2);
  memcpy(&p, b + 2, 2);
  memcpy(&f, b + 4, 4);
  if (charge) *charge = c;
  if (part_ind) *part_ind = p;
  if (wgh) *wgh = f;
}

/* qlh_*_spline, one axis (grid_part_lib.f90:110-143, 212-259): quadratic
 * weights at integer nodes, linear at half-integer nodes, base i = ix-1. */
static inline void qlh_axis(double xp, double a[3], double ah[2], int *i) {
  double xx = 3.0 + xp; /* shx = 3 (grid_part_lib.f90:32) */
  int ix = (int)(xx + 0.5);
  double sx = xx - (double)ix;
  double sx2 = sx * sx;
  a[1] = 0.75 - sx2;
  a[2] = 0.5 * (0.25 + sx2 + sx);
  a[0] = 1. - a[1] - a[2];
  ah[1] = sx + 0.5;
  ah[0] = 1. - ah[1];
  *i = ix - 1;
}

/* qden_*_wgh, one axis (grid_part_lib.f90:337-358, 388-417): ix NOT shifted. */
static inline void qden_axis(double xp, double a[3], int *ix) {
  double xx = 3.0 + xp;
  int i = (int)(xx + 0.5);
  double sx = xx - (double)i;
  double sx2 = sx * sx;
  a[1] = 0.75 - sx2;
  a[2] = 0.5 * (0.25 + sx2 + sx);
  a[0] = 1. - a[1] - a[2];
  *ix = i;
}

void orc_gather(const orc_grid *g, int64_t np, const double *part, int64_t ld,
                const double *ebf, double *ebfp) {
  double ax1[3], axh[2], ay1[3], ayh[2], az1[3], azh[2];
  int i, j, k;
  if (g->ndim == 2) {
    /* set_part2d_hcell_acc, nfield = 3 (grid_part_connect.f90:104-142) */
    for (int64_t n = 0; n < np; n++) {
      double xl = g->dx_inv * (PC(part, n, 1) - g->xmn);
      double yl = g->dy_inv * (PC(part, n, 2) - g->ymn);
      qlh_axis(xl, ax1, axh, &i);
      qlh_axis(yl, ay1, ayh, &j);
      const int ih = i, jh = j;
      double ap1 = 0., ap2 = 0., ap3 = 0.;
      for (int j1 = 0; j1 <= 2; j1++) {
        int j2 = j + j1;
        double dvol = ay1[j1];
        for (int i1 = 0; i1 <= 1; i1++) {
          int i2 = i1 + ih;
          double dvol1 = axh[i1] * dvol;
          ap1 = ap1 + dvol1 * F4(ebf, i2, j2, 1, 1);
        }
      }
      for (int j1 = 0; j1 <= 1; j1++) {
        int j2 = jh + j1;
        double dvol = ayh[j1];
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i + i1;
          double dvol1 = ax1[i1] * dvol;
          ap2 = ap2 + dvol1 * F4(ebf, i2, j2, 1, 2);
        }
        for (int i1 = 0; i1 <= 1; i1++) {
          int i2 = i1 + ih;
          double dvol1 = axh[i1] * dvol;
          ap3 = ap3 + dvol1 * F4(ebf, i2, j2, 1, 3);
        }
      }
      PC(ebfp, n, 1) = ap1;
      PC(ebfp, n, 2) = ap2;
      PC(ebfp, n, 3) = ap3;
    }
    return;
  }
  /* set_part3d_hcell_acc (grid_part_connect.f90:187-285) */
  for (int64_t n = 0; n < np; n++) {
    double xl = g->dx_inv * (PC(part, n, 1) - g->xmn);
    double yl = g->dy_inv * (PC(part, n, 2) - g->ymn);
    double zl = g->dz_inv * (PC(part, n, 3) - g->zmn);
    qlh_axis(xl, ax1, axh, &i);
    qlh_axis(yl, ay1, ayh, &j);
    qlh_axis(zl, az1, azh, &k);
    const int ih = i, jh = j, kh = k;
    double ap[6] = {0., 0., 0., 0., 0., 0.};
    for (int k1 = 0; k1 <= 2; k1++) {
      int k2 = k + k1;
      for (int j1 = 0; j1 <= 2; j1++) {
        int j2 = j + j1;
        double dvol = ay1[j1] * az1[k1];
        for (int i1 = 0; i1 <= 1; i1++) {
          int i2 = i1 + ih;
          ap[0] = ap[0] + axh[i1] * dvol * F4(ebf, i2, j2, k2, 1);
        }
      }
      for (int j1 = 0; j1 <= 1; j1++) {
        int j2 = jh + j1;
        double dvol = ayh[j1] * az1[k1];
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i + i1;
          ap[1] = ap[1] + ax1[i1] * dvol * F4(ebf, i2, j2, k2, 2);
        }
        for (int i1 = 0; i1 <= 1; i1++) {
          int i2 = i1 + ih;
          ap[5] = ap[5] + axh[i1] * dvol * F4(ebf, i2, j2, k2, 6);
        }
      }
    }
    for (int k1 = 0; k1 <= 1; k1++) {
      int k2 = kh + k1;
      for (int j1 = 0; j1 <= 1; j1++) {
        int j2 = jh + j1;
        double dvol = ayh[j1] * azh[k1];
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i1 + i;
          ap[3] = ap[3] + ax1[i1] * dvol * F4(ebf, i2, j2, k2, 4);
        }
      }
      for (int j1 = 0; j1 <= 2; j1++) {
        int j2 = j + j1;
        double dvol = ay1[j1] * azh[k1];
        for (int i1 = 0; i1 <= 1; i1++) {
          int i2 = ih + i1;
          ap[4] = ap[4] + axh[i1] * dvol * F4(ebf, i2, j2, k2, 5);
        }
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i1 + i;
          ap[2] = ap[2] + ax1[i1] * dvol * F4(ebf, i2, j2, k2, 3);
        }
      }
    }
    for (int c = 0; c < 6; c++) PC(ebfp, n, c + 1) = ap[c];
  }
}

void orc_charge_multiply(const orc_grid *g, int64_t np, const double *part,
                         int64_t ld, double *ebfp) {
  const int ch = 2 * g->curr_ndim + 1; /* size(apt, 2) */
  for (int64_t n = 0; n < np; n++) {
    int charge;
    orc_unpack_weight(PC(part, n, ch), &charge, NULL, NULL);
    for (int c = 1; c <= g->nfield; c++)
      PC(ebfp, n, c) = (double)charge * PC(ebfp, n, c);
  }
}

void orc_init_momenta(const orc_grid *g, int64_t np, double *part, int64_t ld,
                      const double *ebfp) {
  const double dth_lp = 0.5 * g->dt;
  const double alp = dth_lp * g->lorentz_fact;
  if (g->curr_ndim == 2) {
    for (int64_t n = 0; n < np; n++) {
      double efp1 = -alp * PC(ebfp, n, 1);
      double efp2 = -alp * PC(ebfp, n, 2);
      double efp3 = -alp * PC(ebfp, n, 3);
      double pp1 = PC(part, n, 3), pp2 = PC(part, n, 4);
      double gam2 = 1. + (pp1 * pp1 + pp2 * pp2);
      double gam_inv = 1. / sqrt(gam2);
      double vp1 = pp1 * gam_inv, vp2 = pp2 * gam_inv;
      PC(part, n, 3) = PC(part, n, 3) + efp1 + vp2 * efp3;
      PC(part, n, 4) = PC(part, n, 4) + efp2 - vp1 * efp3;
    }
    return;
  }
  for (int64_t n = 0; n < np; n++) {
    double pp1 = PC(part, n, 4), pp2 = PC(part, n, 5), pp3 = PC(part, n, 6);
    double efp[6];
    for (int c = 0; c < 6; c++) efp[c] = -alp * PC(ebfp, n, c + 1);
    double gam2 = 1. + (pp1 * pp1 + pp2 * pp2 + pp3 * pp3);
    double gam_inv = 1. / sqrt(gam2);
    double vp1 = gam_inv * pp1, vp2 = gam_inv * pp2, vp3 = gam_inv * pp3;
    PC(part, n, 4) = PC(part, n, 4) + efp[0] + vp2 * efp[5] - vp3 * efp[4];
    PC(part, n, 5) = PC(part, n, 5) + efp[1] + vp3 * efp[3] - vp1 * efp[5];
    PC(part, n, 6) = PC(part, n, 6) + efp[2] + vp1 * efp[4] - vp2 * efp[3];
  }
}

void orc_push(const orc_grid *g, int64_t np, double *part, int64_t ld,
              double *ebfp) {
  const double dt_lp = g->dt;
  const double dth_lp = 0.5 * dt_lp;
  const double alp = dth_lp * g->lorentz_fact;
  int ch;
  if (g->curr_ndim == 2) {
    ch = 5;
    for (int64_t n = 0; n < np; n++) {
      double pp1 = PC(part, n, 3), pp2 = PC(part, n, 4);
      double efp1 = alp * PC(ebfp, n, 1), efp2 = alp * PC(ebfp, n, 2),
             efp3 = alp * PC(ebfp, n, 3);
      double vp1 = pp1 + efp1, vp2 = pp2 + efp2, vp3 = efp3;
      double gam02 = 1. + (vp1 * vp1 + vp2 * vp2);
      double b2 = vp3 * vp3;
      gam02 = gam02 - b2;
      double gam2 = 0.5 * (gam02 + sqrt(gam02 * gam02 + 4. * b2));
      double gam = sqrt(gam2);
      double vph1 = gam2 * vp1 + gam * vp2 * vp3;
      double vph2 = gam2 * vp2 - gam * vp1 * vp3;
      vph1 = vph1 / (gam2 + b2);
      vph2 = vph2 / (gam2 + b2);
      PC(part, n, 3) = 2. * vph1 - pp1;
      PC(part, n, 4) = 2. * vph2 - pp2;
      PC(ebfp, n, 3) = PC(part, n, 1);
      PC(ebfp, n, 4) = PC(part, n, 2);
      pp1 = PC(part, n, 3);
      pp2 = PC(part, n, 4);
      gam2 = 1. + (pp1 * pp1 + pp2 * pp2);
      PC(ebfp, n, 5) = dt_lp / sqrt(gam2);
      vp1 = PC(ebfp, n, 5) * pp1;
      vp2 = PC(ebfp, n, 5) * pp2;
      PC(ebfp, n, 1) = vp1;
      PC(ebfp, n, 2) = vp2;
      PC(part, n, 1) = PC(part, n, 1) + vp1;
      PC(part, n, 2) = PC(part, n, 2) + vp2;
    }
  } else {
    ch = 7;
    for (int64_t n = 0; n < np; n++) {
      double pp[3] = {PC(part, n, 4), PC(part, n, 5), PC(part, n, 6)};
      double efp[6];
      for (int c = 0; c < 6; c++) efp[c] = alp * PC(ebfp, n, c + 1);
      double vp[3] = {pp[0] + efp[0], pp[1] + efp[1], pp[2] + efp[2]};
      double bb[3] = {efp[3], efp[4], efp[5]};
      double gam02 = 1. + (vp[0] * vp[0] + vp[1] * vp[1] + vp[2] * vp[2]);
      double b2 = bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2];
      double bv = bb[0] * vp[0] + bb[1] * vp[1] + bb[2] * vp[2];
      gam02 = gam02 - b2;
      double gam2 =
          0.5 * (gam02 + sqrt(gam02 * gam02 + 4. * (b2 + bv * bv)));
      double gam = sqrt(gam2);
      double vph[3];
      for (int c = 0; c < 3; c++) vph[c] = gam2 * vp[c] + bb[c] * bv;
      vph[0] = vph[0] + gam * (vp[1] * bb[2] - vp[2] * bb[1]);
      vph[1] = vph[1] + gam * (vp[2] * bb[0] - vp[0] * bb[2]);
      vph[2] = vph[2] + gam * (vp[0] * bb[1] - vp[1] * bb[0]);
      for (int c = 0; c < 3; c++) vph[c] = vph[c] / (b2 + gam2);
      for (int c = 0; c < 3; c++) PC(part, n, 4 + c) = 2. * vph[c] - pp[c];
      for (int c = 0; c < 3; c++) PC(ebfp, n, 4 + c) = PC(part, n, 1 + c);
      for (int c = 0; c < 3; c++) pp[c] = PC(part, n, 4 + c);
      gam2 = 1. + (pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2]);
      PC(ebfp, n, 7) = dt_lp / sqrt(gam2);
      for (int c = 0; c < 3; c++) {
        double v = PC(ebfp, n, 7) * pp[c];
        PC(ebfp, n, 1 + c) = v;
        PC(part, n, 1 + c) = PC(part, n, 1 + c) + v;
      }
    }
  }
  /* iform < 2: the packed weight is kept in the stage buffer */
  for (int64_t n = 0; n < np; n++) PC(ebfp, n, ch) = PC(part, n, ch);
}

/* reset_all_part_dist for one axis (mpi_part_interface.f90:467-523). */
static int64_t part_dist_axis(int64_t np, double *part, int64_t ld,
                              double *ebfp, int ndv, int cin, double xl,
                              double xr, int ib) {
  if (ib == 2) {
    double dxp = xr - xl;
    for (int64_t n = 0; n < np; n++) {
      double xp = PC(part, n, cin);
      if (xp < xl) PC(part, n, cin) = xp + dxp;
      xp = PC(part, n, cin);
      if (xp > xr) PC(part, n, cin) = xp - dxp;
    }
    return np;
  }
  int64_t p = 0;
  for (int64_t n = 0; n < np; n++) {
    double xp = PC(part, n, cin);
    if (xp > xl && xp <= xr) {
      if (p != n) {
        for (int c = 1; c <= ndv; c++) {
          PC(part, p, c) = PC(part, n, c);
          PC(ebfp, p, c) = PC(ebfp, n, c);
        }
      }
      p++;
    }
  }
  return p;
}

int64_t orc_part_dist(const orc_grid *g, int64_t np, double *part, int64_t ld,
                      double *ebfp) {
  /* cell_part_dist(.false.) with one rank per direction
   * (mpi_part_interface.f90:525-753): x, then y, then (3-D) z. */
  const int ndv = 2 * g->curr_ndim + 1;
  np = part_dist_axis(np, part, ld, ebfp, ndv, 1, g->xmn, g->xmx, g->ibx);
  np = part_dist_axis(np, part, ld, ebfp, ndv, 2, g->ymn, g->ymx, g->iby);
  if (g->ndim > 2)
    np = part_dist_axis(np, part, ld, ebfp, ndv, 3, g->zmn, g->zmx, g->ibz);
  return np;
}

/* One axis of the Esirkepov inversion shared by esirkepov_2d_curr and
 * esirkepov_3d_curr (grid_part_connect.f90:1007-1025, 1235-1249): W1 placed
 * into the 5-slot window at offset ih = i1-i0+1, curr = -cumsum(W1-W0),
 * scaled by the FP32 weight. */
static int64_t esk_skipped = 0;

int64_t orc_esirkepov_skipped(void) { return esk_skipped; }

static inline void esk_axis(const double a0[3], const double a1[3], int ih,
                            float wght, double wh[5], double cur[5]) {
  for (int s = 0; s < 5; s++) wh[s] = 0.0;
  for (int s = 0; s <= 2; s++) wh[ih + s] = a1[s];
  cur[0] = -wh[0];
  for (int s = 1; s <= 3; s++) cur[s] = cur[s - 1] + a0[s - 1] - wh[s];
  cur[4] = cur[3] - wh[4];
  for (int s = 0; s < 5; s++) cur[s] = (double)wght * cur[s];
}

void orc_esirkepov(const orc_grid *g, int64_t np, const double *part,
                   int64_t ld, double *ebfp, double *jc) {
  if (np <= 0) return;
  if (g->curr_ndim == 2) {
    /* esirkepov_2d_curr, ndim 2, curr_ndim 2 (grid_part_connect.f90:979-1060) */
    for (int64_t n = 0; n < np; n++) {
      int charge;
      float wgh;
      orc_unpack_weight(PC(part, n, 5), &charge, NULL, &wgh);
      float wght = (float)charge * wgh;
      PC(ebfp, n, 5) = (double)wght;
      PC(ebfp, n, 1) = g->dx_inv * (PC(part, n, 1) - g->xmn);
      PC(ebfp, n, 2) = g->dy_inv * (PC(part, n, 2) - g->ymn);
      PC(ebfp, n, 3) = g->dx_inv * (PC(ebfp, n, 3) - g->xmn);
      PC(ebfp, n, 4) = g->dy_inv * (PC(ebfp, n, 4) - g->ymn);
    }
    for (int64_t n = 0; n < np; n++) {
      double ax0[3], ay0[3], ax1[3], ay1[3], axh[5], ayh[5], currx[5],
          curry[5];
      int i, j, ii0, jj0;
      float wght = (float)PC(ebfp, n, 5);
      qden_axis(PC(ebfp, n, 3), ax0, &ii0);
      qden_axis(PC(ebfp, n, 4), ay0, &jj0);
      qden_axis(PC(ebfp, n, 1), ax1, &i);
      qden_axis(PC(ebfp, n, 2), ay1, &j);
      int ih = i - ii0 + 1;
      if (ih < 0 || ih > 2 || j - jj0 + 1 < 0 || j - jj0 + 1 > 2) {
        esk_skipped++; /* moved >= 1 cell: outside the scheme's validity */
        continue;
      }
      esk_axis(ax0, ax1, ih, wght, axh, currx);
      for (int s = 1; s <= 3; s++) axh[s] = axh[s] + ax0[s - 1];
      int x0 = imin(ih, 1), x1 = imax(ih + 2, 3);
      int jh = j - jj0 + 1;
      esk_axis(ay0, ay1, jh, wght, ayh, curry);
      for (int s = 1; s <= 3; s++) ayh[s] = ayh[s] + ay0[s - 1];
      int y0 = imin(jh, 1), y1 = imax(jh + 2, 3);
      jj0 = jj0 - 1;
      jh = jj0 - 1;
      ii0 = ii0 - 1;
      ih = ii0 - 1;
      for (int j1 = y0; j1 <= y1; j1++) {
        int j2 = jh + j1;
        for (int i1 = x0; i1 <= x1; i1++) {
          int i2 = ih + i1;
          F4(jc, i2, j2, 1, 1) = F4(jc, i2, j2, 1, 1) + ayh[j1] * currx[i1];
        }
      }
      for (int j1 = y0; j1 <= y1; j1++) {
        int j2 = jh + j1;
        for (int i1 = x0; i1 <= x1; i1++) {
          int i2 = ih + i1;
          F4(jc, i2, j2, 1, 2) = F4(jc, i2, j2, 1, 2) + axh[i1] * curry[j1];
        }
      }
    }
    return;
  }
  /* esirkepov_3d_curr (grid_part_connect.f90:1168-1334) */
  for (int64_t n = 0; n < np; n++) {
    int charge;
    float wgh;
    orc_unpack_weight(PC(part, n, 7), &charge, NULL, &wgh);
    float wght = (float)charge * wgh;
    PC(ebfp, n, 7) = (double)wght;
    PC(ebfp, n, 1) = g->dx_inv * (PC(part, n, 1) - g->xmn);
    PC(ebfp, n, 2) = g->dy_inv * (PC(part, n, 2) - g->ymn);
    PC(ebfp, n, 3) = g->dz_inv * (PC(part, n, 3) - g->zmn);
    PC(ebfp, n, 4) = g->dx_inv * (PC(ebfp, n, 4) - g->xmn);
    PC(ebfp, n, 5) = g->dy_inv * (PC(ebfp, n, 5) - g->ymn);
    PC(ebfp, n, 6) = g->dz_inv * (PC(ebfp, n, 6) - g->zmn);
  }
  for (int64_t n = 0; n < np; n++) {
    double ax0[3], ay0[3], az0[3], ax1[3], ay1[3], az1[3];
    double axh[5], ayh[5], azh[5], axh0[5], axh1[5], ayh0[5], ayh1[5];
    double currx[5], curry[5], currz[5];
    int i, j, k, ii0, jj0, kk0;
    float wght = (float)PC(ebfp, n, 7);
    qden_axis(PC(ebfp, n, 4), ax0, &ii0);
    qden_axis(PC(ebfp, n, 5), ay0, &jj0);
    qden_axis(PC(ebfp, n, 6), az0, &kk0);
    qden_axis(PC(ebfp, n, 1), ax1, &i);
    qden_axis(PC(ebfp, n, 2), ay1, &j);
    qden_axis(PC(ebfp, n, 3), az1, &k);
    int ih = i - ii0 + 1;
    if (ih < 0 || ih > 2 || j - jj0 + 1 < 0 || j - jj0 + 1 > 2 ||
        k - kk0 + 1 < 0 || k - kk0 + 1 > 2) {
      esk_skipped++; /* moved >= 1 cell: outside the scheme's validity */
      continue;
    }
    esk_axis(ax0, ax1, ih, wght, axh, currx);
    for (int s = 0; s < 5; s++) {
      axh0[s] = 0.5 * axh[s];
      axh1[s] = axh[s];
    }
    for (int s = 1; s <= 3; s++) {
      axh0[s] = axh0[s] + ax0[s - 1];
      axh1[s] = axh1[s] + 0.5 * ax0[s - 1];
    }
    int x0 = imin(ih, 1), x1 = imax(ih + 2, 3);
    i = i - 1;
    ii0 = ii0 - 1;
    int jh = j - jj0 + 1;
    esk_axis(ay0, ay1, jh, wght, ayh, curry);
    for (int s = 0; s < 5; s++) {
      ayh0[s] = 0.5 * ayh[s];
      ayh1[s] = ayh[s];
    }
    for (int s = 1; s <= 3; s++) {
      ayh0[s] = ayh0[s] + ay0[s - 1];
      ayh1[s] = ayh1[s] + 0.5 * ay0[s - 1];
    }
    int y0 = imin(jh, 1), y1 = imax(jh + 2, 3);
    jj0 = jj0 - 1;
    j = j - 1;
    int kh = k - kk0 + 1;
    esk_axis(az0, az1, kh, wght, azh, currz);
    kk0 = kk0 - 1;
    k = k - 1;
    int z0 = imin(kh, 1), z1 = imax(kh + 2, 3);
    jh = jj0 - 1;
    ih = ii0 - 1;
    for (int k1 = 0; k1 <= 2; k1++) {
      for (int j1 = y0; j1 <= y1; j1++) {
        int j2 = jh + j1;
        double dvol = ayh0[j1] * az0[k1];
        double dvolh = ayh1[j1] * az1[k1];
        for (int i1 = x0; i1 <= x1; i1++) {
          int i2 = ih + i1;
          F4(jc, i2, j2, kk0 + k1, 1) =
              F4(jc, i2, j2, kk0 + k1, 1) + dvol * currx[i1];
          F4(jc, i2, j2, k + k1, 1) =
              F4(jc, i2, j2, k + k1, 1) + dvolh * currx[i1];
        }
      }
    }
    for (int k1 = 0; k1 <= 2; k1++) {
      for (int j1 = y0; j1 <= y1; j1++) {
        int j2 = jh + j1;
        double dvol = curry[j1] * az0[k1];
        double dvolh = curry[j1] * az1[k1];
        for (int i1 = x0; i1 <= x1; i1++) {
          int i2 = ih + i1;
          F4(jc, i2, j2, kk0 + k1, 2) =
              F4(jc, i2, j2, kk0 + k1, 2) + axh0[i1] * dvol;
          F4(jc, i2, j2, k + k1, 2) =
              F4(jc, i2, j2, k + k1, 2) + axh1[i1] * dvolh;
        }
      }
    }
    kh = kk0 - 1;
    for (int k1 = z0; k1 <= z1; k1++) {
      int k2 = kh + k1;
      for (int j1 = 0; j1 <= 2; j1++) {
        double dvol = ay0[j1] * currz[k1];
        double dvolh = ay1[j1] * currz[k1];
        for (int i1 = x0; i1 <= x1; i1++) {
          int i2 = ih + i1;
          F4(jc, i2, jj0 + j1, k2, 3) =
              F4(jc, i2, jj0 + j1, k2, 3) + axh0[i1] * dvol;
          F4(jc, i2, j + j1, k2, 3) =
              F4(jc, i2, j + j1, k2, 3) + axh1[i1] * dvolh;
        }
      }
    }
  }
}

static const int x_parity[6] = {-1, 1, 1, -1, 1, 1};
static const int y_parity[6] = {1, -1, 1, 1, -1, 1};
static const int z_parity[6] = {1, 1, -1, 1, 1, -1};

void orc_field_xyzbd(const orc_grid *g, double *ebf, int nc) {
  /* field_xyzbd on a single rank: pex0, pex1 and all *_bd are true
   * (mpi_field_interface.f90:53-198). */
  int i1 = g->ix1, i2 = g->ix2, j1 = g->jy1, j2 = g->jy2, k1 = g->kz1,
      k2 = g->kz2;
  const int ix1 = g->ix1, ix2 = g->ix2, jy1 = g->jy1, jy2 = g->jy2,
            kz1 = g->kz1, kz2 = g->kz2;
  if (g->ibx < 2)
    for (int ik = 1; ik <= nc; ik++)
      for (int iz = k1; iz <= k2; iz++)
        for (int iy = j1; iy <= j2; iy++)
          F4(ebf, ix1 - 1, iy, iz, ik) =
              2. * F4(ebf, ix1, iy, iz, ik) - F4(ebf, ix1 + 1, iy, iz, ik);
  i1 = i1 - 1;
  if (g->ibx == 0)
    for (int ik = 1; ik <= nc; ik++)
      for (int iz = k1; iz <= k2; iz++)
        for (int iy = j1; iy <= j2; iy++)
          F4(ebf, ix2 + 1, iy, iz, ik) =
              2. * F4(ebf, ix2, iy, iz, ik) - F4(ebf, ix2 - 1, iy, iz, ik);
  if (g->ibx == 1)
    for (int ik = 1; ik <= nc; ik++)
      for (int iz = k1; iz <= k2; iz++)
        for (int iy = j1; iy <= j2; iy++)
          F4(ebf, ix2 + 1, iy, iz, ik) =
              x_parity[ik - 1] * F4(ebf, ix2, iy, iz, ik);
  i2 = i2 + 1;
  if (g->ndim < 2) return;
  if (g->iby == 0)
    for (int ik = 1; ik <= nc; ik++)
      for (int iz = k1; iz <= k2; iz++)
        for (int ix = i1; ix <= i2; ix++)
          F4(ebf, ix, jy1 - 1, iz, ik) =
              2. * F4(ebf, ix, jy1, iz, ik) - F4(ebf, ix, jy1 + 1, iz, ik);
  if (g->iby == 1)
    for (int ik = 1; ik <= nc; ik++)
      for (int iz = k1; iz <= k2; iz++)
        for (int ix = i1; ix <= i2; ix++)
          F4(ebf, ix, jy1 - 1, iz, ik) =
              y_parity[ik - 1] * F4(ebf, ix, jy1 + 1, iz, ik);
  j1 = j1 - 1;
  if (g->iby == 0)
    for (int ik = 1; ik <= nc; ik++)
      for (int iz = k1; iz <= k2; iz++)
        for (int ix = i1; ix <= i2; ix++)
          F4(ebf, ix, jy2 + 1, iz, ik) =
              2. * F4(ebf, ix, jy2, iz, ik) - F4(ebf, ix, jy2 - 1, iz, ik);
  if (g->iby == 1)
    for (int ik = 1; ik <= nc; ik++)
      for (int iz = k1; iz <= k2; iz++)
        for (int ix = i1; ix <= i2; ix++)
          F4(ebf, ix, jy2 + 1, iz, ik) =
              y_parity[ik - 1] * F4(ebf, ix, jy2 - 1, iz, ik);
  j2 = j2 + 1;
  if (g->ndim < 3) return;
  if (g->ibz == 0)
    for (int ik = 1; ik <= nc; ik++)
      for (int iy = j1; iy <= j2; iy++)
        for (int ix = i1; ix <= i2; ix++)
          F4(ebf, ix, iy, kz1 - 1, ik) =
              2. * F4(ebf, ix, iy, kz1, ik) - F4(ebf, ix, iy, kz1 + 1, ik);
  if (g->ibz == 1)
    for (int ik = 1; ik <= nc; ik++)
      for (int iy = j1; iy <= j2; iy++)
        for (int ix = i1; ix <= i2; ix++)
          F4(ebf, ix, iy, kz1 - 1, ik) =
              z_parity[ik - 1] * F4(ebf, ix, iy, kz1 + 1, ik);
  k1 = k1 - 1;
  if (g->ibz == 0)
    for (int ik = 1; ik <= nc; ik++)
      for (int iy = j1; iy <= j2; iy++)
        for (int ix = i1; ix <= i2; ix++)
          F4(ebf, ix, iy, kz2 + 1, ik) =
              2. * F4(ebf, ix, iy, kz2, ik) - F4(ebf, ix, iy, kz2 - 1, ik);
  if (g->ibz == 1) /* the reference writes kz1+1 from kz1-1 here */
    for (int ik = 1; ik <= nc; ik++)
      for (int iy = j1; iy <= j2; iy++)
        for (int ix = i1; ix <= i2; ix++)
          F4(ebf, ix, iy, kz1 + 1, ik) =
              z_parity[ik - 1] * F4(ebf, ix, iy, kz1 - 1, ik);
  k2 = k2 + 1;
  (void)k2;
}

void orc_jc_xyzbd(const orc_grid *g, double *jc) {
  /* jc_xyzbd on a single rank (mpi_curr_interface.f90:551-643). */
  const int i1 = g->ix1, n1 = g->ix2, j1 = g->jy1, n2 = g->jy2, k1 = g->kz1,
            n3 = g->kz2;
  const int j2 = n2, k2 = n3, nc = g->curr_ndim;
  if (g->ibx == 0) {
    for (int ik = 1; ik <= nc; ik++)
      for (int kk = k1; kk <= k2; kk++)
        for (int jj = j1; jj <= j2; jj++) {
          F4(jc, i1, jj, kk, ik) = F4(jc, i1, jj, kk, ik) + F4(jc, i1 - 1, jj, kk, ik);
          F4(jc, i1 - 1, jj, kk, ik) = 0.0;
          F4(jc, n1, jj, kk, ik) = F4(jc, n1, jj, kk, ik) +
                                   F4(jc, n1 + 1, jj, kk, ik) +
                                   F4(jc, n1 + 2, jj, kk, ik);
          F4(jc, n1 + 1, jj, kk, ik) = 0.0;
          F4(jc, n1 + 2, jj, kk, ik) = 0.0;
        }
  }
  if (g->ndim < 2) return;
  if (g->iby == 0) {
    for (int ik = 1; ik <= nc; ik++)
      for (int iz = k1; iz <= k2; iz++)
        for (int ix = i1; ix <= n1; ix++) {
          F4(jc, ix, j1, iz, ik) = F4(jc, ix, j1, iz, ik) +
                                   F4(jc, ix, j1 - 1, iz, ik) +
                                   F4(jc, ix, j1 - 2, iz, ik);
          F4(jc, ix, j1 - 2, iz, ik) = 0.0;
          F4(jc, ix, j1 - 1, iz, ik) = 0.0;
        }
  }
  if (g->iby == 1) {
    for (int iz = k1; iz <= k2; iz++)
      for (int ix = i1; ix <= n1; ix++) {
        F4(jc, ix, j1 + 1, iz, 1) = F4(jc, ix, j1 + 1, iz, 1) + F4(jc, ix, j1 - 1, iz, 1);
        F4(jc, ix, j1, iz, 2) = F4(jc, ix, j1, iz, 2) - F4(jc, ix, j1 - 1, iz, 2);
      }
  }
  if (g->iby == 0) {
    for (int ik = 1; ik <= nc; ik++)
      for (int iz = k1; iz <= k2; iz++)
        for (int ix = i1; ix <= n1; ix++) {
          F4(jc, ix, n2, iz, ik) = F4(jc, ix, n2, iz, ik) +
                                   F4(jc, ix, n2 + 1, iz, ik) +
                                   F4(jc, ix, n2 + 2, iz, ik);
          F4(jc, ix, n2 + 1, iz, ik) = 0.0;
          F4(jc, ix, n2 + 2, iz, ik) = 0.0;
        }
  }
  if (g->ndim < 3) return;
  if (g->ibz == 0) {
    for (int ik = 1; ik <= nc; ik++)
      for (int iy = j1; iy <= j2; iy++)
        for (int ix = i1; ix <= n1; ix++) {
          F4(jc, ix, iy, k1, ik) = F4(jc, ix, iy, k1, ik) + F4(jc, ix, iy, k1 - 1, ik);
          F4(jc, ix, iy, k1 - 1, ik) = 0.0;
        }
    for (int ik = 1; ik <= nc; ik++)
      for (int iy = j1; iy <= j2; iy++)
        for (int ix = i1; ix <= n1; ix++) {
          F4(jc, ix, iy, n3, ik) = F4(jc, ix, iy, n3, ik) +
                                   F4(jc, ix, iy, n3 + 1, ik) +
                                   F4(jc, ix, iy, n3 + 2, ik);
          F4(jc, ix, iy, n3 + 1, ik) = 0.0;
          F4(jc, ix, iy, n3 + 2, ik) = 0.0;
        }
  }
}

void orc_curr_collect(const orc_grid *g, double *jc) {
  /* curr_mpi_collect (curr_and_fields_util.f90:104-152), unstretched. */
  orc_jc_xyzbd(g, jc);
  const int64_t ncell = (int64_t)g->n1p * g->n2p * g->n3p;
  for (int i = 0; i < g->ndim; i++)
    for (int64_t c = 0; c < ncell; c++)
      jc[(int64_t)i * ncell + c] = g->djc[i] * jc[(int64_t)i * ncell + c];
}

void orc_ef_bds(const orc_grid *g, double *ebf, double dtl) {
  /* ef_bds(ef, dtl, 0) on a single rank (grid_fields.f90:1204-1354). */
  const int nf = g->nfield, ix2 = g->ix2, jy2 = g->jy2, kz2 = g->kz2;
  double aphx = 1, aphy = 1, aphz = 1;
  if (g->ibx < 2) {
    if (g->ibx == 0) {
      aphx = 1.0 * g->dx_inv * dtl;
      for (int k = g->kz1; k <= g->kz2; k++)
        for (int j = g->jy1; j <= g->jy2; j++)
          F4(ebf, ix2 + 1, j, k, 2) =
              (2. * F4(ebf, ix2, j, k, nf) - (1. - aphx) * F4(ebf, ix2, j, k, 2)) /
              (1. + aphx);
    } else {
      for (int k = g->kz1; k <= g->kz2; k++)
        for (int j = g->jy1; j <= g->jy2; j++)
          F4(ebf, ix2 + 1, j, k, 2) = F4(ebf, ix2, j, k, 2);
    }
    if (nf > 3) {
      if (g->ibx == 0) {
        for (int k = g->kz1; k <= g->kz2; k++)
          for (int j = g->jy1; j <= g->jy2; j++)
            F4(ebf, ix2 + 1, j, k, 3) =
                -(2. * F4(ebf, ix2, j, k, 5) + (1. - aphx) * F4(ebf, ix2, j, k, 3)) /
                (1. + aphx);
      } else {
        for (int k = g->kz1; k <= g->kz2; k++)
          for (int j = g->jy1; j <= g->jy2; j++)
            F4(ebf, ix2 + 1, j, k, 3) = F4(ebf, ix2, j, k, 3);
      }
    }
  }
  if (g->ndim < 2) return;
  if (g->iby < 2) {
    aphy = 1.0 * g->dy_inv * dtl;
    for (int k = g->kz1; k <= g->kz2; k++)
      for (int i = g->ix1; i <= g->ix2; i++)
        F4(ebf, i, jy2 + 1, k, 1) =
            -(2. * F4(ebf, i, jy2, k, nf) + (1. - aphy) * F4(ebf, i, jy2, k, 1)) /
            (1. + aphy);
    if (nf > 3)
      for (int k = g->kz1; k <= g->kz2; k++)
        for (int i = g->ix1; i <= g->ix2; i++)
          F4(ebf, i, jy2 + 1, k, 3) =
              (2. * F4(ebf, i, jy2, k, 4) - (1. - aphy) * F4(ebf, i, jy2, k, 3)) /
              (1. + aphy);
  }
  if (g->ndim < 3) return;
  if (g->ibz < 2) {
    aphz = 1.0 * g->dz_inv * dtl;
    for (int j = g->jy1; j <= g->jy2; j++)
      for (int i = g->ix1; i <= g->ix2; i++) {
        F4(ebf, i, j, kz2 + 1, 1) =
            (2. * F4(ebf, i, j, kz2, 5) - (1. - aphz) * F4(ebf, i, j, kz2, 1)) /
            (1. + aphz);
        F4(ebf, i, j, kz2 + 1, 2) =
            -(2. * F4(ebf, i, j, kz2, 4) + (1. - aphz) * F4(ebf, i, j, kz2, 2)) /
            (1. + aphz);
      }
  }
}

void orc_bf_bds(const orc_grid *g, double *ebf, double dtl) {
  /* bf_bds(ef, dtl, 0) on a single rank (grid_fields.f90:1069-1202). */
  const int nf = g->nfield, ix1 = g->ix1, jy1 = g->jy1, kz1 = g->kz1;
  if (g->ibx < 2) {
    double aphx = 1.0 * g->dx_inv * dtl;
    for (int k = g->kz1; k <= g->kz2; k++)
      for (int j = g->jy1; j <= g->jy2; j++)
        F4(ebf, ix1 - 1, j, k, nf) =
            -(2. * F4(ebf, ix1, j, k, 2) + (1. - aphx) * F4(ebf, ix1, j, k, nf)) /
            (1. + aphx);
    if (nf > 3)
      for (int k = g->kz1; k <= g->kz2; k++)
        for (int j = g->jy1; j <= g->jy2; j++)
          F4(ebf, ix1 - 1, j, k, 5) =
              (2. * F4(ebf, ix1, j, k, 3) - (1. - aphx) * F4(ebf, ix1, j, k, 5)) /
              (1. + aphx);
  }
  if (g->ndim < 2) return;
  if (g->iby < 2) {
    double aphy = 1.0 * g->dy_inv * dtl;
    for (int k = g->kz1; k <= g->kz2; k++)
      for (int i = g->ix1; i <= g->ix2; i++) {
        double v = 2. * F4(ebf, i, jy1, k, 1) - (1. - aphy) * F4(ebf, i, jy1, k, nf);
        F4(ebf, i, jy1 - 1, k, nf) = v / (1. + aphy);
      }
    if (nf > 3)
      for (int k = g->kz1; k <= g->kz2; k++)
        for (int i = g->ix1; i <= g->ix2; i++) {
          double v = -2. * F4(ebf, i, jy1, k, 3) - (1. - aphy) * F4(ebf, i, jy1, k, 4);
          F4(ebf, i, jy1 - 1, k, 4) = v / (1. + aphy);
        }
  }
  if (g->ndim < 3) return;
  if (g->ibz < 2) {
    double aphz = 1.0 * g->dz_inv * dtl;
    for (int j = g->jy1; j <= g->jy2; j++)
      for (int i = g->ix1; i <= g->ix2; i++) {
        double v = 2. * F4(ebf, i, j, kz1, 2) - (1. - aphz) * F4(ebf, i, j, kz1, 4);
        F4(ebf, i, j, kz1 - 1, 4) = v / (1. + aphz);
        v = -2. * F4(ebf, i, j, kz1, 1) - (1. - aphz) * F4(ebf, i, j, kz1, 5);
        F4(ebf, i, j, kz1 - 1, 5) = v / (1. + aphz);
      }
  }
}

void orc_rote(const orc_grid *g, double *ebf, double dtf) {
  /* rote (grid_fields.f90:1405-1478), unstretched metric = 1. */
  const double aphx = dtf * g->dx_inv, aphy = dtf * g->dy_inv,
               aphz = dtf * g->dz_inv;
  const double aph1 = aphx * g->se_coeff[0], aph2 = aphx * g->se_coeff[1];
  const int nf = g->nfield;
  for (int k = g->kz1; k <= g->kz2; k++)
    for (int j = g->jy1; j <= g->jy2; j++) {
      const double sdhy = 1.0 * aphy;
      for (int i = g->ix1; i <= g->ix2; i++)
        F4(ebf, i, j, k, nf) =
            F4(ebf, i, j, k, nf) -
            aph1 * (F4(ebf, i + 1, j, k, 2) - F4(ebf, i, j, k, 2)) +
            sdhy * (F4(ebf, i, j + 1, k, 1) - F4(ebf, i, j, k, 1)) -
            aph2 * (F4(ebf, i + 2, j, k, 2) - F4(ebf, i - 1, j, k, 2));
    }
  if (nf < 6) return;
  for (int k = g->kz1; k <= g->kz2; k++) {
    const double sdhz = 1.0 * aphz;
    for (int j = g->jy1; j <= g->jy2; j++) {
      const double sdhy = 1.0 * aphy;
      for (int i = g->ix1; i <= g->ix2; i++) {
        F4(ebf, i, j, k, 4) =
            F4(ebf, i, j, k, 4) -
            sdhy * (F4(ebf, i, j + 1, k, 3) - F4(ebf, i, j, k, 3)) +
            sdhz * (F4(ebf, i, j, k + 1, 2) - F4(ebf, i, j, k, 2));
        F4(ebf, i, j, k, 5) =
            F4(ebf, i, j, k, 5) +
            aph1 * (F4(ebf, i + 1, j, k, 3) - F4(ebf, i, j, k, 3)) -
            sdhz * (F4(ebf, i, j, k + 1, 1) - F4(ebf, i, j, k, 1)) +
            aph2 * (F4(ebf, i + 2, j, k, 3) - F4(ebf, i - 1, j, k, 3));
      }
    }
  }
}

void orc_rotb(const orc_grid *g, double *ebf, double dtf) {
  /* rotb (grid_fields.f90:1480-1552), unstretched metric = 1. */
  const double aphx = dtf * g->dx_inv, aphy = dtf * g->dy_inv,
               aphz = dtf * g->dz_inv;
  const double aph1 = aphx * g->se_coeff[0], aph2 = aphx * g->se_coeff[1];
  const int nf = g->nfield;
  for (int k = g->kz1; k <= g->kz2; k++)
    for (int j = g->jy1; j <= g->jy2; j++) {
      const double sdy = 1.0 * aphy;
      for (int i = g->ix1; i <= g->ix2; i++) {
        F4(ebf, i, j, k, 1) =
            F4(ebf, i, j, k, 1) +
            sdy * (F4(ebf, i, j, k, nf) - F4(ebf, i, j - 1, k, nf));
        F4(ebf, i, j, k, 2) =
            F4(ebf, i, j, k, 2) -
            aph1 * (F4(ebf, i, j, k, nf) - F4(ebf, i - 1, j, k, nf)) -
            aph2 * (F4(ebf, i + 1, j, k, nf) - F4(ebf, i - 2, j, k, nf));
      }
    }
  if (nf < 6) return;
  for (int k = g->kz1; k <= g->kz2; k++) {
    const double sdz = aphz * 1.0;
    for (int j = g->jy1; j <= g->jy2; j++) {
      const double sdy = aphy * 1.0;
      for (int i = g->ix1; i <= g->ix2; i++) {
        F4(ebf, i, j, k, 1) =
            F4(ebf, i, j, k, 1) -
            sdz * (F4(ebf, i, j, k, 5) - F4(ebf, i, j, k - 1, 5));
        F4(ebf, i, j, k, 2) =
            F4(ebf, i, j, k, 2) +
            sdz * (F4(ebf, i, j, k, 4) - F4(ebf, i, j, k - 1, 4));
        F4(ebf, i, j, k, 3) =
            F4(ebf, i, j, k, 3) +
            aph1 * (F4(ebf, i, j, k, 5) - F4(ebf, i - 1, j, k, 5)) -
            sdy * (F4(ebf, i, j, k, 4) - F4(ebf, i, j - 1, k, 4)) +
            aph2 * (F4(ebf, i + 1, j, k, 5) - F4(ebf, i - 2, j, k, 5));
      }
    }
  }
}

void orc_advance_fields(const orc_grid *g, double *ebf, const double *jc) {
  /* advance_lpf_fields(ef, curr, 0) (curr_and_fields_util.f90:176-260). */
  const double dt_lp = g->dt;
  const double dth = 0.5 * dt_lp;
  orc_ef_bds(g, ebf, 0.0);
  orc_rote(g, ebf, dth);
  orc_bf_bds(g, ebf, dt_lp);
  orc_rotb(g, ebf, dt_lp);
  for (int ik = 1; ik <= g->curr_ndim; ik++)
    for (int k = g->kz1; k <= g->kz2; k++)
      for (int j = g->jy1; j <= g->jy2; j++)
        for (int i = g->ix1; i <= g->ix2; i++)
          F4(ebf, i, j, k, ik) = F4(ebf, i, j, k, ik) - g->ompe * F4(jc, i, j, k, ik);
  orc_ef_bds(g, ebf, dt_lp);
  orc_rote(g, ebf, dth);
}

int64_t orc_particle_stage(const orc_grid *g, int64_t np, double *part,
                           int64_t ld, double *ebfp, const double *ebf,
                           double *jc, int initial_time) {
  /* pic_evolve_in_time.f90:75-91 with part = .true. */
  orc_gather(g, np, part, ld, ebf, ebfp);
  orc_charge_multiply(g, np, part, ld, ebfp);
  if (initial_time) orc_init_momenta(g, np, part, ld, ebfp);
  orc_push(g, np, part, ld, ebfp);
  np = orc_part_dist(g, np, part, ld, ebfp);
  orc_esirkepov(g, np, part, ld, ebfp, jc);
  return np;
}

int64_t orc_lpf2_evolve(const orc_grid *g, int64_t np, double *part,
                        int64_t ld, double *ebfp, double *ebf, double *jc,
                        int initial_time) {
  /* lpf2_evolve (pic_evolve_in_time.f90:37-109), one species, part=.true. */
  orc_field_xyzbd(g, ebf, g->nfield);
  const int64_t njc = (int64_t)g->n1p * g->n2p * g->n3p * g->curr_ndim;
  memset(jc, 0, (size_t)njc * sizeof(double));
  np = orc_particle_stage(g, np, part, ld, ebfp, ebf, jc, initial_time);
  orc_curr_collect(g, jc);
  orc_advance_fields(g, ebf, jc);
  return np;
}

void orc_grid_charge(const orc_grid *g, int64_t np, const double *part,
                     int64_t ld, double *rho) {
  /* set_grid_charge (grid_part_util.f90:240-335), den(:,:,:,1). */
  const int ch = 2 * g->curr_ndim + 1;
  for (int64_t n = 0; n < np; n++) {
    int charge;
    float wgh;
    orc_unpack_weight(PC(part, n, ch), &charge, NULL, &wgh);
    float wght = (float)charge * wgh;
    double ax0[3], ay0[3], az0[3];
    int i, j, k;
    qden_axis(g->dx_inv * (PC(part, n, 1) - g->xmn), ax0, &i);
    qden_axis(g->dy_inv * (PC(part, n, 2) - g->ymn), ay0, &j);
    for (int s = 0; s < 3; s++) ax0[s] = (double)wght * ax0[s];
    i = i - 1;
    j = j - 1;
    if (g->ndim == 2) {
      for (int j1 = 0; j1 <= 2; j1++)
        for (int i1 = 0; i1 <= 2; i1++) {
          double dvol = ax0[i1] * ay0[j1];
          F4(rho, i + i1, j + j1, 1, 1) = F4(rho, i + i1, j + j1, 1, 1) + dvol;
        }
    } else {
      qden_axis(g->dz_inv * (PC(part, n, 3) - g->zmn), az0, &k);
      k = k - 1;
      for (int k1 = 0; k1 <= 2; k1++)
        for (int j1 = 0; j1 <= 2; j1++) {
          double dvol = az0[k1] * ay0[j1];
          for (int i1 = 0; i1 <= 2; i1++)
            F4(rho, i + i1, j + j1, k + k1, 1) =
                F4(rho, i + i1, j + j1, k + k1, 1) + ax0[i1] * dvol;
        }
    }
  }
}

void orc_energy_momenta(const orc_grid *g, int64_t np, const double *part,
                        int64_t ld, double ek[7], double *ekmax) {
  /* energy_momenta (diag_part_and_fields.f90:172-219). */
  for (int c = 0; c < 7; c++) ek[c] = 0.0;
  double emax = 0.0;
  if (g->curr_ndim < 3) {
    for (int64_t n = 0; n < np; n++) {
      double vp1 = PC(part, n, 3), vp2 = PC(part, n, 4);
      double gamm = sqrt(1. + vp1 * vp1 + vp2 * vp2);
      int charge;
      float wgh;
      orc_unpack_weight(PC(part, n, 5), &charge, NULL, &wgh);
      double gam1 = gamm - 1.;
      ek[0] = ek[0] + (double)wgh * vp1;
      ek[1] = ek[1] + (double)wgh * vp2;
      ek[5] = ek[5] + (double)((float)charge * wgh);
      ek[6] = ek[6] + (double)wgh * gam1;
      if (gam1 > emax) emax = gam1;
    }
  } else {
    for (int64_t n = 0; n < np; n++) {
      double xp2 = PC(part, n, 2), xp3 = PC(part, n, 3);
      double vp1 = PC(part, n, 4), vp2 = PC(part, n, 5), vp3 = PC(part, n, 6);
      double gamm = sqrt(1. + vp1 * vp1 + vp2 * vp2 + vp3 * vp3);
      int charge;
      float wgh;
      orc_unpack_weight(PC(part, n, 7), &charge, NULL, &wgh);
      double gam1 = gamm - 1.;
      ek[0] = ek[0] + vp1;
      ek[1] = ek[1] + vp2;
      ek[2] = ek[2] + vp3;
      ek[3] = ek[3] + (double)wgh * (xp2 * vp3 - xp3 * vp2);
      ek[5] = ek[5] + (double)((float)charge * wgh);
      ek[6] = ek[6] + (double)wgh * gam1;
      if (gam1 > emax) emax = gam1;
    }
  }
  if (ekmax) *ekmax = emax;
}

void orc_field_energy(const orc_grid *g, const double *ebf, double out[6]) {
  /* Envar field energy, unstretched branch (diag_part_and_fields.f90:499-670):
   * i1 = ix1 .. nxp = ix2, j1 = jy1 .. nyp = jy2, k1 = kz1 .. nzp = kz2. */
  double dgvol = g->dx * g->dy * g->dz;
  if (g->ndim == 2) dgvol = g->dx * g->dy * g->dy;
  for (int c = 0; c < 6; c++) out[c] = 0.0;
  for (int ik = 1; ik <= g->nfield; ik++) {
    double s = 0.0;
    for (int iz = g->kz1; iz <= g->kz2; iz++)
      for (int iy = g->jy1; iy <= g->jy2; iy++)
        for (int ix = g->ix1; ix <= g->ix2; ix++) {
          double v = F4(ebf, ix, iy, iz, ik);
          s = s + v * v;
        }
    out[ik - 1] = dgvol * s;
  }
}

void orc_fields_left_xshift(const orc_grid *g, double *fld, int ncomp, int nshx) {
  /* fields_left_xshift(fld, i1, i2, 1, ncomp, nshx), one rank
   * (mpi_field_interface.f90:299-359): (i1+nshx : i2+nshx) -> (i1 : i2) on the
   * interior rows, then the last nshx columns zero */
  if (nshx <= 0) return;
  const int i1 = g->ix1, i2 = g->ix2 - nshx;
  for (int ic = 1; ic <= ncomp; ic++)
    for (int iz = g->kz1; iz <= g->kz2; iz++)
      for (int iy = g->jy1; iy <= g->jy2; iy++) {
        for (int ix = i1; ix <= i2; ix++) F4(fld, ix, iy, iz, ic) = F4(fld, ix + nshx, iy, iz, ic);
        for (int ix = i2 + 1; ix <= i2 + nshx; ix++) F4(fld, ix, iy, iz, ic) = 0.0;
      }
}

int orc_window_xshift(orc_grid *g, orc_window *w, int witr, int init_iter,
                      double w_speed, double *ebf, int64_t *np, double *part,
                      int64_t ld) {
  /* lp_window_xshift (window.f90:276-347), one rank, no fluid; the envelope
   * arrays are shifted by the caller with orc_fields_left_xshift (:327-331) */
  if (init_iter == 0) {
    w->xlapse = 0.0;
    w->wi1 = 0;
    return 0;
  }
  w->xlapse = w->xlapse + w_speed * g->dt * witr;
  const int wi2 = (int)lround(g->dx_inv * w->xlapse); /* nint */
  const int nshx = wi2 - w->wi1;
  w->wi1 = wi2;
  g->xmn = g->xmn + g->dx * nshx;
  g->xmx = g->xmx + g->dx * nshx;
  if (g->ix2 - nshx <= 0) return -1; /* wi2 = n1p - nshx <= 0 (window.f90:314-319) */
  orc_fields_left_xshift(g, ebf, g->nfield, nshx);
  /* cell_part_dist(.true.): x only, (xmn, xmx], stable */
  const int ndv = 2 * g->curr_ndim + 1;
  int64_t p = 0;
  for (int64_t n = 0; n < *np; n++) {
    const double xp = PC(part, n, 1);
    if (xp > g->xmn && xp <= g->xmx) {
      if (p != n)
        for (int c = 1; c <= ndv; c++) PC(part, p, c) = PC(part, n, c);
      p++;
    }
  }
  *np = p;
  return nshx;
}

int64_t orc_inject_select(const double *xpt, int64_t nxp, int64_t *nptx, double xmx,
                          double targ_in, double targ_end, int64_t *c0, int64_t *c1) {
  /* lp_window_xshift's gate (window.f90:339-341), then particles_inject
   * (window.f90:106-139) in the reference's 1-based indices */
  *c0 = 0;
  *c1 = -1;
  if (!(targ_in <= xmx && targ_end > xmx)) return 0;
  const int64_t i1 = *nptx + 1;
  int64_t i2;
  if (i1 <= nxp) {
    int64_t ix = i1;
    for (; ix <= nxp; ix++)
      if (xpt[ix - 1] > xmx) break;
    i2 = ix - 1;
    if (ix == nxp) i2 = ix;
  } else {
    i2 = i1 - 1;
  }
  *nptx = i2;
  if (!(i2 > i1)) return 0;
  int64_t last = i2;
  while (last >= i1 && xpt[last - 1] > xmx) last--;
  if (last < i1) return 0;
  *c0 = i1 - 1;
  *c1 = last - 1;
  return last - i1 + 1;
}

int64_t orc_add_particles(const orc_grid *g, const double *xpt, int64_t c0, int64_t c1,
                          const double *ypt, int nyp, const double *zpt, int nzp,
                          const double *wghpt, const double *wghyz, int charge,
                          double *part, int64_t ld, int64_t np) {
  /* add_particles (window.f90:39-104) with t0_pl = 0 */
  const int D = g->curr_ndim;
  int64_t n = np;
  if (D > 2) {
    for (int k = 0; k < nzp; k++)
      for (int j = 0; j < nyp; j++)
        for (int64_t ix = c0; ix <= c1; ix++) {
          const double whz = wghpt[ix] * wghyz[j + (int64_t)nyp * k];
          PC(part, n, 1) = xpt[ix];
          PC(part, n, 2) = ypt[j];
          PC(part, n, 3) = zpt[k];
          PC(part, n, 4) = 0.0;
          PC(part, n, 5) = 0.0;
          PC(part, n, 6) = 0.0;
          PC(part, n, 7) = orc_pack_weight(charge, 0, (float)whz);
          n++;
        }
  } else {
    for (int j = 0; j < nyp; j++)
      for (int64_t ix = c0; ix <= c1; ix++) {
        PC(part, n, 1) = xpt[ix];
        PC(part, n, 2) = ypt[j];
        PC(part, n, 3) = 0.0;
        PC(part, n, 4) = 0.0;
        PC(part, n, 5) = orc_pack_weight(charge, 0, (float)(wghyz[j] * wghpt[ix]));
        n++;
      }
  }
  return n;
}

/* ===================== envelope (ponderomotive) particle routines ======== */

/* qqh_*_spline, one axis (grid_part_lib.f90:145-210, 261-330): quadratic
 * weights at integer nodes (base i = int(xx+1/2) - 1) and at half-integer
 * nodes (base ih = int(xx) - 1, offset xx - int(xx) - 1/2). */
static inline void qqh_axis(double xp, double a[3], double ah[3], int *i, int *ih) {
  double xx = 3.0 + xp; /* shx = 3 */
  int ix = (int)(xx + 0.5);
  double sx = xx - (double)ix;
  double sx2 = sx * sx;
  a[1] = 0.75 - sx2;
  a[2] = 0.5 * (0.25 + sx2 + sx);
  a[0] = 1. - a[1] - a[2];
  int ihx = (int)xx;
  sx = xx - (double)ihx - 0.5;
  sx2 = sx * sx;
  ah[1] = 0.75 - sx2;
  ah[2] = 0.5 * (0.25 + sx2 + sx);
  ah[0] = 1. - ah[1] - ah[2];
  *i = ix - 1;
  *ih = ihx - 1;
}

/* local grid coordinate of part(n, d), d = 1..3 (set_local_*d_positions,
 * grid_part_lib.f90:458-519, unstretched) */
static inline double loc_coord(const orc_grid *g, double x, int d) {
  if (d == 1) return g->dx_inv * (x - g->xmn);
  if (d == 2) return g->dy_inv * (x - g->ymn);
  return g->dz_inv * (x - g->zmn);
}

void orc_env_acc(const orc_grid *g, int64_t np, const double *part, int64_t ld,
                 const double *ebf, const double *av, double dt_step, double *ebfp) {
  /* set_env_acc (grid_part_connect.f90:414-615): E, B, F = |a|^2/2 and
   * grad F at the particle (quadratic shapes at integer and half-integer
   * nodes), the approximate gamma of the ponderomotive scheme, exit
   * pt(1:D) = q E - q^2 grad F/(2 gamp), pt(D+1 ..) = q B/gamp,
   * pt(2D+1) = wgh/gamp.  av(:,:,:,1) = F at integer nodes,
   * av(:,:,:,2:D+1) = grad F at the staggered points of E. */
  const double dth = 0.5 * dt_step;
  double ax1[3], axh1[3], ay1[3], ayh1[3], az1[3], azh1[3];
  int i, ih, j, jh, k, kh;
  if (g->ndim == 2) {
    const int k2 = 1;
    for (int64_t n = 0; n < np; n++) {
      double ap[6] = {0, 0, 0, 0, 0, 0};
      qqh_axis(loc_coord(g, PC(part, n, 1), 1), ax1, axh1, &i, &ih);
      qqh_axis(loc_coord(g, PC(part, n, 2), 2), ay1, ayh1, &j, &jh);
      const double u1 = PC(part, n, 3), u2 = PC(part, n, 4);
      int charge;
      float wgh;
      orc_unpack_weight(PC(part, n, 5), &charge, NULL, &wgh);
      for (int j1 = 0; j1 <= 2; j1++) {
        int j2 = j + j1;
        double dvol = ay1[j1];
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i + i1;
          ap[5] = ap[5] + ax1[i1] * dvol * F4(av, i2, j2, k2, 1);
        }
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = ih + i1;
          double dvol1 = dvol * axh1[i1];
          ap[0] = ap[0] + dvol1 * F4(ebf, i2, j2, k2, 1);
          ap[3] = ap[3] + dvol1 * F4(av, i2, j2, k2, 2);
        }
      }
      for (int j1 = 0; j1 <= 2; j1++) {
        int j2 = jh + j1;
        double dvol = ayh1[j1];
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i + i1;
          double dvol1 = dvol * ax1[i1];
          ap[1] = ap[1] + dvol1 * F4(ebf, i2, j2, k2, 2);
          ap[4] = ap[4] + dvol1 * F4(av, i2, j2, k2, 3);
        }
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = ih + i1;
          ap[2] = ap[2] + axh1[i1] * dvol * F4(ebf, i2, j2, k2, 3);
        }
      }
      const double q = (double)charge;
      const double gam2 = 1. + u1 * u1 + u2 * u2 + ap[5];
      for (int c = 0; c < 3; c++) ap[c] = q * ap[c];
      for (int c = 3; c < 5; c++) ap[c] = 0.5 * q * q * ap[c];
      const double aa1 = dth * (ap[0] * u1 + ap[1] * u2);
      const double b1 = dth * (ap[3] * u1 + ap[4] * u2);
      const double gam = sqrt(gam2);
      const double dgam = (aa1 * gam - b1) / gam2;
      const double gam_inv = (gam - dgam) / gam2;
      for (int c = 2; c < 5; c++) ap[c] = ap[c] * gam_inv;
      PC(ebfp, n, 1) = ap[0] - ap[3];
      PC(ebfp, n, 2) = ap[1] - ap[4];
      PC(ebfp, n, 3) = ap[2];
      PC(ebfp, n, 5) = (double)wgh * gam_inv;
    }
    return;
  }
  for (int64_t n = 0; n < np; n++) {
    double ap[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    qqh_axis(loc_coord(g, PC(part, n, 1), 1), ax1, axh1, &i, &ih);
    qqh_axis(loc_coord(g, PC(part, n, 2), 2), ay1, ayh1, &j, &jh);
    qqh_axis(loc_coord(g, PC(part, n, 3), 3), az1, azh1, &k, &kh);
    const double u[3] = {PC(part, n, 4), PC(part, n, 5), PC(part, n, 6)};
    int charge;
    float wgh;
    orc_unpack_weight(PC(part, n, 7), &charge, NULL, &wgh);
    for (int k1 = 0; k1 <= 2; k1++) {
      int k2 = k + k1;
      for (int j1 = 0; j1 <= 2; j1++) {
        int j2 = j + j1;
        double dvol = ay1[j1] * az1[k1];
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i1 + i;
          ap[9] = ap[9] + ax1[i1] * dvol * F4(av, i2, j2, k2, 1);
        }
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i1 + ih;
          double dvol1 = dvol * axh1[i1];
          ap[0] = ap[0] + dvol1 * F4(ebf, i2, j2, k2, 1);
          ap[6] = ap[6] + dvol1 * F4(av, i2, j2, k2, 2);
        }
      }
      for (int j1 = 0; j1 <= 2; j1++) {
        int j2 = jh + j1;
        double dvol = ayh1[j1] * az1[k1];
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i + i1;
          double dvol1 = dvol * ax1[i1];
          ap[1] = ap[1] + dvol1 * F4(ebf, i2, j2, k2, 2);
          ap[7] = ap[7] + dvol1 * F4(av, i2, j2, k2, 3);
        }
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i1 + ih;
          ap[5] = ap[5] + axh1[i1] * dvol * F4(ebf, i2, j2, k2, 6);
        }
      }
    }
    for (int k1 = 0; k1 <= 2; k1++) {
      int k2 = kh + k1;
      for (int j1 = 0; j1 <= 2; j1++) {
        int j2 = jh + j1;
        double dvol = ayh1[j1] * azh1[k1];
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i1 + i;
          ap[3] = ap[3] + ax1[i1] * dvol * F4(ebf, i2, j2, k2, 4);
        }
      }
      for (int j1 = 0; j1 <= 2; j1++) {
        int j2 = j + j1;
        double dvol = ay1[j1] * azh1[k1];
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = ih + i1;
          ap[4] = ap[4] + axh1[i1] * dvol * F4(ebf, i2, j2, k2, 5);
        }
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i1 + i;
          double dvol1 = dvol * ax1[i1];
          ap[2] = ap[2] + dvol1 * F4(ebf, i2, j2, k2, 3);
          ap[8] = ap[8] + dvol1 * F4(av, i2, j2, k2, 4);
        }
      }
    }
    const double q = (double)charge;
    const double gam2 = 1. + u[0] * u[0] + u[1] * u[1] + u[2] * u[2] + ap[9];
    for (int c = 0; c < 6; c++) ap[c] = q * ap[c];
    for (int c = 6; c < 9; c++) ap[c] = 0.5 * q * q * ap[c];
    const double aa1 = dth * (ap[0] * u[0] + ap[1] * u[1] + ap[2] * u[2]);
    const double b1 = dth * (ap[6] * u[0] + ap[7] * u[1] + ap[8] * u[2]);
    const double gam = sqrt(gam2);
    const double dgam = (aa1 * gam - b1) / gam2;
    const double gam_inv = (gam - dgam) / gam2;
    for (int c = 3; c < 9; c++) ap[c] = ap[c] * gam_inv;
    for (int c = 0; c < 3; c++) PC(ebfp, n, c + 1) = ap[c] - ap[c + 6];
    for (int c = 3; c < 6; c++) PC(ebfp, n, c + 1) = ap[c];
    PC(ebfp, n, 7) = (double)wgh * gam_inv;
  }
}

void orc_env_momenta(const orc_grid *g, int64_t np, double *part, int64_t ld, double *ebfp) {
  /* lpf_env_momenta (boris_push.f90:170-223): Boris rotation with the
   * particle's gamp fixed, ebfp(1:D) = q(E - q grad F/(2 gamp)),
   * ebfp(D+1..) = qB/gamp; exit p^{n+1/2} and ebfp(1:D) = x^n. */
  const double dt_lp = g->dt, alp = 0.5 * dt_lp * g->lorentz_fact;
  if (g->ndim == 2) {
    for (int64_t p = 0; p < np; p++) {
      double pp[2] = {PC(part, p, 3), PC(part, p, 4)}, efp[3], vp[2], vph[2];
      for (int c = 0; c < 3; c++) efp[c] = alp * PC(ebfp, p, c + 1);
      vp[0] = pp[0] + efp[0];
      vp[1] = pp[1] + efp[1];
      const double bb1 = efp[2];
      const double b2 = 1. + bb1 * bb1;
      vph[0] = vp[0] + vp[1] * bb1;
      vph[1] = vp[1] - vp[0] * bb1;
      vph[0] = vph[0] / b2;
      vph[1] = vph[1] / b2;
      PC(part, p, 3) = 2. * vph[0] - pp[0];
      PC(part, p, 4) = 2. * vph[1] - pp[1];
      PC(ebfp, p, 1) = PC(part, p, 1);
      PC(ebfp, p, 2) = PC(part, p, 2);
    }
    return;
  }
  for (int64_t p = 0; p < np; p++) {
    double pp[3], efp[6], vp[3], bb[3], vph[3];
    for (int c = 0; c < 3; c++) pp[c] = PC(part, p, c + 4);
    for (int c = 0; c < 6; c++) efp[c] = alp * PC(ebfp, p, c + 1);
    for (int c = 0; c < 3; c++) vp[c] = efp[c] + pp[c];
    for (int c = 0; c < 3; c++) bb[c] = efp[c + 3];
    const double b2 = 1. + (bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2]);
    const double bv = bb[0] * vp[0] + bb[1] * vp[1] + bb[2] * vp[2];
    vph[0] = vp[0] + vp[1] * bb[2] - vp[2] * bb[1] + bb[0] * bv;
    vph[1] = vp[1] + vp[2] * bb[0] - vp[0] * bb[2] + bb[1] * bv;
    vph[2] = vp[2] + vp[0] * bb[1] - vp[1] * bb[0] + bb[2] * bv;
    for (int c = 0; c < 3; c++) vph[c] = vph[c] / b2;
    for (int c = 0; c < 3; c++) PC(part, p, c + 4) = 2. * vph[c] - pp[c];
    for (int c = 0; c < 3; c++) PC(ebfp, p, c + 1) = PC(part, p, c + 1);
  }
}

void orc_env_density(const orc_grid *g, int64_t np, double *ebfp, int64_t ld, double *den) {
  /* set_env_density (grid_part_connect.f90:884-952): den += wgh/gamp with
   * the quadratic density shapes at x^n; ebfp(1:D) = x^n on entry, the
   * local coordinates on exit, ebfp(2D+1) = wgh/gamp. */
  double ax1[3], ay1[3], az1[3] = {1., 0., 0.};
  int i, j, k;
  const int D = g->ndim;
  const int cw = 2 * D + 1;
  for (int64_t n = 0; n < np; n++) {
    for (int d = 1; d <= D; d++) PC(ebfp, n, d) = loc_coord(g, PC(ebfp, n, d), d);
    const double wghp = PC(ebfp, n, cw);
    qden_axis(PC(ebfp, n, 1), ax1, &i);
    qden_axis(PC(ebfp, n, 2), ay1, &j);
    i = i - 1;
    j = j - 1;
    if (D == 2) {
      for (int j1 = 0; j1 <= 2; j1++) {
        int j2 = j + j1;
        double dvol = ay1[j1] * wghp;
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i1 + i;
          double dvol1 = dvol * ax1[i1];
          F4(den, i2, j2, 1, 1) = F4(den, i2, j2, 1, 1) + dvol1;
        }
      }
    } else {
      qden_axis(PC(ebfp, n, 3), az1, &k);
      k = k - 1;
      for (int k1 = 0; k1 <= 2; k1++) {
        int k2 = k + k1;
        for (int j1 = 0; j1 <= 2; j1++) {
          int j2 = j + j1;
          double dvol = ay1[j1] * az1[k1] * wghp;
          for (int i1 = 0; i1 <= 2; i1++) {
            int i2 = i1 + i;
            double dvol1 = dvol * ax1[i1];
            F4(den, i2, j2, k2, 1) = F4(den, i2, j2, k2, 1) + dvol1;
          }
        }
      }
    }
  }
}

void orc_env_grad_interp(const orc_grid *g, int64_t np, const double *part, int64_t ld,
                         const double *av, double *ebfp) {
  /* set_env_grad_interp (grid_part_connect.f90:760-882): grad F and F
   * (F = |a|^2/2 at t^{n+1/2}) at x^n: ebfp(1:D) = grad F, ebfp(D+1) = F. */
  double ax1[3], axh1[3], ay1[3], ayh1[3], az1[3], azh1[3];
  int i, ih, j, jh, k, kh;
  if (g->ndim == 2) {
    const int k2 = 1;
    for (int64_t n = 0; n < np; n++) {
      double ap[4] = {0, 0, 0, 0};
      qqh_axis(loc_coord(g, PC(part, n, 1), 1), ax1, axh1, &i, &ih);
      qqh_axis(loc_coord(g, PC(part, n, 2), 2), ay1, ayh1, &j, &jh);
      for (int j1 = 0; j1 <= 2; j1++) {
        int j2 = j + j1;
        double dvol = ay1[j1];
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i1 + ih;
          double dvol1 = dvol * axh1[i1];
          ap[0] = ap[0] + dvol1 * F4(av, i2, j2, k2, 2);
          i2 = i1 + i;
          ap[2] = ap[2] + ax1[i1] * dvol * F4(av, i2, j2, k2, 1);
        }
      }
      for (int j1 = 0; j1 <= 2; j1++) {
        int j2 = jh + j1;
        double dvol = ayh1[j1];
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i + i1;
          double dvol1 = dvol * ax1[i1];
          ap[1] = ap[1] + dvol1 * F4(av, i2, j2, k2, 3);
        }
      }
      for (int c = 0; c < 3; c++) PC(ebfp, n, c + 1) = ap[c];
    }
    return;
  }
  for (int64_t n = 0; n < np; n++) {
    double ap[4] = {0, 0, 0, 0};
    qqh_axis(loc_coord(g, PC(part, n, 1), 1), ax1, axh1, &i, &ih);
    qqh_axis(loc_coord(g, PC(part, n, 2), 2), ay1, ayh1, &j, &jh);
    qqh_axis(loc_coord(g, PC(part, n, 3), 3), az1, azh1, &k, &kh);
    for (int k1 = 0; k1 <= 2; k1++) {
      int k2 = k + k1;
      for (int j1 = 0; j1 <= 2; j1++) {
        int j2 = j + j1;
        double dvol = ay1[j1] * az1[k1];
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i1 + ih;
          double dvol1 = dvol * axh1[i1];
          ap[0] = ap[0] + dvol1 * F4(av, i2, j2, k2, 2);
          i2 = i1 + i;
          ap[3] = ap[3] + ax1[i1] * dvol * F4(av, i2, j2, k2, 1);
        }
      }
      for (int j1 = 0; j1 <= 2; j1++) {
        int j2 = jh + j1;
        double dvol = ayh1[j1] * az1[k1];
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i + i1;
          double dvol1 = dvol * ax1[i1];
          ap[1] = ap[1] + dvol1 * F4(av, i2, j2, k2, 3);
        }
      }
      k2 = kh + k1;
      for (int j1 = 0; j1 <= 2; j1++) {
        int j2 = j + j1;
        double dvol = ay1[j1] * azh1[k1];
        for (int i1 = 0; i1 <= 2; i1++) {
          int i2 = i + i1;
          double dvol1 = dvol * ax1[i1];
          ap[2] = ap[2] + dvol1 * F4(av, i2, j2, k2, 4);
        }
      }
    }
    for (int c = 0; c < 4; c++) PC(ebfp, n, c + 1) = ap[c];
  }
}

void orc_env_positions(const orc_grid *g, int64_t np, double *part, int64_t ld, double *ebfp) {
  /* lpf_env_positions (boris_push.f90:225-295), iform < 2, no comoving
   * shift: gamp at t^{n+1/2} from p^{n+1/2} and F, corrected by grad F;
   * x^{n+1} = x^n + dt p/gamp; exit ebfp(1:D) = dt v, ebfp(D+1:2D) = x^n,
   * ebfp(2D+1) = the weight slot (the Esirkepov deposit's input). */
  const double dt_lp = g->dt;
  const int D = g->ndim, ch = 2 * D + 1;
  for (int64_t p = 0; p < np; p++) {
    double pp[3] = {0, 0, 0}, vp[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) {
      pp[d] = PC(part, p, D + d + 1);
      vp[d] = PC(ebfp, p, d + 1);
    }
    const double gam2 = D == 2 ? 1. + (pp[0] * pp[0] + pp[1] * pp[1]) + PC(ebfp, p, 3)
                               : 1. + (pp[0] * pp[0] + pp[1] * pp[1] + pp[2] * pp[2]) +
                                     PC(ebfp, p, 4);
    const double gam = sqrt(gam2);
    const double gam3 = gam2 * gam;
    const double b2 = D == 2 ? 0.25 * (pp[0] * vp[0] + pp[1] * vp[1])
                             : 0.25 * (pp[0] * vp[0] + pp[1] * vp[1] + pp[2] * vp[2]);
    double gam_inv = D == 2 ? 1. / gam : 1. / sqrt(gam2);
    gam_inv = gam_inv * (1. - dt_lp * b2 / gam3);
    for (int d = 0; d < D; d++) vp[d] = dt_lp * gam_inv * pp[d];
    for (int d = 0; d < D; d++) PC(ebfp, p, D + d + 1) = PC(part, p, d + 1);
    PC(ebfp, p, ch) = dt_lp * gam_inv;
    for (int d = 0; d < D; d++) PC(part, p, d + 1) = PC(part, p, d + 1) + vp[d];
    for (int d = 0; d < D; d++) PC(ebfp, p, d + 1) = vp[d];
    PC(ebfp, p, ch) = PC(part, p, ch);
  }
}

/* ------------------------------------------------------------------------
 * Envelope model, grid side (one rank, all faces physical, unstretched).
 * ---------------------------------------------------------------------- */

void orc_env_setup(const orc_grid *g, double cfl, double om0, orc_env *e) {
  /* set_field_param (grid_field_param.f90:36-99): nu, hord_der2 (der_ord 3:
   * opt_der2 = -(1-nu^2)/12; der_ord 4: -1/12), opt_der1 (der_ord 2: 1;
   * der_ord 3: (4-nu^2)/3).  hord_der2 is left 0 and opt_der1 1 where the
   * reference does not set them (der_ord 2 / 4): neither is read there. */
  double nu = cfl;
  const double yx_rat = g->dy / g->dx;
  if (g->ndim > 1) nu = cfl * yx_rat / sqrt(yx_rat * yx_rat + (double)(float)g->ndim - 1.);
  e->om0 = om0;
  e->hord_der2 = 0.0;
  e->opt_der1 = 1.0;
  if (g->der_ord == 3) {
    e->hord_der2 = -(1. - nu * nu) / 12.;
    e->opt_der1 = (4. - nu * nu) / 3.;
  } else if (g->der_ord == 4) {
    e->hord_der2 = -1. / 12.;
  }
}

void orc_env_grad(const orc_grid *g, double *av) {
  /* env_grad (grid_fields.f90:426-586) with every face physical (xl_bd ..
   * zr_bd true on one rank) and loc_yg = loc_zg = 1: av(.,1) = F on the
   * interior; exit av(.,2:1+D) = staggered grad F on the interior, with the
   * one-sided forms at the faces, and F extrapolated into the z ghost
   * planes kz1-1, kz2+1 (3-D). */
  const int ix1 = g->ix1, ix2 = g->ix2, jy1 = g->jy1, jy2 = g->jy2, kz1 = g->kz1, kz2 = g->kz2;
  const double a_hcd = 13. / 12., b_hcd = -1. / 24.;
  double ax1, ay1, az1, ax2, ay2, az2;
  if (g->der_ord < 4) {
    ax1 = g->dx_inv;
    ay1 = g->dy_inv;
    az1 = g->dz_inv;
    ax2 = 0.;
    ay2 = 0.;
    az2 = 0.;
  } else {
    ax1 = g->dx_inv * a_hcd;
    ay1 = g->dy_inv * a_hcd;
    az1 = g->dz_inv * a_hcd;
    ax2 = g->dx_inv * b_hcd;
    ay2 = g->dy_inv * b_hcd;
    az2 = g->dz_inv * b_hcd;
  }
  int i01 = ix1, i02 = ix2, j01 = jy1, j02 = jy2, k01 = kz1, k02 = kz2;
  /* x: left face (xl_bd) */
  for (int k = kz1; k <= kz2; k++)
    for (int j = jy1; j <= jy2; j++)
      F4(av, ix1, j, k, 2) = g->dx_inv * (F4(av, ix1 + 1, j, k, 1) - F4(av, ix1, j, k, 1));
  i01 = ix1 + 1;
  /* x: right face (xr_bd) */
  for (int k = kz1; k <= kz2; k++)
    for (int j = jy1; j <= jy2; j++) {
      const int i = ix2 - 1;
      F4(av, i, j, k, 2) = g->dx_inv * (F4(av, i + 1, j, k, 1) - F4(av, i, j, k, 1));
      F4(av, i + 1, j, k, 2) = g->dx_inv * (2. * F4(av, i, j, k, 1) - 3. * F4(av, i - 1, j, k, 1) +
                                            F4(av, i - 2, j, k, 1));
    }
  i02 = ix2 - 2;
  for (int k = kz1; k <= kz2; k++)
    for (int j = jy1; j <= jy2; j++)
      for (int i = i01; i <= i02; i++)
        F4(av, i, j, k, 2) = ax1 * (F4(av, i + 1, j, k, 1) - F4(av, i, j, k, 1));
  if (g->der_ord == 4)
    for (int k = kz1; k <= kz2; k++)
      for (int j = jy1; j <= jy2; j++)
        for (int i = i01; i <= i02; i++)
          F4(av, i, j, k, 2) = F4(av, i, j, k, 2) +
                               ax2 * (F4(av, i + 2, j, k, 1) - F4(av, i + 1, j, k, 1) +
                                      F4(av, i, j, k, 1) - F4(av, i - 1, j, k, 1));
  /* y: right face (yr_bd) first, as the reference orders them */
  for (int k = kz1; k <= kz2; k++) {
    int j = jy2;
    double shy = 1.0 * g->dy_inv;
    for (int i = ix1; i <= ix2; i++)
      F4(av, i, j, k, 3) = shy * (2. * F4(av, i, j, k, 1) - 3. * F4(av, i, j - 1, k, 1) +
                                  F4(av, i, j - 2, k, 1));
    j = jy2 - 1;
    shy = 1.0 * g->dy_inv;
    for (int i = ix1; i <= ix2; i++)
      F4(av, i, j, k, 3) = shy * (F4(av, i, j + 1, k, 1) - F4(av, i, j, k, 1));
  }
  j02 = jy2 - 2;
  {
    const int j = jy1;
    const double shy = 1.0 * g->dy_inv;
    for (int k = kz1; k <= kz2; k++)
      for (int i = ix1; i <= ix2; i++)
        F4(av, i, j, k, 3) = shy * (F4(av, i, j + 1, k, 1) - F4(av, i, j, k, 1));
  }
  j01 = jy1 + 1;
  for (int k = kz1; k <= kz2; k++)
    for (int j = j01; j <= j02; j++) {
      const double shy = 1.0 * ay1;
      for (int i = ix1; i <= ix2; i++)
        F4(av, i, j, k, 3) = shy * (F4(av, i, j + 1, k, 1) - F4(av, i, j, k, 1));
    }
  if (g->der_ord == 4)
    for (int k = kz1; k <= kz2; k++)
      for (int j = j01; j <= j02; j++) {
        const double shp = 1.0 * ay2, shm = 1.0 * ay2;
        for (int i = ix1; i <= ix2; i++)
          F4(av, i, j, k, 3) = F4(av, i, j, k, 3) +
                               shp * (F4(av, i, j + 2, k, 1) - F4(av, i, j + 1, k, 1)) +
                               shm * (F4(av, i, j, k, 1) - F4(av, i, j - 1, k, 1));
      }
  if (g->ndim == 2) return;
  /* z: right face (zr_bd) writes the ghost plane kz2+1 of F first */
  {
    const int k = kz2;
    const double shz = 1.0 * g->dz_inv;
    for (int j = jy1; j <= jy2; j++)
      for (int i = ix1; i <= ix2; i++) {
        F4(av, i, j, k + 1, 1) = 2. * F4(av, i, j, k, 1) - F4(av, i, j, k - 1, 1);
        F4(av, i, j, k, 4) = shz * (F4(av, i, j, k + 1, 1) - F4(av, i, j, k, 1));
      }
  }
  k02 = kz2 - 1;
  {
    const int k = kz1;
    const double shz = 1.0 * g->dz_inv;
    for (int j = jy1; j <= jy2; j++)
      for (int i = ix1; i <= ix2; i++) {
        F4(av, i, j, k - 1, 1) = 2. * F4(av, i, j, k, 1) - F4(av, i, j, k + 1, 1);
        F4(av, i, j, k, 4) = shz * (F4(av, i, j, k + 1, 1) - F4(av, i, j, k, 1));
      }
  }
  k01 = kz1 + 1;
  for (int k = k01; k <= k02; k++) {
    const double shz = 1.0 * az1;
    for (int j = jy1; j <= jy2; j++)
      for (int i = ix1; i <= ix2; i++)
        F4(av, i, j, k, 4) = shz * (F4(av, i, j, k + 1, 1) - F4(av, i, j, k, 1));
  }
  if (g->der_ord == 4)
    for (int k = k01; k <= k02; k++) {
      const double shp = 1.0 * az2, shm = 1.0 * az2;
      for (int j = jy1; j <= jy2; j++)
        for (int i = ix1; i <= ix2; i++)
          F4(av, i, j, k, 4) = F4(av, i, j, k, 4) +
                               shp * (F4(av, i, j, k + 2, 1) - F4(av, i, j, k + 1, 1)) +
                               shm * (F4(av, i, j, k, 1) - F4(av, i, j, k - 1, 1));
    }
}

void orc_env_amp_prepare(const orc_grid *g, const double *env, double *av) {
  /* env_amp_prepare (env_evolve_in_time.f90:145-170): av(.,1) = |A^n|^2/2
   * on the interior, then env_grad. */
  for (int k = g->kz1; k <= g->kz2; k++)
    for (int j = g->jy1; j <= g->jy2; j++)
      for (int i = g->ix1; i <= g->ix2; i++)
        F4(av, i, j, k, 1) = 0.5 * (F4(env, i, j, k, 1) * F4(env, i, j, k, 1) +
                                    F4(env, i, j, k, 2) * F4(env, i, j, k, 2));
  orc_env_grad(g, av);
}

void orc_env_fields_average(const orc_grid *g, const double *env, double *av) {
  /* env_fields_average (env_evolve_in_time.f90:117-143): av(.,1) =
   * |(A^{n+1} + A^n)/2|^2/2 = ((A1+A3)^2 + (A2+A4)^2)/8 on the interior,
   * then env_grad. */
  const double frac = 1.0 / 8.;
  for (int k = g->kz1; k <= g->kz2; k++)
    for (int j = g->jy1; j <= g->jy2; j++)
      for (int i = g->ix1; i <= g->ix2; i++) {
        const double r = F4(env, i, j, k, 1) + F4(env, i, j, k, 3);
        const double m = F4(env, i, j, k, 2) + F4(env, i, j, k, 4);
        F4(av, i, j, k, 1) = frac * (r * r + m * m);
      }
  orc_env_grad(g, av);
}

void orc_env_bds(const orc_grid *g, double *ef, int ptrght, int ptlft, int c1, int c2) {
  /* env_bds (grid_fields.f90:864-1067), open boundaries (ibx = iby = ibz =
   * 0), unstretched (alpha = 1): x ghosts copy the edge value (the one-point
   * stencil COEFF_0); y and z ghosts take the quadratic extrapolation
   * 3 f(1) - 3 f(2) + f(3) written as alpha*(f(3)-f(2)) - 2 f(2) + 3 f(1),
   * the same value in every ghost layer.  The y pass covers the x ghosts
   * just written, the z pass the x and y ghosts. */
  int i1 = g->ix1, i2 = g->ix2, j1 = g->jy1, j2 = g->jy2, k1 = g->kz1, k2 = g->kz2;
  for (int ic = c1; ic <= c2; ic++)
    for (int k = k1; k <= k2; k++)
      for (int j = j1; j <= j2; j++)
        for (int point = ptlft; point >= 1; point--) F4(ef, i1 - point, j, k, ic) = F4(ef, i1, j, k, ic);
  i1 = i1 - ptlft;
  for (int ic = c1; ic <= c2; ic++)
    for (int k = k1; k <= k2; k++)
      for (int j = j1; j <= j2; j++)
        for (int point = 1; point <= ptrght; point++) F4(ef, i2 + point, j, k, ic) = F4(ef, i2, j, k, ic);
  i2 = i2 + ptrght;
  const double alpha = 1.0 / 1.0;
  for (int j = j1 - ptlft; j <= j1 - 1; j++)
    for (int ic = c1; ic <= c2; ic++)
      for (int k = k1; k <= k2; k++)
        for (int i = i1; i <= i2; i++)
          F4(ef, i, j, k, ic) = alpha * (F4(ef, i, j1 + 2, k, ic) - F4(ef, i, j1 + 1, k, ic)) -
                                2 * F4(ef, i, j1 + 1, k, ic) + 3 * F4(ef, i, j1, k, ic);
  j1 = j1 - ptlft;
  for (int j = j2 + 1; j <= j2 + ptrght; j++)
    for (int ic = c1; ic <= c2; ic++)
      for (int k = k1; k <= k2; k++)
        for (int i = i1; i <= i2; i++)
          F4(ef, i, j, k, ic) = alpha * (F4(ef, i, j2 - 2, k, ic) - F4(ef, i, j2 - 1, k, ic)) -
                                2 * F4(ef, i, j2 - 1, k, ic) + 3 * F4(ef, i, j2, k, ic);
  j2 = j2 + ptrght;
  if (g->ndim < 3) return;
  for (int k = k1 - ptlft; k <= k1 - 1; k++)
    for (int ic = c1; ic <= c2; ic++)
      for (int j = j1; j <= j2; j++)
        for (int i = i1; i <= i2; i++)
          F4(ef, i, j, k, ic) = alpha * (F4(ef, i, j, k1 + 2, ic) - F4(ef, i, j, k1 + 1, ic)) -
                                2 * F4(ef, i, j, k1 + 1, ic) + 3 * F4(ef, i, j, k1, ic);
  for (int k = k2 + 1; k <= k2 + ptrght; k++)
    for (int ic = c1; ic <= c2; ic++)
      for (int j = j1; j <= j2; j++)
        for (int i = i1; i <= i2; i++)
          F4(ef, i, j, k, ic) = alpha * (F4(ef, i, j, k2 - 2, ic) - F4(ef, i, j, k2 - 1, ic)) -
                                2 * F4(ef, i, j, k2 - 1, ic) + 3 * F4(ef, i, j, k2, ic);
}

static void orc_pp_lapl(const orc_grid *g, const double *av, double *source, int ic1, int ic2) {
  /* pp_lapl (grid_fields.f90:326-424), unstretched: the transverse
   * Laplacian added to source, y then z, each with the der_ord 4
   * wide-stencil term after the compact one. */
  double cf[2] = {1., 0.};
  if (g->der_ord == 4) {
    cf[0] = 4. / 3.;
    cf[1] = -1. / 12.;
  }
  const double dy2_inv = g->dy_inv * g->dy_inv, dz2_inv = g->dz_inv * g->dz_inv;
  const double dy4_inv[2] = {cf[0] * dy2_inv, cf[1] * dy2_inv};
  const double dz4_inv[2] = {cf[0] * dz2_inv, cf[1] * dz2_inv};
  for (int ic = ic1; ic <= ic2; ic++)
    for (int k = g->kz1; k <= g->kz2; k++)
      for (int j = g->jy1; j <= g->jy2; j++) {
        const double shy = dy4_inv[0] * 1.0, sphy = 1.0, smhy = 1.0;
        for (int i = g->ix1; i <= g->ix2; i++)
          F4(source, i, j, k, ic) =
              F4(source, i, j, k, ic) +
              shy * (sphy * (F4(av, i, j + 1, k, ic) - F4(av, i, j, k, ic)) -
                     smhy * (F4(av, i, j, k, ic) - F4(av, i, j - 1, k, ic)));
      }
  if (g->der_ord == 4)
    for (int ic = ic1; ic <= ic2; ic++)
      for (int k = g->kz1; k <= g->kz2; k++)
        for (int j = g->jy1; j <= g->jy2; j++) {
          const double shy = dy4_inv[1] * 1.0, sphy = 1.0, smhy = 1.0;
          for (int i = g->ix1; i <= g->ix2; i++)
            F4(source, i, j, k, ic) =
                F4(source, i, j, k, ic) +
                shy * (sphy * (F4(av, i, j + 2, k, ic) - F4(av, i, j, k, ic)) -
                       smhy * (F4(av, i, j, k, ic) - F4(av, i, j - 2, k, ic)));
        }
  if (g->ndim < 3) return;
  for (int ic = ic1; ic <= ic2; ic++)
    for (int k = g->kz1; k <= g->kz2; k++) {
      const double shz = dz4_inv[0] * 1.0, sphz = 1.0, smhz = 1.0;
      for (int j = g->jy1; j <= g->jy2; j++)
        for (int i = g->ix1; i <= g->ix2; i++)
          F4(source, i, j, k, ic) =
              F4(source, i, j, k, ic) +
              shz * (sphz * (F4(av, i, j, k + 1, ic) - F4(av, i, j, k, ic)) -
                     smhz * (F4(av, i, j, k, ic) - F4(av, i, j, k - 1, ic)));
    }
  if (g->der_ord == 4)
    for (int ic = ic1; ic <= ic2; ic++)
      for (int k = g->kz1; k <= g->kz2; k++) {
        const double shz = dz4_inv[1] * 1.0, sphz = 1.0, smhz = 1.0;
        for (int j = g->jy1; j <= g->jy2; j++)
          for (int i = g->ix1; i <= g->ix2; i++)
            F4(source, i, j, k, ic) =
                F4(source, i, j, k, ic) +
                shz * (sphz * (F4(av, i, j, k + 2, ic) - F4(av, i, j, k, ic)) -
                       smhz * (F4(av, i, j, k, ic) - F4(av, i, j, k - 2, ic)));
      }
}

void orc_potential_lapl(const orc_grid *g, const orc_env *e, const double *apf, double *curr,
                        int ic1, int ic2) {
  /* potential_lapl (grid_fields.f90:1356-1403): curr += Laplacian(apf),
   * x part with cf = (1 - 4 hord_der2, hord_der2) for der_ord > 2, then
   * pp_lapl for the transverse part. */
  const double dx2 = g->dx_inv * g->dx_inv;
  double cf[2] = {1., 0.0};
  if (g->der_ord > 2) {
    cf[0] = 1. - 4. * e->hord_der2;
    cf[1] = e->hord_der2;
  }
  const double dx4[2] = {cf[0] * dx2, cf[1] * dx2};
  for (int ic = ic1; ic <= ic2; ic++)
    for (int k = g->kz1; k <= g->kz2; k++)
      for (int j = g->jy1; j <= g->jy2; j++)
        for (int i = g->ix1; i <= g->ix2; i++)
          F4(curr, i, j, k, ic) =
              F4(curr, i, j, k, ic) +
              dx4[0] * (F4(apf, i + 1, j, k, ic) + F4(apf, i - 1, j, k, ic) - 2. * F4(apf, i, j, k, ic));
  if (g->der_ord > 2)
    for (int ic = ic1; ic <= ic2; ic++)
      for (int k = g->kz1; k <= g->kz2; k++)
        for (int j = g->jy1; j <= g->jy2; j++)
          for (int i = g->ix1; i <= g->ix2; i++)
            F4(curr, i, j, k, ic) =
                F4(curr, i, j, k, ic) +
                dx4[1] * (F4(apf, i + 2, j, k, ic) + F4(apf, i - 2, j, k, ic) - 2. * F4(apf, i, j, k, ic));
  if (g->ndim > 1) orc_pp_lapl(g, apf, curr, ic1, ic2);
}

void orc_env_maxw_solve(const orc_grid *g, const orc_env *e, double *curr, double *evf,
                        double dtl) {
  /* env_maxw_solve (grid_fields.f90:588-694): on entry curr(1:2) =
   * -ompe chi A^n; S(A) = curr + Laplacian A^n + the centred 2 om0 D_x term
   * (first_ader); F = dt^2 S + 2 A^n - A^{n-1} -/+ kfact A^{n-1}_{I/R};
   * A^{n+1} = (F -/+ kfact F_{I/R})/(1 + kfact^2), A^{n-1} <- A^n.
   * evf(1:2) = A^n, evf(3:4) = A^{n-1}. */
  const double dt2 = dtl * dtl, om0 = e->om0;
  const double kfact = om0 * dtl, k2_fact = 1. / (1. + kfact * kfact), skfact = om0;
  const double dx1_inv = skfact * g->dx_inv;
  double aph_opt[2] = {1., 0.};
  if (g->der_ord == 3) {
    aph_opt[0] = dx1_inv * e->opt_der1;
    aph_opt[1] = dx1_inv * 0.5 * (1. - e->opt_der1);
  }
  orc_potential_lapl(g, e, evf, curr, 1, 2);
  /* first_ader */
  for (int k = g->kz1; k <= g->kz2; k++)
    for (int j = g->jy1; j <= g->jy2; j++)
      for (int i = g->ix1; i <= g->ix2; i++) {
        if (g->der_ord < 3) {
          F4(curr, i, j, k, 1) = F4(curr, i, j, k, 1) -
                                 dx1_inv * (F4(evf, i + 1, j, k, 2) - F4(evf, i - 1, j, k, 2));
          F4(curr, i, j, k, 2) = F4(curr, i, j, k, 2) +
                                 dx1_inv * (F4(evf, i + 1, j, k, 1) - F4(evf, i - 1, j, k, 1));
        } else {
          F4(curr, i, j, k, 1) = F4(curr, i, j, k, 1) -
                                 aph_opt[0] * (F4(evf, i + 1, j, k, 2) - F4(evf, i - 1, j, k, 2)) -
                                 aph_opt[1] * (F4(evf, i + 2, j, k, 2) - F4(evf, i - 2, j, k, 2));
          F4(curr, i, j, k, 2) = F4(curr, i, j, k, 2) +
                                 aph_opt[0] * (F4(evf, i + 1, j, k, 1) - F4(evf, i - 1, j, k, 1)) +
                                 aph_opt[1] * (F4(evf, i + 2, j, k, 1) - F4(evf, i - 2, j, k, 1));
        }
      }
  for (int k = g->kz1; k <= g->kz2; k++)
    for (int j = g->jy1; j <= g->jy2; j++)
      for (int i = g->ix1; i <= g->ix2; i++) {
        F4(curr, i, j, k, 1) = dt2 * F4(curr, i, j, k, 1) + 2. * F4(evf, i, j, k, 1) -
                               F4(evf, i, j, k, 3) + kfact * F4(evf, i, j, k, 4);
        F4(curr, i, j, k, 2) = dt2 * F4(curr, i, j, k, 2) + 2. * F4(evf, i, j, k, 2) -
                               F4(evf, i, j, k, 4) - kfact * F4(evf, i, j, k, 3);
      }
  for (int k = g->kz1; k <= g->kz2; k++)
    for (int j = g->jy1; j <= g->jy2; j++)
      for (int i = g->ix1; i <= g->ix2; i++) {
        F4(evf, i, j, k, 3) = F4(evf, i, j, k, 1);
        F4(evf, i, j, k, 4) = F4(evf, i, j, k, 2);
        F4(evf, i, j, k, 1) = k2_fact * (F4(curr, i, j, k, 1) - kfact * F4(curr, i, j, k, 2));
        F4(evf, i, j, k, 2) = k2_fact * (F4(curr, i, j, k, 2) + kfact * F4(curr, i, j, k, 1));
      }
}

void orc_advance_lpf_envelope(const orc_grid *g, const orc_env *e, double *curr, double *evf) {
  /* advance_lpf_envelope (curr_and_fields_util.f90:262-299) outside the
   * comoving frame (ib = 2: env_maxw_solve): env_bds(evf, 2, 2) on A^n,
   * curr(1:2) = -ompe curr(3) A^n on the interior, the solve. */
  orc_env_bds(g, evf, 2, 2, 1, 2);
  for (int k = g->kz1; k <= g->kz2; k++)
    for (int j = g->jy1; j <= g->jy2; j++)
      for (int i = g->ix1; i <= g->ix2; i++) {
        F4(curr, i, j, k, 1) = -g->ompe * F4(curr, i, j, k, 3) * F4(evf, i, j, k, 1);
        F4(curr, i, j, k, 2) = -g->ompe * F4(curr, i, j, k, 3) * F4(evf, i, j, k, 2);
      }
  orc_env_maxw_solve(g, e, curr, evf, g->dt);
}

void orc_den_zyxbd(const orc_grid *g, double *rho, int ik) {
  /* den_zyxbd (mpi_curr_interface.f90:646-720), open boundaries, one rank:
   * per face the first ghost layer is added to the second interior layer,
   * which is then copied onto the first (z, then y, then x). */
  const int i1 = g->ix1, i2 = g->ix2, j1 = g->jy1, j2 = g->jy2, k1 = g->kz1, k2 = g->kz2;
  if (g->ndim > 2) {
    for (int iy = j1; iy <= j2; iy++)
      for (int ix = i1; ix <= i2; ix++) {
        F4(rho, ix, iy, k1 + 1, ik) = F4(rho, ix, iy, k1 + 1, ik) + F4(rho, ix, iy, k1 - 1, ik);
        F4(rho, ix, iy, k1, ik) = F4(rho, ix, iy, k1 + 1, ik);
      }
    for (int iy = j1; iy <= j2; iy++)
      for (int ix = i1; ix <= i2; ix++) {
        F4(rho, ix, iy, k2 - 1, ik) = F4(rho, ix, iy, k2 - 1, ik) + F4(rho, ix, iy, k2 + 1, ik);
        F4(rho, ix, iy, k2, ik) = F4(rho, ix, iy, k2 - 1, ik);
      }
  }
  for (int k = k1; k <= k2; k++)
    for (int i = i1; i <= i2; i++) {
      F4(rho, i, j1 + 1, k, ik) = F4(rho, i, j1 + 1, k, ik) + F4(rho, i, j1 - 1, k, ik);
      F4(rho, i, j1, k, ik) = F4(rho, i, j1 + 1, k, ik);
    }
  for (int k = k1; k <= k2; k++)
    for (int i = i1; i <= i2; i++) {
      F4(rho, i, j2 - 1, k, ik) = F4(rho, i, j2 - 1, k, ik) + F4(rho, i, j2 + 1, k, ik);
      F4(rho, i, j2, k, ik) = F4(rho, i, j2 - 1, k, ik);
    }
  for (int k = k1; k <= k2; k++)
    for (int j = j1; j <= j2; j++) {
      F4(rho, i1 + 1, j, k, ik) = F4(rho, i1 + 1, j, k, ik) + F4(rho, i1 - 1, j, k, ik);
      F4(rho, i1, j, k, ik) = F4(rho, i1 + 1, j, k, ik);
    }
  for (int k = k1; k <= k2; k++)
    for (int j = j1; j <= j2; j++) {
      F4(rho, i2 - 1, j, k, ik) = F4(rho, i2 - 1, j, k, ik) + F4(rho, i2 + 1, j, k, ik);
      F4(rho, i2, j, k, ik) = F4(rho, i2 - 1, j, k, ik);
    }
}

int64_t orc_env_lpf2_evolve(const orc_grid *g, const orc_env *e, int64_t np, double *part,
                            int64_t ld, double *ebfp, double *ebf, double *env, double *jc) {
  /* env_lpf2_evolve (env_evolve_in_time.f90:196-361) for one species, one
   * rank, no ionization, one colour, no fluid: jc is the reference's
   * jc(n1p, n2p, n3p, curr_ndim + 1), reused as it reuses it (|A|^2/2 and
   * its gradient, the density source, the solver's scratch, the current).
   * Returns the new particle count. */
  const int D = g->curr_ndim;
  const int64_t nn = (int64_t)g->n1p * g->n2p * g->n3p;
  double *c1 = jc, *c3 = jc + 2 * nn;
  memset(jc, 0, sizeof(double) * nn * (D + 1));
  orc_env_amp_prepare(g, env, jc);
  memset(ebfp, 0, sizeof(double) * ld * (2 * D + 1));
  orc_env_acc(g, np, part, ld, ebf, jc, g->dt, ebfp);
  orc_env_momenta(g, np, part, ld, ebfp);
  memset(c1, 0, sizeof(double) * nn);
  orc_env_density(g, np, ebfp, ld, c1);
  orc_den_zyxbd(g, jc, 1);
  memcpy(c3, c1, sizeof(double) * nn);
  orc_advance_lpf_envelope(g, e, jc, env);
  orc_env_fields_average(g, env, jc);
  orc_env_grad_interp(g, np, part, ld, jc, ebfp);
  orc_env_positions(g, np, part, ld, ebfp);
  np = orc_part_dist(g, np, part, ld, ebfp);
  memset(jc, 0, sizeof(double) * nn * (D + 1));
  orc_esirkepov(g, np, part, ld, ebfp, jc);
  orc_curr_collect(g, jc);
  orc_advance_fields(g, ebf, jc);
  return np;
}

/* ---------------------------------------------------------------------------
 * Bunch fields at the start of a run (init_beam_part_distrib.f90:538-602 on a
 * uniform grid, one rank): the bunch charge density in jc(1) becomes its
 * potential by fft_3d_psolv / fft_2d_psolv with sine transforms (ft_mod = 2,
 * sym = 1, psolve.f90:117-247), jc(2) = bet0*jc(1), and initial_beam_fields
 * (init_grid_fields.f90:32-131) makes the bunch fields, which are added to
 * ebf with Bx = 0.
 * ------------------------------------------------------------------------- */

/* FFTW_RODFT00 of length n along a strided line (the r2r plans of
 * modern_fft_lib.F90:196-224): y_k = 2 sum_j x_j sin(pi (j+1)(k+1)/(n+1)),
 * times sc. */
static void orc_dst1(double *v, int n, int64_t st, double sc, double *tmp) {
  for (int k = 0; k < n; k++) {
    double s = 0.0;
    for (int j = 0; j < n; j++)
      s += v[j * st] * sin(acos(-1.0) * (double)(j + 1) * (double)(k + 1) / (double)(n + 1));
    tmp[k] = 2.0 * s;
  }
  for (int k = 0; k < n; k++) v[k * st] = sc * tmp[k];
}

/* ftw1d_sc along axis dir (1, 2, 3) of wb(n1, n2, n3); is < 0 scales by
 * 1/logical_dimension: 2((n1+1)+1) along x (the x call passes n1+1,
 * modern_fft_lib.F90:689), 2(n+1) along y and z (:732, :748). */
static void orc_ftw1d_sc(double *wb, int n1, int n2, int n3, int is, int dir, double *tmp) {
  const int n = dir == 1 ? n1 : (dir == 2 ? n2 : n3);
  const int64_t st = dir == 1 ? 1 : (dir == 2 ? n1 : (int64_t)n1 * n2);
  const double sc = is < 0 ? 1.0 / (double)(2 * ((dir == 1 ? n1 + 1 : n) + 1)) : 1.0;
  for (int k = 0; k < (dir == 3 ? 1 : n3); k++)
    for (int j = 0; j < (dir == 2 ? 1 : n2); j++)
      for (int i = 0; i < (dir == 1 ? 1 : n1); i++)
        orc_dst1(wb + i + (int64_t)n1 * (j + (int64_t)n2 * k), n, st, sc, tmp);
}

/* sine-transform wavenumbers sk(i, 2) = 2 sin(dh*ak/2)/dh, ak = i*pi/l_box
 * (set_grid_param.f90:959-978), i = 1 .. n; l_box = x(nx) - x(1) = (nx-1)dx
 * for an open x (set_grid_param.f90:104-108), ny*dy and nz*dz (:170, :223). */
static double orc_sk(int i, double dh, double l_box) {
  const double wk = acos(-1.0) / l_box;
  const double ak = wk * (double)i;
  return 2. * sin(0.5 * dh * ak) / dh;
}

void orc_sine_transform(double *v, int n) {
  if (n <= 0) return;
  double *tmp = (double *)malloc((size_t)n * sizeof(double));
  orc_dst1(v, n, 1, 1.0, tmp);
  free(tmp);
}

void orc_beam_potential(const orc_grid *g, double gam2, double *jc) {
  const int n1 = g->nx, n2 = g->ny, n3 = g->ndim == 3 ? g->nz : 1;
  const int64_t nw = (int64_t)n1 * n2 * n3;
  double *wb = (double *)calloc((size_t)nw, sizeof(double));
  double *tmp = (double *)malloc((size_t)imax(n1, imax(n2, n3)) * sizeof(double));
  /* x(nx) - x(1), y(ny+1) - y(1), z(nz+1) - z(1) with x(i) = dx*(i-1),
   * y(i) = dy*(i-1-ny/2), z(i) = dz*(i-1-nz/2) (set_grid_param.f90:76-78,
   * 128-130, 187-189) */
  const double lx = g->dx * (double)(g->nx - 1) - g->dx * 0.0;
  const double ly = g->dy * (double)(g->ny - g->ny / 2) - g->dy * (double)(-(g->ny / 2));
  const double lz = g->dz * (double)(g->nz - g->nz / 2) - g->dz * (double)(-(g->nz / 2));
#define WB(i, j, k) wb[(i)-1 + (int64_t)n1 * ((j)-1 + (int64_t)n2 * ((k)-1))]
  /* wb(1:n1, 1:n2, 1:n3) = rho(i1:i2, j1:j2, k1:k2, 1) */
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++)
        WB(i, j, k) = F4(jc, g->ix1 + i - 1, g->jy1 + j - 1, g->kz1 + k - 1, 1);
  if (g->ndim == 3) {
    /* fft_3d_psolv: pftw3d_sc(-1) = x, then pftw2d_sc(-1) = y, z */
    orc_ftw1d_sc(wb, n1, n2, n3, -1, 1, tmp);
    if (n2 > 2) {
      orc_ftw1d_sc(wb, n1, n2, n3, -1, 2, tmp);
      if (n3 > 2) orc_ftw1d_sc(wb, n1, n2, n3, -1, 3, tmp);
    }
    for (int64_t q = 0; q < nw; q++) wb[q] = g->ompe * wb[q];
    /* beam_potential (psolve.f90:66-115), n3_loc > 1 */
    for (int iz = 1; iz <= n3; iz++)
      for (int iy = 1; iy <= n2; iy++) {
        const double sz = orc_sk(iz, g->dz, lz), sy = orc_sk(iy, g->dy, ly);
        const double ak2p = sz * sz + sy * sy;
        if (ak2p > 0.0) {
          for (int ix = 1; ix <= n1; ix++) {
            const double sx = orc_sk(ix, g->dx, lx);
            const double ak2 = ak2p + sx * sx / gam2;
            WB(ix, iy, iz) = WB(ix, iy, iz) / ak2;
          }
        } else {
          for (int ix = 2; ix <= n1; ix++) {
            const double sx = orc_sk(ix, g->dx, lx);
            const double ak2 = sx * sx / gam2;
            WB(ix, iy, iz) = WB(ix, iy, iz) / ak2;
          }
        }
      }
    /* pftw3d_sc(+1) = pftw2d_sc(+1) (z, then y), then x */
    if (n3 > 1) orc_ftw1d_sc(wb, n1, n2, n3, 1, 3, tmp);
    if (n2 > 1) orc_ftw1d_sc(wb, n1, n2, n3, 1, 2, tmp);
    orc_ftw1d_sc(wb, n1, n2, n3, 1, 1, tmp);
  } else {
    /* fft_2d_psolv: pftw2d_sc(-1) = y (n3 = 1: no z) */
    if (n2 > 2) orc_ftw1d_sc(wb, n1, n2, n3, -1, 2, tmp);
    for (int64_t q = 0; q < nw; q++) wb[q] = g->ompe * wb[q];
    /* beam_2d_potential (psolve.f90:38-64); skz = 0 for nz = 1 */
    for (int iy = 1; iy <= n2; iy++) {
      const double sy = orc_sk(iy, g->dy, ly);
      const double ak2p = 0.0 * 0.0 + sy * sy;
      for (int ix = 1; ix <= n1; ix++)
        WB(ix, iy, 1) = ak2p > 0.0 ? WB(ix, iy, 1) / ak2p : 0.0;
    }
    if (n2 > 1) orc_ftw1d_sc(wb, n1, n2, n3, 1, 2, tmp);
  }
  for (int k = 1; k <= n3; k++)
    for (int j = 1; j <= n2; j++)
      for (int i = 1; i <= n1; i++)
        F4(jc, g->ix1 + i - 1, g->jy1 + j - 1, g->kz1 + k - 1, 1) = WB(i, j, k);
#undef WB
  free(tmp);
  free(wb);
}

void orc_initial_beam_fields(const orc_grid *g, double *jc, double *efb, double gam2,
                             double bet0) {
  const int64_t nc = (int64_t)g->n1p * g->n2p * g->n3p;
  const int ix1 = g->ix1, ix2 = g->ix2, jy1 = g->jy1, jy2 = g->jy2, kz1 = g->kz1,
            kz2 = g->kz2;
#define EF(i, j, k, c) efb[((int64_t)(c)-1) * nc + ((i)-1) + (int64_t)g->n1p * (((j)-1) + (int64_t)g->n2p * ((k)-1))]
  const double dx_inv = g->dx_inv, dy_inv = g->dy_inv, dz_inv = g->dz_inv;
  for (int k = kz1; k <= kz2; k++) /* pe1y */
    for (int i = ix1; i <= ix2; i++)
      F4(jc, i, jy2 + 1, k, 1) = 3. * (F4(jc, i, jy2, k, 1) - F4(jc, i, jy2 - 1, k, 1)) +
                                 F4(jc, i, jy2 - 2, k, 1);
  for (int k = kz1; k <= kz2; k++) /* pe1x */
    for (int j = jy1; j <= jy2; j++) {
      F4(jc, ix2 + 1, j, k, 1) = F4(jc, ix2, j, k, 1);
      F4(jc, ix2 + 1, j, k, 2) = 2. * F4(jc, ix2, j, k, 2) - F4(jc, ix2 - 1, j, k, 2);
    }
  for (int k = kz1; k <= kz2; k++) /* jx(i+1/2): component 4, zeroed by the caller */
    for (int j = jy1; j <= jy2; j++)
      for (int i = ix1; i <= ix2; i++)
        EF(i, j, k, 4) = 0.5 * (F4(jc, i, j, k, 2) + F4(jc, i + 1, j, k, 2));
  for (int k = kz1; k <= kz2; k++)
    for (int j = jy1; j <= jy2; j++) {
      const double sdhy = 1.0 * dy_inv; /* loc_yg(jj, 3, imody) = 1, unstretched */
      for (int i = ix1; i <= ix2 + 1; i++)
        EF(i, j, k, 2) = -sdhy * (F4(jc, i, j + 1, k, 1) - F4(jc, i, j, k, 1));
      for (int i = ix1; i <= ix2; i++)
        EF(i, j, k, 1) = -dx_inv * (F4(jc, i + 1, j, k, 1) - F4(jc, i, j, k, 1)) / gam2;
    }
  if (g->ndim == 2) {
    for (int k = kz1; k <= kz2; k++)
      for (int j = jy1; j <= jy2; j++)
        for (int i = ix1; i <= ix2; i++)
          EF(i, j, k, 3) = 0.5 * bet0 * (EF(i, j, k, 2) + EF(i + 1, j, k, 2));
  } else {
    for (int j = jy1; j <= jy2; j++) /* pe1z */
      for (int i = ix1; i <= ix2; i++)
        F4(jc, i, j, kz2 + 1, 1) = 2. * F4(jc, i, j, kz2, 1) - F4(jc, i, j, kz2 - 1, 1);
    for (int k = kz1; k <= kz2; k++)
      for (int j = jy1; j <= jy2; j++)
        for (int i = ix1; i <= ix2; i++)
          EF(i, j, k, 6) = 0.5 * bet0 * (EF(i, j, k, 2) + EF(i + 1, j, k, 2));
    for (int k = kz1; k <= kz2; k++) {
      const double sdhz = 1.0 * dz_inv;
      for (int j = jy1; j <= jy2; j++)
        for (int i = ix1; i <= ix2; i++)
          EF(i, j, k, 3) = -sdhz * (F4(jc, i, j, k + 1, 1) - F4(jc, i, j, k, 1));
    }
    for (int k = kz1; k <= kz2; k++)
      for (int j = jy1; j <= jy2; j++)
        for (int i = ix1; i <= ix2; i++) {
          F4(jc, i, j, k, 2) = 0.5 * (EF(i, j, k, 3) + EF(i + 1, j, k, 3));
          EF(i, j, k, 5) = -bet0 * F4(jc, i, j, k, 2);
        }
  }
#undef EF
}

void orc_beam_fields(const orc_grid *g, double gam2, double bet0, double *jc, double *ebf) {
  orc_beam_potential(g, gam2, jc);
  /* jc(:, :, :, 2) = bet0*jc(:, :, :, 1) */
  const int64_t nc = (int64_t)g->n1p * g->n2p * g->n3p;
  for (int64_t q = 0; q < nc; q++) jc[nc + q] = bet0 * jc[q];
  /* initial_beam_fields(jc, efb, gam2, bet0) into efb = 0 (ebf_bunch is
   * zeroed when allocated, init_beam_part_distrib.f90:488-491) */
  if (nc <= 0) return;
  double *efb = (double *)calloc((size_t)nc * 6, sizeof(double));
  orc_initial_beam_fields(g, jc, efb, gam2, bet0);
  /* ebf_bunch(:,:,:,4) = 0; ebf(1:nfield) += ebf_bunch(1:nfield) */
  for (int64_t q = 0; q < nc; q++) efb[3 * nc + q] = 0.0;
  for (int c = 0; c < g->nfield; c++)
    for (int64_t q = 0; q < nc; q++) ebf[(int64_t)c * nc + q] += efb[(int64_t)c * nc + q];
  free(efb);
}
