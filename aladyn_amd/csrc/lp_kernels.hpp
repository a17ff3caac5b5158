// lp_kernels.hpp -- host-side launchers of the LP-step kernels and the
// device functions shared by the tile kernel and the overflow kernel.
#pragma once

#include "lp_common.hpp"

namespace aladyn {

// ---------------------------------------------------------------- fields.hip
template <typename T> void field_xyzbd(const DevGrid& g, T* ebf, int nc, hipStream_t s);
// jc_xyzbd, then jc *= djc unless the fold that wrote jc applied it already
template <typename T>
void curr_collect(const DevGrid& g, T* jc, hipStream_t s, bool normalised = false);
template <typename T> void advance_fields(const DevGrid& g, T* ebf, const T* jc, hipStream_t s);
template <typename T> void fill(T* a, int64_t n, T v, hipStream_t s);
// ncomp components of the grid's layout (0: the grid's nfield)
template <typename T>
void fields_left_xshift(const DevGrid& g, T* ebf, int shx, hipStream_t s, int ncomp = 0);

// Batched launches: one launch covers the same kernel of many independent
// runs, blockIdx.y selecting the run.  FieldRun is one run's grids; the
// BatchDims are the largest per-run work counts, which size the x extent.
template <typename T>
struct FieldRun {
  DevGrid g;
  T* ebf;
  T* jc;
};
struct BatchDims {
  int64_t fx, fy, fz;   // field ghost faces (times nfield)
  int64_t jx, jy, jz;   // current ghost folds (times curr_ndim)
  int64_t ncell, nface, nint;
  bool any_wide;        // some run has the wide (der_ord > 2) x stencil
};
BatchDims batch_dims(const DevGrid* g, int n);
template <typename T>
void field_xyzbd_batch(const FieldRun<T>* runs, int nruns, int ndim, const BatchDims& d,
                       hipStream_t s);
// jc_xyzbd only: the batched fold applies djc as it writes jc
template <typename T>
void curr_collect_batch(const FieldRun<T>* runs, int nruns, int ndim, const BatchDims& d,
                        hipStream_t s);
template <typename T>
void advance_fields_batch(const FieldRun<T>* runs, int nruns, const BatchDims& d, hipStream_t s);

// ----------------------------------------------------------------- beam.hip
// The bunch fields of charge density rho (FP64, ncell; its interior becomes
// the potential) added to ebf (init_beam_part_distrib.f90:538-602); waits
// for the stream.
template <typename R>
hipError_t beam_fields(const DevGrid& g, double* rho, double gam2, double bet0, R* ebf,
                       hipStream_t s);

// ---------------------------------------------------------- reductions.hip
template <typename T>
void field_energy(const DevGrid& g, const T* ebf, double* out, hipStream_t s);
// r = rho1 - rho0 + (1/curr_ndim) sum_d D_d J_d on the interior; out[0] =
// max|r|, out[1] = max|rho1 - rho0| (atomic max on non-negative doubles).
template <typename T>
void continuity(const DevGrid& g, const double* rho0, const double* rho1, const T* jraw,
                double* out, hipStream_t s);

// ------------------------------------------------ shared device functions
// Reference qlh/qden arithmetic on an absolute FP64 coordinate:
// xx = 3 + dinv*(x - xmn); ix = int(xx + 1/2); s = xx - ix.
__device__ inline void ref_node(double x, double xmn, double dinv, int& ix, double& s) {
  double xx = 3.0 + dinv * (x - xmn);
  ix = (int)(xx + 0.5);
  s = xx - (double)ix;
}

// Esirkepov deposit for one particle moving from (node io[d], offset so[d])
// to (in[d], sn[d]) with charge*weight wq, in the reference's algebra
// (grid_part_connect.f90:979-1060 for 2-D, 1168-1334 for 3-D), the old and new
// transverse windows merged into one union window so that every grid cell
// receives one accumulate per particle.  acc(c, i, j, k, v) adds v to
// component c (0-based) at Fortran node (i, j, k).
template <int D, typename T, class Acc>
__device__ inline void esirkepov(const int io[D], const T so[D], const int in[D], const T sn[D],
                                 T wq, Acc& acc) {
  T a0[D][3], a1[D][3];
#pragma unroll
  for (int d = 0; d < D; d++) {
    qweights<T>(so[d], a0[d]);
    qweights<T>(sn[d], a1[d]);
  }
  T wh[D][5], cur[D][5];
  int lo[D], hi[D];
#pragma unroll
  for (int d = 0; d < D; d++) {
    const int ih = in[d] - io[d] + 1;  // 0, 1 or 2
    esk_axis<T>(a0[d], a1[d], ih, wh[d], cur[d]);
#pragma unroll
    for (int s = 0; s < 5; s++) cur[d][s] = wq * cur[d][s];
    lo[d] = ih < 1 ? ih : 1;
    hi[d] = ih + 2 > 3 ? ih + 2 : 3;
  }
  if constexpr (D == 2) {
    // 2-D, two current components: J_x = currx (x) (W0y + W1y),
    // J_y = curry (x) (W0x + W1x), both on the 5x5 window at (i0-2, j0-2).
    T sx[5], sy[5];
#pragma unroll
    for (int s = 0; s < 5; s++) {
      sx[s] = wh[0][s];
      sy[s] = wh[1][s];
    }
#pragma unroll
    for (int s = 1; s <= 3; s++) {
      sx[s] = sx[s] + a0[0][s - 1];
      sy[s] = sy[s] + a0[1][s - 1];
    }
    const int ib = io[0] - 2, jb = io[1] - 2;
#pragma unroll
    for (int j1 = 0; j1 < 5; j1++) {
      if (j1 < lo[1] || j1 > hi[1]) continue;
#pragma unroll
      for (int i1 = 0; i1 < 5; i1++) {
        if (i1 < lo[0] || i1 > hi[0]) continue;
        acc(0, ib + i1, jb + j1, 1, sy[j1] * cur[0][i1]);
        acc(1, ib + i1, jb + j1, 1, sx[i1] * cur[1][j1]);
      }
    }
  } else {
    // 3-D: half-weighted transverse windows (axh0/axh1, ayh0/ayh1) as in
    // esirkepov_3d_curr, z (for J_x, J_y) and y (for J_z) merged over the
    // union of the old and new 3-node windows.
    T h0[2][5], h1[2][5];  // [x|y][slot]
#pragma unroll
    for (int d = 0; d < 2; d++) {
#pragma unroll
      for (int s = 0; s < 5; s++) {
        h0[d][s] = T(0.5) * wh[d][s];
        h1[d][s] = wh[d][s];
      }
#pragma unroll
      for (int s = 1; s <= 3; s++) {
        h0[d][s] = h0[d][s] + a0[d][s - 1];
        h1[d][s] = h1[d][s] + T(0.5) * a0[d][s - 1];
      }
    }
    // union windows: z for J_x/J_y, y for J_z; base = min(old,new) - 1
    const int zb = (io[2] < in[2] ? io[2] : in[2]) - 1;
    const int zo0 = io[2] - 1 - zb, zo1 = in[2] - 1 - zb;  // 0 or 1
    const int yb = (io[1] < in[1] ? io[1] : in[1]) - 1;
    const int yo0 = io[1] - 1 - yb, yo1 = in[1] - 1 - yb;
    T Z0[4], Z1[4], Y0[4], Y1[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
      Z0[m] = zo0 == 0 ? (m < 3 ? a0[2][m] : T(0)) : (m > 0 ? a0[2][m - 1] : T(0));
      Z1[m] = zo1 == 0 ? (m < 3 ? a1[2][m] : T(0)) : (m > 0 ? a1[2][m - 1] : T(0));
      Y0[m] = yo0 == 0 ? (m < 3 ? a0[1][m] : T(0)) : (m > 0 ? a0[1][m - 1] : T(0));
      Y1[m] = yo1 == 0 ? (m < 3 ? a1[1][m] : T(0)) : (m > 0 ? a1[1][m - 1] : T(0));
    }
    const int zn = (io[2] == in[2]) ? 3 : 4;
    const int yn = (io[1] == in[1]) ? 3 : 4;
    const int ib = io[0] - 2, jb = io[1] - 2, kb = io[2] - 2;
    // J_x(ib+i1, jb+j1, zb+m) += currx[i1]*(ayh0[j1]*Z0[m] + ayh1[j1]*Z1[m])
    // J_y(ib+i1, jb+j1, zb+m) += curry[j1]*(axh0[i1]*Z0[m] + axh1[i1]*Z1[m])
#pragma unroll
    for (int m = 0; m < 4; m++) {
      if (m >= zn) continue;
#pragma unroll
      for (int j1 = 0; j1 < 5; j1++) {
        if (j1 < lo[1] || j1 > hi[1]) continue;
        const T cx = h0[1][j1] * Z0[m] + h1[1][j1] * Z1[m];
        const T cy0 = cur[1][j1] * Z0[m], cy1 = cur[1][j1] * Z1[m];
#pragma unroll
        for (int i1 = 0; i1 < 5; i1++) {
          if (i1 < lo[0] || i1 > hi[0]) continue;
          acc(0, ib + i1, jb + j1, zb + m, cx * cur[0][i1]);
          acc(1, ib + i1, jb + j1, zb + m, h0[0][i1] * cy0 + h1[0][i1] * cy1);
        }
      }
    }
    // J_z(ib+i1, yb+m, kb+k1) += currz[k1]*(axh0[i1]*Y0[m] + axh1[i1]*Y1[m])
#pragma unroll
    for (int k1 = 0; k1 < 5; k1++) {
      if (k1 < lo[2] || k1 > hi[2]) continue;
#pragma unroll
      for (int m = 0; m < 4; m++) {
        if (m >= yn) continue;
        const T cz0 = Y0[m] * cur[2][k1], cz1 = Y1[m] * cur[2][k1];
#pragma unroll
        for (int i1 = 0; i1 < 5; i1++) {
          if (i1 < lo[0] || i1 > hi[0]) continue;
          acc(2, ib + i1, yb + m, kb + k1, h0[0][i1] * cz0 + h1[0][i1] * cz1);
        }
      }
    }
  }
}

}  // namespace aladyn
