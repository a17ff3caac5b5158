// lp_common.hpp -- device-side constants and index helpers shared by the
// field, particle and reduction kernels of the LP step.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aladyn {

// Run constants passed by value to every kernel.  Index bounds are the
// reference's 1-based Fortran ones (ix1 = 3 .. ix2 = nx+2, code_util.f90:37).
struct DevGrid {
  int ndim, nfield, curr_ndim;
  int nx, ny, nz;
  int n1p, n2p, n3p;
  int ix1, ix2, jy1, jy2, kz1, kz2;
  int ibx, iby, ibz;
  int64_t ncell;  // n1p*n2p*n3p: stride between components of ebf / jc
  double dx, dy, dz, dx_inv, dy_inv, dz_inv;
  double xmn, ymn, zmn, xmx, ymx, zmx;
  double dt, ompe;
  double djc[3];
  double se[2];
  double dgvol;
};

// Fortran (i,j,k) 1-based -> linear offset inside one component.
__host__ __device__ inline int64_t fofs(const DevGrid& g, int i, int j, int k) {
  return (int64_t)(i - 1) + (int64_t)g.n1p * ((int64_t)(j - 1) + (int64_t)g.n2p * (int64_t)(k - 1));
}

// Component c (1-based) of an ebf / jc style array.
template <typename T>
__device__ inline T* comp(T* a, const DevGrid& g, int c) {
  return a + (int64_t)(c - 1) * g.ncell;
}

// Quadratic weights at integer nodes from the offset s = xx - ix in [-1/2,1/2)
// (grid_part_lib.f90:117-121): a1 = 3/4 - s^2, a2 = (1/4 + s^2 + s)/2,
// a0 = 1 - a1 - a2.
template <typename T>
__device__ inline void qweights(T s, T a[3]) {
  T s2 = s * s;
  a[1] = T(0.75) - s2;
  a[2] = T(0.5) * (T(0.25) + s2 + s);
  a[0] = T(1) - a[1] - a[2];
}

// Linear weights at half-integer nodes (grid_part_lib.f90:123-124).
template <typename T>
__device__ inline void lweights(T s, T ah[2]) {
  ah[1] = s + T(0.5);
  ah[0] = T(1) - ah[1];
}

// One axis of the Esirkepov inversion (grid_part_connect.f90:1007-1025):
// W1 placed in the 5-slot window at ih = i_new - i_old + 1 in {0,1,2},
// cur[s] = -sum_{m<=s} (W1 - W0)[m], not yet scaled by the weight.
template <typename T>
__device__ inline void esk_axis(const T a0[3], const T a1[3], int ih, T wh[5], T cur[5]) {
#pragma unroll
  for (int s = 0; s < 5; s++) wh[s] = T(0);
#pragma unroll
  for (int s = 0; s < 3; s++) {
    // ih is 0,1 or 2: select without dynamic register indexing
    if (ih == 0) wh[s] = a1[s];
    if (ih == 1) wh[s + 1] = a1[s];
    if (ih == 2) wh[s + 2] = a1[s];
  }
  cur[0] = -wh[0];
#pragma unroll
  for (int s = 1; s <= 3; s++) cur[s] = cur[s - 1] + a0[s - 1] - wh[s];
  cur[4] = cur[3] - wh[4];
}

}  // namespace aladyn
