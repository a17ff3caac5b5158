// fields.hip -- grid kernels of the LP step: ghost-layer extrapolation
// (field_xyzbd), current ghost fold + normalisation (jc_xyzbd, djc), Mur-type
// absorbing boundaries (ef_bds/bf_bds, imbd = 0) and the Yee curl updates
// (rote, rotb + current source).  Templated on the storage type (double for
// the FP64 path, float for the FP32 path); all arrays are in the
// reference's Fortran layout ebf(n1p,n2p,n3p,nc), x fastest, so consecutive
// lanes touch consecutive addresses.
#include <algorithm>

#include "lp_kernels.hpp"

namespace aladyn {

static constexpr int kFieldBlock = 256;

static inline unsigned nblocks(int64_t n) {
  int64_t b = (n + kFieldBlock - 1) / kFieldBlock;
  return (unsigned)(b < 1 ? 1 : (b > 65535 * 16 ? 65535 * 16 : b));
}

// ---------------------------------------------------------------- field_xyzbd
// mpi_field_interface.f90:53-198 on one rank.  Three launches (x, y, z faces)
// because the y faces extend over the x ghosts and z over both.
template <typename T>
__device__ inline void b_field_xbd(DevGrid g, T* ebf, int nc) {
  const int nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
  const int64_t n = (int64_t)nj * nk * nc;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    int j = g.jy1 + (int)(t % nj);
    int k = g.kz1 + (int)((t / nj) % nk);
    int ik = 1 + (int)(t / ((int64_t)nj * nk));
    T* f = comp(ebf, g, ik);
    if (g.ibx < 2)
      f[fofs(g, g.ix1 - 1, j, k)] = T(2) * f[fofs(g, g.ix1, j, k)] - f[fofs(g, g.ix1 + 1, j, k)];
    if (g.ibx == 0)
      f[fofs(g, g.ix2 + 1, j, k)] = T(2) * f[fofs(g, g.ix2, j, k)] - f[fofs(g, g.ix2 - 1, j, k)];
    if (g.ibx == 1) {
      const T par = (ik == 1 || ik == 4) ? T(-1) : T(1);  // x_parity
      f[fofs(g, g.ix2 + 1, j, k)] = par * f[fofs(g, g.ix2, j, k)];
    }
  }
}
template <typename T>
__global__ void k_field_xbd(DevGrid g, T* ebf, int nc) { b_field_xbd<T>(g, ebf, nc); }
template <typename T>
__global__ void kb_field_xbd(const FieldRun<T>* runs) {
  const FieldRun<T>& r = runs[blockIdx.y];
  b_field_xbd<T>(r.g, r.ebf, r.g.nfield);
}

template <typename T>
__device__ inline void b_field_ybd(DevGrid g, T* ebf, int nc) {
  const int i1 = g.ix1 - 1, ni = g.ix2 - g.ix1 + 3, nk = g.kz2 - g.kz1 + 1;
  const int64_t n = (int64_t)ni * nk * nc;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    int i = i1 + (int)(t % ni);
    int k = g.kz1 + (int)((t / ni) % nk);
    int ik = 1 + (int)(t / ((int64_t)ni * nk));
    T* f = comp(ebf, g, ik);
    if (g.iby == 0) {
      f[fofs(g, i, g.jy1 - 1, k)] = T(2) * f[fofs(g, i, g.jy1, k)] - f[fofs(g, i, g.jy1 + 1, k)];
      f[fofs(g, i, g.jy2 + 1, k)] = T(2) * f[fofs(g, i, g.jy2, k)] - f[fofs(g, i, g.jy2 - 1, k)];
    } else if (g.iby == 1) {
      const T par = (ik == 2 || ik == 5) ? T(-1) : T(1);  // y_parity
      f[fofs(g, i, g.jy1 - 1, k)] = par * f[fofs(g, i, g.jy1 + 1, k)];
      f[fofs(g, i, g.jy2 + 1, k)] = par * f[fofs(g, i, g.jy2 - 1, k)];
    }
  }
}
template <typename T>
__global__ void k_field_ybd(DevGrid g, T* ebf, int nc) { b_field_ybd<T>(g, ebf, nc); }
template <typename T>
__global__ void kb_field_ybd(const FieldRun<T>* runs) {
  const FieldRun<T>& r = runs[blockIdx.y];
  b_field_ybd<T>(r.g, r.ebf, r.g.nfield);
}

template <typename T>
__device__ inline void b_field_zbd(DevGrid g, T* ebf, int nc) {
  const int i1 = g.ix1 - 1, ni = g.ix2 - g.ix1 + 3;
  const int j1 = g.jy1 - 1, nj = g.jy2 - g.jy1 + 3;
  const int64_t n = (int64_t)ni * nj * nc;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    int i = i1 + (int)(t % ni);
    int j = j1 + (int)((t / ni) % nj);
    int ik = 1 + (int)(t / ((int64_t)ni * nj));
    T* f = comp(ebf, g, ik);
    if (g.ibz == 0) {
      f[fofs(g, i, j, g.kz1 - 1)] = T(2) * f[fofs(g, i, j, g.kz1)] - f[fofs(g, i, j, g.kz1 + 1)];
      f[fofs(g, i, j, g.kz2 + 1)] = T(2) * f[fofs(g, i, j, g.kz2)] - f[fofs(g, i, j, g.kz2 - 1)];
    } else if (g.ibz == 1) {
      const T par = (ik == 3 || ik == 6) ? T(-1) : T(1);  // z_parity
      f[fofs(g, i, j, g.kz1 - 1)] = par * f[fofs(g, i, j, g.kz1 + 1)];
      // the reference writes kz1+1 from kz1-1 here (mpi_field_interface.f90:191)
      f[fofs(g, i, j, g.kz1 + 1)] = par * f[fofs(g, i, j, g.kz1 - 1)];
    }
  }
}
template <typename T>
__global__ void k_field_zbd(DevGrid g, T* ebf, int nc) { b_field_zbd<T>(g, ebf, nc); }
template <typename T>
__global__ void kb_field_zbd(const FieldRun<T>* runs) {
  const FieldRun<T>& r = runs[blockIdx.y];
  b_field_zbd<T>(r.g, r.ebf, r.g.nfield);
}

template <typename T>
void field_xyzbd(const DevGrid& g, T* ebf, int nc, hipStream_t s) {
  int64_t nx = (int64_t)(g.jy2 - g.jy1 + 1) * (g.kz2 - g.kz1 + 1) * nc;
  k_field_xbd<T><<<nblocks(nx), kFieldBlock, 0, s>>>(g, ebf, nc);
  if (g.ndim < 2) return;
  int64_t ny = (int64_t)(g.ix2 - g.ix1 + 3) * (g.kz2 - g.kz1 + 1) * nc;
  k_field_ybd<T><<<nblocks(ny), kFieldBlock, 0, s>>>(g, ebf, nc);
  if (g.ndim < 3) return;
  int64_t nz = (int64_t)(g.ix2 - g.ix1 + 3) * (g.jy2 - g.jy1 + 3) * nc;
  k_field_zbd<T><<<nblocks(nz), kFieldBlock, 0, s>>>(g, ebf, nc);
}

// ------------------------------------------------------------------- jc_xyzbd
// mpi_curr_interface.f90:551-643 on one rank (ib* = 0 folds).
template <typename T>
__device__ inline void b_jc_xfold(DevGrid g, T* jc) {
  const int nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1, nc = g.curr_ndim;
  const int64_t n = (int64_t)nj * nk * nc;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    int j = g.jy1 + (int)(t % nj);
    int k = g.kz1 + (int)((t / nj) % nk);
    int ik = 1 + (int)(t / ((int64_t)nj * nk));
    T* c = comp(jc, g, ik);
    c[fofs(g, g.ix1, j, k)] = c[fofs(g, g.ix1, j, k)] + c[fofs(g, g.ix1 - 1, j, k)];
    c[fofs(g, g.ix1 - 1, j, k)] = T(0);
    c[fofs(g, g.ix2, j, k)] =
        c[fofs(g, g.ix2, j, k)] + c[fofs(g, g.ix2 + 1, j, k)] + c[fofs(g, g.ix2 + 2, j, k)];
    c[fofs(g, g.ix2 + 1, j, k)] = T(0);
    c[fofs(g, g.ix2 + 2, j, k)] = T(0);
  }
}
template <typename T>
__global__ void k_jc_xfold(DevGrid g, T* jc) { b_jc_xfold<T>(g, jc); }
template <typename T>
__global__ void kb_jc_xfold(const FieldRun<T>* runs) {
  const FieldRun<T>& r = runs[blockIdx.y];
  b_jc_xfold<T>(r.g, r.jc);
}

template <typename T>
__device__ inline void b_jc_yfold(DevGrid g, T* jc) {
  const int ni = g.ix2 - g.ix1 + 1, nk = g.kz2 - g.kz1 + 1, nc = g.curr_ndim;
  const int64_t n = (int64_t)ni * nk * nc;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    int i = g.ix1 + (int)(t % ni);
    int k = g.kz1 + (int)((t / ni) % nk);
    int ik = 1 + (int)(t / ((int64_t)ni * nk));
    T* c = comp(jc, g, ik);
    const int j1 = g.jy1, n2 = g.jy2;
    c[fofs(g, i, j1, k)] =
        c[fofs(g, i, j1, k)] + c[fofs(g, i, j1 - 1, k)] + c[fofs(g, i, j1 - 2, k)];
    c[fofs(g, i, j1 - 2, k)] = T(0);
    c[fofs(g, i, j1 - 1, k)] = T(0);
    c[fofs(g, i, n2, k)] =
        c[fofs(g, i, n2, k)] + c[fofs(g, i, n2 + 1, k)] + c[fofs(g, i, n2 + 2, k)];
    c[fofs(g, i, n2 + 1, k)] = T(0);
    c[fofs(g, i, n2 + 2, k)] = T(0);
  }
}
template <typename T>
__global__ void k_jc_yfold(DevGrid g, T* jc) { b_jc_yfold<T>(g, jc); }
template <typename T>
__global__ void kb_jc_yfold(const FieldRun<T>* runs) {
  const FieldRun<T>& r = runs[blockIdx.y];
  b_jc_yfold<T>(r.g, r.jc);
}

template <typename T>
__device__ inline void b_jc_zfold(DevGrid g, T* jc) {
  const int ni = g.ix2 - g.ix1 + 1, nj = g.jy2 - g.jy1 + 1, nc = g.curr_ndim;
  const int64_t n = (int64_t)ni * nj * nc;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    int i = g.ix1 + (int)(t % ni);
    int j = g.jy1 + (int)((t / ni) % nj);
    int ik = 1 + (int)(t / ((int64_t)ni * nj));
    T* c = comp(jc, g, ik);
    const int k1 = g.kz1, n3 = g.kz2;
    c[fofs(g, i, j, k1)] = c[fofs(g, i, j, k1)] + c[fofs(g, i, j, k1 - 1)];
    c[fofs(g, i, j, k1 - 1)] = T(0);
    c[fofs(g, i, j, n3)] =
        c[fofs(g, i, j, n3)] + c[fofs(g, i, j, n3 + 1)] + c[fofs(g, i, j, n3 + 2)];
    c[fofs(g, i, j, n3 + 1)] = T(0);
    c[fofs(g, i, j, n3 + 2)] = T(0);
  }
}
template <typename T>
__global__ void k_jc_zfold(DevGrid g, T* jc) { b_jc_zfold<T>(g, jc); }
template <typename T>
__global__ void kb_jc_zfold(const FieldRun<T>* runs) {
  const FieldRun<T>& r = runs[blockIdx.y];
  b_jc_zfold<T>(r.g, r.jc);
}

// curr(:,:,:,i) = djc(i)*curr(:,:,:,i), i = 1..ndim (curr_and_fields_util.f90:115-119)
template <typename T>
__device__ inline void b_jc_scale(DevGrid g, T* jc) {
  const int64_t n = g.ncell;
  for (int c = 0; c < g.ndim; c++) {
    const T f = T(g.djc[c]);
    T* a = jc + c * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x)
      a[t] = f * a[t];
  }
}
template <typename T>
__global__ void k_jc_scale(DevGrid g, T* jc) { b_jc_scale<T>(g, jc); }
template <typename T>
__global__ void kb_jc_scale(const FieldRun<T>* runs) {
  const FieldRun<T>& r = runs[blockIdx.y];
  b_jc_scale<T>(r.g, r.jc);
}

template <typename T>
void curr_collect(const DevGrid& g, T* jc, hipStream_t s) {
  if (g.ibx == 0) {
    int64_t n = (int64_t)(g.jy2 - g.jy1 + 1) * (g.kz2 - g.kz1 + 1) * g.curr_ndim;
    k_jc_xfold<T><<<nblocks(n), kFieldBlock, 0, s>>>(g, jc);
  }
  if (g.iby == 0) {
    int64_t n = (int64_t)(g.ix2 - g.ix1 + 1) * (g.kz2 - g.kz1 + 1) * g.curr_ndim;
    k_jc_yfold<T><<<nblocks(n), kFieldBlock, 0, s>>>(g, jc);
  }
  if (g.ndim > 2 && g.ibz == 0) {
    int64_t n = (int64_t)(g.ix2 - g.ix1 + 1) * (g.jy2 - g.jy1 + 1) * g.curr_ndim;
    k_jc_zfold<T><<<nblocks(n), kFieldBlock, 0, s>>>(g, jc);
  }
  k_jc_scale<T><<<nblocks(g.ncell), kFieldBlock, 0, s>>>(g, jc);
}

// ------------------------------------------------------------ ef_bds / bf_bds
// grid_fields.f90:1204-1354 (E ghosts on the right faces) and 1069-1202
// (B ghosts on the left faces), imbd = 0, unstretched metric = 1.  Each
// thread owns one face cell; faces do not read what another face writes.
template <typename T>
__device__ inline void b_ef_bds(DevGrid g, T* ebf, double dtl) {
  const int nf = g.nfield;
  const int ni = g.ix2 - g.ix1 + 1, nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
  const int64_t nxf = (g.ibx < 2) ? (int64_t)nj * nk : 0;
  const int64_t nyf = (g.iby < 2) ? (int64_t)ni * nk : 0;
  const int64_t nzf = (g.ndim > 2 && g.ibz < 2) ? (int64_t)ni * nj : 0;
  T* ex = comp(ebf, g, 1);
  T* ey = comp(ebf, g, 2);
  T* ez = comp(ebf, g, 3);
  T* bx = comp(ebf, g, 4);
  T* by = comp(ebf, g, 5);
  T* bz = comp(ebf, g, nf);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nxf + nyf + nzf;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (t < nxf) {
      int j = g.jy1 + (int)(t % nj), k = g.kz1 + (int)(t / nj);
      const int i = g.ix2;
      if (g.ibx == 0) {
        const T a = T(1.0 * g.dx_inv * dtl);
        ey[fofs(g, i + 1, j, k)] =
            (T(2) * bz[fofs(g, i, j, k)] - (T(1) - a) * ey[fofs(g, i, j, k)]) / (T(1) + a);
        if (nf > 3)
          ez[fofs(g, i + 1, j, k)] =
              -(T(2) * by[fofs(g, i, j, k)] + (T(1) - a) * ez[fofs(g, i, j, k)]) / (T(1) + a);
      } else {
        ey[fofs(g, i + 1, j, k)] = ey[fofs(g, i, j, k)];
        if (nf > 3) ez[fofs(g, i + 1, j, k)] = ez[fofs(g, i, j, k)];
      }
    } else if (t < nxf + nyf) {
      int64_t u = t - nxf;
      int i = g.ix1 + (int)(u % ni), k = g.kz1 + (int)(u / ni);
      const int j = g.jy2;
      const T a = T(1.0 * g.dy_inv * dtl);
      ex[fofs(g, i, j + 1, k)] =
          -(T(2) * bz[fofs(g, i, j, k)] + (T(1) - a) * ex[fofs(g, i, j, k)]) / (T(1) + a);
      if (nf > 3)
        ez[fofs(g, i, j + 1, k)] =
            (T(2) * bx[fofs(g, i, j, k)] - (T(1) - a) * ez[fofs(g, i, j, k)]) / (T(1) + a);
    } else {
      int64_t u = t - nxf - nyf;
      int i = g.ix1 + (int)(u % ni), j = g.jy1 + (int)(u / ni);
      const int k = g.kz2;
      const T a = T(1.0 * g.dz_inv * dtl);
      ex[fofs(g, i, j, k + 1)] =
          (T(2) * by[fofs(g, i, j, k)] - (T(1) - a) * ex[fofs(g, i, j, k)]) / (T(1) + a);
      ey[fofs(g, i, j, k + 1)] =
          -(T(2) * bx[fofs(g, i, j, k)] + (T(1) - a) * ey[fofs(g, i, j, k)]) / (T(1) + a);
    }
  }
}
template <typename T>
__global__ void k_ef_bds(DevGrid g, T* ebf, double dtl) { b_ef_bds<T>(g, ebf, dtl); }
template <typename T>
__global__ void kb_ef_bds(const FieldRun<T>* runs, int mode) {
  const FieldRun<T>& r = runs[blockIdx.y];
  b_ef_bds<T>(r.g, r.ebf, mode ? r.g.dt : 0.0);
}

template <typename T>
__device__ inline void b_bf_bds(DevGrid g, T* ebf, double dtl) {
  const int nf = g.nfield;
  const int ni = g.ix2 - g.ix1 + 1, nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
  const int64_t nxf = (g.ibx < 2) ? (int64_t)nj * nk : 0;
  const int64_t nyf = (g.iby < 2) ? (int64_t)ni * nk : 0;
  const int64_t nzf = (g.ndim > 2 && g.ibz < 2) ? (int64_t)ni * nj : 0;
  T* ex = comp(ebf, g, 1);
  T* ey = comp(ebf, g, 2);
  T* ez = comp(ebf, g, 3);
  T* bx = comp(ebf, g, 4);
  T* by = comp(ebf, g, 5);
  T* bz = comp(ebf, g, nf);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nxf + nyf + nzf;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (t < nxf) {
      int j = g.jy1 + (int)(t % nj), k = g.kz1 + (int)(t / nj);
      const int i = g.ix1;
      const T a = T(1.0 * g.dx_inv * dtl);
      bz[fofs(g, i - 1, j, k)] =
          -(T(2) * ey[fofs(g, i, j, k)] + (T(1) - a) * bz[fofs(g, i, j, k)]) / (T(1) + a);
      if (nf > 3)
        by[fofs(g, i - 1, j, k)] =
            (T(2) * ez[fofs(g, i, j, k)] - (T(1) - a) * by[fofs(g, i, j, k)]) / (T(1) + a);
    } else if (t < nxf + nyf) {
      int64_t u = t - nxf;
      int i = g.ix1 + (int)(u % ni), k = g.kz1 + (int)(u / ni);
      const int j = g.jy1;
      const T a = T(1.0 * g.dy_inv * dtl);
      T v = T(2) * ex[fofs(g, i, j, k)] - (T(1) - a) * bz[fofs(g, i, j, k)];
      bz[fofs(g, i, j - 1, k)] = v / (T(1) + a);
      if (nf > 3) {
        v = T(-2) * ez[fofs(g, i, j, k)] - (T(1) - a) * bx[fofs(g, i, j, k)];
        bx[fofs(g, i, j - 1, k)] = v / (T(1) + a);
      }
    } else {
      int64_t u = t - nxf - nyf;
      int i = g.ix1 + (int)(u % ni), j = g.jy1 + (int)(u / ni);
      const int k = g.kz1;
      const T a = T(1.0 * g.dz_inv * dtl);
      T v = T(2) * ey[fofs(g, i, j, k)] - (T(1) - a) * bx[fofs(g, i, j, k)];
      bx[fofs(g, i, j, k - 1)] = v / (T(1) + a);
      v = T(-2) * ex[fofs(g, i, j, k)] - (T(1) - a) * by[fofs(g, i, j, k)];
      by[fofs(g, i, j, k - 1)] = v / (T(1) + a);
    }
  }
}
template <typename T>
__global__ void k_bf_bds(DevGrid g, T* ebf, double dtl) { b_bf_bds<T>(g, ebf, dtl); }
template <typename T>
__global__ void kb_bf_bds(const FieldRun<T>* runs) {
  const FieldRun<T>& r = runs[blockIdx.y];
  b_bf_bds<T>(r.g, r.ebf, r.g.dt);
}

// --------------------------------------------------------------- rote / rotb
// B -= dtf*curl E on the interior (grid_fields.f90:1405-1478).
template <typename T>
__device__ inline void b_rote(DevGrid g, T* ebf, double dtf) {
  const int nf = g.nfield;
  const int ni = g.ix2 - g.ix1 + 1, nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
  const int64_t n = (int64_t)ni * nj * nk;
  const T aphx = T(dtf * g.dx_inv), aphy = T(dtf * g.dy_inv), aphz = T(dtf * g.dz_inv);
  const T aph1 = T(double(dtf * g.dx_inv) * g.se[0]), aph2 = T(double(dtf * g.dx_inv) * g.se[1]);
  (void)aphx;
  const T* ex = comp(ebf, g, 1);
  const T* ey = comp(ebf, g, 2);
  const T* ez = comp(ebf, g, 3);
  T* bx = comp(ebf, g, 4);
  T* by = comp(ebf, g, 5);
  T* bz = comp(ebf, g, nf);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = g.ix1 + (int)(t % ni);
    const int j = g.jy1 + (int)((t / ni) % nj);
    const int k = g.kz1 + (int)(t / ((int64_t)ni * nj));
    const int64_t o = fofs(g, i, j, k);
    const int64_t sy = g.n1p, sz = (int64_t)g.n1p * g.n2p;
    const T sdhy = aphy;
    bz[o] = bz[o] - aph1 * (ey[o + 1] - ey[o]) + sdhy * (ex[o + sy] - ex[o]) -
            aph2 * (ey[o + 2] - ey[o - 1]);
    if (nf >= 6) {
      const T sdhz = aphz;
      bx[o] = bx[o] - sdhy * (ez[o + sy] - ez[o]) + sdhz * (ey[o + sz] - ey[o]);
      by[o] = by[o] + aph1 * (ez[o + 1] - ez[o]) - sdhz * (ex[o + sz] - ex[o]) +
              aph2 * (ez[o + 2] - ez[o - 1]);
    }
  }
}
template <typename T>
__global__ void k_rote(DevGrid g, T* ebf, double dtf) { b_rote<T>(g, ebf, dtf); }
template <typename T>
__global__ void kb_rote(const FieldRun<T>* runs) {
  const FieldRun<T>& r = runs[blockIdx.y];
  b_rote<T>(r.g, r.ebf, 0.5 * r.g.dt);
}

// E += dtf*curl B on the interior (grid_fields.f90:1480-1552) followed by the
// source term E(1:curr_ndim) -= ompe*jc (curr_and_fields_util.f90:233-236).
template <typename T>
__device__ inline void b_rotb_source(DevGrid g, T* ebf, const T* jc, double dtf) {
  const int nf = g.nfield;
  const int ni = g.ix2 - g.ix1 + 1, nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
  const int64_t n = (int64_t)ni * nj * nk;
  const T aphy = T(dtf * g.dy_inv), aphz = T(dtf * g.dz_inv);
  const T aph1 = T(double(dtf * g.dx_inv) * g.se[0]), aph2 = T(double(dtf * g.dx_inv) * g.se[1]);
  const T ompe = T(g.ompe);
  T* ex = comp(ebf, g, 1);
  T* ey = comp(ebf, g, 2);
  T* ez = comp(ebf, g, 3);
  const T* bx = comp(ebf, g, 4);
  const T* by = comp(ebf, g, 5);
  const T* bz = comp(ebf, g, nf);
  const T* jx = jc;
  const T* jy = jc + g.ncell;
  const T* jz = jc + 2 * g.ncell;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = g.ix1 + (int)(t % ni);
    const int j = g.jy1 + (int)((t / ni) % nj);
    const int k = g.kz1 + (int)(t / ((int64_t)ni * nj));
    const int64_t o = fofs(g, i, j, k);
    const int64_t sy = g.n1p, sz = (int64_t)g.n1p * g.n2p;
    const T sdy = aphy;
    T vx = ex[o] + sdy * (bz[o] - bz[o - sy]);
    T vy = ey[o] - aph1 * (bz[o] - bz[o - 1]) - aph2 * (bz[o + 1] - bz[o - 2]);
    if (nf >= 6) {
      const T sdz = aphz;
      vx = vx - sdz * (by[o] - by[o - sz]);
      vy = vy + sdz * (bx[o] - bx[o - sz]);
      T vz = ez[o] + aph1 * (by[o] - by[o - 1]) - sdy * (bx[o] - bx[o - sy]) +
             aph2 * (by[o + 1] - by[o - 2]);
      if (g.curr_ndim >= 3) vz = vz - ompe * jz[o];
      ez[o] = vz;
    }
    ex[o] = vx - ompe * jx[o];
    ey[o] = vy - ompe * jy[o];
  }
}
template <typename T>
__global__ void k_rotb_source(DevGrid g, T* ebf, const T* jc, double dtf) {
  b_rotb_source<T>(g, ebf, jc, dtf);
}
template <typename T>
__global__ void kb_rotb_source(const FieldRun<T>* runs) {
  const FieldRun<T>& r = runs[blockIdx.y];
  b_rotb_source<T>(r.g, r.ebf, r.jc, r.g.dt);
}

template <typename T>
void advance_fields(const DevGrid& g, T* ebf, const T* jc, hipStream_t s) {
  const double dt_lp = g.dt, dth = 0.5 * dt_lp;
  const int ni = g.ix2 - g.ix1 + 1, nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
  const int64_t nint = (int64_t)ni * nj * nk;
  const int64_t nface = (int64_t)nj * nk + (int64_t)ni * nk + (g.ndim > 2 ? (int64_t)ni * nj : 0);
  k_ef_bds<T><<<nblocks(nface), kFieldBlock, 0, s>>>(g, ebf, 0.0);
  k_rote<T><<<nblocks(nint), kFieldBlock, 0, s>>>(g, ebf, dth);
  k_bf_bds<T><<<nblocks(nface), kFieldBlock, 0, s>>>(g, ebf, dt_lp);
  k_rotb_source<T><<<nblocks(nint), kFieldBlock, 0, s>>>(g, ebf, jc, dt_lp);
  k_ef_bds<T><<<nblocks(nface), kFieldBlock, 0, s>>>(g, ebf, dt_lp);
  k_rote<T><<<nblocks(nint), kFieldBlock, 0, s>>>(g, ebf, dth);
}

// ------------------------------------------------------- batched launches
BatchDims batch_dims(const DevGrid* gs, int n) {
  BatchDims d{};
  for (int r = 0; r < n; r++) {
    const DevGrid& g = gs[r];
    const int64_t ni = g.ix2 - g.ix1 + 1, nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
    const int64_t nf = g.nfield, nc = g.curr_ndim;
    d.fx = std::max(d.fx, nj * nk * nf);
    d.fy = std::max(d.fy, (ni + 2) * nk * nf);
    d.fz = std::max(d.fz, (ni + 2) * (nj + 2) * nf);
    d.jx = std::max(d.jx, nj * nk * nc);
    d.jy = std::max(d.jy, ni * nk * nc);
    d.jz = std::max(d.jz, ni * nj * nc);
    d.ncell = std::max(d.ncell, g.ncell);
    d.nface = std::max(d.nface, nj * nk + ni * nk + (g.ndim > 2 ? ni * nj : 0));
    d.nint = std::max(d.nint, ni * nj * nk);
  }
  return d;
}

// x extent of a batched launch: enough workgroups for the largest run, at
// most 16384 in all (each run's workgroups stride over its own work), so the
// many small runs of a scan do not each schedule the largest run's grid
static inline dim3 bgrid(int64_t n, int nruns) {
  const unsigned cap = (unsigned)std::max(16, 16384 / std::max(nruns, 1));
  return dim3(std::min(nblocks(n), cap), nruns);
}

template <typename T>
void field_xyzbd_batch(const FieldRun<T>* runs, int nruns, int ndim, const BatchDims& d,
                       hipStream_t s) {
  kb_field_xbd<T><<<bgrid(d.fx, nruns), kFieldBlock, 0, s>>>(runs);
  kb_field_ybd<T><<<bgrid(d.fy, nruns), kFieldBlock, 0, s>>>(runs);
  if (ndim > 2) kb_field_zbd<T><<<bgrid(d.fz, nruns), kFieldBlock, 0, s>>>(runs);
}

// open boundaries on every face (the only ones a context accepts)
template <typename T>
void curr_collect_batch(const FieldRun<T>* runs, int nruns, int ndim, const BatchDims& d,
                        hipStream_t s) {
  kb_jc_xfold<T><<<bgrid(d.jx, nruns), kFieldBlock, 0, s>>>(runs);
  kb_jc_yfold<T><<<bgrid(d.jy, nruns), kFieldBlock, 0, s>>>(runs);
  if (ndim > 2) kb_jc_zfold<T><<<bgrid(d.jz, nruns), kFieldBlock, 0, s>>>(runs);
  kb_jc_scale<T><<<bgrid(d.ncell, nruns), kFieldBlock, 0, s>>>(runs);
}

template <typename T>
void advance_fields_batch(const FieldRun<T>* runs, int nruns, const BatchDims& d, hipStream_t s) {
  kb_ef_bds<T><<<bgrid(d.nface, nruns), kFieldBlock, 0, s>>>(runs, 0);
  kb_rote<T><<<bgrid(d.nint, nruns), kFieldBlock, 0, s>>>(runs);
  kb_bf_bds<T><<<bgrid(d.nface, nruns), kFieldBlock, 0, s>>>(runs);
  kb_rotb_source<T><<<bgrid(d.nint, nruns), kFieldBlock, 0, s>>>(runs);
  kb_ef_bds<T><<<bgrid(d.nface, nruns), kFieldBlock, 0, s>>>(runs, 1);
  kb_rote<T><<<bgrid(d.nint, nruns), kFieldBlock, 0, s>>>(runs);
}

template <typename T>
__global__ void k_fill(T* a, int64_t n, T v) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    a[t] = v;
}

template <typename T>
void fill(T* a, int64_t n, T v, hipStream_t s) {
  k_fill<T><<<nblocks(n), kFieldBlock, 0, s>>>(a, n, v);
}

// fields_left_xshift (mpi_field_interface.f90:299-359), one rank: for every
// component and interior row (jy1..jy2, kz1..kz2), f(ix) = f(ix + shx) for
// ix = ix1 .. ix2 - shx and f = 0 on the last shx columns ix2-shx+1 .. ix2.
// One workgroup per row walks it left to right in blocks: a block of
// sources is read before its destinations (shx columns to the left) are
// written, and later blocks only read columns further right.
template <typename T>
__global__ __launch_bounds__(kFieldBlock) void k_left_xshift(DevGrid g, T* ebf, int shx) {
  const int ny = g.jy2 - g.jy1 + 1, nz = g.kz2 - g.kz1 + 1;
  const int row = blockIdx.x;
  const int iy = g.jy1 + row % ny, iz = g.kz1 + (row / ny) % nz, ic = row / (ny * nz) + 1;
  T* f = comp(ebf, g, ic) + fofs(g, 1, iy, iz) - 1;  // f[i] = ebf(i, iy, iz, ic)
  const int i1 = g.ix1, i2 = g.ix2 - shx;
  for (int b = i1; b <= g.ix2; b += kFieldBlock) {
    const int ix = b + (int)threadIdx.x;
    T v = T(0);
    if (ix <= i2) v = f[ix + shx];
    __syncthreads();
    if (ix <= g.ix2) f[ix] = v;
    __syncthreads();
  }
}

template <typename T>
void fields_left_xshift(const DevGrid& g, T* ebf, int shx, hipStream_t s, int ncomp) {
  if (shx <= 0) return;
  if (ncomp <= 0) ncomp = g.nfield;
  const int rows = (g.jy2 - g.jy1 + 1) * (g.kz2 - g.kz1 + 1) * ncomp;
  k_left_xshift<T><<<rows, kFieldBlock, 0, s>>>(g, ebf, shx);
}

#define ALADYN_INST_FIELDS(T)                                              \
  template void fields_left_xshift<T>(const DevGrid&, T*, int, hipStream_t, int); \
  template void field_xyzbd<T>(const DevGrid&, T*, int, hipStream_t);      \
  template void curr_collect<T>(const DevGrid&, T*, hipStream_t);          \
  template void advance_fields<T>(const DevGrid&, T*, const T*, hipStream_t); \
  template void fill<T>(T*, int64_t, T, hipStream_t);                                        \
  template void field_xyzbd_batch<T>(const FieldRun<T>*, int, int, const BatchDims&,          \
                                     hipStream_t);                                             \
  template void curr_collect_batch<T>(const FieldRun<T>*, int, int, const BatchDims&,         \
                                      hipStream_t);                                            \
  template void advance_fields_batch<T>(const FieldRun<T>*, int, const BatchDims&, hipStream_t);

ALADYN_INST_FIELDS(double)
ALADYN_INST_FIELDS(float)
template void fill<int>(int*, int64_t, int, hipStream_t);

}  // namespace aladyn
