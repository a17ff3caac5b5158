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

// field_xyzbd with open boundaries on every axis in one launch: one thread
// per ghost node the three face kernels above write (the x ghosts at interior
// j, k; the y ghosts over the x-extended i range at interior k; the z ghosts
// over the x- and y-extended ranges), which forms its value from interior
// nodes in the order the face kernels do (x extrapolations, then y, then
// z), so the result is the same to the bit and no thread reads a node
// another writes.
template <typename T>
__device__ inline T ghost_x(const T* f, const DevGrid& g, int i, int j, int k) {
  if (i == g.ix1 - 1) return T(2) * f[fofs(g, g.ix1, j, k)] - f[fofs(g, g.ix1 + 1, j, k)];
  if (i == g.ix2 + 1) return T(2) * f[fofs(g, g.ix2, j, k)] - f[fofs(g, g.ix2 - 1, j, k)];
  return f[fofs(g, i, j, k)];
}
template <typename T>
__device__ inline T ghost_xy(const T* f, const DevGrid& g, int i, int j, int k) {
  if (j == g.jy1 - 1) return T(2) * ghost_x(f, g, i, g.jy1, k) - ghost_x(f, g, i, g.jy1 + 1, k);
  if (j == g.jy2 + 1) return T(2) * ghost_x(f, g, i, g.jy2, k) - ghost_x(f, g, i, g.jy2 - 1, k);
  return ghost_x(f, g, i, j, k);
}
template <typename T>
__global__ __launch_bounds__(kFieldBlock) void k_field_bd_open(DevGrid g, T* ebf, int nc) {
  const int ni = g.ix2 - g.ix1 + 1, nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
  const int64_t nxs = 2LL * nj * nk, nys = g.ndim > 1 ? 2LL * (ni + 2) * nk : 0;
  const int64_t nzs = g.ndim > 2 ? 2LL * (ni + 2) * (nj + 2) : 0;
  const int64_t per = nxs + nys + nzs;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < per * nc;
       t += (int64_t)gridDim.x * blockDim.x) {
    T* f = comp(ebf, g, 1 + (int)(t / per));
    int64_t r = t % per;
    int i, j, k;
    T v;
    if (r < nxs) {  // x ghosts: (ix1-1 | ix2+1, j, k)
      const int side = (int)(r / ((int64_t)nj * nk));
      r -= (int64_t)side * nj * nk;
      j = g.jy1 + (int)(r % nj);
      k = g.kz1 + (int)(r / nj);
      i = side == 0 ? g.ix1 - 1 : g.ix2 + 1;
      v = ghost_x(f, g, i, j, k);
    } else if ((r -= nxs) < nys) {  // y ghosts over i = ix1-1 .. ix2+1
      const int side = (int)(r / ((int64_t)(ni + 2) * nk));
      r -= (int64_t)side * (ni + 2) * nk;
      i = g.ix1 - 1 + (int)(r % (ni + 2));
      k = g.kz1 + (int)(r / (ni + 2));
      j = side == 0 ? g.jy1 - 1 : g.jy2 + 1;
      v = ghost_xy(f, g, i, j, k);
    } else {  // z ghosts over i = ix1-1 .. ix2+1, j = jy1-1 .. jy2+1
      r -= nys;
      const int side = (int)(r / ((int64_t)(ni + 2) * (nj + 2)));
      r -= (int64_t)side * (ni + 2) * (nj + 2);
      i = g.ix1 - 1 + (int)(r % (ni + 2));
      j = g.jy1 - 1 + (int)(r / (ni + 2));
      if (side == 0) {
        k = g.kz1 - 1;
        v = T(2) * ghost_xy(f, g, i, j, g.kz1) - ghost_xy(f, g, i, j, g.kz1 + 1);
      } else {
        k = g.kz2 + 1;
        v = T(2) * ghost_xy(f, g, i, j, g.kz2) - ghost_xy(f, g, i, j, g.kz2 - 1);
      }
    }
    f[fofs(g, i, j, k)] = v;
  }
}

template <typename T>
void field_xyzbd(const DevGrid& g, T* ebf, int nc, hipStream_t s) {
  if (g.ibx == 0 && g.iby == 0 && (g.ndim < 3 || g.ibz == 0) && g.ndim > 1) {
    const int64_t ni = g.ix2 - g.ix1 + 1, nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
    const int64_t n = (2 * nj * nk + 2 * (ni + 2) * nk +
                       (g.ndim > 2 ? 2 * (ni + 2) * (nj + 2) : 0)) * nc;
    k_field_bd_open<T><<<nblocks(n), kFieldBlock, 0, s>>>(g, ebf, nc);
    return;
  }
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

// jc_xyzbd's three folds in one launch: one thread per node on the faces of
// the interior box (x faces whole; y faces without the x-face nodes; z faces
// without either), which adds the ghosts of every fold that reaches it in
// the order of the three kernels above (x, then y, then z) and zeroes them.
// Each ghost is read by one such node only, so no thread reads what another
// writes, and the sums are the same to the bit.
template <typename T>
__global__ __launch_bounds__(kFieldBlock) void k_jc_fold_all(DevGrid g, T* jc, int fx, int fy,
                                                             int fz) {
  const int ni = g.ix2 - g.ix1 + 1, nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
  const int xi = ni > 1 ? 2 : 1, yj = nj > 1 ? 2 : 1, zk = g.ndim > 2 && nk > 1 ? 2 : 1;
  const int mi = max(ni - 2, 0), mj = max(nj - 2, 0);
  const int64_t nxs = (int64_t)xi * nj * nk, nys = (int64_t)yj * mi * nk;
  const int64_t nzs = g.ndim > 2 ? (int64_t)zk * mi * mj : 0;
  const int64_t per = nxs + nys + nzs;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < per * g.curr_ndim;
       t += (int64_t)gridDim.x * blockDim.x) {
    T* c = comp(jc, g, 1 + (int)(t / per));
    int64_t r = t % per;
    int i, j, k;
    if (r < nxs) {  // x faces, i = ix1 | ix2
      const int side = (int)(r / ((int64_t)nj * nk));
      r -= (int64_t)side * nj * nk;
      i = side == 0 ? g.ix1 : g.ix2;
      j = g.jy1 + (int)(r % nj);
      k = g.kz1 + (int)(r / nj);
    } else if ((r -= nxs) < nys) {  // y faces, j = jy1 | jy2, ix1 < i < ix2
      const int side = (int)(r / ((int64_t)mi * nk));
      r -= (int64_t)side * mi * nk;
      j = side == 0 ? g.jy1 : g.jy2;
      i = g.ix1 + 1 + (int)(r % mi);
      k = g.kz1 + (int)(r / mi);
    } else {  // z faces, k = kz1 | kz2, interior of the y and x faces
      r -= nys;
      const int side = (int)(r / ((int64_t)mi * mj));
      r -= (int64_t)side * mi * mj;
      k = side == 0 ? g.kz1 : g.kz2;
      i = g.ix1 + 1 + (int)(r % mi);
      j = g.jy1 + 1 + (int)(r / mi);
    }
    const int64_t o = fofs(g, i, j, k);
    T v = c[o];
    const int64_t sy = g.n1p, sz = (int64_t)g.n1p * g.n2p;
    if (fx) {
      if (i == g.ix1) {
        v = v + c[o - 1];
        c[o - 1] = T(0);
      }
      if (i == g.ix2) {
        v = v + c[o + 1] + c[o + 2];
        c[o + 1] = T(0);
        c[o + 2] = T(0);
      }
    }
    if (fy) {
      if (j == g.jy1) {
        v = v + c[o - sy] + c[o - 2 * sy];
        c[o - 2 * sy] = T(0);
        c[o - sy] = T(0);
      }
      if (j == g.jy2) {
        v = v + c[o + sy] + c[o + 2 * sy];
        c[o + sy] = T(0);
        c[o + 2 * sy] = T(0);
      }
    }
    if (fz) {
      if (k == g.kz1) {
        v = v + c[o - sz];
        c[o - sz] = T(0);
      }
      if (k == g.kz2) {
        v = v + c[o + sz] + c[o + 2 * sz];
        c[o + sz] = T(0);
        c[o + 2 * sz] = T(0);
      }
    }
    c[o] = v;
  }
}

template <typename T>
void curr_collect(const DevGrid& g, T* jc, hipStream_t s, bool normalised) {
  if (g.ndim > 1) {
    const int fx = g.ibx == 0, fy = g.iby == 0, fz = g.ndim > 2 && g.ibz == 0;
    const int64_t ni = g.ix2 - g.ix1 + 1, nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
    const int64_t mi = std::max<int64_t>(ni - 2, 0), mj = std::max<int64_t>(nj - 2, 0);
    const int64_t n = ((ni > 1 ? 2 : 1) * nj * nk + (nj > 1 ? 2 : 1) * mi * nk +
                       (g.ndim > 2 ? (nk > 1 ? 2 : 1) * mi * mj : 0)) *
                      g.curr_ndim;
    if (fx || fy || fz) k_jc_fold_all<T><<<nblocks(n), kFieldBlock, 0, s>>>(g, jc, fx, fy, fz);
    if (!normalised) k_jc_scale<T><<<nblocks(g.ncell), kFieldBlock, 0, s>>>(g, jc);
    return;
  }
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
  if (!normalised) k_jc_scale<T><<<nblocks(g.ncell), kFieldBlock, 0, s>>>(g, jc);
}

// ------------------------------------------------------------ ef_bds / bf_bds
// grid_fields.f90:1204-1354 (E ghosts on the right faces) and 1069-1202
// (B ghosts on the left faces), imbd = 0, unstretched metric = 1.  Every
// ghost value is one of two first-order Mur forms of the interior node's B
// and E (or E and B) and a = dtl/dh:
//   mur_p(u, v) = (2u - (1-a)v)/(1+a),  mur_m(u, v) = -(2u + (1-a)v)/(1+a).
template <typename T>
__device__ inline T mur_p(T u, T v, T a) {
  return (T(2) * u - (T(1) - a) * v) / (T(1) + a);
}
template <typename T>
__device__ inline T mur_m(T u, T v, T a) {
  return -(T(2) * u + (T(1) - a) * v) / (T(1) + a);
}

// ef_bds alone: one thread per face cell; faces do not read what another
// face writes.  Used before the first Yee half-step when the x stencil
// is wide (der_ord > 2), where a node reads the ghost two nodes to its right.
template <typename T>
__device__ inline void b_ef_bds(DevGrid g, T* ebf, double dtl) {
  const int nf = g.nfield;
  const int ni = g.ix2 - g.ix1 + 1, nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
  const int64_t nxf = (int64_t)nj * nk;
  const int64_t nyf = (int64_t)ni * nk;
  const int64_t nzf = g.ndim > 2 ? (int64_t)ni * nj : 0;
  T* ex = comp(ebf, g, 1);
  T* ey = comp(ebf, g, 2);
  T* ez = comp(ebf, g, 3);
  T* bx = comp(ebf, g, 4);
  T* by = comp(ebf, g, 5);
  T* bz = comp(ebf, g, nf);
  const int64_t sy = g.n1p, sz = (int64_t)g.n1p * g.n2p;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nxf + nyf + nzf;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (t < nxf) {
      const int64_t o = fofs(g, g.ix2, g.jy1 + (int)(t % nj), g.kz1 + (int)(t / nj));
      const T a = T(1.0 * g.dx_inv * dtl);
      ey[o + 1] = mur_p(bz[o], ey[o], a);
      if (nf > 3) ez[o + 1] = mur_m(by[o], ez[o], a);
    } else if (t < nxf + nyf) {
      const int64_t u = t - nxf;
      const int64_t o = fofs(g, g.ix1 + (int)(u % ni), g.jy2, g.kz1 + (int)(u / ni));
      const T a = T(1.0 * g.dy_inv * dtl);
      ex[o + sy] = mur_m(bz[o], ex[o], a);
      if (nf > 3) ez[o + sy] = mur_p(bx[o], ez[o], a);
    } else {
      const int64_t u = t - nxf - nyf;
      const int64_t o = fofs(g, g.ix1 + (int)(u % ni), g.jy1 + (int)(u / ni), g.kz2);
      const T a = T(1.0 * g.dz_inv * dtl);
      ex[o + sz] = mur_p(by[o], ex[o], a);
      ey[o + sz] = mur_m(bx[o], ey[o], a);
    }
  }
}
template <typename T>
__global__ void k_ef_bds(DevGrid g, T* ebf, double dtl) { b_ef_bds<T>(g, ebf, dtl); }
// ef_bds(0) of the runs with the wide x stencil
template <typename T>
__global__ void kb_ef_bds(const FieldRun<T>* runs) {
  const FieldRun<T>& r = runs[blockIdx.y];
  if (r.g.se[1] != 0.0) b_ef_bds<T>(r.g, r.ebf, 0.0);
}

// --------------------------------------------------------------- rote / rotb
// advance_lpf_fields (curr_and_fields_util.f90:176-260) with imbd = 0 is
//   ef_bds(0); rote(dt/2); bf_bds(dt); rotb(dt) + source; ef_bds(dt); rote(dt/2)
// Each boundary call only reads nodes of the interior layer next to its
// face and writes the ghost beyond it, so it joins the curl sweep whose
// thread owns that interior node:
//   * ef_bds(0) (S0): the thread of a node on a right face forms the ghost
//     E its forward difference reads from its own E and B before it updates
//     B.  ef_bds(dt) overwrites every such ghost in the same step, so S0's
//     values are not stored.  With the wide x stencil a node also reads the
//     ghost two nodes to its right, so S0 then runs alone first.
//   * bf_bds(dt) (S2): the thread of a node on a left face writes the ghost
//     B from its updated B and its E (the B sweep reads no B ghost).
//   * ef_bds(dt) (S4): the thread of a node on a right face writes the
//     ghost E from its updated E and B (the E sweep reads E at its own node
//     only).

// B -= dtf*curl E on the interior (grid_fields.f90:1405-1478).  FIRST: the
// first half-step, with S0 (when s0) and S2 (a = dt/dh).
template <typename T, bool FIRST>
__device__ inline void b_rote(DevGrid g, T* ebf, double dtf, bool s0) {
  const int nf = g.nfield;
  const int ni = g.ix2 - g.ix1 + 1, nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
  const int64_t n = (int64_t)ni * nj * nk;
  const T aphy = T(dtf * g.dy_inv), aphz = T(dtf * g.dz_inv);
  const T aph1 = T(double(dtf * g.dx_inv) * g.se[0]), aph2 = T(double(dtf * g.dx_inv) * g.se[1]);
  const T zero = T(1.0 * g.dx_inv * 0.0);  // a of ef_bds(0)
  const T ax = T(1.0 * g.dx_inv * g.dt), ay = T(1.0 * g.dy_inv * g.dt),
          az = T(1.0 * g.dz_inv * g.dt);
  const T* ex = comp(ebf, g, 1);
  const T* ey = comp(ebf, g, 2);
  const T* ez = comp(ebf, g, 3);
  T* bx = comp(ebf, g, 4);
  T* by = comp(ebf, g, 5);
  T* bz = comp(ebf, g, nf);
  const int64_t sy = g.n1p, sz = (int64_t)g.n1p * g.n2p;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = g.ix1 + (int)(t % ni);
    const int j = g.jy1 + (int)((t / ni) % nj);
    const int k = g.kz1 + (int)(t / ((int64_t)ni * nj));
    const int64_t o = fofs(g, i, j, k);
    const bool rx = FIRST && s0 && i == g.ix2, ry = FIRST && s0 && j == g.jy2,
               rz = FIRST && s0 && k == g.kz2;
    const T ex0 = ex[o], ey0 = ey[o], bz0 = bz[o];
    const T eyx = rx ? mur_p(bz0, ey0, zero) : ey[o + 1];
    const T exy = ry ? mur_m(bz0, ex0, zero) : ex[o + sy];
    const T bz1 = bz0 - aph1 * (eyx - ey0) + aphy * (exy - ex0) - aph2 * (ey[o + 2] - ey[o - 1]);
    bz[o] = bz1;
    if (nf >= 6) {
      const T ez0 = ez[o], bx0 = bx[o], by0 = by[o];
      const T ezy = ry ? mur_p(bx0, ez0, zero) : ez[o + sy];
      const T eyz = rz ? mur_m(bx0, ey0, zero) : ey[o + sz];
      const T ezx = rx ? mur_m(by0, ez0, zero) : ez[o + 1];
      const T exz = rz ? mur_p(by0, ex0, zero) : ex[o + sz];
      const T bx1 = bx0 - aphy * (ezy - ez0) + aphz * (eyz - ey0);
      const T by1 = by0 + aph1 * (ezx - ez0) - aphz * (exz - ex0) + aph2 * (ez[o + 2] - ez[o - 1]);
      bx[o] = bx1;
      by[o] = by1;
      if (FIRST) {  // S2
        if (i == g.ix1) by[o - 1] = mur_p(ez0, by1, ax);
        if (j == g.jy1) bx[o - sy] = mur_m(ez0, bx1, ay);
        if (k == g.kz1) {
          bx[o - sz] = mur_p(ey0, bx1, az);
          by[o - sz] = mur_m(ex0, by1, az);
        }
      }
    }
    if (FIRST) {
      if (i == g.ix1) bz[o - 1] = mur_m(ey0, bz1, ax);
      if (j == g.jy1) bz[o - sy] = mur_p(ex0, bz1, ay);
    }
  }
}
template <typename T, bool FIRST>
__global__ void k_rote(DevGrid g, T* ebf, double dtf, int s0) {
  b_rote<T, FIRST>(g, ebf, dtf, s0 != 0);
}
template <typename T, bool FIRST>
__global__ void kb_rote(const FieldRun<T>* runs) {
  const FieldRun<T>& r = runs[blockIdx.y];
  b_rote<T, FIRST>(r.g, r.ebf, 0.5 * r.g.dt, r.g.se[1] == 0.0);
}

// E += dt*curl B on the interior (grid_fields.f90:1480-1552) followed by the
// source term E(1:curr_ndim) -= ompe*jc (curr_and_fields_util.f90:233-236)
// and S4, ef_bds(dt) (a = dt/dh).
template <typename T>
__device__ inline void b_rotb_source(DevGrid g, T* ebf, const T* jc, double dtf) {
  const int nf = g.nfield;
  const int ni = g.ix2 - g.ix1 + 1, nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
  const int64_t n = (int64_t)ni * nj * nk;
  const T aphy = T(dtf * g.dy_inv), aphz = T(dtf * g.dz_inv);
  const T aph1 = T(double(dtf * g.dx_inv) * g.se[0]), aph2 = T(double(dtf * g.dx_inv) * g.se[1]);
  const T ax = T(1.0 * g.dx_inv * g.dt), ay = T(1.0 * g.dy_inv * g.dt),
          az = T(1.0 * g.dz_inv * g.dt);
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
  const int64_t sy = g.n1p, sz = (int64_t)g.n1p * g.n2p;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = g.ix1 + (int)(t % ni);
    const int j = g.jy1 + (int)((t / ni) % nj);
    const int k = g.kz1 + (int)(t / ((int64_t)ni * nj));
    const int64_t o = fofs(g, i, j, k);
    const T bz0 = bz[o];
    T vx = ex[o] + aphy * (bz0 - bz[o - sy]);
    T vy = ey[o] - aph1 * (bz0 - bz[o - 1]) - aph2 * (bz[o + 1] - bz[o - 2]);
    if (nf >= 6) {
      const T bx0 = bx[o], by0 = by[o];
      vx = vx - aphz * (by0 - by[o - sz]);
      vy = vy + aphz * (bx0 - bx[o - sz]);
      T vz = ez[o] + aph1 * (by0 - by[o - 1]) - aphy * (bx0 - bx[o - sy]) +
             aph2 * (by[o + 1] - by[o - 2]);
      if (g.curr_ndim >= 3) vz = vz - ompe * jz[o];
      ez[o] = vz;
      vx = vx - ompe * jx[o];
      vy = vy - ompe * jy[o];
      if (i == g.ix2) ez[o + 1] = mur_m(by0, vz, ax);
      if (j == g.jy2) ez[o + sy] = mur_p(bx0, vz, ay);
      if (k == g.kz2) {
        ex[o + sz] = mur_p(by0, vx, az);
        ey[o + sz] = mur_m(bx0, vy, az);
      }
    } else {
      vx = vx - ompe * jx[o];
      vy = vy - ompe * jy[o];
    }
    ex[o] = vx;
    ey[o] = vy;
    if (i == g.ix2) ey[o + 1] = mur_p(bz0, vy, ax);
    if (j == g.jy2) ex[o + sy] = mur_m(bz0, vx, ay);
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
  const bool s0 = g.se[1] == 0.0;  // compact x stencil: S0 joins the first sweep
  if (!s0) k_ef_bds<T><<<nblocks(nface), kFieldBlock, 0, s>>>(g, ebf, 0.0);
  k_rote<T, true><<<nblocks(nint), kFieldBlock, 0, s>>>(g, ebf, dth, s0 ? 1 : 0);
  k_rotb_source<T><<<nblocks(nint), kFieldBlock, 0, s>>>(g, ebf, jc, dt_lp);
  k_rote<T, false><<<nblocks(nint), kFieldBlock, 0, s>>>(g, ebf, dth, 0);
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
    d.any_wide = d.any_wide || g.se[1] != 0.0;
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
}

template <typename T>
void advance_fields_batch(const FieldRun<T>* runs, int nruns, const BatchDims& d, hipStream_t s) {
  if (d.any_wide) kb_ef_bds<T><<<bgrid(d.nface, nruns), kFieldBlock, 0, s>>>(runs);
  kb_rote<T, true><<<bgrid(d.nint, nruns), kFieldBlock, 0, s>>>(runs);
  kb_rotb_source<T><<<bgrid(d.nint, nruns), kFieldBlock, 0, s>>>(runs);
  kb_rote<T, false><<<bgrid(d.nint, nruns), kFieldBlock, 0, s>>>(runs);
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
  template void curr_collect<T>(const DevGrid&, T*, hipStream_t, bool);    \
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
