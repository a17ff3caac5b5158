// beam.hip -- the bunch fields at the start of a run on the device: the
// reference's init_beam_part_distrib.f90:538-602 on a uniform grid and one
// rank.  The bunch charge density (set_grid_charge of the bunch species)
// becomes its potential by the sine-transform Poisson solve of
// fft_3d_psolv / fft_2d_psolv (ft_mod = 2, sym = 1, psolve.f90:117-247), and
// initial_beam_fields (init_grid_fields.f90:32-131) makes E and B from it,
// which are added to ebf (Bx = 0).  Everything is computed in FP64 whatever
// the precision of the fields it is added to.
//
// Each one-dimensional sine transform (FFTW_RODFT00 of length n: y_k =
// 2 sum_j x_j sin(pi (j+1)(k+1)/(n+1))) along an axis of the n1 x n2 x n3
// box is the product of the n x n sine matrix with the box seen as an
// n x (n1 n2 n3 / n) matrix, computed by an LDS-tiled FP64 kernel: for the
// box sizes of a run (n up to a few hundred) this is a few GFLOP, once.
#include "lp_kernels.hpp"

namespace aladyn {

static constexpr int kBB = 256;

static unsigned bblocks(int64_t n) {
  int64_t b = (n + kBB - 1) / kBB;
  if (b < 1) b = 1;
  if (b > 65535LL * 8) b = 65535LL * 8;
  return (unsigned)b;
}

// S(k, j) = 2 sin(pi (j+1)(k+1)/(n+1)), row-major
__global__ void k_sine_matrix(double* S, int n) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < (int64_t)n * n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(t / n), j = (int)(t % n);
    S[t] = 2.0 * sin(acos(-1.0) * (double)(j + 1) * (double)(k + 1) / (double)(n + 1));
  }
}

// out(k, o) = sc sum_j S(k, j) in(j, o) along axis AX of the n1 x n2 x n3
// box (x fastest): element j of line o is at base(o) + j*st.  A workgroup
// computes a 64 x 64 tile (4 x 4 per thread) over j in chunks of 16 staged
// in LDS; the sum runs over j in order.
template <int AX>
__global__ __launch_bounds__(kBB) void k_axis_sine(const double* __restrict__ S,
                                                   const double* __restrict__ in,
                                                   double* __restrict__ out, int n1, int n2,
                                                   int n3, double sc) {
  const int n = AX == 0 ? n1 : (AX == 1 ? n2 : n3);
  const int64_t no = (int64_t)n1 * n2 * n3 / n;
  const int64_t st = AX == 0 ? 1 : (AX == 1 ? (int64_t)n1 : (int64_t)n1 * n2);
  const auto base = [&](int64_t o) -> int64_t {
    if (AX == 0) return o * n1;
    if (AX == 1) return (o % n1) + (int64_t)n1 * n2 * (o / n1);
    return o;
  };
  __shared__ double Ss[64][17];
  __shared__ double As[16][65];
  const int k0 = blockIdx.y * 64;
  const int64_t o0 = (int64_t)blockIdx.x * 64;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  double acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) acc[r][c] = 0.0;
  for (int j0 = 0; j0 < n; j0 += 16) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int e = threadIdx.x + kBB * u;
      const int kk = e / 16, jj = e % 16, k = k0 + kk, j = j0 + jj;
      Ss[kk][jj] = (k < n && j < n) ? S[(int64_t)k * n + j] : 0.0;
      const int jb = e / 64, ob = e % 64;
      const int64_t o = o0 + ob;
      As[jb][ob] = (o < no && j0 + jb < n) ? in[base(o) + (int64_t)(j0 + jb) * st] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int jj = 0; jj < 16; jj++) {
      double a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; r++) a[r] = Ss[ty * 4 + r][jj];
#pragma unroll
      for (int c = 0; c < 4; c++) b[c] = As[jj][tx * 4 + c];
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int c = 0; c < 4; c++) acc[r][c] += a[r] * b[c];
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int k = k0 + ty * 4 + r;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const int64_t o = o0 + tx * 4 + c;
      if (k < n && o < no) out[base(o) + (int64_t)k * st] = sc * acc[r][c];
    }
  }
}

// wb(1:n1, 1:n2, 1:n3) <-> rho(ix1:ix2, jy1:jy2, kz1:kz2)
__global__ void k_box_copy(DevGrid g, double* rho, double* wb, int to_box) {
  const int n1 = g.nx, n2 = g.ny, n3 = g.ndim == 3 ? g.nz : 1;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < (int64_t)n1 * n2 * n3;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % n1), j = (int)((t / n1) % n2), k = (int)(t / ((int64_t)n1 * n2));
    const int64_t o = fofs(g, g.ix1 + i, g.jy1 + j, g.kz1 + k);
    if (to_box)
      wb[t] = rho[o];
    else
      rho[o] = wb[t];
  }
}

// sine-transform wavenumber 2 sin(dh ak/2)/dh, ak = m pi/l_box
// (set_grid_param.f90:959-978)
__device__ inline double sine_k(int m, double dh, double l_box) {
  const double wk = acos(-1.0) / l_box;
  const double ak = wk * (double)m;
  return 2. * sin(0.5 * dh * ak) / dh;
}

// wb = ompe*wb, then the division of beam_potential (psolve.f90:66-115) in
// 3-D or beam_2d_potential (:38-64) in 2-D
__global__ void k_beam_scale(DevGrid g, double* wb, double gam2, double lx, double ly,
                             double lz) {
  const int n1 = g.nx, n2 = g.ny, n3 = g.ndim == 3 ? g.nz : 1;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < (int64_t)n1 * n2 * n3;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int ix = (int)(t % n1) + 1, iy = (int)((t / n1) % n2) + 1,
              iz = (int)(t / ((int64_t)n1 * n2)) + 1;
    double w = g.ompe * wb[t];
    const double sy = sine_k(iy, g.dy, ly);
    if (g.ndim == 3) {
      const double sz = sine_k(iz, g.dz, lz), sx = sine_k(ix, g.dx, lx);
      const double ak2p = sz * sz + sy * sy;
      if (ak2p > 0.0)
        w = w / (ak2p + sx * sx / gam2);
      else if (ix >= 2)
        w = w / (sx * sx / gam2);
    } else {
      const double ak2p = 0.0 * 0.0 + sy * sy;
      w = ak2p > 0.0 ? w / ak2p : 0.0;
    }
    wb[t] = w;
  }
}

// initial_beam_fields, first part: the ghost values of the potential it
// reads (init_grid_fields.f90:48-64, 96-103): y above jy2 by quadratic
// extrapolation, x beyond ix2 a copy of ix2, z above kz2 linear.  The cells
// written are disjoint and read only interior cells.
__global__ void k_beam_ghosts(DevGrid g, double* pot) {
  const int ni = g.ix2 - g.ix1 + 1, nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
  const int64_t ny_f = (int64_t)ni * nk, nx_f = (int64_t)nj * nk,
                nz_f = g.ndim == 3 ? (int64_t)ni * nj : 0;
  const int64_t sy = g.n1p, sz = (int64_t)g.n1p * g.n2p;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < ny_f + nx_f + nz_f;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (t < ny_f) {
      const int64_t o = fofs(g, g.ix1 + (int)(t % ni), g.jy2, g.kz1 + (int)(t / ni));
      pot[o + sy] = 3. * (pot[o] - pot[o - sy]) + pot[o - 2 * sy];
    } else if (t < ny_f + nx_f) {
      const int64_t u = t - ny_f;
      const int64_t o = fofs(g, g.ix2, g.jy1 + (int)(u % nj), g.kz1 + (int)(u / nj));
      pot[o + 1] = pot[o];
    } else {
      const int64_t u = t - ny_f - nx_f;
      const int64_t o = fofs(g, g.ix1 + (int)(u % ni), g.jy1 + (int)(u / ni), g.kz2);
      pot[o + sz] = 2. * pot[o] - pot[o - sz];
    }
  }
}

// initial_beam_fields, E (init_grid_fields.f90:71-84, 107-115): efb(1) =
// Ex on ix1..ix2, efb(2) = Ey on ix1..ix2+1, efb(3) = Ez (3-D) on
// ix1..ix2, over the interior j, k
__global__ void k_beam_e(DevGrid g, const double* pot, double* efb, double gam2) {
  const int ni = g.ix2 - g.ix1 + 2, nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
  const int64_t sy = g.n1p, sz = (int64_t)g.n1p * g.n2p;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < (int64_t)ni * nj * nk;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = g.ix1 + (int)(t % ni), j = g.jy1 + (int)((t / ni) % nj),
              k = g.kz1 + (int)(t / ((int64_t)ni * nj));
    const int64_t o = fofs(g, i, j, k);
    const double sdhy = 1.0 * g.dy_inv;  // unstretched metric
    efb[g.ncell + o] = -sdhy * (pot[o + sy] - pot[o]);
    if (i <= g.ix2) {
      efb[o] = -g.dx_inv * (pot[o + 1] - pot[o]) / gam2;
      if (g.ndim == 3) {
        const double sdhz = 1.0 * g.dz_inv;
        efb[2 * g.ncell + o] = -sdhz * (pot[o + sz] - pot[o]);
      }
    }
  }
}

// initial_beam_fields, B (init_grid_fields.f90:85-95, 104-128): Bz =
// bet <Ey>_x into efb(3) (2-D) or efb(6) (3-D), By = -bet <Ez>_x into
// efb(5) (Ez at ix2+1 is not set: 0)
__global__ void k_beam_b(DevGrid g, double* efb, double bet) {
  const int ni = g.ix2 - g.ix1 + 1, nj = g.jy2 - g.jy1 + 1, nk = g.kz2 - g.kz1 + 1;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < (int64_t)ni * nj * nk;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = g.ix1 + (int)(t % ni), j = g.jy1 + (int)((t / ni) % nj),
              k = g.kz1 + (int)(t / ((int64_t)ni * nj));
    const int64_t o = fofs(g, i, j, k);
    const double* ey = efb + g.ncell;
    const double bz = 0.5 * bet * (ey[o] + ey[o + 1]);
    if (g.ndim == 2) {
      efb[2 * g.ncell + o] = bz;
    } else {
      const double* ez = efb + 2 * g.ncell;
      efb[5 * g.ncell + o] = bz;
      efb[4 * g.ncell + o] = -bet * (0.5 * (ez[o] + ez[o + 1]));
    }
  }
}

// ebf(1:nfield) += efb(1:nfield), Bx left as it is (ebf_bunch(4) = 0)
template <typename R>
__global__ void k_beam_add(DevGrid g, const double* efb, R* ebf) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.ncell * g.nfield;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(t / g.ncell);
    if (g.ndim == 3 && c == 3) continue;
    ebf[t] = (R)((double)ebf[t] + efb[t]);
  }
}

template <typename R>
hipError_t beam_fields(const DevGrid& g, double* rho, double gam2, double bet0, R* ebf,
                       hipStream_t s) {
  const int n1 = g.nx, n2 = g.ny, n3 = g.ndim == 3 ? g.nz : 1;
  const int64_t nw = (int64_t)n1 * n2 * n3;
  double *wb = nullptr, *tmp = nullptr, *S1 = nullptr, *S2 = nullptr, *S3 = nullptr,
         *efb = nullptr;
  hipError_t e = hipSuccess;
  const auto ck = [&](hipError_t r) {
    if (e == hipSuccess && r != hipSuccess) e = r;
    return e == hipSuccess;
  };
  ck(hipMalloc(&wb, nw * sizeof(double)));
  ck(hipMalloc(&tmp, nw * sizeof(double)));
  ck(hipMalloc(&S1, (size_t)n1 * n1 * sizeof(double)));
  ck(hipMalloc(&S2, (size_t)n2 * n2 * sizeof(double)));
  ck(hipMalloc(&S3, (size_t)n3 * n3 * sizeof(double)));
  ck(hipMalloc(&efb, (size_t)g.ncell * 6 * sizeof(double)));
  if (e == hipSuccess) {
    k_sine_matrix<<<bblocks((int64_t)n1 * n1), kBB, 0, s>>>(S1, n1);
    k_sine_matrix<<<bblocks((int64_t)n2 * n2), kBB, 0, s>>>(S2, n2);
    k_sine_matrix<<<bblocks((int64_t)n3 * n3), kBB, 0, s>>>(S3, n3);
    k_box_copy<<<bblocks(nw), kBB, 0, s>>>(g, rho, wb, 1);
    double* cur = wb;
    double* oth = tmp;
    const auto axis = [&](int ax, double sc) {
      const int n = ax == 0 ? n1 : (ax == 1 ? n2 : n3);
      const dim3 grid((unsigned)((nw / n + 63) / 64), (unsigned)((n + 63) / 64));
      if (ax == 0) k_axis_sine<0><<<grid, kBB, 0, s>>>(S1, cur, oth, n1, n2, n3, sc);
      if (ax == 1) k_axis_sine<1><<<grid, kBB, 0, s>>>(S2, cur, oth, n1, n2, n3, sc);
      if (ax == 2) k_axis_sine<2><<<grid, kBB, 0, s>>>(S3, cur, oth, n1, n2, n3, sc);
      double* x = cur;
      cur = oth;
      oth = x;
    };
    // forward scales 1/logical_dimension: 2((n1+1)+1) along x (the x call
    // passes n1+1, modern_fft_lib.F90:689), 2(n+1) along y and z
    const double fx = 1.0 / (double)(2 * ((n1 + 1) + 1)), fy = 1.0 / (double)(2 * (n2 + 1)),
                 fz = 1.0 / (double)(2 * (n3 + 1));
    // box lengths x(nx) - x(1), y(ny+1) - y(1), z(nz+1) - z(1)
    // (set_grid_param.f90:76-78, 104-108, 128-130, 170, 187-189, 223)
    const double lx = g.dx * (double)(g.nx - 1) - g.dx * 0.0;
    const double ly = g.dy * (double)(g.ny - g.ny / 2) - g.dy * (double)(-(g.ny / 2));
    const double lz = g.dz * (double)(g.nz - g.nz / 2) - g.dz * (double)(-(g.nz / 2));
    if (g.ndim == 3) {
      axis(0, fx);  // pftw3d_sc(-1): x, then pftw2d_sc(-1): y, z
      if (n2 > 2) {
        axis(1, fy);
        if (n3 > 2) axis(2, fz);
      }
      k_beam_scale<<<bblocks(nw), kBB, 0, s>>>(g, cur, gam2, lx, ly, lz);
      if (n3 > 1) axis(2, 1.0);  // pftw3d_sc(+1): z, y, then x
      if (n2 > 1) axis(1, 1.0);
      axis(0, 1.0);
    } else {
      if (n2 > 2) axis(1, fy);  // pftw2d_sc(-1): y
      k_beam_scale<<<bblocks(nw), kBB, 0, s>>>(g, cur, gam2, lx, ly, lz);
      if (n2 > 1) axis(1, 1.0);
    }
    k_box_copy<<<bblocks(nw), kBB, 0, s>>>(g, rho, cur, 0);
    k_beam_ghosts<<<bblocks(g.ncell), kBB, 0, s>>>(g, rho);
    fill<double>(efb, g.ncell * 6, 0.0, s);
    k_beam_e<<<bblocks(g.ncell), kBB, 0, s>>>(g, rho, efb, gam2);
    k_beam_b<<<bblocks(g.ncell), kBB, 0, s>>>(g, efb, bet0);
    k_beam_add<R><<<bblocks(g.ncell * g.nfield), kBB, 0, s>>>(g, efb, ebf);
    ck(hipGetLastError());
    ck(hipStreamSynchronize(s));
  }
  for (double* p : {wb, tmp, S1, S2, S3, efb}) (void)hipFree(p);
  return e;
}

template hipError_t beam_fields<float>(const DevGrid&, double*, double, double, float*,
                                       hipStream_t);
template hipError_t beam_fields<double>(const DevGrid&, double*, double, double, double*,
                                        hipStream_t);

}  // namespace aladyn
