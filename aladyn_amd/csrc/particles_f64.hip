// particles_f64.hip -- the FP64 particle path on the reference's own layout
// part(ld, 2D+1) (struct_def.f90:26-28): one thread per particle runs the
// whole species-loop body of lpf2_evolve (pic_evolve_in_time.f90:77-89) in
// registers -- alternating-order gather, charge multiply, first-step
// back-shift, exact-gamma push, box test, Esirkepov deposit -- so the
// reference's per-stage scratch array ebfp never exists.  The deposit uses
// hardware FP64 atomics into jc; this path is the FP64 baseline, the FP32 tile
// path (particles_f32.hip) is the fast one.
#include "lp_kernels.hpp"

namespace aladyn {

static constexpr int kPBlock = 256;

struct GlobalAcc64 {
  double* jc;
  const DevGrid* g;
  __device__ void operator()(int c, int i, int j, int k, double v) const {
    unsafeAtomicAdd(jc + (int64_t)c * g->ncell + fofs(*g, i, j, k), v);
  }
};

__device__ inline void unpack_weight(double w, int& charge, float& wgh) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(w);
  charge = (int)(short)(b & 0xffffull);
  wgh = __uint_as_float((unsigned)(b >> 32));
}

template <int D>
__global__ __launch_bounds__(kPBlock) void k_particles_f64(DevGrid g, double* part, int64_t ld,
                                                           int64_t np, const double* ebf,
                                                           double* jc, double alp, int initial,
                                                           int* ndead) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= np) return;
  const double mn[3] = {g.xmn, g.ymn, g.zmn};
  const double mx[3] = {g.xmx, g.ymx, g.zmx};
  const double inv[3] = {g.dx_inv, g.dy_inv, g.dz_inv};
  double x[D], p[D];
#pragma unroll
  for (int d = 0; d < D; d++) {
    x[d] = part[d * ld + n];
    p[d] = part[(D + d) * ld + n];
  }
  int charge;
  float wgh;
  unpack_weight(part[2 * D * ld + n], charge, wgh);

  // ---- gather (set_part2d/3d_hcell_acc): quadratic at integer nodes, linear
  // at half-integer nodes; base index i = ix - 1 (grid_part_lib.f90:137-141)
  int io[D];
  double so[D], a[D][3], ah[D][2];
#pragma unroll
  for (int d = 0; d < D; d++) {
    ref_node(x[d], mn[d], inv[d], io[d], so[d]);
    qweights<double>(so[d], a[d]);
    lweights<double>(so[d], ah[d]);
  }
  double ef[6];
  if constexpr (D == 2) {
    const double* Ex = ebf;
    const double* Ey = ebf + g.ncell;
    const double* Bz = ebf + 2 * g.ncell;
    const int i = io[0] - 1, j = io[1] - 1;
    double ap1 = 0., ap2 = 0., ap3 = 0.;
#pragma unroll
    for (int j1 = 0; j1 <= 2; j1++) {
      const double dvol = a[1][j1];
#pragma unroll
      for (int i1 = 0; i1 <= 1; i1++) ap1 = ap1 + (ah[0][i1] * dvol) * Ex[fofs(g, i + i1, j + j1, 1)];
    }
#pragma unroll
    for (int j1 = 0; j1 <= 1; j1++) {
      const double dvol = ah[1][j1];
#pragma unroll
      for (int i1 = 0; i1 <= 2; i1++) ap2 = ap2 + (a[0][i1] * dvol) * Ey[fofs(g, i + i1, j + j1, 1)];
#pragma unroll
      for (int i1 = 0; i1 <= 1; i1++) ap3 = ap3 + (ah[0][i1] * dvol) * Bz[fofs(g, i + i1, j + j1, 1)];
    }
    ef[0] = ap1;
    ef[1] = ap2;
    ef[2] = ap3;
  } else {
    const int i = io[0] - 1, j = io[1] - 1, k = io[2] - 1;
    double ap[6] = {0., 0., 0., 0., 0., 0.};
    const double* F[6];
#pragma unroll
    for (int c = 0; c < 6; c++) F[c] = ebf + c * g.ncell;
#pragma unroll
    for (int k1 = 0; k1 <= 2; k1++) {
#pragma unroll
      for (int j1 = 0; j1 <= 2; j1++) {
        const double dvol = a[1][j1] * a[2][k1];
#pragma unroll
        for (int i1 = 0; i1 <= 1; i1++)
          ap[0] = ap[0] + ah[0][i1] * dvol * F[0][fofs(g, i + i1, j + j1, k + k1)];
      }
#pragma unroll
      for (int j1 = 0; j1 <= 1; j1++) {
        const double dvol = ah[1][j1] * a[2][k1];
#pragma unroll
        for (int i1 = 0; i1 <= 2; i1++)
          ap[1] = ap[1] + a[0][i1] * dvol * F[1][fofs(g, i + i1, j + j1, k + k1)];
#pragma unroll
        for (int i1 = 0; i1 <= 1; i1++)
          ap[5] = ap[5] + ah[0][i1] * dvol * F[5][fofs(g, i + i1, j + j1, k + k1)];
      }
    }
#pragma unroll
    for (int k1 = 0; k1 <= 1; k1++) {
#pragma unroll
      for (int j1 = 0; j1 <= 1; j1++) {
        const double dvol = ah[1][j1] * ah[2][k1];
#pragma unroll
        for (int i1 = 0; i1 <= 2; i1++)
          ap[3] = ap[3] + a[0][i1] * dvol * F[3][fofs(g, i + i1, j + j1, k + k1)];
      }
#pragma unroll
      for (int j1 = 0; j1 <= 2; j1++) {
        const double dvol = a[1][j1] * ah[2][k1];
#pragma unroll
        for (int i1 = 0; i1 <= 1; i1++)
          ap[4] = ap[4] + ah[0][i1] * dvol * F[4][fofs(g, i + i1, j + j1, k + k1)];
#pragma unroll
        for (int i1 = 0; i1 <= 2; i1++)
          ap[2] = ap[2] + a[0][i1] * dvol * F[2][fofs(g, i + i1, j + j1, k + k1)];
      }
    }
#pragma unroll
    for (int c = 0; c < 6; c++) ef[c] = ap[c];
  }
  // ---- field_charge_multiply (curr_and_fields_util.f90:58-72)
  const int nf = D == 2 ? 3 : 6;
#pragma unroll
  for (int c = 0; c < nf; c++) ef[c] = (double)charge * ef[c];

  if constexpr (D == 2) {
    // ---- init_lpf_momenta (boris_push.f90:45-54)
    if (initial) {
      const double e1 = -alp * ef[0], e2 = -alp * ef[1], e3 = -alp * ef[2];
      const double gam_inv = 1. / sqrt(1. + (p[0] * p[0] + p[1] * p[1]));
      const double v1 = p[0] * gam_inv, v2 = p[1] * gam_inv;
      p[0] = p[0] + e1 + v2 * e3;
      p[1] = p[1] + e2 - v1 * e3;
    }
    // ---- lpf_momenta_and_positions (boris_push.f90:92-119)
    const double e1 = alp * ef[0], e2 = alp * ef[1], e3 = alp * ef[2];
    const double v1 = p[0] + e1, v2 = p[1] + e2, v3 = e3;
    double gam02 = 1. + (v1 * v1 + v2 * v2);
    const double b2 = v3 * v3;
    gam02 = gam02 - b2;
    const double gam2 = 0.5 * (gam02 + sqrt(gam02 * gam02 + 4. * b2));
    const double gam = sqrt(gam2);
    double h1 = gam2 * v1 + gam * v2 * v3;
    double h2 = gam2 * v2 - gam * v1 * v3;
    h1 = h1 / (gam2 + b2);
    h2 = h2 / (gam2 + b2);
    p[0] = 2. * h1 - p[0];
    p[1] = 2. * h2 - p[1];
    const double dtg = g.dt / sqrt(1. + (p[0] * p[0] + p[1] * p[1]));
    x[0] = x[0] + dtg * p[0];
    x[1] = x[1] + dtg * p[1];
  } else {
    if (initial) {
      // boris_push.f90:56-70
      double e[6];
#pragma unroll
      for (int c = 0; c < 6; c++) e[c] = -alp * ef[c];
      const double gam_inv = 1. / sqrt(1. + (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]));
      const double v1 = gam_inv * p[0], v2 = gam_inv * p[1], v3 = gam_inv * p[2];
      p[0] = p[0] + e[0] + v2 * e[5] - v3 * e[4];
      p[1] = p[1] + e[1] + v3 * e[3] - v1 * e[5];
      p[2] = p[2] + e[2] + v1 * e[4] - v2 * e[3];
    }
    // boris_push.f90:124-150
    double e[6];
#pragma unroll
    for (int c = 0; c < 6; c++) e[c] = alp * ef[c];
    const double vp[3] = {p[0] + e[0], p[1] + e[1], p[2] + e[2]};
    const double bb[3] = {e[3], e[4], e[5]};
    double gam02 = 1. + (vp[0] * vp[0] + vp[1] * vp[1] + vp[2] * vp[2]);
    const double b2 = bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2];
    const double bv = bb[0] * vp[0] + bb[1] * vp[1] + bb[2] * vp[2];
    gam02 = gam02 - b2;
    const double gam2 = 0.5 * (gam02 + sqrt(gam02 * gam02 + 4. * (b2 + bv * bv)));
    const double gam = sqrt(gam2);
    double vph[3];
#pragma unroll
    for (int c = 0; c < 3; c++) vph[c] = gam2 * vp[c] + bb[c] * bv;
    vph[0] = vph[0] + gam * (vp[1] * bb[2] - vp[2] * bb[1]);
    vph[1] = vph[1] + gam * (vp[2] * bb[0] - vp[0] * bb[2]);
    vph[2] = vph[2] + gam * (vp[0] * bb[1] - vp[1] * bb[0]);
#pragma unroll
    for (int c = 0; c < 3; c++) p[c] = 2. * (vph[c] / (b2 + gam2)) - p[c];
    const double dtg = g.dt / sqrt(1. + (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]));
#pragma unroll
    for (int c = 0; c < 3; c++) x[c] = x[c] + dtg * p[c];
  }
  double xold[D];
#pragma unroll
  for (int d = 0; d < D; d++) {
    xold[d] = part[d * ld + n];
    part[d * ld + n] = x[d];
    part[(D + d) * ld + n] = p[d];
  }
  // ---- cell_part_dist(.false.): open boundaries drop x <= gmin or x > gmax
  // (mpi_part_interface.f90:494-498); written so that NaN also drops
  bool keep = true;
#pragma unroll
  for (int d = 0; d < D; d++) keep = keep && (x[d] > mn[d]) && (x[d] <= mx[d]);
  if (!keep) {
    atomicAdd(ndead, 1);
    return;
  }
  // ---- curr_accumulate -> esirkepov_*_curr with wght = real(charge*wgh, sp)
  int in[D];
  double sn[D];
#pragma unroll
  for (int d = 0; d < D; d++) {
    ref_node(xold[d], mn[d], inv[d], io[d], so[d]);
    ref_node(x[d], mn[d], inv[d], in[d], sn[d]);
  }
  const double wq = (double)((float)charge * wgh);
  GlobalAcc64 acc{jc, &g};
  esirkepov<D, double>(io, so, in, sn, wq, acc);
}

void particles_f64(const DevGrid& g, double* part, int64_t ld, int64_t np, const double* ebf,
                   double* jc, double alp, int initial, int* ndead, hipStream_t s) {
  if (np <= 0) return;
  const unsigned nb = (unsigned)((np + kPBlock - 1) / kPBlock);
  if (g.ndim == 2)
    k_particles_f64<2><<<nb, kPBlock, 0, s>>>(g, part, ld, np, ebf, jc, alp, initial, ndead);
  else
    k_particles_f64<3><<<nb, kPBlock, 0, s>>>(g, part, ld, np, ebf, jc, alp, initial, ndead);
}

// ------------------------------------------------------------- compaction
// Stable: chunk c of kChunk particles keeps its survivors in order at
// offset(c) = sum of survivors of chunks < c.
static constexpr int kCBlock = 256, kCPer = 8, kChunk = kCBlock * kCPer;

int64_t nblk_compact(int64_t np) { return (np + kChunk - 1) / kChunk; }

__device__ inline bool in_box(const DevGrid& g, const double* part, int64_t ld, int64_t n, int D) {
  const double mn[3] = {g.xmn, g.ymn, g.zmn};
  const double mx[3] = {g.xmx, g.ymx, g.zmx};
  bool keep = true;
  for (int d = 0; d < D; d++) {
    const double x = part[d * ld + n];
    keep = keep && (x > mn[d]) && (x <= mx[d]);
  }
  return keep;
}

__device__ inline int block_excl_scan(int v, int* sh, int& total) {
  // sh: kCBlock ints of LDS; Hillis-Steele inclusive scan
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kCBlock; off <<= 1) {
    int add = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  total = sh[kCBlock - 1];
  int excl = sh[t] - v;
  __syncthreads();
  return excl;
}

__global__ __launch_bounds__(kCBlock) void k_count_keep(DevGrid g, const double* part, int64_t ld,
                                                        int64_t np, int* blk) {
  __shared__ int sh[kCBlock];
  const int D = g.ndim;
  int cnt = 0;
  const int64_t base = blockIdx.x * (int64_t)kChunk + threadIdx.x * (int64_t)kCPer;
  for (int q = 0; q < kCPer; q++) {
    const int64_t n = base + q;
    if (n < np && in_box(g, part, ld, n, D)) cnt++;
  }
  int total;
  block_excl_scan(cnt, sh, total);
  if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// exclusive scan of the chunk counts by one workgroup; total -> *d_count
__global__ __launch_bounds__(kCBlock) void k_scan_blocks(int* blk, int64_t nblk, int64_t* d_count) {
  __shared__ int sh[kCBlock];
  int64_t carry = 0;
  for (int64_t b0 = 0; b0 < nblk; b0 += kCBlock) {
    const int64_t b = b0 + threadIdx.x;
    int v = b < nblk ? blk[b] : 0;
    int total;
    int ex = block_excl_scan(v, sh, total);
    if (b < nblk) blk[b] = (int)(carry + ex);
    carry += total;
  }
  if (threadIdx.x == 0) *d_count = carry;
}

__global__ __launch_bounds__(kCBlock) void k_compact(DevGrid g, const double* part, double* out,
                                                     int64_t ld, int64_t np, const int* blk) {
  __shared__ int sh[kCBlock];
  const int D = g.ndim, ncol = 2 * D + 1;
  const int64_t base = blockIdx.x * (int64_t)kChunk + threadIdx.x * (int64_t)kCPer;
  unsigned keepmask = 0;
  int cnt = 0;
  for (int q = 0; q < kCPer; q++) {
    const int64_t n = base + q;
    if (n < np && in_box(g, part, ld, n, D)) {
      keepmask |= 1u << q;
      cnt++;
    }
  }
  int total;
  int dst = blk[blockIdx.x] + block_excl_scan(cnt, sh, total);
  for (int q = 0; q < kCPer; q++) {
    if (!(keepmask & (1u << q))) continue;
    const int64_t n = base + q;
    for (int c = 0; c < ncol; c++) out[c * ld + dst] = part[c * ld + n];
    dst++;
  }
}

void compact_f64(const DevGrid& g, const double* part, double* out, int64_t ld, int64_t np,
                 int* blk, int64_t* d_count, hipStream_t s) {
  const int64_t nb = nblk_compact(np);
  if (nb == 0) return;
  k_count_keep<<<(unsigned)nb, kCBlock, 0, s>>>(g, part, ld, np, blk);
  k_scan_blocks<<<1, kCBlock, 0, s>>>(blk, nb, d_count);
  k_compact<<<(unsigned)nb, kCBlock, 0, s>>>(g, part, out, ld, np, blk);
}

// ------------------------------------------------------------ diagnostics
template <int D>
__global__ void k_grid_charge_f64(DevGrid g, const double* part, int64_t ld, int64_t np,
                                  double* rho) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= np) return;
  const double mn[3] = {g.xmn, g.ymn, g.zmn};
  const double inv[3] = {g.dx_inv, g.dy_inv, g.dz_inv};
  int charge;
  float wgh;
  unpack_weight(part[2 * D * ld + n], charge, wgh);
  const double wq = (double)((float)charge * wgh);
  int ix[3] = {1, 1, 1};
  double a[3][3] = {{0, 0, 0}, {0, 0, 0}, {1, 0, 0}};
#pragma unroll
  for (int d = 0; d < D; d++) {
    double s;
    ref_node(part[d * ld + n], mn[d], inv[d], ix[d], s);
    qweights<double>(s, a[d]);
    ix[d] -= 1;
  }
  const int nk = D == 3 ? 3 : 1;
  for (int k1 = 0; k1 < nk; k1++)
    for (int j1 = 0; j1 < 3; j1++)
      for (int i1 = 0; i1 < 3; i1++)
        unsafeAtomicAdd(rho + fofs(g, ix[0] + i1, ix[1] + j1, D == 3 ? ix[2] + k1 : 1),
                        (wq * a[0][i1]) * (a[1][j1] * a[2][k1]));
}

void grid_charge_f64(const DevGrid& g, const double* part, int64_t ld, int64_t np, double* rho,
                     hipStream_t s) {
  if (np <= 0) return;
  const unsigned nb = (unsigned)((np + kPBlock - 1) / kPBlock);
  if (g.ndim == 2)
    k_grid_charge_f64<2><<<nb, kPBlock, 0, s>>>(g, part, ld, np, rho);
  else
    k_grid_charge_f64<3><<<nb, kPBlock, 0, s>>>(g, part, ld, np, rho);
}

__device__ inline void atomic_max_nonneg(double* addr, double v) {
  // non-negative doubles order like their bit patterns as signed integers
  atomicMax((unsigned long long*)addr, (unsigned long long)__double_as_longlong(v));
}

template <int D>
__global__ __launch_bounds__(kPBlock) void k_energy_f64(DevGrid g, const double* part, int64_t ld,
                                                        int64_t np, double* out) {
  __shared__ double sh[8][kPBlock / 64];
  double ek[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < np;
       n += (int64_t)gridDim.x * blockDim.x) {
    int charge;
    float wgh;
    unpack_weight(part[2 * D * ld + n], charge, wgh);
    double p[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; d++) p[d] = part[(D + d) * ld + n];
    const double gam1 = sqrt(1. + p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) - 1.;
    if constexpr (D == 2) {
      ek[0] += (double)wgh * p[0];
      ek[1] += (double)wgh * p[1];
    } else {
      ek[0] += p[0];
      ek[1] += p[1];
      ek[2] += p[2];
      ek[3] += (double)wgh * (part[ld + n] * p[2] - part[2 * ld + n] * p[1]);
    }
    ek[5] += (double)((float)charge * wgh);
    ek[6] += (double)wgh * gam1;
    ek[7] = fmax(ek[7], gam1);
  }
  for (int c = 0; c < 8; c++) {
    double v = ek[c];
    for (int off = 32; off > 0; off >>= 1) {
      double o = __shfl_down(v, off, 64);
      v = c == 7 ? fmax(v, o) : v + o;
    }
    if ((threadIdx.x & 63) == 0) sh[c][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int c = threadIdx.x;
    double v = sh[c][0];
    for (int w = 1; w < kPBlock / 64; w++) v = c == 7 ? fmax(v, sh[c][w]) : v + sh[c][w];
    if (c == 7)
      atomic_max_nonneg(out + 7, v);
    else
      unsafeAtomicAdd(out + c, v);
  }
}

void energy_momenta_f64(const DevGrid& g, const double* part, int64_t ld, int64_t np,
                        double* out, hipStream_t s) {
  if (np <= 0) return;
  int64_t nb = (np + kPBlock - 1) / kPBlock;
  if (nb > 4096) nb = 4096;
  if (g.ndim == 2)
    k_energy_f64<2><<<(unsigned)nb, kPBlock, 0, s>>>(g, part, ld, np, out);
  else
    k_energy_f64<3><<<(unsigned)nb, kPBlock, 0, s>>>(g, part, ld, np, out);
}

}  // namespace aladyn
