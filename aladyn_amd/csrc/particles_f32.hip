// particles_f32.hip -- the FP32 tile path (see tile_store.hpp for the store).
//
// One workgroup per tile:
//   1. stages the tile's field patch (nodes T0-1 .. T0+T per axis, all
//      nfield components) from HBM into LDS and zeroes an LDS current patch
//      (nodes T0-2 .. T0+T+1: the 5-slot Esirkepov window of any particle
//      whose nearest node lies in the tile, grid_part_connect.f90:1008);
//   2. streams the tile's particles: gather from LDS with the alternating-
//      order shapes, charge multiply, [first-step back-shift], exact-gamma
//      push, box test, Esirkepov deposit into the LDS patch (LDS atomics,
//      no global atomics in this loop), and writes each survivor into the
//      segment of its new tile in the other buffer;
//   3. writes its current patch to a per-tile partial-sum buffer.
// A fold kernel then sums, for every grid node, the (at most 2 per axis)
// tile patches covering it with compensated (Kahan) summation and writes jc.
// Overflowed particles (destination segment full) take a global-memory
// kernel the next step; it deposits with global atomics after the fold.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "lp_kernels.hpp"
#include "tile_store.hpp"

namespace aladyn {

// ------------------------------------------------------------ geometry
static int bits_for(int n) {  // bits to hold values 0..n
  int b = 0;
  while ((1 << b) <= n) b++;
  return b;
}

const char* tile_geom_error(int code) {
  switch (code) {
    case 1: return "tile extent must be >= 4 nodes per tiled axis";
    case 2: return "grid too large for the 31-bit packed cell index";
    case 3: return "tile patches exceed 64 KiB of LDS; use smaller tiles";
    case 4: return "block must be a multiple of 64 in [64, 1024]";
    default: return "invalid tile geometry";
  }
}

int tile_geom_init(const DevGrid& g, const aladyn_config& cfg, TileGeom* tg) {
  std::memset(tg, 0, sizeof *tg);
  const int D = g.ndim;
  const int def[3] = {D == 3 ? 8 : 16, D == 3 ? 8 : 16, D == 3 ? 8 : 1};
  for (int d = 0; d < 3; d++) tg->T[d] = cfg.tile[d] > 0 ? cfg.tile[d] : def[d];
  if (D == 2) tg->T[2] = 1;
  for (int d = 0; d < D; d++)
    if (tg->T[d] < 4) return 1;
  // particle nearest nodes span Fortran indices 3 .. n+3 (x_loc in (0, n])
  const int n[3] = {g.nx, g.ny, g.nz};
  for (int d = 0; d < 3; d++) tg->nt[d] = d < D ? (n[d] + 1 + tg->T[d] - 1) / tg->T[d] : 1;
  tg->ntiles = tg->nt[0] * tg->nt[1] * tg->nt[2];
  for (int d = 0; d < 3; d++) {
    tg->G[d] = d < D ? tg->T[d] + 2 : 1;
    tg->J[d] = d < D ? tg->T[d] + 4 : 1;
  }
  tg->gvol = tg->G[0] * tg->G[1] * tg->G[2];
  tg->pvol = tg->J[0] * tg->J[1] * tg->J[2];
  tg->bx = bits_for(g.n1p);
  tg->by = bits_for(g.n2p);
  const int bz = D == 3 ? bits_for(g.n3p) : 0;
  if (tg->bx + tg->by + bz > 31) return 2;
  tg->block = cfg.block > 0 ? cfg.block : 256;
  if (tg->block % 64 || tg->block < 64 || tg->block > 1024) return 4;
  tg->sort_every = cfg.sort_every > 0 ? cfg.sort_every : 16;
  tg->lds_bytes = (g.nfield * tg->gvol + g.curr_ndim * tg->pvol) * (int)sizeof(float);
  if (tg->lds_bytes > 64 * 1024) return 3;
  return 0;
}

// ------------------------------------------------------- record helpers
struct Rec {
  int N[3];
  float s[3];
  float p[3];
  float w;
  int id;
};

template <int D>
__device__ inline void load_rec(const RecPtr& r, int64_t n, Rec& q, int bx, int by) {
  const int cell = r.cell[n];
  q.N[0] = cell & ((1 << bx) - 1);
  q.N[1] = (cell >> bx) & ((1 << by) - 1);
  q.N[2] = D == 3 ? (cell >> (bx + by)) : 1;
#pragma unroll
  for (int d = 0; d < D; d++) {
    q.s[d] = r.s[d][n];
    q.p[d] = r.p[d][n];
  }
  q.w = r.w[n];
  q.id = r.id ? r.id[n] : 0;
}

template <int D>
__device__ inline void store_rec(const RecPtr& r, int64_t n, const Rec& q, int bx, int by) {
  r.cell[n] = q.N[0] | (q.N[1] << bx) | (D == 3 ? (q.N[2] << (bx + by)) : 0);
#pragma unroll
  for (int d = 0; d < D; d++) {
    r.s[d][n] = q.s[d];
    r.p[d][n] = q.p[d];
  }
  r.w[n] = q.w;
  if (r.id) r.id[n] = q.id;
}

// Fields at the particle with the reference's alternating-order shapes
// (set_part2d_hcell_acc / set_part3d_hcell_acc, grid_part_connect.f90:89-285).
// F(c, i, j, k) reads component c at local node (i, j, k) where (i, j, k)
// are already relative to the accessor's origin; (bi, bj, bk) is the local
// index of node N - 1.
template <int D, class FA>
__device__ inline void gather(int bi, int bj, int bk, const float s[3], const FA& F, float ef[6]) {
  float a[3][3], ah[3][2];
#pragma unroll
  for (int d = 0; d < D; d++) {
    qweights<float>(s[d], a[d]);
    lweights<float>(s[d], ah[d]);
  }
  if constexpr (D == 2) {
    float e0 = 0.f, e1 = 0.f, e2 = 0.f;
#pragma unroll
    for (int j1 = 0; j1 <= 2; j1++)
#pragma unroll
      for (int i1 = 0; i1 <= 1; i1++) e0 += (ah[0][i1] * a[1][j1]) * F(0, bi + i1, bj + j1, bk);
#pragma unroll
    for (int j1 = 0; j1 <= 1; j1++) {
#pragma unroll
      for (int i1 = 0; i1 <= 2; i1++) e1 += (a[0][i1] * ah[1][j1]) * F(1, bi + i1, bj + j1, bk);
#pragma unroll
      for (int i1 = 0; i1 <= 1; i1++) e2 += (ah[0][i1] * ah[1][j1]) * F(2, bi + i1, bj + j1, bk);
    }
    ef[0] = e0;
    ef[1] = e1;
    ef[2] = e2;
  } else {
    float ap[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k1 = 0; k1 <= 2; k1++) {
#pragma unroll
      for (int j1 = 0; j1 <= 2; j1++) {
        const float dv = a[1][j1] * a[2][k1];
#pragma unroll
        for (int i1 = 0; i1 <= 1; i1++) ap[0] += ah[0][i1] * dv * F(0, bi + i1, bj + j1, bk + k1);
      }
#pragma unroll
      for (int j1 = 0; j1 <= 1; j1++) {
        const float dv = ah[1][j1] * a[2][k1];
#pragma unroll
        for (int i1 = 0; i1 <= 2; i1++) ap[1] += a[0][i1] * dv * F(1, bi + i1, bj + j1, bk + k1);
#pragma unroll
        for (int i1 = 0; i1 <= 1; i1++) ap[5] += ah[0][i1] * dv * F(5, bi + i1, bj + j1, bk + k1);
      }
    }
#pragma unroll
    for (int k1 = 0; k1 <= 1; k1++) {
#pragma unroll
      for (int j1 = 0; j1 <= 1; j1++) {
        const float dv = ah[1][j1] * ah[2][k1];
#pragma unroll
        for (int i1 = 0; i1 <= 2; i1++) ap[3] += a[0][i1] * dv * F(3, bi + i1, bj + j1, bk + k1);
      }
#pragma unroll
      for (int j1 = 0; j1 <= 2; j1++) {
        const float dv = a[1][j1] * ah[2][k1];
#pragma unroll
        for (int i1 = 0; i1 <= 1; i1++) ap[4] += ah[0][i1] * dv * F(4, bi + i1, bj + j1, bk + k1);
#pragma unroll
        for (int i1 = 0; i1 <= 2; i1++) ap[2] += a[0][i1] * dv * F(2, bi + i1, bj + j1, bk + k1);
      }
    }
#pragma unroll
    for (int c = 0; c < 6; c++) ef[c] = ap[c];
  }
}

// Charge multiply + [init_lpf_momenta] + lpf_momenta_and_positions in FP32
// (curr_and_fields_util.f90:58-72, boris_push.f90:34-168).  Returns
// dt/gamma^{n+1/2}.
template <int D>
__device__ inline float push(float p[3], float ef[6], float qch, float alp, float dt, int initial) {
  const int nf = D == 2 ? 3 : 6;
#pragma unroll
  for (int c = 0; c < nf; c++) ef[c] = qch * ef[c];
  if constexpr (D == 2) {
    if (initial) {
      const float e1 = -alp * ef[0], e2 = -alp * ef[1], e3 = -alp * ef[2];
      const float gi = 1.f / sqrtf(1.f + (p[0] * p[0] + p[1] * p[1]));
      const float v1 = p[0] * gi, v2 = p[1] * gi;
      p[0] = p[0] + e1 + v2 * e3;
      p[1] = p[1] + e2 - v1 * e3;
    }
    const float v1 = p[0] + alp * ef[0], v2 = p[1] + alp * ef[1], v3 = alp * ef[2];
    const float b2 = v3 * v3;
    const float gam02 = (1.f + (v1 * v1 + v2 * v2)) - b2;
    const float gam2 = 0.5f * (gam02 + sqrtf(gam02 * gam02 + 4.f * b2));
    const float gam = sqrtf(gam2);
    const float r = 1.f / (gam2 + b2);
    const float h1 = (gam2 * v1 + gam * v2 * v3) * r;
    const float h2 = (gam2 * v2 - gam * v1 * v3) * r;
    p[0] = 2.f * h1 - p[0];
    p[1] = 2.f * h2 - p[1];
    return dt / sqrtf(1.f + (p[0] * p[0] + p[1] * p[1]));
  } else {
    if (initial) {
      float e[6];
#pragma unroll
      for (int c = 0; c < 6; c++) e[c] = -alp * ef[c];
      const float gi = 1.f / sqrtf(1.f + (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]));
      const float v1 = gi * p[0], v2 = gi * p[1], v3 = gi * p[2];
      p[0] = p[0] + e[0] + v2 * e[5] - v3 * e[4];
      p[1] = p[1] + e[1] + v3 * e[3] - v1 * e[5];
      p[2] = p[2] + e[2] + v1 * e[4] - v2 * e[3];
    }
    const float vp[3] = {p[0] + alp * ef[0], p[1] + alp * ef[1], p[2] + alp * ef[2]};
    const float bb[3] = {alp * ef[3], alp * ef[4], alp * ef[5]};
    const float b2 = bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2];
    const float bv = bb[0] * vp[0] + bb[1] * vp[1] + bb[2] * vp[2];
    const float gam02 = (1.f + (vp[0] * vp[0] + vp[1] * vp[1] + vp[2] * vp[2])) - b2;
    const float gam2 = 0.5f * (gam02 + sqrtf(gam02 * gam02 + 4.f * (b2 + bv * bv)));
    const float gam = sqrtf(gam2);
    const float r = 1.f / (b2 + gam2);
    float vph[3];
    vph[0] = gam2 * vp[0] + bb[0] * bv + gam * (vp[1] * bb[2] - vp[2] * bb[1]);
    vph[1] = gam2 * vp[1] + bb[1] * bv + gam * (vp[2] * bb[0] - vp[0] * bb[2]);
    vph[2] = gam2 * vp[2] + bb[2] * bv + gam * (vp[0] * bb[1] - vp[1] * bb[0]);
#pragma unroll
    for (int c = 0; c < 3; c++) p[c] = 2.f * (vph[c] * r) - p[c];
    return dt / sqrtf(1.f + (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]));
  }
}

// x^{n+1} = x^n + dt v: offset update with an exact integer carry; the box
// test is (x - gmin) > 0 and (x - gmin) <= n*dx, i.e. node 3 + (0, n] in
// local units (mpi_part_interface.f90:494-498), evaluated without rounding.
template <int D>
__device__ inline bool move_and_test(Rec& q, float dtg, const float inv[3], const int nn[3]) {
  bool keep = true;
#pragma unroll
  for (int d = 0; d < D; d++) {
    float s = q.s[d] + (dtg * inv[d]) * q.p[d];
    int N = q.N[d];
    if (s >= 0.5f) {
      N += 1;
      s -= 1.f;
    } else if (s < -0.5f) {
      N -= 1;
      s += 1.f;
    }
    q.N[d] = N;
    q.s[d] = s;
    const bool fin = fabsf(s) <= 0.5f;  // false for NaN
    const bool lo = (N > 3) || (N == 3 && s > 0.f);
    const bool hi = (N < nn[d] + 3) || (N == nn[d] + 3 && s <= 0.f);
    keep = keep && fin && lo && hi;
  }
  return keep;
}

struct LdsAcc {
  float* j;
  int ox, oy, oz, jx, jxy, pvol;
  __device__ void operator()(int c, int i, int jj, int k, float v) const {
    atomicAdd(j + c * pvol + (k - oz) * jxy + (jj - oy) * jx + (i - ox), v);
  }
};

struct GlobalAcc32 {
  float* jc;
  const DevGrid* g;
  __device__ void operator()(int c, int i, int j, int k, float v) const {
    unsafeAtomicAdd(jc + (int64_t)c * g->ncell + fofs(*g, i, j, k), v);
  }
};

struct LdsField {
  const float* f;
  int gx, gxy, gvol;
  __device__ float operator()(int c, int i, int j, int k) const {
    return f[c * gvol + k * gxy + j * gx + i];
  }
};

struct GlobalField {
  const float* f;
  const DevGrid* g;
  __device__ float operator()(int c, int i, int j, int k) const {
    return f[(int64_t)c * g->ncell + fofs(*g, i, j, k)];
  }
};

struct StepConst {
  float alp, qch, dt;
  float inv[3];
  int nn[3];
  int initial;
  int64_t ovf_base;
};

// --------------------------------------------------------- tile kernel
template <int D>
__global__ __launch_bounds__(256) void k_tile_particles(
    TileGeom tg, DevGrid g, StepConst sc, const float* __restrict__ ebf, RecPtr A, RecPtr B,
    const int* __restrict__ seg_start, const int* __restrict__ seg_cap,
    const int* __restrict__ cnt_cur, int* __restrict__ cnt_next, int* __restrict__ ovf_next,
    float* __restrict__ patch) {
  extern __shared__ float lds[];
  __shared__ int hist[27];
  __shared__ int hbase[27];
  constexpr int nf = D == 3 ? 6 : 3;
  float* Fl = lds;
  float* Jl = lds + nf * tg.gvol;
  const int t = blockIdx.x;
  const int tx = t % tg.nt[0], ty = (t / tg.nt[0]) % tg.nt[1], tz = t / (tg.nt[0] * tg.nt[1]);
  const int T0[3] = {3 + tx * tg.T[0], 3 + ty * tg.T[1], D == 3 ? 3 + tz * tg.T[2] : 1};
  const int G0[3] = {T0[0] - 1, T0[1] - 1, D == 3 ? T0[2] - 1 : 1};
  const int J0[3] = {T0[0] - 2, T0[1] - 2, D == 3 ? T0[2] - 2 : 1};
  const int gxy = tg.G[0] * tg.G[1];
  for (int q = threadIdx.x; q < nf * tg.gvol; q += blockDim.x) {
    const int c = q / tg.gvol, r = q - c * tg.gvol;
    const int li = r % tg.G[0], lj = (r / tg.G[0]) % tg.G[1], lk = r / gxy;
    const int i = G0[0] + li, j = G0[1] + lj, k = G0[2] + lk;
    float v = 0.f;
    if (i <= g.n1p && j <= g.n2p && k <= g.n3p) v = ebf[(int64_t)c * g.ncell + fofs(g, i, j, k)];
    Fl[q] = v;
  }
  for (int q = threadIdx.x; q < D * tg.pvol; q += blockDim.x) Jl[q] = 0.f;
  __syncthreads();
  const LdsField F{Fl, tg.G[0], gxy, tg.gvol};
  const LdsAcc acc{Jl, J0[0], J0[1], J0[2], tg.J[0], tg.J[0] * tg.J[1], tg.pvol};
  const int cnt = min(cnt_cur[t], seg_cap[t]);
  const int64_t base = seg_start[t];
  for (int it = 0; it < cnt; it += blockDim.x) {
    if (threadIdx.x < 27) hist[threadIdx.x] = 0;
    __syncthreads();
    const int n = it + threadIdx.x;
    bool keep = false;
    int code = 0, lrank = 0;
    Rec q;
    if (n < cnt) {
      load_rec<D>(A, base + n, q, tg.bx, tg.by);
      float ef[6];
      gather<D>(q.N[0] - 1 - G0[0], q.N[1] - 1 - G0[1], D == 3 ? q.N[2] - 1 - G0[2] : 0, q.s, F, ef);
      const int io[3] = {q.N[0], q.N[1], q.N[2]};
      const float so[3] = {q.s[0], q.s[1], q.s[2]};
      const float dtg = push<D>(q.p, ef, sc.qch, sc.alp, sc.dt, sc.initial);
      keep = move_and_test<D>(q, dtg, sc.inv, sc.nn);
      if (keep) {
        esirkepov<D, float>(io, so, q.N, q.s, sc.qch * q.w, acc);
        int dc[3] = {0, 0, 0};
        dc[0] = (q.N[0] - 3) / tg.T[0] - tx;
        dc[1] = (q.N[1] - 3) / tg.T[1] - ty;
        if (D == 3) dc[2] = (q.N[2] - 3) / tg.T[2] - tz;
        code = (dc[0] + 1) + 3 * (dc[1] + 1) + 9 * (dc[2] + 1);
        lrank = atomicAdd(&hist[code], 1);
      }
    }
    __syncthreads();
    if (threadIdx.x < 27) {
      const int h = hist[threadIdx.x];
      if (h > 0) {
        const int c = threadIdx.x;
        const int dtile = (tx + c % 3 - 1) + tg.nt[0] * ((ty + (c / 3) % 3 - 1) + tg.nt[1] * (tz + c / 9 - 1));
        hbase[c] = atomicAdd(&cnt_next[dtile], h);
      }
    }
    __syncthreads();
    if (keep) {
      const int dtile = (tx + code % 3 - 1) +
                        tg.nt[0] * ((ty + (code / 3) % 3 - 1) + tg.nt[1] * (tz + code / 9 - 1));
      const int slot = hbase[code] + lrank;
      const int64_t dst = slot < seg_cap[dtile] ? (int64_t)seg_start[dtile] + slot
                                                : sc.ovf_base + atomicAdd(ovf_next, 1);
      store_rec<D>(B, dst, q, tg.bx, tg.by);
    }
  }
  __syncthreads();
  float* P = patch + (int64_t)t * D * tg.pvol;
  for (int q = threadIdx.x; q < D * tg.pvol; q += blockDim.x) P[q] = Jl[q];
}

// ----------------------------------------------------- overflow kernel
template <int D>
__global__ __launch_bounds__(256) void k_ovf_particles(
    TileGeom tg, DevGrid g, StepConst sc, const float* __restrict__ ebf, RecPtr A, RecPtr B,
    const int* __restrict__ seg_start, const int* __restrict__ seg_cap,
    const int* __restrict__ ovf_cur, int* __restrict__ cnt_next, int* __restrict__ ovf_next,
    float* __restrict__ jc) {
  const int cnt = *ovf_cur;
  const GlobalField F{ebf, &g};
  const GlobalAcc32 acc{jc, &g};
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < cnt; n += gridDim.x * blockDim.x) {
    Rec q;
    load_rec<D>(A, sc.ovf_base + n, q, tg.bx, tg.by);
    float ef[6];
    gather<D>(q.N[0] - 1, q.N[1] - 1, D == 3 ? q.N[2] - 1 : 1, q.s, F, ef);
    const int io[3] = {q.N[0], q.N[1], q.N[2]};
    const float so[3] = {q.s[0], q.s[1], q.s[2]};
    const float dtg = push<D>(q.p, ef, sc.qch, sc.alp, sc.dt, sc.initial);
    if (!move_and_test<D>(q, dtg, sc.inv, sc.nn)) continue;
    esirkepov<D, float>(io, so, q.N, q.s, sc.qch * q.w, acc);
    const int dtile = (q.N[0] - 3) / tg.T[0] +
                      tg.nt[0] * ((q.N[1] - 3) / tg.T[1] + tg.nt[1] * (D == 3 ? (q.N[2] - 3) / tg.T[2] : 0));
    const int slot = atomicAdd(&cnt_next[dtile], 1);
    const int64_t dst = slot < seg_cap[dtile] ? (int64_t)seg_start[dtile] + slot
                                              : sc.ovf_base + atomicAdd(ovf_next, 1);
    store_rec<D>(B, dst, q, tg.bx, tg.by);
  }
}

// ------------------------------------------------------------ fold kernel
// jc(node) [+]= sum over the tile patches covering node (at most two tiles
// per axis since T >= 4), compensated summation.
__global__ __launch_bounds__(256) void k_fold(TileGeom tg, DevGrid g, const float* __restrict__ patch,
                                              float* __restrict__ jc, int accumulate) {
  const int D = g.ndim;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < g.ncell;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int idx[3] = {(int)(t % g.n1p) + 1, (int)((t / g.n1p) % g.n2p) + 1,
                        (int)(t / ((int64_t)g.n1p * g.n2p)) + 1};
    int tl[3][2], loc[3][2], nt[3];
    for (int d = 0; d < 3; d++) {
      nt[d] = 0;
      if (d >= D) {
        tl[d][0] = 0;
        loc[d][0] = 0;
        nt[d] = 1;
        continue;
      }
      const int u = idx[d] - 3;  // >= -2
      int thi = (u + 2) / tg.T[d];
      if (thi > tg.nt[d] - 1) thi = tg.nt[d] - 1;
      for (int tt = (thi > 0 ? thi - 1 : 0); tt <= thi; tt++) {
        const int o = 1 + tt * tg.T[d];  // patch origin J0 = T0 - 2
        const int l = idx[d] - o;
        if (l >= 0 && l < tg.J[d]) {
          tl[d][nt[d]] = tt;
          loc[d][nt[d]] = l;
          nt[d]++;
        }
      }
    }
    for (int c = 0; c < D; c++) {
      float sum = accumulate ? jc[(int64_t)c * g.ncell + t] : 0.f, comp = 0.f;
      for (int a = 0; a < nt[2]; a++)
        for (int b = 0; b < nt[1]; b++)
          for (int e = 0; e < nt[0]; e++) {
            const int tile = tl[0][e] + tg.nt[0] * (tl[1][b] + tg.nt[1] * tl[2][a]);
            const int l = loc[0][e] + tg.J[0] * (loc[1][b] + tg.J[1] * loc[2][a]);
            const float v = patch[((int64_t)tile * D + c) * tg.pvol + l];
            const float y = v - comp;
            const float s2 = sum + y;
            comp = (s2 - sum) - y;
            sum = s2;
          }
      jc[(int64_t)c * g.ncell + t] = sum;
    }
  }
}

// ------------------------------------------------- relayout / transfers
static constexpr int kB = 256;

static unsigned grid_for(int64_t n, int64_t cap = 65535LL * 8) {
  int64_t b = (n + kB - 1) / kB;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

__device__ inline int tile_of(const TileGeom& tg, int cell, int D) {
  const int i = cell & ((1 << tg.bx) - 1);
  const int j = (cell >> tg.bx) & ((1 << tg.by) - 1);
  const int k = D == 3 ? (cell >> (tg.bx + tg.by)) : 3;
  return (i - 3) / tg.T[0] + tg.nt[0] * ((j - 3) / tg.T[1] + tg.nt[1] * (D == 3 ? (k - 3) / tg.T[2] : 0));
}

__device__ inline int block_scan_excl(int v, int* sh, int& total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kB; off <<= 1) {
    int add = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  total = sh[kB - 1];
  const int ex = sh[t] - v;
  __syncthreads();
  return ex;
}

// out[t] = exclusive scan of f(in[t]) for t < n, out[n] = total; one workgroup.
// mode 0: f = min(cnt, cap); mode 1: f = capacity for count (cnt + cnt/4 + 16)
__global__ __launch_bounds__(kB) void k_scan_tiles(const int* cnt, const int* cap, int n, int mode,
                                                    int* out, int* cap_out) {
  __shared__ int sh[kB];
  long long carry = 0;
  for (int b0 = 0; b0 < n; b0 += kB) {
    const int t = b0 + threadIdx.x;
    int v = 0;
    if (t < n) {
      if (mode == 0) {
        v = min(cnt[t], cap[t]);
      } else {
        v = cnt[t] + cnt[t] / 4 + 16;
        cap_out[t] = v;
      }
    }
    int total;
    const int ex = block_scan_excl(v, sh, total);
    if (t < n) out[t] = (int)(carry + ex);
    carry += total;
  }
  if (threadIdx.x == 0) out[n] = (int)carry;
}

template <int D>
__device__ inline void copy_rec(const RecPtr& a, int64_t i, const RecPtr& b, int64_t o) {
  b.cell[o] = a.cell[i];
#pragma unroll
  for (int d = 0; d < D; d++) {
    b.s[d][o] = a.s[d][i];
    b.p[d][o] = a.p[d][i];
  }
  b.w[o] = a.w[i];
  if (a.id && b.id) b.id[o] = a.id[i];
}

// live records of A (tile segments, then overflow) -> contiguous U
template <int D>
__global__ __launch_bounds__(kB) void k_gather_live(RecPtr A, RecPtr U, const int* seg_start,
                                                    const int* seg_cap, const int* cnt,
                                                    const int* off, int ntiles, int64_t ovf_base,
                                                    const int* ovf_cnt) {
  const int t = blockIdx.x;
  if (t < ntiles) {
    const int n = min(cnt[t], seg_cap[t]);
    for (int q = threadIdx.x; q < n; q += blockDim.x)
      copy_rec<D>(A, (int64_t)seg_start[t] + q, U, (int64_t)off[t] + q);
  } else {
    const int n = *ovf_cnt;
    const int64_t o0 = off[ntiles];
    for (int q = (t - ntiles) * blockDim.x + threadIdx.x; q < n; q += (gridDim.x - ntiles) * blockDim.x)
      copy_rec<D>(A, ovf_base + q, U, o0 + q);
  }
}

template <int D>
__global__ void k_hist(TileGeom tg, RecPtr U, int64_t n, int* hcount) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    atomicAdd(&hcount[tile_of(tg, U.cell[q], D)], 1);
}

template <int D>
__global__ void k_scatter(TileGeom tg, RecPtr U, int64_t n, RecPtr A, const int* seg_start,
                          int* fillc) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int t = tile_of(tg, U.cell[q], D);
    const int slot = atomicAdd(&fillc[t], 1);
    copy_rec<D>(U, q, A, (int64_t)seg_start[t] + slot);
  }
}

// reference layout part(ld, 2D+1) -> records (ids = input index)
template <int D>
__global__ void k_convert_in(DevGrid g, int bx, int by, const double* part, int64_t ld,
                             int64_t np, RecPtr U) {
  const double mn[3] = {g.xmn, g.ymn, g.zmn};
  const double inv[3] = {g.dx_inv, g.dy_inv, g.dz_inv};
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < np;
       n += (int64_t)gridDim.x * blockDim.x) {
    int N[3] = {1, 1, 1};
#pragma unroll
    for (int d = 0; d < D; d++) {
      double s;
      ref_node(part[d * ld + n], mn[d], inv[d], N[d], s);
      float sf = (float)s;
      if (sf >= 0.5f) {  // keep the [-1/2, 1/2) convention after rounding
        N[d] += 1;
        sf -= 1.f;
      }
      U.s[d][n] = sf;
      U.p[d][n] = (float)part[(D + d) * ld + n];
    }
    U.cell[n] = N[0] | (N[1] << bx) | (D == 3 ? (N[2] << (bx + by)) : 0);
    const unsigned long long b = (unsigned long long)__double_as_longlong(part[2 * D * ld + n]);
    U.w[n] = __uint_as_float((unsigned)(b >> 32));
    if (U.id) U.id[n] = (int)n;
  }
}

// records -> reference layout (weights repacked on the host)
template <int D>
__global__ void k_convert_out(DevGrid g, int bx, int by, RecPtr U, int64_t n, double* out,
                              int64_t ld) {
  const double mn[3] = {g.xmn, g.ymn, g.zmn};
  const double dd[3] = {g.dx, g.dy, g.dz};
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int cell = U.cell[q];
    const int N[3] = {cell & ((1 << bx) - 1), (cell >> bx) & ((1 << by) - 1),
                      D == 3 ? (cell >> (bx + by)) : 1};
#pragma unroll
    for (int d = 0; d < D; d++) {
      out[d * ld + q] = mn[d] + ((double)(N[d] - 3) + (double)U.s[d][q]) * dd[d];
      out[(D + d) * ld + q] = (double)U.p[d][q];
    }
    out[2 * D * ld + q] = (double)U.w[q];
  }
}

static void set_ptrs(TileStore* ts, int b, int D) {
  // layout of one buffer: cell | s[D] | p[D] | w | id
  char* m = (char*)ts->mem[b];
  const int64_t n = ts->alloc;
  RecPtr& r = ts->buf[b];
  r.cell = (int*)m;
  m += n * 4;
  for (int d = 0; d < 3; d++) {
    r.s[d] = d < D ? (float*)m : nullptr;
    if (d < D) m += n * 4;
  }
  for (int d = 0; d < 3; d++) {
    r.p[d] = d < D ? (float*)m : nullptr;
    if (d < D) m += n * 4;
  }
  r.w = (float*)m;
  m += n * 4;
  r.id = ts->track_ids ? (int*)m : nullptr;
}

#define TS_CK(call)                                                       \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) {                                               \
      *msg = std::string(#call " failed: ") + hipGetErrorString(e_);      \
      return ALADYN_EHIP;                                                 \
    }                                                                     \
  } while (0)

void tile_store_free(TileStore* ts) {
  for (int b = 0; b < 2; b++) {
    (void)hipFree(ts->mem[b]);
    ts->mem[b] = nullptr;
    (void)hipFree(ts->count[b]);
    ts->count[b] = nullptr;
  }
  (void)hipFree(ts->seg_start);
  (void)hipFree(ts->seg_cap);
  (void)hipFree(ts->ovf_count);
  (void)hipFree(ts->scratch);
  (void)hipFree(ts->d_total);
  if (ts->h_ovf) (void)hipHostFree(ts->h_ovf);
  if (ts->ovf_ev) (void)hipEventDestroy(ts->ovf_ev);
  *ts = TileStore();
}

// Counting sort of the n records of U (stored in buffer 1-cur) into buffer
// cur with capacities cnt + cnt/4 + 16 per tile.
template <int D>
static int sort_into_cur(TileStore* ts, const TileGeom& tg, int64_t n, hipStream_t s,
                         std::string* msg) {
  const int cur = ts->cur, oth = 1 - cur;
  int* h = ts->count[cur];
  fill<int>(h, tg.ntiles, 0, s);
  if (n > 0) k_hist<D><<<grid_for(n), kB, 0, s>>>(tg, ts->buf[oth], n, h);
  k_scan_tiles<<<1, kB, 0, s>>>(h, nullptr, tg.ntiles, 1, ts->seg_start, ts->seg_cap);
  int total = 0;
  TS_CK(hipMemcpyAsync(&total, ts->seg_start + tg.ntiles, sizeof(int), hipMemcpyDeviceToHost, s));
  TS_CK(hipStreamSynchronize(s));
  if ((int64_t)total > ts->ovf_base) {
    *msg = "tile store: segment capacities exceed the allocation";
    return ALADYN_ENOMEM;
  }
  fill<int>(ts->scratch, tg.ntiles, 0, s);
  if (n > 0)
    k_scatter<D><<<grid_for(n), kB, 0, s>>>(tg, ts->buf[oth], n, ts->buf[cur], ts->seg_start,
                                            ts->scratch);
  fill<int>(ts->ovf_count, 2, 0, s);
  TS_CK(hipGetLastError());
  ts->np = n;
  ts->steps_since_sort = 0;
  return 0;
}

int tile_store_put(TileStore* ts, const TileGeom& tg, const DevGrid& g, const double* part,
                   int64_t ld, int64_t np, hipStream_t s, bool track_ids, std::string* msg) {
  const int D = g.ndim;
  // the FP32 record keeps the FP32 weight; charge and part_ind are per species
  int charge = 0, pind = 0;
  for (int64_t n = 0; n < np; n++) {
    unsigned long long b;
    std::memcpy(&b, part + 2 * D * ld + n, 8);
    const int c = (int)(short)(b & 0xffff), p = (int)(short)((b >> 16) & 0xffff);
    if (n == 0) {
      charge = c;
      pind = p;
    } else if (c != charge || p != pind) {
      *msg = "FP32 tile path needs one charge and part_ind per species";
      return ALADYN_EINVAL;
    }
  }
  const int64_t need = np + np / 4 + 16LL * tg.ntiles + 64;  // segments
  const int64_t alloc = need + (np > 0 ? np : 1);            // + overflow region
  if (alloc >= (1LL << 31)) {
    *msg = "FP32 tile path: particle count exceeds the 31-bit slot range";
    return ALADYN_EINVAL;
  }
  if (!ts->loaded || ts->alloc < alloc || ts->track_ids != track_ids) {
    tile_store_free(ts);
    ts->track_ids = track_ids;
    ts->alloc = alloc;
    const size_t per = (size_t)(2 + 2 * D + (track_ids ? 1 : 0)) * 4;
    for (int b = 0; b < 2; b++) {
      TS_CK(hipMalloc(&ts->mem[b], per * (size_t)alloc));
      set_ptrs(ts, b, D);
      TS_CK(hipMalloc(&ts->count[b], sizeof(int) * (tg.ntiles + 1)));
    }
    TS_CK(hipMalloc(&ts->seg_start, sizeof(int) * (tg.ntiles + 1)));
    TS_CK(hipMalloc(&ts->seg_cap, sizeof(int) * (tg.ntiles + 1)));
    TS_CK(hipMalloc(&ts->ovf_count, sizeof(int) * 2));
    TS_CK(hipMalloc(&ts->scratch, sizeof(int) * (tg.ntiles + 1)));
    TS_CK(hipMalloc(&ts->d_total, 64));
    TS_CK(hipHostMalloc(&ts->h_ovf, 2 * sizeof(int), hipHostMallocDefault));
    TS_CK(hipEventCreateWithFlags(&ts->ovf_ev, hipEventDisableTiming));
    ts->h_ovf[0] = ts->h_ovf[1] = 0;
  }
  ts->ovf_cap = np > 0 ? np : 1;
  ts->ovf_base = ts->alloc - ts->ovf_cap;
  ts->charge = charge;
  ts->part_ind = pind;
  ts->cur = 0;
  double* dpart = nullptr;
  const int ncol = 2 * D + 1;
  if (np > 0) {
    TS_CK(hipMalloc(&dpart, (size_t)np * ncol * sizeof(double)));
    for (int c = 0; c < ncol; c++)
      TS_CK(hipMemcpyAsync(dpart + c * np, part + c * ld, (size_t)np * sizeof(double),
                           hipMemcpyHostToDevice, s));
    if (D == 2)
      k_convert_in<2><<<grid_for(np), kB, 0, s>>>(g, tg.bx, tg.by, dpart, np, np, ts->buf[1]);
    else
      k_convert_in<3><<<grid_for(np), kB, 0, s>>>(g, tg.bx, tg.by, dpart, np, np, ts->buf[1]);
    TS_CK(hipGetLastError());
  }
  int r = D == 2 ? sort_into_cur<2>(ts, tg, np, s, msg) : sort_into_cur<3>(ts, tg, np, s, msg);
  TS_CK(hipStreamSynchronize(s));
  (void)hipFree(dpart);
  if (r) return r;
  ts->loaded = true;
  return 0;
}

// gathers the live records of buffer cur into buffer 1-cur; returns count
template <int D>
static int gather_live(TileStore* ts, const TileGeom& tg, hipStream_t s, int64_t* n,
                       std::string* msg) {
  const int cur = ts->cur, oth = 1 - cur;
  k_scan_tiles<<<1, kB, 0, s>>>(ts->count[cur], ts->seg_cap, tg.ntiles, 0, ts->scratch, nullptr);
  k_gather_live<D><<<tg.ntiles + 64, kB, 0, s>>>(ts->buf[cur], ts->buf[oth], ts->seg_start,
                                                 ts->seg_cap, ts->count[cur], ts->scratch,
                                                 tg.ntiles, ts->ovf_base, ts->ovf_count + cur);
  TS_CK(hipGetLastError());
  int h[2] = {0, 0};
  TS_CK(hipMemcpyAsync(&h[0], ts->scratch + tg.ntiles, sizeof(int), hipMemcpyDeviceToHost, s));
  TS_CK(hipMemcpyAsync(&h[1], ts->ovf_count + cur, sizeof(int), hipMemcpyDeviceToHost, s));
  TS_CK(hipStreamSynchronize(s));
  *n = (int64_t)h[0] + h[1];
  return 0;
}

int tile_store_relayout(TileStore* ts, const TileGeom& tg, const DevGrid& g, hipStream_t s,
                        std::string* msg) {
  if (!ts->loaded) return 0;
  int64_t n = 0;
  const int D = g.ndim;
  int r = D == 2 ? gather_live<2>(ts, tg, s, &n, msg) : gather_live<3>(ts, tg, s, &n, msg);
  if (r) return r;
  return D == 2 ? sort_into_cur<2>(ts, tg, n, s, msg) : sort_into_cur<3>(ts, tg, n, s, msg);
}

int tile_store_count(TileStore* ts, const TileGeom& tg, hipStream_t s, int64_t* np,
                     std::string* msg) {
  *np = 0;
  if (!ts->loaded) return 0;
  // live = sum over tiles of min(count, cap) + overflow
  const int cur = ts->cur;
  k_scan_tiles<<<1, kB, 0, s>>>(ts->count[cur], ts->seg_cap, tg.ntiles, 0, ts->scratch, nullptr);
  TS_CK(hipGetLastError());
  int h[2] = {0, 0};
  TS_CK(hipMemcpyAsync(&h[0], ts->scratch + tg.ntiles, sizeof(int), hipMemcpyDeviceToHost, s));
  TS_CK(hipMemcpyAsync(&h[1], ts->ovf_count + cur, sizeof(int), hipMemcpyDeviceToHost, s));
  TS_CK(hipStreamSynchronize(s));
  *np = (int64_t)h[0] + h[1];
  return 0;
}

int tile_store_get(TileStore* ts, const TileGeom& tg, const DevGrid& g, double* part, int64_t ld,
                   hipStream_t s, std::string* msg) {
  if (!ts->loaded) return 0;
  const int D = g.ndim, ncol = 2 * D + 1;
  int64_t n = 0;
  int r = D == 2 ? gather_live<2>(ts, tg, s, &n, msg) : gather_live<3>(ts, tg, s, &n, msg);
  if (r) return r;
  if (n == 0) return 0;
  if (ld < n) {
    *msg = "get_species: ld smaller than the particle count";
    return ALADYN_EINVAL;
  }
  double* d = nullptr;
  TS_CK(hipMalloc(&d, (size_t)n * ncol * sizeof(double)));
  const RecPtr& U = ts->buf[1 - ts->cur];
  if (D == 2)
    k_convert_out<2><<<grid_for(n), kB, 0, s>>>(g, tg.bx, tg.by, U, n, d, n);
  else
    k_convert_out<3><<<grid_for(n), kB, 0, s>>>(g, tg.bx, tg.by, U, n, d, n);
  std::vector<double> h((size_t)n * ncol);
  std::vector<int> ids;
  TS_CK(hipMemcpyAsync(h.data(), d, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  if (ts->track_ids) {
    ids.resize((size_t)n);
    TS_CK(hipMemcpyAsync(ids.data(), U.id, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, s));
  }
  TS_CK(hipStreamSynchronize(s));
  (void)hipFree(d);
  std::vector<int64_t> order((size_t)n);
  std::iota(order.begin(), order.end(), 0);
  if (ts->track_ids)
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return ids[a] < ids[b]; });
  // repack the weight slot: int16 charge | int16 part_ind | FP32 wgh
  for (int64_t q = 0; q < n; q++) {
    const int64_t src = order[q];
    for (int c = 0; c < 2 * D; c++) part[c * ld + q] = h[c * n + src];
    const float w = (float)h[2 * D * n + src];
    unsigned int wb;
    std::memcpy(&wb, &w, 4);
    const unsigned long long b = (unsigned long long)(unsigned short)(short)ts->charge |
                                 ((unsigned long long)(unsigned short)(short)ts->part_ind << 16) |
                                 ((unsigned long long)wb << 32);
    std::memcpy(part + 2 * D * ld + q, &b, 8);
  }
  return 0;
}

int tile_store_step(TileStore* ts, const TileGeom& tg, const DevGrid& g, const float* ebf,
                    float* jc, float* patch, float alp, int initial, bool j_is_zero,
                    hipStream_t s, std::string* msg) {
  if (!ts->loaded) return 0;
  const int D = g.ndim, cur = ts->cur, nxt = 1 - cur;
  StepConst sc;
  sc.alp = alp;
  sc.qch = (float)ts->charge;
  sc.dt = (float)g.dt;
  sc.inv[0] = (float)g.dx_inv;
  sc.inv[1] = (float)g.dy_inv;
  sc.inv[2] = (float)g.dz_inv;
  sc.nn[0] = g.nx;
  sc.nn[1] = g.ny;
  sc.nn[2] = g.nz;
  sc.initial = initial;
  sc.ovf_base = ts->ovf_base;
  fill<int>(ts->count[nxt], tg.ntiles, 0, s);
  fill<int>(ts->ovf_count + nxt, 1, 0, s);
  if (D == 2)
    k_tile_particles<2><<<tg.ntiles, tg.block, tg.lds_bytes, s>>>(
        tg, g, sc, ebf, ts->buf[cur], ts->buf[nxt], ts->seg_start, ts->seg_cap, ts->count[cur],
        ts->count[nxt], ts->ovf_count + nxt, patch);
  else
    k_tile_particles<3><<<tg.ntiles, tg.block, tg.lds_bytes, s>>>(
        tg, g, sc, ebf, ts->buf[cur], ts->buf[nxt], ts->seg_start, ts->seg_cap, ts->count[cur],
        ts->count[nxt], ts->ovf_count + nxt, patch);
  TS_CK(hipGetLastError());
  k_fold<<<grid_for(g.ncell, 8192), kB, 0, s>>>(tg, g, patch, jc, j_is_zero ? 0 : 1);
  if (D == 2)
    k_ovf_particles<2><<<512, kB, 0, s>>>(tg, g, sc, ebf, ts->buf[cur], ts->buf[nxt], ts->seg_start,
                                          ts->seg_cap, ts->ovf_count + cur, ts->count[nxt],
                                          ts->ovf_count + nxt, jc);
  else
    k_ovf_particles<3><<<512, kB, 0, s>>>(tg, g, sc, ebf, ts->buf[cur], ts->buf[nxt], ts->seg_start,
                                          ts->seg_cap, ts->ovf_count + cur, ts->count[nxt],
                                          ts->ovf_count + nxt, jc);
  TS_CK(hipGetLastError());
  ts->cur = nxt;
  ts->steps_since_sort++;
  // lazily watch the overflow count; re-layout on cadence or on growth
  bool relayout = ts->steps_since_sort >= tg.sort_every;
  if (hipEventQuery(ts->ovf_ev) == hipSuccess && ts->h_ovf[0] > (int)(ts->ovf_cap / 64 + 1024))
    relayout = true;
  TS_CK(hipMemcpyAsync(ts->h_ovf, ts->ovf_count + nxt, sizeof(int), hipMemcpyDeviceToHost, s));
  TS_CK(hipEventRecord(ts->ovf_ev, s));
  if (relayout) return tile_store_relayout(ts, tg, g, s, msg);
  return 0;
}

// ------------------------------------------------------------ diagnostics
template <int D>
__global__ void k_grid_charge_f32(TileGeom tg, RecPtr U, int64_t n, float qch, double* rho,
                                  DevGrid g) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int cell = U.cell[q];
    const int N[3] = {cell & ((1 << tg.bx) - 1), (cell >> tg.bx) & ((1 << tg.by) - 1),
                      D == 3 ? (cell >> (tg.bx + tg.by)) : 1};
    double a[3][3] = {{0, 0, 0}, {0, 0, 0}, {1, 0, 0}};
#pragma unroll
    for (int d = 0; d < D; d++) qweights<double>((double)U.s[d][q], a[d]);
    const double wq = (double)(qch * U.w[q]);
    const int nk = D == 3 ? 3 : 1;
    for (int k1 = 0; k1 < nk; k1++)
      for (int j1 = 0; j1 < 3; j1++)
        for (int i1 = 0; i1 < 3; i1++)
          unsafeAtomicAdd(rho + fofs(g, N[0] - 1 + i1, N[1] - 1 + j1, D == 3 ? N[2] - 1 + k1 : 1),
                          (wq * a[0][i1]) * (a[1][j1] * a[2][k1]));
  }
}

void tile_store_grid_charge(TileStore* ts, const TileGeom& tg, const DevGrid& g, double* rho,
                            hipStream_t s) {
  if (!ts->loaded) return;
  std::string msg;
  int64_t n = 0;
  const int D = g.ndim;
  if ((D == 2 ? gather_live<2>(ts, tg, s, &n, &msg) : gather_live<3>(ts, tg, s, &n, &msg)) || n == 0)
    return;
  const RecPtr& U = ts->buf[1 - ts->cur];
  if (D == 2)
    k_grid_charge_f32<2><<<grid_for(n), kB, 0, s>>>(tg, U, n, (float)ts->charge, rho, g);
  else
    k_grid_charge_f32<3><<<grid_for(n), kB, 0, s>>>(tg, U, n, (float)ts->charge, rho, g);
}

__device__ inline void amax_nonneg(double* addr, double v) {
  atomicMax((unsigned long long*)addr, (unsigned long long)__double_as_longlong(v));
}

template <int D>
__global__ __launch_bounds__(kB) void k_energy_f32(TileGeom tg, DevGrid g, RecPtr U, int64_t n,
                                                   float qch, double* out) {
  __shared__ double sh[8][kB / 64];
  double ek[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    double p[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; d++) p[d] = (double)U.p[d][q];
    const double w = (double)U.w[q];
    const double gam1 = sqrt(1. + p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) - 1.;
    if constexpr (D == 2) {
      ek[0] += w * p[0];
      ek[1] += w * p[1];
    } else {
      const int cell = U.cell[q];
      const int Nj = (cell >> tg.bx) & ((1 << tg.by) - 1), Nk = cell >> (tg.bx + tg.by);
      const double y = g.ymn + ((double)(Nj - 3) + (double)U.s[1][q]) * g.dy;
      const double z = g.zmn + ((double)(Nk - 3) + (double)U.s[2][q]) * g.dz;
      ek[0] += p[0];
      ek[1] += p[1];
      ek[2] += p[2];
      ek[3] += w * (y * p[2] - z * p[1]);
    }
    ek[5] += (double)(qch * U.w[q]);
    ek[6] += w * gam1;
    ek[7] = fmax(ek[7], gam1);
  }
  for (int c = 0; c < 8; c++) {
    double v = ek[c];
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_down(v, off, 64);
      v = c == 7 ? fmax(v, o) : v + o;
    }
    if ((threadIdx.x & 63) == 0) sh[c][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int c = threadIdx.x;
    double v = sh[c][0];
    for (int w = 1; w < kB / 64; w++) v = c == 7 ? fmax(v, sh[c][w]) : v + sh[c][w];
    if (c == 7)
      amax_nonneg(out + 7, v);
    else
      unsafeAtomicAdd(out + c, v);
  }
}

void tile_store_energy(TileStore* ts, const TileGeom& tg, const DevGrid& g, double* out,
                       hipStream_t s) {
  if (!ts->loaded) return;
  std::string msg;
  int64_t n = 0;
  const int D = g.ndim;
  if ((D == 2 ? gather_live<2>(ts, tg, s, &n, &msg) : gather_live<3>(ts, tg, s, &n, &msg)) || n == 0)
    return;
  const RecPtr& U = ts->buf[1 - ts->cur];
  if (D == 2)
    k_energy_f32<2><<<grid_for(n, 4096), kB, 0, s>>>(tg, g, U, n, (float)ts->charge, out);
  else
    k_energy_f32<3><<<grid_for(n, 4096), kB, 0, s>>>(tg, g, U, n, (float)ts->charge, out);
}

}  // namespace aladyn
