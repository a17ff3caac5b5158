// envelope.hip -- the envelope (ponderomotive) model's step on the device:
// env_lpf2_evolve (env_evolve_in_time.f90:196-361) for the first species on
// one rank, every face open, one laser colour, no ionization and no fluid.
//
// The reference reuses its current array jc(n1p, n2p, n3p, curr_ndim + 1)
// through the step; the device keeps that array (here `je`, in the grid's
// precision) and reuses it the same way, so every ghost value the particle
// gathers read is the one the reference's would read:
//   je(1) = |A^n|^2/2, je(2:D+1) = its staggered gradient   (env_amp_prepare)
//   particles: fields at x^n, the envelope pusher's momenta, and the density
//   wgh/gamp deposited at x^n into a separate FP64 array `den` (the gather
//   still reads je(1) while particles deposit)
//   den_zyxbd(den); je(1) = den; je(3) = den                (set_env_density)
//   je(1:2) = -ompe je(3) A^n, + Laplacian, + 2 om0 D_x     (advance_lpf_envelope)
//   A^{n+1}, A^{n-1} <- A^n                                 (env_maxw_solve)
//   je(1) = |(A^{n+1}+A^n)/2|^2/2, je(2:D+1) gradient       (env_fields_average)
//   particles: grad F, F at x^n, positions (lpf_env_positions), removal at
//   the faces, Esirkepov deposit, re-binning; then the current fold and the
//   field update of the LP step.
#include "env.hpp"
#include "tile_device.hpp"

namespace aladyn {

static constexpr int kEB = 256;

static unsigned blocks_for(int64_t n) {
  int64_t b = (n + kEB - 1) / kEB;
  if (b < 1) b = 1;
  if (b > 65535LL * 8) b = 65535LL * 8;
  return (unsigned)b;
}

// sum_{k1,j1,i1} wx[i1] wy[j1] wz[k1] f(i0+i1, j0+j1, k0+k1), in the
// reference's loop order (x innermost, dvol = wy*wz first)
template <int D, typename R>
__device__ inline R contract3(const DevGrid& g, const R* f, const R wx[3], const R wy[3],
                              const R wz[3], int i0, int j0, int k0) {
  R r = R(0);
  const int nk = D == 3 ? 3 : 1;
  for (int k1 = 0; k1 < nk; k1++)
#pragma unroll
    for (int j1 = 0; j1 < 3; j1++) {
      const R dvol = D == 3 ? wy[j1] * wz[k1] : wy[j1];
      const R* row = f + fofs(g, i0, j0 + j1, D == 3 ? k0 + k1 : 1);
#pragma unroll
      for (int i1 = 0; i1 < 3; i1++) r = r + (dvol * wx[i1]) * row[i1];
    }
  return r;
}

template <typename R>
__device__ inline const R* comp_c(const R* a, const DevGrid& g, int c) {
  return a + (int64_t)c * g.ncell;
}

// iterates over the records of the overflow list (the tiles' segments are
// the tile workgroups' below)
template <typename F>
__device__ inline void for_overflow(int64_t ovf_base, const int* ovf_cnt, F&& f) {
  const int n = *ovf_cnt;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
    f(ovf_base + q);
}

// ---------------------------------------------------- particles, phase 1
// set_env_acc (grid_part_connect.f90:414-615) + lpf_env_momenta
// (boris_push.f90:170-223) + set_env_density (grid_part_connect.f90:884-952)
// per particle: E, B, F = je(1) and grad F = je(2:D+1) at x^n with qqh
// shapes; gamp from p^{n-1/2}, F and the half-step energy change; the
// rotation with gamp fixed gives p^{n+1/2} (written back in place); den +=
// (wgh/gamp) x the quadratic shapes at x^n.  This kernel takes the overflow
// list, reading the grids in HBM and adding to den with global atomics.
template <int D, typename R>
__global__ __launch_bounds__(kEB) void k_env_push(TileGeom tg, DevGrid g, RecPtr A,
                                                  int64_t ovf_base, const int* ovf_cnt,
                                                  const R* ebf, const R* je, R qch, R dth, R alp,
                                                  double* den) {
  for_overflow(ovf_base, ovf_cnt, [&](int64_t n) {
    Rec<R> q;
    load_rec<D, R>(A, n, q, tg.bx, tg.by);
    R ax[3], axh[3], ay[3], ayh[3], az[3] = {R(1), R(0), R(0)}, azh[3] = {R(1), R(0), R(0)};
    int hx, hy, hz = 0;
    qqh<R>(q.s[0], ax, axh, hx);
    qqh<R>(q.s[1], ay, ayh, hy);
    if (D == 3) qqh<R>(q.s[2], az, azh, hz);
    const int i = q.N[0] - 1, j = q.N[1] - 1, k = D == 3 ? q.N[2] - 1 : 1;
    const int ih = i + hx, jh = j + hy, kh = D == 3 ? k + hz : 1;
    const R one = R(1), half = R(0.5);
    if constexpr (D == 2) {
      // ap: Ex, Ey, Bz, dFx, dFy, F
      R ap[6];
      ap[5] = contract3<2, R>(g, comp_c(je, g, 0), ax, ay, az, i, j, 1);
      ap[0] = contract3<2, R>(g, comp_c(ebf, g, 0), axh, ay, az, ih, j, 1);
      ap[3] = contract3<2, R>(g, comp_c(je, g, 1), axh, ay, az, ih, j, 1);
      ap[1] = contract3<2, R>(g, comp_c(ebf, g, 1), ax, ayh, az, i, jh, 1);
      ap[4] = contract3<2, R>(g, comp_c(je, g, 2), ax, ayh, az, i, jh, 1);
      ap[2] = contract3<2, R>(g, comp_c(ebf, g, 2), axh, ayh, az, ih, jh, 1);
      const R u1 = q.p[0], u2 = q.p[1];
      const R gam2 = one + u1 * u1 + u2 * u2 + ap[5];
      for (int c = 0; c < 3; c++) ap[c] = qch * ap[c];
      for (int c = 3; c < 5; c++) ap[c] = half * qch * qch * ap[c];
      const R aa1 = dth * (ap[0] * u1 + ap[1] * u2);
      const R b1 = dth * (ap[3] * u1 + ap[4] * u2);
      const R gam = sqrt(gam2);
      const R dgam = (aa1 * gam - b1) / gam2;
      const R gam_inv = (gam - dgam) / gam2;
      for (int c = 2; c < 5; c++) ap[c] = ap[c] * gam_inv;
      const R e1 = alp * (ap[0] - ap[3]), e2 = alp * (ap[1] - ap[4]), bb1 = alp * ap[2];
      const R vp0 = u1 + e1, vp1 = u2 + e2;
      const R b2 = one + bb1 * bb1;
      const R vh0 = (vp0 + vp1 * bb1) / b2, vh1 = (vp1 - vp0 * bb1) / b2;
      q.p[0] = R(2) * vh0 - u1;
      q.p[1] = R(2) * vh1 - u2;
      const R wghp = (R)q.w * gam_inv;
      for (int j1 = 0; j1 < 3; j1++) {
        const R dvol = ay[j1] * wghp;
        for (int i1 = 0; i1 < 3; i1++)
          unsafeAtomicAdd(den + fofs(g, i + i1, j + j1, 1), (double)(dvol * ax[i1]));
      }
    } else {
      // ap: Ex, Ey, Ez, Bx, By, Bz, dFx, dFy, dFz, F
      R ap[10];
      ap[9] = contract3<3, R>(g, comp_c(je, g, 0), ax, ay, az, i, j, k);
      ap[0] = contract3<3, R>(g, comp_c(ebf, g, 0), axh, ay, az, ih, j, k);
      ap[6] = contract3<3, R>(g, comp_c(je, g, 1), axh, ay, az, ih, j, k);
      ap[1] = contract3<3, R>(g, comp_c(ebf, g, 1), ax, ayh, az, i, jh, k);
      ap[7] = contract3<3, R>(g, comp_c(je, g, 2), ax, ayh, az, i, jh, k);
      ap[5] = contract3<3, R>(g, comp_c(ebf, g, 5), axh, ayh, az, ih, jh, k);
      ap[3] = contract3<3, R>(g, comp_c(ebf, g, 3), ax, ayh, azh, i, jh, kh);
      ap[4] = contract3<3, R>(g, comp_c(ebf, g, 4), axh, ay, azh, ih, j, kh);
      ap[2] = contract3<3, R>(g, comp_c(ebf, g, 2), ax, ay, azh, i, j, kh);
      ap[8] = contract3<3, R>(g, comp_c(je, g, 3), ax, ay, azh, i, j, kh);
      const R u[3] = {q.p[0], q.p[1], q.p[2]};
      const R gam2 = one + u[0] * u[0] + u[1] * u[1] + u[2] * u[2] + ap[9];
      for (int c = 0; c < 6; c++) ap[c] = qch * ap[c];
      for (int c = 6; c < 9; c++) ap[c] = half * qch * qch * ap[c];
      const R aa1 = dth * (ap[0] * u[0] + ap[1] * u[1] + ap[2] * u[2]);
      const R b1 = dth * (ap[6] * u[0] + ap[7] * u[1] + ap[8] * u[2]);
      const R gam = sqrt(gam2);
      const R dgam = (aa1 * gam - b1) / gam2;
      const R gam_inv = (gam - dgam) / gam2;
      for (int c = 3; c < 9; c++) ap[c] = ap[c] * gam_inv;
      R vp[3], bb[3], vph[3];
      for (int c = 0; c < 3; c++) vp[c] = alp * (ap[c] - ap[c + 6]) + u[c];
      for (int c = 0; c < 3; c++) bb[c] = alp * ap[c + 3];
      const R b2 = one + (bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2]);
      const R bv = bb[0] * vp[0] + bb[1] * vp[1] + bb[2] * vp[2];
      vph[0] = vp[0] + vp[1] * bb[2] - vp[2] * bb[1] + bb[0] * bv;
      vph[1] = vp[1] + vp[2] * bb[0] - vp[0] * bb[2] + bb[1] * bv;
      vph[2] = vp[2] + vp[0] * bb[1] - vp[1] * bb[0] + bb[2] * bv;
      for (int c = 0; c < 3; c++) q.p[c] = R(2) * (vph[c] / b2) - u[c];
      const R wghp = (R)q.w * gam_inv;
      for (int k1 = 0; k1 < 3; k1++)
        for (int j1 = 0; j1 < 3; j1++) {
          const R dvol = ay[j1] * az[k1] * wghp;
          for (int i1 = 0; i1 < 3; i1++)
            unsafeAtomicAdd(den + fofs(g, i + i1, j + j1, k + k1), (double)(dvol * ax[i1]));
        }
    }
    put_rec<D, R>(A, n, pack_cell<D>(q.N, tg.bx, tg.by), q.s, q.p, q.w);
  });
}

// ---------------------------------------------------- particles, phase 2
// Adds a particle's current to the FP64 patch of the tile its old node lies
// in (nodes T0-2 .. T0+T+1 per axis).
template <int D>
struct EnvPatchAcc {
  double* P;
  int o[3];
  int J0, J1, pvol;
  template <typename R>
  __device__ void operator()(int c, int i, int j, int k, R v) const {
    const int l = (i - o[0]) + J0 * ((j - o[1]) + J1 * (D == 3 ? k - o[2] : 0));
    unsafeAtomicAdd(P + c * pvol + l, (double)v);
  }
};

// set_env_grad_interp (grid_part_connect.f90:760-882) + lpf_env_positions
// (boris_push.f90:225-295) + cell_part_dist + the Esirkepov deposit
// (grid_part_connect.f90:979-1060, 1168-1334) + re-binning into B: grad F
// and F = je(1:D+1) at x^n with qqh shapes; gamp at t^{n+1/2} corrected by
// grad F; x^{n+1} = x^n + dt p/gamp; particles leaving the box are dropped
// before they deposit.  This kernel takes the overflow list (global atomics
// into the patches the tile workgroups wrote).
template <int D, typename R>
__global__ __launch_bounds__(kEB) void k_env_move(TileGeom tg, DevGrid g, StepConst<R> sc,
                                                  RecPtr A, RecPtr B, const int* seg_start,
                                                  const int* seg_cap, const int* ovf_cur,
                                                  const R* je, int* cnt_next, int* ovf_next,
                                                  double* patch, int* grow) {
  for_overflow(sc.ovf_base, ovf_cur, [&](int64_t n) {
    Rec<R> q;
    load_rec<D, R>(A, n, q, tg.bx, tg.by);
    R ax[3], axh[3], ay[3], ayh[3], az[3] = {R(1), R(0), R(0)}, azh[3] = {R(1), R(0), R(0)};
    int hx, hy, hz = 0;
    qqh<R>(q.s[0], ax, axh, hx);
    qqh<R>(q.s[1], ay, ayh, hy);
    if (D == 3) qqh<R>(q.s[2], az, azh, hz);
    const int i = q.N[0] - 1, j = q.N[1] - 1, k = D == 3 ? q.N[2] - 1 : 1;
    const int ih = i + hx, jh = j + hy, kh = D == 3 ? k + hz : 1;
    // ap: grad F (D), F
    R ap[4] = {R(0), R(0), R(0), R(0)};
    ap[0] = contract3<D, R>(g, comp_c(je, g, 1), axh, ay, az, ih, j, k);
    ap[D] = contract3<D, R>(g, comp_c(je, g, 0), ax, ay, az, i, j, k);
    ap[1] = contract3<D, R>(g, comp_c(je, g, 2), ax, ayh, az, i, jh, k);
    if (D == 3) ap[2] = contract3<D, R>(g, comp_c(je, g, 3), ax, ay, azh, i, j, kh);
    const R one = R(1);
    R pp2 = R(0), pv = R(0);
#pragma unroll
    for (int d = 0; d < D; d++) {
      pp2 = pp2 + q.p[d] * q.p[d];
      pv = pv + q.p[d] * ap[d];
    }
    const R gam2 = one + pp2 + ap[D];
    const R gam = sqrt(gam2);
    const R gam3 = gam2 * gam;
    const R b2 = R(0.25) * pv;
    const R gam_inv = (one / gam) * (one - sc.dt * b2 / gam3);
    const int io[3] = {q.N[0], q.N[1], q.N[2]};
    const R so[3] = {q.s[0], q.s[1], q.s[2]};
    if (!move_and_test<D, R>(q, sc.dt * gam_inv, sc.inv, sc.nn)) return;
    int tt[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; d++) tt[d] = (io[d] - 3) / tg.T[d];
    const int t0 = tt[0] + tg.nt[0] * (tt[1] + tg.nt[1] * tt[2]);
    const EnvPatchAcc<D> acc{patch + (int64_t)t0 * D * tg.pvol,
                             {1 + tt[0] * tg.T[0], 1 + tt[1] * tg.T[1], 1 + tt[2] * tg.T[2]},
                             tg.J[0], tg.J[1], tg.pvol};
    // real(sp) :: wght = charge*wgh (grid_part_connect.f90:983, 1205)
    esirkepov<D, R>(io, so, q.N, q.s, (R)((float)sc.qch * q.w), acc);
    const int dtile = tile_of(tg, pack_cell<D>(q.N, tg.bx, tg.by), D);
    const int slot = reserve_slot(dtile, cnt_next);
    int64_t dst = (int64_t)seg_start[dtile] + slot;
    if (slot >= seg_cap[dtile]) {
      dst = sc.ovf_base + atomicAdd(ovf_next, 1);
      grow[dtile] = 1;
    }
    store_rec<D, R>(B, dst, q, tg.bx, tg.by);
  });
}

// ------------------------------------------------- tile workgroups
// The same two phases with one workgroup per tile of the store: the tile's
// fields (ebf and je, or je alone) are staged in LDS over nodes T0-2 ..
// T0+T per axis (the reach of the qqh shapes of particles whose nearest
// node lies in the tile's nodes T0 .. T0+T-1), and the density or current
// is summed in an FP64 LDS patch over nodes T0-2 .. T0+T+1 (the current
// patch's extent tg.J, so that the fold of the LP step sums it).  Nodes
// of the staged box beyond the grid (the last tile per axis may reach past
// it) are zero and no shape reaches them.  The overflow list is left to
// the global-memory kernels above, launched after these.
struct EnvTile {
  int T0[3];  // first node of the tile per axis (1-based; 1 for z in 2-D)
  int G[3];   // staged box extent: T + 3 (1 for z in 2-D)
  int gvol;
};

template <int D>
__device__ inline EnvTile env_tile(const TileGeom& tg, int t) {
  EnvTile e;
  const int tt[3] = {t % tg.nt[0], (t / tg.nt[0]) % tg.nt[1], t / (tg.nt[0] * tg.nt[1])};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    e.T0[d] = d < D ? 3 + tt[d] * tg.T[d] : 1;
    e.G[d] = d < D ? tg.T[d] + 3 : 1;
  }
  e.gvol = e.G[0] * e.G[1] * e.G[2];
  return e;
}

// Stages n0 components of a, then n1 components of b (ncell stride each)
// over the tile's box into dst (component stride gvol).  The threads take
// whole x rows of the box, kStageRows rows each per pass, and every load of
// a pass is issued before its LDS stores, so a pass pays the HBM latency
// once instead of once per node.
static constexpr int kStageRows = 8;
template <int D, typename R>
__device__ inline void env_stage_rows(const DevGrid& g, const EnvTile& e, const R* a, int n0,
                                      const R* b, int n1, R* dst) {
  const int G0 = e.G[0], G1 = e.G[1], G12 = e.G[1] * e.G[2];
  // rows per pass, one node per thread (the compiled tile shapes have
  // G0 <= 35 < blockDim)
  const int rpp = max(1, (int)blockDim.x / G0);
  const int li = threadIdx.x % G0, r0 = threadIdx.x / G0;
  const int rows = (n0 + n1) * G12;
  const int i = e.T0[0] - 2 + li;
  const bool xin = r0 < rpp && i <= g.n1p;
  for (int rb = r0; rb < rows; rb += kStageRows * rpp) {
    R v[kStageRows];
#pragma unroll
    for (int u = 0; u < kStageRows; u++) {
      const int row = rb + u * rpp;
      v[u] = R(0);
      if (xin && row < rows) {
        const int c = row / G12, rem = row - c * G12;
        const int lk = rem / G1, lj = rem - lk * G1;
        const int j = e.T0[1] - 2 + lj, k = D == 3 ? e.T0[2] - 2 + lk : 1;
        const R* src = c < n0 ? comp_c(a, g, c) : comp_c(b, g, c - n0);
        if (j <= g.n2p && k <= g.n3p) v[u] = src[fofs(g, i, j, k)];
      }
    }
#pragma unroll
    for (int u = 0; u < kStageRows; u++) {
      const int row = rb + u * rpp;
      if (r0 < rpp && row < rows) dst[row * G0 + li] = v[u];
    }
  }
}

// Adds to an FP64 LDS patch over nodes o .. o+J-1 (component stride pvol).
template <int D>
struct EnvLdsAcc {
  double* P;
  int o[3];
  int J0, J1, pvol;
  template <typename R>
  __device__ void operator()(int c, int i, int j, int k, R v) const {
    const int l = (i - o[0]) + J0 * ((j - o[1]) + J1 * (D == 3 ? k - o[2] : 0));
    atomicAdd(P + c * pvol + l, (double)v);
  }
};

// Copies nc components of the tile's LDS patch to its slot of the global
// patch array (tile stride D components).
template <int D>
__device__ inline void env_patch_out(const TileGeom& tg, int t, const double* P, int nc,
                                     double* patch) {
  double* dst = patch + (int64_t)t * D * tg.pvol;
  for (int l = threadIdx.x; l < nc * tg.pvol; l += blockDim.x) dst[l] = P[l];
}

// Phase 1 per tile: set_env_acc + lpf_env_momenta + set_env_density, the
// density summed into component 0 of the tile's patch.
template <int D, typename R>
__global__ __launch_bounds__(kEB) void k_env_push_tile(TileGeom tg, DevGrid g, RecPtr A,
                                                       const int* seg_start, const int* seg_cap,
                                                       const int* cnt, const R* ebf, const R* je,
                                                       R qch, R dth, R alp, double* patch) {
  extern __shared__ double lds_d[];
  const int t = blockIdx.x;
  const EnvTile e = env_tile<D>(tg, t);
  const int gvol = e.gvol, pvol = tg.pvol;
  const int nf = g.nfield;
  double* Pd = lds_d;              // density patch, pvol
  R* Fl = (R*)(lds_d + pvol);      // nf ebf components, then je(1:D+1)
  const int n = min(cnt[t], seg_cap[t]);
  const int64_t base = seg_start[t];
  // the thread's first record is loaded with the staging loads
  Rec<R> nxt = {};
  if ((int)threadIdx.x < n) load_raw<D, R>(A, base + threadIdx.x, nxt);
  for (int l = threadIdx.x; l < pvol; l += blockDim.x) Pd[l] = 0.0;
  env_stage_rows<D, R>(g, e, ebf, nf, je, D + 1, Fl);
  __syncthreads();
  const int GX = e.G[0], GXY = e.G[0] * e.G[1];
  const R* EB = Fl;
  const R* JE = Fl + nf * gvol;
  for (int q0 = threadIdx.x; q0 < n; q0 += blockDim.x) {
    const int64_t src = base + q0;
    Rec<R> q = nxt;
    // the next record's load is in flight during this one's push
    if (q0 + (int)blockDim.x < n) load_raw<D, R>(A, src + blockDim.x, nxt);
    if (q.N[0] == kHole) continue;
    decode_cell<D, R>(q, tg.bx, tg.by);
    R ax[3], axh[3], ay[3], ayh[3], az[3] = {R(1), R(0), R(0)}, azh[3] = {R(1), R(0), R(0)};
    int hx, hy, hz = 0;
    qqh<R>(q.s[0], ax, axh, hx);
    qqh<R>(q.s[1], ay, ayh, hy);
    if (D == 3) qqh<R>(q.s[2], az, azh, hz);
    // local box indices of the integer (i) and half-integer (ih) shape bases
    const int i = q.N[0] - e.T0[0] + 1, j = q.N[1] - e.T0[1] + 1, k = D == 3 ? q.N[2] - e.T0[2] + 1 : 0;
    const int ih = i + hx, jh = j + hy, kh = D == 3 ? k + hz : 0;
    const R one = R(1), half = R(0.5);
    const EnvLdsAcc<D> acc{Pd, {e.T0[0] - 2, e.T0[1] - 2, e.T0[2] - (D == 3 ? 2 : 0)},
                           tg.J[0], tg.J[1], pvol};
    if constexpr (D == 2) {
      R ap[6];
      ap[5] = contract3l<2, R>(JE, GX, GXY, ax, ay, az, i, j, 0);
      ap[0] = contract3l<2, R>(EB, GX, GXY, axh, ay, az, ih, j, 0);
      ap[3] = contract3l<2, R>(JE + gvol, GX, GXY, axh, ay, az, ih, j, 0);
      ap[1] = contract3l<2, R>(EB + gvol, GX, GXY, ax, ayh, az, i, jh, 0);
      ap[4] = contract3l<2, R>(JE + 2 * gvol, GX, GXY, ax, ayh, az, i, jh, 0);
      ap[2] = contract3l<2, R>(EB + 2 * gvol, GX, GXY, axh, ayh, az, ih, jh, 0);
      const R u1 = q.p[0], u2 = q.p[1];
      const R gam2 = one + u1 * u1 + u2 * u2 + ap[5];
      for (int c = 0; c < 3; c++) ap[c] = qch * ap[c];
      for (int c = 3; c < 5; c++) ap[c] = half * qch * qch * ap[c];
      const R aa1 = dth * (ap[0] * u1 + ap[1] * u2);
      const R b1 = dth * (ap[3] * u1 + ap[4] * u2);
      const R gam = sqrt(gam2);
      const R dgam = (aa1 * gam - b1) / gam2;
      const R gam_inv = (gam - dgam) / gam2;
      for (int c = 2; c < 5; c++) ap[c] = ap[c] * gam_inv;
      const R e1 = alp * (ap[0] - ap[3]), e2 = alp * (ap[1] - ap[4]), bb1 = alp * ap[2];
      const R vp0 = u1 + e1, vp1 = u2 + e2;
      const R b2 = one + bb1 * bb1;
      const R vh0 = (vp0 + vp1 * bb1) / b2, vh1 = (vp1 - vp0 * bb1) / b2;
      q.p[0] = R(2) * vh0 - u1;
      q.p[1] = R(2) * vh1 - u2;
      const R wghp = (R)q.w * gam_inv;
      for (int j1 = 0; j1 < 3; j1++) {
        const R dvol = ay[j1] * wghp;
        for (int i1 = 0; i1 < 3; i1++)
          acc(0, q.N[0] - 1 + i1, q.N[1] - 1 + j1, 1, dvol * ax[i1]);
      }
    } else {
      R ap[10];
      ap[9] = contract3l<3, R>(JE, GX, GXY, ax, ay, az, i, j, k);
      ap[0] = contract3l<3, R>(EB, GX, GXY, axh, ay, az, ih, j, k);
      ap[6] = contract3l<3, R>(JE + gvol, GX, GXY, axh, ay, az, ih, j, k);
      ap[1] = contract3l<3, R>(EB + gvol, GX, GXY, ax, ayh, az, i, jh, k);
      ap[7] = contract3l<3, R>(JE + 2 * gvol, GX, GXY, ax, ayh, az, i, jh, k);
      ap[5] = contract3l<3, R>(EB + 5 * gvol, GX, GXY, axh, ayh, az, ih, jh, k);
      ap[3] = contract3l<3, R>(EB + 3 * gvol, GX, GXY, ax, ayh, azh, i, jh, kh);
      ap[4] = contract3l<3, R>(EB + 4 * gvol, GX, GXY, axh, ay, azh, ih, j, kh);
      ap[2] = contract3l<3, R>(EB + 2 * gvol, GX, GXY, ax, ay, azh, i, j, kh);
      ap[8] = contract3l<3, R>(JE + 3 * gvol, GX, GXY, ax, ay, azh, i, j, kh);
      const R u[3] = {q.p[0], q.p[1], q.p[2]};
      const R gam2 = one + u[0] * u[0] + u[1] * u[1] + u[2] * u[2] + ap[9];
      for (int c = 0; c < 6; c++) ap[c] = qch * ap[c];
      for (int c = 6; c < 9; c++) ap[c] = half * qch * qch * ap[c];
      const R aa1 = dth * (ap[0] * u[0] + ap[1] * u[1] + ap[2] * u[2]);
      const R b1 = dth * (ap[6] * u[0] + ap[7] * u[1] + ap[8] * u[2]);
      const R gam = sqrt(gam2);
      const R dgam = (aa1 * gam - b1) / gam2;
      const R gam_inv = (gam - dgam) / gam2;
      for (int c = 3; c < 9; c++) ap[c] = ap[c] * gam_inv;
      R vp[3], bb[3], vph[3];
      for (int c = 0; c < 3; c++) vp[c] = alp * (ap[c] - ap[c + 6]) + u[c];
      for (int c = 0; c < 3; c++) bb[c] = alp * ap[c + 3];
      const R b2 = one + (bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2]);
      const R bv = bb[0] * vp[0] + bb[1] * vp[1] + bb[2] * vp[2];
      vph[0] = vp[0] + vp[1] * bb[2] - vp[2] * bb[1] + bb[0] * bv;
      vph[1] = vp[1] + vp[2] * bb[0] - vp[0] * bb[2] + bb[1] * bv;
      vph[2] = vp[2] + vp[0] * bb[1] - vp[1] * bb[0] + bb[2] * bv;
      for (int c = 0; c < 3; c++) q.p[c] = R(2) * (vph[c] / b2) - u[c];
      const R wghp = (R)q.w * gam_inv;
      for (int k1 = 0; k1 < 3; k1++)
        for (int j1 = 0; j1 < 3; j1++) {
          const R dvol = ay[j1] * az[k1] * wghp;
          for (int i1 = 0; i1 < 3; i1++)
            acc(0, q.N[0] - 1 + i1, q.N[1] - 1 + j1, q.N[2] - 1 + k1, dvol * ax[i1]);
        }
    }
    put_rec<D, R>(A, src, pack_cell<D>(q.N, tg.bx, tg.by), q.s, q.p, q.w);
  }
  __syncthreads();
  env_patch_out<D>(tg, t, Pd, 1, patch);
}

// Phase 2 per tile: set_env_grad_interp + lpf_env_positions + removal +
// Esirkepov deposit into the tile's LDS current patch + re-binning into B.
template <int D, typename R>
__global__ __launch_bounds__(kEB) void k_env_move_tile(TileGeom tg, DevGrid g, StepConst<R> sc,
                                                       RecPtr A, RecPtr B, const int* seg_start,
                                                       const int* seg_cap, const int* cnt,
                                                       const R* je, int* cnt_next, int* ovf_next,
                                                       double* patch, int* grow) {
  extern __shared__ double lds_d[];
  const int t = blockIdx.x;
  const EnvTile e = env_tile<D>(tg, t);
  double* Jl = lds_d;                      // D current components, pvol each
  R* JE = (R*)(lds_d + D * tg.pvol);       // je(1:D+1)
  for (int l = threadIdx.x; l < D * tg.pvol; l += blockDim.x) Jl[l] = 0.0;
  env_stage_rows<D, R>(g, e, je, D + 1, je, 0, JE);
  __syncthreads();
  const int GX = e.G[0], GXY = e.G[0] * e.G[1];
  const EnvLdsAcc<D> acc{Jl, {e.T0[0] - 2, e.T0[1] - 2, e.T0[2] - (D == 3 ? 2 : 0)},
                         tg.J[0], tg.J[1], tg.pvol};
  const int n = min(cnt[t], seg_cap[t]);
  const int64_t base = seg_start[t];
  for (int q0 = threadIdx.x; q0 < n; q0 += blockDim.x) {
    const int64_t src = base + q0;
    if (rec_cell<R>(A, src) == kHole) continue;
    Rec<R> q;
    load_rec<D, R>(A, src, q, tg.bx, tg.by);
    R ax[3], axh[3], ay[3], ayh[3], az[3] = {R(1), R(0), R(0)}, azh[3] = {R(1), R(0), R(0)};
    int hx, hy, hz = 0;
    qqh<R>(q.s[0], ax, axh, hx);
    qqh<R>(q.s[1], ay, ayh, hy);
    if (D == 3) qqh<R>(q.s[2], az, azh, hz);
    const int i = q.N[0] - e.T0[0] + 1, j = q.N[1] - e.T0[1] + 1, k = D == 3 ? q.N[2] - e.T0[2] + 1 : 0;
    const int ih = i + hx, jh = j + hy, kh = D == 3 ? k + hz : 0;
    R ap[4] = {R(0), R(0), R(0), R(0)};
    ap[0] = contract3l<D, R>(JE + e.gvol, GX, GXY, axh, ay, az, ih, j, k);
    ap[D] = contract3l<D, R>(JE, GX, GXY, ax, ay, az, i, j, k);
    ap[1] = contract3l<D, R>(JE + 2 * e.gvol, GX, GXY, ax, ayh, az, i, jh, k);
    if (D == 3) ap[2] = contract3l<D, R>(JE + 3 * e.gvol, GX, GXY, ax, ay, azh, i, j, kh);
    const R one = R(1);
    R pp2 = R(0), pv = R(0);
#pragma unroll
    for (int d = 0; d < D; d++) {
      pp2 = pp2 + q.p[d] * q.p[d];
      pv = pv + q.p[d] * ap[d];
    }
    const R gam2 = one + pp2 + ap[D];
    const R gam = sqrt(gam2);
    const R gam3 = gam2 * gam;
    const R b2 = R(0.25) * pv;
    const R gam_inv = (one / gam) * (one - sc.dt * b2 / gam3);
    const int io[3] = {q.N[0], q.N[1], q.N[2]};
    const R so[3] = {q.s[0], q.s[1], q.s[2]};
    if (!move_and_test<D, R>(q, sc.dt * gam_inv, sc.inv, sc.nn)) continue;
    esirkepov<D, R>(io, so, q.N, q.s, (R)((float)sc.qch * q.w), acc);
    const int dtile = tile_of(tg, pack_cell<D>(q.N, tg.bx, tg.by), D);
    const int slot = reserve_slot(dtile, cnt_next);
    int64_t dst = (int64_t)seg_start[dtile] + slot;
    if (slot >= seg_cap[dtile]) {
      dst = sc.ovf_base + atomicAdd(ovf_next, 1);
      grow[dtile] = 1;
    }
    store_rec<D, R>(B, dst, q, tg.bx, tg.by);
  }
  __syncthreads();
  env_patch_out<D>(tg, t, Jl, D, patch);
}

// ------------------------------------------------------------ grid side
// je(1) on the interior: |A^n|^2/2 (env_amp_prepare,
// env_evolve_in_time.f90:145-170) or |(A^{n+1}+A^n)/2|^2/2
// (env_fields_average, :117-143).
template <typename R>
__global__ void k_env_amp(DevGrid g, const R* env, R* je, int average) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + g.ix1;
  const int j = blockIdx.y + g.jy1, k = blockIdx.z + g.kz1;
  if (i > g.ix2) return;
  const int64_t o = fofs(g, i, j, k);
  const R a1 = env[o], a2 = env[g.ncell + o];
  if (average) {
    const R r = a1 + env[2 * g.ncell + o], m = a2 + env[3 * g.ncell + o];
    je[o] = R(1.0 / 8.) * (r * r + m * m);
  } else {
    je[o] = R(0.5) * (a1 * a1 + a2 * a2);
  }
}

// env_grad (grid_fields.f90:426-586) with every face physical and an
// unstretched grid: je(2:D+1) = the staggered gradient of F = je(1) on the
// interior, one-sided at the faces; in 3-D the ghost planes kz1-1, kz2+1 of
// F are extrapolated linearly (the threads of the face planes write them;
// the gradient reads its own copy of the value).
template <typename R>
__global__ void k_env_grad(DevGrid g, EnvConst ec, R* je) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + g.ix1;
  const int j = blockIdx.y + g.jy1, k = blockIdx.z + g.kz1;
  if (i > g.ix2) return;
  const R* F = je;
  const int64_t o = fofs(g, i, j, k);
  const int64_t sy = g.n1p, sz = (int64_t)g.n1p * g.n2p;
  const bool d4 = ec.der_ord == 4;
  const R a_hcd = R(13. / 12.), b_hcd = R(-1. / 24.);
  const R dxi = (R)g.dx_inv, dyi = (R)g.dy_inv, dzi = (R)g.dz_inv;
  // x: faces ix1, ix2-1 (forward difference), ix2 (one-sided)
  R gx;
  if (i == g.ix1 || i == g.ix2 - 1) {
    gx = dxi * (F[o + 1] - F[o]);
  } else if (i == g.ix2) {
    gx = dxi * (R(2) * F[o - 1] - R(3) * F[o - 2] + F[o - 3]);
  } else {
    gx = (d4 ? dxi * a_hcd : dxi) * (F[o + 1] - F[o]);
    if (d4) gx = gx + dxi * b_hcd * (F[o + 2] - F[o + 1] + F[o] - F[o - 1]);
  }
  je[g.ncell + o] = gx;
  // y: faces jy2 (one-sided), jy2-1 and jy1 (forward difference)
  R gy;
  if (j == g.jy2) {
    gy = dyi * (R(2) * F[o] - R(3) * F[o - sy] + F[o - 2 * sy]);
  } else if (j == g.jy2 - 1 || j == g.jy1) {
    gy = dyi * (F[o + sy] - F[o]);
  } else {
    gy = (d4 ? dyi * a_hcd : dyi) * (F[o + sy] - F[o]);
    if (d4) {
      const R s2 = dyi * b_hcd;
      gy = gy + s2 * (F[o + 2 * sy] - F[o + sy]) + s2 * (F[o] - F[o - sy]);
    }
  }
  je[2 * g.ncell + o] = gy;
  if (g.ndim < 3) return;
  // z: faces kz2 and kz1 extrapolate F into the ghost plane, interior
  // forward differences up to kz2-1
  const R fu = k == g.kz2 ? R(2) * F[o] - F[o - sz] : F[o + sz];
  if (k == g.kz2) je[o + sz] = fu;
  if (k == g.kz1) je[o - sz] = R(2) * F[o] - F[o + sz];
  R gz;
  if (k == g.kz2 || k == g.kz1) {
    gz = dzi * (fu - F[o]);
  } else {
    gz = (d4 ? dzi * a_hcd : dzi) * (fu - F[o]);
    if (d4) {
      const R s2 = dzi * b_hcd;
      const R fuu = k + 1 == g.kz2 ? R(2) * F[o + sz] - F[o] : F[o + 2 * sz];
      gz = gz + s2 * (fuu - fu) + s2 * (F[o] - F[o - sz]);
    }
  }
  je[3 * g.ncell + o] = gz;
}

// den_zyxbd (mpi_curr_interface.f90:646-720), open faces, one pass per axis
// (axis 2, then 1, then 0): the first ghost layer is added to the second
// interior layer, which is then copied onto the face layer.
__global__ void k_den_fold(DevGrid g, double* rho, int axis) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  const int side = blockIdx.z;  // 0: low face, 1: high face
  int i, j, k;
  int64_t st;
  if (axis == 2) {
    i = g.ix1 + a;
    j = g.jy1 + b;
    if (i > g.ix2 || j > g.jy2) return;
    k = side ? g.kz2 : g.kz1;
    st = (int64_t)g.n1p * g.n2p;
  } else if (axis == 1) {
    i = g.ix1 + a;
    k = g.kz1 + b;
    if (i > g.ix2 || k > g.kz2) return;
    j = side ? g.jy2 : g.jy1;
    st = g.n1p;
  } else {
    j = g.jy1 + a;
    k = g.kz1 + b;
    if (j > g.jy2 || k > g.kz2) return;
    i = side ? g.ix2 : g.ix1;
    st = 1;
  }
  const int64_t o = fofs(g, i, j, k);
  const int64_t in = side ? o - st : o + st, gh = side ? o + st : o - st;
  rho[in] = rho[in] + rho[gh];
  rho[o] = rho[in];
}

// je(1) = je(3) = den on every node (set_env_density's output into jc(1),
// then jc(3) = jc(1), env_evolve_in_time.f90:304-312)
template <typename R>
__global__ void k_env_den_copy(DevGrid g, const double* den, R* je) {
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < g.ncell;
       n += (int64_t)gridDim.x * blockDim.x) {
    const R v = (R)den[n];
    je[n] = v;
    je[2 * g.ncell + n] = v;
  }
}

// env_bds(evf, 2, 2) (grid_fields.f90:864-1067), open faces, unstretched, on
// components 1..2 of env: pass 0 copies the edge value into the two x ghost
// layers; pass 1 (y) and pass 2 (z) put 3 f(1) - 3 f(2) + f(3), written as
// (f(3) - f(2)) - 2 f(2) + 3 f(1), into both ghost layers of the range the
// earlier passes extended.
template <typename R>
__global__ void k_env_bds(DevGrid g, R* env, int pass) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  const int c = blockIdx.z >> 2, gsel = blockIdx.z & 3;  // component, ghost layer (4 per comp)
  R* ef = env + (int64_t)c * g.ncell;
  const int side = gsel >> 1, lay = (gsel & 1) + 1;  // 1 or 2 layers out
  if (pass == 0) {
    const int j = g.jy1 + a, k = g.kz1 + b;
    if (j > g.jy2 || k > g.kz2) return;
    const int ie = side ? g.ix2 : g.ix1;
    ef[fofs(g, side ? ie + lay : ie - lay, j, k)] = ef[fofs(g, ie, j, k)];
  } else if (pass == 1) {
    const int i = g.ix1 - 2 + a, k = g.kz1 + b;
    if (i > g.ix2 + 2 || k > g.kz2) return;
    const int je = side ? g.jy2 : g.jy1, s = side ? -1 : 1;
    const R f1 = ef[fofs(g, i, je, k)], f2 = ef[fofs(g, i, je + s, k)],
            f3 = ef[fofs(g, i, je + 2 * s, k)];
    ef[fofs(g, i, side ? je + lay : je - lay, k)] = (f3 - f2) - R(2) * f2 + R(3) * f1;
  } else {
    const int i = g.ix1 - 2 + a, j = g.jy1 - 2 + b;
    if (i > g.ix2 + 2 || j > g.jy2 + 2) return;
    const int ke = side ? g.kz2 : g.kz1, s = side ? -1 : 1;
    const R f1 = ef[fofs(g, i, j, ke)], f2 = ef[fofs(g, i, j, ke + s)],
            f3 = ef[fofs(g, i, j, ke + 2 * s)];
    ef[fofs(g, i, j, side ? ke + lay : ke - lay)] = (f3 - f2) - R(2) * f2 + R(3) * f1;
  }
}

// advance_lpf_envelope's source and env_maxw_solve's explicit step
// (curr_and_fields_util.f90:262-299, grid_fields.f90:588-694) on the
// interior, written into je(1:2): S = -ompe chi A^n + Laplacian A^n
// (potential_lapl: x, then the transverse pp_lapl, grid_fields.f90:326-424,
// 1356-1403) + the first-derivative term of 2 i om0 D_x, then
// F = dt^2 S + 2 A^n - A^{n-1} -/+ kfact A^{n-1}_{I/R}.
template <typename R>
__global__ void k_env_source(DevGrid g, EnvConst ec, const R* env, R* je) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + g.ix1;
  const int j = blockIdx.y + g.jy1, k = blockIdx.z + g.kz1;
  if (i > g.ix2) return;
  const int64_t o = fofs(g, i, j, k), nc = g.ncell;
  const int64_t sy = g.n1p, sz = (int64_t)g.n1p * g.n2p;
  const R chi = je[2 * nc + o];
  R cur[2];
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const R* A = env + c * nc + o;
    R v = (-(R)g.ompe * chi) * A[0];
    v = v + (R)ec.dx4[0] * (A[1] + A[-1] - R(2) * A[0]);
    if (ec.der_ord > 2) v = v + (R)ec.dx4[1] * (A[2] + A[-2] - R(2) * A[0]);
    v = v + (R)ec.dy4[0] * ((A[sy] - A[0]) - (A[0] - A[-sy]));
    if (ec.der_ord == 4) v = v + (R)ec.dy4[1] * ((A[2 * sy] - A[0]) - (A[0] - A[-2 * sy]));
    if (g.ndim == 3) {
      v = v + (R)ec.dz4[0] * ((A[sz] - A[0]) - (A[0] - A[-sz]));
      if (ec.der_ord == 4) v = v + (R)ec.dz4[1] * ((A[2 * sz] - A[0]) - (A[0] - A[-2 * sz]));
    }
    cur[c] = v;
  }
  const R* A1 = env + o;
  const R* A2 = env + nc + o;
  if (ec.der_ord < 3) {
    const R d1 = (R)ec.dx1_inv;
    cur[0] = cur[0] - d1 * (A2[1] - A2[-1]);
    cur[1] = cur[1] + d1 * (A1[1] - A1[-1]);
  } else {
    const R p0 = (R)ec.aph[0], p1 = (R)ec.aph[1];
    cur[0] = cur[0] - p0 * (A2[1] - A2[-1]) - p1 * (A2[2] - A2[-2]);
    cur[1] = cur[1] + p0 * (A1[1] - A1[-1]) + p1 * (A1[2] - A1[-2]);
  }
  const R dt2 = (R)ec.dt2, kf = (R)ec.kfact;
  const R a3 = env[2 * nc + o], a4 = env[3 * nc + o];
  je[o] = dt2 * cur[0] + R(2) * A1[0] - a3 + kf * a4;
  je[nc + o] = dt2 * cur[1] + R(2) * A2[0] - a4 - kf * a3;
}

// A^{n-1} <- A^n, A^{n+1} = (F -/+ kfact F_{I/R})/(1 + kfact^2)
// (grid_fields.f90:660-669)
template <typename R>
__global__ void k_env_update(DevGrid g, EnvConst ec, R* env, const R* je) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + g.ix1;
  const int j = blockIdx.y + g.jy1, k = blockIdx.z + g.kz1;
  if (i > g.ix2) return;
  const int64_t o = fofs(g, i, j, k), nc = g.ncell;
  const R kf = (R)ec.kfact, k2 = (R)ec.k2_fact;
  const R f1 = je[o], f2 = je[nc + o];
  env[2 * nc + o] = env[o];
  env[3 * nc + o] = env[nc + o];
  env[o] = k2 * (f1 - kf * f2);
  env[nc + o] = k2 * (f2 + kf * f1);
}

// ------------------------------------------------------------ launchers
EnvConst env_const(const DevGrid& g, int der_ord, double om0, double cfl) {
  EnvConst e{};
  e.der_ord = der_ord;
  // set_field_param (grid_field_param.f90:36-99): nu, hord_der2, opt_der1
  const double yx_rat = g.dy / g.dx;
  const double nu = cfl * yx_rat / std::sqrt(yx_rat * yx_rat + (double)(float)g.ndim - 1.);
  double hord_der2 = 0.0, opt_der1 = 1.0;
  if (der_ord == 3) {
    hord_der2 = -(1. - nu * nu) / 12.;
    opt_der1 = (4. - nu * nu) / 3.;
  } else if (der_ord == 4) {
    hord_der2 = -1. / 12.;
  }
  // potential_lapl's x coefficients (grid_fields.f90:1356-1403) and
  // pp_lapl's transverse ones (:326-424)
  const double dx2 = g.dx_inv * g.dx_inv;
  e.dx4[0] = der_ord > 2 ? (1. - 4. * hord_der2) * dx2 : dx2;
  e.dx4[1] = der_ord > 2 ? hord_der2 * dx2 : 0.0;
  const double cf0 = der_ord == 4 ? 4. / 3. : 1., cf1 = der_ord == 4 ? -1. / 12. : 0.;
  const double dy2 = g.dy_inv * g.dy_inv, dz2 = g.dz_inv * g.dz_inv;
  e.dy4[0] = cf0 * dy2;
  e.dy4[1] = cf1 * dy2;
  e.dz4[0] = cf0 * dz2;
  e.dz4[1] = cf1 * dz2;
  // env_maxw_solve (grid_fields.f90:588-620)
  const double dtl = g.dt;
  e.dt2 = dtl * dtl;
  e.kfact = om0 * dtl;
  e.k2_fact = 1. / (1. + e.kfact * e.kfact);
  e.dx1_inv = om0 * g.dx_inv;
  e.aph[0] = 1.;
  e.aph[1] = 0.;
  if (der_ord == 3) {
    e.aph[0] = e.dx1_inv * opt_der1;
    e.aph[1] = e.dx1_inv * 0.5 * (1. - opt_der1);
  }
  return e;
}

template <typename R>
hipError_t env_grid_advance(const DevGrid& g, const EnvConst& ec, R* env, R* je, double* den,
                            hipStream_t s) {
  const dim3 bi((g.nx + 63) / 64, g.ny, g.ndim == 3 ? g.nz : 1);
  // den_zyxbd, then jc(1) = jc(3) = den
  if (g.ndim == 3) k_den_fold<<<dim3((g.nx + 63) / 64, g.ny, 2), 64, 0, s>>>(g, den, 2);
  k_den_fold<<<dim3((g.nx + 63) / 64, g.ndim == 3 ? g.nz : 1, 2), 64, 0, s>>>(g, den, 1);
  k_den_fold<<<dim3((g.ny + 63) / 64, g.ndim == 3 ? g.nz : 1, 2), 64, 0, s>>>(g, den, 0);
  k_env_den_copy<R><<<blocks_for(g.ncell), kEB, 0, s>>>(g, den, je);
  // advance_lpf_envelope: env_bds, source + solve, update
  k_env_bds<R><<<dim3((g.ny + 63) / 64, g.ndim == 3 ? g.nz : 1, 8), 64, 0, s>>>(g, env, 0);
  k_env_bds<R><<<dim3((g.nx + 4 + 63) / 64, g.ndim == 3 ? g.nz : 1, 8), 64, 0, s>>>(g, env, 1);
  if (g.ndim == 3)
    k_env_bds<R><<<dim3((g.nx + 4 + 63) / 64, g.ny + 4, 8), 64, 0, s>>>(g, env, 2);
  k_env_source<R><<<bi, 64, 0, s>>>(g, ec, env, je);
  k_env_update<R><<<bi, 64, 0, s>>>(g, ec, env, je);
  // env_fields_average
  k_env_amp<R><<<bi, 64, 0, s>>>(g, env, je, 1);
  k_env_grad<R><<<bi, 64, 0, s>>>(g, ec, je);
  return hipGetLastError();
}

template <typename R>
hipError_t env_amp_prepare(const DevGrid& g, const EnvConst& ec, const R* env, R* je,
                           hipStream_t s) {
  const dim3 bi((g.nx + 63) / 64, g.ny, g.ndim == 3 ? g.nz : 1);
  fill<R>(je, g.ncell * (g.curr_ndim + 1), R(0), s);
  k_env_amp<R><<<bi, 64, 0, s>>>(g, env, je, 0);
  k_env_grad<R><<<bi, 64, 0, s>>>(g, ec, je);
  return hipGetLastError();
}

// LDS bytes of the tile workgroups
template <typename R>
static size_t env_push_lds(const TileGeom& tg, const DevGrid& g) {
  const int G = (tg.T[0] + 3) * (tg.T[1] + 3) * (g.ndim == 3 ? tg.T[2] + 3 : 1);
  return (size_t)tg.pvol * sizeof(double) + (size_t)(g.nfield + g.ndim + 1) * G * sizeof(R);
}
template <typename R>
static size_t env_move_lds(const TileGeom& tg, const DevGrid& g) {
  const int G = (tg.T[0] + 3) * (tg.T[1] + 3) * (g.ndim == 3 ? tg.T[2] + 3 : 1);
  return (size_t)g.ndim * tg.pvol * sizeof(double) + (size_t)(g.ndim + 1) * G * sizeof(R);
}

template <typename R>
void env_particles_push(TileStore* ts, const TileGeom& tg, const DevGrid& g, const R* ebf,
                        const R* je, double* den, double* patch, double lorentz_fact,
                        hipStream_t s, hipEvent_t ev_begin, hipEvent_t ev_end) {
  if (!ts->loaded) {
    fill<double>(den, g.ncell, 0.0, s);
    return;
  }
  if (ev_begin) (void)hipEventRecord(ev_begin, s);
  const int cur = ts->cur;
  const RecPtr& A = ts->buf[cur];
  const R qch = (R)ts->charge, dth = (R)(0.5 * g.dt), alp = (R)(0.5 * g.dt * lorentz_fact);
  const size_t lds = env_push_lds<R>(tg, g);
  // tiles: density into the patches, folded onto den; then the overflow
  // list adds to den directly
  if (g.ndim == 2)
    k_env_push_tile<2, R><<<tg.ntiles, kEB, lds, s>>>(tg, g, A, ts->seg_start, ts->seg_cap,
                                                      ts->count[cur], ebf, je, qch, dth, alp,
                                                      patch);
  else
    k_env_push_tile<3, R><<<tg.ntiles, kEB, lds, s>>>(tg, g, A, ts->seg_start, ts->seg_cap,
                                                      ts->count[cur], ebf, je, qch, dth, alp,
                                                      patch);
  tile_store_fold_density(tg, g, patch, den, s);
  if (g.ndim == 2)
    k_env_push<2, R><<<64, kEB, 0, s>>>(tg, g, A, ts->ovf_base, ts->ovf_count + cur, ebf, je,
                                        qch, dth, alp, den);
  else
    k_env_push<3, R><<<64, kEB, 0, s>>>(tg, g, A, ts->ovf_base, ts->ovf_count + cur, ebf, je,
                                        qch, dth, alp, den);
  if (ev_end) (void)hipEventRecord(ev_end, s);
}

template <typename R>
int env_particles_move(TileStore* ts, const TileGeom& tg, const DevGrid& g, const R* je, R* jc,
                       double* patch, hipStream_t s, std::string* msg, hipEvent_t ev_begin,
                       hipEvent_t ev_end) {
  const int cur = ts->cur, nxt = 1 - cur;
  if (ts->loaded) {
    StepConst<R> sc{};
    sc.qch = (R)ts->charge;
    sc.dt = (R)g.dt;
    sc.inv[0] = (R)g.dx_inv;
    sc.inv[1] = (R)g.dy_inv;
    sc.inv[2] = (R)g.dz_inv;
    sc.nn[0] = g.nx;
    sc.nn[1] = g.ny;
    sc.nn[2] = g.nz;
    sc.ovf_base = ts->ovf_base;
    tile_store_clear_next(ts, tg, s);
    const size_t lds = env_move_lds<R>(tg, g);
    if (ev_begin) (void)hipEventRecord(ev_begin, s);
    // every tile writes its whole patch; the overflow list adds to them.
    // The tile kernel of the LP step with the envelope's advance (cell-sorted
    // chunks, register sums of the current) where it is compiled for the
    // workgroup size, else the simpler tile workgroups above.
    const hipError_t tk = launch_env_move_kernel<R>(
        tg, g, sc, je, ts->buf[cur], ts->buf[nxt], ts->seg_start, ts->seg_cap, ts->count[cur],
        ts->count[nxt], ts->ovf_count + cur, ts->ovf_count + nxt, patch, ts->grow, s);
    if (tk != hipSuccess && tk != hipErrorInvalidConfiguration) {
      *msg = "envelope tile kernel launch failed";
      return ALADYN_EHIP;
    }
    const bool tiled = tk == hipSuccess;
    if (g.ndim == 2) {
      if (!tiled)
        k_env_move_tile<2, R><<<tg.ntiles, kEB, lds, s>>>(
            tg, g, sc, ts->buf[cur], ts->buf[nxt], ts->seg_start, ts->seg_cap, ts->count[cur],
            je, ts->count[nxt], ts->ovf_count + nxt, patch, ts->grow);
      k_env_move<2, R><<<64, kEB, 0, s>>>(tg, g, sc, ts->buf[cur], ts->buf[nxt], ts->seg_start,
                                          ts->seg_cap, ts->ovf_count + cur, je, ts->count[nxt],
                                          ts->ovf_count + nxt, patch, ts->grow);
    } else {
      if (!tiled)
        k_env_move_tile<3, R><<<tg.ntiles, kEB, lds, s>>>(
            tg, g, sc, ts->buf[cur], ts->buf[nxt], ts->seg_start, ts->seg_cap, ts->count[cur],
            je, ts->count[nxt], ts->ovf_count + nxt, patch, ts->grow);
      k_env_move<3, R><<<64, kEB, 0, s>>>(tg, g, sc, ts->buf[cur], ts->buf[nxt], ts->seg_start,
                                          ts->seg_cap, ts->ovf_count + cur, je, ts->count[nxt],
                                          ts->ovf_count + nxt, patch, ts->grow);
    }
    if (ev_end) (void)hipEventRecord(ev_end, s);
    if (hipGetLastError() != hipSuccess) {
      *msg = "envelope particle kernel launch failed";
      return ALADYN_EHIP;
    }
  } else {
    fill<double>(patch, (int64_t)tg.ntiles * tg.pvol * g.curr_ndim, 0.0, s);
  }
  tile_store_fold<R>(tg, g, patch, jc, 0, s, true);
  if (ts->loaded) {
    if (hipMemcpyAsync(ts->h_ovf + ts->ovf_slot, ts->ovf_count + nxt, sizeof(int),
                       hipMemcpyDeviceToHost, s) != hipSuccess) {
      *msg = "envelope step: overflow count copy failed";
      return ALADYN_EHIP;
    }
    return tile_store_step_finish(ts, tg, g, s, msg);
  }
  return 0;
}

#define ALADYN_ENV_INST(R)                                                                       \
  template hipError_t env_grid_advance<R>(const DevGrid&, const EnvConst&, R*, R*, double*,     \
                                          hipStream_t);                                          \
  template hipError_t env_amp_prepare<R>(const DevGrid&, const EnvConst&, const R*, R*,         \
                                         hipStream_t);                                           \
  template void env_particles_push<R>(TileStore*, const TileGeom&, const DevGrid&, const R*,    \
                                      const R*, double*, double*, double, hipStream_t,           \
                                      hipEvent_t, hipEvent_t);                                   \
  template int env_particles_move<R>(TileStore*, const TileGeom&, const DevGrid&, const R*, R*, \
                                     double*, hipStream_t, std::string*, hipEvent_t, hipEvent_t);
ALADYN_ENV_INST(float)
ALADYN_ENV_INST(double)
#undef ALADYN_ENV_INST

}  // namespace aladyn
