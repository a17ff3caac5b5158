// tile_kernel.hip -- the particle kernel of the tile path (FP32 and FP64).
//
// One workgroup per tile of TX x TY x TZ node-cells, one thread per
// node-cell (the tile shape is a template parameter, so LDS offsets inside a
// stencil are immediates; R is float or double).  Per workgroup:
//   1. stage the field patch (nodes T0-1 .. T0+T per axis, all components)
//      from HBM into LDS and zero an FP64 LDS current patch (nodes
//      T0-2 .. T0+T+1, the reach of the 5-slot Esirkepov window of
//      grid_part_connect.f90:1008);
//   2. counting-sort the tile's particles by cell in LDS (chunks of kChunk);
//   3. split the sorted chunk into equal consecutive slices, one per thread,
//      and walk them in lock step: gather from LDS, charge multiply,
//      [back-shift], exact-gamma push, carry, box test; a particle whose
//      nearest node moved on at most one axis adds the part of its Esirkepov
//      current on its old cell's node block to register sums of that cell
//      (flushed into the LDS patch when the thread's slice moves to the next
//      cell); the layer beyond the block of one whose node moved on one axis
//      waits in a per-wave edge queue (64 entries) and is added, one particle
//      per lane, when the next round's entries would not fit; one whose node
//      moved on two or three axes is queued likewise and deposited through
//      the union window; survivors are written to the
//      other buffer: stayers into slots reserved per chunk, movers into
//      their new tile with wave-aggregated atomics, overflow into the global
//      list;
//   4. write the LDS current patch to HBM for the fold.
// LDS accumulation is FP64 because LDS FP32 atomic adds are far slower than
// FP64 ones on gfx950; there are no global float atomics.
#include <type_traits>

#include "tile_device.hpp"


namespace aladyn {

// Union-window Esirkepov deposit (grid_part_connect.f90:979-1060, 1168-1334)
// for a particle whose nearest node changed on two or three axes.  Per axis
// the old and new 3-node windows merge into 4 slots starting at
// min(N_old, N_new) - 1; the 3 fluxes between them are cumulative sums of
// W0 - W1 (the fourth vanishes by the partition of unity).  jb points at the
// LDS patch node of the union base; offsets are compile-time.
template <int D, int JX, int JY, int PVOL, typename R>
__device__ inline void deposit_general(double* jb, const R s0[3], const R s1[3],
                                       const int o0[3], const int o1[3], R wq) {
  const R zero = R(0), half = R(0.5);
  R W0[3][4], W1[3][4];
#pragma unroll
  for (int d = 0; d < D; d++) {
    R a0[3], a1[3];
    qweights<R>(s0[d], a0);
    qweights<R>(s1[d], a1);
#pragma unroll
    for (int m = 0; m < 4; m++) {
      W0[d][m] = o0[d] == 0 ? (m < 3 ? a0[m] : zero) : (m > 0 ? a0[m - 1] : zero);
      W1[d][m] = o1[d] == 0 ? (m < 3 ? a1[m] : zero) : (m > 0 ? a1[m - 1] : zero);
    }
  }
  R F[3][3];
#pragma unroll
  for (int d = 0; d < D; d++) {
    R c = zero;
#pragma unroll
    for (int f = 0; f < 3; f++) {
      c = c + W0[d][f] - W1[d][f];
      F[d][f] = wq * c;
    }
  }
  if constexpr (D == 2) {
#pragma unroll
    for (int f = 0; f < 3; f++)
#pragma unroll
      for (int m = 0; m < 4; m++) {
        atomicAdd(jb + (f + JX * m), (double)(F[0][f] * (W0[1][m] + W1[1][m])));
        atomicAdd(jb + PVOL + (m + JX * f), (double)(F[1][f] * (W0[0][m] + W1[0][m])));
      }
  } else {
    // J_x: flux_x (x) [W0z (W0y + W1y/2) + W1z (W1y + W0y/2)], etc.
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const R y0 = W0[1][m] + half * W1[1][m], y1 = W1[1][m] + half * W0[1][m];
      const R x0 = W0[0][m] + half * W1[0][m], x1 = W1[0][m] + half * W0[0][m];
#pragma unroll
      for (int l = 0; l < 4; l++) {
        const R tyz = y0 * W0[2][l] + y1 * W1[2][l];
        const R txz = x0 * W0[2][l] + x1 * W1[2][l];
        const R txy = x0 * W0[1][l] + x1 * W1[1][l];
#pragma unroll
        for (int f = 0; f < 3; f++) {
          atomicAdd(jb + (f + JX * (m + JY * l)), (double)(F[0][f] * tyz));
          atomicAdd(jb + PVOL + (m + JX * (f + JY * l)), (double)(F[1][f] * txz));
          atomicAdd(jb + 2 * PVOL + (m + JX * (l + JY * f)), (double)(F[2][f] * txy));
        }
      }
    }
  }
}

// Window data of one axis for a particle whose nearest node moved by sg
// (-1, 0 or +1) from N0, in the frame of the old node block N0-1 .. N0+1:
// a0 = old quadratic weights (given), w1 = the new weights that fall inside the
// block, fl = the two in-block Esirkepov fluxes (N0-1 -> N0, N0 -> N0+1),
// and for sg != 0 the new weight wo and the flux fo just outside the block
// (at N0+2 / N0+1 -> N0+2 for sg = +1, at N0-2 / N0-2 -> N0-1 for sg = -1).
// The fluxes are the cumulative sums of W0 - W1 of grid_part_connect.f90
// (1008-1030 and 1200-1240); for sg = 0 the two nonzero ones are written in
// closed form, (s1-s0)(1-s0-s1)/2 and (s1-s0)(1+s0+s1)/2 times wq, which
// avoids the cancellation of W0 - W1 for short moves.
template <typename R>
__device__ inline void axis_window(R s0, R s1, int sg, R wq, const R a0[3], R w1[3], R fl[2],
                                   R& wo, R& fo) {
  const R zero = R(0), half = R(0.5), one = R(1);
  R a1[3];
  qweights<R>(s1, a1);
  const bool up = sg > 0, dn = sg < 0;
  w1[0] = up ? zero : (dn ? a1[1] : a1[0]);
  w1[1] = up ? a1[0] : (dn ? a1[2] : a1[1]);
  w1[2] = up ? a1[1] : (dn ? zero : a1[2]);
  wo = up ? a1[2] : (dn ? a1[0] : zero);
  // cumulative sums over the 5-slot frame N0-2 .. N0+2
  const R lo = dn ? a1[0] : zero;  // new weight at N0-2
  const R c0 = -lo;
  const R c1 = c0 + a0[0] - w1[0];
  const R c2 = c1 + a0[1] - w1[1];
  const R c3 = c2 + a0[2] - w1[2];
  const R h = half * wq * (s1 - s0);
  fl[0] = sg == 0 ? h * ((one - s0) - s1) : wq * c1;
  fl[1] = sg == 0 ? h * ((one + s0) + s1) : wq * c2;
  fo = up ? wq * c3 : (dn ? wq * c0 : zero);
}

// Esirkepov current (grid_part_connect.f90:979-1060, 1168-1334) of the
// particles of one cell, summed in registers: the part of each particle's
// current on the cell's node block N0-1 .. N0+1, which is all of it for a
// particle that keeps its nearest node (2x3x3 values per component, 2x3 in
// 2-D, offsets from node N0-1).  A particle whose node moved on one axis
// also has a layer just outside the block on that axis, added to LDS by
// edge().
template <int D, typename R>
struct CellAcc {
  R jx[2][3][3], jy[3][2][3], jz[3][3][2];

  __device__ inline void zero() {
#pragma unroll
    for (int a = 0; a < 18; a++) {
      (&jx[0][0][0])[a] = R(0);
      (&jy[0][0][0])[a] = R(0);
      (&jz[0][0][0])[a] = R(0);
    }
  }

  // a0, w1, fl per axis as axis_window returns them
  __device__ inline void add(const R a0[3][3], const R w1[3][3], const R fl[3][2]) {
    const R half = R(0.5);
    if constexpr (D == 2) {
#pragma unroll
      for (int mm = 0; mm < 3; mm++) {
        const R ty2 = a0[1][mm] + w1[1][mm];
        const R tx2 = a0[0][mm] + w1[0][mm];
#pragma unroll
        for (int f = 0; f < 2; f++) {
          jx[f][mm][0] = fma_of<R>(fl[0][f], ty2, jx[f][mm][0]);
          jy[mm][f][0] = fma_of<R>(fl[1][f], tx2, jy[mm][f][0]);
        }
      }
    } else {
#pragma unroll
      for (int mm = 0; mm < 3; mm++) {
        const R y0 = a0[1][mm] + half * w1[1][mm], y1 = w1[1][mm] + half * a0[1][mm];
        const R x0 = a0[0][mm] + half * w1[0][mm], x1 = w1[0][mm] + half * a0[0][mm];
#pragma unroll
        for (int l = 0; l < 3; l++) {
          const R tyz = y0 * a0[2][l] + y1 * w1[2][l];
          const R txz = x0 * a0[2][l] + x1 * w1[2][l];
          const R txy = x0 * a0[1][l] + x1 * w1[1][l];
#pragma unroll
          for (int f = 0; f < 2; f++) {
            jx[f][mm][l] = fma_of<R>(fl[0][f], tyz, jx[f][mm][l]);
            jy[mm][f][l] = fma_of<R>(fl[1][f], txz, jy[mm][f][l]);
            jz[mm][l][f] = fma_of<R>(fl[2][f], txy, jz[mm][l][f]);
          }
        }
      }
    }
  }

  // The layer outside the block of a particle whose node moved on one axis
  // (sg nonzero on that axis only), added to the FP64 LDS patch; j1 points at
  // node N0-1.  On the moved axis the layer holds the outer flux fo (its own
  // component) and the outer node's weight wo (the transverse components);
  // the other axes are inside the block.
  template <int JX, int JY, int PVOL>
  __device__ static inline void edge(double* j1, const int sg[3], const R a0[3][3],
                                     const R w1[3][3], const R fl[3][2], const R wo[3],
                                     const R fo[3]) {
    const R half = R(0.5);
    if constexpr (D == 2) {
      if (sg[0] != 0) {
        const int xf = sg[0] > 0 ? 2 : -1, xs = sg[0] > 0 ? 3 : -1;
#pragma unroll
        for (int m = 0; m < 3; m++)
          atomicAdd(j1 + xf + JX * m, (double)(fo[0] * (a0[1][m] + w1[1][m])));
#pragma unroll
        for (int f = 0; f < 2; f++) atomicAdd(j1 + PVOL + xs + JX * f, (double)(fl[1][f] * wo[0]));
      }
      if (sg[1] != 0) {
        const int yf = sg[1] > 0 ? 2 : -1, ys = sg[1] > 0 ? 3 : -1;
#pragma unroll
        for (int m = 0; m < 3; m++)
          atomicAdd(j1 + PVOL + m + JX * yf, (double)(fo[1] * (a0[0][m] + w1[0][m])));
#pragma unroll
        for (int f = 0; f < 2; f++) atomicAdd(j1 + f + JX * ys, (double)(fl[0][f] * wo[1]));
      }
    } else {
      if (sg[0] != 0) {
        const int xf = sg[0] > 0 ? 2 : -1, xs = sg[0] > 0 ? 3 : -1;
#pragma unroll
        for (int m = 0; m < 3; m++) {
          const R y0 = a0[1][m] + half * w1[1][m], y1 = w1[1][m] + half * a0[1][m];
#pragma unroll
          for (int l = 0; l < 3; l++)
            atomicAdd(j1 + xf + JX * (m + JY * l), (double)(fo[0] * (y0 * a0[2][l] + y1 * w1[2][l])));
        }
#pragma unroll
        for (int l = 0; l < 3; l++) {
          const R tz = wo[0] * (half * a0[2][l] + w1[2][l]);
          const R ty = wo[0] * (half * a0[1][l] + w1[1][l]);
#pragma unroll
          for (int f = 0; f < 2; f++) {
            atomicAdd(j1 + PVOL + xs + JX * (f + JY * l), (double)(fl[1][f] * tz));
            atomicAdd(j1 + 2 * PVOL + xs + JX * (l + JY * f), (double)(fl[2][f] * ty));
          }
        }
      }
      if (sg[1] != 0) {
        const int yf = sg[1] > 0 ? 2 : -1, ys = sg[1] > 0 ? 3 : -1;
#pragma unroll
        for (int m = 0; m < 3; m++) {
          const R x0 = a0[0][m] + half * w1[0][m], x1 = w1[0][m] + half * a0[0][m];
#pragma unroll
          for (int l = 0; l < 3; l++)
            atomicAdd(j1 + PVOL + m + JX * (yf + JY * l),
                      (double)(fo[1] * (x0 * a0[2][l] + x1 * w1[2][l])));
        }
#pragma unroll
        for (int l = 0; l < 3; l++) {
          const R tz = wo[1] * (half * a0[2][l] + w1[2][l]);
          const R tx = wo[1] * (w1[0][l] + half * a0[0][l]);
#pragma unroll
          for (int f = 0; f < 2; f++) {
            atomicAdd(j1 + f + JX * (ys + JY * l), (double)(fl[0][f] * tz));
            atomicAdd(j1 + 2 * PVOL + l + JX * (ys + JY * f), (double)(fl[2][f] * tx));
          }
        }
      }
      if (sg[2] != 0) {
        const int zf = sg[2] > 0 ? 2 : -1, zs = sg[2] > 0 ? 3 : -1;
#pragma unroll
        for (int m = 0; m < 3; m++) {
          const R x0 = a0[0][m] + half * w1[0][m], x1 = w1[0][m] + half * a0[0][m];
#pragma unroll
          for (int l = 0; l < 3; l++)
            atomicAdd(j1 + 2 * PVOL + m + JX * (l + JY * zf),
                      (double)(fo[2] * (x0 * a0[1][l] + x1 * w1[1][l])));
        }
#pragma unroll
        for (int l = 0; l < 3; l++) {
          const R ty = wo[2] * (w1[1][l] + half * a0[1][l]);
          const R tx = wo[2] * (w1[0][l] + half * a0[0][l]);
#pragma unroll
          for (int f = 0; f < 2; f++) {
            atomicAdd(j1 + f + JX * (l + JY * zs), (double)(fl[0][f] * ty));
            atomicAdd(j1 + PVOL + l + JX * (f + JY * zs), (double)(fl[1][f] * tx));
          }
        }
      }
    }
  }

  // adds the sums into the FP64 LDS patch; j1 points at node N-1
  template <int JX, int JY, int PVOL>
  __device__ inline void flush(double* j1) const {
    if constexpr (D == 2) {
#pragma unroll
      for (int f = 0; f < 2; f++)
#pragma unroll
        for (int mm = 0; mm < 3; mm++) {
          atomicAdd(j1 + (f + JX * mm), (double)jx[f][mm][0]);
          atomicAdd(j1 + PVOL + (mm + JX * f), (double)jy[mm][f][0]);
        }
    } else {
#pragma unroll
      for (int mm = 0; mm < 3; mm++)
#pragma unroll
        for (int l = 0; l < 3; l++)
#pragma unroll
          for (int f = 0; f < 2; f++) {
            atomicAdd(j1 + (f + JX * (mm + JY * l)), (double)jx[f][mm][l]);
            atomicAdd(j1 + PVOL + (mm + JX * (f + JY * l)), (double)jy[mm][f][l]);
            atomicAdd(j1 + 2 * PVOL + (mm + JX * (l + JY * f)), (double)jz[mm][l][f]);
          }
    }
  }
};

// LDS layout of the tile kernel (all sizes compile-time; R = float or double):
//   Fl  R      [NF][GVOL]   field patch, nodes T0-1 .. T0+T per axis
//   Jl  double [D][PVOL]    current patch, nodes T0-2 .. T0+T+1 per axis
//   Q   per wave: R s[6][QCAP], float wq[QCAP], int meta[QCAP]: the queue of
//       particles whose node moved on two or three axes
//   E   per wave: the same for particles whose node moved on one axis (the
//       layer of their current outside the old cell's node block)
//   M   per wave: the record words of MCAP movers (word-major), their
//       destination tiles and ids
//   offs, fillc int [NCL+1], [NCL]; misc int [4]; wsum int [NW]; perm u16 [kChunk]
//
// What the kernel does to a particle before it moves, and so what its field
// patch holds:
//   LpPush   the LP step's gather + exact-gamma push: alp*q*F of the nfield
//            components over nodes T0-1 .. T0+T (the reach of the qlh shapes)
//   EnvMove  the envelope step's second half (set_env_grad_interp and the
//            gamma of lpf_env_positions): je(1:D+1) = |A|^2/2 and its
//            gradient over nodes T0-2 .. T0+T (the reach of the qqh shapes)
// Everything after (removal, deposit, re-binning) is the same.
// SPREAD: whether the lanes of a wave take slices of the sorted chunk far
// enough apart that no two of them sit in one cell (fewer same-address LDS
// atomics, more bank conflicts in the gather); without it thread me takes
// slice me.
struct LpPush {
  template <int D> static constexpr int NF = D == 3 ? 6 : 3;
  static constexpr int LO = 1, EXT = 2;  // patch nodes T0-LO .. T0-LO+T+EXT-1
  static constexpr bool SPREAD = true;
};
struct EnvMove {
  template <int D> static constexpr int NF = D + 1;
  static constexpr int LO = 2, EXT = 3;
  static constexpr bool SPREAD = false;
};

template <int D, int TX, int TY, int TZ, int NT, typename R, class P = LpPush>
struct TileShape {
  static constexpr int NCL = TX * TY * TZ, NF = P::template NF<D>, NW = NT / 64;
  static constexpr int GX = TX + P::EXT, GY = TY + P::EXT, GZ = D == 3 ? TZ + P::EXT : 1;
  static constexpr int GVOL = GX * GY * GZ;
  static constexpr int JX = TX + 4, JY = TY + 4, JZ = D == 3 ? TZ + 4 : 1, PVOL = JX * JY * JZ;
  static constexpr int QCAP = 64;  // one entry per lane: a drain takes all
  static constexpr int F_BYTES = (NF * GVOL * (int)sizeof(R) + 15) & ~15;
  static constexpr int J_BYTES = 8 * D * PVOL;
  static constexpr int QW_BYTES = QCAP * (6 * (int)sizeof(R) + 8);  // one wave's queue
  static constexpr int Q_BYTES = NW * QW_BYTES;
  static constexpr int MCAP = 64;  // movers staged per wave
  static constexpr int MW_BYTES = MCAP * (8 * (int)sizeof(R) + 8);
  static constexpr int M_BYTES = NW * MW_BYTES;
  static constexpr int I_BYTES = 4 * (2 * NCL + 5 + NW);
  static constexpr int BYTES = F_BYTES + J_BYTES + 2 * Q_BYTES + M_BYTES + I_BYTES + 2 * kChunk;
  static_assert(NT % 64 == 0 && NT >= NCL, "threads: whole wavefronts, at least one per cell");
  static_assert(kChunk % NT == 0, "a chunk is a whole number of records per thread");
};

// One wave's queue in LDS.
template <typename R, int QCAP>
struct WaveQueue {
  R* s;        // [6][QCAP]: s0[3], s1[3]
  float* wq;   // [QCAP]
  int* meta;   // [QCAP]: local cell | (sign+1) << (10 + 2d) per axis
  __device__ WaveQueue(char* base)
      : s((R*)base), wq((float*)(base + 6 * QCAP * sizeof(R))),
        meta((int*)(base + 6 * QCAP * sizeof(R) + 4 * QCAP)) {}
};

// Drains the n <= 64 queued particles, one per lane, through the
// union-window deposit.
template <int D, int TX, int TY, int TZ, int NT, typename R>
__device__ inline void drain_queue(const WaveQueue<R, 64>& qu, int n, double* Jl) {
  using S = TileShape<D, TX, TY, TZ, NT, R>;
  const int lane = (int)__lane_id();
  if (lane < n) {
    const int e = lane;
    R s0[3] = {R(0), R(0), R(0)}, s1[3] = {R(0), R(0), R(0)};
#pragma unroll
    for (int d = 0; d < D; d++) {
      s0[d] = qu.s[d * S::QCAP + e];
      s1[d] = qu.s[(3 + d) * S::QCAP + e];
    }
    const R wq = (R)qu.wq[e];
    const int meta = qu.meta[e];
    const int lc = meta & 1023;
    int o0[3] = {0, 0, 0}, o1[3] = {0, 0, 0}, sh[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; d++) {
      const int sig = ((meta >> (10 + 2 * d)) & 3) - 1;
      o0[d] = sig < 0 ? 1 : 0;
      o1[d] = sig > 0 ? 1 : 0;
      sh[d] = sig < 0 ? -1 : 0;
    }
    const int cx = lc % TX, cy = (lc / TX) % TY, cz = D == 3 ? lc / (TX * TY) : 0;
    // union base N-1+min(0,sig) sits at patch offset 1+min(0,sig) from node N-2
    double* ub = Jl + (cx + 1 + sh[0]) +
                 S::JX * ((cy + 1 + sh[1]) + S::JY * (D == 3 ? cz + 1 + sh[2] : 0));
    deposit_general<D, S::JX, S::JY, S::PVOL, R>(ub, s0, s1, o0, o1, wq);
  }
}

// One axis on which the nearest node stayed: quadratic weights of the old
// and new offsets and the two in-block fluxes, as axis_window gives them for
// sg = 0.
template <typename R>
__device__ inline void axis_kept(R s0, R s1, R wq, R a0[3], R w1[3], R fl[2]) {
  const R half = R(0.5), one = R(1);
  qweights<R>(s0, a0);
  qweights<R>(s1, w1);
  const R h = half * wq * (s1 - s0);
  fl[0] = h * ((one - s0) - s1);
  fl[1] = h * ((one + s0) + s1);
}

// Drains the n <= 64 queued single-axis movers, one per lane: the layer of
// each one's Esirkepov current just outside its old cell's node block (the
// rest went into the register sums), the terms CellAcc::edge adds.  The
// moved axis a and the other two b < c are per-lane values (LDS strides and
// component offsets), so every lane runs the same instructions whichever
// axis its particle crossed on:
//   J_a, outer flux:  fo * (P0_b[m] a0_c[l] + P1_b[m] w1_c[l]), P0 = a0 + w1/2,
//                     P1 = w1 + a0/2 (the transverse factor of J_a)
//   J_b, outer node:  fl_b[f] * (wo * (a0_c[l]/2 + w1_c[l]))
//   J_c, outer node:  fl_c[f] * (wo * (a0_b[l]/2 + w1_b[l]))
// (in 2-D: fo * (a0_b[m] + w1_b[m]) and fl_b[f] * wo).
template <int D, int TX, int TY, int TZ, int NT, typename R>
__device__ inline void drain_edges(const WaveQueue<R, 64>& qu, int n, double* Jl) {
  using S = TileShape<D, TX, TY, TZ, NT, R>;
  constexpr int JX = S::JX, JXY = S::JX * S::JY, PVOL = S::PVOL;
  const int lane = (int)__lane_id();
  if (lane < n) {
    const int e = lane;
    const R wq = (R)qu.wq[e];
    const int meta = qu.meta[e];
    const int lc = meta & 1023;
    const int sgy = ((meta >> 12) & 3) - 1, sgz = D == 3 ? ((meta >> 14) & 3) - 1 : 0;
    const int ax = sgy != 0 ? 1 : (sgz != 0 ? 2 : 0);  // the moved axis
    const int sg = ((meta >> (10 + 2 * ax)) & 3) - 1;
    const int b = ax == 0 ? 1 : 0, c = ax == 2 ? 1 : 2;
    const int sa = ax == 0 ? 1 : (ax == 1 ? JX : JXY);
    const int sb = b == 0 ? 1 : JX, sc = c == 1 ? JX : JXY;
    const int af = sg > 0 ? 2 : -1, as = sg > 0 ? 3 : -1;
    const R half = R(0.5);
    R a0a[3], w1a[3], fla[2], wo, fo;
    const R s0a = qu.s[ax * S::QCAP + e];
    qweights<R>(s0a, a0a);
    axis_window<R>(s0a, qu.s[(3 + ax) * S::QCAP + e], sg, wq, a0a, w1a, fla, wo, fo);
    R a0b[3], w1b[3], flb[2];
    axis_kept<R>(qu.s[b * S::QCAP + e], qu.s[(3 + b) * S::QCAP + e], wq, a0b, w1b, flb);
    const int cx = lc % TX, cy = (lc / TX) % TY, cz = D == 3 ? lc / (TX * TY) : 0;
    double* j1 = Jl + (cx + 1) + JX * ((cy + 1) + S::JY * (D == 3 ? cz + 1 : 0));
    double* pa = j1 + ax * PVOL + af * sa;
    double* pb = j1 + b * PVOL + as * sa;
    if constexpr (D == 2) {
#pragma unroll
      for (int m = 0; m < 3; m++) atomicAdd(pa + m * sb, (double)(fo * (a0b[m] + w1b[m])));
#pragma unroll
      for (int f = 0; f < 2; f++) atomicAdd(pb + f * sb, (double)(flb[f] * wo));
    } else {
      R a0c[3], w1c[3], flc[2];
      axis_kept<R>(qu.s[c * S::QCAP + e], qu.s[(3 + c) * S::QCAP + e], wq, a0c, w1c, flc);
      double* pc = j1 + c * PVOL + as * sa;
#pragma unroll
      for (int m = 0; m < 3; m++) {
        const R p0 = a0b[m] + half * w1b[m], p1 = w1b[m] + half * a0b[m];
#pragma unroll
        for (int l = 0; l < 3; l++)
          atomicAdd(pa + m * sb + l * sc, (double)(fo * (p0 * a0c[l] + p1 * w1c[l])));
      }
#pragma unroll
      for (int l = 0; l < 3; l++) {
        const R tc = wo * (half * a0c[l] + w1c[l]), tb = wo * (half * a0b[l] + w1b[l]);
#pragma unroll
        for (int f = 0; f < 2; f++) {
          atomicAdd(pb + f * sb + l * sc, (double)(flb[f] * tc));
          atomicAdd(pc + l * sb + f * sc, (double)(flc[f] * tb));
        }
      }
    }
  }
}

// One wave's staging buffer of particles that move to another tile: their
// slots in the destination segments are reserved when the buffer is written
// out, one atomic per destination for up to 64 of them, instead of one
// atomic round trip in every round.
template <typename R, int MCAP>
struct MoverStage {
  using V = typename RecWord<R>::V;
  static constexpr int NWD = RecWord<R>::N;
  V* w;       // [NWD][MCAP]
  int* tile;  // [MCAP]
  int* id;    // [MCAP]
  __device__ MoverStage(char* base)
      : w((V*)base), tile((int*)(base + NWD * MCAP * sizeof(V))),
        id((int*)(base + NWD * MCAP * sizeof(V) + 4 * MCAP)) {}
};

// Writes the n <= 64 staged movers, one per lane, into their destination
// segments of B (the overflow list when a segment is full).
template <typename R, int MCAP>
__device__ inline void flush_movers(const MoverStage<R, MCAP>& ms, int n, const RecPtr& B,
                                    const int* __restrict__ seg_start,
                                    const int* __restrict__ seg_cap, int* __restrict__ cnt_next,
                                    int* __restrict__ ovf_next, int* __restrict__ grow,
                                    int64_t ovf_base) {
  const int lane = (int)__lane_id();
  if (lane < n) {
    const int dtile = ms.tile[lane];
    const int cap = seg_cap[dtile];
    const int64_t start = seg_start[dtile];
    const int slot = reserve_slot(dtile, cnt_next);
    int64_t dst = start + slot;
    if (slot >= cap) {  // segment full: overflow list, and a larger segment next re-layout
      dst = ovf_base + atomicAdd(ovf_next, 1);
      grow[dtile] = 1;
    }
    typename RecWord<R>::V* d = rec_at<R>(B, dst);
#pragma unroll
    for (int k = 0; k < RecWord<R>::N; k++) d[k] = ms.w[k * MCAP + lane];
    if (B.id) B.id[dst] = ms.id[lane];
  }
}

// The envelope step's advance of one particle (set_env_grad_interp,
// grid_part_connect.f90:760-882, and the gamma of lpf_env_positions,
// boris_push.f90:225-295): F and grad F at x^n with qqh shapes from the
// patch at node N-2 (fb), gamp = sqrt(1 + p^2 + F) corrected by the
// half-step change of F along p; returns dt/gamp.  a0 receives the quadratic
// weights at x^n that the deposit reuses.
template <int D, int GX, int GXY, int GVOL, typename R>
__device__ inline R env_advance(const Rec<R>& q, const R* fb, R a0[3][3], R dt) {
  R axh[3], ayh[3], azh[3] = {R(1), R(0), R(0)};
  int hx, hy, hz = 0;
  qqh<R>(q.s[0], a0[0], axh, hx);
  qqh<R>(q.s[1], a0[1], ayh, hy);
  if (D == 3) qqh<R>(q.s[2], a0[2], azh, hz);
  const int k1 = D == 3 ? 1 : 0;  // integer-shape base N-1 is local 1
  R ap[4] = {R(0), R(0), R(0), R(0)};
  ap[0] = contract3l<D, R>(fb + GVOL, GX, GXY, axh, a0[1], a0[2], 1 + hx, 1, k1);
  ap[D] = contract3l<D, R>(fb, GX, GXY, a0[0], a0[1], a0[2], 1, 1, k1);
  ap[1] = contract3l<D, R>(fb + 2 * GVOL, GX, GXY, a0[0], ayh, a0[2], 1, 1 + hy, k1);
  if (D == 3) ap[2] = contract3l<D, R>(fb + 3 * GVOL, GX, GXY, a0[0], a0[1], azh, 1, 1, 1 + hz);
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
  const R gam_inv = (one / gam) * (one - dt * b2 / gam3);
  return dt * gam_inv;
}

// The workgroup of tile t.
template <int D, int TX, int TY, int TZ, int NT, typename R, class P>
__device__ inline void tile_body(const TileArgs& ta, const StepConst<R>& sc,
                                 const R* __restrict__ ebf, const RecPtr& A, const RecPtr& B,
                                 const int* __restrict__ seg_start, const int* __restrict__ seg_cap,
                                 const int* __restrict__ cnt_cur, int* __restrict__ cnt_next,
                                 int* __restrict__ ovf_cur, int* __restrict__ ovf_next,
                                 double* __restrict__ patch, int* __restrict__ grow, const int t) {
  using S = TileShape<D, TX, TY, TZ, NT, R, P>;
  constexpr int NCL = S::NCL, NF = S::NF, GX = S::GX, GY = S::GY, GVOL = S::GVOL;
  constexpr int LO = P::LO;
  constexpr int JX = S::JX, JY = S::JY, PVOL = S::PVOL, QCAP = S::QCAP;
  extern __shared__ double lds_d[];
  char* lds = (char*)lds_d;
  R* Fl = (R*)lds;
  double* Jl = (double*)(lds + S::F_BYTES);
  char* Q = lds + S::F_BYTES + S::J_BYTES;
  char* E = Q + S::Q_BYTES;
  int* offs = (int*)(E + S::Q_BYTES + S::M_BYTES);  // NCL + 1
  int* fillc = offs + NCL + 1;  // NCL
  int* misc = fillc + NCL;      // [0] chunk slot base, [1] stayers, [2] max per cell
  int* wsum = misc + 4;          // per-wave totals of the cell-count scan
  unsigned short* perm = (unsigned short*)(wsum + S::NW);

  const int tx = t % ta.nt0, ty = (t / ta.nt0) % ta.nt1, tz = t / (ta.nt0 * ta.nt1);
  const int T0x = 3 + tx * TX, T0y = 3 + ty * TY, T0z = D == 3 ? 3 + tz * TZ : 1;
  const int me = threadIdx.x;
  const int lane = (int)__lane_id();
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  const int cap_t = seg_cap[t];
  const int cnt = min(cnt_cur[t], cap_t);
  const int64_t base = seg_start[t];
  const int mi = (1 << ta.bx) - 1, mj = (1 << ta.by) - 1, shk = ta.bx + ta.by;
  // the packed cells of a chunk, loaded once (all loads in flight together)
  // and kept in registers for the histogram and the permutation
  constexpr int CPT = kChunk / NT;
  int lcv[CPT];
  const auto load_cells = [&](int c0) {
    const int m = min(kChunk, cnt - c0);
#pragma unroll
    for (int u = 0; u < CPT; u++) {
      const int q = me + u * NT;
      lcv[u] = q < m ? rec_cell<R>(A, base + c0 + q) : kHole;
    }
  };
  // field patch: nodes T0-LO .. T0-LO+G-1; every load of the thread, and
  // those of the first chunk's cells, is issued before the first LDS store,
  // so the HBM latency is paid once per tile
  {
    // rows of GX consecutive x nodes: row = (c*GZ + lk)*GY + lj holds LDS
    // entries row*GX .. row*GX+GX-1; a pass covers RPP rows, one node per
    // thread, so the index arithmetic is per row and mostly per thread
    constexpr int GZ = S::GZ, ROWS = NF * GZ * GY, RPP = NT / GX;
    constexpr int PASSES = (ROWS + RPP - 1) / RPP;
    const int li = me % GX, r0 = me / GX;
    const int i = T0x - LO + li;
    const bool xin = r0 < RPP && i <= ta.n1p;
    // node (T0x-LO+li, T0y-LO, T0z-LO) of component 0
    const R* eb = ebf + (i - 1) +
                  (int64_t)ta.n1p *
                      ((T0y - LO - 1) + (int64_t)ta.n2p * (D == 3 ? T0z - LO - 1 : 0));
    R fv[PASSES];
#pragma unroll
    for (int u = 0; u < PASSES; u++) {
      const int row = r0 + u * RPP;
      const int c = row / (GZ * GY), rem = row - c * (GZ * GY);
      const int lk = rem / GY, lj = rem - lk * GY;
      const int j = T0y - LO + lj, k = D == 3 ? T0z - LO + lk : 1;
      fv[u] = R(0);
      if (xin && row < ROWS && j <= ta.n2p && k <= ta.n3p)
        fv[u] = eb[(int64_t)c * ta.ncell + ta.n1p * (lj + ta.n2p * lk)];
    }
    load_cells(0);
    // the LP patch holds alp*charge*F: the push's charge multiply and alp
    // factor (field_charge_multiply, boris_push.f90:75-168) are done here
    // once per node instead of once per particle
    const R fs = std::is_same<P, LpPush>::value ? sc.alp * sc.qch : R(1);
#pragma unroll
    for (int u = 0; u < PASSES; u++) {
      const int row = r0 + u * RPP;
      if (r0 < RPP && row < ROWS) Fl[row * GX + li] = fs * fv[u];
    }
  }
  // the FP64 patch moves as 16-byte pairs (D*PVOL is even: every patch
  // extent T+4 is)
  static_assert((D * PVOL) % 2 == 0, "patch of whole double pairs");
  for (int q = threadIdx.x; q < D * PVOL / 2; q += NT)
    reinterpret_cast<double2*>(Jl)[q] = make_double2(0.0, 0.0);

  const WaveQueue<R, QCAP> qu(Q + (me >> 6) * S::QW_BYTES);
  int qcount = 0;  // wave-uniform: entries 0 .. qcount-1 are waiting
  const WaveQueue<R, QCAP> equ(E + (me >> 6) * S::QW_BYTES);
  int ecount = 0;  // wave-uniform: single-axis movers waiting for their layer
  const MoverStage<R, S::MCAP> mst(E + S::Q_BYTES + (me >> 6) * S::MW_BYTES);
  int mcount = 0;  // wave-uniform: staged movers
  // LDS current patch node N-1 of local cell lc (the patch starts at T0-2)
  const auto patch_node = [&](int lc) -> double* {
    const int li = lc % TX, lj = (lc / TX) % TY, lk = D == 3 ? lc / (TX * TY) : 0;
    return Jl + (li + 1) + JX * ((lj + 1) + JY * (D == 3 ? lk + 1 : 0));
  };
  CellAcc<D, R> acc;
  acc.zero();
  int acc_lc = -1;

  for (int c0 = 0; c0 < cnt; c0 += kChunk) {
    __syncthreads();  // patches staged / previous chunk consumed
    if (me < NCL) offs[me] = 0;
    if (P::SPREAD && me == 0) misc[2] = 0;
    // local cell index, kHole for a hole, -2 for a record of another tile
#pragma unroll
    for (int u = 0; u < CPT; u++) {
      const int cell = lcv[u];
      if (cell == kHole) continue;
      const int li = (cell & mi) - T0x, lj = ((cell >> ta.bx) & mj) - T0y;
      const int lk = D == 3 ? (cell >> shk) - T0z : 0;
      lcv[u] = (li >= 0 && li < TX && lj >= 0 && lj < TY && lk >= 0 && lk < TZ)
                   ? li + TX * (lj + TY * lk) : -2;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CPT; u++)
      if (lcv[u] >= 0) atomicAdd(&offs[lcv[u]], 1);
    __syncthreads();
    {  // exclusive scan of offs[0..NCL): within each wave by shuffles, then
       // across the waves through wsum
      const int v = me < NCL ? offs[me] : 0;
      int x = v;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
      }
      const int vmax = P::SPREAD ? wave_max(v) : 0;
      if (lane == 63) {
        wsum[me >> 6] = x;
        if (vmax > 0) atomicMax(&misc[2], vmax);
      }
      if (me < NCL) fillc[me] = 0;
      __syncthreads();
      int pre = 0;
      for (int w = 0; w < (me >> 6); w++) pre += wsum[w];
      const int incl = x + pre;
      if (me < NCL) offs[me] = incl - v;
      if (me == NCL - 1) {
        offs[NCL] = incl;
        // slots for every valid particle of the chunk in this tile
        misc[0] = atomicAdd(&cnt_next[t], incl);
        misc[1] = 0;
      }
      __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < CPT; u++) {
      const int q = me + u * NT, lc = lcv[u];
      if (lc >= 0) {
        perm[offs[lc] + atomicAdd(&fillc[lc], 1)] = (unsigned short)q;
      } else if (lc == -2) {
        // not in this tile (the store never does this): hand it to the
        // global-memory kernel that runs after this one
        copy_rec<R>(A, base + c0 + q, A, sc.ovf_base + atomicAdd(ovf_cur, 1));
      }
    }
    __syncthreads();
    const int chunk_base = misc[0];
    const int nval = offs[NCL];
    // balanced split of the cell-sorted chunk into NT slices of W consecutive
    // sorted positions, so every thread has the same trip count whatever the
    // density.  The lanes of a wave take slices sigma apart (sigma a power of
    // two dividing NW), so their particles are sigma*W sorted positions apart:
    // on different cells when no cell holds more than sigma*W (no two lanes
    // of an LDS atomic on one address), and on nearby cells when sigma is as
    // small as that allows (stencil addresses spread over the LDS banks).
    // sigma is the smallest such stride for the fullest cell of the chunk
    // (1 for a policy without SPREAD).
    constexpr int NW = S::NW;
    const int W = (nval + NT - 1) / NT;
    int sigma = 1;
    if constexpr (P::SPREAD)
      while (sigma < NW && sigma * W < misc[2]) sigma <<= 1;
    const int wv = me >> 6;
    const int sl = (wv / sigma) * (64 * sigma) + (wv % sigma) + sigma * lane;
    const int q0 = sl * W, q1 = min(q0 + W, nval);
    const int64_t cb = base + c0;
    Rec<R> nxt = {};
    if (q0 < q1) load_raw<D, R>(A, cb + perm[q0], nxt);
    for (int r = 0; r < W; r++) {
      const int qi = q0 + r;
      Rec<R> q = nxt;
      if (qi + 1 < q1) load_raw<D, R>(A, cb + perm[qi + 1], nxt);
      bool keep = false, cross = false;
      int nmv = 0;  // axes on which the nearest node moved
      R so[3] = {R(0), R(0), R(0)}, wq = R(0);
      R a0[3][3] = {{R(1), R(0), R(0)}, {R(1), R(0), R(0)}, {R(1), R(0), R(0)}};
      int lc = 0, N0[3] = {0, 0, 0};
      if (qi < q1) {
        decode_cell<D, R>(q, ta.bx, ta.by);
        const int li = q.N[0] - T0x, lj = q.N[1] - T0y, lk = D == 3 ? q.N[2] - T0z : 0;
        lc = li + TX * (lj + TY * lk);
        // LDS field patch base at node N-LO of the particle's cell
        const R* fb = Fl + li + GX * (lj + GY * lk);
#pragma unroll
        for (int d = 0; d < 3; d++) {
          so[d] = q.s[d];
          N0[d] = q.N[d];
        }
        R dtg;
        if constexpr (std::is_same<P, LpPush>::value) {
          const auto F = [&](int c, int i1, int j1, int k1) -> R {
            return fb[c * GVOL + i1 + GX * (j1 + GY * k1)];
          };
          R ef[6];
          gather<D, R>(q.s, F, ef, a0);
          // the fields are already alp*charge*F
          dtg = push<D, R>(q.p, ef, R(1), R(1), sc.dt, sc.initial);
        } else {
          dtg = env_advance<D, GX, GX * GY, GVOL, R>(q, fb, a0, sc.dt);
        }
        keep = move_and_test<D, R>(q, dtg, sc.inv, sc.nn);
        // real(sp) :: wght = charge*wgh (grid_part_connect.f90:983, 1205):
        // the product is rounded to FP32 in both paths
        wq = (R)((float)sc.qch * q.w);
        int sg[3] = {0, 0, 0}, nmoved = 0;
#pragma unroll
        for (int d = 0; d < D; d++) {
          sg[d] = q.N[d] - N0[d];
          nmoved += sg[d] != 0 ? 1 : 0;
        }
        nmv = nmoved;
        cross = nmoved > 1;  // node moved on two or three axes: the queue
        if (keep && !cross) {
          // a thread's slice runs through the cells in order: sum in
          // registers while the cell stays, add to LDS when it changes
          if (lc != acc_lc) {
            if (acc_lc >= 0) acc.template flush<JX, JY, PVOL>(patch_node(acc_lc));
            acc.zero();
            acc_lc = lc;
          }
          R w1[3][3] = {{R(1), R(0), R(0)}, {R(1), R(0), R(0)}, {R(1), R(0), R(0)}};
          R fl[3][2] = {}, wo[3] = {}, fo[3] = {};
#pragma unroll
          for (int d = 0; d < D; d++) axis_window<R>(so[d], q.s[d], sg[d], wq, a0[d], w1[d], fl[d], wo[d], fo[d]);
          acc.add(a0, w1, fl);
        }
      }
      // particles whose node moved on one axis: the layer of their current
      // outside the node block goes through the wave's edge queue
      const bool emove = keep && nmv == 1;
      const unsigned long long em = __ballot(emove);
      const int enew = __popcll(em);
      if (ecount + enew > QCAP) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        drain_edges<D, TX, TY, TZ, NT, R>(equ, ecount, Jl);
        ecount = 0;
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      if (emove) {
        const int e = ecount + __popcll(em & lt_mask);
        int meta = lc;
#pragma unroll
        for (int d = 0; d < D; d++) {
          equ.s[d * QCAP + e] = so[d];
          equ.s[(3 + d) * QCAP + e] = q.s[d];
          meta |= (q.N[d] - N0[d] + 1) << (10 + 2 * d);
        }
        equ.wq[e] = (float)wq;
        equ.meta[e] = meta;
      }
      ecount += enew;
      // particles whose node moved on two or three axes go to the wave's queue
      const unsigned long long cm = __ballot(keep && cross);
      const int nnew = __popcll(cm);
      if (qcount + nnew > QCAP) {  // no room for this round's entries: drain first
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        drain_queue<D, TX, TY, TZ, NT, R>(qu, qcount, Jl);
        qcount = 0;
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      if (keep && cross) {
        const int e = qcount + __popcll(cm & lt_mask);
        int meta = lc;
#pragma unroll
        for (int d = 0; d < D; d++) {
          qu.s[d * QCAP + e] = so[d];
          qu.s[(3 + d) * QCAP + e] = q.s[d];
          meta |= (q.N[d] - N0[d] + 1) << (10 + 2 * d);
        }
        qu.wq[e] = (float)wq;
        qu.meta[e] = meta;
      }
      qcount += nnew;
      // Survivors go to the other buffer: those that stay into the slots the
      // chunk reserved in this tile's segment, the others into their new
      // tile's segment; reserved slots left over become holes after the loop.
      const int dtile = (q.N[0] - 3) / TX +
                        ta.nt0 * ((q.N[1] - 3) / TY + ta.nt1 * (D == 3 ? (q.N[2] - 3) / TZ : 0));
      const bool stay = keep && dtile == t;
      const bool mover = keep && !stay;
      const unsigned long long mm = __ballot(mover);
      const int mnew = __popcll(mm);
      if (mcount + mnew > S::MCAP) {  // no room for this round's movers: write out first
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        flush_movers<R, S::MCAP>(mst, mcount, B, seg_start, seg_cap, cnt_next, ovf_next, grow,
                                 sc.ovf_base);
        mcount = 0;
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      if (mover) {
        const int e = mcount + __popcll(mm & lt_mask);
        typename RecWord<R>::V wd[RecWord<R>::N];
        rec_words<D, R>(pack_cell<D>(q.N, ta.bx, ta.by), q.s, q.p, q.w, wd);
#pragma unroll
        for (int k = 0; k < RecWord<R>::N; k++) mst.w[k * S::MCAP + e] = wd[k];
        mst.tile[e] = dtile;
        mst.id[e] = q.id;
      }
      mcount += mnew;
      // stayers of a round take consecutive slots of the chunk's reserved
      // range (one wave-aggregated LDS atomic), so a wave's record stores
      // are contiguous
      if (stay) {
        const int slot = chunk_base + reserve_one(&misc[1]);
        int64_t dst = base + slot;
        if (slot >= cap_t) {  // segment full: overflow list, and a larger segment next re-layout
          dst = sc.ovf_base + atomicAdd(ovf_next, 1);
          grow[t] = 1;
        }
        store_rec<D, R>(B, dst, q, ta.bx, ta.by);
      }
    }
    __syncthreads();
    {  // reserved slots not taken by stayers become holes
      const int nstay = misc[1], nval = offs[NCL];
      for (int q = nstay + me; q < nval; q += NT) {
        const int slot = chunk_base + q;
        if (slot < cap_t) set_rec_cell<R>(B, base + slot, kHole);
      }
    }
    if (c0 + kChunk < cnt) load_cells(c0 + kChunk);
  }
  if (mcount > 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    flush_movers<R, S::MCAP>(mst, mcount, B, seg_start, seg_cap, cnt_next, ovf_next, grow,
                             sc.ovf_base);
  }
  if (qcount > 0 || ecount > 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (qcount > 0) drain_queue<D, TX, TY, TZ, NT, R>(qu, qcount, Jl);
    if (ecount > 0) drain_edges<D, TX, TY, TZ, NT, R>(equ, ecount, Jl);
  }
  if (acc_lc >= 0) acc.template flush<JX, JY, PVOL>(patch_node(acc_lc));
  __syncthreads();
  double* Pt = patch + (int64_t)t * D * PVOL;
  for (int q = threadIdx.x; q < D * PVOL / 2; q += NT)
    reinterpret_cast<double2*>(Pt)[q] = reinterpret_cast<const double2*>(Jl)[q];
}

template <int D, int TX, int TY, int TZ, int NT, typename R, class P = LpPush>
__global__ __launch_bounds__(NT, sizeof(R) == 4 ? 512 / NT : 1) void k_tile(
    TileArgs ta, StepConst<R> sc, const R* __restrict__ ebf, RecPtr A, RecPtr B,
    const int* __restrict__ seg_start, const int* __restrict__ seg_cap,
    const int* __restrict__ cnt_cur, int* __restrict__ cnt_next, int* __restrict__ ovf_cur,
    int* __restrict__ ovf_next, double* __restrict__ patch, int* __restrict__ grow) {
  tile_body<D, TX, TY, TZ, NT, R, P>(ta, sc, ebf, A, B, seg_start, seg_cap, cnt_cur, cnt_next,
                                  ovf_cur, ovf_next, patch, grow, (int)blockIdx.x);
}

// Many independent runs in one launch: workgroup b takes tile work[b].y of
// run work[b].x, whose stores and grids the run's descriptor holds.
template <int D, int TX, int TY, int TZ, int NT, typename R>
__global__ __launch_bounds__(NT, sizeof(R) == 4 ? 512 / NT : 1) void k_tile_batch(
    const TileRun<R>* __restrict__ runs, const int2* __restrict__ work) {
  const int2 w = work[blockIdx.x];
  const TileRun<R>& r = runs[w.x];
  tile_body<D, TX, TY, TZ, NT, R, LpPush>(r.ta, r.sc, r.ebf, r.A, r.B, r.seg_start, r.seg_cap,
                                  r.cnt_cur, r.cnt_next, r.ovf_cur, r.ovf_next, r.patch, r.grow,
                                  w.y);
}

// Threads per workgroup: one per node-cell of the tile by default (4
// wavefronts for an 8x8x4 tile), or two per node-cell (NT = 2 NCL, the
// aladyn_config.block knob, compiled for the 8x8x4 and 16x16 shapes): twice
// the threads split the same sorted chunk into half as long slices, and the
// workgroup holds twice the per-wave queues in LDS.  The register sums of
// the current need more than the 128 VGPRs a 1024-thread workgroup would
// allow, so NT stays at most 512.
template <int D, int TX, int TY, int TZ, int NT, typename R, class P = LpPush>
static int lds_bytes_for() {
  return TileShape<D, TX, TY, TZ, NT, R, P>::BYTES;
}

template <int D, int TX, int TY, int TZ, int NT, typename R, class P = LpPush>
static hipError_t launch(const TileGeom& tg, const TileArgs& ta, const StepConst<R>& sc,
                         const R* ebf, const RecPtr& A, const RecPtr& B, const int* ss,
                         const int* sc_, const int* cc, int* cn, int* oc, int* on, double* patch,
                         int* grow, hipStream_t s) {
  k_tile<D, TX, TY, TZ, NT, R, P>
      <<<tg.ntiles, NT, lds_bytes_for<D, TX, TY, TZ, NT, R, P>(), s>>>(
          ta, sc, ebf, A, B, ss, sc_, cc, cn, oc, on, patch, grow);
  return hipGetLastError();
}

// shapes and workgroup sizes with a compiled kernel; tile_geom_init accepts
// only these (and only those whose LDS fits, see tile_lds_bytes)
bool tile_shape_supported(int D, int tx, int ty, int tz, int nt, int real_bytes) {
  const int ncl = tx * ty * (D == 3 ? tz : 1);
  if (D == 3) {
    // the FP64 kernel's registers do not fit two wavefronts per SIMD
    if (tx == 8 && ty == 8 && tz == 4) return nt == ncl || (nt == 2 * ncl && real_bytes == 4);
    return nt == ncl && ((tx == 8 && ty == 8 && tz == 8) || (tx == 16 && ty == 8 && tz == 4));
  }
  if (tx == 16 && ty == 16) return nt == ncl || nt == 2 * ncl;
  return nt == ncl && tx == 32 && ty == 16;
}

template <typename R>
static int lds_bytes_r(int D, int tx, int ty, int tz, int nt) {
  if (D == 3) {
    if (tx == 8 && ty == 8 && tz == 8) return lds_bytes_for<3, 8, 8, 8, 512, R>();
    if (tx == 8 && ty == 8 && tz == 4)
      return nt == 512 ? lds_bytes_for<3, 8, 8, 4, 512, R>() : lds_bytes_for<3, 8, 8, 4, 256, R>();
    return lds_bytes_for<3, 16, 8, 4, 512, R>();
  }
  if (tx == 16)
    return nt == 512 ? lds_bytes_for<2, 16, 16, 1, 512, R>() : lds_bytes_for<2, 16, 16, 1, 256, R>();
  return lds_bytes_for<2, 32, 16, 1, 512, R>();
}

int tile_lds_bytes(int D, int tx, int ty, int tz, int nt, int real_bytes) {
  return real_bytes == 8 ? lds_bytes_r<double>(D, tx, ty, tz, nt)
                         : lds_bytes_r<float>(D, tx, ty, tz, nt);
}

template <typename R>
hipError_t launch_tile_kernel(const TileGeom& tg, const DevGrid& g, const StepConst<R>& sc,
                              const R* ebf, const RecPtr& A, const RecPtr& B,
                              const int* seg_start, const int* seg_cap, const int* cnt_cur,
                              int* cnt_next, int* ovf_cur, int* ovf_next, double* patch,
                              int* grow, hipStream_t s) {
  TileArgs ta{tg.nt[0], tg.nt[1], tg.bx, tg.by, g.n1p, g.n2p, g.n3p, g.ncell};
  const int* T = tg.T;
#define ALADYN_TILE_CASE(DD, X, Y, Z, NT)                                                    \
  if (g.ndim == DD && T[0] == X && T[1] == Y && (DD == 2 || T[2] == Z) && tg.block == NT)    \
    return launch<DD, X, Y, Z, NT, R>(tg, ta, sc, ebf, A, B, seg_start, seg_cap, cnt_cur,    \
                                      cnt_next, ovf_cur, ovf_next, patch, grow, s);
  ALADYN_TILE_CASE(3, 8, 8, 8, 512)
  ALADYN_TILE_CASE(3, 8, 8, 4, 256)
  if constexpr (sizeof(R) == 4) {  // the FP64 queues and patches of these exceed the LDS
    ALADYN_TILE_CASE(3, 8, 8, 4, 512)
    ALADYN_TILE_CASE(3, 16, 8, 4, 512)
  }
  ALADYN_TILE_CASE(2, 16, 16, 1, 256)
  ALADYN_TILE_CASE(2, 16, 16, 1, 512)
  ALADYN_TILE_CASE(2, 32, 16, 1, 512)
#undef ALADYN_TILE_CASE
  return hipErrorInvalidConfiguration;
}

// The envelope step's second half on the tile kernel (P = EnvMove): je
// holds je(1:D+1); compiled for one thread per node-cell of each tile shape
// (hipErrorInvalidConfiguration for two).
template <typename R>
hipError_t launch_env_move_kernel(const TileGeom& tg, const DevGrid& g, const StepConst<R>& sc,
                                  const R* je, const RecPtr& A, const RecPtr& B,
                                  const int* seg_start, const int* seg_cap, const int* cnt_cur,
                                  int* cnt_next, int* ovf_cur, int* ovf_next, double* patch,
                                  int* grow, hipStream_t s) {
  TileArgs ta{tg.nt[0], tg.nt[1], tg.bx, tg.by, g.n1p, g.n2p, g.n3p, g.ncell};
  const int* T = tg.T;
#define ALADYN_TILE_CASE(DD, X, Y, Z, NT)                                                    \
  if (g.ndim == DD && T[0] == X && T[1] == Y && (DD == 2 || T[2] == Z) && tg.block == NT)    \
    return launch<DD, X, Y, Z, NT, R, EnvMove>(tg, ta, sc, je, A, B, seg_start, seg_cap,     \
                                               cnt_cur, cnt_next, ovf_cur, ovf_next, patch,  \
                                               grow, s);
  ALADYN_TILE_CASE(3, 8, 8, 8, 512)
  ALADYN_TILE_CASE(3, 8, 8, 4, 256)
  if constexpr (sizeof(R) == 4) ALADYN_TILE_CASE(3, 16, 8, 4, 512)
  ALADYN_TILE_CASE(2, 16, 16, 1, 256)
  ALADYN_TILE_CASE(2, 32, 16, 1, 512)
#undef ALADYN_TILE_CASE
  return hipErrorInvalidConfiguration;
}
template hipError_t launch_env_move_kernel<float>(const TileGeom&, const DevGrid&,
                                                  const StepConst<float>&, const float*,
                                                  const RecPtr&, const RecPtr&, const int*,
                                                  const int*, const int*, int*, int*, int*,
                                                  double*, int*, hipStream_t);
template hipError_t launch_env_move_kernel<double>(const TileGeom&, const DevGrid&,
                                                   const StepConst<double>&, const double*,
                                                   const RecPtr&, const RecPtr&, const int*,
                                                   const int*, const int*, int*, int*, int*,
                                                   double*, int*, hipStream_t);

template <int D, int TX, int TY, int TZ, int NT, typename R>
static hipError_t launch_b(const TileRun<R>* runs, const int2* work, int nwork, hipStream_t s) {
  k_tile_batch<D, TX, TY, TZ, NT, R><<<nwork, NT, lds_bytes_for<D, TX, TY, TZ, NT, R>(), s>>>(
      runs, work);
  return hipGetLastError();
}

// The batched launch: every run has the geometry kind of tg (dimension, tile
// shape, threads per tile) and the work list holds nwork (run, tile) pairs.
template <typename R>
hipError_t launch_tile_batch(const TileGeom& tg, int ndim, const TileRun<R>* runs,
                             const int2* work, int nwork, hipStream_t s) {
  const int* T = tg.T;
#define ALADYN_TILE_CASE(DD, X, Y, Z, NT)                                                  \
  if (ndim == DD && T[0] == X && T[1] == Y && (DD == 2 || T[2] == Z) && tg.block == NT)    \
    return launch_b<DD, X, Y, Z, NT, R>(runs, work, nwork, s);
  ALADYN_TILE_CASE(3, 8, 8, 8, 512)
  ALADYN_TILE_CASE(3, 8, 8, 4, 256)
  if constexpr (sizeof(R) == 4) {
    ALADYN_TILE_CASE(3, 8, 8, 4, 512)
    ALADYN_TILE_CASE(3, 16, 8, 4, 512)
  }
  ALADYN_TILE_CASE(2, 16, 16, 1, 256)
  ALADYN_TILE_CASE(2, 16, 16, 1, 512)
  ALADYN_TILE_CASE(2, 32, 16, 1, 512)
#undef ALADYN_TILE_CASE
  return hipErrorInvalidConfiguration;
}
template hipError_t launch_tile_batch<float>(const TileGeom&, int, const TileRun<float>*,
                                             const int2*, int, hipStream_t);
template hipError_t launch_tile_batch<double>(const TileGeom&, int, const TileRun<double>*,
                                              const int2*, int, hipStream_t);

template hipError_t launch_tile_kernel<float>(const TileGeom&, const DevGrid&,
                                              const StepConst<float>&, const float*,
                                              const RecPtr&, const RecPtr&, const int*, const int*,
                                              const int*, int*, int*, int*, double*, int*,
                                              hipStream_t);
template hipError_t launch_tile_kernel<double>(const TileGeom&, const DevGrid&,
                                               const StepConst<double>&, const double*,
                                               const RecPtr&, const RecPtr&, const int*,
                                               const int*, const int*, int*, int*, int*, double*,
                                               int*, hipStream_t);

}  // namespace aladyn
