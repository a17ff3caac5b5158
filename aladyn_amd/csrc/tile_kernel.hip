// tile_kernel.hip -- the FP32 particle kernel of the tile path.
//
// One workgroup per tile of TX x TY x TZ node-cells and one thread per cell
// (the tile shape is a template parameter, so every LDS address below is a
// per-thread base plus an immediate offset).  Per workgroup:
//   1. stage the field patch (nodes T0-1 .. T0+T per axis, all components)
//      from HBM into LDS and zero an LDS current patch (nodes T0-2 .. T0+T+1,
//      the reach of the 5-slot Esirkepov window of grid_part_connect.f90:1008);
//   2. counting-sort the tile's particles by cell in LDS (chunks of kChunk);
//   3. each thread walks the particles of its own cell: gather from LDS,
//      charge multiply, [back-shift], exact-gamma push, carry, box test;
//      a particle that keeps its nearest node on every axis adds its
//      Esirkepov current to registers (54 values in 3-D, 12 in 2-D), one that
//      changes node goes through the union-window deposit into LDS;
//      survivors are written to the other buffer: stayers into slots
//      reserved for the chunk with one global atomic, movers into their new
//      tile with wave-aggregated atomics, overflow into the global list;
//   4. flush the register currents to LDS and the LDS patch to HBM.
// No global float atomics are used.
#include "tile_device.hpp"

namespace aladyn {

struct TileArgs {
  int nt0, nt1, bx, by;
  int n1p, n2p, n3p;
  int64_t ncell;
};

// Union-window Esirkepov deposit (grid_part_connect.f90:979-1060, 1168-1334)
// for a particle whose nearest node changed on at least one axis.  Per axis
// the old and new 3-node windows merge into 4 slots starting at
// min(N_old, N_new) - 1; the 3 fluxes between them are cumulative sums of
// W0 - W1 (the fourth vanishes by the partition of unity).  jb points at the
// LDS patch node of the union base; offsets are compile-time.
template <int D, int JX, int JY, int PVOL>
__device__ inline void deposit_general(float* jb, const float s0[3], const float s1[3],
                                       const int o0[3], const int o1[3], float wq) {
  float W0[3][4], W1[3][4];
#pragma unroll
  for (int d = 0; d < D; d++) {
    float a0[3], a1[3];
    qweights<float>(s0[d], a0);
    qweights<float>(s1[d], a1);
#pragma unroll
    for (int m = 0; m < 4; m++) {
      W0[d][m] = o0[d] == 0 ? (m < 3 ? a0[m] : 0.f) : (m > 0 ? a0[m - 1] : 0.f);
      W1[d][m] = o1[d] == 0 ? (m < 3 ? a1[m] : 0.f) : (m > 0 ? a1[m - 1] : 0.f);
    }
  }
  float F[3][3];
#pragma unroll
  for (int d = 0; d < D; d++) {
    float c = 0.f;
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
        atomicAdd(jb + (f + JX * m), F[0][f] * (W0[1][m] + W1[1][m]));
        atomicAdd(jb + PVOL + (m + JX * f), F[1][f] * (W0[0][m] + W1[0][m]));
      }
  } else {
    // J_x: flux_x (x) [W0z (W0y + W1y/2) + W1z (W1y + W0y/2)], etc.
#pragma unroll
    for (int m = 0; m < 4; m++) {
      const float y0 = W0[1][m] + 0.5f * W1[1][m], y1 = W1[1][m] + 0.5f * W0[1][m];
      const float x0 = W0[0][m] + 0.5f * W1[0][m], x1 = W1[0][m] + 0.5f * W0[0][m];
#pragma unroll
      for (int l = 0; l < 4; l++) {
        const float tyz = y0 * W0[2][l] + y1 * W1[2][l];
        const float txz = x0 * W0[2][l] + x1 * W1[2][l];
        const float txy = x0 * W0[1][l] + x1 * W1[1][l];
#pragma unroll
        for (int f = 0; f < 3; f++) {
          atomicAdd(jb + (f + JX * (m + JY * l)), F[0][f] * tyz);
          atomicAdd(jb + PVOL + (m + JX * (f + JY * l)), F[1][f] * txz);
          atomicAdd(jb + 2 * PVOL + (m + JX * (l + JY * f)), F[2][f] * txy);
        }
      }
    }
  }
}

template <int D, int TX, int TY, int TZ>
__global__ __launch_bounds__(TX* TY* TZ) void k_tile(
    TileArgs ta, StepConst sc, const float* __restrict__ ebf, RecPtr A, RecPtr B,
    const int* __restrict__ seg_start, const int* __restrict__ seg_cap,
    const int* __restrict__ cnt_cur, int* __restrict__ cnt_next, int* __restrict__ ovf_cur,
    int* __restrict__ ovf_next, float* __restrict__ patch) {
  constexpr int NCL = TX * TY * TZ, NF = D == 3 ? 6 : 3;
  constexpr int GX = TX + 2, GY = TY + 2, GZ = D == 3 ? TZ + 2 : 1, GVOL = GX * GY * GZ;
  constexpr int JX = TX + 4, JY = TY + 4, JZ = D == 3 ? TZ + 4 : 1, PVOL = JX * JY * JZ;
  extern __shared__ float lds[];
  float* Fl = lds;
  float* Jl = Fl + NF * GVOL;
  unsigned short* perm = (unsigned short*)(Jl + D * PVOL);
  int* offs = (int*)(perm + kChunk);  // NCL + 1
  int* fillc = offs + NCL + 1;        // NCL
  int* misc = fillc + NCL;            // [0] chunk slot base, [1] stayers

  const int t = blockIdx.x;
  const int tx = t % ta.nt0, ty = (t / ta.nt0) % ta.nt1, tz = t / (ta.nt0 * ta.nt1);
  const int T0x = 3 + tx * TX, T0y = 3 + ty * TY, T0z = D == 3 ? 3 + tz * TZ : 1;
  // field patch: nodes T0-1 .. T0+T
  for (int q = threadIdx.x; q < NF * GVOL; q += NCL) {
    const int c = q / GVOL, r = q - c * GVOL;
    const int li = r % GX, lj = (r / GX) % GY, lk = r / (GX * GY);
    const int i = T0x - 1 + li, j = T0y - 1 + lj, k = D == 3 ? T0z - 1 + lk : 1;
    float v = 0.f;
    if (i <= ta.n1p && j <= ta.n2p && k <= ta.n3p)
      v = ebf[(int64_t)c * ta.ncell + (i - 1) + (int64_t)ta.n1p * ((j - 1) + (int64_t)ta.n2p * (k - 1))];
    Fl[q] = v;
  }
  for (int q = threadIdx.x; q < D * PVOL; q += NCL) Jl[q] = 0.f;

  const int me = threadIdx.x;
  const int cx = me % TX, cy = (me / TX) % TY, cz = D == 3 ? me / (TX * TY) : 0;
  const int Nc[3] = {T0x + cx, T0y + cy, D == 3 ? T0z + cz : 1};
  // LDS field patch base at node N-1 of this cell, current patch base at N-2
  const float* fb = Fl + cx + GX * (cy + GY * cz);
  float* jb2 = Jl + cx + JX * (cy + JY * cz);
  const auto F = [&](int c, int i1, int j1, int k1) -> float {
    return fb[c * GVOL + i1 + GX * (j1 + GY * k1)];
  };
  // register currents of non-crossing particles, node offsets from N-1
  float jx[2][3][3], jy[3][2][3], jz[3][3][2];
#pragma unroll
  for (int a = 0; a < 18; a++) {
    (&jx[0][0][0])[a] = 0.f;
    (&jy[0][0][0])[a] = 0.f;
    (&jz[0][0][0])[a] = 0.f;
  }
  const int cap_t = seg_cap[t];
  const int cnt = min(cnt_cur[t], cap_t);
  const int64_t base = seg_start[t];
  const int mi = (1 << ta.bx) - 1, mj = (1 << ta.by) - 1, shk = ta.bx + ta.by;

  for (int c0 = 0; c0 < cnt; c0 += kChunk) {
    const int m = min(kChunk, cnt - c0);
    __syncthreads();  // patches staged / previous chunk consumed
    offs[me] = 0;
    __syncthreads();
    for (int q = me; q < m; q += NCL) {
      const int cell = A.cell[base + c0 + q];
      if (cell == kHole) continue;
      const int li = (cell & mi) - T0x, lj = ((cell >> ta.bx) & mj) - T0y;
      const int lk = D == 3 ? (cell >> shk) - T0z : 0;
      if (li >= 0 && li < TX && lj >= 0 && lj < TY && lk >= 0 && lk < TZ)
        atomicAdd(&offs[li + TX * (lj + TY * lk)], 1);
    }
    __syncthreads();
    {  // exclusive scan of offs[0..NCL) (Hillis-Steele through fillc)
      const int v = offs[me];
      fillc[me] = v;
      __syncthreads();
#pragma unroll 1
      for (int off = 1; off < NCL; off <<= 1) {
        const int add = me >= off ? fillc[me - off] : 0;
        __syncthreads();
        fillc[me] += add;
        __syncthreads();
      }
      offs[me] = fillc[me] - v;
      if (me == NCL - 1) {
        offs[NCL] = fillc[me];
        // slots for every valid particle of the chunk in this tile
        misc[0] = atomicAdd(&cnt_next[t], fillc[me]);
        misc[1] = 0;
      }
      __syncthreads();
      fillc[me] = 0;
      __syncthreads();
    }
    for (int q = me; q < m; q += NCL) {
      const int cell = A.cell[base + c0 + q];
      if (cell == kHole) continue;
      const int li = (cell & mi) - T0x, lj = ((cell >> ta.bx) & mj) - T0y;
      const int lk = D == 3 ? (cell >> shk) - T0z : 0;
      if (li >= 0 && li < TX && lj >= 0 && lj < TY && lk >= 0 && lk < TZ) {
        const int lc = li + TX * (lj + TY * lk);
        perm[offs[lc] + atomicAdd(&fillc[lc], 1)] = (unsigned short)q;
      } else {
        // not in this tile (the store never does this): hand it to the
        // global-memory kernel that runs after this one
        copy_rec<D>(A, base + c0 + q, A, sc.ovf_base + atomicAdd(ovf_cur, 1));
      }
    }
    __syncthreads();
    const int chunk_base = misc[0];
    const int p_end = offs[me + 1];
    for (int pi = offs[me]; pi < p_end; pi++) {
      Rec q;
      load_rec<D>(A, base + c0 + perm[pi], q, ta.bx, ta.by);
      float ef[6];
      gather<D>(q.s, F, ef);
      const float so[3] = {q.s[0], q.s[1], q.s[2]};
      const float dtg = push<D>(q.p, ef, sc.qch, sc.alp, sc.dt, sc.initial);
      if (!move_and_test<D>(q, dtg, sc.inv, sc.nn)) continue;
      const float wq = sc.qch * q.w;
      bool cross = q.N[0] != Nc[0] || q.N[1] != Nc[1];
      if (D == 3) cross = cross || q.N[2] != Nc[2];
      if (!cross) {
        float a0[3][3], a1[3][3], fl[3][2];
#pragma unroll
        for (int d = 0; d < D; d++) {
          qweights<float>(so[d], a0[d]);
          qweights<float>(q.s[d], a1[d]);
          // W0[0]-W1[0] = (s1-s0)(1-s0-s1)/2, -(W0[2]-W1[2]) = (s1-s0)(1+s0+s1)/2
          const float h = 0.5f * wq * (q.s[d] - so[d]);
          fl[d][0] = h * ((1.f - so[d]) - q.s[d]);
          fl[d][1] = h * ((1.f + so[d]) + q.s[d]);
        }
        if constexpr (D == 2) {
#pragma unroll
          for (int mm = 0; mm < 3; mm++) {
            const float ty2 = a0[1][mm] + a1[1][mm];
            const float tx2 = a0[0][mm] + a1[0][mm];
#pragma unroll
            for (int f = 0; f < 2; f++) {
              jx[f][mm][0] += fl[0][f] * ty2;
              jy[mm][f][0] += fl[1][f] * tx2;
            }
          }
        } else {
#pragma unroll
          for (int mm = 0; mm < 3; mm++) {
            const float y0 = a0[1][mm] + 0.5f * a1[1][mm], y1 = a1[1][mm] + 0.5f * a0[1][mm];
            const float x0 = a0[0][mm] + 0.5f * a1[0][mm], x1 = a1[0][mm] + 0.5f * a0[0][mm];
#pragma unroll
            for (int l = 0; l < 3; l++) {
              const float tyz = y0 * a0[2][l] + y1 * a1[2][l];
              const float txz = x0 * a0[2][l] + x1 * a1[2][l];
              const float txy = x0 * a0[1][l] + x1 * a1[1][l];
#pragma unroll
              for (int f = 0; f < 2; f++) {
                jx[f][mm][l] += fl[0][f] * tyz;
                jy[mm][f][l] += fl[1][f] * txz;
                jz[mm][l][f] += fl[2][f] * txy;
              }
            }
          }
        }
      } else {
        int o0[3] = {0, 0, 0}, o1[3] = {0, 0, 0}, sh[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < D; d++) {
          const int sig = q.N[d] - Nc[d];  // -1, 0, +1
          o0[d] = sig < 0 ? 1 : 0;
          o1[d] = sig > 0 ? 1 : 0;
          sh[d] = sig < 0 ? -1 : 0;  // union base = N-1 + min(0, sig) -> patch N-2 offset
        }
        float* ub = jb2 + (1 + sh[0]) + JX * ((1 + sh[1]) + JY * (D == 3 ? 1 + sh[2] : 0));
        deposit_general<D, JX, JY, PVOL>(ub, so, q.s, o0, o1, wq);
      }
      // write the survivor into the other buffer
      const int dtile = (q.N[0] - 3) / TX +
                        ta.nt0 * ((q.N[1] - 3) / TY + ta.nt1 * (D == 3 ? (q.N[2] - 3) / TZ : 0));
      int64_t dst;
      if (dtile == t) {
        const int slot = chunk_base + reserve_one(&misc[1]);
        dst = slot < cap_t ? base + slot : sc.ovf_base + atomicAdd(ovf_next, 1);
      } else {
        const int slot = reserve_slot(dtile, cnt_next);
        dst = slot < seg_cap[dtile] ? (int64_t)seg_start[dtile] + slot
                                    : sc.ovf_base + atomicAdd(ovf_next, 1);
      }
      store_rec<D>(B, dst, q, ta.bx, ta.by);
    }
    __syncthreads();
    // reserved slots not taken by stayers become holes
    {
      const int nstay = misc[1], nval = offs[NCL];
      for (int q = nstay + me; q < nval; q += NCL) {
        const int slot = chunk_base + q;
        if (slot < cap_t) B.cell[base + slot] = kHole;
      }
    }
  }
  // flush the register currents (node N-1 is patch offset +1 from jb2)
  {
    float* j1 = jb2 + 1 + JX * (1 + JY * (D == 3 ? 1 : 0));
    if constexpr (D == 2) {
#pragma unroll
      for (int f = 0; f < 2; f++)
#pragma unroll
        for (int mm = 0; mm < 3; mm++) {
          atomicAdd(j1 + (f + JX * mm), jx[f][mm][0]);
          atomicAdd(j1 + PVOL + (mm + JX * f), jy[mm][f][0]);
        }
    } else {
#pragma unroll
      for (int mm = 0; mm < 3; mm++)
#pragma unroll
        for (int l = 0; l < 3; l++)
#pragma unroll
          for (int f = 0; f < 2; f++) {
            atomicAdd(j1 + (f + JX * (mm + JY * l)), jx[f][mm][l]);
            atomicAdd(j1 + PVOL + (mm + JX * (f + JY * l)), jy[mm][f][l]);
            atomicAdd(j1 + 2 * PVOL + (mm + JX * (l + JY * f)), jz[mm][l][f]);
          }
    }
  }
  __syncthreads();
  float* P = patch + (int64_t)t * D * PVOL;
  for (int q = threadIdx.x; q < D * PVOL; q += NCL) P[q] = Jl[q];
}

template <int D, int TX, int TY, int TZ>
static int lds_bytes_for() {
  constexpr int NCL = TX * TY * TZ, NF = D == 3 ? 6 : 3;
  constexpr int GVOL = (TX + 2) * (TY + 2) * (D == 3 ? TZ + 2 : 1);
  constexpr int PVOL = (TX + 4) * (TY + 4) * (D == 3 ? TZ + 4 : 1);
  return (NF * GVOL + D * PVOL) * 4 + kChunk * 2 + (2 * NCL + 3) * 4;
}

template <int D, int TX, int TY, int TZ>
static hipError_t launch(const TileGeom& tg, const TileArgs& ta, const StepConst& sc,
                         const float* ebf, const RecPtr& A, const RecPtr& B, const int* ss,
                         const int* sc_, const int* cc, int* cn, int* oc, int* on, float* patch,
                         hipStream_t s) {
  k_tile<D, TX, TY, TZ><<<tg.ntiles, TX * TY * TZ, lds_bytes_for<D, TX, TY, TZ>(), s>>>(
      ta, sc, ebf, A, B, ss, sc_, cc, cn, oc, on, patch);
  return hipGetLastError();
}

// shapes with a compiled kernel; tile_geom_init accepts only these
bool tile_shape_supported(int D, int tx, int ty, int tz) {
  if (D == 3) return (tx == 8 && ty == 8 && tz == 8) || (tx == 8 && ty == 8 && tz == 4) ||
                     (tx == 16 && ty == 8 && tz == 4);
  return (tx == 16 && ty == 16) || (tx == 32 && ty == 16);
}

int tile_lds_bytes(int D, int tx, int ty, int tz) {
  if (D == 3) {
    if (tx == 8 && ty == 8 && tz == 8) return lds_bytes_for<3, 8, 8, 8>();
    if (tx == 8 && ty == 8 && tz == 4) return lds_bytes_for<3, 8, 8, 4>();
    return lds_bytes_for<3, 16, 8, 4>();
  }
  if (tx == 16) return lds_bytes_for<2, 16, 16, 1>();
  return lds_bytes_for<2, 32, 16, 1>();
}

hipError_t launch_tile_kernel(const TileGeom& tg, const DevGrid& g, const StepConst& sc,
                              const float* ebf, const RecPtr& A, const RecPtr& B,
                              const int* seg_start, const int* seg_cap, const int* cnt_cur,
                              int* cnt_next, int* ovf_cur, int* ovf_next, float* patch,
                              hipStream_t s) {
  TileArgs ta{tg.nt[0], tg.nt[1], tg.bx, tg.by, g.n1p, g.n2p, g.n3p, g.ncell};
  const int* T = tg.T;
#define ALADYN_TILE_CASE(DD, X, Y, Z)                                                        \
  if (g.ndim == DD && T[0] == X && T[1] == Y && (DD == 2 || T[2] == Z))                      \
    return launch<DD, X, Y, Z>(tg, ta, sc, ebf, A, B, seg_start, seg_cap, cnt_cur, cnt_next, \
                               ovf_cur, ovf_next, patch, s);
  ALADYN_TILE_CASE(3, 8, 8, 8)
  ALADYN_TILE_CASE(3, 8, 8, 4)
  ALADYN_TILE_CASE(3, 16, 8, 4)
  ALADYN_TILE_CASE(2, 16, 16, 1)
  ALADYN_TILE_CASE(2, 32, 16, 1)
#undef ALADYN_TILE_CASE
  return hipErrorInvalidConfiguration;
}

}  // namespace aladyn
