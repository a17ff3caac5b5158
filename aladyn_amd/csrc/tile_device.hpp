// tile_device.hpp -- device code shared by the FP32 tile path kernels:
// particle records, the alternating-order gather, the FP32 exact-gamma push,
// the offset/carry position update and wave-aggregated slot reservation.
#pragma once

#include "lp_kernels.hpp"
#include "tile_store.hpp"

namespace aladyn {

// a reserved but unused slot of a tile segment
static constexpr int kHole = -1;
// particles counting-sorted per pass of the tile kernel (uint16 permutation)
static constexpr int kChunk = 4096;

struct Rec {
  int N[3];
  float s[3];
  float p[3];
  float w;
  int id;
};

struct StepConst {
  float alp, qch, dt;
  float inv[3];
  int nn[3];
  int initial;
  int64_t ovf_base;
};

// accessors of one field of a record
__device__ inline int rec_cell(const RecPtr& r, int64_t n) {
  return __float_as_int(rec_at(r, n)[0].x);
}
__device__ inline void set_rec_cell(const RecPtr& r, int64_t n, int cell) {
  reinterpret_cast<int*>(rec_at(r, n))[0] = cell;
}
__device__ inline float rec_s(const RecPtr& r, int64_t n, int d) {
  return reinterpret_cast<const float*>(rec_at(r, n))[1 + d];
}
__device__ inline float rec_p(const RecPtr& r, int64_t n, int d) {
  return reinterpret_cast<const float*>(rec_at(r, n))[4 + d];
}
__device__ inline float rec_w(const RecPtr& r, int64_t n) { return rec_at(r, n)[1].w; }

// The record with the packed cell left undecoded (q.N[0] holds the bits), so
// that a load can be issued one iteration ahead of its use.
template <int D>
__device__ inline void load_raw(const RecPtr& r, int64_t n, Rec& q) {
  const float4 a = rec_at(r, n)[0], b = rec_at(r, n)[1];
  q.N[0] = __float_as_int(a.x);
  q.s[0] = a.y;
  q.s[1] = a.z;
  q.s[2] = a.w;
  q.p[0] = b.x;
  q.p[1] = b.y;
  q.p[2] = b.z;
  q.w = b.w;
  q.id = r.id ? r.id[n] : 0;
}

template <int D>
__device__ inline void decode_cell(Rec& q, int bx, int by) {
  const int cell = q.N[0];
  q.N[0] = cell & ((1 << bx) - 1);
  q.N[1] = (cell >> bx) & ((1 << by) - 1);
  q.N[2] = D == 3 ? (cell >> (bx + by)) : 1;
}

template <int D>
__device__ inline void load_rec(const RecPtr& r, int64_t n, Rec& q, int bx, int by) {
  load_raw<D>(r, n, q);
  decode_cell<D>(q, bx, by);
}

template <int D>
__device__ inline void store_rec(const RecPtr& r, int64_t n, const Rec& q, int bx, int by) {
  const int cell = q.N[0] | (q.N[1] << bx) | (D == 3 ? (q.N[2] << (bx + by)) : 0);
  float4* o = rec_at(r, n);
  o[0] = make_float4(__int_as_float(cell), q.s[0], q.s[1], D == 3 ? q.s[2] : 0.f);
  o[1] = make_float4(q.p[0], q.p[1], D == 3 ? q.p[2] : 0.f, q.w);
  if (r.id) r.id[n] = q.id;
}

template <int D>
__device__ inline void copy_rec(const RecPtr& a, int64_t i, const RecPtr& b, int64_t o) {
  const float4 x = rec_at(a, i)[0], y = rec_at(a, i)[1];
  rec_at(b, o)[0] = x;
  rec_at(b, o)[1] = y;
  if (a.id && b.id) b.id[o] = a.id[i];
}

// Separable contraction sum_{i,j,k} wx[i] wy[j] wz[k] F(c, i, j, k):
// x first, then y, then z (NX*NY*NZ + NY*NZ + NZ multiply-adds instead of
// forming every 3-D weight product).
template <int NX, int NY, int NZ, class FA>
__device__ inline float contract(const FA& F, int c, const float* wx, const float* wy,
                                 const float* wz) {
  float r = 0.f;
#pragma unroll
  for (int k = 0; k < NZ; k++) {
    float u = 0.f;
#pragma unroll
    for (int j = 0; j < NY; j++) {
      float t = wx[0] * F(c, 0, j, k);
#pragma unroll
      for (int i = 1; i < NX; i++) t = fmaf(wx[i], F(c, i, j, k), t);
      u = j == 0 ? wy[0] * t : fmaf(wy[j], t, u);
    }
    r = k == 0 ? (NZ == 1 ? u : wz[0] * u) : fmaf(wz[k], u, r);
  }
  return r;
}

// Fields at the particle with the reference's alternating-order shapes
// (set_part2d_hcell_acc / set_part3d_hcell_acc, grid_part_connect.f90:89-285):
// quadratic weights on integer-node axes, linear on half-node axes
// (Ex: half x; Ey: half y; Ez: half z; Bx: half y,z; By: half x,z;
// Bz: half x,y).  F(c, i1, j1, k1) returns component c at node
// (N-1+i1, N-1+j1, N-1+k1).
template <int D, class FA>
__device__ inline void gather(const float s[3], const FA& F, float ef[6]) {
  float a[3][3], ah[3][2];
#pragma unroll
  for (int d = 0; d < D; d++) {
    qweights<float>(s[d], a[d]);
    lweights<float>(s[d], ah[d]);
  }
  if constexpr (D == 2) {
    const float one = 1.f;
    ef[0] = contract<2, 3, 1>(F, 0, ah[0], a[1], &one);
    ef[1] = contract<3, 2, 1>(F, 1, a[0], ah[1], &one);
    ef[2] = contract<2, 2, 1>(F, 2, ah[0], ah[1], &one);
  } else {
    ef[0] = contract<2, 3, 3>(F, 0, ah[0], a[1], a[2]);
    ef[1] = contract<3, 2, 3>(F, 1, a[0], ah[1], a[2]);
    ef[2] = contract<3, 3, 2>(F, 2, a[0], a[1], ah[2]);
    ef[3] = contract<3, 2, 2>(F, 3, a[0], ah[1], ah[2]);
    ef[4] = contract<2, 3, 2>(F, 4, ah[0], a[1], ah[2]);
    ef[5] = contract<2, 2, 3>(F, 5, ah[0], ah[1], a[2]);
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

// Slot reservation aggregated over a wavefront: among the active lanes, one
// atomicAdd on counters[key] per distinct key; returns this lane's slot.
__device__ inline int reserve_slot(int key, int* counters) {
  const int lane = (int)__lane_id();
  unsigned long long todo = __ballot(1);
  int slot = 0;
  while (todo) {
    const int leader = __ffsll((unsigned long long)todo) - 1;
    const int k = __shfl(key, leader);
    const unsigned long long same = __ballot(key == k);
    if (key == k) {
      int b = 0;
      if (lane == leader) b = atomicAdd(&counters[k], __popcll(same));
      b = __shfl(b, leader);
      slot = b + __popcll(same & ((1ull << lane) - 1ull));
    }
    todo &= ~same;
  }
  return slot;
}

// The same for one counter shared by all active lanes (LDS or global).
__device__ inline int reserve_one(int* counter) {
  const int lane = (int)__lane_id();
  const unsigned long long act = __ballot(1);
  const int leader = __ffsll((unsigned long long)act) - 1;
  int b = 0;
  if (lane == leader) b = atomicAdd(counter, __popcll(act));
  b = __shfl(b, leader);
  return b + __popcll(act & ((1ull << lane) - 1ull));
}

// Maximum of v over the wavefront.
__device__ inline int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

__device__ inline int tile_of(const TileGeom& tg, int cell, int D) {
  const int i = cell & ((1 << tg.bx) - 1);
  const int j = (cell >> tg.bx) & ((1 << tg.by) - 1);
  const int k = D == 3 ? (cell >> (tg.bx + tg.by)) : 3;
  return (i - 3) / tg.T[0] + tg.nt[0] * ((j - 3) / tg.T[1] + tg.nt[1] * (D == 3 ? (k - 3) / tg.T[2] : 0));
}

}  // namespace aladyn
