// tile_device.hpp -- device code shared by the tile path kernels (FP32 and
// FP64): particle records, the alternating-order gather, the exact-gamma
// push, the offset/carry position update and wave-aggregated slot
// reservation.
#pragma once

#include "lp_kernels.hpp"
#include "tile_store.hpp"

namespace aladyn {

// a reserved but unused slot of a tile segment
static constexpr int kHole = -1;
// particles counting-sorted per pass of the tile kernel (uint16 permutation)
static constexpr int kChunk = 4096;

// Real-type helpers: R is the precision of the particle state and fields
// (float for the FP32 path, double for the FP64 path).
// FP64: correctly rounded.  FP32: the hardware square root, reciprocal and
// reciprocal square root (v_sqrt_f32, v_rcp_f32, v_rsq_f32, 1 ulp).  The
// push's arguments are 1 + p^2, gamma^2 + b^2 (>= 1) and
// (1 + u^2 - b^2)^2 + 4(b^2 + (b.u)^2) (>= 1/4), far from denormals.
template <typename R> __device__ inline R rsqrt_of(R x) { return R(1) / sqrt(x); }
template <> __device__ inline float rsqrt_of<float>(float x) { return __builtin_amdgcn_rsqf(x); }
template <typename R> __device__ inline R sqrt_of(R x) { return sqrt(x); }
template <> __device__ inline float sqrt_of<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <typename R> __device__ inline R rcp_of(R x) { return R(1) / x; }
template <> __device__ inline float rcp_of<float>(float x) { return __builtin_amdgcn_rcpf(x); }
// a / sqrt(x)
template <typename R> __device__ inline R div_sqrt(R a, R x) { return a / sqrt(x); }
template <> __device__ inline float div_sqrt<float>(float a, float x) { return a * __builtin_amdgcn_rsqf(x); }
template <typename R> __device__ inline R fma_of(R a, R b, R c) { return fma(a, b, c); }
template <> __device__ inline float fma_of<float>(float a, float b, float c) { return fmaf(a, b, c); }

template <typename R>
struct Rec {
  int N[3];
  R s[3];
  R p[3];
  float w;  // the reference's FP32 weight (precision_def.F90:33-40)
  int id;
};

// Record layouts (the RecPtr comment in tile_store.hpp):
//   float : 2 x float4  = [cell, s0, s1, s2 | p0, p1, p2, w]            32 B
//   double: 4 x double2 = [cell w | s0 | s1 | s2 | p0 | p1 | p2 | 0]     64 B
// (cell and w share the first 8 bytes of the double record).
template <typename R> struct RecWord;
template <> struct RecWord<float> { using V = float4; static constexpr int N = 2; };
template <> struct RecWord<double> { using V = double2; static constexpr int N = 4; };

template <typename R>
__host__ __device__ inline typename RecWord<R>::V* rec_at(const RecPtr& a, int64_t n) {
  return static_cast<typename RecWord<R>::V*>(a.r) + RecWord<R>::N * n;
}

template <typename R>
__device__ inline int rec_cell(const RecPtr& r, int64_t n) {
  return reinterpret_cast<const int*>(rec_at<R>(r, n))[0];
}
template <typename R>
__device__ inline void set_rec_cell(const RecPtr& r, int64_t n, int cell) {
  reinterpret_cast<int*>(rec_at<R>(r, n))[0] = cell;
}
template <typename R>
__device__ inline R rec_s(const RecPtr& r, int64_t n, int d) {
  if constexpr (sizeof(R) == 4) return reinterpret_cast<const float*>(rec_at<R>(r, n))[1 + d];
  else return reinterpret_cast<const double*>(rec_at<R>(r, n))[1 + d];
}
template <typename R>
__device__ inline R rec_p(const RecPtr& r, int64_t n, int d) {
  if constexpr (sizeof(R) == 4) return reinterpret_cast<const float*>(rec_at<R>(r, n))[4 + d];
  else return reinterpret_cast<const double*>(rec_at<R>(r, n))[4 + d];
}
template <typename R>
__device__ inline float rec_w(const RecPtr& r, int64_t n) {
  if constexpr (sizeof(R) == 4) return rec_at<R>(r, n)[1].w;
  else return reinterpret_cast<const float*>(rec_at<R>(r, n))[1];
}

// The record with the packed cell left undecoded (q.N[0] holds the bits), so
// that a load can be issued one iteration ahead of its use.
template <int D, typename R>
__device__ inline void load_raw(const RecPtr& r, int64_t n, Rec<R>& q) {
  if constexpr (sizeof(R) == 4) {
    const float4 a = rec_at<R>(r, n)[0], b = rec_at<R>(r, n)[1];
    q.N[0] = __float_as_int(a.x);
    q.s[0] = a.y;
    q.s[1] = a.z;
    q.s[2] = a.w;
    q.p[0] = b.x;
    q.p[1] = b.y;
    q.p[2] = b.z;
    q.w = b.w;
  } else {
    const double2* v = rec_at<R>(r, n);
    const double2 a = v[0], b = v[1], c = v[2], e = v[3];
    const long long cw = __double_as_longlong(a.x);
    q.N[0] = (int)(cw & 0xffffffffll);
    q.w = __int_as_float((int)(cw >> 32));
    q.s[0] = a.y;
    q.s[1] = b.x;
    q.s[2] = b.y;
    q.p[0] = c.x;
    q.p[1] = c.y;
    q.p[2] = e.x;
  }
  q.id = r.id ? r.id[n] : 0;
}

template <int D, typename R>
__device__ inline void decode_cell(Rec<R>& q, int bx, int by) {
  const int cell = q.N[0];
  q.N[0] = cell & ((1 << bx) - 1);
  q.N[1] = (cell >> bx) & ((1 << by) - 1);
  q.N[2] = D == 3 ? (cell >> (bx + by)) : 1;
}

template <int D, typename R>
__device__ inline void load_rec(const RecPtr& r, int64_t n, Rec<R>& q, int bx, int by) {
  load_raw<D, R>(r, n, q);
  decode_cell<D, R>(q, bx, by);
}

// the words of a record from its fields (packed cell bits given)
template <int D, typename R>
__device__ inline void rec_words(int cell, const R s[3], const R p[3], float w,
                                 typename RecWord<R>::V o[RecWord<R>::N]) {
  if constexpr (sizeof(R) == 4) {
    o[0] = make_float4(__int_as_float(cell), s[0], s[1], D == 3 ? s[2] : 0.f);
    o[1] = make_float4(p[0], p[1], D == 3 ? p[2] : 0.f, w);
  } else {
    const long long cw = (long long)(unsigned)cell | ((long long)__float_as_int(w) << 32);
    o[0] = make_double2(__longlong_as_double(cw), s[0]);
    o[1] = make_double2(s[1], D == 3 ? s[2] : 0.0);
    o[2] = make_double2(p[0], p[1]);
    o[3] = make_double2(D == 3 ? p[2] : 0.0, 0.0);
  }
}

// writes a record from its fields (packed cell bits given)
template <int D, typename R>
__device__ inline void put_rec(const RecPtr& r, int64_t n, int cell, const R s[3], const R p[3],
                               float w) {
  typename RecWord<R>::V o[RecWord<R>::N];
  rec_words<D, R>(cell, s, p, w, o);
  typename RecWord<R>::V* d = rec_at<R>(r, n);
#pragma unroll
  for (int k = 0; k < RecWord<R>::N; k++) d[k] = o[k];
}

template <int D>
__device__ inline int pack_cell(const int N[3], int bx, int by) {
  return N[0] | (N[1] << bx) | (D == 3 ? (N[2] << (bx + by)) : 0);
}

template <int D, typename R>
__device__ inline void store_rec(const RecPtr& r, int64_t n, const Rec<R>& q, int bx, int by) {
  const int cell = q.N[0] | (q.N[1] << bx) | (D == 3 ? (q.N[2] << (bx + by)) : 0);
  put_rec<D, R>(r, n, cell, q.s, q.p, q.w);
  if (r.id) r.id[n] = q.id;
}

template <typename R>
__device__ inline void copy_rec(const RecPtr& a, int64_t i, const RecPtr& b, int64_t o) {
#pragma unroll
  for (int k = 0; k < RecWord<R>::N; k++) rec_at<R>(b, o)[k] = rec_at<R>(a, i)[k];
  if (a.id && b.id) b.id[o] = a.id[i];
}

// Separable contraction sum_{i,j,k} wx[i] wy[j] wz[k] F(c, i, j, k):
// x first, then y, then z (NX*NY*NZ + NY*NZ + NZ multiply-adds instead of
// forming every 3-D weight product).
template <int NX, int NY, int NZ, typename R, class FA>
__device__ inline R contract(const FA& F, int c, const R* wx, const R* wy, const R* wz) {
  R r = R(0);
#pragma unroll
  for (int k = 0; k < NZ; k++) {
    R u = R(0);
#pragma unroll
    for (int j = 0; j < NY; j++) {
      R t = wx[0] * F(c, 0, j, k);
#pragma unroll
      for (int i = 1; i < NX; i++) t = fma_of<R>(wx[i], F(c, i, j, k), t);
      u = j == 0 ? wy[0] * t : fma_of<R>(wy[j], t, u);
    }
    r = k == 0 ? (NZ == 1 ? u : wz[0] * u) : fma_of<R>(wz[k], u, r);
  }
  return r;
}

// Fields at the particle with the reference's alternating-order shapes
// (set_part2d_hcell_acc / set_part3d_hcell_acc, grid_part_connect.f90:89-285):
// quadratic weights on integer-node axes, linear on half-node axes
// (Ex: half x; Ey: half y; Ez: half z; Bx: half y,z; By: half x,z;
// Bz: half x,y).  F(c, i1, j1, k1) returns component c at node
// (N-1+i1, N-1+j1, N-1+k1); a[d] receives the quadratic weights of s[d].
template <int D, typename R, class FA>
__device__ inline void gather(const R s[3], const FA& F, R ef[6], R a[3][3]) {
  R ah[3][2];
#pragma unroll
  for (int d = 0; d < D; d++) {
    qweights<R>(s[d], a[d]);
    lweights<R>(s[d], ah[d]);
  }
  if constexpr (D == 2) {
    const R one = R(1);
    ef[0] = contract<2, 3, 1, R>(F, 0, ah[0], a[1], &one);
    ef[1] = contract<3, 2, 1, R>(F, 1, a[0], ah[1], &one);
    ef[2] = contract<2, 2, 1, R>(F, 2, ah[0], ah[1], &one);
  } else {
    ef[0] = contract<2, 3, 3, R>(F, 0, ah[0], a[1], a[2]);
    ef[1] = contract<3, 2, 3, R>(F, 1, a[0], ah[1], a[2]);
    ef[2] = contract<3, 3, 2, R>(F, 2, a[0], a[1], ah[2]);
    ef[3] = contract<3, 2, 2, R>(F, 3, a[0], ah[1], ah[2]);
    ef[4] = contract<2, 3, 2, R>(F, 4, ah[0], a[1], ah[2]);
    ef[5] = contract<2, 2, 3, R>(F, 5, ah[0], ah[1], a[2]);
  }
}

// The same without the quadratic weights (a[d] = the shape of s[d]).
template <int D, typename R, class FA>
__device__ inline void gather(const R s[3], const FA& F, R ef[6]) {
  R a[3][3];
  gather<D, R>(s, F, ef, a);
}

// Charge multiply + [init_lpf_momenta] + lpf_momenta_and_positions
// (curr_and_fields_util.f90:58-72, boris_push.f90:34-168).  Returns
// dt/gamma^{n+1/2}.
template <int D, typename R>
__device__ inline R push(R p[3], R ef[6], R qch, R alp, R dt, int initial) {
  const int nf = D == 2 ? 3 : 6;
  const R one = R(1), half = R(0.5), two = R(2);
#pragma unroll
  for (int c = 0; c < nf; c++) ef[c] = qch * ef[c];
  if constexpr (D == 2) {
    if (initial) {
      const R e1 = -alp * ef[0], e2 = -alp * ef[1], e3 = -alp * ef[2];
      const R gi = rsqrt_of<R>(one + (p[0] * p[0] + p[1] * p[1]));
      const R v1 = p[0] * gi, v2 = p[1] * gi;
      p[0] = p[0] + e1 + v2 * e3;
      p[1] = p[1] + e2 - v1 * e3;
    }
    const R v1 = p[0] + alp * ef[0], v2 = p[1] + alp * ef[1], v3 = alp * ef[2];
    const R b2 = v3 * v3;
    const R gam02 = (one + (v1 * v1 + v2 * v2)) - b2;
    const R gam2 = half * (gam02 + sqrt_of<R>(gam02 * gam02 + R(4) * b2));
    const R gam = sqrt_of<R>(gam2);
    const R r = rcp_of<R>(gam2 + b2);
    const R h1 = (gam2 * v1 + gam * v2 * v3) * r;
    const R h2 = (gam2 * v2 - gam * v1 * v3) * r;
    p[0] = two * h1 - p[0];
    p[1] = two * h2 - p[1];
    return div_sqrt<R>(dt, one + (p[0] * p[0] + p[1] * p[1]));
  } else {
    if (initial) {
      R e[6];
#pragma unroll
      for (int c = 0; c < 6; c++) e[c] = -alp * ef[c];
      const R gi = rsqrt_of<R>(one + (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]));
      const R v1 = gi * p[0], v2 = gi * p[1], v3 = gi * p[2];
      p[0] = p[0] + e[0] + v2 * e[5] - v3 * e[4];
      p[1] = p[1] + e[1] + v3 * e[3] - v1 * e[5];
      p[2] = p[2] + e[2] + v1 * e[4] - v2 * e[3];
    }
    const R vp[3] = {p[0] + alp * ef[0], p[1] + alp * ef[1], p[2] + alp * ef[2]};
    const R bb[3] = {alp * ef[3], alp * ef[4], alp * ef[5]};
    const R b2 = bb[0] * bb[0] + bb[1] * bb[1] + bb[2] * bb[2];
    const R bv = bb[0] * vp[0] + bb[1] * vp[1] + bb[2] * vp[2];
    const R gam02 = (one + (vp[0] * vp[0] + vp[1] * vp[1] + vp[2] * vp[2])) - b2;
    const R gam2 = half * (gam02 + sqrt_of<R>(gam02 * gam02 + R(4) * (b2 + bv * bv)));
    const R gam = sqrt_of<R>(gam2);
    const R r = rcp_of<R>(b2 + gam2);
    R vph[3];
    vph[0] = gam2 * vp[0] + bb[0] * bv + gam * (vp[1] * bb[2] - vp[2] * bb[1]);
    vph[1] = gam2 * vp[1] + bb[1] * bv + gam * (vp[2] * bb[0] - vp[0] * bb[2]);
    vph[2] = gam2 * vp[2] + bb[2] * bv + gam * (vp[0] * bb[1] - vp[1] * bb[0]);
#pragma unroll
    for (int c = 0; c < 3; c++) p[c] = two * (vph[c] * r) - p[c];
    return div_sqrt<R>(dt, one + (p[0] * p[0] + p[1] * p[1] + p[2] * p[2]));
  }
}

// x^{n+1} = x^n + dt v: offset update with an exact integer carry; the box
// test is (x - gmin) > 0 and (x - gmin) <= n*dx, i.e. node 3 + (0, n] in
// local units (mpi_part_interface.f90:494-498), evaluated without rounding.
template <int D, typename R>
__device__ inline bool move_and_test(Rec<R>& q, R dtg, const R inv[3], const int nn[3]) {
  bool keep = true;
  const R half = R(0.5), one = R(1), zero = R(0);
#pragma unroll
  for (int d = 0; d < D; d++) {
    R s = q.s[d] + (dtg * inv[d]) * q.p[d];
    int N = q.N[d];
    if (s >= half) {
      N += 1;
      s -= one;
    } else if (s < -half) {
      N -= 1;
      s += one;
    }
    q.N[d] = N;
    q.s[d] = s;
    const bool fin = fabs(s) <= half;  // false for NaN
    const bool lo = (N > 3) || (N == 3 && s > zero);
    const bool hi = (N < nn[d] + 3) || (N == nn[d] + 3 && s <= zero);
    keep = keep && fin && lo && hi;
  }
  return keep;
}

// Slot reservation aggregated over a wavefront: among the active lanes, one
// atomicAdd on counters[key] per distinct key; returns this lane's slot.
__device__ inline int reserve_slot(int key, int* counters) {
  const int lane = (int)__lane_id();
  // group the active lanes by key with ballots only (no memory traffic)
  unsigned long long todo = __ballot(1);
  int leader = lane, rank = 0, count = 1;
  while (todo) {
    const int l = __ffsll((unsigned long long)todo) - 1;
    const int k = __shfl(key, l);
    const unsigned long long same = __ballot(key == k);
    if (key == k) {
      leader = l;
      rank = __popcll(same & ((1ull << lane) - 1ull));
      count = __popcll(same);
    }
    todo &= ~same;
  }
  // one atomic per group, all groups in the same instruction
  int b = 0;
  if (lane == leader) b = atomicAdd(&counters[key], count);
  return __shfl(b, leader) + rank;
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

// Sum of v over the wavefront (every lane active).
__device__ inline int wave_sum(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// qqh_*_spline for one axis (grid_part_lib.f90:145-210, 261-330) from the
// nearest node N and offset s (xx = N + s, s in [-1/2, 1/2)): quadratic
// weights at integer nodes from node N-1, and at half-integer nodes from
// node ih = int(xx) - 1, i.e. N-2 for s < 0 and N-1 otherwise, with offset
// xx - int(xx) - 1/2 = s + 1/2 or s - 1/2.  hb = ih - (N-1).
template <typename R>
__device__ inline void qqh(R s, R a[3], R ah[3], int& hb) {
  qweights<R>(s, a);
  const bool neg = s < R(0);
  qweights<R>(neg ? s + R(0.5) : s - R(0.5), ah);
  hb = neg ? -1 : 0;
}

// sum_{k1,j1,i1} wx[i1] wy[j1] wz[k1] f(i0+i1, j0+j1, k0+k1) over an LDS box
// (x stride 1, y stride GX, z stride GXY), in the reference's loop order
// (x innermost, dvol = wy*wz first)
template <int D, typename R>
__device__ inline R contract3l(const R* f, int GX, int GXY, const R wx[3], const R wy[3],
                               const R wz[3], int i0, int j0, int k0) {
  R r = R(0);
  const int nk = D == 3 ? 3 : 1;
  for (int k1 = 0; k1 < nk; k1++)
#pragma unroll
    for (int j1 = 0; j1 < 3; j1++) {
      const R dvol = D == 3 ? wy[j1] * wz[k1] : wy[j1];
      const R* row = f + i0 + GX * (j0 + j1) + GXY * (k0 + k1);
#pragma unroll
      for (int i1 = 0; i1 < 3; i1++) r = r + (dvol * wx[i1]) * row[i1];
    }
  return r;
}

__device__ inline int tile_of(const TileGeom& tg, int cell, int D) {
  const int i = cell & ((1 << tg.bx) - 1);
  const int j = (cell >> tg.bx) & ((1 << tg.by) - 1);
  const int k = D == 3 ? (cell >> (tg.bx + tg.by)) : 3;
  return (i - 3) / tg.T[0] + tg.nt[0] * ((j - 3) / tg.T[1] + tg.nt[1] * (D == 3 ? (k - 3) / tg.T[2] : 0));
}

}  // namespace aladyn
