// tile_store.hip -- host side of the FP32 tile path (see tile_store.hpp):
// geometry, the per-step launch sequence, the fold of the per-tile current
// patches, the global-memory kernel for overflowed particles, the re-layout
// (counting sort with fresh capacities), transfers and FP64 diagnostics.
#include <algorithm>
#include <climits>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>

#include "tile_device.hpp"

namespace aladyn {

// tile_kernel.hip
bool tile_shape_supported(int D, int tx, int ty, int tz, int nt, int real_bytes);
int tile_lds_bytes(int D, int tx, int ty, int tz, int nt, int real_bytes);
template <typename R>
hipError_t launch_tile_batch(const TileGeom& tg, int ndim, const TileRun<R>* runs,
                             const int2* work, int nwork, hipStream_t s);
template <typename R>
hipError_t launch_tile_kernel(const TileGeom& tg, const DevGrid& g, const StepConst<R>& sc,
                              const R* ebf, const RecPtr& A, const RecPtr& B,
                              const int* seg_start, const int* seg_cap, const int* cnt_cur,
                              int* cnt_next, int* ovf_cur, int* ovf_next, double* patch,
                              int* grow, hipStream_t s);

// ------------------------------------------------------------ geometry
static int bits_for(int n) {  // bits to hold values 0..n
  int b = 0;
  while ((1 << b) <= n) b++;
  return b;
}

const char* tile_geom_error(int code) {
  switch (code) {
    case 1: return "tile shape or threads per tile (block) has no compiled kernel (3-D: 8x8x8, "
                   "8x8x4, 16x8x4; 2-D: 16x16, 32x16; block 0 or one thread per node-cell, or "
                   "two per node-cell for 8x8x4 and 16x16)";
    case 2: return "grid too large for the 31-bit packed cell index";
    case 3: return "tile patches exceed the LDS of one workgroup";
    default: return "invalid tile geometry";
  }
}

int tile_geom_init(const DevGrid& g, const aladyn_config& cfg, TileGeom* tg) {
  std::memset(tg, 0, sizeof *tg);
  const int D = g.ndim;
  // 8x8x4 in 3-D: 256 threads and 68 KB of LDS, so two tiles share a CU and
  // one's staging latency hides behind the other's particles
  const int def[3] = {D == 3 ? 8 : 16, D == 3 ? 8 : 16, D == 3 ? 4 : 1};
  for (int d = 0; d < 3; d++) tg->T[d] = cfg.tile[d] > 0 ? cfg.tile[d] : def[d];
  if (D == 2) tg->T[2] = 1;
  // threads per workgroup (aladyn_config.block): one per node-cell unless
  // the caller asks for two
  const int ncl = tg->T[0] * tg->T[1] * tg->T[2];
  tg->block = cfg.block > 0 ? cfg.block : ncl;
  const int rbytes = cfg.precision == ALADYN_PREC_F64 ? 8 : 4;
  if (!tile_shape_supported(D, tg->T[0], tg->T[1], tg->T[2], tg->block, rbytes)) return 1;
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
  tg->sort_every = cfg.sort_every > 0 ? cfg.sort_every : 64;
  tg->rbytes = cfg.precision == ALADYN_PREC_F64 ? 8 : 4;
  tg->lds_bytes = tile_lds_bytes(D, tg->T[0], tg->T[1], tg->T[2], tg->block, tg->rbytes);
  if (tg->lds_bytes > 160 * 1024) return 3;
  return 0;
}

// Adds a particle's current to the FP64 patch of its tile (nodes T0-2 ..
// T0+T+1 per axis cover the Esirkepov window of every particle whose old
// node lies in the tile).
template <int D>
struct PatchAcc {
  double* P;        // the tile's patch, component stride pvol
  int o[3];         // Fortran node of patch index 0 per axis (T0 - 2)
  int J0, J1, pvol;
  template <typename R>
  __device__ void operator()(int c, int i, int j, int k, R v) const {
    const int l = (i - o[0]) + J0 * ((j - o[1]) + J1 * (D == 3 ? k - o[2] : 0));
    unsafeAtomicAdd(P + c * pvol + l, (double)v);
  }
};

// ----------------------------------------------------- overflow kernel
// Particles whose destination segment was full: gather from HBM, push,
// deposit into the FP64 patch of the tile they start in (global FP64
// atomics, before the fold sums the patches onto jc, so their current is
// rounded to the grid's type once like everyone else's), re-bin like the
// tile kernel.  Runs every step; usually there is nothing to do.
template <int D, typename R>
__device__ inline void ovf_body(const TileGeom& tg, const DevGrid& g, const StepConst<R>& sc,
                                const R* __restrict__ ebf, const RecPtr& A, const RecPtr& B,
                                const int* __restrict__ seg_start, const int* __restrict__ seg_cap,
                                const int* __restrict__ ovf_cur, int* __restrict__ cnt_next,
                                int* __restrict__ ovf_next, double* __restrict__ patch,
                                int* __restrict__ grow) {
  const int cnt = *ovf_cur;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < cnt; n += gridDim.x * blockDim.x) {
    Rec<R> q;
    load_rec<D, R>(A, sc.ovf_base + n, q, tg.bx, tg.by);
    const int i0 = q.N[0] - 1, j0 = q.N[1] - 1, k0 = D == 3 ? q.N[2] - 1 : 1;
    const auto F = [&](int c, int i1, int j1, int k1) -> R {
      return ebf[(int64_t)c * g.ncell + fofs(g, i0 + i1, j0 + j1, D == 3 ? k0 + k1 : 1)];
    };
    R ef[6];
    gather<D, R>(q.s, F, ef);
    const int io[3] = {q.N[0], q.N[1], q.N[2]};
    const R so[3] = {q.s[0], q.s[1], q.s[2]};
    // the tile the particle starts in and its patch
    int tt[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; d++) tt[d] = (io[d] - 3) / tg.T[d];
    const int t0 = tt[0] + tg.nt[0] * (tt[1] + tg.nt[1] * tt[2]);
    const PatchAcc<D> acc{patch + (int64_t)t0 * D * tg.pvol,
                          {1 + tt[0] * tg.T[0], 1 + tt[1] * tg.T[1], 1 + tt[2] * tg.T[2]},
                          tg.J[0], tg.J[1], tg.pvol};
    const R dtg = push<D, R>(q.p, ef, sc.qch, sc.alp, sc.dt, sc.initial);
    if (!move_and_test<D, R>(q, dtg, sc.inv, sc.nn)) continue;
    esirkepov<D, R>(io, so, q.N, q.s, (R)((float)sc.qch * q.w), acc);  // FP32 charge*wgh
    const int dtile = (q.N[0] - 3) / tg.T[0] +
                      tg.nt[0] * ((q.N[1] - 3) / tg.T[1] + tg.nt[1] * (D == 3 ? (q.N[2] - 3) / tg.T[2] : 0));
    const int slot = reserve_slot(dtile, cnt_next);
    int64_t dst = (int64_t)seg_start[dtile] + slot;
    if (slot >= seg_cap[dtile]) {
      dst = sc.ovf_base + atomicAdd(ovf_next, 1);
      grow[dtile] = 1;
    }
    store_rec<D, R>(B, dst, q, tg.bx, tg.by);
  }
}

template <int D, typename R>
__global__ __launch_bounds__(256) void k_ovf_particles(
    TileGeom tg, DevGrid g, StepConst<R> sc, const R* __restrict__ ebf, RecPtr A, RecPtr B,
    const int* __restrict__ seg_start, const int* __restrict__ seg_cap,
    const int* __restrict__ ovf_cur, int* __restrict__ cnt_next, int* __restrict__ ovf_next,
    double* __restrict__ patch, int* __restrict__ grow) {
  ovf_body<D, R>(tg, g, sc, ebf, A, B, seg_start, seg_cap, ovf_cur, cnt_next, ovf_next, patch,
                 grow);
}

// the same for many runs in one launch, blockIdx.y selecting the run
template <int D, typename R>
__global__ __launch_bounds__(256) void kb_ovf_particles(const TileRun<R>* __restrict__ runs) {
  const TileRun<R>& r = runs[blockIdx.y];
  ovf_body<D, R>(r.tg, r.g, r.sc, r.ebf, r.A, r.B, r.seg_start, r.seg_cap, r.ovf_cur,
                 r.cnt_next, r.ovf_next, r.patch, r.grow);
}

// ------------------------------------------------------------ fold kernel
// jc(node) [+]= the sum of the tile patches covering node, for the first nc
// of the D components each tile's patch slot holds: at most two tiles
// per axis (T >= 4), the one whose interior holds the node and the neighbour
// whose halo reaches it.  One thread per node of a 64 x 4 x 1 block of the
// grid (no index division); the up to 8 FP64 patch values of a component are
// all loaded before they are summed in FP64, in the order of increasing tile
// index, multiplied by the component's factor, and rounded to the grid's
// type once.
template <typename R>
__device__ inline void fold_node(const TileGeom& tg, const DevGrid& g,
                                 const double* __restrict__ patch, R* __restrict__ jc,
                                 int accumulate, const int idx[3], const JScale& js,
                                 int nc) {
  const int D = g.ndim;
  // per axis: candidate tiles t0 = t1 - 1 and t1, local patch indices, validity
  int tl[3][2], loc[3][2];
  bool ok[3][2];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (d >= D) {
      tl[d][0] = tl[d][1] = 0;
      loc[d][0] = loc[d][1] = 0;
      ok[d][0] = true;
      ok[d][1] = false;
      continue;
    }
    const int sh = __ffs(tg.T[d]) - 1;  // T is a power of two
    const int hi = min((idx[d] - 1) >> sh, tg.nt[d] - 1);  // (u + 2) / T, u = idx - 3
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const int tt = hi - 1 + e;
      const int l = idx[d] - (1 + tt * tg.T[d]);  // patch origin T0 - 2
      ok[d][e] = tt >= 0 && l >= 0 && l < tg.J[d];
      tl[d][e] = max(tt, 0);
      loc[d][e] = ok[d][e] ? l : 0;
    }
  }
  const int64_t o = (idx[0] - 1) + (int64_t)g.n1p * ((idx[1] - 1) + (int64_t)g.n2p * (idx[2] - 1));
  for (int c = 0; c < nc; c++) {
    double v[8];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int e = 0; e < 2; e++) {
          const bool use = ok[0][e] && ok[1][b] && ok[2][a];
          const int tile = tl[0][e] + tg.nt[0] * (tl[1][b] + tg.nt[1] * tl[2][a]);
          const int l = loc[0][e] + tg.J[0] * (loc[1][b] + tg.J[1] * loc[2][a]);
          v[a * 4 + b * 2 + e] = use ? patch[((int64_t)tile * D + c) * tg.pvol + l] : 0.0;
        }
    double sum = 0.0;
#pragma unroll
    for (int q = 0; q < 8; q++) sum += v[q];
    // the factor by selects, not a dynamic index into the argument
    sum *= c == 0 ? js.s[0] : (c == 1 ? js.s[1] : js.s[2]);
    if (accumulate) sum += (double)jc[(int64_t)c * g.ncell + o];
    jc[(int64_t)c * g.ncell + o] = (R)sum;
  }
}

template <typename R>
__global__ __launch_bounds__(256) void k_fold(TileGeom tg, DevGrid g,
                                              const double* __restrict__ patch,
                                              R* __restrict__ jc, int accumulate, JScale js,
                                              int nc) {
  const int idx[3] = {(int)(blockIdx.x * 64 + (threadIdx.x & 63)) + 1,
                      (int)(blockIdx.y * 4 + (threadIdx.x >> 6)) + 1, (int)blockIdx.z + 1};
  if (idx[0] > g.n1p || idx[1] > g.n2p) return;
  fold_node<R>(tg, g, patch, jc, accumulate, idx, js, nc);
}

// jc = djc x the folded patches of many runs in one launch (blockIdx.y =
// run), one node per thread in storage order
template <typename R>
__global__ __launch_bounds__(256) void kb_fold(const TileRun<R>* __restrict__ runs) {
  const TileRun<R>& r = runs[blockIdx.y];
  const DevGrid& g = r.g;
  const JScale js = jscale_of(g, true);
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < g.ncell;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t jk = q / g.n1p;
    const int idx[3] = {(int)(q - jk * g.n1p) + 1, (int)(jk % g.n2p) + 1, (int)(jk / g.n2p) + 1};
    fold_node<R>(r.tg, g, r.patch, r.jc, 0, idx, js, g.ndim);
  }
}

// the counters the step's kernels fill (next buffer's tile counts and
// overflow count) start at zero
template <typename R>
__global__ __launch_bounds__(256) void kb_clear(const TileRun<R>* __restrict__ runs) {
  const TileRun<R>& r = runs[blockIdx.y];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < r.tg.ntiles; t += gridDim.x * blockDim.x)
    r.cnt_next[t] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *r.ovf_next = 0;
}

template <typename R>
__global__ void kb_ovf_out(const TileRun<R>* __restrict__ runs, int nruns, int* out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nruns) out[r] = *runs[r].ovf_next;
}

template <typename R>
void tile_store_fold(const TileGeom& tg, const DevGrid& g, const double* patch, R* jc,
                     int accumulate, hipStream_t s, bool normalise) {
  const dim3 fb((g.n1p + 63) / 64, (g.n2p + 3) / 4, g.n3p);
  k_fold<R><<<fb, 256, 0, s>>>(tg, g, patch, jc, accumulate, jscale_of(g, normalise), g.ndim);
}

void tile_store_fold_density(const TileGeom& tg, const DevGrid& g, const double* patch,
                             double* den, hipStream_t s) {
  const dim3 fb((g.n1p + 63) / 64, (g.n2p + 3) / 4, g.n3p);
  k_fold<double><<<fb, 256, 0, s>>>(tg, g, patch, den, 0, jscale_of(g, false), 1);
}
template void tile_store_fold<float>(const TileGeom&, const DevGrid&, const double*, float*, int,
                                     hipStream_t, bool);
template void tile_store_fold<double>(const TileGeom&, const DevGrid&, const double*, double*,
                                      int, hipStream_t, bool);

// ------------------------------------------------- relayout / transfers
static constexpr int kB = 256;

static unsigned grid_for(int64_t n, int64_t cap = 65535LL * 8) {
  int64_t b = (n + kB - 1) / kB;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
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

// Segment capacities at a re-layout.  The segments of one buffer share a
// budget of kSegBudget records per particle (HBM is plentiful and slack that
// no particle uses is never read): a tile holding v particles gets
// v + v/2 + 64 slots, or 2v + 256 if it spilled since the last re-layout,
// and the rest of the budget is spread evenly over all tiles.  The even share
// matters where the density moves fast: a tile swept by the dense rear of a
// plasma wake gains thousands of particles per step from a low count.
static constexpr int64_t kSegBudgetNum = 5, kSegBudgetDen = 2;  // 2.5 slots per particle
__host__ __device__ inline int64_t tile_base_capacity(int64_t v, bool grown) {
  return grown ? 2 * v + 256 : v + v / 2 + 64;
}
// slots of the segment area for np particles in ntiles tiles (bounds every
// sum of base capacities, since sum v <= np)
static inline int64_t seg_budget(int64_t np, int64_t ntiles) {
  return np * kSegBudgetNum / kSegBudgetDen + 256 * ntiles + 64;
}

// nb[t] = the largest count among tile t and its face, edge and corner
// neighbours.  Plasma streams from tile to tile: a tile that a dense region
// (a bunch, the rear of a wake) is about to enter holds few particles at a
// re-layout but needs room for its neighbour's.
__global__ __launch_bounds__(kB) void k_nbmax(TileGeom tg, const int* __restrict__ v,
                                              int* __restrict__ nb) {
  const int t = blockIdx.x * kB + threadIdx.x;
  if (t >= tg.ntiles) return;
  const int x = t % tg.nt[0], y = (t / tg.nt[0]) % tg.nt[1], z = t / (tg.nt[0] * tg.nt[1]);
  int m = 0;
  for (int c = max(z - 1, 0); c <= min(z + 1, tg.nt[2] - 1); c++)
    for (int b = max(y - 1, 0); b <= min(y + 1, tg.nt[1] - 1); b++)
      for (int a = max(x - 1, 0); a <= min(x + 1, tg.nt[0] - 1); a++)
        m = max(m, v[a + tg.nt[0] * (b + tg.nt[1] * c)]);
  nb[t] = m;
}

// out[t] = exclusive scan of v[t], out[n] = total (one workgroup).  With
// cap_out, v[t] is first turned into the segment capacity of tile t (base
// capacity plus an even share of what the budget leaves; the tile's grow
// flag is consumed).  With nb (k_nbmax of v), the base capacity is that of
// the neighbourhood's largest count when all of those fit the budget, else
// the tile's own.
// One workgroup of kScanB threads: the sums are wave shuffles with one
// barrier per pass of kScanB tiles, and the next pass's loads are issued
// before that barrier, so the pass time is not the HBM latency.
static constexpr int kScanB = 1024, kScanW = kScanB / 64;
__device__ inline long long wave_sum_ll(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__global__ __launch_bounds__(kScanB) void k_scan_tiles(const int* v_in, int n, int* out,
                                                       int* cap_out, int* grow = nullptr,
                                                       int64_t budget = 0,
                                                       const int* nb = nullptr) {
  __shared__ long long red[2][kScanW];
  __shared__ int wtot[2][kScanW];
  const int me = threadIdx.x, lane = me & 63, wv = me >> 6;
  bool use_nb = false;
  long long base_sum = 0;
  if (cap_out) {
    long long own = 0, wide = 0;
#pragma unroll 4
    for (int t = me; t < n; t += kScanB) {
      const bool gr = grow && grow[t];
      own += tile_base_capacity(v_in[t], gr);
      if (nb) wide += tile_base_capacity(nb[t], gr);
    }
    own = wave_sum_ll(own);  // block sums of the base capacities
    wide = wave_sum_ll(wide);
    if (lane == 0) {
      red[0][wv] = own;
      red[1][wv] = wide;
    }
    __syncthreads();
    long long so = 0, sw = 0;
    for (int w = 0; w < kScanW; w++) {
      so += red[0][w];
      sw += red[1][w];
    }
    use_nb = nb && sw <= budget;
    base_sum = use_nb ? sw : so;
  }
  const long long left = cap_out ? budget - base_sum : 0;
  const long long extra = left > 0 ? left / n : 0;
  // the tile's value: its count, or with cap_out its segment capacity
  const auto value = [&](int t, int raw, int nbv, bool gr) -> int {
    if (t >= n) return 0;
    return cap_out ? (int)(tile_base_capacity(use_nb ? nbv : raw, gr) + extra) : raw;
  };
  int raw = 0, nbv = 0;
  bool gr = false;
  const auto fetch = [&](int t) {
    if (t < n) {
      raw = v_in[t];
      if (cap_out && use_nb) nbv = nb[t];
      gr = cap_out && grow && grow[t];
    }
  };
  fetch(me);
  long long carry = 0;
  for (int b0 = 0, p = 0; b0 < n; b0 += kScanB, p ^= 1) {
    const int t = b0 + me;
    const int v = value(t, raw, nbv, gr);
    if (t < n && cap_out) {
      if (grow) grow[t] = 0;
      cap_out[t] = v;
    }
    fetch(t + kScanB);  // the next pass's loads, in flight across the barrier
    int x = v;  // inclusive scan within the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wtot[p][wv] = x;
    __syncthreads();  // wtot[p] complete; wtot[p^1] is free for the next pass
    int pre = 0, tot = 0;
    for (int w = 0; w < kScanW; w++) {
      const int s = wtot[p][w];
      pre += w < wv ? s : 0;
      tot += s;
    }
    if (t < n) out[t] = (int)(carry + pre + x - v);
    carry += tot;
  }
  if (me == 0) out[n] = (int)carry;
}

// live (non-hole) records per tile: live[t], and the overflow count at [ntiles]
template <typename R>
__global__ __launch_bounds__(kB) void k_count_live(RecPtr A, const int* seg_start, const int* seg_cap,
                                                   const int* cnt, int ntiles, const int* ovf_cnt,
                                                   int* live) {
  __shared__ int sh[kB];
  const int t = blockIdx.x;
  if (t == ntiles) {
    if (threadIdx.x == 0) live[ntiles] = *ovf_cnt;
    return;
  }
  const int n = min(cnt[t], seg_cap[t]);
  const int64_t s0 = seg_start[t];
  int c = 0;
#pragma unroll 4
  for (int q = threadIdx.x; q < n; q += kB) c += rec_cell<R>(A, s0 + q) != kHole;
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int total = 0;
    for (int w = 0; w < kB / 64; w++) total += sh[w];
    live[t] = total;
  }
}

// live records of A (tile segments in order, then the overflow list) -> U
template <int D, typename R>
__global__ __launch_bounds__(kB) void k_gather_live(RecPtr A, RecPtr U, const int* seg_start,
                                                    const int* seg_cap, const int* cnt,
                                                    const int* off, int ntiles, int64_t ovf_base,
                                                    const int* ovf_cnt) {
  __shared__ int sh[kB];
  const int t = blockIdx.x;
  if (t < ntiles) {
    const int n = min(cnt[t], seg_cap[t]);
    int dst = off[t];
    for (int q0 = 0; q0 < n; q0 += kB) {
      const int q = q0 + threadIdx.x;
      const bool ok = q < n && rec_cell<R>(A, (int64_t)seg_start[t] + q) != kHole;
      int total;
      const int ex = block_scan_excl(ok ? 1 : 0, sh, total);
      if (ok) copy_rec<R>(A, (int64_t)seg_start[t] + q, U, (int64_t)dst + ex);
      dst += total;
    }
  } else {
    const int n = *ovf_cnt;
    const int64_t o0 = off[ntiles];
    for (int q = (t - ntiles) * blockDim.x + threadIdx.x; q < n; q += (gridDim.x - ntiles) * blockDim.x)
      copy_rec<R>(A, ovf_base + q, U, o0 + q);
  }
}

template <int D, typename R>
__global__ void k_hist(TileGeom tg, RecPtr U, int64_t n, int* hcount) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    (void)reserve_slot(tile_of(tg, rec_cell<R>(U, q), D), hcount);
}

template <int D, typename R>
__global__ void k_scatter(TileGeom tg, RecPtr U, int64_t n, RecPtr A, const int* seg_start,
                          int* fillc) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int t = tile_of(tg, rec_cell<R>(U, q), D);
    const int slot = reserve_slot(t, fillc);
    copy_rec<R>(U, q, A, (int64_t)seg_start[t] + slot);
  }
}

// One-pass re-layout: counts per destination tile straight from the tile
// segments of A (holes skipped) and its overflow list ...
// A window shift of shx cells moves every record shx nodes to the left
// (an exact change of the packed node index) and drops the records now at or
// behind the box's left face, x <= xmin: node < 3, or node 3 with s <= 0
// (the test of reset_all_part_dist, mpi_part_interface.f90:494-498).
// The overflow list's records from index ovf_keep on were appended after the
// shift was decided (the plasma injected at the same move) and are already in
// the new frame: they are not shifted.
template <typename R>
__device__ inline int shift_cell(const TileGeom& tg, int cell, R s0, int shx) {
  if (shx == 0 || cell == kHole) return cell;
  const int i = (cell & ((1 << tg.bx) - 1)) - shx;
  if (i < 3 || (i == 3 && !(s0 > R(0)))) return kHole;
  return cell - shx;
}
template <typename R>
__device__ inline int shifted_cell(const TileGeom& tg, const RecPtr& A, int64_t src, int cell,
                                   int shx) {
  if (shx == 0 || cell == kHole) return cell;
  const int i = (cell & ((1 << tg.bx) - 1)) - shx;
  if (i < 3 || (i == 3 && !(rec_s<R>(A, src, 0) > 0.f))) return kHole;
  return cell - shx;
}

// Records per thread and pass of the re-layout kernels: their loads are
// issued together, so a pass pays the HBM latency once.
static constexpr int kRU = 4;

template <int D, typename R>
__global__ __launch_bounds__(kB) void k_hist_seg(TileGeom tg, RecPtr A, const int* seg_start,
                                                 const int* seg_cap, const int* cnt,
                                                 int64_t ovf_base, const int* ovf_cnt,
                                                 int* hcount, int shx, int ovf_keep) {
  const int t = blockIdx.x;
  if (t < tg.ntiles) {
    // records that stay in tile t are counted per thread and added with one
    // atomic per wave; the others go through the wave-aggregated counters
    const int n = min(cnt[t], seg_cap[t]);
    const int64_t s0 = seg_start[t];
    int own = 0;
    for (int q0 = threadIdx.x; q0 < n; q0 += kRU * kB) {
      int cell[kRU];
#pragma unroll
      for (int u = 0; u < kRU; u++) {
        const int q = q0 + u * kB;
        cell[u] = q < n ? rec_cell<R>(A, s0 + q) : kHole;
      }
#pragma unroll
      for (int u = 0; u < kRU; u++) {
        const int c = shifted_cell<R>(tg, A, s0 + q0 + u * kB, cell[u], shx);
        if (c == kHole) continue;
        const int d = tile_of(tg, c, D);
        if (d == t) own++;
        else (void)reserve_slot(d, hcount);
      }
    }
    own = wave_sum(own);
    if ((threadIdx.x & 63) == 0 && own > 0) atomicAdd(&hcount[t], own);
  } else {
    const int n = *ovf_cnt;
    const int stride = (gridDim.x - tg.ntiles) * kB;
    for (int q0 = (t - tg.ntiles) * kB + threadIdx.x; q0 < n; q0 += kRU * stride) {
      int cell[kRU];
#pragma unroll
      for (int u = 0; u < kRU; u++) {
        const int q = q0 + u * stride;
        cell[u] = q < n ? rec_cell<R>(A, ovf_base + q) : kHole;
      }
#pragma unroll
      for (int u = 0; u < kRU; u++) {
        const int q = q0 + u * stride;
        const int c = shifted_cell<R>(tg, A, ovf_base + q, cell[u], q < ovf_keep ? shx : 0);
        if (c != kHole) (void)reserve_slot(tile_of(tg, c, D), hcount);
      }
    }
  }
}

// the packed cell and the x offset of a record's first word
template <typename R>
__device__ inline int word_cell(const typename RecWord<R>::V& w) {
  if constexpr (sizeof(R) == 4) return __float_as_int(w.x);
  else return (int)(__double_as_longlong(w.x) & 0xffffffffll);
}
template <typename R>
__device__ inline void set_word_cell(typename RecWord<R>::V& w, int cell) {
  if constexpr (sizeof(R) == 4) {
    w.x = __int_as_float(cell);
  } else {
    const long long cw = __double_as_longlong(w.x);
    w.x = __longlong_as_double((cw & ~0xffffffffll) | (long long)(unsigned)cell);
  }
}

// ... and moves every live record of A into its tile's new segment of U.
template <int D, typename R>
__global__ __launch_bounds__(kB) void k_scatter_seg(TileGeom tg, RecPtr A, const int* seg_start,
                                                    const int* seg_cap, const int* cnt,
                                                    int64_t ovf_base, const int* ovf_cnt, RecPtr U,
                                                    const int* new_start, int* fillc, int shx,
                                                    int ovf_keep) {
  using V = typename RecWord<R>::V;
  constexpr int NW = RecWord<R>::N;
  const int t = blockIdx.x;
  const int lane = (int)__lane_id();
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  const bool seg = t < tg.ntiles;
  const int n = seg ? min(cnt[t], seg_cap[t]) : *ovf_cnt;
  const int64_t s0 = seg ? (int64_t)seg_start[t] : ovf_base;
  const int first = seg ? threadIdx.x : (t - tg.ntiles) * kB + threadIdx.x;
  const int stride = seg ? kB : (gridDim.x - tg.ntiles) * kB;
  for (int q0 = first; q0 < n; q0 += kRU * stride) {
    V w[kRU][NW];
    int id[kRU], d[kRU];
#pragma unroll
    for (int u = 0; u < kRU; u++) {
      const int q = q0 + u * stride;
      d[u] = -1;
      id[u] = 0;
      if (q < n) {
        const V* src = rec_at<R>(A, s0 + q);
#pragma unroll
        for (int k = 0; k < NW; k++) w[u][k] = src[k];
        if (A.id && U.id) id[u] = A.id[s0 + q];
      }
    }
#pragma unroll
    for (int u = 0; u < kRU; u++) {
      const int q = q0 + u * stride;
      if (q >= n) continue;
      const int c = shift_cell<R>(tg, word_cell<R>(w[u][0]), (R)w[u][0].y,
                                  seg || q < ovf_keep ? shx : 0);
      if (c == kHole) continue;
      set_word_cell<R>(w[u][0], c);
      d[u] = tile_of(tg, c, D);
    }
    // records that stay in tile t: one atomic per wave for all of them
    int base = 0, nown = 0;
    unsigned long long ownm[kRU];
#pragma unroll
    for (int u = 0; u < kRU; u++) {
      ownm[u] = __ballot(seg && d[u] == t);
      nown += __popcll(ownm[u]);
    }
    if (nown > 0) {
      const int leader = __ffsll((unsigned long long)__ballot(1)) - 1;
      if (lane == leader) base = atomicAdd(&fillc[t], nown);
      base = __shfl(base, leader);
    }
#pragma unroll
    for (int u = 0; u < kRU; u++) {
      int64_t dst = -1;
      if (seg && d[u] == t) {
        dst = (int64_t)new_start[t] + base + __popcll(ownm[u] & lt_mask);
      } else if (d[u] >= 0) {
        dst = (int64_t)new_start[d[u]] + reserve_slot(d[u], fillc);
      }
      base += __popcll(ownm[u]);
      if (dst >= 0) {
        V* o = rec_at<R>(U, dst);
#pragma unroll
        for (int k = 0; k < NW; k++) o[k] = w[u][k];
        if (A.id && U.id) U.id[dst] = id[u];
      }
    }
  }
}

// reference layout part(ld, 2D+1) -> records (ids = id0 + input index)
template <int D, typename R>
__global__ void k_convert_in(DevGrid g, int bx, int by, const double* part, int64_t ld,
                             int64_t np, RecPtr U, int64_t id0 = 0) {
  const double mn[3] = {g.xmn, g.ymn, g.zmn};
  const double inv[3] = {g.dx_inv, g.dy_inv, g.dz_inv};
  for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < np;
       n += (int64_t)gridDim.x * blockDim.x) {
    int N[3] = {1, 1, 1};
    R sv[3] = {R(0), R(0), R(0)}, pv[3] = {R(0), R(0), R(0)};
#pragma unroll
    for (int d = 0; d < D; d++) {
      double s;
      ref_node(part[d * ld + n], mn[d], inv[d], N[d], s);
      R sf = (R)s;
      if (sf >= R(0.5)) {  // keep the [-1/2, 1/2) convention after rounding
        N[d] += 1;
        sf -= R(1);
      }
      sv[d] = sf;
      pv[d] = (R)part[(D + d) * ld + n];
    }
    const int cell = N[0] | (N[1] << bx) | (D == 3 ? (N[2] << (bx + by)) : 0);
    const unsigned long long b = (unsigned long long)__double_as_longlong(part[2 * D * ld + n]);
    put_rec<D, R>(U, n, cell, sv, pv, __uint_as_float((unsigned)(b >> 32)));
    if (U.id) U.id[n] = (int)(id0 + n);
  }
}

// records -> reference layout (weights repacked on the host)
template <int D, typename R>
__global__ void k_convert_out(DevGrid g, int bx, int by, RecPtr U, int64_t n, double* out,
                              int64_t ld) {
  const double mn[3] = {g.xmn, g.ymn, g.zmn};
  const double dd[3] = {g.dx, g.dy, g.dz};
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int cell = rec_cell<R>(U, q);
    const int N[3] = {cell & ((1 << bx) - 1), (cell >> bx) & ((1 << by) - 1),
                      D == 3 ? (cell >> (bx + by)) : 1};
#pragma unroll
    for (int d = 0; d < D; d++) {
      out[d * ld + q] = mn[d] + ((double)(N[d] - 3) + (double)rec_s<R>(U, q, d)) * dd[d];
      out[(D + d) * ld + q] = (double)rec_p<R>(U, q, d);
    }
    out[2 * D * ld + q] = (double)rec_w<R>(U, q);
  }
}

// bytes of one record (RecWord in tile_device.hpp): 32 in the FP32 path,
// 64 in the FP64 path
static size_t rec_size(const TileGeom& tg) { return tg.rbytes == 8 ? 64 : 32; }
// bytes of one record slot: the record plus the tracked id
static size_t rec_bytes(const TileGeom& tg, bool track_ids) {
  return rec_size(tg) + (track_ids ? 4 : 0);
}

static void set_ptrs(TileStore* ts, const TileGeom& tg, int b) {
  // layout of one buffer: alloc records, then alloc ids
  char* m = (char*)ts->mem[b];
  RecPtr& r = ts->buf[b];
  r.r = m;
  r.id = ts->track_ids ? (int*)(m + rec_size(tg) * ts->alloc) : nullptr;
}

// Calls f(dim, real) with dim a std::integral_constant<int, D> and real a
// value of the record's real type, for the store's dimension and precision.
template <typename F>
static int with_dr(const TileGeom& tg, int D, F&& f) {
  using D2 = std::integral_constant<int, 2>;
  using D3 = std::integral_constant<int, 3>;
  if (D == 2) return tg.rbytes == 8 ? f(D2{}, 0.0) : f(D2{}, 0.0f);
  return tg.rbytes == 8 ? f(D3{}, 0.0) : f(D3{}, 0.0f);
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
  (void)hipFree(ts->seg_start2);
  (void)hipFree(ts->seg_cap2);
  (void)hipFree(ts->ovf_count);
  (void)hipFree(ts->scratch);
  (void)hipFree(ts->d_total);
  (void)hipFree(ts->grow);
  if (ts->h_ovf) (void)hipHostFree(ts->h_ovf);
  for (int b = 0; b < 2; b++)
    if (ts->ovf_ev[b]) (void)hipEventDestroy(ts->ovf_ev[b]);
  *ts = TileStore();
}

// Counting sort of the n records of U (stored in buffer 1-cur) into buffer
// cur with the re-layout capacities of k_scan_tiles.
template <int D, typename R>
static int sort_into_cur(TileStore* ts, const TileGeom& tg, int64_t n, hipStream_t s,
                         std::string* msg) {
  const int cur = ts->cur, oth = 1 - cur;
  int* h = ts->count[cur];
  fill<int>(h, tg.ntiles, 0, s);
  if (n > 0) k_hist<D, R><<<grid_for(n), kB, 0, s>>>(tg, ts->buf[oth], n, h);
  k_nbmax<<<(tg.ntiles + kB - 1) / kB, kB, 0, s>>>(tg, h, ts->scratch);
  k_scan_tiles<<<1, kScanB, 0, s>>>(h, tg.ntiles, ts->seg_start, ts->seg_cap, ts->grow,
                                seg_budget(ts->ovf_cap, tg.ntiles), ts->scratch);
  int total = 0;
  TS_CK(hipMemcpyAsync(&total, ts->seg_start + tg.ntiles, sizeof(int), hipMemcpyDeviceToHost, s));
  TS_CK(hipStreamSynchronize(s));
  if ((int64_t)total > ts->ovf_base) {
    *msg = "tile store: segment capacities exceed the allocation";
    return ALADYN_ENOMEM;
  }
  fill<int>(ts->scratch, tg.ntiles, 0, s);
  if (n > 0)
    k_scatter<D, R><<<grid_for(n), kB, 0, s>>>(tg, ts->buf[oth], n, ts->buf[cur], ts->seg_start,
                                            ts->scratch);
  fill<int>(ts->ovf_count, 2, 0, s);
  TS_CK(hipGetLastError());
  ts->live_bound = n;
  ts->ovf_valid[0] = ts->ovf_valid[1] = false;
  ts->steps_since_sort = 0;
  return 0;
}

// The path's name in messages.
static const char* path_name(const TileGeom& tg) {
  return tg.rbytes == 8 ? "FP64 tile path" : "FP32 tile path";
}

int tile_store_put(TileStore* ts, const TileGeom& tg, const DevGrid& g, const double* part,
                   int64_t ld, int64_t np, hipStream_t s, bool track_ids, std::string* msg) {
  const int D = g.ndim;
  ts->layout_version++;
  ts->pending_shx = 0;  // the particles given are in the current box
  // the record keeps the FP32 weight; charge and part_ind are per species
  int charge = 0, pind = 0;
  for (int64_t n = 0; n < np; n++) {
    unsigned long long b;
    std::memcpy(&b, part + 2 * D * ld + n, 8);
    const int c = (int)(short)(b & 0xffff), p = (int)(short)((b >> 16) & 0xffff);
    if (n == 0) {
      charge = c;
      pind = p;
    } else if (c != charge || p != pind) {
      *msg = std::string(path_name(tg)) + " needs one charge and part_ind per species";
      return ALADYN_EINVAL;
    }
  }
  const int64_t need = seg_budget(np, tg.ntiles);  // segments
  const int64_t alloc = need + (np > 0 ? np : 1);   // + overflow region
  if (alloc >= (1LL << 31)) {
    *msg = std::string(path_name(tg)) + ": particle count exceeds the 31-bit slot range";
    return ALADYN_EINVAL;
  }
  if (!ts->loaded || ts->alloc < alloc || ts->track_ids != track_ids) {
    tile_store_free(ts);
    ts->track_ids = track_ids;
    ts->alloc = alloc;
    const size_t per = rec_bytes(tg, track_ids);
    for (int b = 0; b < 2; b++) {
      TS_CK(hipMalloc(&ts->mem[b], per * (size_t)alloc));
      set_ptrs(ts, tg, b);
      TS_CK(hipMalloc(&ts->count[b], sizeof(int) * (tg.ntiles + 1)));
    }
    TS_CK(hipMalloc(&ts->seg_start, sizeof(int) * (tg.ntiles + 1)));
    TS_CK(hipMalloc(&ts->seg_cap, sizeof(int) * (tg.ntiles + 1)));
    TS_CK(hipMalloc(&ts->seg_start2, sizeof(int) * (tg.ntiles + 1)));
    TS_CK(hipMalloc(&ts->seg_cap2, sizeof(int) * (tg.ntiles + 1)));
    TS_CK(hipMalloc(&ts->ovf_count, sizeof(int) * 2));
    TS_CK(hipMalloc(&ts->scratch, sizeof(int) * (tg.ntiles + 2)));
    TS_CK(hipMalloc(&ts->d_total, 64));
    TS_CK(hipMalloc(&ts->grow, sizeof(int) * tg.ntiles));
    TS_CK(hipHostMalloc(&ts->h_ovf, 2 * sizeof(int), hipHostMallocDefault));
    for (int b = 0; b < 2; b++)
      TS_CK(hipEventCreateWithFlags(&ts->ovf_ev[b], hipEventDisableTiming));
    ts->h_ovf[0] = ts->h_ovf[1] = 0;
  }
  ts->ovf_cap = np > 0 ? np : 1;
  ts->ovf_base = ts->alloc - ts->ovf_cap;
  ts->charge = charge;
  ts->part_ind = pind;
  ts->cur = 0;
  ts->next_id = np;
  // The host columns go through a device staging area of at most kPutChunk
  // particles, converted chunk by chunk into buffer 1 (stream order keeps a
  // chunk's copy behind the previous chunk's conversion), so a put needs
  // little device memory beyond the store itself.
  constexpr int64_t kPutChunk = 1 << 22;
  double* dpart = nullptr;
  const int ncol = 2 * D + 1;
  const int64_t chunk = std::min<int64_t>(np, kPutChunk);
  if (np > 0) TS_CK(hipMalloc(&dpart, (size_t)chunk * ncol * sizeof(double)));
  fill<int>(ts->grow, tg.ntiles, 0, s);
  const int r = with_dr(tg, D, [&](auto dc, auto rv) {
    constexpr int DD = decltype(dc)::value;
    using R = decltype(rv);
    for (int64_t c0 = 0; c0 < np; c0 += chunk) {
      const int64_t m = std::min<int64_t>(chunk, np - c0);
      for (int c = 0; c < ncol; c++)
        TS_CK(hipMemcpyAsync(dpart + c * chunk, part + c * ld + c0, (size_t)m * sizeof(double),
                             hipMemcpyHostToDevice, s));
      RecPtr out = ts->buf[1];
      out.r = rec_at<R>(out, c0);
      if (out.id) out.id += c0;
      k_convert_in<DD, R><<<grid_for(m), kB, 0, s>>>(g, tg.bx, tg.by, dpart, chunk, m, out, c0);
      TS_CK(hipGetLastError());
    }
    return sort_into_cur<DD, R>(ts, tg, np, s, msg);
  });
  TS_CK(hipStreamSynchronize(s));
  (void)hipFree(dpart);
  if (r) return r;
  ts->loaded = true;
  return 0;
}

// gathers the live records of buffer cur into buffer 1-cur; returns count
template <int D, typename R>
static int gather_live(TileStore* ts, const TileGeom& tg, hipStream_t s, int64_t* n,
                       std::string* msg) {
  const int cur = ts->cur, oth = 1 - cur;
  int* live = ts->count[oth];  // free between steps
  k_count_live<R><<<tg.ntiles + 1, kB, 0, s>>>(ts->buf[cur], ts->seg_start, ts->seg_cap,
                                               ts->count[cur], tg.ntiles, ts->ovf_count + cur,
                                               live);
  k_scan_tiles<<<1, kScanB, 0, s>>>(live, tg.ntiles + 1, ts->scratch, nullptr);
  k_gather_live<D, R><<<tg.ntiles + 64, kB, 0, s>>>(ts->buf[cur], ts->buf[oth], ts->seg_start,
                                                    ts->seg_cap, ts->count[cur], ts->scratch,
                                                    tg.ntiles, ts->ovf_base, ts->ovf_count + cur);
  TS_CK(hipGetLastError());
  int tot = 0;
  TS_CK(hipMemcpyAsync(&tot, ts->scratch + tg.ntiles + 1, sizeof(int), hipMemcpyDeviceToHost, s));
  TS_CK(hipStreamSynchronize(s));
  *n = tot;
  return 0;
}

// One-pass re-layout of buffer cur into buffer 1-cur (which becomes cur):
// count per destination tile from the segments, fresh capacities, scatter.
// Fully asynchronous: the capacities fit the segment area by construction
// (live particles never exceed the count given to put).
template <int D, typename R>
static int relayout_seg(TileStore* ts, const TileGeom& tg, hipStream_t s, std::string* msg,
                        int shx = 0, int ovf_keep = INT_MAX) {
  const int cur = ts->cur, oth = 1 - cur;
  int* h = ts->count[oth];  // per-tile counts of the new layout
  fill<int>(h, tg.ntiles, 0, s);
  k_hist_seg<D, R><<<tg.ntiles + 64, kB, 0, s>>>(tg, ts->buf[cur], ts->seg_start, ts->seg_cap,
                                                 ts->count[cur], ts->ovf_base, ts->ovf_count + cur,
                                                 h, shx, ovf_keep);
  k_nbmax<<<(tg.ntiles + kB - 1) / kB, kB, 0, s>>>(tg, h, ts->scratch);
  k_scan_tiles<<<1, kScanB, 0, s>>>(h, tg.ntiles, ts->seg_start2, ts->seg_cap2, ts->grow,
                                seg_budget(ts->ovf_cap, tg.ntiles), ts->scratch);
  fill<int>(ts->scratch, tg.ntiles, 0, s);
  k_scatter_seg<D, R><<<tg.ntiles + 64, kB, 0, s>>>(tg, ts->buf[cur], ts->seg_start, ts->seg_cap,
                                                    ts->count[cur], ts->ovf_base,
                                                    ts->ovf_count + cur, ts->buf[oth],
                                                    ts->seg_start2, ts->scratch, shx, ovf_keep);
  fill<int>(ts->ovf_count, 2, 0, s);
  TS_CK(hipGetLastError());
  std::swap(ts->seg_start, ts->seg_start2);
  std::swap(ts->seg_cap, ts->seg_cap2);
  ts->cur = oth;
  ts->ovf_valid[0] = ts->ovf_valid[1] = false;
  ts->steps_since_sort = 0;
  return 0;
}

__global__ __launch_bounds__(kB) void k_clear_next(int* __restrict__ cnt, int n,
                                                   int* __restrict__ ovf) {
  for (int t = blockIdx.x * kB + threadIdx.x; t < n; t += gridDim.x * kB) cnt[t] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *ovf = 0;
}

void tile_store_clear_next(TileStore* ts, const TileGeom& tg, hipStream_t s) {
  const int nxt = 1 - ts->cur;
  k_clear_next<<<grid_for(tg.ntiles), kB, 0, s>>>(ts->count[nxt], tg.ntiles, ts->ovf_count + nxt);
}

int tile_store_relayout(TileStore* ts, const TileGeom& tg, const DevGrid& g, hipStream_t s,
                        std::string* msg) {
  if (!ts->loaded) return 0;
  ts->layout_version++;
  return with_dr(tg, g.ndim, [&](auto dc, auto rv) {
    return relayout_seg<decltype(dc)::value, decltype(rv)>(ts, tg, s, msg);
  });
}

// Moves the box window_cells cells to the right (window.f90:276-347): every
// record's node index drops by window_cells, records left behind the new
// left face are dropped, and the store is re-laid out in the same pass (or,
// deferred, in the pass of the injection that follows).
int tile_store_window(TileStore* ts, const TileGeom& tg, const DevGrid& g, int window_cells,
                      hipStream_t s, std::string* msg, bool defer) {
  if (!ts->loaded || window_cells == 0) return 0;
  ts->layout_version++;
  ts->pending_shx += window_cells;
  return defer ? 0 : tile_store_window_apply(ts, tg, g, s, msg);
}

int tile_store_window_apply(TileStore* ts, const TileGeom& tg, const DevGrid& g, hipStream_t s,
                            std::string* msg) {
  const int shx = ts->pending_shx;
  ts->pending_shx = 0;
  if (!ts->loaded || shx == 0) return 0;
  return with_dr(tg, g.ndim, [&](auto dc, auto rv) {
    return relayout_seg<decltype(dc)::value, decltype(rv)>(ts, tg, s, msg, shx);
  });
}

// Re-allocates the two buffers for new_cap particles, keeping the live ones.
template <int D, typename R>
static int grow_store(TileStore* ts, const TileGeom& tg, hipStream_t s, int64_t new_cap,
                      std::string* msg) {
  int64_t n = 0;
  if (int r = gather_live<D, R>(ts, tg, s, &n, msg)) return r;  // dense, in buffer 1-cur
  const int64_t alloc = seg_budget(new_cap, tg.ntiles) + new_cap;
  if (alloc >= (1LL << 31)) {
    *msg = std::string(path_name(tg)) + ": particle count exceeds the 31-bit slot range";
    return ALADYN_EINVAL;
  }
  const size_t per = rec_bytes(tg, ts->track_ids);
  void* fresh[2] = {nullptr, nullptr};
  for (int b = 0; b < 2; b++) {
    if (hipMalloc(&fresh[b], per * (size_t)alloc) != hipSuccess) {
      (void)hipFree(fresh[0]);
      *msg = "tile store: out of device memory growing the particle buffers";
      return ALADYN_ENOMEM;
    }
  }
  const RecPtr old = ts->buf[1 - ts->cur];
  void* oldmem[2] = {ts->mem[0], ts->mem[1]};
  ts->mem[0] = fresh[0];
  ts->mem[1] = fresh[1];
  ts->alloc = alloc;
  set_ptrs(ts, tg, 0);
  set_ptrs(ts, tg, 1);
  const RecPtr& nu = ts->buf[1];
  if (n > 0) {
    TS_CK(hipMemcpyAsync(nu.r, old.r, (size_t)n * rec_size(tg), hipMemcpyDeviceToDevice, s));
    if (nu.id) TS_CK(hipMemcpyAsync(nu.id, old.id, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
  }
  TS_CK(hipStreamSynchronize(s));
  (void)hipFree(oldmem[0]);
  (void)hipFree(oldmem[1]);
  ts->ovf_cap = new_cap;
  ts->ovf_base = alloc - new_cap;
  ts->cur = 0;
  return sort_into_cur<D, R>(ts, tg, n, s, msg);
}

// Appends n records to the species: fill(dst, id0) writes them into the
// overflow list of the current buffer (dst points at its first free record,
// id0 is the first new id) and the store is re-laid out, so the next step
// moves them in their tiles.
template <int D, typename R, typename Fill>
static int append_impl(TileStore* ts, const TileGeom& tg, int64_t n, hipStream_t s,
                       std::string* msg, Fill&& fill) {
  // the exact count (a pass over the store) only when the bound does not fit
  int64_t live = ts->live_bound;
  if (live + n > ts->ovf_cap) {
    if (int r = tile_store_count(ts, tg, s, &live, msg)) return r;
    ts->live_bound = live;
    if (live + n > ts->ovf_cap)
      if (int r = grow_store<D, R>(ts, tg, s, std::max(2 * ts->ovf_cap, live + n), msg)) return r;
  }
  const int cur = ts->cur;
  int ovf = 0;
  TS_CK(hipMemcpyAsync(&ovf, ts->ovf_count + cur, sizeof(int), hipMemcpyDeviceToHost, s));
  TS_CK(hipStreamSynchronize(s));
  // overflow records are live ones, so ovf + n <= live + n <= ovf_cap
  RecPtr dst = ts->buf[cur];
  const int64_t o = ts->ovf_base + ovf;
  dst.r = (char*)dst.r + rec_size(tg) * (size_t)o;
  if (dst.id) dst.id += o;
  fill(dst, ts->next_id);
  TS_CK(hipGetLastError());
  ts->next_id += n;
  ts->live_bound += n;
  const int total = ovf + (int)n;
  TS_CK(hipMemcpyAsync(ts->ovf_count + cur, &total, sizeof(int), hipMemcpyHostToDevice, s));
  TS_CK(hipStreamSynchronize(s));
  // a deferred window shift moves every record but the ones just written
  // (they were placed in the shifted box)
  const int shx = ts->pending_shx;
  ts->pending_shx = 0;
  return relayout_seg<D, R>(ts, tg, s, msg, shx, ovf);
}

// Appends n particles (reference layout, inside the box) to the species
// (add_particles / particles_inject, window.f90:39-270).
template <int D, typename R>
static int inject_impl(TileStore* ts, const TileGeom& tg, const DevGrid& g, const double* part,
                       int64_t ld, int64_t n, hipStream_t s, std::string* msg) {
  for (int64_t q = 0; q < n; q++) {
    unsigned long long b;
    std::memcpy(&b, part + 2 * D * ld + q, 8);
    const int c = (int)(short)(b & 0xffff), pi = (int)(short)((b >> 16) & 0xffff);
    if (c != ts->charge || pi != ts->part_ind) {
      *msg = std::string(path_name(tg)) +
             ": injected particles need the species' charge and part_ind";
      return ALADYN_EINVAL;
    }
  }
  const int ncol = 2 * D + 1;
  double* dpart = nullptr;
  TS_CK(hipMalloc(&dpart, (size_t)n * ncol * sizeof(double)));
  for (int c = 0; c < ncol; c++)
    TS_CK(hipMemcpyAsync(dpart + c * n, part + c * ld, (size_t)n * sizeof(double),
                         hipMemcpyHostToDevice, s));
  const int r = append_impl<D, R>(ts, tg, n, s, msg, [&](RecPtr dst, int64_t id0) {
    k_convert_in<D, R><<<grid_for(n), kB, 0, s>>>(g, tg.bx, tg.by, dpart, n, n, dst, id0);
  });
  (void)hipStreamSynchronize(s);
  (void)hipFree(dpart);
  return r;
}

// add_particles (window.f90:39-104) on the device: record q of the columns
// c0 .. c0+ncol-1 of the target table, in the reference's order (column
// fastest, then y row j, then z row k), at x = xpt, y = ypt[j], z = zpt[k],
// p = 0, with the FP32 weight real(wghpt*wghyz(j,k), sp).
template <int D, typename R>
__global__ void k_table_in(DevGrid g, int bx, int by, InjTableDev tb, int64_t c0, int64_t ncol,
                           int64_t n, RecPtr U, int64_t id0) {
  const double mn[3] = {g.xmn, g.ymn, g.zmn};
  const double inv[3] = {g.dx_inv, g.dy_inv, g.dz_inv};
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ix = c0 + q % ncol, r = q / ncol;
    const int j = (int)(r % tb.nyp), k = (int)(r / tb.nyp);
    const double x[3] = {tb.xpt[ix], tb.ypt[j], D == 3 ? tb.zpt[k] : 0.0};
    const float w = D == 3 ? (float)(tb.wghpt[ix] * tb.wghyz[j + (int64_t)tb.nyp * k])
                           : (float)(tb.wghyz[j] * tb.wghpt[ix]);
    int N[3] = {1, 1, 1};
    R sv[3] = {R(0), R(0), R(0)}, pv[3] = {R(0), R(0), R(0)};
#pragma unroll
    for (int d = 0; d < D; d++) {
      double sd;
      ref_node(x[d], mn[d], inv[d], N[d], sd);
      R sf = (R)sd;
      if (sf >= R(0.5)) {  // keep the [-1/2, 1/2) convention after rounding
        N[d] += 1;
        sf -= R(1);
      }
      sv[d] = sf;
    }
    const int cell = N[0] | (N[1] << bx) | (D == 3 ? (N[2] << (bx + by)) : 0);
    put_rec<D, R>(U, q, cell, sv, pv, w);
    if (U.id) U.id[q] = (int)(id0 + q);
  }
}

int tile_store_inject_table(TileStore* ts, const TileGeom& tg, const DevGrid& g,
                            const InjTableDev& tb, int64_t c0, int64_t ncol, hipStream_t s,
                            std::string* msg) {
  const int64_t n = ncol * tb.nyp * (g.ndim == 3 ? tb.nzp : 1);
  if (n <= 0) return 0;
  if (!ts->loaded) {
    *msg = "tile store: table injection needs a species that was put first";
    return ALADYN_ESTATE;
  }
  ts->layout_version++;
  return with_dr(tg, g.ndim, [&](auto dc, auto rv) {
    constexpr int DD = decltype(dc)::value;
    using R = decltype(rv);
    return append_impl<DD, R>(ts, tg, n, s, msg, [&](RecPtr dst, int64_t id0) {
      k_table_in<DD, R><<<grid_for(n), kB, 0, s>>>(g, tg.bx, tg.by, tb, c0, ncol, n, dst, id0);
    });
  });
}

int tile_store_inject(TileStore* ts, const TileGeom& tg, const DevGrid& g, const double* part,
                      int64_t ld, int64_t n, hipStream_t s, bool track_ids, std::string* msg) {
  if (n <= 0) return 0;
  if (!ts->loaded) return tile_store_put(ts, tg, g, part, ld, n, s, track_ids, msg);
  ts->layout_version++;
  return with_dr(tg, g.ndim, [&](auto dc, auto rv) {
    return inject_impl<decltype(dc)::value, decltype(rv)>(ts, tg, g, part, ld, n, s, msg);
  });
}

int tile_store_count(TileStore* ts, const TileGeom& tg, hipStream_t s, int64_t* np,
                     std::string* msg) {
  *np = 0;
  if (!ts->loaded) return 0;
  const int cur = ts->cur, oth = 1 - cur;
  int* live = ts->count[oth];
  if (tg.rbytes == 8)
    k_count_live<double><<<tg.ntiles + 1, kB, 0, s>>>(ts->buf[cur], ts->seg_start, ts->seg_cap,
                                                      ts->count[cur], tg.ntiles,
                                                      ts->ovf_count + cur, live);
  else
    k_count_live<float><<<tg.ntiles + 1, kB, 0, s>>>(ts->buf[cur], ts->seg_start, ts->seg_cap,
                                                     ts->count[cur], tg.ntiles,
                                                     ts->ovf_count + cur, live);
  k_scan_tiles<<<1, kScanB, 0, s>>>(live, tg.ntiles + 1, ts->scratch, nullptr);
  TS_CK(hipGetLastError());
  int tot = 0;
  TS_CK(hipMemcpyAsync(&tot, ts->scratch + tg.ntiles + 1, sizeof(int), hipMemcpyDeviceToHost, s));
  TS_CK(hipStreamSynchronize(s));
  *np = tot;
  return 0;
}

int tile_store_get(TileStore* ts, const TileGeom& tg, const DevGrid& g, double* part, int64_t ld,
                   hipStream_t s, std::string* msg) {
  if (!ts->loaded) return 0;
  const int D = g.ndim, ncol = 2 * D + 1;
  int64_t n = 0;
  int r = with_dr(tg, D, [&](auto dc, auto rv) {
    return gather_live<decltype(dc)::value, decltype(rv)>(ts, tg, s, &n, msg);
  });
  if (r) return r;
  if (n == 0) return 0;
  if (ld < n) {
    *msg = "get_species: ld smaller than the particle count";
    return ALADYN_EINVAL;
  }
  double* d = nullptr;
  TS_CK(hipMalloc(&d, (size_t)n * ncol * sizeof(double)));
  const RecPtr& U = ts->buf[1 - ts->cur];
  with_dr(tg, D, [&](auto dc, auto rv) {
    k_convert_out<decltype(dc)::value, decltype(rv)>
        <<<grid_for(n), kB, 0, s>>>(g, tg.bx, tg.by, U, n, d, n);
    return 0;
  });
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

template <typename R>
int tile_store_step_enqueue(TileStore* ts, const TileGeom& tg, const DevGrid& g, const R* ebf,
                            R* jc, double* patch, R alp, int initial, bool j_is_zero,
                            double* jraw, hipStream_t s, hipEvent_t ev_begin, hipEvent_t ev_end,
                            std::string* msg, bool normalise) {
  if (!ts->loaded) return 0;
  if ((int)sizeof(R) != tg.rbytes) {
    *msg = "tile store: step called with the other precision's arrays";
    return ALADYN_EINVAL;
  }
  const int D = g.ndim, cur = ts->cur, nxt = 1 - cur;
  StepConst<R> sc;
  sc.alp = alp;
  sc.qch = (R)ts->charge;
  sc.dt = (R)g.dt;
  sc.inv[0] = (R)g.dx_inv;
  sc.inv[1] = (R)g.dy_inv;
  sc.inv[2] = (R)g.dz_inv;
  sc.nn[0] = g.nx;
  sc.nn[1] = g.ny;
  sc.nn[2] = g.nz;
  sc.initial = initial;
  sc.ovf_base = ts->ovf_base;
  tile_store_clear_next(ts, tg, s);
  if (ev_begin) TS_CK(hipEventRecord(ev_begin, s));
  TS_CK(launch_tile_kernel<R>(tg, g, sc, ebf, ts->buf[cur], ts->buf[nxt], ts->seg_start,
                              ts->seg_cap, ts->count[cur], ts->count[nxt], ts->ovf_count + cur,
                              ts->ovf_count + nxt, patch, ts->grow, s));
  if (ev_end) TS_CK(hipEventRecord(ev_end, s));
  // the overflowed particles add their current to the patches, then the
  // patches are summed onto jc
  if (D == 2)
    k_ovf_particles<2, R><<<512, kB, 0, s>>>(tg, g, sc, ebf, ts->buf[cur], ts->buf[nxt],
                                             ts->seg_start, ts->seg_cap, ts->ovf_count + cur,
                                             ts->count[nxt], ts->ovf_count + nxt, patch, ts->grow);
  else
    k_ovf_particles<3, R><<<512, kB, 0, s>>>(tg, g, sc, ebf, ts->buf[cur], ts->buf[nxt],
                                             ts->seg_start, ts->seg_cap, ts->ovf_count + cur,
                                             ts->count[nxt], ts->ovf_count + nxt, patch, ts->grow);
  {
    const dim3 fb((g.n1p + 63) / 64, (g.n2p + 3) / 4, g.n3p);
    k_fold<R><<<fb, kB, 0, s>>>(tg, g, patch, jc, j_is_zero ? 0 : 1, jscale_of(g, normalise), D);
    if (jraw) k_fold<double><<<fb, kB, 0, s>>>(tg, g, patch, jraw, 0, jscale_of(g, false), D);
  }
  TS_CK(hipGetLastError());
  // the overflow count of this step goes to the pinned slot of this step
  TS_CK(hipMemcpyAsync(ts->h_ovf + ts->ovf_slot, ts->ovf_count + nxt, sizeof(int),
                       hipMemcpyDeviceToHost, s));
  return 0;
}

// The host side of a step whose work tile_store_step_enqueue queued (and
// that may have run inside a graph launched since).
int tile_store_step_finish(TileStore* ts, const TileGeom& tg, const DevGrid& g, hipStream_t s,
                           std::string* msg) {
  if (!ts->loaded) return 0;
  ts->cur = 1 - ts->cur;
  ts->steps_since_sort++;
  // Re-layout on cadence, or as soon as the overflow list is long enough to
  // cost more than a re-layout amortises (its particles deposit with global
  // atomics).  The count of this step was copied into one pinned slot; the
  // host waits only for the previous step's slot, so one step stays queued
  // on the GPU while the host decides.
  const int slot = ts->ovf_slot, prev = 1 - slot;
  TS_CK(hipEventRecord(ts->ovf_ev[slot], s));
  ts->ovf_valid[slot] = true;
  ts->ovf_slot = prev;
  bool relayout = ts->steps_since_sort >= tg.sort_every;
  if (!relayout && ts->ovf_valid[prev]) {
    TS_CK(hipEventSynchronize(ts->ovf_ev[prev]));
    if (ts->h_ovf[prev] > (int)(ts->ovf_cap / 8192 + 2048)) relayout = true;
  }
  if (relayout) {
    ts->ovf_valid[0] = ts->ovf_valid[1] = false;  // counts of the old layout
    return tile_store_relayout(ts, tg, g, s, msg);
  }
  return 0;
}

template <typename R>
int tile_store_step(TileStore* ts, const TileGeom& tg, const DevGrid& g, const R* ebf, R* jc,
                    double* patch, R alp, int initial, bool j_is_zero, double* jraw,
                    hipStream_t s, hipEvent_t ev_begin, hipEvent_t ev_end, std::string* msg,
                    bool normalise) {
  if (int r = tile_store_step_enqueue<R>(ts, tg, g, ebf, jc, patch, alp, initial, j_is_zero,
                                         jraw, s, ev_begin, ev_end, msg, normalise))
    return r;
  return tile_store_step_finish(ts, tg, g, s, msg);
}
#define ALADYN_STEP_INST(R)                                                                    \
  template int tile_store_step<R>(TileStore*, const TileGeom&, const DevGrid&, const R*, R*,   \
                                  double*, R, int, bool, double*, hipStream_t, hipEvent_t,     \
                                  hipEvent_t, std::string*, bool);                             \
  template int tile_store_step_enqueue<R>(TileStore*, const TileGeom&, const DevGrid&,         \
                                          const R*, R*, double*, R, int, bool, double*,        \
                                          hipStream_t, hipEvent_t, hipEvent_t, std::string*,   \
                                          bool);
ALADYN_STEP_INST(float)
ALADYN_STEP_INST(double)
#undef ALADYN_STEP_INST

// ------------------------------------------------------ batched runs
template <typename R>
void tile_store_batch_desc(TileStore* ts, const TileGeom& tg, const DevGrid& g, const R* ebf,
                           R* jc, double* patch, R alp, int initial, TileRun<R>* d) {
  const int cur = ts->cur, nxt = 1 - cur;
  d->tg = tg;
  d->g = g;
  d->ta = TileArgs{tg.nt[0], tg.nt[1], tg.bx, tg.by, g.n1p, g.n2p, g.n3p, g.ncell};
  StepConst<R>& sc = d->sc;
  sc.alp = alp;
  sc.qch = (R)ts->charge;
  sc.dt = (R)g.dt;
  sc.inv[0] = (R)g.dx_inv;
  sc.inv[1] = (R)g.dy_inv;
  sc.inv[2] = (R)g.dz_inv;
  sc.nn[0] = g.nx;
  sc.nn[1] = g.ny;
  sc.nn[2] = g.nz;
  sc.initial = initial;
  sc.ovf_base = ts->ovf_base;
  d->ebf = ebf;
  d->jc = jc;
  d->A = ts->buf[cur];
  d->B = ts->buf[nxt];
  d->seg_start = ts->seg_start;
  d->seg_cap = ts->seg_cap;
  d->cnt_cur = ts->count[cur];
  d->cnt_next = ts->count[nxt];
  d->ovf_cur = ts->ovf_count + cur;
  d->ovf_next = ts->ovf_count + nxt;
  d->patch = patch;
  d->grow = ts->grow;
}

template <typename R>
hipError_t tile_store_batch_launch(const TileGeom& tg0, int ndim, const TileRun<R>* runs,
                                   int nruns, const int2* work, int nwork, int max_ntiles,
                                   int64_t max_ncell, int* ovf_out, hipStream_t s) {
  kb_clear<R><<<dim3(grid_for(max_ntiles, 64), nruns), kB, 0, s>>>(runs);
  if (hipError_t e = launch_tile_batch<R>(tg0, ndim, runs, work, nwork, s)) return e;
  if (ndim == 2)
    kb_ovf_particles<2, R><<<dim3(64, nruns), kB, 0, s>>>(runs);
  else
    kb_ovf_particles<3, R><<<dim3(64, nruns), kB, 0, s>>>(runs);
  // at most 16384 workgroups in all (each run's stride over its own nodes)
  const int64_t fcap = std::max(16, 16384 / std::max(nruns, 1));
  kb_fold<R><<<dim3(grid_for(max_ncell, fcap), nruns), kB, 0, s>>>(runs);
  kb_ovf_out<R><<<(nruns + 63) / 64, 64, 0, s>>>(runs, nruns, ovf_out);
  return hipGetLastError();
}

int tile_store_batch_finish(TileStore* ts, const TileGeom& tg, const DevGrid& g, int ovf_prev,
                            hipStream_t s, std::string* msg) {
  if (!ts->loaded) return 0;
  ts->cur = 1 - ts->cur;
  ts->steps_since_sort++;
  // the single-run path's pinned overflow slots do not hold this step
  ts->ovf_valid[0] = ts->ovf_valid[1] = false;
  const bool relayout = ts->steps_since_sort >= tg.sort_every ||
                        (ovf_prev >= 0 && ovf_prev > (int)(ts->ovf_cap / 8192 + 2048));
  return relayout ? tile_store_relayout(ts, tg, g, s, msg) : 0;
}

#define ALADYN_BATCH_INST(R)                                                                     \
  template void tile_store_batch_desc<R>(TileStore*, const TileGeom&, const DevGrid&, const R*, \
                                         R*, double*, R, int, TileRun<R>*);                      \
  template hipError_t tile_store_batch_launch<R>(const TileGeom&, int, const TileRun<R>*, int,  \
                                                 const int2*, int, int, int64_t, int*,           \
                                                 hipStream_t);
ALADYN_BATCH_INST(float)
ALADYN_BATCH_INST(double)
#undef ALADYN_BATCH_INST

// ------------------------------------------------------------ diagnostics
template <int D, typename R>
__global__ void k_grid_charge_rec(TileGeom tg, RecPtr U, int64_t n, double qch, double* rho,
                                  DevGrid g) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int cell = rec_cell<R>(U, q);
    const int N[3] = {cell & ((1 << tg.bx) - 1), (cell >> tg.bx) & ((1 << tg.by) - 1),
                      D == 3 ? (cell >> (tg.bx + tg.by)) : 1};
    double a[3][3] = {{0, 0, 0}, {0, 0, 0}, {1, 0, 0}};
#pragma unroll
    for (int d = 0; d < D; d++) qweights<double>((double)rec_s<R>(U, q, d), a[d]);
    // wght = charge*wgh in FP32 (grid_part_util.f90:280-286)
    const double wq = (double)((float)qch * rec_w<R>(U, q));
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
  with_dr(tg, D, [&](auto dc, auto rv) {
    constexpr int DD = decltype(dc)::value;
    using R = decltype(rv);
    if (gather_live<DD, R>(ts, tg, s, &n, &msg) || n == 0) return 0;
    const RecPtr& U = ts->buf[1 - ts->cur];
    k_grid_charge_rec<DD, R><<<grid_for(n), kB, 0, s>>>(tg, U, n, (double)ts->charge, rho, g);
    return 0;
  });
}

__device__ inline void amax_nonneg(double* addr, double v) {
  atomicMax((unsigned long long*)addr, (unsigned long long)__double_as_longlong(v));
}

template <int D, typename R>
__global__ __launch_bounds__(kB) void k_energy_rec(TileGeom tg, DevGrid g, RecPtr U, int64_t n,
                                                   double qch, double* out) {
  __shared__ double sh[8][kB / 64];
  double ek[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    double p[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; d++) p[d] = (double)rec_p<R>(U, q, d);
    const double w = (double)rec_w<R>(U, q);
    const double gam1 = sqrt(1. + p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) - 1.;
    if constexpr (D == 2) {
      ek[0] += w * p[0];
      ek[1] += w * p[1];
    } else {
      const int cell = rec_cell<R>(U, q);
      const int Nj = (cell >> tg.bx) & ((1 << tg.by) - 1), Nk = cell >> (tg.bx + tg.by);
      const double y = g.ymn + ((double)(Nj - 3) + (double)rec_s<R>(U, q, 1)) * g.dy;
      const double z = g.zmn + ((double)(Nk - 3) + (double)rec_s<R>(U, q, 2)) * g.dz;
      ek[0] += p[0];
      ek[1] += p[1];
      ek[2] += p[2];
      ek[3] += w * (y * p[2] - z * p[1]);
    }
    ek[5] += (double)((float)qch * (float)w);
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
  with_dr(tg, D, [&](auto dc, auto rv) {
    constexpr int DD = decltype(dc)::value;
    using R = decltype(rv);
    if (gather_live<DD, R>(ts, tg, s, &n, &msg) || n == 0) return 0;
    const RecPtr& U = ts->buf[1 - ts->cur];
    k_energy_rec<DD, R><<<grid_for(n, 4096), kB, 0, s>>>(tg, g, U, n, (double)ts->charge, out);
    return 0;
  });
}

}  // namespace aladyn
