// tile_store.hpp -- the particle store of the tile path (FP32 and FP64).
//
// Particles of one species live in records grouped by tile: tile t owns
// the segment [seg_start[t], seg_start[t] + seg_cap[t]) of which the first
// count[t] slots are live.  A particle is stored as
//   cell  int32  packed nearest-node indices (i | j << bx | k << (bx+by)),
//                Fortran node numbering of ebf (interior starts at 3)
//   s[d]  R      offset from that node, in cells, in [-1/2, 1/2)
//   p[d]  R      momentum p/(mc)
//   w     FP32   weight (the FP32 `wgh` of the packed slot)
// with R = float (32-byte record) or double (64-byte record)
// so x = xmn + (i - 3 + s_x)*dx exactly as the reference's local coordinate
// xx = 3 + (x - xmn)/dx = i + s_x (grid_part_lib.f90:114-116); a moving-window
// shift is an integer subtraction.  Two such buffers are used in turn: the
// particle kernel reads the tile's segment from one and writes every
// surviving particle into the segment of its new tile in the other (slots
// reserved with per-workgroup aggregated atomics), so the store stays
// tile-sorted every step at the cost of one read and one write per particle.
// Particles that find their destination segment full go to an overflow list
// at the end of the buffer, which a global-memory kernel processes the next
// step; a re-layout (counting sort with fresh capacities) runs every
// sort_every steps and whenever the overflow grows.
#pragma once

#include <string>

#include "../../include/aladyn_gpu.h"
#include "lp_common.hpp"

namespace aladyn {

struct TileGeom {
  int T[3];    // tile extent in nodes (T[2] = 1 in 2-D)
  int nt[3];   // tiles per axis
  int ntiles;
  int G[3];    // field patch extent: T + 2 (nodes T0-1 .. T0+T)
  int J[3];    // current patch extent: T + 4 (nodes T0-2 .. T0+T+1)
  int gvol, pvol;
  int bx, by;  // bit widths of i and j in the packed cell
  int block;   // threads per workgroup
  int sort_every;
  int lds_bytes;
  int rbytes;  // bytes of the record's real type: 4 (FP32 path) or 8 (FP64 path)
};

// Particle records of the tile store: one record per particle, 32 bytes in
// the FP32 path and 64 in the FP64 path (RecWord in tile_device.hpp gives
// the layout), so a record moves with two or four 16-byte loads or stores.
struct RecPtr {
  void* r;  // records
  int* id;  // optional original index (tests), nullptr when not tracked
};

struct TileStore {
  int64_t alloc = 0;     // records per buffer
  int64_t ovf_base = 0;  // first record of the overflow region
  int64_t ovf_cap = 0;
  void* mem[2] = {nullptr, nullptr};
  RecPtr buf[2] = {};
  int cur = 0;
  int* seg_start = nullptr;  // ntiles + 1
  int* seg_cap = nullptr;    // ntiles
  int* seg_start2 = nullptr;  // the next layout's, swapped in at a re-layout
  int* seg_cap2 = nullptr;
  int* count[2] = {nullptr, nullptr};
  int* ovf_count = nullptr;  // [2]
  int* scratch = nullptr;    // ntiles + 1 ints for scans / fills
  int64_t* d_total = nullptr;
  int charge = 0, part_ind = 0;
  bool track_ids = false;
  bool loaded = false;
  // bound on the live records: exact when the store is binned from a dense
  // list (put, growth), plus every append since (between them particles only
  // leave)
  int64_t live_bound = 0;
  int64_t next_id = 0;  // id of the next injected particle (ids tracked)
  int steps_since_sort = 0;
  // a window shift (cells) deferred to the re-layout of the injection that
  // follows it at the same move, so that one pass does both
  int pending_shx = 0;
  // bumped by every change of the layout or of the buffers (put, injection,
  // window, re-layout): work captured into a graph is valid for one version
  int64_t layout_version = 0;
  int* grow = nullptr;   // ntiles flags: the tile spilled since the last re-layout
  // overflow count of recent steps, read back through two pinned slots: the
  // host waits only for the step before the one just queued
  int* h_ovf = nullptr;  // pinned [2]
  hipEvent_t ovf_ev[2] = {nullptr, nullptr};
  bool ovf_valid[2] = {false, false};
  int ovf_slot = 0;
};

// Constants of one species' step (alp = dt/2 * lorentz_fact, the charge,
// the cell inverses and counts, the first-step flag, the overflow region).
template <typename R>
struct StepConst {
  R alp, qch, dt;
  R inv[3];
  int nn[3];
  int initial;
  int64_t ovf_base;
};

struct TileArgs {
  int nt0, nt1, bx, by;
  int n1p, n2p, n3p;
  int64_t ncell;
};

// Everything one run's particle kernels of a step read: the batched launches
// (many independent runs in one launch) take an array of these.  A and B are
// the buffers read and written this step; cnt/ovf the matching counters.
template <typename R>
struct TileRun {
  TileGeom tg;
  DevGrid g;
  TileArgs ta;
  StepConst<R> sc;
  const R* ebf;
  R* jc;
  RecPtr A, B;
  const int* seg_start;
  const int* seg_cap;
  const int* cnt_cur;
  int* cnt_next;
  int* ovf_cur;
  int* ovf_next;
  double* patch;
  int* grow;
};

// Factors the fold applies to each current component as it writes jc: all
// ones for the raw current, djc when the fold stands in for the
// normalisation of curr_mpi_collect (curr_and_fields_util.f90:115-119).
struct JScale {
  double s[3];
};
__host__ __device__ inline JScale jscale_of(const DevGrid& g, bool normalise) {
  return normalise ? JScale{{g.djc[0], g.djc[1], g.djc[2]}} : JScale{{1.0, 1.0, 1.0}};
}

int tile_geom_init(const DevGrid& g, const aladyn_config& cfg, TileGeom* tg);
// The tile kernel with the envelope step's advance in place of the LP push
// (je = je(1:D+1)); hipErrorInvalidConfiguration for a workgroup size it is
// not compiled for.
template <typename R>
hipError_t launch_env_move_kernel(const TileGeom& tg, const DevGrid& g, const StepConst<R>& sc,
                                  const R* je, const RecPtr& A, const RecPtr& B,
                                  const int* seg_start, const int* seg_cap, const int* cnt_cur,
                                  int* cnt_next, int* ovf_cur, int* ovf_next, double* patch,
                                  int* grow, hipStream_t s);
const char* tile_geom_error(int code);

int tile_store_put(TileStore* ts, const TileGeom& tg, const DevGrid& g, const double* part,
                   int64_t ld, int64_t np, hipStream_t s, bool track_ids, std::string* msg);
int tile_store_count(TileStore* ts, const TileGeom& tg, hipStream_t s, int64_t* np,
                     std::string* msg);
// With defer, the shift is only recorded: the next append (injection) or
// tile_store_window_apply re-lays the store out with it.
int tile_store_window(TileStore* ts, const TileGeom& tg, const DevGrid& g, int window_cells,
                      hipStream_t s, std::string* msg, bool defer = false);
int tile_store_window_apply(TileStore* ts, const TileGeom& tg, const DevGrid& g, hipStream_t s,
                            std::string* msg);
int tile_store_inject(TileStore* ts, const TileGeom& tg, const DevGrid& g, const double* part,
                      int64_t ld, int64_t n, hipStream_t s, bool track_ids, std::string* msg);
// Device copy of a target table for injection at window moves (xpt,
// wghpt: ncolumns; ypt: nyp; zpt: nzp; wghyz: nyp*nzp, j fastest).
struct InjTableDev {
  const double* xpt;
  const double* ypt;
  const double* zpt;
  const double* wghpt;
  const double* wghyz;
  int nyp, nzp;
};
// Appends the particles of the ncol table columns from c0 (add_particles,
// window.f90:39-104, cold): charge and part_ind are the species'.
int tile_store_inject_table(TileStore* ts, const TileGeom& tg, const DevGrid& g,
                            const InjTableDev& tb, int64_t c0, int64_t ncol, hipStream_t s,
                            std::string* msg);
int tile_store_get(TileStore* ts, const TileGeom& tg, const DevGrid& g, double* part, int64_t ld,
                   hipStream_t s, std::string* msg);
// One species-loop body: gather + push + box test + deposit + re-binning.
// jc(raw) = (j_is_zero ? 0 : jc) + this species' Esirkepov current.  R is
// float or double and must match the precision the geometry was made for.
// The per-tile current patches are FP64 in both paths.  jraw, when given,
// receives this species' raw current alone, summed in FP64 (the
// continuity diagnostic).  With normalise, this species' current goes onto
// jc multiplied by djc (the normalisation a collect without it then skips).
template <typename R>
int tile_store_step(TileStore* ts, const TileGeom& tg, const DevGrid& g, const R* ebf, R* jc,
                    double* patch, R alp, int initial, bool j_is_zero, double* jraw,
                    hipStream_t s, hipEvent_t ev_begin, hipEvent_t ev_end, std::string* msg,
                    bool normalise = false);
// The two halves of tile_store_step: the device work (kernels and the copy
// of the overflow count, nothing that waits, so it may be captured into a
// graph) and the host side (buffer swap, overflow event, re-layout
// decision), which runs after the work was launched.
template <typename R>
int tile_store_step_enqueue(TileStore* ts, const TileGeom& tg, const DevGrid& g, const R* ebf,
                            R* jc, double* patch, R alp, int initial, bool j_is_zero,
                            double* jraw, hipStream_t s, hipEvent_t ev_begin, hipEvent_t ev_end,
                            std::string* msg, bool normalise = false);
int tile_store_step_finish(TileStore* ts, const TileGeom& tg, const DevGrid& g, hipStream_t s,
                           std::string* msg);
int tile_store_relayout(TileStore* ts, const TileGeom& tg, const DevGrid& g, hipStream_t s,
                        std::string* msg);
// jc = (accumulate != 0 ? jc : 0) + the sum of the per-tile FP64 current
// patches (times djc with normalise), rounded to the grid's type once per node
template <typename R>
void tile_store_fold(const TileGeom& tg, const DevGrid& g, const double* patch, R* jc,
                     int accumulate, hipStream_t s, bool normalise = false);
// The counters a step's particle kernels fill (the next buffer's tile counts
// and its overflow count) set to zero, in one launch.
void tile_store_clear_next(TileStore* ts, const TileGeom& tg, hipStream_t s);
// den = the sum of component 0 of the per-tile patches (a density)
void tile_store_fold_density(const TileGeom& tg, const DevGrid& g, const double* patch,
                             double* den, hipStream_t s);
// Batched runs: the step's descriptor of one store (the buffers it reads and
// writes this step), the batched particle-side launches of a step (counter
// clear, tile kernel over the work list of (run, tile) pairs, overflow
// kernel, fold of the patches onto jc times djc, the overflow counts into ovf_out),
// and the host side of one store after such a step (buffer swap, re-layout
// on cadence or when ovf_prev, the previous step's overflow count, is large).
template <typename R>
void tile_store_batch_desc(TileStore* ts, const TileGeom& tg, const DevGrid& g, const R* ebf,
                           R* jc, double* patch, R alp, int initial, TileRun<R>* d);
template <typename R>
hipError_t tile_store_batch_launch(const TileGeom& tg0, int ndim, const TileRun<R>* runs,
                                   int nruns, const int2* work, int nwork, int max_ntiles,
                                   int64_t max_ncell, int* ovf_out, hipStream_t s);
int tile_store_batch_finish(TileStore* ts, const TileGeom& tg, const DevGrid& g, int ovf_prev,
                            hipStream_t s, std::string* msg);
void tile_store_grid_charge(TileStore* ts, const TileGeom& tg, const DevGrid& g, double* rho,
                            hipStream_t s);
void tile_store_energy(TileStore* ts, const TileGeom& tg, const DevGrid& g, double* out,
                       hipStream_t s);
void tile_store_free(TileStore* ts);

}  // namespace aladyn
