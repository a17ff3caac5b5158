// capi.hip -- the C-ABI of include/aladyn_gpu.h: context, device memory,
// host<->device transfers in the reference's Fortran layouts and the step
// orchestration of lpf2_evolve (pic_evolve_in_time.f90:37-109).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/aladyn_gpu.h"
#include "env.hpp"
#include "lp_kernels.hpp"
#include "tile_store.hpp"

using namespace aladyn;

struct TimerRing {
  static constexpr int kRing = 256;
  hipEvent_t a[kRing] = {}, b[kRing] = {};
  int64_t n = 0;      // pairs recorded since reset
  int64_t base = 0;   // total pairs recorded (ring position)
  int create() {
    for (int i = 0; i < kRing; i++) {
      if (hipEventCreate(&a[i]) != hipSuccess || hipEventCreate(&b[i]) != hipSuccess) return 1;
    }
    return 0;
  }
  void destroy() {
    for (int i = 0; i < kRing; i++) {
      if (a[i]) (void)hipEventDestroy(a[i]);
      if (b[i]) (void)hipEventDestroy(b[i]);
    }
  }
  hipEvent_t begin() const { return a[base % kRing]; }
  hipEvent_t end() const { return b[base % kRing]; }
  void commit() {
    base++;
    n++;
  }
  // average ms over the last min(n, kRing) pairs (stream must be idle)
  double average() const {
    const int64_t m = n < kRing ? n : kRing;
    double sum = 0;
    int64_t cnt = 0;
    for (int64_t q = 1; q <= m; q++) {
      const int i = (int)((base - q) % kRing);
      float ms = 0;
      if (hipEventElapsedTime(&ms, a[i], b[i]) == hipSuccess) {
        sum += ms;
        cnt++;
      }
    }
    return cnt ? sum / cnt : 0.0;
  }
};

// Target table of one species for injection at window moves (the arrays
// particles_inject / add_particles read, window.f90:39-270).
struct InjTable {
  bool active = false;
  std::vector<double> xpt;  // ascending column positions
  int nyp = 0, nzp = 0;
  double* dev = nullptr;    // xpt | ypt | zpt | wghpt | wghyz on the device
  InjTableDev tb{};
  int64_t nptx = 0;         // columns already consumed (nptx(ic))
  double targ_in = 0.0, targ_end = 0.0;
};

struct aladyn_ctx {
  aladyn_config cfg{};
  DevGrid g{};
  hipStream_t stream = nullptr;
  std::string err;
  int dev = 0;
  bool have_fields = false;
  // grids of the configured precision (the other pair stays null)
  double* ebf64 = nullptr;
  double* jc64 = nullptr;
  float* ebf32 = nullptr;
  float* jc32 = nullptr;
  // tile-sorted particle stores, one per species
  TileStore ts[ALADYN_MAX_SPECIES];
  TileGeom tg{};
  double* patch = nullptr;  // per-tile J partial sums (FP64 in both paths)
  bool j_is_zero = true;
  double* d_red = nullptr;      // reduction outputs (16 doubles)
  double* h_red = nullptr;      // pinned mirror
  // diagnostics (continuity)
  bool diag = false;
  double* rho0 = nullptr;
  double* rho1 = nullptr;
  double* jraw = nullptr;  // the last stage's raw current, summed in FP64
  bool have_resid = false;
  // lp_window_xshift saved state (window.f90:278-280)
  double win_xlapse = 0.0;
  int win_wi1 = 0;
  InjTable inj[ALADYN_MAX_SPECIES];
  // envelope model (aladyn_put_envelope): env(n1p, n2p, n3p, 4) = A_R^n,
  // A_I^n, A_R^{n-1}, A_I^{n-1}; je(n1p, n2p, n3p, curr_ndim + 1) the
  // envelope step's reuse of jc; den the FP64 density source
  bool have_env = false;
  double env_oml = 0.0;
  double* env64 = nullptr;
  float* env32 = nullptr;
  double* je64 = nullptr;
  float* je32 = nullptr;
  double* den = nullptr;
  // timing: event pairs around the particle kernel and the field solve,
  // recorded on the context stream (ring of kRing steps)
  TimerRing tpart, tfield;
  // envelope steps: the two particle kernels (push, move)
  TimerRing tenv_push, tenv_move;
  // captured lpf2_evolve steps (aladyn_set_graphs): one executable graph per
  // state of the stores and of the box, see step_key
  bool use_graphs = false;
  int64_t grid_version = 0;  // bumped when the box moves
  struct StepGraph {
    std::vector<int64_t> key;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<StepGraph> graphs;
};

static void clear_graphs(aladyn_ctx* c) {
  for (auto& gr : c->graphs)
    if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
  c->graphs.clear();
}

static int fail(aladyn_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCK(ctx, call)                                                                  \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(ctx, ALADYN_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                                    \
  } while (0)

static bool f64(const aladyn_ctx* c) { return c->cfg.precision == ALADYN_PREC_F64; }
static int noalloc_scratch_init(aladyn_ctx* c);

// ----------------------------------------------------------------- set-up
static int setup_grid(aladyn_ctx* c) {
  const aladyn_config& a = c->cfg;
  DevGrid& g = c->g;
  std::memset(&g, 0, sizeof g);
  g.ndim = a.ndim;
  g.nfield = a.ndim == 3 ? 6 : 3;
  g.curr_ndim = a.ndim == 3 ? 3 : 2;
  g.nx = a.nx;
  g.ny = a.ny;
  g.nz = a.ndim == 3 ? a.nz : 1;
  // array_alloc.f90:62-69: n1p = nxp + ihx with nxp = nx + 2, ihx = 3
  g.n1p = g.nx + 5;
  g.n2p = g.ny + 5;
  g.n3p = a.ndim == 3 ? g.nz + 5 : 1;
  g.ix1 = 3;
  g.ix2 = g.nx + 2;
  g.jy1 = 3;
  g.jy2 = g.ny + 2;
  g.kz1 = a.ndim == 3 ? 3 : 1;
  g.kz2 = a.ndim == 3 ? g.nz + 2 : 1;
  g.ibx = a.ibx;
  g.iby = a.iby;
  g.ibz = a.ibz;
  g.ncell = (int64_t)g.n1p * g.n2p * g.n3p;
  // set_grid (set_grid_param.f90:68-71, 107-113, 158-163)
  g.dx = a.dx;
  g.dx_inv = 1.0 / g.dx;
  g.dy = a.dy;
  g.dy_inv = 1.0 / g.dy;
  g.dz = a.ndim == 3 ? a.dz : 1.0;
  g.dz_inv = a.ndim == 3 ? 1.0 / g.dz : 1.0;
  g.xmn = 0.0;
  g.xmx = g.dx * (double)g.nx;
  g.ymn = g.dy * (double)(-(g.ny / 2));
  g.ymx = g.dy * (double)(g.ny - g.ny / 2);
  g.zmn = a.ndim == 3 ? g.dz * (double)(-(g.nz / 2)) : 0.0;
  g.zmx = a.ndim == 3 ? g.dz * (double)(g.nz - g.nz / 2) : 0.0;
  g.dt = a.dt;
  g.ompe = a.ompe;
  // djc (set_init_param.f90:92-102)
  if (a.ndim == 2) {
    g.djc[0] = 0.5 * g.dx;
    g.djc[1] = 0.5 * g.dy;
    g.djc[2] = 0.0;
  } else {
    g.djc[0] = g.dx / 3.;
    g.djc[1] = g.dy / 3.;
    g.djc[2] = g.dz / 3.;
  }
  // se_coeff (grid_field_param.f90:36-99)
  const double yx_rat = g.dy / g.dx;
  const double nu = a.cfl * yx_rat / std::sqrt(yx_rat * yx_rat + (double)a.ndim - 1.);
  if (a.der_ord == 3) {
    g.se[0] = 1. + 0.125 * (1. - nu * nu);
    g.se[1] = (1. - g.se[0]) / 3.;
  } else if (a.der_ord == 4) {
    g.se[0] = 1.125;
    g.se[1] = (1. - g.se[0]) / 3.;
  } else {
    g.se[0] = 1.;
    g.se[1] = 0.;
  }
  g.dgvol = a.ndim == 2 ? g.dx * g.dy * g.dy : g.dx * g.dy * g.dz;
  return 0;
}

extern "C" int aladyn_create(const aladyn_config* cfg, aladyn_ctx** out) {
  if (!cfg || !out) return ALADYN_EINVAL;
  *out = nullptr;
  aladyn_ctx* c = new aladyn_ctx();
  c->cfg = *cfg;
  const aladyn_config& a = c->cfg;
  auto bad = [&](const char* m) {
    *out = c;  // keep the context so the caller can read the message
    return fail(c, ALADYN_EINVAL, "%s", m);
  };
  if (a.ndim != 2 && a.ndim != 3) return bad("ndim must be 2 or 3");
  if (a.nx < 4 || a.ny < 4 || (a.ndim == 3 && a.nz < 4)) return bad("grid too small (>= 4 cells per axis)");
  if (a.ibx != 0 || a.iby != 0 || (a.ndim == 3 && a.ibz != 0))
    return bad("only open boundaries (ibx = iby = ibz = 0) are supported");
  if (a.der_ord != 2 && a.der_ord != 3 && a.der_ord != 4) return bad("der_ord must be 2, 3 or 4");
  if (a.precision != ALADYN_PREC_F64 && a.precision != ALADYN_PREC_F32)
    return bad("precision must be 64 or 32");
  if (a.nsp < 1 || a.nsp > ALADYN_MAX_SPECIES) return bad("nsp must be 1..4");
  if (!(a.dx > 0) || !(a.dy > 0) || (a.ndim == 3 && !(a.dz > 0)) || !(a.dt > 0))
    return bad("dx, dy, dz, dt must be positive");
  // |v| < c must keep every particle within one cell per step (the
  // Esirkepov window of grid_part_connect.f90:1008 holds 5 slots)
  if (a.dt >= a.dx || a.dt >= a.dy || (a.ndim == 3 && a.dt >= a.dz))
    return bad("dt must be smaller than every cell size (one-cell-per-step limit)");
  setup_grid(c);
  *out = c;
  HIPCK(c, hipGetDevice(&c->dev));
  if (int r = noalloc_scratch_init(c)) return r;
  HIPCK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  if (c->tpart.create() || c->tfield.create() || c->tenv_push.create() || c->tenv_move.create())
    return fail(c, ALADYN_EHIP, "event creation failed");
  const DevGrid& g = c->g;
  const size_t nf = (size_t)g.ncell * g.nfield, nj = (size_t)g.ncell * g.curr_ndim;
  if (f64(c)) {
    HIPCK(c, hipMalloc(&c->ebf64, nf * sizeof(double)));
    HIPCK(c, hipMalloc(&c->jc64, nj * sizeof(double)));
    HIPCK(c, hipMemsetAsync(c->ebf64, 0, nf * sizeof(double), c->stream));
    HIPCK(c, hipMemsetAsync(c->jc64, 0, nj * sizeof(double), c->stream));
  } else {
    HIPCK(c, hipMalloc(&c->ebf32, nf * sizeof(float)));
    HIPCK(c, hipMalloc(&c->jc32, nj * sizeof(float)));
    HIPCK(c, hipMemsetAsync(c->ebf32, 0, nf * sizeof(float), c->stream));
    HIPCK(c, hipMemsetAsync(c->jc32, 0, nj * sizeof(float), c->stream));
  }
  if (int r = tile_geom_init(g, a, &c->tg)) return fail(c, ALADYN_EINVAL, "%s", tile_geom_error(r));
  const size_t npatch = (size_t)c->tg.ntiles * c->tg.pvol * g.curr_ndim;
  HIPCK(c, hipMalloc(&c->patch, npatch * sizeof(double)));
  HIPCK(c, hipMalloc(&c->d_red, 32 * sizeof(double)));
  HIPCK(c, hipHostMalloc(&c->h_red, 32 * sizeof(double), hipHostMallocDefault));
  HIPCK(c, hipStreamSynchronize(c->stream));
  return ALADYN_OK;
}

extern "C" int aladyn_destroy(aladyn_ctx* c) {
  if (!c) return ALADYN_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  (void)hipFree(c->ebf64);
  (void)hipFree(c->jc64);
  (void)hipFree(c->ebf32);
  (void)hipFree(c->jc32);
  (void)hipFree(c->patch);
  for (auto& t : c->ts) tile_store_free(&t);
  for (auto& t : c->inj) (void)hipFree(t.dev);
  (void)hipFree(c->d_red);
  if (c->h_red) (void)hipHostFree(c->h_red);
  (void)hipFree(c->rho0);
  (void)hipFree(c->rho1);
  (void)hipFree(c->jraw);
  (void)hipFree(c->env64);
  (void)hipFree(c->env32);
  (void)hipFree(c->je64);
  (void)hipFree(c->je32);
  (void)hipFree(c->den);
  clear_graphs(c);
  c->tpart.destroy();
  c->tfield.destroy();
  c->tenv_push.destroy();
  c->tenv_move.destroy();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return ALADYN_OK;
}

extern "C" const char* aladyn_last_error(const aladyn_ctx* c) {
  return c ? c->err.c_str() : "null context";
}

// ---------------------------------------------------------------- species
static int check_species(aladyn_ctx* c, int32_t ic) {
  if (!c) return ALADYN_EINVAL;
  if (ic < 0 || ic >= c->cfg.nsp) return fail(c, ALADYN_EINVAL, "species %d out of range", ic);
  return 0;
}

// Positions must lie in the particle box (x > gmin, x <= gmax): every kernel
// index is derived from them, so this is the bounds guarantee.  Momenta
// must be finite.
static int check_particles(aladyn_ctx* c, const double* part, int64_t ld, int64_t np) {
  const DevGrid& g = c->g;
  const int D = g.curr_ndim;
  const double mn[3] = {g.xmn, g.ymn, g.zmn}, mx[3] = {g.xmx, g.ymx, g.zmx};
  for (int d = 0; d < D; d++)
    for (int64_t n = 0; n < np; n++) {
      const double x = part[d * ld + n];
      if (!(x > mn[d] && x <= mx[d]))
        return fail(c, ALADYN_ERANGE, "particle %lld coordinate %d = %g outside (%g, %g]",
                    (long long)n, d + 1, x, mn[d], mx[d]);
    }
  for (int d = D; d < 2 * D; d++)
    for (int64_t n = 0; n < np; n++)
      if (!std::isfinite(part[d * ld + n]))
        return fail(c, ALADYN_EINVAL, "particle %lld momentum is not finite", (long long)n);
  return 0;
}

extern "C" int aladyn_put_species(aladyn_ctx* c, int32_t ic, const double* part, int64_t ld,
                                  int64_t np) {
  if (int r = check_species(c, ic)) return r;
  if (np < 0 || (np > 0 && (!part || ld < np)))
    return fail(c, ALADYN_EINVAL, "put_species: need part != NULL and ld >= np");
  const DevGrid& g = c->g;
  if (int r = check_particles(c, part, ld, np)) return r;
  std::string msg;
  int r = tile_store_put(&c->ts[ic], c->tg, g, part, ld, np, c->stream, c->cfg.reserved[0] != 0,
                         &msg);
  if (r) return fail(c, r, "%s", msg.c_str());
  return ALADYN_OK;
}

extern "C" int aladyn_species_count(aladyn_ctx* c, int32_t ic, int64_t* np) {
  if (int r = check_species(c, ic)) return r;
  if (!np) return fail(c, ALADYN_EINVAL, "np is NULL");
  std::string msg;
  int r = tile_store_count(&c->ts[ic], c->tg, c->stream, np, &msg);
  if (r) return fail(c, r, "%s", msg.c_str());
  return ALADYN_OK;
}

// particles_inject (window.f90:106-139) after a move, gated as in
// lp_window_xshift (window.f90:339-341): with 1-based columns, i1 = nptx+1,
// the scan stops at the first xpt(ix) > xmax, i2 = ix-1 (i2 = ix when it
// stopped at the last column), nptx = i2, and columns i1..i2 are added only
// if i2 > i1.  Columns outside (xmin, xmax] are consumed but not added: the
// reference would remove them before their first deposit.
static int inject_from_table(aladyn_ctx* c, int ic) {
  InjTable& t = c->inj[ic];
  if (!t.active) return 0;
  const DevGrid& g = c->g;
  if (!(t.targ_in <= g.xmx && t.targ_end > g.xmx)) return 0;
  const int64_t nxp = (int64_t)t.xpt.size(), i1 = t.nptx + 1;
  int64_t i2 = i1 - 1;
  if (i1 <= nxp) {
    int64_t ix = i1;
    for (; ix <= nxp; ix++)
      if (t.xpt[ix - 1] > g.xmx) break;
    i2 = ix - 1;
    if (ix == nxp) i2 = ix;
  }
  t.nptx = i2;
  if (!(i2 > i1)) return 0;
  int64_t a = i1, b = i2;  // 1-based, inclusive
  while (a <= b && !(t.xpt[a - 1] > g.xmn)) a++;
  while (b >= a && t.xpt[b - 1] > g.xmx) b--;
  if (b < a) return 0;
  std::string msg;
  if (int r = tile_store_inject_table(&c->ts[ic], c->tg, g, t.tb, a - 1, b - a + 1, c->stream, &msg))
    return fail(c, r, "%s", msg.c_str());
  return 0;
}

extern "C" int aladyn_set_injection(aladyn_ctx* c, int32_t ic, const double* xpt, int64_t nxp,
                                    const double* ypt, int32_t nyp, const double* zpt,
                                    int32_t nzp, const double* wghpt, const double* wghyz,
                                    int64_t nptx, double targ_in, double targ_end) {
  if (int r = check_species(c, ic)) return r;
  InjTable& t = c->inj[ic];
  (void)hipFree(t.dev);
  t = InjTable();
  if (nxp == 0) return ALADYN_OK;  // table cleared
  const DevGrid& g = c->g;
  const int D = g.curr_ndim;
  if (D == 2) nzp = 1;
  if (nxp < 0 || nyp < 1 || nzp < 1 || !xpt || !ypt || !wghpt || !wghyz || (D == 3 && !zpt))
    return fail(c, ALADYN_EINVAL, "set_injection: empty or missing table arrays");
  if (nptx < 0 || nptx > nxp) return fail(c, ALADYN_EINVAL, "set_injection: nptx outside 0..nxp");
  for (int64_t i = 0; i < nxp; i++)
    if (!std::isfinite(xpt[i]) || !std::isfinite(wghpt[i]) || (i > 0 && xpt[i] < xpt[i - 1]))
      return fail(c, ALADYN_EINVAL, "set_injection: xpt must be finite and ascending");
  for (int j = 0; j < nyp; j++)
    if (!(ypt[j] > g.ymn && ypt[j] <= g.ymx))
      return fail(c, ALADYN_ERANGE, "set_injection: ypt[%d] = %g outside (%g, %g]", j, ypt[j],
                  g.ymn, g.ymx);
  if (D == 3)
    for (int k = 0; k < nzp; k++)
      if (!(zpt[k] > g.zmn && zpt[k] <= g.zmx))
        return fail(c, ALADYN_ERANGE, "set_injection: zpt[%d] = %g outside (%g, %g]", k, zpt[k],
                    g.zmn, g.zmx);
  for (int64_t q = 0; q < (int64_t)nyp * nzp; q++)
    if (!std::isfinite(wghyz[q])) return fail(c, ALADYN_EINVAL, "set_injection: wghyz not finite");
  t.xpt.assign(xpt, xpt + nxp);
  t.nyp = nyp;
  t.nzp = nzp;
  const size_t ny = (size_t)nyp, nz = (size_t)nzp, nx = (size_t)nxp;
  const size_t tot = nx + ny + nz + nx + ny * nz;
  HIPCK(c, hipMalloc(&t.dev, tot * sizeof(double)));
  double* d = t.dev;
  HIPCK(c, hipMemcpyAsync(d, xpt, nx * 8, hipMemcpyHostToDevice, c->stream));
  HIPCK(c, hipMemcpyAsync(d + nx, ypt, ny * 8, hipMemcpyHostToDevice, c->stream));
  if (D == 3) HIPCK(c, hipMemcpyAsync(d + nx + ny, zpt, nz * 8, hipMemcpyHostToDevice, c->stream));
  HIPCK(c, hipMemcpyAsync(d + nx + ny + nz, wghpt, nx * 8, hipMemcpyHostToDevice, c->stream));
  HIPCK(c, hipMemcpyAsync(d + 2 * nx + ny + nz, wghyz, ny * nz * 8, hipMemcpyHostToDevice,
                          c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));  // the host arrays are only borrowed
  t.tb = InjTableDev{d, d + nx, d + nx + ny, d + nx + ny + nz, d + 2 * nx + ny + nz, nyp, nzp};
  t.nptx = nptx;
  t.targ_in = targ_in;
  t.targ_end = targ_end;
  t.active = true;
  return ALADYN_OK;
}

extern "C" int aladyn_injection_state(aladyn_ctx* c, int32_t ic, int64_t* nptx) {
  if (int r = check_species(c, ic)) return r;
  if (!nptx) return fail(c, ALADYN_EINVAL, "nptx is NULL");
  *nptx = c->inj[ic].nptx;
  return ALADYN_OK;
}

extern "C" int aladyn_lp_window_xshift(aladyn_ctx* c, int32_t witr, int32_t init_iter,
                                       double w_speed, int32_t* nshx_out) {
  if (!c) return ALADYN_EINVAL;
  if (!nshx_out) return fail(c, ALADYN_EINVAL, "nshx is NULL");
  *nshx_out = 0;
  if (init_iter == 0) {  // window.f90:284-288
    c->win_xlapse = 0.0;
    c->win_wi1 = 0;
    return ALADYN_OK;
  }
  if (!(w_speed >= 0.0) || witr < 0)
    return fail(c, ALADYN_EINVAL, "window: w_speed and witr must be >= 0");
  DevGrid& g = c->g;
  const double xlapse = c->win_xlapse + w_speed * g.dt * witr;
  const long wi2 = std::lround(g.dx_inv * xlapse);
  const long nshx = wi2 - c->win_wi1;
  if (nshx >= g.ix2)  // the reference sets ier = 2 (window.f90:314-319)
    return fail(c, ALADYN_ERANGE, "window shift of %ld cells exceeds the box", nshx);
  c->win_xlapse = xlapse;
  c->win_wi1 = (int)wi2;
  *nshx_out = (int32_t)nshx;
  if (nshx != 0) {
    g.xmn += g.dx * nshx;
    g.xmx += g.dx * nshx;
    c->grid_version++;
    if (c->have_fields) {
      if (f64(c))
        fields_left_xshift<double>(g, c->ebf64, (int)nshx, c->stream);
      else
        fields_left_xshift<float>(g, c->ebf32, (int)nshx, c->stream);
      HIPCK(c, hipGetLastError());
    }
    if (c->have_env) {  // the envelope's four components (window.f90:327-331)
      if (f64(c))
        fields_left_xshift<double>(g, c->env64, (int)nshx, c->stream, 4);
      else
        fields_left_xshift<float>(g, c->env32, (int)nshx, c->stream, 4);
      HIPCK(c, hipGetLastError());
    }
    // particles behind the new left face leave (cell_part_dist(.true.)); the
    // shift waits for the re-layout of the injection below, so that one pass
    // over the store does both
    for (int ic = 0; ic < c->cfg.nsp; ic++) {
      std::string msg;
      if (int r = tile_store_window(&c->ts[ic], c->tg, g, (int)nshx, c->stream, &msg, true))
        return fail(c, r, "%s", msg.c_str());
    }
  }
  // plasma entering on the right from the registered target tables
  int rc = ALADYN_OK;
  for (int ic = 0; ic < c->cfg.nsp && rc == ALADYN_OK; ic++) rc = inject_from_table(c, ic);
  // species with nothing injected take the shift on their own
  for (int ic = 0; ic < c->cfg.nsp; ic++) {
    std::string msg;
    if (int r = tile_store_window_apply(&c->ts[ic], c->tg, g, c->stream, &msg))
      return rc != ALADYN_OK ? rc : fail(c, r, "%s", msg.c_str());
  }
  return rc;
}

extern "C" int aladyn_inject_species(aladyn_ctx* c, int32_t ic, const double* part, int64_t ld,
                                     int64_t np) {
  if (int r = check_species(c, ic)) return r;
  if (np < 0 || (np > 0 && (!part || ld < np)))
    return fail(c, ALADYN_EINVAL, "inject_species: need part != NULL and ld >= np");
  if (np == 0) return ALADYN_OK;
  if (int r = check_particles(c, part, ld, np)) return r;
  std::string msg;
  if (int r = tile_store_inject(&c->ts[ic], c->tg, c->g, part, ld, np, c->stream,
                                c->cfg.reserved[0] != 0, &msg))
    return fail(c, r, "%s", msg.c_str());
  return ALADYN_OK;
}

extern "C" int aladyn_set_box_origin(aladyn_ctx* c, double xmin) {
  if (!c) return ALADYN_EINVAL;
  if (!std::isfinite(xmin)) return fail(c, ALADYN_EINVAL, "set_box_origin: xmin not finite");
  for (int ic = 0; ic < ALADYN_MAX_SPECIES; ic++)
    if (c->ts[ic].loaded)
      return fail(c, ALADYN_ESTATE, "set_box_origin: call it before put_species");
  DevGrid& g = c->g;
  g.xmn = xmin;
  g.xmx = xmin + g.dx * (double)g.nx;
  c->grid_version++;
  return ALADYN_OK;
}

extern "C" int aladyn_box(aladyn_ctx* c, double out[6]) {
  if (!c || !out) return ALADYN_EINVAL;
  const DevGrid& g = c->g;
  const double v[6] = {g.xmn, g.xmx, g.ymn, g.ymx, g.zmn, g.zmx};
  for (int i = 0; i < 6; i++) out[i] = v[i];
  return ALADYN_OK;
}

extern "C" int aladyn_get_species(aladyn_ctx* c, int32_t ic, double* part, int64_t ld,
                                  int64_t* np) {
  if (int r = check_species(c, ic)) return r;
  int64_t n = 0;
  if (int r = aladyn_species_count(c, ic, &n)) return r;
  if (np) *np = n;
  if (n > 0 && (!part || ld < n))
    return fail(c, ALADYN_EINVAL, "get_species: ld (%lld) < particle count (%lld)",
                (long long)ld, (long long)n);
  std::string msg;
  int r = tile_store_get(&c->ts[ic], c->tg, c->g, part, ld, c->stream, &msg);
  if (r) return fail(c, r, "%s", msg.c_str());
  return ALADYN_OK;
}

// ----------------------------------------------------------------- fields
template <typename T>
static int put_field_array(aladyn_ctx* c, T* dst, const double* src, int64_t n) {
  if (sizeof(T) == sizeof(double)) {
    HIPCK(c, hipMemcpyAsync(dst, src, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));  // the host array is only borrowed
  } else {
    std::vector<float> h((size_t)n);
    for (int64_t i = 0; i < n; i++) h[i] = (float)src[i];
    HIPCK(c, hipMemcpyAsync(dst, h.data(), (size_t)n * sizeof(float), hipMemcpyHostToDevice,
                            c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
  }
  return 0;
}

template <typename T>
static int get_field_array(aladyn_ctx* c, double* dst, const T* src, int64_t n) {
  if (sizeof(T) == sizeof(double)) {
    HIPCK(c, hipMemcpyAsync(dst, src, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> h((size_t)n);
    HIPCK(c, hipMemcpyAsync(h.data(), src, (size_t)n * sizeof(float), hipMemcpyDeviceToHost,
                            c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < n; i++) dst[i] = (double)h[i];
  }
  return 0;
}

extern "C" int aladyn_put_fields(aladyn_ctx* c, const double* ebf) {
  if (!c || !ebf) return fail(c, ALADYN_EINVAL, "put_fields: NULL argument");
  const int64_t n = c->g.ncell * c->g.nfield;
  for (int64_t i = 0; i < n; i++)
    if (!std::isfinite(ebf[i])) return fail(c, ALADYN_EINVAL, "put_fields: non-finite value");
  int r = f64(c) ? put_field_array(c, c->ebf64, ebf, n) : put_field_array(c, c->ebf32, ebf, n);
  if (r) return r;
  c->have_fields = true;
  return ALADYN_OK;
}

extern "C" int aladyn_get_fields(aladyn_ctx* c, double* ebf) {
  if (!c || !ebf) return fail(c, ALADYN_EINVAL, "get_fields: NULL argument");
  const int64_t n = c->g.ncell * c->g.nfield;
  return f64(c) ? get_field_array(c, ebf, c->ebf64, n) : get_field_array(c, ebf, c->ebf32, n);
}

extern "C" int aladyn_get_current(aladyn_ctx* c, double* jc) {
  if (!c || !jc) return fail(c, ALADYN_EINVAL, "get_current: NULL argument");
  const int64_t n = c->g.ncell * c->g.curr_ndim;
  return f64(c) ? get_field_array(c, jc, c->jc64, n) : get_field_array(c, jc, c->jc32, n);
}

// ----------------------------------------------------------------- stages
extern "C" int aladyn_pfields_prepare(aladyn_ctx* c) {
  if (!c) return ALADYN_EINVAL;
  if (f64(c))
    field_xyzbd<double>(c->g, c->ebf64, c->g.nfield, c->stream);
  else
    field_xyzbd<float>(c->g, c->ebf32, c->g.nfield, c->stream);
  HIPCK(c, hipGetLastError());
  return ALADYN_OK;
}

extern "C" int aladyn_curr_clean(aladyn_ctx* c) {
  if (!c) return ALADYN_EINVAL;
  const int64_t n = c->g.ncell * c->g.curr_ndim;
  if (f64(c))
    fill<double>(c->jc64, n, 0.0, c->stream);
  else
    fill<float>(c->jc32, n, 0.0f, c->stream);
  c->j_is_zero = true;
  HIPCK(c, hipGetLastError());
  return ALADYN_OK;
}

static int ensure_diag_buffers(aladyn_ctx* c) {
  const size_t nc = (size_t)c->g.ncell;
  if (!c->rho0) HIPCK(c, hipMalloc(&c->rho0, nc * sizeof(double)));
  if (!c->rho1) HIPCK(c, hipMalloc(&c->rho1, nc * sizeof(double)));
  if (!c->jraw) HIPCK(c, hipMalloc(&c->jraw, nc * c->g.curr_ndim * sizeof(double)));
  return 0;
}

// rho(ncell) += grid charge of species ic (device FP64)
static int grid_charge_dev(aladyn_ctx* c, int ic, double* rho) {
  tile_store_grid_charge(&c->ts[ic], c->tg, c->g, rho, c->stream);
  HIPCK(c, hipGetLastError());
  return 0;
}

// One species' stage; with normalise its current goes onto jc times djc
// (the fused step, whose collect then skips the normalisation).
static int particle_stage(aladyn_ctx* c, int ic, int initial_time, bool normalise) {
  if (!c->have_fields) return fail(c, ALADYN_ESTATE, "particle_stage before put_fields");
  const DevGrid& g = c->g;
  const double alp = 0.5 * g.dt * c->cfg.lorentz_fact[ic];
  if (c->diag) {
    if (int r = ensure_diag_buffers(c)) return r;
    fill<double>(c->rho0, g.ncell, 0.0, c->stream);
    if (int r = grid_charge_dev(c, ic, c->rho0)) return r;
    // the stage writes its own raw current, summed in FP64, into jraw
    fill<double>(c->jraw, g.ncell * g.curr_ndim, 0.0, c->stream);
    c->have_resid = false;
  }
  std::string msg;
  const int init = initial_time ? 1 : 0;
  double* jraw = c->diag ? c->jraw : nullptr;
  const int r =
      f64(c) ? tile_store_step<double>(&c->ts[ic], c->tg, g, c->ebf64, c->jc64, c->patch, alp,
                                       init, c->j_is_zero, jraw, c->stream, c->tpart.begin(),
                                       c->tpart.end(), &msg, normalise)
             : tile_store_step<float>(&c->ts[ic], c->tg, g, c->ebf32, c->jc32, c->patch,
                                      (float)alp, init, c->j_is_zero, jraw, c->stream,
                                      c->tpart.begin(), c->tpart.end(), &msg, normalise);
  if (r) return fail(c, r, "%s", msg.c_str());
  if (c->ts[ic].loaded) c->tpart.commit();
  c->j_is_zero = false;
  if (c->diag) {
    fill<double>(c->rho1, g.ncell, 0.0, c->stream);
    if (int r = grid_charge_dev(c, ic, c->rho1)) return r;
    c->have_resid = true;
  }
  return ALADYN_OK;
}

extern "C" int aladyn_particle_stage(aladyn_ctx* c, int32_t ic, int32_t initial_time) {
  if (int r = check_species(c, ic)) return r;
  return particle_stage(c, ic, initial_time, false);
}

static int curr_mpi_collect(aladyn_ctx* c, bool normalised) {
  if (f64(c))
    curr_collect<double>(c->g, c->jc64, c->stream, normalised);
  else
    curr_collect<float>(c->g, c->jc32, c->stream, normalised);
  HIPCK(c, hipGetLastError());
  return ALADYN_OK;
}

extern "C" int aladyn_curr_mpi_collect(aladyn_ctx* c) {
  if (!c) return ALADYN_EINVAL;
  return curr_mpi_collect(c, false);
}

extern "C" int aladyn_advance_lpf_fields(aladyn_ctx* c) {
  if (!c) return ALADYN_EINVAL;
  if (f64(c))
    advance_fields<double>(c->g, c->ebf64, c->jc64, c->stream);
  else
    advance_fields<float>(c->g, c->ebf32, c->jc32, c->stream);
  HIPCK(c, hipGetLastError());
  return ALADYN_OK;
}

// ------------------------------------------------------- captured steps
// What a captured step depends on besides the configuration: the box
// position (kernel arguments) and, per species, the buffer in use, the
// pinned slot of the overflow count and the layout version (segment
// arrays and buffers).
static std::vector<int64_t> step_key(const aladyn_ctx* c) {
  std::vector<int64_t> k{c->grid_version};
  for (int ic = 0; ic < c->cfg.nsp; ic++) {
    const TileStore& t = c->ts[ic];
    k.insert(k.end(), {(int64_t)t.loaded, (int64_t)t.cur, (int64_t)t.ovf_slot, t.layout_version});
  }
  return k;
}

// The device work of one step after the first (field ghosts, jc = 0, the
// species' particle kernels, fold and normalisation, Yee update), with
// nothing that waits: what a graph holds.
static int enqueue_step(aladyn_ctx* c) {
  const DevGrid& g = c->g;
  const int64_t nj = g.ncell * g.curr_ndim;
  std::string msg;
  // jc = 0: the first species' fold writes every node of jc (it does not
  // add), so the fill is needed only when that species holds no store
  const bool fold_sets_jc = c->ts[0].loaded;
  if (f64(c)) {
    field_xyzbd<double>(g, c->ebf64, g.nfield, c->stream);
    if (!fold_sets_jc) fill<double>(c->jc64, nj, 0.0, c->stream);
  } else {
    field_xyzbd<float>(g, c->ebf32, g.nfield, c->stream);
    if (!fold_sets_jc) fill<float>(c->jc32, nj, 0.0f, c->stream);
  }
  for (int ic = 0; ic < c->cfg.nsp; ic++) {
    const double alp = 0.5 * g.dt * c->cfg.lorentz_fact[ic];
    const int r =
        f64(c) ? tile_store_step_enqueue<double>(&c->ts[ic], c->tg, g, c->ebf64, c->jc64,
                                                 c->patch, alp, 0, ic == 0, nullptr, c->stream,
                                                 nullptr, nullptr, &msg, true)
               : tile_store_step_enqueue<float>(&c->ts[ic], c->tg, g, c->ebf32, c->jc32,
                                                c->patch, (float)alp, 0, ic == 0, nullptr,
                                                c->stream, nullptr, nullptr, &msg, true);
    if (r) return fail(c, r, "%s", msg.c_str());
  }
  if (f64(c)) {
    curr_collect<double>(g, c->jc64, c->stream, true);
    advance_fields<double>(g, c->ebf64, c->jc64, c->stream);
  } else {
    curr_collect<float>(g, c->jc32, c->stream, true);
    advance_fields<float>(g, c->ebf32, c->jc32, c->stream);
  }
  HIPCK(c, hipGetLastError());
  return ALADYN_OK;
}

// One step through a graph: captured on first use for the current key,
// replayed after; the host side of each species' step (buffer swap,
// overflow count, re-layout) follows the launch.
static int lpf2_evolve_graph(aladyn_ctx* c) {
  const std::vector<int64_t> key = step_key(c);
  hipGraphExec_t ex = nullptr;
  for (auto& gr : c->graphs)
    if (gr.key == key) ex = gr.exec;
  if (!ex) {
    if (c->graphs.size() >= 8) clear_graphs(c);
    HIPCK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int r = enqueue_step(c);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (r) {
      if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
      return r;
    }
    HIPCK(c, e);
    const hipError_t ei = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIPCK(c, ei);
    c->graphs.push_back({key, ex});
  }
  HIPCK(c, hipGraphLaunch(ex, c->stream));
  std::string msg;
  for (int ic = 0; ic < c->cfg.nsp; ic++)
    if (int r = tile_store_step_finish(&c->ts[ic], c->tg, c->g, c->stream, &msg))
      return fail(c, r, "%s", msg.c_str());
  c->j_is_zero = false;
  return ALADYN_OK;
}

// Lp_cycle's trim of the last step, dt_loc = tmax - tnow
// (ALaDyn.F90:101): every dt-dependent quantity (push constants, Yee and
// Mur coefficients, the window's lapse, the envelope solver's constants) is
// formed from g.dt at each call, so only the captured steps, whose kernel
// arguments hold it, are dropped.
extern "C" int aladyn_set_dt(aladyn_ctx* c, double dt) {
  if (!c) return ALADYN_EINVAL;
  const DevGrid& g = c->g;
  if (!(dt > 0.0)) return fail(c, ALADYN_EINVAL, "set_dt: dt must be positive");
  if (dt >= g.dx || dt >= g.dy || (g.ndim == 3 && dt >= g.dz))
    return fail(c, ALADYN_EINVAL, "set_dt: dt must be smaller than every cell size");
  c->cfg.dt = dt;
  c->g.dt = dt;
  clear_graphs(c);
  return ALADYN_OK;
}

extern "C" int aladyn_set_graphs(aladyn_ctx* c, int32_t enable) {
  if (!c) return ALADYN_EINVAL;
  c->use_graphs = enable != 0;
  if (!c->use_graphs) clear_graphs(c);
  return ALADYN_OK;
}

extern "C" int aladyn_lpf2_evolve(aladyn_ctx* c, int32_t iter, int32_t initial_time) {
  (void)iter;
  if (!c) return ALADYN_EINVAL;
  if (!c->have_fields) return fail(c, ALADYN_ESTATE, "lpf2_evolve before put_fields");
  if (c->use_graphs && !initial_time && !c->diag) return lpf2_evolve_graph(c);
  if (int r = aladyn_pfields_prepare(c)) return r;
  if (c->ts[0].loaded) {
    // jc = 0 without the fill: the first species' fold writes every node
    c->j_is_zero = true;
  } else if (int r = aladyn_curr_clean(c)) {
    return r;
  }
  // the folds apply djc, so the collect does not
  for (int ic = 0; ic < c->cfg.nsp; ic++)
    if (int r = particle_stage(c, ic, initial_time, true)) return r;
  if (int r = curr_mpi_collect(c, true)) return r;
  HIPCK(c, hipEventRecord(c->tfield.begin(), c->stream));
  if (int r = aladyn_advance_lpf_fields(c)) return r;
  HIPCK(c, hipEventRecord(c->tfield.end(), c->stream));
  c->tfield.commit();
  return ALADYN_OK;
}

// --------------------------------------------------------- envelope model
extern "C" int aladyn_put_envelope(aladyn_ctx* c, double oml, const double* env) {
  if (!c || !env) return fail(c, ALADYN_EINVAL, "put_envelope: NULL argument");
  if (!(oml > 0.0) || !std::isfinite(oml))
    return fail(c, ALADYN_EINVAL, "put_envelope: the laser frequency oml must be positive");
  const DevGrid& g = c->g;
  const int64_t n = g.ncell * 4;
  for (int64_t i = 0; i < n; i++)
    if (!std::isfinite(env[i])) return fail(c, ALADYN_EINVAL, "put_envelope: non-finite value");
  const size_t nje = (size_t)g.ncell * (g.curr_ndim + 1);
  if (f64(c)) {
    if (!c->env64) HIPCK(c, hipMalloc(&c->env64, (size_t)n * sizeof(double)));
    if (!c->je64) HIPCK(c, hipMalloc(&c->je64, nje * sizeof(double)));
    if (int r = put_field_array(c, c->env64, env, n)) return r;
  } else {
    if (!c->env32) HIPCK(c, hipMalloc(&c->env32, (size_t)n * sizeof(float)));
    if (!c->je32) HIPCK(c, hipMalloc(&c->je32, nje * sizeof(float)));
    if (int r = put_field_array(c, c->env32, env, n)) return r;
  }
  if (!c->den) HIPCK(c, hipMalloc(&c->den, (size_t)g.ncell * sizeof(double)));
  c->env_oml = oml;
  c->have_env = true;
  return ALADYN_OK;
}

extern "C" int aladyn_get_envelope(aladyn_ctx* c, double* env) {
  if (!c || !env) return fail(c, ALADYN_EINVAL, "get_envelope: NULL argument");
  if (!c->have_env) return fail(c, ALADYN_ESTATE, "get_envelope before put_envelope");
  const int64_t n = c->g.ncell * 4;
  return f64(c) ? get_field_array(c, env, c->env64, n) : get_field_array(c, env, c->env32, n);
}

template <typename R>
static int env_step(aladyn_ctx* c, R* ebf, R* jc, R* env, R* je) {
  const DevGrid& g = c->g;
  const EnvConst ec = env_const(g, c->cfg.der_ord, c->env_oml, c->cfg.cfl);
  std::string msg;
  // jc = 0; jc(1) = |A^n|^2/2, jc(2:D+1) its gradient (env_amp_prepare)
  HIPCK(c, env_amp_prepare<R>(g, ec, env, je, c->stream));
  // set_env_acc, lpf_env_momenta, set_env_density for species 1 (den is
  // written whole, zero without particles)
  const bool timed = c->ts[0].loaded;
  env_particles_push<R>(&c->ts[0], c->tg, g, ebf, je, c->den, c->patch,
                        c->cfg.lorentz_fact[0], c->stream, timed ? c->tenv_push.begin() : nullptr,
                        timed ? c->tenv_push.end() : nullptr);
  HIPCK(c, hipGetLastError());
  if (timed) c->tenv_push.commit();
  // env_den_collect, advance_lpf_envelope, env_fields_average
  HIPCK(c, env_grid_advance<R>(g, ec, env, je, c->den, c->stream));
  // set_env_grad_interp, lpf_env_positions, cell_part_dist, curr_accumulate
  if (int r = env_particles_move<R>(&c->ts[0], c->tg, g, je, jc, c->patch, c->stream, &msg,
                                    timed ? c->tenv_move.begin() : nullptr,
                                    timed ? c->tenv_move.end() : nullptr))
    return fail(c, r, "%s", msg.c_str());
  if (timed) c->tenv_move.commit();
  // curr_mpi_collect, advance_lpf_fields (env_evolve_in_time.f90:345-358)
  curr_collect<R>(g, jc, c->stream, true);
  advance_fields<R>(g, ebf, jc, c->stream);
  HIPCK(c, hipGetLastError());
  c->j_is_zero = false;
  return ALADYN_OK;
}

extern "C" int aladyn_env_lpf2_evolve(aladyn_ctx* c, int32_t iter) {
  if (!c) return ALADYN_EINVAL;
  (void)iter;
  if (!c->have_fields) return fail(c, ALADYN_ESTATE, "env_lpf2_evolve before put_fields");
  if (!c->have_env) return fail(c, ALADYN_ESTATE, "env_lpf2_evolve before put_envelope");
  return f64(c) ? env_step<double>(c, c->ebf64, c->jc64, c->env64, c->je64)
                : env_step<float>(c, c->ebf32, c->jc32, c->env32, c->je32);
}

// ------------------------------------------------------------ batched runs
// Many independent contexts stepped together: each kernel of the step runs
// once for all of them (the particle kernel over a work list of (run, tile)
// pairs, the grid kernels with one row of workgroups per run), the step's
// launches replayed as one captured HIP graph.  The kernels read each run's
// buffers from a descriptor array refreshed before every step, so the graph
// stays valid across buffer swaps and re-layouts.
struct aladyn_batch {
  std::vector<aladyn_ctx*> ctx;
  hipStream_t stream = nullptr;
  bool f64 = false, graphs = true;
  int ndim = 2, nruns = 0, nwork = 0, max_ntiles = 0;
  int64_t max_ncell = 0;
  BatchDims dims{};
  TileGeom tg0{};
  void* d_tile = nullptr;    // TileRun<R>[nruns]
  void* d_field = nullptr;   // FieldRun<R>[nruns]
  void* h_tile[2] = {nullptr, nullptr};   // pinned, one per step parity
  void* h_field[2] = {nullptr, nullptr};
  int2* d_work = nullptr;
  int* d_ovf = nullptr;
  int* h_ovf[2] = {nullptr, nullptr};     // pinned overflow counts per step parity
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool ev_valid[2] = {false, false};
  int64_t steps = 0;
  hipGraphExec_t graph = nullptr;
  std::string err;
};

static int bfail(aladyn_batch* b, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (b) b->err = buf;
  return code;
}

#define BHIPCK(b, call)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return bfail(b, ALADYN_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),  \
                   __FILE__, __LINE__);                                                    \
  } while (0)

extern "C" const char* aladyn_batch_last_error(const aladyn_batch* b) {
  return b ? b->err.c_str() : "null batch";
}

extern "C" int aladyn_batch_destroy(aladyn_batch* b) {
  if (!b) return ALADYN_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  (void)hipFree(b->d_tile);
  (void)hipFree(b->d_field);
  (void)hipFree(b->d_work);
  (void)hipFree(b->d_ovf);
  for (int k = 0; k < 2; k++) {
    if (b->h_tile[k]) (void)hipHostFree(b->h_tile[k]);
    if (b->h_field[k]) (void)hipHostFree(b->h_field[k]);
    if (b->h_ovf[k]) (void)hipHostFree(b->h_ovf[k]);
    if (b->ev[k]) (void)hipEventDestroy(b->ev[k]);
  }
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return ALADYN_OK;
}

extern "C" int aladyn_batch_create(aladyn_ctx* const* ctxs, int32_t n, int32_t use_graphs,
                                   aladyn_batch** out) {
  if (!out) return ALADYN_EINVAL;
  *out = nullptr;
  aladyn_batch* b = new aladyn_batch();
  *out = b;
  if (!ctxs || n < 1) return bfail(b, ALADYN_EINVAL, "batch_create: no contexts");
  b->graphs = use_graphs != 0;
  for (int r = 0; r < n; r++) {
    aladyn_ctx* c = ctxs[r];
    if (!c) return bfail(b, ALADYN_EINVAL, "batch_create: context %d is NULL", r);
    const aladyn_ctx* c0 = ctxs[0];
    if (c->cfg.nsp != 1)
      return bfail(b, ALADYN_EINVAL, "batch_create: context %d runs %d species (1 supported)", r,
                   c->cfg.nsp);
    if (!c->ts[0].loaded || !c->have_fields)
      return bfail(b, ALADYN_ESTATE, "batch_create: context %d has no species or fields", r);
    if (c->cfg.precision != c0->cfg.precision || c->g.ndim != c0->g.ndim ||
        c->tg.T[0] != c0->tg.T[0] || c->tg.T[1] != c0->tg.T[1] || c->tg.T[2] != c0->tg.T[2] ||
        c->tg.block != c0->tg.block)
      return bfail(b, ALADYN_EINVAL,
                   "batch_create: context %d differs from context 0 in precision, dimension, "
                   "tile shape or threads per tile", r);
    if (c->dev != c0->dev) return bfail(b, ALADYN_EINVAL, "batch_create: contexts on two devices");
    BHIPCK(b, hipStreamSynchronize(c->stream));
    b->ctx.push_back(c);
  }
  const aladyn_ctx* c0 = ctxs[0];
  b->f64 = c0->cfg.precision == ALADYN_PREC_F64;
  b->ndim = c0->g.ndim;
  b->nruns = n;
  b->tg0 = c0->tg;
  std::vector<DevGrid> gs;
  std::vector<int2> work;
  for (aladyn_ctx* c : b->ctx) {
    gs.push_back(c->g);
    b->max_ntiles = std::max(b->max_ntiles, c->tg.ntiles);
    b->max_ncell = std::max(b->max_ncell, c->g.ncell);
  }
  // the work list: the runs' tiles, run-major
  for (int r = 0; r < n; r++)
    for (int t = 0; t < b->ctx[r]->tg.ntiles; t++) work.push_back(make_int2(r, t));
  b->nwork = (int)work.size();
  b->dims = batch_dims(gs.data(), n);
  const size_t tsz = b->f64 ? sizeof(TileRun<double>) : sizeof(TileRun<float>);
  const size_t fsz = b->f64 ? sizeof(FieldRun<double>) : sizeof(FieldRun<float>);
  BHIPCK(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  BHIPCK(b, hipMalloc(&b->d_tile, tsz * n));
  BHIPCK(b, hipMalloc(&b->d_field, fsz * n));
  BHIPCK(b, hipMalloc(&b->d_work, sizeof(int2) * work.size()));
  BHIPCK(b, hipMalloc(&b->d_ovf, sizeof(int) * n));
  for (int k = 0; k < 2; k++) {
    BHIPCK(b, hipHostMalloc(&b->h_tile[k], tsz * n, hipHostMallocDefault));
    BHIPCK(b, hipHostMalloc(&b->h_field[k], fsz * n, hipHostMallocDefault));
    BHIPCK(b, hipHostMalloc(&b->h_ovf[k], sizeof(int) * n, hipHostMallocDefault));
    BHIPCK(b, hipEventCreateWithFlags(&b->ev[k], hipEventDisableTiming));
  }
  BHIPCK(b, hipMemcpyAsync(b->d_work, work.data(), sizeof(int2) * work.size(),
                           hipMemcpyHostToDevice, b->stream));
  BHIPCK(b, hipStreamSynchronize(b->stream));
  return ALADYN_OK;
}

// The device work of one batched step (what the graph holds).
template <typename R>
static hipError_t batch_enqueue(aladyn_batch* b) {
  const auto* tr = static_cast<const TileRun<R>*>(b->d_tile);
  const auto* fr = static_cast<const FieldRun<R>*>(b->d_field);
  field_xyzbd_batch<R>(fr, b->nruns, b->ndim, b->dims, b->stream);
  if (hipError_t e = tile_store_batch_launch<R>(b->tg0, b->ndim, tr, b->nruns, b->d_work,
                                                 b->nwork, b->max_ntiles, b->max_ncell, b->d_ovf,
                                                 b->stream))
    return e;
  curr_collect_batch<R>(fr, b->nruns, b->ndim, b->dims, b->stream);
  advance_fields_batch<R>(fr, b->nruns, b->dims, b->stream);
  return hipGetLastError();
}

template <typename R>
static int batch_step(aladyn_batch* b, int initial) {
  const int par = (int)(b->steps & 1), prev = 1 - par;
  // the pinned descriptors of this parity were last read by the copy of two
  // steps ago, which the event wait of the previous step has seen complete
  auto* ht = static_cast<TileRun<R>*>(b->h_tile[par]);
  auto* hf = static_cast<FieldRun<R>*>(b->h_field[par]);
  for (int r = 0; r < b->nruns; r++) {
    aladyn_ctx* c = b->ctx[r];
    R* ebf = f64(c) ? (R*)c->ebf64 : (R*)c->ebf32;
    R* jc = f64(c) ? (R*)c->jc64 : (R*)c->jc32;
    const double alp = 0.5 * c->g.dt * c->cfg.lorentz_fact[0];
    tile_store_batch_desc<R>(&c->ts[0], c->tg, c->g, ebf, jc, c->patch, (R)alp, initial,
                             &ht[r]);
    hf[r] = FieldRun<R>{c->g, ebf, jc};
  }
  BHIPCK(b, hipMemcpyAsync(b->d_tile, ht, sizeof(TileRun<R>) * b->nruns, hipMemcpyHostToDevice,
                           b->stream));
  BHIPCK(b, hipMemcpyAsync(b->d_field, hf, sizeof(FieldRun<R>) * b->nruns,
                           hipMemcpyHostToDevice, b->stream));
  if (b->graphs) {
    if (!b->graph) {
      BHIPCK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      const hipError_t e = batch_enqueue<R>(b);
      hipGraph_t graph = nullptr;
      const hipError_t ee = hipStreamEndCapture(b->stream, &graph);
      if (e != hipSuccess || ee != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        BHIPCK(b, e != hipSuccess ? e : ee);
      }
      const hipError_t ei = hipGraphInstantiate(&b->graph, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      BHIPCK(b, ei);
    }
    BHIPCK(b, hipGraphLaunch(b->graph, b->stream));
  } else {
    BHIPCK(b, batch_enqueue<R>(b));
  }
  BHIPCK(b, hipMemcpyAsync(b->h_ovf[par], b->d_ovf, sizeof(int) * b->nruns,
                           hipMemcpyDeviceToHost, b->stream));
  BHIPCK(b, hipEventRecord(b->ev[par], b->stream));
  b->ev_valid[par] = true;
  // the previous step's overflow counts decide this step's re-layouts
  const int* oprev = nullptr;
  if (b->ev_valid[prev]) {
    BHIPCK(b, hipEventSynchronize(b->ev[prev]));
    oprev = b->h_ovf[prev];
  }
  b->steps++;
  std::string msg;
  for (int r = 0; r < b->nruns; r++) {
    aladyn_ctx* c = b->ctx[r];
    if (int e = tile_store_batch_finish(&c->ts[0], c->tg, c->g, oprev ? oprev[r] : -1, b->stream,
                                        &msg))
      return bfail(b, e, "run %d: %s", r, msg.c_str());
    c->j_is_zero = false;
  }
  return ALADYN_OK;
}

extern "C" int aladyn_batch_lpf2_evolve(aladyn_batch* b, int32_t iter, int32_t initial_time) {
  if (!b || b->nruns < 1) return ALADYN_EINVAL;
  (void)iter;
  const int init = initial_time ? 1 : 0;
  return b->f64 ? batch_step<double>(b, init) : batch_step<float>(b, init);
}

extern "C" int aladyn_batch_synchronize(aladyn_batch* b) {
  if (!b) return ALADYN_EINVAL;
  BHIPCK(b, hipStreamSynchronize(b->stream));
  return ALADYN_OK;
}

// ------------------------------------------------------------ diagnostics
extern "C" int aladyn_energy_momenta(aladyn_ctx* c, int32_t ic, double ek[7], double* ekmax) {
  if (int r = check_species(c, ic)) return r;
  HIPCK(c, hipMemsetAsync(c->d_red, 0, 8 * sizeof(double), c->stream));
  tile_store_energy(&c->ts[ic], c->tg, c->g, c->d_red, c->stream);
  HIPCK(c, hipGetLastError());
  HIPCK(c, hipMemcpyAsync(c->h_red, c->d_red, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < 7; i++) ek[i] = c->h_red[i];
  if (ekmax) *ekmax = c->h_red[7];
  return ALADYN_OK;
}

extern "C" int aladyn_field_energy(aladyn_ctx* c, double out[6]) {
  if (!c || !out) return fail(c, ALADYN_EINVAL, "field_energy: NULL argument");
  HIPCK(c, hipMemsetAsync(c->d_red, 0, 8 * sizeof(double), c->stream));
  if (f64(c))
    field_energy<double>(c->g, c->ebf64, c->d_red, c->stream);
  else
    field_energy<float>(c->g, c->ebf32, c->d_red, c->stream);
  HIPCK(c, hipGetLastError());
  HIPCK(c, hipMemcpyAsync(c->h_red, c->d_red, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < 6; i++) out[i] = i < c->g.nfield ? c->g.dgvol * c->h_red[i] : 0.0;
  return ALADYN_OK;
}

extern "C" int aladyn_grid_charge(aladyn_ctx* c, int32_t ic, double* rho) {
  if (int r = check_species(c, ic)) return r;
  if (!rho) return fail(c, ALADYN_EINVAL, "grid_charge: NULL rho");
  if (int r = ensure_diag_buffers(c)) return r;
  const int64_t n = c->g.ncell;
  HIPCK(c, hipMemcpyAsync(c->rho1, rho, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (int r = grid_charge_dev(c, ic, c->rho1)) return r;
  HIPCK(c, hipMemcpyAsync(rho, c->rho1, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  c->have_resid = false;  // rho1 was reused
  return ALADYN_OK;
}

extern "C" int aladyn_beam_fields(aladyn_ctx* c, int32_t ic, double gam2, double bet0) {
  if (int r = check_species(c, ic)) return r;
  if (!c->have_fields) return fail(c, ALADYN_ESTATE, "beam_fields before put_fields");
  if (!(gam2 >= 1.0) || !std::isfinite(gam2) || !(bet0 >= 0.0 && bet0 < 1.0))
    return fail(c, ALADYN_EINVAL, "beam_fields: need gam2 >= 1 and 0 <= bet0 < 1");
  const DevGrid& g = c->g;
  double* rho = nullptr;
  HIPCK(c, hipMalloc(&rho, (size_t)g.ncell * sizeof(double)));
  fill<double>(rho, g.ncell, 0.0, c->stream);
  int r = grid_charge_dev(c, ic, rho);
  hipError_t e = hipSuccess;
  if (r == 0)
    e = f64(c) ? beam_fields<double>(g, rho, gam2, bet0, c->ebf64, c->stream)
               : beam_fields<float>(g, rho, gam2, bet0, c->ebf32, c->stream);
  (void)hipStreamSynchronize(c->stream);
  (void)hipFree(rho);
  if (r) return r;
  HIPCK(c, e);
  return ALADYN_OK;
}

// ------------------------------------------- allocation-free grid charge
// Device staging of aladyn_grid_charge_noalloc: one array per GPU, in the
// code object's data (no allocation call ever makes it).  Its host-side name
// is not a device address: host code uses only the address that
// hipGetSymbolAddress returns for the current device (noalloc_scratch),
// looked up once per device at aladyn_create.  Every call on one GPU shares
// the array: calls that overlap on one GPU (several host threads or
// streams) race, and it keeps the last call's contents.
static constexpr int64_t kNoAllocCells = int64_t(1) << 23;  // 64 MiB of FP64
static __device__ __align__(256) double scratch[kNoAllocCells];
static constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Looks up this device's copy of the array (the device must be current).
static int noalloc_scratch_init(aladyn_ctx* c) {
  if (c->dev < 0 || c->dev >= kMaxDevices)
    return fail(c, ALADYN_EINVAL, "device %d beyond the %d the library tracks", c->dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[c->dev]) return 0;
  void* p = nullptr;
  HIPCK(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[c->dev] = static_cast<double*>(p);
  return 0;
}

static double* noalloc_scratch(const aladyn_ctx* c) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_ptr[c->dev];
}

extern "C" int aladyn_grid_charge_noalloc(aladyn_ctx* c, int32_t ic, double* rho) {
  if (int r = check_species(c, ic)) return r;
  if (!rho) return fail(c, ALADYN_EINVAL, "grid_charge_noalloc: NULL rho");
  const int64_t n = c->g.ncell;
  if (n > kNoAllocCells)
    return fail(c, ALADYN_EINVAL,
                "grid_charge_noalloc: %lld grid nodes exceed its fixed device buffer of %lld "
                "(use aladyn_grid_charge)",
                (long long)n, (long long)kNoAllocCells);
  double* d = noalloc_scratch(c);
  if (!d) return fail(c, ALADYN_ESTATE, "grid_charge_noalloc: device buffer not set up");
  // the whole used range is overwritten by the caller's rho: nothing of an
  // earlier call survives into this one
  HIPCK(c, hipMemcpyAsync(d, rho, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (int r = grid_charge_dev(c, ic, d)) return r;
  HIPCK(c, hipMemcpyAsync(rho, d, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  return ALADYN_OK;
}

extern "C" int aladyn_set_diagnostics(aladyn_ctx* c, int32_t enable) {
  if (!c) return ALADYN_EINVAL;
  c->diag = enable != 0;
  c->have_resid = false;
  return ALADYN_OK;
}

extern "C" int aladyn_continuity_residual(aladyn_ctx* c, double out[2]) {
  if (!c || !out) return fail(c, ALADYN_EINVAL, "continuity_residual: NULL argument");
  if (!c->have_resid)
    return fail(c, ALADYN_ESTATE, "continuity_residual needs set_diagnostics(1) then a particle_stage");
  HIPCK(c, hipMemsetAsync(c->d_red, 0, 2 * sizeof(double), c->stream));
  continuity<double>(c->g, c->rho0, c->rho1, c->jraw, c->d_red, c->stream);
  HIPCK(c, hipGetLastError());
  HIPCK(c, hipMemcpyAsync(c->h_red, c->d_red, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  out[0] = c->h_red[0];
  out[1] = c->h_red[1];
  c->have_resid = false;
  return ALADYN_OK;
}

extern "C" int aladyn_synchronize(aladyn_ctx* c) {
  if (!c) return ALADYN_EINVAL;
  HIPCK(c, hipStreamSynchronize(c->stream));
  return ALADYN_OK;
}

extern "C" int aladyn_env_kernel_timing(aladyn_ctx* c, int32_t reset, double* push_ms,
                                        double* move_ms, int64_t* launches) {
  if (!c) return ALADYN_EINVAL;
  HIPCK(c, hipStreamSynchronize(c->stream));
  if (push_ms) *push_ms = c->tenv_push.average();
  if (move_ms) *move_ms = c->tenv_move.average();
  if (launches) *launches = c->tenv_move.n;
  if (reset) {
    c->tenv_push.n = 0;
    c->tenv_move.n = 0;
  }
  return ALADYN_OK;
}

extern "C" int aladyn_kernel_timing(aladyn_ctx* c, int32_t reset, double* particle_ms,
                                    double* field_ms, int64_t* launches) {
  if (!c) return ALADYN_EINVAL;
  HIPCK(c, hipStreamSynchronize(c->stream));
  if (particle_ms) *particle_ms = c->tpart.average();
  if (field_ms) *field_ms = c->tfield.average();
  if (launches) *launches = c->tpart.n;
  if (reset) {
    c->tpart.n = 0;
    c->tfield.n = 0;
  }
  return ALADYN_OK;
}
