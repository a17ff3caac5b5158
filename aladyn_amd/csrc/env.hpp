// env.hpp -- host launchers of the envelope model's step (envelope.hip).
#pragma once

#include <string>

#include "lp_kernels.hpp"
#include "tile_store.hpp"

namespace aladyn {

// Constants of the envelope solver, derived as set_field_param
// (grid_field_param.f90:36-99), potential_lapl / pp_lapl
// (grid_fields.f90:326-424, 1356-1403) and env_maxw_solve (:588-620) derive
// them for an unstretched grid.
struct EnvConst {
  int der_ord;
  double dx4[2], dy4[2], dz4[2];  // Laplacian coefficients (compact, wide)
  double dt2, kfact, k2_fact;     // dt^2, om0 dt, 1/(1 + kfact^2)
  double dx1_inv;                 // om0/dx (first-derivative term)
  double aph[2];                  // der_ord 3 first-derivative coefficients
};

EnvConst env_const(const DevGrid& g, int der_ord, double om0, double cfl);

// je = 0; je(1) = |A^n|^2/2 on the interior; je(2:D+1) = its gradient
// (env_amp_prepare, env_evolve_in_time.f90:145-170).  env holds A_R^n,
// A_I^n, A_R^{n-1}, A_I^{n-1}; je has curr_ndim + 1 components.
template <typename R>
hipError_t env_amp_prepare(const DevGrid& g, const EnvConst& ec, const R* env, R* je,
                           hipStream_t s);

// The particles' first half (set_env_acc, lpf_env_momenta, set_env_density)
// on species store ts: momenta to p^{n+1/2} in place, den = the wgh/gamp
// shapes (summed per tile in the patch array, then folded).  The events,
// when given, are recorded around the particle kernels and the fold.
template <typename R>
void env_particles_push(TileStore* ts, const TileGeom& tg, const DevGrid& g, const R* ebf,
                        const R* je, double* den, double* patch, double lorentz_fact,
                        hipStream_t s, hipEvent_t ev_begin = nullptr,
                        hipEvent_t ev_end = nullptr);

// den_zyxbd, je(1) = je(3) = den, advance_lpf_envelope (A^n -> A^{n+1}) and
// env_fields_average (je(1:D+1) = F at t^{n+1/2} and its gradient).
template <typename R>
hipError_t env_grid_advance(const DevGrid& g, const EnvConst& ec, R* env, R* je, double* den,
                            hipStream_t s);

// The particles' second half (set_env_grad_interp, lpf_env_positions,
// cell_part_dist, Esirkepov deposit) and the re-binning of the store; jc =
// djc x the folded current (the normalisation of curr_mpi_collect, which the
// collect that follows then skips).  Ends with the store's host-side step
// bookkeeping.
template <typename R>
int env_particles_move(TileStore* ts, const TileGeom& tg, const DevGrid& g, const R* je, R* jc,
                       double* patch, hipStream_t s, std::string* msg,
                       hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);

}  // namespace aladyn
