!-----------------------------------------------------------------------
! aladyn_gpu.f90 -- Fortran bind(C) interfaces to libaladyn_gpu.so, the
! MI355X implementation of the LP PIC step (include/aladyn_gpu.h).
!
! The reference calls its hot path as plain module procedures inside
! lpf2_evolve (src/dynamics/pic_evolve_in_time.f90:43-107).  A caller that
! switches to the GPU `use`s this module and calls the entry points below
! instead; arrays are passed in the reference's own layouts:
!   part(ld, 2*curr_ndim+1)    spec(ic)%part          (struct_def.f90:26-28)
!   ebf(n1p, n2p, n3p, nfield)                         (fstruct_data.f90:27)
!   jc(n1p, n2p, n3p, curr_ndim)                       (fstruct_data.f90:28)
! Every function returns 0 on success; a nonzero value is an ALADYN_E*
! code, and aladyn_error_message(ctx) returns the library's explanation.
!-----------------------------------------------------------------------
module aladyn_gpu
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: ALADYN_OK = 0, ALADYN_EINVAL = 1, &
    ALADYN_EHIP = 2, ALADYN_ENOMEM = 3, ALADYN_ESTATE = 4, ALADYN_ERANGE = 5
  integer(c_int32_t), parameter, public :: ALADYN_PREC_F64 = 64, ALADYN_PREC_F32 = 32
  integer, parameter, public :: ALADYN_MAX_SPECIES = 4

  ! run constants of one single-rank sub-domain (struct aladyn_config)
  type, bind(C), public :: aladyn_config
    integer(c_int32_t) :: ndim = 2
    integer(c_int32_t) :: nx = 0, ny = 0, nz = 1
    integer(c_int32_t) :: ibx = 0, iby = 0, ibz = 0
    integer(c_int32_t) :: der_ord = 2
    integer(c_int32_t) :: precision = ALADYN_PREC_F64
    integer(c_int32_t) :: nsp = 1
    real(c_double) :: dx = 0, dy = 0, dz = 1
    real(c_double) :: dt = 0
    real(c_double) :: cfl = 0
    real(c_double) :: ompe = 0
    real(c_double) :: lorentz_fact(4) = 0
    integer(c_int32_t) :: tile(3) = 0
    integer(c_int32_t) :: block = 0
    integer(c_int32_t) :: sort_every = 0
    integer(c_int32_t) :: reserved(5) = 0
  end type aladyn_config

  public :: aladyn_create, aladyn_destroy, aladyn_last_error, aladyn_error_message
  public :: aladyn_put_species, aladyn_species_count, aladyn_get_species
  public :: aladyn_put_fields, aladyn_get_fields, aladyn_get_current
  public :: aladyn_lpf2_evolve, aladyn_set_graphs, aladyn_pfields_prepare, aladyn_curr_clean
  public :: aladyn_particle_stage, aladyn_curr_mpi_collect, aladyn_advance_lpf_fields
  public :: aladyn_lp_window_xshift, aladyn_inject_species, aladyn_box
  public :: aladyn_set_injection, aladyn_injection_state, aladyn_set_box_origin
  public :: aladyn_energy_momenta, aladyn_field_energy, aladyn_grid_charge
  public :: aladyn_grid_charge_noalloc
  public :: aladyn_continuity_residual, aladyn_set_diagnostics
  public :: aladyn_synchronize, aladyn_kernel_timing

  interface
    integer(c_int) function aladyn_create(cfg, ctx) bind(C, name="aladyn_create")
      import :: c_int, c_ptr, aladyn_config
      type(aladyn_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function aladyn_destroy(ctx) bind(C, name="aladyn_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function aladyn_last_error(ctx) bind(C, name="aladyn_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function aladyn_put_species(ctx, ic, part, ld, np) &
        bind(C, name="aladyn_put_species")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ic
      real(c_double), intent(in) :: part(*)
      integer(c_int64_t), value :: ld, np
    end function

    integer(c_int) function aladyn_species_count(ctx, ic, np) &
        bind(C, name="aladyn_species_count")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ic
      integer(c_int64_t), intent(out) :: np
    end function

    integer(c_int) function aladyn_get_species(ctx, ic, part, ld, np) &
        bind(C, name="aladyn_get_species")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ic
      real(c_double), intent(inout) :: part(*)
      integer(c_int64_t), value :: ld
      integer(c_int64_t), intent(out) :: np
    end function

    integer(c_int) function aladyn_put_fields(ctx, ebf) bind(C, name="aladyn_put_fields")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ebf(*)
    end function

    integer(c_int) function aladyn_get_fields(ctx, ebf) bind(C, name="aladyn_get_fields")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: ebf(*)
    end function

    integer(c_int) function aladyn_get_current(ctx, jc) bind(C, name="aladyn_get_current")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: jc(*)
    end function

    integer(c_int) function aladyn_lpf2_evolve(ctx, iter, initial_time) &
        bind(C, name="aladyn_lpf2_evolve")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: iter, initial_time
    end function

    ! Lp_cycle's dt_loc for the next steps (ALaDyn.F90:89-107)
    integer(c_int) function aladyn_set_dt(ctx, dt) bind(C, name="aladyn_set_dt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function

    ! many contexts stepped as one batch (one launch per kernel for all runs)
    integer(c_int) function aladyn_batch_create(ctxs, n, use_graphs, batch) &
        bind(C, name="aladyn_batch_create")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), intent(in) :: ctxs(*)
      integer(c_int32_t), value :: n, use_graphs
      type(c_ptr), intent(out) :: batch
    end function

    integer(c_int) function aladyn_batch_lpf2_evolve(batch, iter, initial_time) &
        bind(C, name="aladyn_batch_lpf2_evolve")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: iter, initial_time
    end function

    integer(c_int) function aladyn_batch_synchronize(batch) &
        bind(C, name="aladyn_batch_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function

    integer(c_int) function aladyn_batch_destroy(batch) bind(C, name="aladyn_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function

    integer(c_int) function aladyn_set_graphs(ctx, enable) bind(C, name="aladyn_set_graphs")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: enable
    end function

    integer(c_int) function aladyn_pfields_prepare(ctx) bind(C, name="aladyn_pfields_prepare")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function aladyn_curr_clean(ctx) bind(C, name="aladyn_curr_clean")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function aladyn_particle_stage(ctx, ic, initial_time) &
        bind(C, name="aladyn_particle_stage")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ic, initial_time
    end function

    integer(c_int) function aladyn_curr_mpi_collect(ctx) bind(C, name="aladyn_curr_mpi_collect")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function aladyn_advance_lpf_fields(ctx) &
        bind(C, name="aladyn_advance_lpf_fields")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function aladyn_lp_window_xshift(ctx, witr, init_iter, w_speed, nshx) &
        bind(C, name="aladyn_lp_window_xshift")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: witr, init_iter
      real(c_double), value :: w_speed
      integer(c_int32_t), intent(out) :: nshx
    end function

    integer(c_int) function aladyn_inject_species(ctx, ic, part, ld, np) &
        bind(C, name="aladyn_inject_species")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ic
      real(c_double), intent(in) :: part(*)
      integer(c_int64_t), value :: ld, np
    end function

    integer(c_int) function aladyn_set_injection(ctx, ic, xpt, nxp, ypt, nyp, zpt, nzp, &
        wghpt, wghyz, nptx, targ_in, targ_end) bind(C, name="aladyn_set_injection")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ic, nyp, nzp
      integer(c_int64_t), value :: nxp, nptx
      real(c_double), intent(in) :: xpt(*), ypt(*), zpt(*), wghpt(*), wghyz(*)
      real(c_double), value :: targ_in, targ_end
    end function

    integer(c_int) function aladyn_injection_state(ctx, ic, nptx) &
        bind(C, name="aladyn_injection_state")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ic
      integer(c_int64_t), intent(out) :: nptx
    end function

    integer(c_int) function aladyn_set_box_origin(ctx, xmin) &
        bind(C, name="aladyn_set_box_origin")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: xmin
    end function

    integer(c_int) function aladyn_box(ctx, out) bind(C, name="aladyn_box")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(6)
    end function

    integer(c_int) function aladyn_energy_momenta(ctx, ic, ek, ekmax) &
        bind(C, name="aladyn_energy_momenta")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ic
      real(c_double), intent(out) :: ek(7), ekmax
    end function

    integer(c_int) function aladyn_field_energy(ctx, out) bind(C, name="aladyn_field_energy")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(6)
    end function

    integer(c_int) function aladyn_grid_charge(ctx, ic, rho) bind(C, name="aladyn_grid_charge")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ic
      real(c_double), intent(inout) :: rho(*)
    end function

    ! no counterpart in the reference: aladyn_grid_charge for callers that
    ! cannot allocate device memory (a fixed 2**23-node device array per GPU)
    integer(c_int) function aladyn_grid_charge_noalloc(ctx, ic, rho) &
        bind(C, name="aladyn_grid_charge_noalloc")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ic
      real(c_double), intent(inout) :: rho(*)
    end function

    ! replaces the bunch-field part of init_beam_part_distrib
    ! (set_grid_charge, fft_3d_psolv, initial_beam_fields, ebf += ebf_bunch)
    integer(c_int) function aladyn_beam_fields(ctx, ic, gam2, bet0) &
        bind(C, name="aladyn_beam_fields")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ic
      real(c_double), value :: gam2, bet0
    end function

    integer(c_int) function aladyn_continuity_residual(ctx, out) &
        bind(C, name="aladyn_continuity_residual")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(2)
    end function

    integer(c_int) function aladyn_set_diagnostics(ctx, enable) &
        bind(C, name="aladyn_set_diagnostics")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: enable
    end function

    integer(c_int) function aladyn_synchronize(ctx) bind(C, name="aladyn_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    ! the envelope model: env(n1p, n2p, n3p, 4), oml the laser frequency
    integer(c_int) function aladyn_put_envelope(ctx, oml, env) &
        bind(C, name="aladyn_put_envelope")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: oml
      real(c_double), intent(in) :: env(*)
    end function

    integer(c_int) function aladyn_get_envelope(ctx, env) bind(C, name="aladyn_get_envelope")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: env(*)
    end function

    ! replaces env_lpf2_evolve(iter) (env_evolve_in_time.f90:196-361)
    integer(c_int) function aladyn_env_lpf2_evolve(ctx, iter) &
        bind(C, name="aladyn_env_lpf2_evolve")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: iter
    end function

    integer(c_int) function aladyn_env_kernel_timing(ctx, reset, push_ms, move_ms, launches) &
        bind(C, name="aladyn_env_kernel_timing")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: reset
      real(c_double), intent(out) :: push_ms, move_ms
      integer(c_int64_t), intent(out) :: launches
    end function

    integer(c_int) function aladyn_kernel_timing(ctx, reset, particle_ms, field_ms, launches) &
        bind(C, name="aladyn_kernel_timing")
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: reset
      real(c_double), intent(out) :: particle_ms, field_ms
      integer(c_int64_t), intent(out) :: launches
    end function
  end interface

  interface
    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! The library's last error text as a Fortran string.
  function aladyn_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = aladyn_last_error(ctx)
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate (character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function aladyn_error_message

end module aladyn_gpu
