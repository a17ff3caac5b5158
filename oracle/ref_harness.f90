!=============================================================================
! ref_harness.f90 -- TEST INFRASTRUCTURE ONLY.
!
! A bind(C) harness around the reference's own hot-path routines, compiled
! together with the unmodified reference sources (see oracle/Makefile) into
! oracle/_ref/libaladyn_ref.so.  It is used only as a checker (to pin the C
! oracle) and as the CPU baseline; no product code links it.
!
! It sets the module globals that set_initial_param / set_grid /
! set_loc_grid_param would set for a single-rank, unstretched run
! (set_init_param.f90:76-102,151-162; set_grid_param.f90:52-225,811-829),
! copies the caller's arrays into the reference's data structures
! (spec(1)%part, ebfp, ebf, jc) and calls:
!   set_part2d_hcell_acc / set_part3d_hcell_acc (grid_part_connect.f90:89-285)
!   init_lpf_momenta, lpf_momenta_and_positions (boris_push.f90:34-168)
!   esirkepov_2d_curr / esirkepov_3d_curr      (grid_part_connect.f90:956-1334)
!   set_grid_charge                            (grid_part_util.f90:240-335)
!   set_field_param                            (grid_field_param.f90:36-99)
!   set_env_acc, set_env_grad_interp, set_env_density
!                                              (grid_part_connect.f90:414-952)
!   lpf_env_momenta, lpf_env_positions         (boris_push.f90:170-295)
! Routines in modules that need an MPI library (field_charge_multiply,
! reset_all_part_dist, the Maxwell solver) are not reachable here.
!=============================================================================
module ref_harness
  use iso_c_binding
  use precision_def
  use common_param
  use grid_param
  use mpi_var
  use pstruct_data
  use fstruct_data
  use grid_part_lib
  use grid_part_connect
  use boris_push
  use grid_part_util
  use grid_field_param
  implicit none

  integer :: h_n1p = 0, h_n2p = 0, h_n3p = 0

contains

  subroutine ref_setup(c_ndim, c_nx, c_ny, c_nz, c_dx, c_dy, c_dz, c_dt, &
                       c_lfact) bind(C, name='ref_setup')
    integer(c_int), value :: c_ndim, c_nx, c_ny, c_nz
    real(c_double), value :: c_dx, c_dy, c_dz, c_dt, c_lfact

    ndim = c_ndim
    nx = c_nx
    ny = c_ny
    nz = c_nz
    nfield = 3
    curr_ndim = 2
    if (ndim > 2) then
      nfield = 6
      curr_ndim = ndim
    end if
    nd2 = 2*curr_ndim
    iform = 0
    vbeam = 0.0
    stretch = .false.
    n_str = 0
    imody = 0
    imodz = 0
    dx = c_dx
    dx_inv = 1.0/dx
    dy = c_dy
    dy_inv = 1.0/dy
    dz = 1.
    dz_inv = 1.
    if (ndim > 2) then
      dz = c_dz
      dz_inv = 1./dz
    end if
    ! local origin = x(1), y(1), z(1) of set_grid on a single rank
    xmn = dx*real(1 - 1, dp)
    ymn = dy*real(1 - 1 - ny/2, dp)
    zmn = 0.0
    if (ndim > 2) zmn = dz*real(1 - 1 - nz/2, dp)
    h_n1p = nx + 5
    h_n2p = ny + 5
    h_n3p = 1
    if (ndim > 2) h_n3p = nz + 5
    dt_loc = c_dt
    lorentz_fact(1:4) = c_lfact
  end subroutine

  ! set_part*_hcell_acc: ebfp_out(1:np, 1:nfield) = gathered fields
  subroutine ref_gather(np, c_part, ld, c_ebf, c_ebfp) bind(C, name='ref_gather')
    integer(c_int64_t), value :: np, ld
    real(c_double), intent(in) :: c_part(ld, nd2 + 1)
    real(c_double), intent(in) :: c_ebf(h_n1p, h_n2p, h_n3p, nfield)
    real(c_double), intent(inout) :: c_ebfp(ld, nd2 + 1)
    integer :: n

    call load_part(int(np), c_part, ld)
    if (allocated(ebf)) deallocate (ebf)
    allocate (ebf(h_n1p, h_n2p, h_n3p, nfield))
    ebf = c_ebf
    if (ndim == 2) then
      call set_part2d_hcell_acc(ebf, spec(1), ebfp, int(np), nfield)
    else
      call set_part3d_hcell_acc(ebf, spec(1), ebfp, int(np))
    end if
    do n = 1, int(np)
      c_ebfp(n, 1:nfield) = ebfp(n, 1:nfield)
    end do
  end subroutine

  ! init_lpf_momenta (if initial /= 0) then lpf_momenta_and_positions.
  ! In: part, ebfp(1:nfield) = charge-multiplied fields.  Out: part, ebfp.
  subroutine ref_push(np, c_part, ld, c_ebfp, initial) bind(C, name='ref_push')
    integer(c_int64_t), value :: np, ld
    integer(c_int), value :: initial
    real(c_double), intent(inout) :: c_part(ld, nd2 + 1)
    real(c_double), intent(inout) :: c_ebfp(ld, nd2 + 1)
    integer :: n

    call load_part(int(np), c_part, ld)
    do n = 1, int(np)
      ebfp(n, 1:nd2 + 1) = c_ebfp(n, 1:nd2 + 1)
    end do
    if (initial /= 0) call init_lpf_momenta(spec(1), ebfp, int(np), 1)
    call lpf_momenta_and_positions(spec(1), ebfp, int(np), 1)
    do n = 1, int(np)
      c_part(n, 1:nd2 + 1) = spec(1)%part(n, 1:nd2 + 1)
      c_ebfp(n, 1:nd2 + 1) = ebfp(n, 1:nd2 + 1)
    end do
  end subroutine

  ! esirkepov_*_curr: c_jc += raw current.  In: part (new positions),
  ! ebfp(nd2/2+1 .. nd2) = old positions.
  subroutine ref_esirkepov(np, c_part, ld, c_ebfp, c_jc) bind(C, name='ref_esirkepov')
    integer(c_int64_t), value :: np, ld
    real(c_double), intent(in) :: c_part(ld, nd2 + 1)
    real(c_double), intent(inout) :: c_ebfp(ld, nd2 + 1)
    real(c_double), intent(inout) :: c_jc(h_n1p, h_n2p, h_n3p, curr_ndim)
    integer :: n

    call load_part(int(np), c_part, ld)
    do n = 1, int(np)
      ebfp(n, 1:nd2 + 1) = c_ebfp(n, 1:nd2 + 1)
    end do
    if (allocated(jc)) deallocate (jc)
    allocate (jc(h_n1p, h_n2p, h_n3p, curr_ndim))
    jc = c_jc
    if (ndim < 3) then
      call esirkepov_2d_curr(spec(1), ebfp, jc, int(np))
    else
      call esirkepov_3d_curr(spec(1), ebfp, jc, int(np))
    end if
    c_jc = jc
  end subroutine

  ! set_grid_charge into den(:,:,:,1)
  subroutine ref_grid_charge(np, c_part, ld, c_rho) bind(C, name='ref_grid_charge')
    integer(c_int64_t), value :: np, ld
    real(c_double), intent(in) :: c_part(ld, nd2 + 1)
    real(c_double), intent(inout) :: c_rho(h_n1p, h_n2p, h_n3p)
    real(dp), allocatable :: den(:, :, :, :)

    call load_part(int(np), c_part, ld)
    allocate (den(h_n1p, h_n2p, h_n3p, 1))
    den(:, :, :, 1) = c_rho
    call set_grid_charge(spec(1), ebfp, den, int(np), 1)
    c_rho = den(:, :, :, 1)
    deallocate (den)
  end subroutine

  ! The metric's loop body on the reference routines, for CPU timing:
  ! gather, charge multiply, push, deposit over nsteps with fixed fields.
  ! Particle removal is not applied (the caller keeps particles inside).
  subroutine ref_gather_push_deposit(np, c_part, ld, c_ebf, c_jc, nsteps) &
    bind(C, name='ref_gather_push_deposit')
    integer(c_int64_t), value :: np, ld
    integer(c_int), value :: nsteps
    real(c_double), intent(inout) :: c_part(ld, nd2 + 1)
    real(c_double), intent(in) :: c_ebf(h_n1p, h_n2p, h_n3p, nfield)
    real(c_double), intent(inout) :: c_jc(h_n1p, h_n2p, h_n3p, curr_ndim)
    integer :: n, it, ch

    call load_part(int(np), c_part, ld)
    if (allocated(ebf)) deallocate (ebf)
    allocate (ebf(h_n1p, h_n2p, h_n3p, nfield))
    ebf = c_ebf
    if (allocated(jc)) deallocate (jc)
    allocate (jc(h_n1p, h_n2p, h_n3p, curr_ndim))
    jc = 0.0
    ch = nd2 + 1
    do it = 1, nsteps
      if (ndim == 2) then
        call set_part2d_hcell_acc(ebf, spec(1), ebfp, int(np), nfield)
      else
        call set_part3d_hcell_acc(ebf, spec(1), ebfp, int(np))
      end if
      do n = 1, int(np)
        wgh_cmp = spec(1)%part(n, ch)
        ebfp(n, 1:nfield) = charge*ebfp(n, 1:nfield)
      end do
      call lpf_momenta_and_positions(spec(1), ebfp, int(np), 1)
      if (ndim < 3) then
        call esirkepov_2d_curr(spec(1), ebfp, jc, int(np))
      else
        call esirkepov_3d_curr(spec(1), ebfp, jc, int(np))
      end if
    end do
    do n = 1, int(np)
      c_part(n, 1:nd2 + 1) = spec(1)%part(n, 1:nd2 + 1)
    end do
    c_jc = jc
  end subroutine

  ! set_field_param: the finite-difference coefficients se_coeff(1:2) of the
  ! Yee x-derivative for der_ord 2, 3 or 4 at the given cfl and cell ratio
  subroutine ref_se_coeff(c_ndim, c_der_ord, c_cfl, c_yx_rat, c_out) bind(C, name='ref_se_coeff')
    integer(c_int), value :: c_ndim, c_der_ord
    real(c_double), value :: c_cfl, c_yx_rat
    real(c_double), intent(out) :: c_out(2)

    ndim = c_ndim
    der_ord = c_der_ord
    cfl = c_cfl
    yx_rat = c_yx_rat
    comoving = .false.
    nx = 8
    ny = 8
    nz = 8
    ny_loc = ny
    nz_loc = nz
    if (allocated(ww1)) deallocate (ww1)
    if (allocated(ww2)) deallocate (ww2)
    if (allocated(ww0)) deallocate (ww0)
    if (allocated(wr)) deallocate (wr)
    if (allocated(wl)) deallocate (wl)
    if (allocated(var)) deallocate (var)
    se_coeff = 0.0
    call set_field_param
    c_out = se_coeff
  end subroutine

  ! set_field_param's envelope-solver constants besides se_coeff:
  ! c_out = (hord_der2, opt_der1), read after the call for der_ord 3 or 4
  subroutine ref_field_param(c_ndim, c_der_ord, c_cfl, c_yx_rat, c_out) bind(C, name='ref_field_param')
    integer(c_int), value :: c_ndim, c_der_ord
    real(c_double), value :: c_cfl, c_yx_rat
    real(c_double), intent(out) :: c_out(2)
    real(c_double) :: se(2)

    hord_der2 = 0.0
    opt_der1 = 1.0
    call ref_se_coeff(c_ndim, c_der_ord, c_cfl, c_yx_rat, se)
    c_out(1) = hord_der2
    c_out(2) = opt_der1
  end subroutine

  ! set_env_acc(ebf, av, spec(1), ebfp, np, dt): ebfp_out(1:2D+1) as it exits
  subroutine ref_env_acc(np, c_part, ld, c_ebf, c_av, dt_step, c_ebfp) bind(C, name='ref_env_acc')
    integer(c_int64_t), value :: np, ld
    real(c_double), value :: dt_step
    real(c_double), intent(in) :: c_part(ld, nd2 + 1)
    real(c_double), intent(in) :: c_ebf(h_n1p, h_n2p, h_n3p, nfield)
    real(c_double), intent(in) :: c_av(h_n1p, h_n2p, h_n3p, curr_ndim + 1)
    real(c_double), intent(inout) :: c_ebfp(ld, nd2 + 1)
    real(dp), allocatable :: av(:, :, :, :)
    integer :: n

    call load_part(int(np), c_part, ld)
    do n = 1, int(np)
      ebfp(n, 1:nd2 + 1) = c_ebfp(n, 1:nd2 + 1)
    end do
    if (allocated(ebf)) deallocate (ebf)
    allocate (ebf(h_n1p, h_n2p, h_n3p, nfield))
    ebf = c_ebf
    allocate (av(h_n1p, h_n2p, h_n3p, curr_ndim + 1))
    av = c_av
    call set_env_acc(ebf, av, spec(1), ebfp, int(np), dt_step)
    do n = 1, int(np)
      c_ebfp(n, 1:nd2 + 1) = ebfp(n, 1:nd2 + 1)
    end do
    deallocate (av)
  end subroutine

  ! lpf_env_momenta(spec(1), ebfp, np, 1): part and ebfp in/out
  subroutine ref_env_momenta(np, c_part, ld, c_ebfp) bind(C, name='ref_env_momenta')
    integer(c_int64_t), value :: np, ld
    real(c_double), intent(inout) :: c_part(ld, nd2 + 1)
    real(c_double), intent(inout) :: c_ebfp(ld, nd2 + 1)
    integer :: n

    call load_part(int(np), c_part, ld)
    do n = 1, int(np)
      ebfp(n, 1:nd2 + 1) = c_ebfp(n, 1:nd2 + 1)
    end do
    call lpf_env_momenta(spec(1), ebfp, int(np), 1)
    do n = 1, int(np)
      c_part(n, 1:nd2 + 1) = spec(1)%part(n, 1:nd2 + 1)
      c_ebfp(n, 1:nd2 + 1) = ebfp(n, 1:nd2 + 1)
    end do
  end subroutine

  ! set_env_density(ebfp, den, np, 1): den(:,:,:) += wgh/gamp shapes
  subroutine ref_env_density(np, ld, c_ebfp, c_den) bind(C, name='ref_env_density')
    integer(c_int64_t), value :: np, ld
    real(c_double), intent(inout) :: c_ebfp(ld, nd2 + 1)
    real(c_double), intent(inout) :: c_den(h_n1p, h_n2p, h_n3p)
    real(dp), allocatable :: efp(:, :), den(:, :, :, :)
    integer :: n

    allocate (efp(max(int(np), 1), nd2 + 1))
    do n = 1, int(np)
      efp(n, 1:nd2 + 1) = c_ebfp(n, 1:nd2 + 1)
    end do
    allocate (den(h_n1p, h_n2p, h_n3p, 1))
    den(:, :, :, 1) = c_den
    call set_env_density(efp, den, int(np), 1)
    c_den = den(:, :, :, 1)
    do n = 1, int(np)
      c_ebfp(n, 1:nd2 + 1) = efp(n, 1:nd2 + 1)
    end do
    deallocate (efp, den)
  end subroutine

  ! set_env_grad_interp(av, spec(1), ebfp, np, curr_ndim)
  subroutine ref_env_grad_interp(np, c_part, ld, c_av, c_ebfp) bind(C, name='ref_env_grad_interp')
    integer(c_int64_t), value :: np, ld
    real(c_double), intent(in) :: c_part(ld, nd2 + 1)
    real(c_double), intent(in) :: c_av(h_n1p, h_n2p, h_n3p, curr_ndim + 1)
    real(c_double), intent(inout) :: c_ebfp(ld, nd2 + 1)
    real(dp), allocatable :: av(:, :, :, :)
    integer :: n

    call load_part(int(np), c_part, ld)
    do n = 1, int(np)
      ebfp(n, 1:nd2 + 1) = c_ebfp(n, 1:nd2 + 1)
    end do
    allocate (av(h_n1p, h_n2p, h_n3p, curr_ndim + 1))
    av = c_av
    call set_env_grad_interp(av, spec(1), ebfp, int(np), curr_ndim)
    do n = 1, int(np)
      c_ebfp(n, 1:nd2 + 1) = ebfp(n, 1:nd2 + 1)
    end do
    deallocate (av)
  end subroutine

  ! lpf_env_positions(spec(1), ebfp, np): part and ebfp in/out
  subroutine ref_env_positions(np, c_part, ld, c_ebfp) bind(C, name='ref_env_positions')
    integer(c_int64_t), value :: np, ld
    real(c_double), intent(inout) :: c_part(ld, nd2 + 1)
    real(c_double), intent(inout) :: c_ebfp(ld, nd2 + 1)
    integer :: n

    call load_part(int(np), c_part, ld)
    do n = 1, int(np)
      ebfp(n, 1:nd2 + 1) = c_ebfp(n, 1:nd2 + 1)
    end do
    call lpf_env_positions(spec(1), ebfp, int(np))
    do n = 1, int(np)
      c_part(n, 1:nd2 + 1) = spec(1)%part(n, 1:nd2 + 1)
      c_ebfp(n, 1:nd2 + 1) = ebfp(n, 1:nd2 + 1)
    end do
  end subroutine

  subroutine load_part(np, c_part, ld)
    integer, intent(in) :: np
    integer(c_int64_t), intent(in) :: ld
    real(c_double), intent(in) :: c_part(ld, nd2 + 1)
    integer :: n

    if (allocated(spec(1)%part)) then
      if (size(spec(1)%part, 1) /= max(np, 1) .or. &
          size(spec(1)%part, 2) /= nd2 + 1) deallocate (spec(1)%part)
    end if
    if (.not. allocated(spec(1)%part)) allocate (spec(1)%part(max(np, 1), nd2 + 1))
    if (allocated(ebfp)) then
      if (size(ebfp, 1) /= max(np, 1) .or. size(ebfp, 2) /= nd2 + 1) &
        deallocate (ebfp)
    end if
    if (.not. allocated(ebfp)) allocate (ebfp(max(np, 1), nd2 + 1))
    do n = 1, np
      spec(1)%part(n, 1:nd2 + 1) = c_part(n, 1:nd2 + 1)
    end do
  end subroutine

end module
