!-----------------------------------------------------------------------
! drive_lp.f90 -- Fortran caller of libaladyn_gpu.so through the bind(C)
! module aladyn_gpu, used by tests/test_fortran_binding.py.
!
! It stands where the reference's lpf2_evolve loop stands: it owns
! spec-style part(ld, 2*curr_ndim+1) and ebf(n1p,n2p,n3p,nfield) arrays,
! hands them to the library, runs nsteps of lpf2_evolve and reads the
! particles, fields and current back.
!
! Usage: drive_lp <input file> <output file>
! Input (stream, native endian): int32 ndim, nx, ny, nz, precision, nsteps,
!   track_ids; int64 np; float64 dx, dy, dz, dt, ompe, lorentz_fact;
!   part(np, 2*curr_ndim+1); ebf(n1p, n2p, n3p, nfield).
! Output: int32 status; int64 np_out; part(np_out, ...); ebf; jc.
!-----------------------------------------------------------------------
program drive_lp
  use, intrinsic :: iso_c_binding
  use aladyn_gpu
  implicit none
  type(aladyn_config) :: cfg
  type(c_ptr) :: ctx
  integer(c_int32_t) :: hdr(7), it
  integer(c_int64_t) :: np, np_out
  real(c_double) :: rh(6)
  real(c_double), allocatable :: part(:, :), ebf(:, :, :, :), jc(:, :, :, :)
  integer :: ncol, cnd, nf, n1p, n2p, n3p, u, rc
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u) hdr, np, rh
  cfg%ndim = hdr(1)
  cfg%nx = hdr(2)
  cfg%ny = hdr(3)
  cfg%nz = hdr(4)
  cfg%precision = hdr(5)
  cfg%reserved(1) = hdr(7)
  cfg%nsp = 1
  cfg%dx = rh(1)
  cfg%dy = rh(2)
  cfg%dz = rh(3)
  cfg%dt = rh(4)
  cfg%ompe = rh(5)
  cfg%lorentz_fact(1) = rh(6)
  cnd = merge(3, 2, cfg%ndim == 3)
  nf = merge(6, 3, cfg%ndim == 3)
  ncol = 2*cnd + 1
  n1p = cfg%nx + 5
  n2p = cfg%ny + 5
  n3p = merge(cfg%nz + 5, 1, cfg%ndim == 3)
  allocate (part(np, ncol), ebf(n1p, n2p, n3p, nf), jc(n1p, n2p, n3p, cnd))
  read (u) part, ebf
  close (u)

  rc = aladyn_create(cfg, ctx)
  if (rc == ALADYN_OK) rc = aladyn_put_species(ctx, 0_c_int32_t, part, np, np)
  if (rc == ALADYN_OK) rc = aladyn_put_fields(ctx, ebf)
  do it = 0, hdr(6) - 1
    if (rc /= ALADYN_OK) exit
    rc = aladyn_lpf2_evolve(ctx, it, merge(1_c_int32_t, 0_c_int32_t, it == 0))
  end do
  np_out = 0
  if (rc == ALADYN_OK) rc = aladyn_get_species(ctx, 0_c_int32_t, part, np, np_out)
  if (rc == ALADYN_OK) rc = aladyn_get_fields(ctx, ebf)
  if (rc == ALADYN_OK) rc = aladyn_get_current(ctx, jc)
  if (rc /= ALADYN_OK) then
    write (*, '(a,i0,2a)') 'drive_lp: error ', rc, ': ', aladyn_error_message(ctx)
  end if

  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) int(rc, c_int32_t), np_out
  if (rc == ALADYN_OK) write (u) part(1:np_out, :), ebf, jc
  close (u)
  if (c_associated(ctx)) rc = aladyn_destroy(ctx)
end program drive_lp
