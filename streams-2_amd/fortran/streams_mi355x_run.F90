! streams_mi355x_run.F90 -- Fortran driver of the MI355X hot path.
!
! Mirrors the shape of the reference run() (singleideal_cpu.F90:1519-1693)
! for one restart-style case: the scalar parameters, grid and initial field
! come from a binary case file (written by the host tooling in the layout of
! the C ABI), the loop runs on the GPU through the bind(c) backend object,
! and the final conserved field and run scalars are written back.
!
! usage: streams_mi355x_run case.bin out.bin
!        streams_mi355x_run --sizes          byte sizes of the bind(c) records
!        streams_mi355x_run --read case.bin  what was read, no device work
program streams_mi355x_run
  use, intrinsic :: iso_c_binding
  use streams_mi355x
  implicit none
  type(st_params_t) :: p
  type(st_scalars_t) :: s
  type(equation_singleideal_mi355x_object) :: eq
  integer(c_int32_t) :: nsteps
  integer :: sx, sy, sz, nvol, u
  real(c_double), allocatable :: x(:), y(:), z(:), yn(:)
  real(c_double), allocatable :: dcsidx(:), dcsidxs(:), dcsidx2(:), detady(:), detadys(:), detady2(:)
  real(c_double), allocatable :: dzitdz(:), dzitdzs(:), dzitdz2(:), w(:)
  character(len=512) :: fin, fout
  real(c_double) :: dt

  call get_command_argument(1, fin)
  if (trim(fin) == '--sizes') then
    ! byte sizes of the bind(c) records, checked against the C header
    write(*,'(i0,1x,i0)') storage_size(p)/8, storage_size(s)/8
    stop
  endif
  call get_command_argument(2, fout)
  if (trim(fin) == '--read') fin = fout
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) p
  read(u) nsteps
  sx = p%nx + 2*p%ng
  sy = p%ny + 2*p%ng
  sz = p%nz + 2*p%ng
  nvol = sx*sy*sz
  allocate(x(sx), y(sy), z(sz), yn(p%ny+1))
  allocate(dcsidx(p%nx), dcsidxs(p%nx), dcsidx2(p%nx), detady(p%ny), detadys(p%ny), detady2(p%ny))
  allocate(dzitdz(p%nz), dzitdzs(p%nz), dzitdz2(p%nz), w(5*nvol))
  read(u) x, y, z, yn, dcsidx, dcsidxs, dcsidx2, detady, detadys, detady2, dzitdz, dzitdzs, dzitdz2
  read(u) w
  close(u)
  call get_command_argument(1, fout)
  if (trim(fout) == '--read') then
    write(*,'(5(i0,1x),3es24.16)') p%nx, p%ny, p%nz, p%ng, nsteps, p%cfl, sum(yn), sum(w)
    stop
  endif
  call get_command_argument(2, fout)

  call eq%initialize(p, x, y, z, yn, dcsidx, dcsidxs, dcsidx2, detady, detadys, detady2, dzitdz, dzitdzs, dzitdz2)
  call eq%copy_from_field(w)
  call eq%init_run()
  call eq%compute_dt(dt)
  write(*,*) 'dt =', dt
  call eq%advance(int(nsteps))
  call eq%compute_residual(s)
  call eq%copy_to_field(w)
  call eq%finalize()
  write(*,'(a,i0,a,es20.10)') 'steps: ', s%icyc, '  residual_rhou: ', s%residual_rhou

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) w
  write(u) s
  close(u)
end program streams_mi355x_run
