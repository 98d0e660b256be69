! streams_mi355x_run.F90 -- Fortran driver of the MI355X hot path.
!
! Mirrors the shape of the reference run() (singleideal_cpu.F90:1519-1693)
! for one restart-style case: the scalar parameters, grid and initial field
! come from a binary case file (written by the host tooling in the layout of
! the C ABI), the loop runs on the GPU through the bind(c) backend object,
! and the final conserved field and run scalars are written back.
!
! usage: streams_mi355x_run case.bin out.bin [dir [mode]]
!        streams_mi355x_run --sizes          byte sizes of the bind(c) records
!        streams_mi355x_run --read case.bin  what was read, no device work
! mode "procedure" runs each step through rk_sync_by_procedure (the stage
! procedure by procedure, euler_x in two ranges) instead of the fused loop.
! With dir, the end of the run is also written there as the reference's
! restart and statistics files: rst.bin (write_field, field.F90:1923-1966),
! field_info0.dat (write_field_info, singleideal.F90:2620-2659) and stat.bin
! (write_stats, singleideal.F90:1779-1822) of one compute_stats sample taken
! on the device.
program streams_mi355x_run
  use, intrinsic :: iso_c_binding
  use streams_mi355x
  use streams_equation_singleideal_mi355x_object
  implicit none
  type(st_params_t) :: p
  type(st_scalars_t) :: s
  type(equation_singleideal_mi355x_object) :: eq
  integer(c_int32_t) :: nsteps
  integer :: sx, sy, sz, u, l
  real(c_double), allocatable :: x(:), y(:), z(:), yn(:)
  real(c_double), allocatable :: dcsidx(:), dcsidxs(:), dcsidx2(:), detady(:), detadys(:), detady2(:)
  real(c_double), allocatable :: dzitdz(:), dzitdzs(:), dzitdz2(:), w(:,:,:,:), w_stat(:,:,:)
  character(len=512) :: fin, fout, dir, mode
  real(c_double) :: dt
  integer :: n

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
  allocate(x(sx), y(sy), z(sz), yn(p%ny+1))
  allocate(dcsidx(p%nx), dcsidxs(p%nx), dcsidx2(p%nx), detady(p%ny), detadys(p%ny), detady2(p%ny))
  allocate(dzitdz(p%nz), dzitdzs(p%nz), dzitdz2(p%nz))
  ! the ghosted conserved field with the reference's bounds, w(1-ng:nx+ng, ..., 1:5)
  allocate(w(1-p%ng:p%nx+p%ng, 1-p%ng:p%ny+p%ng, 1-p%ng:p%nz+p%ng, 5))
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
  mode = 'fused'
  if (command_argument_count() >= 4) call get_command_argument(4, mode)
  if (trim(mode) == 'procedure') then
    do n = 1, nsteps
      call eq%rk_sync_by_procedure()
    enddo
  else
    call eq%run(int(nsteps), 0, 6)
  endif
  call eq%compute_residual(s)
  call eq%copy_to_field(w)
  write(*,'(a,i0,a,es20.10)') 'steps: ', s%icyc, '  residual_rhou: ', s%residual_rhou

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) w
  write(u) s
  close(u)

  if (command_argument_count() >= 3) then
    call get_command_argument(3, dir)
    ! rst.bin: the interior of each conserved variable, x fastest
    open(newunit=u, file=trim(dir)//'/rst.bin', access='stream', form='unformatted', status='replace')
    write(u) w(1:p%nx, 1:p%ny, 1:p%nz, 1:5)
    close(u)
    ! field_info0.dat: the [field_info] section of the run's end
    open(newunit=u, file=trim(dir)//'/field_info0.dat', status='replace')
    write(u, '(a)') '[field_info]'
    write(u, '(a,i0)') 'icyc0 = ', s%icyc
    write(u, '(a,es25.17)') 'time0 = ', s%time
    write(u, '(a)') 'itav = 1'
    close(u)
    ! stat.bin: one statistics sample of the final state, one block per variable
    call eq2_stats(w_stat)
    open(newunit=u, file=trim(dir)//'/stat.bin', access='stream', form='unformatted', status='replace')
    do l = 1, 70
      write(u) w_stat(l, 1:p%nx, 1:p%ny)
    enddo
    close(u)
  endif
  call eq%finalize()

contains

  subroutine eq2_stats(ws)
    real(c_double), allocatable, intent(out) :: ws(:,:,:)
    allocate(ws(70, p%nx, p%ny))
    call eq%compute_stats(0)
    call eq%get_stats(ws)
  end subroutine eq2_stats
end program streams_mi355x_run
