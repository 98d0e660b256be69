! singleideal_mi355x.F90 -- the MI355X backend object of the single-ideal-gas
! equation, a drop-in for equation_singleideal_cpu_object
! (code/src/singleideal/singleideal_cpu.F90:29-138).
!
! The device state of one solve lives in libstreams_mi355x behind a handle;
! every type-bound procedure below carries the name and the arguments of the
! reference procedure it replaces and calls the matching entry point of the
! C ABI (streams_mi355x.F90).  rk_sync runs the fused stage kernels;
! rk_sync_by_procedure runs the reference's stage procedure by procedure,
! with euler_x split into two ranges as rk_async splits it
! (singleideal_cpu.F90:427-526), for backends that interleave their own
! work.  run is the integration loop of run() (singleideal_cpu.F90:
! 1519-1693) without the output manager, which stays on the host side.
module streams_equation_singleideal_mi355x_object
  use, intrinsic :: iso_c_binding
  use streams_mi355x
  implicit none
  private
  public :: equation_singleideal_mi355x_object

  type :: equation_singleideal_mi355x_object
    type(c_ptr) :: handle = c_null_ptr
    type(st_params_t) :: p
    integer :: icyc0 = 0
    integer :: ndim = 3
    ! eul_imin, eul_imax, eul_jmin, eul_jmax, eul_kmin, eul_kmax (bc_preproc)
    integer(c_int32_t) :: eul(6)
  contains
    procedure, pass(self) :: initialize
    procedure, pass(self) :: finalize
    procedure, pass(self) :: copy_from_field
    procedure, pass(self) :: copy_to_field
    procedure, pass(self) :: compute_dt
    procedure, pass(self) :: compute_residual
    procedure, pass(self) :: print_progress
    procedure, pass(self) :: run
    procedure, pass(self) :: rk_sync
    procedure, pass(self) :: rk_sync_by_procedure
    procedure, pass(self) :: update_ghost
    procedure, pass(self) :: force_rhs
    procedure, pass(self) :: force_var
    procedure, pass(self) :: euler_x
    procedure, pass(self) :: euler_y
    procedure, pass(self) :: euler_z
    procedure, pass(self) :: visflx
    procedure, pass(self) :: compute_aux
    procedure, pass(self) :: recyc_exchange
    procedure, pass(self) :: bc_nr
    procedure, pass(self) :: bcextr_var
    procedure, pass(self) :: bcswap_var
    procedure, pass(self) :: count_weno
    procedure, pass(self) :: set_recycling
    procedure, pass(self) :: set_wmean
    procedure, pass(self) :: compute_stats
    procedure, pass(self) :: get_stats
    procedure, pass(self) :: set_stats
    procedure, pass(self) :: init_run
    procedure, pass(self) :: advance
    procedure, pass(self) :: scalars
  end type equation_singleideal_mi355x_object

contains

  ! alloc + the setup of bc_preproc / correct_bc_order on the device
  ! (singleideal.F90:4703-4750, 946-993) and the grid of set_local_grid
  subroutine initialize(self, p, x, y, z, yn, dcsidx, dcsidxs, dcsidx2, detady, detadys, detady2, &
      dzitdz, dzitdzs, dzitdz2)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    type(st_params_t), intent(in) :: p
    real(c_double), intent(in) :: x(*), y(*), z(*), yn(*)
    real(c_double), intent(in) :: dcsidx(*), dcsidxs(*), dcsidx2(*)
    real(c_double), intent(in) :: detady(*), detadys(*), detady2(*)
    real(c_double), intent(in) :: dzitdz(*), dzitdzs(*), dzitdz2(*)
    self%p = p
    self%ndim = 3
    if (p%nz == 1) self%ndim = 2
    call check(st_solver_create(p, self%handle), 'st_solver_create')
    call check(st_set_grid(self%handle, x, y, z, yn, dcsidx, dcsidxs, dcsidx2, detady, detadys, detady2, &
      dzitdz, dzitdzs, dzitdz2), 'st_set_grid')
    call check(st_get_eul_range(self%handle, self%eul), 'st_get_eul_range')
  end subroutine initialize

  subroutine finalize(self)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    if (c_associated(self%handle)) call check(st_solver_destroy(self%handle), 'st_solver_destroy')
    self%handle = c_null_ptr
  end subroutine finalize

  ! copy_from_field / copy_to_field (base_cpu.F90:1914-1929)
  subroutine copy_from_field(self, w)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    real(c_double), intent(in) :: w(*)
    call check(st_copy_from_field(self%handle, w), 'copy_from_field')
  end subroutine copy_from_field

  subroutine copy_to_field(self, w)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    real(c_double), intent(inout) :: w(*)
    call check(st_copy_to_field(self%handle, w), 'copy_to_field')
  end subroutine copy_to_field

  ! compute_dt (singleideal_cpu.F90:1476-1517)
  subroutine compute_dt(self, dt)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    real(c_double), intent(out) :: dt
    call check(st_compute_dt(self%handle, dt), 'compute_dt')
  end subroutine compute_dt

  ! compute_residual (singleideal_cpu.F90:1339-1364)
  subroutine compute_residual(self, s)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    type(st_scalars_t), intent(out) :: s
    call check(st_compute_residual(self%handle, s), 'compute_residual')
  end subroutine compute_residual

  ! print_progress (singleideal_cpu.F90:1397-1474), the channel columns:
  ! icyc, dt, time, residual, -dpdx/dt, bulk rho/u/T, extrema of rho/T/p
  subroutine print_progress(self, unit)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in) :: unit
    type(st_scalars_t) :: s
    call self%compute_residual(s)
    write(unit, '(i0,13es20.10)') s%icyc, s%dt, s%time, s%residual_rhou, -s%dpdx/s%dt, s%rhobulk, s%ubulk, &
      s%tbulk, s%rhomin, s%rhomax, s%tmin, s%tmax, s%pmin, s%pmax
  end subroutine print_progress

  ! The integration loop of run() (singleideal_cpu.F90:1599-1645): dt every
  ! iter_dt_recompute steps, rk_sync, time += dt, with the progress line
  ! every print_every steps; the output manager is the host's.
  subroutine run(self, nsteps, print_every, unit)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in) :: nsteps, print_every, unit
    integer :: done, n
    done = 0
    do while (done < nsteps)
      n = nsteps - done
      if (print_every > 0) n = min(n, print_every)
      call self%advance(n)
      done = done + n
      if (print_every > 0) call self%print_progress(unit)
    enddo
  end subroutine run

  ! rk_sync (singleideal_cpu.F90:332-409), fused stage kernels
  subroutine rk_sync(self)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    call check(st_rk_sync(self%handle), 'rk_sync')
  end subroutine rk_sync

  ! rk_sync procedure by procedure (singleideal_cpu.F90:352-406), one full
  ! Wray RK3 step including the dt of the step start and time += dt, with
  ! euler_x split at the middle of its range into two calls whose face
  ! fluxes the second call's update combines
  subroutine rk_sync_by_procedure(self)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    real(c_double), parameter :: rhork(3) = [0._c_double, -17._c_double/60._c_double, -5._c_double/12._c_double]
    real(c_double), parameter :: gamrk(3) = [8._c_double/15._c_double, 5._c_double/12._c_double, 3._c_double/4._c_double]
    type(st_scalars_t) :: s
    real(c_double) :: dt
    integer :: istep, imid, idt
    call self%scalars(s)
    idt = max(self%p%iter_dt_recompute, 1)
    if (self%p%cfl > 0._c_double .and. mod(int(s%icyc + 1) - self%icyc0, idt) == 0) then
      call self%compute_dt(dt)
    else
      dt = s%dt
    endif
    if (self%p%channel_case /= 0) call check(st_reset_forcing(self%handle), 'reset_forcing')
    imid = (self%eul(1) + self%eul(2)) / 2
    do istep = 1, 3
      call check(st_init_flux(self%handle, rhork(istep)*dt), 'init_flux')
      call check(st_bcswap(self%handle), 'bcswap')
      call self%compute_aux()
      call self%euler_x(int(self%eul(1)), imid, int(self%eul(1)) - 1, imid, .false.)
      call self%euler_x(imid + 1, int(self%eul(2)), imid, int(self%eul(2)), .true.)
      if (self%p%conservative_viscous == 1) then
        call self%visflx(6)
        call self%visflx(5)
      else
        call self%visflx(4)
      endif
      call self%bcextr_var(10)
      call self%bcswap_var(10)
      call self%euler_y(int(self%eul(3)), int(self%eul(4)))
      if (self%ndim == 3) call self%euler_z(int(self%eul(5)), int(self%eul(6)))
      if (istep == 3) then
        call self%visflx(3)
        call self%bcswap_var(8)
      endif
      if (self%p%conservative_viscous == 1) then
        call self%visflx(7)
      else
        call self%visflx(1)
      endif
      call self%bc_nr()
      call check(st_update_flux(self%handle, gamrk(istep)*dt), 'update_flux')
      if (self%p%channel_case /= 0) call self%force_rhs()
      call check(st_update_field(self%handle), 'update_field')
      if (self%p%channel_case /= 0) call self%force_var()
      call self%update_ghost(0)
    enddo
    call self%scalars(s)
    call check(st_set_time(self%handle, s%time + dt, dt, s%icyc + 1), 'set_time')
  end subroutine rk_sync_by_procedure

  ! update_ghost (singleideal_cpu.F90:1064-1138)
  subroutine update_ghost(self, do_swap)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in), optional :: do_swap
    integer(c_int32_t) :: ds
    ds = 1
    if (present(do_swap)) ds = do_swap
    call check(st_update_ghost(self%handle, ds), 'update_ghost')
  end subroutine update_ghost

  ! force_rhs (singleideal_cpu.F90:845-889)
  subroutine force_rhs(self)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    call check(st_force_rhs(self%handle), 'force_rhs')
  end subroutine force_rhs

  ! force_var (singleideal_cpu.F90:890-933)
  subroutine force_var(self)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    call check(st_force_var(self%handle), 'force_var')
  end subroutine force_var

  ! euler_x(istart, iend, istart_face, iend_face, do_update)
  ! (singleideal_cpu.F90:674-747)
  subroutine euler_x(self, istart, iend, istart_face, iend_face, do_update)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in) :: istart, iend
    integer, optional :: istart_face, iend_face
    logical, optional :: do_update
    integer(c_int32_t) :: f0, f1, du
    f0 = istart - 1
    if (present(istart_face)) f0 = istart_face
    f1 = iend
    if (present(iend_face)) f1 = iend_face
    du = 1
    if (present(do_update)) then
      if (.not. do_update) du = 0
    endif
    call check(st_euler_x(self%handle, int(istart, c_int32_t), int(iend, c_int32_t), f0, f1, du), 'euler_x')
  end subroutine euler_x

  ! euler_y(eul_jmin, eul_jmax) (singleideal_cpu.F90:749-801)
  subroutine euler_y(self, eul_jmin, eul_jmax)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in) :: eul_jmin, eul_jmax
    call check(st_euler_y(self%handle, int(eul_jmin, c_int32_t), int(eul_jmax, c_int32_t)), 'euler_y')
  end subroutine euler_y

  ! euler_z(eul_kmin, eul_kmax) (singleideal_cpu.F90:803-843), with the
  ! random inflow plane of a recycling inflow redrawn
  subroutine euler_z(self, eul_kmin, eul_kmax)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in) :: eul_kmin, eul_kmax
    call check(st_euler_z(self%handle, int(eul_kmin, c_int32_t), int(eul_kmax, c_int32_t)), 'euler_z')
  end subroutine euler_z

  ! visflx(mode) (singleideal_cpu.F90:186-279)
  subroutine visflx(self, mode)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in) :: mode
    call check(st_visflx(self%handle, int(mode, c_int32_t)), 'visflx')
  end subroutine visflx

  ! compute_aux(central, ghost) (singleideal_cpu.F90:281-330)
  subroutine compute_aux(self, central, ghost)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in), optional :: central, ghost
    integer(c_int32_t) :: c_, g_
    c_ = 1
    g_ = 1
    if (present(central)) c_ = central
    if (present(ghost)) g_ = ghost
    call check(st_compute_aux(self%handle, c_, g_), 'compute_aux')
  end subroutine compute_aux

  ! recyc_exchange (singleideal_cpu.F90:935-999)
  subroutine recyc_exchange(self)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    call check(st_recyc_exchange(self%handle), 'recyc_exchange')
  end subroutine recyc_exchange

  ! bc_nr (singleideal_cpu.F90:1001-1062)
  subroutine bc_nr(self)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    call check(st_bc_nr(self%handle), 'bc_nr')
  end subroutine bc_nr

  ! bcextr_var (singleideal_cpu.F90:411-426, mode 1) of w_aux variable var;
  ! the w_aux array lives on the device, so the variable is named by index
  subroutine bcextr_var(self, var)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in) :: var
    call check(st_bcextr_var(self%handle, int(var, c_int32_t)), 'bcextr_var')
  end subroutine bcextr_var

  ! bcswap_var of the base object (base_cpu.F90:1835-1912)
  subroutine bcswap_var(self, var)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in) :: var
    call check(st_bcswap_var(self%handle, int(var, c_int32_t)), 'bcswap_var')
  end subroutine bcswap_var

  ! count_weno (singleideal_cpu.F90:1366-1395)
  subroutine count_weno(self, frac)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    real(c_double), intent(out) :: frac(3)
    call check(st_count_weno(self%handle, frac), 'count_weno')
  end subroutine count_weno

  ! recyc_prepare's arrays (singleideal.F90:1512-1588) for the xmin
  ! recycling inflow: wmean_inflow(1-ng:0,1:ny,1:4), weta/map_j_*(1:ny),
  ! the y+ and eta coordinates of inflow and station (1-ng:ny+ng)
  subroutine set_recycling(self, i_recyc, betarecyc, glund1, wmean_inflow, weta, map_j_inn, map_j_out, &
                           yplus_inflow, eta_inflow, yplus_recyc, eta_recyc)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in) :: i_recyc
    real(c_double), intent(in) :: betarecyc, glund1
    real(c_double), intent(in) :: wmean_inflow(:,:,:), weta(:), yplus_inflow(:), eta_inflow(:)
    real(c_double), intent(in) :: yplus_recyc(:), eta_recyc(:)
    integer(c_int32_t), intent(in) :: map_j_inn(:), map_j_out(:)
    call check(st_set_recycling(self%handle, int(i_recyc, c_int32_t), betarecyc, glund1, wmean_inflow, weta, &
                                map_j_inn, map_j_out, yplus_inflow, eta_inflow, yplus_recyc, eta_recyc), &
               'set_recycling')
  end subroutine set_recycling

  ! wmean(1-ng:0,1:ny,1:4) of the laminar inflow (bctags 9)
  subroutine set_wmean(self, wmean)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    real(c_double), intent(in) :: wmean(:,:,:)
    call check(st_set_wmean(self%handle, wmean), 'set_wmean')
  end subroutine set_wmean

  ! compute_stats (singleideal.F90:2003-2179) on the device: the itav-th
  ! sample of the running statistics; get/set move w_stat(1:70,1:nx,1:ny)
  ! for write_stats / read_stats
  subroutine compute_stats(self, itav)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in) :: itav
    call check(st_compute_stats(self%handle, int(itav, c_int32_t)), 'compute_stats')
  end subroutine compute_stats

  subroutine get_stats(self, w_stat)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    real(c_double), intent(out) :: w_stat(:,:,:)
    call check(st_get_stats(self%handle, w_stat), 'get_stats')
  end subroutine get_stats

  subroutine set_stats(self, w_stat)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    real(c_double), intent(in) :: w_stat(:,:,:)
    call check(st_set_stats(self%handle, w_stat), 'set_stats')
  end subroutine set_stats

  ! the pre-loop of run() (singleideal_cpu.F90:1547-1594)
  subroutine init_run(self)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    type(st_scalars_t) :: s
    call check(st_init_run(self%handle), 'init_run')
    call self%scalars(s)
    self%icyc0 = int(s%icyc)
  end subroutine init_run

  ! nsteps of the loop body of run() with no host synchronisation
  subroutine advance(self, nsteps)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in) :: nsteps
    call check(st_advance(self%handle, int(nsteps, c_int32_t)), 'advance')
  end subroutine advance

  subroutine scalars(self, s)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    type(st_scalars_t), intent(out) :: s
    call check(st_get_scalars(self%handle, s), 'get_scalars')
  end subroutine scalars

end module streams_equation_singleideal_mi355x_object
