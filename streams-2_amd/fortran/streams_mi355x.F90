! streams_mi355x.F90 -- Fortran binding of libstreams_mi355x (include/streams_mi355x.h).
!
! The module mirrors, with iso_c_binding interfaces, the C ABI that replaces
! the per-kernel bind(c) interfaces of the reference AMD backend
! (code/src/singleideal/kernels_amd.F90) and wraps it in a backend object
! whose type-bound procedures carry the names of the reference backend
! object equation_singleideal_cpu_object (singleideal_cpu.F90:29-138):
! update_ghost, compute_aux, visflx, euler_x/y/z, compute_dt, rk_sync,
! compute_residual, count_weno, plus copy_from_field / copy_to_field of the
! base object (base_cpu.F90:1914-1929).  A non-zero status of any call stops
! the program with the library's message, as fail_input_any does for the
! reference's input errors (parameters.F90:166-171).
module streams_mi355x
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer, parameter, public :: st_max_cp_coeff = 16

  ! st_params of the C header, field by field
  type, bind(c), public :: st_params_t
    integer(c_int32_t) :: nx, ny, nz, ng
    integer(c_int32_t) :: ep_order, visc_order, nkeep, weno_scheme, weno_version, flux_splitting
    integer(c_int32_t) :: conservative_viscous, visc_model, calorically_perfect, indx_cp_l, indx_cp_r
    integer(c_int32_t) :: bctags(6), bctags_nr(6)
    integer(c_int32_t) :: is_y_staggered, channel_case, correct_bound_ord, iter_dt_recompute
    integer(c_int32_t) :: precision, pad0, pad1
    real(c_double) :: rgas0, prandtl, t0, rho0, u0, l0, mu0
    real(c_double) :: sutherland_s, t_ref_dim, powerlaw_vtexp
    real(c_double) :: sensor_threshold, t_wall, theta_wall, t_bulk_target, volchan, cfl, tol_iter_nr
    real(c_double) :: cp_coeff(st_max_cp_coeff), cv_coeff(st_max_cp_coeff)
    real(c_double) :: winf(5), winf_past_shock(5)
    real(c_double) :: xshock_imp, shock_angle, tanhfacs
  end type st_params_t

  type, bind(c), public :: st_scalars_t
    real(c_double) :: dt, time, dpdx, rhobulk, ubulk, tbulk, residual_rhou
    real(c_double) :: rhomin, rhomax, tmin, tmax, pmin, pmax
    integer(c_int64_t) :: icyc
  end type st_scalars_t

  interface
    integer(c_int) function st_solver_create(p, s) bind(c, name="st_solver_create")
      import :: c_int, c_ptr, st_params_t
      type(st_params_t), intent(in) :: p
      type(c_ptr), intent(out) :: s
    end function
    integer(c_int) function st_solver_destroy(s) bind(c, name="st_solver_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int64_t) function st_solver_footprint(p) bind(c, name="st_solver_footprint")
      import :: c_int64_t, st_params_t
      type(st_params_t), intent(in) :: p
    end function
    integer(c_int) function st_set_grid(s, x, y, z, yn, dcsidx, dcsidxs, dcsidx2, detady, detadys, detady2, &
        dzitdz, dzitdzs, dzitdz2) bind(c, name="st_set_grid")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: s
      real(c_double), intent(in) :: x(*), y(*), z(*), yn(*)
      real(c_double), intent(in) :: dcsidx(*), dcsidxs(*), dcsidx2(*)
      real(c_double), intent(in) :: detady(*), detadys(*), detady2(*)
      real(c_double), intent(in) :: dzitdz(*), dzitdzs(*), dzitdz2(*)
    end function
    integer(c_int) function st_copy_from_field(s, w) bind(c, name="st_copy_from_field")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: s
      real(c_double), intent(in) :: w(*)
    end function
    integer(c_int) function st_copy_to_field(s, w) bind(c, name="st_copy_to_field")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: s
      real(c_double), intent(inout) :: w(*)
    end function
    integer(c_int) function st_get_scalars(s, out) bind(c, name="st_get_scalars")
      import :: c_int, c_ptr, st_scalars_t
      type(c_ptr), value :: s
      type(st_scalars_t), intent(out) :: out
    end function
    integer(c_int) function st_update_ghost(s, do_swap) bind(c, name="st_update_ghost")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: s
      integer(c_int32_t), value :: do_swap
    end function
    integer(c_int) function st_compute_aux(s, central, ghost) bind(c, name="st_compute_aux")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: s
      integer(c_int32_t), value :: central, ghost
    end function
    integer(c_int) function st_visflx(s, mode) bind(c, name="st_visflx")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: s
      integer(c_int32_t), value :: mode
    end function
    integer(c_int) function st_euler_x(s) bind(c, name="st_euler_x")
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function st_euler_y(s) bind(c, name="st_euler_y")
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function st_euler_z(s) bind(c, name="st_euler_z")
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function st_bc_nr(s) bind(c, name="st_bc_nr")
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function st_compute_stats(s, itav) bind(c, name="st_compute_stats")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: s
      integer(c_int32_t), value :: itav
    end function
    integer(c_int) function st_get_stats(s, w_stat) bind(c, name="st_get_stats")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: s
      real(c_double), intent(out) :: w_stat(*)
    end function
    integer(c_int) function st_set_stats(s, w_stat) bind(c, name="st_set_stats")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: s
      real(c_double), intent(in) :: w_stat(*)
    end function
    integer(c_int) function st_set_recycling(s, i_recyc, betarecyc, glund1, wmean_inflow, weta, map_j_inn, &
                                             map_j_out, yplus_inflow, eta_inflow, yplus_recyc, eta_recyc) &
                                             bind(c, name="st_set_recycling")
      import :: c_int, c_int32_t, c_double, c_ptr
      type(c_ptr), value :: s
      integer(c_int32_t), value :: i_recyc
      real(c_double), value :: betarecyc, glund1
      real(c_double), intent(in) :: wmean_inflow(*), weta(*), yplus_inflow(*), eta_inflow(*)
      real(c_double), intent(in) :: yplus_recyc(*), eta_recyc(*)
      integer(c_int32_t), intent(in) :: map_j_inn(*), map_j_out(*)
    end function
    integer(c_int) function st_init_run(s) bind(c, name="st_init_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function st_compute_dt(s, dt) bind(c, name="st_compute_dt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: s
      real(c_double), intent(out) :: dt
    end function
    integer(c_int) function st_rk_sync(s) bind(c, name="st_rk_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function st_advance(s, nsteps) bind(c, name="st_advance")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: s
      integer(c_int32_t), value :: nsteps
    end function
    integer(c_int) function st_compute_residual(s, out) bind(c, name="st_compute_residual")
      import :: c_int, c_ptr, st_scalars_t
      type(c_ptr), value :: s
      type(st_scalars_t), intent(out) :: out
    end function
    integer(c_int) function st_count_weno(s, frac) bind(c, name="st_count_weno")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: s
      real(c_double), intent(out) :: frac(3)
    end function
    type(c_ptr) function st_last_error() bind(c, name="st_last_error")
      import :: c_ptr
    end function
  end interface

  ! Backend object holding the device state of one flow solve.
  type, public :: equation_singleideal_mi355x_object
    type(c_ptr) :: handle = c_null_ptr
    type(st_params_t) :: p
  contains
    procedure :: initialize => mi_initialize
    procedure :: finalize => mi_finalize
    procedure :: copy_from_field => mi_copy_from_field
    procedure :: copy_to_field => mi_copy_to_field
    procedure :: update_ghost => mi_update_ghost
    procedure :: compute_aux => mi_compute_aux
    procedure :: visflx => mi_visflx
    procedure :: euler_x => mi_euler_x
    procedure :: euler_y => mi_euler_y
    procedure :: euler_z => mi_euler_z
    procedure :: bc_nr => mi_bc_nr
    procedure :: set_recycling => mi_set_recycling
    procedure :: compute_stats => mi_compute_stats
    procedure :: get_stats => mi_get_stats
    procedure :: set_stats => mi_set_stats
    procedure :: init_run => mi_init_run
    procedure :: compute_dt => mi_compute_dt
    procedure :: rk_sync => mi_rk_sync
    procedure :: advance => mi_advance
    procedure :: compute_residual => mi_compute_residual
    procedure :: count_weno => mi_count_weno
    procedure :: scalars => mi_scalars
  end type equation_singleideal_mi355x_object

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc /= 0) then
      call c_f_pointer(st_last_error(), msg, [1024])
      n = 0
      do while (n < 1024)
        if (msg(n+1) == c_null_char) exit
        n = n + 1
      enddo
      write(*,'(a,a,a,i0,a)') 'streams_mi355x: ', what, ' failed (status ', rc, ')'
      if (n > 0) write(*,*) msg(1:n)
      error stop 15
    endif
  end subroutine check

  subroutine mi_initialize(self, p, x, y, z, yn, dcsidx, dcsidxs, dcsidx2, detady, detadys, detady2, &
      dzitdz, dzitdzs, dzitdz2)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    type(st_params_t), intent(in) :: p
    real(c_double), intent(in) :: x(*), y(*), z(*), yn(*)
    real(c_double), intent(in) :: dcsidx(*), dcsidxs(*), dcsidx2(*)
    real(c_double), intent(in) :: detady(*), detadys(*), detady2(*)
    real(c_double), intent(in) :: dzitdz(*), dzitdzs(*), dzitdz2(*)
    self%p = p
    call check(st_solver_create(p, self%handle), 'st_solver_create')
    call check(st_set_grid(self%handle, x, y, z, yn, dcsidx, dcsidxs, dcsidx2, detady, detadys, detady2, &
      dzitdz, dzitdzs, dzitdz2), 'st_set_grid')
  end subroutine mi_initialize

  subroutine mi_finalize(self)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    if (c_associated(self%handle)) call check(st_solver_destroy(self%handle), 'st_solver_destroy')
    self%handle = c_null_ptr
  end subroutine mi_finalize

  subroutine mi_copy_from_field(self, w)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    real(c_double), intent(in) :: w(*)
    call check(st_copy_from_field(self%handle, w), 'copy_from_field')
  end subroutine mi_copy_from_field

  subroutine mi_copy_to_field(self, w)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    real(c_double), intent(inout) :: w(*)
    call check(st_copy_to_field(self%handle, w), 'copy_to_field')
  end subroutine mi_copy_to_field

  subroutine mi_update_ghost(self, do_swap)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in), optional :: do_swap
    integer(c_int32_t) :: ds
    ds = 1
    if (present(do_swap)) ds = do_swap
    call check(st_update_ghost(self%handle, ds), 'update_ghost')
  end subroutine mi_update_ghost

  subroutine mi_compute_aux(self, central, ghost)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in), optional :: central, ghost
    integer(c_int32_t) :: c_, g_
    c_ = 1
    g_ = 1
    if (present(central)) c_ = central
    if (present(ghost)) g_ = ghost
    call check(st_compute_aux(self%handle, c_, g_), 'compute_aux')
  end subroutine mi_compute_aux

  subroutine mi_visflx(self, mode)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in) :: mode
    call check(st_visflx(self%handle, int(mode, c_int32_t)), 'visflx')
  end subroutine mi_visflx

  subroutine mi_euler_x(self)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    call check(st_euler_x(self%handle), 'euler_x')
  end subroutine mi_euler_x

  subroutine mi_euler_y(self)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    call check(st_euler_y(self%handle), 'euler_y')
  end subroutine mi_euler_y

  subroutine mi_euler_z(self)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    call check(st_euler_z(self%handle), 'euler_z')
  end subroutine mi_euler_z

  subroutine mi_bc_nr(self)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    call check(st_bc_nr(self%handle), 'bc_nr')
  end subroutine mi_bc_nr

  ! recyc_prepare's arrays (singleideal.F90:1512-1588) for the xmin
  ! recycling inflow: wmean_inflow(1-ng:0,1:ny,1:4), weta/map_j_*(1:ny),
  ! the y+ and eta coordinates of inflow and station (1-ng:ny+ng)
  subroutine mi_set_recycling(self, i_recyc, betarecyc, glund1, wmean_inflow, weta, map_j_inn, map_j_out, &
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
  end subroutine mi_set_recycling

  ! compute_stats (singleideal.F90:2003-2179) on the device: the itav-th
  ! sample of the running statistics; get/set move w_stat(1:70,1:nx,1:ny)
  ! for write_stats / read_stats
  subroutine mi_compute_stats(self, itav)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in) :: itav
    call check(st_compute_stats(self%handle, int(itav, c_int32_t)), 'compute_stats')
  end subroutine mi_compute_stats

  subroutine mi_get_stats(self, w_stat)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    real(c_double), intent(out) :: w_stat(:,:,:)
    call check(st_get_stats(self%handle, w_stat), 'get_stats')
  end subroutine mi_get_stats

  subroutine mi_set_stats(self, w_stat)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    real(c_double), intent(in) :: w_stat(:,:,:)
    call check(st_set_stats(self%handle, w_stat), 'set_stats')
  end subroutine mi_set_stats

  subroutine mi_init_run(self)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    call check(st_init_run(self%handle), 'init_run')
  end subroutine mi_init_run

  subroutine mi_compute_dt(self, dt)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    real(c_double), intent(out) :: dt
    call check(st_compute_dt(self%handle, dt), 'compute_dt')
  end subroutine mi_compute_dt

  subroutine mi_rk_sync(self)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    call check(st_rk_sync(self%handle), 'rk_sync')
  end subroutine mi_rk_sync

  subroutine mi_advance(self, nsteps)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    integer, intent(in) :: nsteps
    call check(st_advance(self%handle, int(nsteps, c_int32_t)), 'advance')
  end subroutine mi_advance

  subroutine mi_compute_residual(self, s)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    type(st_scalars_t), intent(out) :: s
    call check(st_compute_residual(self%handle, s), 'compute_residual')
  end subroutine mi_compute_residual

  subroutine mi_count_weno(self, frac)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    real(c_double), intent(out) :: frac(3)
    call check(st_count_weno(self%handle, frac), 'count_weno')
  end subroutine mi_count_weno

  subroutine mi_scalars(self, s)
    class(equation_singleideal_mi355x_object), intent(inout) :: self
    type(st_scalars_t), intent(out) :: s
    call check(st_get_scalars(self%handle, s), 'get_scalars')
  end subroutine mi_scalars

end module streams_mi355x
