! streams_mi355x.F90 -- iso_c_binding interfaces of libstreams_mi355x
! (include/streams_mi355x.h).
!
! The module mirrors, one bind(c) interface per entry point, the C ABI that
! replaces the per-kernel bind(c) interfaces of the reference AMD backend
! (code/src/singleideal/kernels_amd.F90).  The backend object with the
! reference's procedure names is in singleideal_mi355x.F90.  check() stops
! the program with the library's message on a non-zero status, as
! fail_input_any does for the reference's input errors (parameters.F90:166-171).
module streams_mi355x
  use, intrinsic :: iso_c_binding
  implicit none
  public

  integer, parameter, public :: st_max_cp_coeff = 16

  ! st_params of the C header, field by field
  type, bind(c), public :: st_params_t
    integer(c_int32_t) :: nx, ny, nz, ng
    integer(c_int32_t) :: ep_order, visc_order, nkeep, weno_scheme, weno_version, flux_splitting
    integer(c_int32_t) :: conservative_viscous, visc_model, calorically_perfect, indx_cp_l, indx_cp_r
    integer(c_int32_t) :: bctags(6), bctags_nr(6)
    integer(c_int32_t) :: is_y_staggered, channel_case, correct_bound_ord, iter_dt_recompute
    integer(c_int32_t) :: precision, rand_type, pad1
    real(c_double) :: rgas0, prandtl, t0, rho0, u0, l0, mu0
    real(c_double) :: sutherland_s, t_ref_dim, powerlaw_vtexp
    real(c_double) :: sensor_threshold, t_wall, theta_wall, t_bulk_target, volchan, cfl, tol_iter_nr
    real(c_double) :: cp_coeff(st_max_cp_coeff), cv_coeff(st_max_cp_coeff)
    real(c_double) :: winf(5), winf_past_shock(5)
    real(c_double) :: xshock_imp, shock_angle, tanhfacs
    real(c_double) :: aoa, ptot0, ttot0
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
    integer(c_int) function st_euler_x(s, istart, iend, istart_face, iend_face, do_update) &
        bind(c, name="st_euler_x")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: s
      integer(c_int32_t), value :: istart, iend, istart_face, iend_face, do_update
    end function
    integer(c_int) function st_euler_y(s, eul_jmin, eul_jmax) bind(c, name="st_euler_y")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: s
      integer(c_int32_t), value :: eul_jmin, eul_jmax
    end function
    integer(c_int) function st_euler_z(s, eul_kmin, eul_kmax) bind(c, name="st_euler_z")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: s
      integer(c_int32_t), value :: eul_kmin, eul_kmax
    end function
    integer(c_int) function st_get_eul_range(s, range) bind(c, name="st_get_eul_range")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: s
      integer(c_int32_t), intent(out) :: range(6)
    end function
    integer(c_int) function st_init_flux(s, rhodt) bind(c, name="st_init_flux")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: s
      real(c_double), value :: rhodt
    end function
    integer(c_int) function st_update_flux(s, gamdt) bind(c, name="st_update_flux")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: s
      real(c_double), value :: gamdt
    end function
    integer(c_int) function st_update_field(s) bind(c, name="st_update_field")
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function st_force_rhs(s) bind(c, name="st_force_rhs")
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function st_force_var(s) bind(c, name="st_force_var")
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function st_reset_forcing(s) bind(c, name="st_reset_forcing")
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function st_bcswap(s) bind(c, name="st_bcswap")
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function st_bcextr_var(s, var) bind(c, name="st_bcextr_var")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: s
      integer(c_int32_t), value :: var
    end function
    integer(c_int) function st_bcswap_var(s, var) bind(c, name="st_bcswap_var")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: s
      integer(c_int32_t), value :: var
    end function
    integer(c_int) function st_recyc_exchange(s) bind(c, name="st_recyc_exchange")
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function st_set_wmean(s, wmean) bind(c, name="st_set_wmean")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: s
      real(c_double), intent(in) :: wmean(*)
    end function
    integer(c_int) function st_set_time(s, time, dt, icyc) bind(c, name="st_set_time")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: s
      real(c_double), value :: time, dt
      integer(c_int64_t), value :: icyc
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

end module streams_mi355x
