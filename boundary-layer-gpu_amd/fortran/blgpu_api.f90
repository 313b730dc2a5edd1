!------------------------------------------------------------------------------!
! ISO_C_BINDING interface to the blgpu C ABI (include/blgpu.h).                 !
!                                                                              !
! This module is the binding a Fortran caller of the reference solver uses to  !
! replace the OpenACC data region (main.f90:61-81), the per-step OpenACC       !
! regions and cuFFT plans (time_integration.f90, projection.f90,               !
! cufft_solver.f90) by calls into the HIP library.                             !
!------------------------------------------------------------------------------!
module blgpu_api
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: BL_FIELD_U = 0, BL_FIELD_V = 1, BL_FIELD_W = 2
  integer(c_int), parameter :: BL_FIELD_P = 3, BL_FIELD_NUT = 4
  integer(c_int), parameter :: BL_PREC_FP64 = 0, BL_PREC_MIXED = 1

  ! struct bl_config (same member order and types as the C declaration)
  type, bind(C) :: bl_config
     integer(c_int) :: nx, ny, nz
     type(c_ptr)    :: x, y, z
     real(c_double) :: nu, dPdx, dPdz, cfl
     integer(c_int) :: scheme, les, precision
     type(c_ptr)    :: U_inlet, V_inlet, W_inlet, U_top, V_top, W_top
     integer(c_int) :: n_modes, m_modes
     type(c_ptr)    :: zmode, tmode
     type(c_ptr)    :: qu_re, qu_im, qv_re, qv_im, qw_re, qw_im
     integer(c_int) :: nut_bc
  end type bl_config

  ! bl_boundary: the general dispatcher's flags (boundary_conditions.f90:22-121)
  type, bind(C) :: bl_boundary
     integer(c_int)       :: inflow_flag, top_flag
     real(c_double)       :: amplitude
     integer(c_long_long) :: seed
     real(c_double)       :: vbs_max, x_bs, sigma_bs, phi_bs
     integer(c_int)       :: i_rescale
     real(c_double)       :: beta_hartree
     real(c_double)       :: delta_inlet, t_resc
  end type bl_boundary

  ! bl_wall_params: the wall model (compute_wall_model wallmodel.f90:25-144)
  type, bind(C) :: bl_wall_params
     integer(c_int) :: iwall_model, istress_model
     real(c_double) :: alpha_mean_x, alpha_mean_y, alpha_mean_z, alpha_std, freq_mult
     integer(c_int) :: iwall_model_nut
     real(c_double) :: frac_vis
  end type bl_wall_params

  interface
     integer(c_int) function bl_create(cfg, h) bind(C, name='bl_create')
       import :: c_int, c_ptr, bl_config
       type(bl_config), intent(in) :: cfg
       type(c_ptr), intent(out)    :: h
     end function bl_create

     subroutine bl_destroy(h) bind(C, name='bl_destroy')
       import :: c_ptr
       type(c_ptr), value :: h
     end subroutine bl_destroy

     type(c_ptr) function bl_last_error_c() bind(C, name='bl_last_error')
       import :: c_ptr
     end function bl_last_error_c

     integer(c_int) function bl_upload(h, field, host) bind(C, name='bl_upload')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value       :: h
       integer(c_int), value    :: field
       real(c_double), intent(in) :: host(*)
     end function bl_upload

     integer(c_int) function bl_download(h, field, host) bind(C, name='bl_download')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value        :: h
       integer(c_int), value     :: field
       real(c_double), intent(out) :: host(*)
     end function bl_download

     integer(c_int) function bl_set_time(h, t) bind(C, name='bl_set_time')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value    :: h
       real(c_double), value :: t
     end function bl_set_time

     integer(c_int) function bl_set_boundary(h, b) bind(C, name='bl_set_boundary')
       import :: c_int, c_ptr, bl_boundary
       type(c_ptr), value            :: h
       type(bl_boundary), intent(in) :: b
     end function bl_set_boundary

     integer(c_int) function bl_set_wall_model(h, p) bind(C, name='bl_set_wall_model')
       import :: c_int, c_ptr, bl_wall_params
       type(c_ptr), value               :: h
       type(bl_wall_params), intent(in) :: p
     end function bl_set_wall_model

     integer(c_int) function bl_get_time(h, t) bind(C, name='bl_get_time')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value          :: h
       real(c_double), intent(out) :: t
     end function bl_get_time

     integer(c_int) function bl_compute_dt(h, dt) bind(C, name='bl_compute_dt')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value          :: h
       real(c_double), intent(out) :: dt
     end function bl_compute_dt

     integer(c_int) function bl_step(h) bind(C, name='bl_step')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function bl_step

     integer(c_int) function bl_synchronize(h) bind(C, name='bl_synchronize')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function bl_synchronize

     integer(c_int) function bl_eddy_viscosity(h) bind(C, name='bl_eddy_viscosity')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function bl_eddy_viscosity

     integer(c_int) function bl_max_divergence(h, div) bind(C, name='bl_max_divergence')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value          :: h
       real(c_double), intent(out) :: div
     end function bl_max_divergence

     integer(c_int) function bl_wall_stats(h, cf) bind(C, name='bl_wall_stats')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value          :: h
       real(c_double), intent(out) :: cf(*)
     end function bl_wall_stats

     integer(c_int) function bl_monitor(h, out) bind(C, name='bl_monitor')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value          :: h
       real(c_double), intent(out) :: out(8)
     end function bl_monitor

     ! compute_statistics' z means on the device (statistics.f90:44-80)
     integer(c_int) function bl_statistics(h, means, nu_t_mean, cf) bind(C, name='bl_statistics')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value          :: h
       real(c_double), intent(out) :: means(*)
       type(c_ptr), value          :: nu_t_mean
       real(c_double), intent(out) :: cf(*)
     end function bl_statistics

     ! snapshot fields copied to pinned host memory behind the time loop
     integer(c_int) function bl_snapshot_begin(h) bind(C, name='bl_snapshot_begin')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function bl_snapshot_begin

     integer(c_int) function bl_snapshot_end(h, pu, pv, pw, pp, pn) bind(C, name='bl_snapshot_end')
       import :: c_int, c_ptr
       type(c_ptr), value        :: h
       type(c_ptr), intent(out)  :: pu, pv, pw, pp, pn
     end function bl_snapshot_end

     integer(c_int) function bl_blasius_tables(nx, ny, x, y, nu, n, eta, f, df, U_inlet, V_inlet, &
          W_inlet, U_top, V_top, W_top) bind(C, name='bl_blasius_tables')
       import :: c_int, c_double
       integer(c_int), value       :: nx, ny, n
       real(c_double), value       :: nu
       real(c_double), intent(in)  :: x(*), y(*), eta(*), f(*), df(*)
       real(c_double), intent(out) :: U_inlet(*), V_inlet(*), W_inlet(*), U_top(*), V_top(*), W_top(*)
     end function bl_blasius_tables

     integer(c_int) function bl_blasius_field(nx, ny, nz, x, y, nu, n, eta, f, df, U, V, W) &
          bind(C, name='bl_blasius_field')
       import :: c_int, c_double
       integer(c_int), value       :: nx, ny, nz, n
       real(c_double), value       :: nu
       real(c_double), intent(in)  :: x(*), y(*), eta(*), f(*), df(*)
       real(c_double), intent(out) :: U(*), V(*), W(*)
     end function bl_blasius_field
  end interface

contains

  ! Message of the last failure, as a Fortran string
  function bl_last_error() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: cs(:)
    integer :: n
    p = bl_last_error_c()
    msg = ''
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [4096])
    n = 0
    do while (n < 4096)
       if (cs(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(cs(1:n), msg)
  end function bl_last_error

  ! Stop with the library's message on failure, as the reference Stops on
  ! errors (cufft_solver.f90:39).
  subroutine bl_check(ierr, what)
    integer(c_int), intent(in) :: ierr
    character(len=*), intent(in) :: what
    if (ierr /= 0) then
       write(*,*) 'blgpu error in ', what, ': ', bl_last_error()
       error stop 1
    end if
  end subroutine bl_check

end module blgpu_api
