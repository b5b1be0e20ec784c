! icls_capi.f90 -- iso_c_binding interface to libicls.so (include/icls.h).
!
! The derived type icls_params is interoperable with struct icls_params; every
! interface below binds one extern "C" entry point of the library.
module icls_capi
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: ICLS_OK = 0, ICLS_ERR_PARAM = 1, ICLS_ERR_HIP = 2, ICLS_ERR_NAN = 3, &
                                   ICLS_ERR_DIVERGENCE = 4, ICLS_ERR_STATE = 5, ICLS_ERR_NODEVICE = 6, &
                                   ICLS_ERR_UNSUPPORTED = 7
  integer(c_int32_t), parameter :: ICLS_MAX_DROPS = 1024
  ! BC_in_y of solver2d_cos (initsolver.f90:84-124)
  integer(c_int32_t), parameter :: ICLS_BCY_WALLS = 1, ICLS_BCY_INOUT = 2, ICLS_BCY_OUTOUT = 3

  ! field ids (enum icls_field)
  integer(c_int32_t), parameter :: F_UNEW = 0, F_VNEW = 1, F_WNEW = 2, F_DUDT = 3, F_DVDT = 4, F_DWDT = 5, &
       F_PNEW = 6, F_POLD = 7, F_P = 8, F_PJUMP = 9, F_DUOLD = 10, F_DVOLD = 11, F_DWOLD = 12, &
       F_PX = 13, F_PY = 14, F_PZ = 15, F_PXOLD = 16, F_PYOLD = 17, F_PZOLD = 18, F_MIU = 19, F_RHO = 20, &
       F_JIEMIAN = 21, F_CURV = 22, F_UCC = 23, F_VCC = 24, F_WCC = 25, F_RHO_U = 26, F_RHO_V = 27, &
       F_RHO_W = 28, F_LVSET = 29, F_NORMAL_M = 30
  integer(c_int32_t), parameter :: S_TIME = 0, S_DT = 1, S_RHO_AVR = 2, S_V_BULK = 3, S_MAXCHK = 4, &
       S_REINIT_RES = 5, S_NORM_OLD = 6

  type, bind(c) :: icls_params
     integer(c_int32_t) :: itot, jtot, ktot, lmax
     real(c_double)     :: lx, ly, lz, timestep, rho1, rho2, miu1, miu2, Re, We, iFr
     real(c_double)     :: g_unit(3)
     integer(c_int32_t) :: ls_adv_space, re_initialization, reinit_step, reinit_iter_max, &
                           mass_correction, inflate_step, ns_np
     real(c_double)     :: wall_speed
     integer(c_int32_t) :: constPoi, iniu, mode, ioutchk
     real(c_double)     :: cfl
     integer(c_int32_t) :: ls_full
     integer(c_int32_t) :: mass_trigger_exp
     integer(c_int32_t) :: reserved(4)
     real(c_double)     :: drop_xc(ICLS_MAX_DROPS), drop_yc(ICLS_MAX_DROPS), drop_zc(ICLS_MAX_DROPS), &
                           drop_r(ICLS_MAX_DROPS)
  end type icls_params

  interface
     function icls_last_error() bind(c, name='icls_last_error') result(msg)
       import :: c_ptr
       type(c_ptr) :: msg
     end function
     integer(c_int32_t) function icls_init(p) bind(c, name='icls_init')
       import :: c_int32_t, icls_params
       type(icls_params), intent(in) :: p
     end function
     integer(c_int32_t) function icls_finalize() bind(c, name='icls_finalize')
       import :: c_int32_t
     end function
     integer(c_int32_t) function icls_init_program() bind(c, name='icls_init_program')
       import :: c_int32_t
     end function
     integer(c_int32_t) function icls_upload(field, l, host) bind(c, name='icls_upload')
       import :: c_int32_t, c_double
       integer(c_int32_t), value :: field, l
       real(c_double), intent(in) :: host(*)
     end function
     integer(c_int32_t) function icls_download(field, l, host) bind(c, name='icls_download')
       import :: c_int32_t, c_double
       integer(c_int32_t), value :: field, l
       real(c_double), intent(out) :: host(*)
     end function
     integer(c_int32_t) function icls_get_scalar(which, v) bind(c, name='icls_get_scalar')
       import :: c_int32_t, c_double
       integer(c_int32_t), value :: which
       real(c_double), intent(out) :: v
     end function
     integer(c_int32_t) function icls_set_scalar(which, v) bind(c, name='icls_set_scalar')
       import :: c_int32_t, c_double
       integer(c_int32_t), value :: which
       real(c_double), value :: v
     end function
     integer(c_int32_t) function icls_get_volume(l, v, v0) bind(c, name='icls_get_volume')
       import :: c_int32_t, c_double
       integer(c_int32_t), value :: l
       real(c_double), intent(out) :: v, v0
     end function
     integer(c_int32_t) function icls_sync() bind(c, name='icls_sync')
       import :: c_int32_t
     end function
     integer(c_int32_t) function icls_ls_main(istep) bind(c, name='icls_ls_main')
       import :: c_int32_t
       integer(c_int32_t), value :: istep
     end function
     integer(c_int32_t) function icls_average_stream_vel(v) bind(c, name='icls_average_stream_vel')
       import :: c_int32_t, c_double
       real(c_double), intent(out) :: v
     end function
     integer(c_int32_t) function icls_ab2() bind(c, name='icls_ab2')
       import :: c_int32_t
     end function
     integer(c_int32_t) function icls_fillps() bind(c, name='icls_fillps')
       import :: c_int32_t
     end function
     integer(c_int32_t) function icls_gfm4p() bind(c, name='icls_gfm4p')
       import :: c_int32_t
     end function
     integer(c_int32_t) function icls_add_pjump() bind(c, name='icls_add_pjump')
       import :: c_int32_t
     end function
     integer(c_int32_t) function icls_solver2d() bind(c, name='icls_solver2d')
       import :: c_int32_t
     end function
     integer(c_int32_t) function icls_solver2d_cos(bc_y) bind(c, name='icls_solver2d_cos')
       import :: c_int32_t
       integer(c_int32_t), value :: bc_y
     end function
     integer(c_int32_t) function icls_boundp(field) bind(c, name='icls_boundp')
       import :: c_int32_t
       integer(c_int32_t), value :: field
     end function
     integer(c_int32_t) function icls_correc() bind(c, name='icls_correc')
       import :: c_int32_t
     end function
     integer(c_int32_t) function icls_bounduvw(which) bind(c, name='icls_bounduvw')
       import :: c_int32_t
       integer(c_int32_t), value :: which
     end function
     integer(c_int32_t) function icls_update_pressure(first) bind(c, name='icls_update_pressure')
       import :: c_int32_t
       integer(c_int32_t), value :: first
     end function
     integer(c_int32_t) function icls_average_vel_at_center() bind(c, name='icls_average_vel_at_center')
       import :: c_int32_t
     end function
     integer(c_int32_t) function icls_check_nan() bind(c, name='icls_check_nan')
       import :: c_int32_t
     end function
     integer(c_int32_t) function icls_chkdiv(istep, divtot, divmax, ijk) bind(c, name='icls_chkdiv')
       import :: c_int32_t, c_double
       integer(c_int32_t), value :: istep
       real(c_double), intent(out) :: divtot, divmax
       integer(c_int32_t), intent(out) :: ijk(3)
     end function
     integer(c_int32_t) function icls_chkdt(dtmax) bind(c, name='icls_chkdt')
       import :: c_int32_t, c_double
       real(c_double), intent(out) :: dtmax
     end function
     integer(c_int32_t) function icls_max_vel_at_center(vmax) bind(c, name='icls_max_vel_at_center')
       import :: c_int32_t, c_double
       real(c_double), intent(out) :: vmax
     end function
     integer(c_int32_t) function icls_spurious_currents(Ca) bind(c, name='icls_spurious_currents')
       import :: c_int32_t, c_double
       real(c_double), intent(out) :: Ca
     end function
     integer(c_int32_t) function icls_step(istep, begin) bind(c, name='icls_step')
       import :: c_int32_t
       integer(c_int32_t), value :: istep, begin
     end function
     integer(c_int32_t) function icls_write_fld(path, nr) bind(c, name='icls_write_fld')
       import :: c_int32_t, c_char
       character(kind=c_char), dimension(*), intent(in) :: path
       integer(c_int32_t), value :: nr
     end function
     integer(c_int32_t) function icls_init_program_restart(path, begin) bind(c, name='icls_init_program_restart')
       import :: c_int32_t, c_char
       character(kind=c_char), dimension(*), intent(in) :: path
       integer(c_int32_t), intent(out) :: begin
     end function
     integer(c_int32_t) function icls_write_vtk(path) bind(c, name='icls_write_vtk')
       import :: c_int32_t, c_char
       character(kind=c_char), dimension(*), intent(in) :: path
     end function
     integer(c_int32_t) function icls_post2d(path) bind(c, name='icls_post2d')
       import :: c_int32_t, c_char
       character(kind=c_char), dimension(*), intent(in) :: path
     end function
  end interface

contains

  ! print the library's message and stop with a nonzero status, as the reference's
  ! error paths print then stop (chkdiv.f90:60-86, main.f90:207-210)
  subroutine icls_check(rc, where)
    integer(c_int32_t), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: s(:)
    integer :: n
    if (rc == ICLS_OK) return
    call c_f_pointer(icls_last_error(), s, [1024])
    n = 0
    do while (n < 1024)
       if (s(n + 1) == c_null_char) exit
       n = n + 1
    end do
    write(6, '(A,A,I3,A,1024A1)') where, ' failed (code ', rc, '): ', s(1:n)
    error stop 1
  end subroutine icls_check

end module icls_capi
