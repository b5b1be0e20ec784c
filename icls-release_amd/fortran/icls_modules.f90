! icls_modules.f90 -- the reference's per-step subroutines under their own module
! and subroutine names, forwarding to libicls.so through icls_capi.
!
! The state lives on the GPU, owned by the library.  The array arguments keep the
! reference's signatures so that a main program written against the reference
! compiles unchanged; by default their contents are not touched.  With the
! environment variable ICLS_SYNC=1 every wrapper copies the device result back into
! the array it was given (debug / A-B mode).
module mod_icls_state
  use iso_c_binding
  use icls_capi
  implicit none
  type(icls_params), save :: prm
  integer, save :: imax = 0, jmax = 0, kmax = 0, i1 = 0, j1 = 0, k1 = 0, lmax = 0
  logical, save :: sync_host = .false.
  ! BC_in_y (param.f90:115): the y rule solver2d_cos applies ('Walls', 'InOut', 'OutOut')
  character(len=8), save :: bc_in_y = 'Periodic'
contains
  subroutine icls_setup(p)
    type(icls_params), intent(in) :: p
    character(len=8) :: env
    integer :: st
    prm = p
    imax = p%itot; jmax = p%jtot; kmax = p%ktot
    i1 = imax + 1; j1 = jmax + 1; k1 = kmax + 1
    lmax = p%lmax
    call get_environment_variable('ICLS_SYNC', env, status=st)
    sync_host = (st == 0 .and. env(1:1) == '1')
    call icls_check(icls_init(prm), 'icls_init')
  end subroutine icls_setup

  ! download one C-halo field into a (0:i1,0:j1,0:k1) array
  subroutine pull_c(field, a)
    integer(c_int32_t), intent(in) :: field
    real(c_double), intent(out) :: a(0:i1, 0:j1, 0:k1)
    call icls_check(icls_download(field, 0_c_int32_t, a), 'icls_download')
  end subroutine pull_c

  subroutine pull_i(field, a)
    integer(c_int32_t), intent(in) :: field
    real(c_double), intent(out) :: a(1:imax, 1:jmax, 1:kmax)
    call icls_check(icls_download(field, 0_c_int32_t, a), 'icls_download')
  end subroutine pull_i
end module mod_icls_state

! init_program (src/init.f90:22-117)
module mod_init
  use mod_icls_state
  implicit none
  private
  public init_program
contains
  subroutine init_program(begin, duold, dvold, dwold, pold, curv_restart)
    integer, intent(in) :: begin
    real, dimension(1:, 1:, 1:), intent(out) :: duold, dvold, dwold
    real, dimension(0:, 0:, 0:), intent(out) :: pold, curv_restart
    character(len=7) :: istepchar
    integer(c_int32_t) :: nr
    if (begin == 0) then
       call icls_check(icls_init_program(), 'init_program')
    else
       ! init.f90:44-51: the fields of data/fld<begin> (loadflds(0,begin)), then the
       ! derived state
       write(istepchar, '(i7.7)') begin
       call icls_check(icls_init_program_restart('data/fld'//istepchar//c_null_char, nr), 'init_program')
       write(6, *) 'Restarting from step ', begin
    end if
    curv_restart = 0.
    if (sync_host) then
       call pull_i(F_DUOLD, duold); call pull_i(F_DVOLD, dvold); call pull_i(F_DWOLD, dwold)
       call pull_c(F_POLD, pold)
    end if
  end subroutine init_program
end module mod_init

! ls_main (src/levelset.f90:28-76)
module mod_levelset
  use mod_icls_state
  implicit none
  private
  public ls_main
contains
  subroutine ls_main(istep)
    integer, intent(in) :: istep
    call icls_check(icls_ls_main(int(istep, c_int32_t)), 'ls_main')
  end subroutine ls_main
end module mod_levelset

! average_vel_at_center, average_stream_vel, max_vel_at_center (src/misc.f90:587-655)
module mod_misc
  use mod_icls_state
  implicit none
  private
  public average_vel_at_center, average_stream_vel, max_vel_at_center, spurious_currents
contains
  subroutine average_vel_at_center
    call icls_check(icls_average_vel_at_center(), 'average_vel_at_center')
  end subroutine average_vel_at_center
  subroutine average_stream_vel(v_avr)
    real, intent(inout) :: v_avr
    call icls_check(icls_average_stream_vel(v_avr), 'average_stream_vel')
  end subroutine average_stream_vel
  subroutine max_vel_at_center(vel_mag_max_all)
    real, intent(out) :: vel_mag_max_all
    call icls_check(icls_max_vel_at_center(vel_mag_max_all), 'max_vel_at_center')
  end subroutine max_vel_at_center
  subroutine spurious_currents(Ca)
    real, intent(out) :: Ca
    call icls_check(icls_spurious_currents(Ca), 'spurious_currents')
  end subroutine spurious_currents
end module mod_misc

! ab2 (src/ab2.f90:21-138)
module mod_ab
  use mod_icls_state
  implicit none
  private
  public ab2
contains
  subroutine ab2(duold, dvold, dwold, pold)
    real, dimension(1:, 1:, 1:), intent(inout) :: duold, dvold, dwold
    real, dimension(0:, 0:, 0:), intent(in) :: pold
    call icls_check(icls_ab2(), 'ab2')
    if (sync_host) then
       call pull_i(F_DUOLD, duold); call pull_i(F_DVOLD, dvold); call pull_i(F_DWOLD, dwold)
    end if
  end subroutine ab2
end module mod_ab

! fillps (src/fillps.f90:13-57)
module mod_fillps
  use mod_icls_state
  implicit none
  private
  public fillps
contains
  subroutine fillps(p)
    real, intent(out), dimension(0:, 0:, 0:) :: p
    call icls_check(icls_fillps(), 'fillps')
    if (sync_host) call pull_c(F_P, p)
  end subroutine fillps
end module mod_fillps

! gfm4p (src/gfm.f90:17-267)
module mod_gfm
  use mod_icls_state
  implicit none
  private
  public gfm4p
contains
  subroutine gfm4p(p_jump)
    real, dimension(0:, 0:, 0:), intent(out) :: p_jump
    call icls_check(icls_gfm4p(), 'gfm4p')
    if (sync_host) call pull_c(F_PJUMP, p_jump)
  end subroutine gfm4p
end module mod_gfm

! solver2d (src/solver.f90:68-196) and solver2d_cos (solver.f90:202-317); pz is the
! interior section of p
module mod_solver
  use mod_icls_state
  implicit none
  private
  public solver2d, solver2d_cos
contains
  subroutine solver2d(pz)
    real, intent(inout), dimension(1:, 1:, 1:) :: pz
    real, allocatable :: tmp(:, :, :)
    call icls_check(icls_solver2d(), 'solver2d')
    if (sync_host) then
       allocate(tmp(0:i1, 0:j1, 0:k1))
       call pull_c(F_P, tmp)
       pz = tmp(1:imax, 1:jmax, 1:kmax)
       deallocate(tmp)
    end if
  end subroutine solver2d

  ! solver2d_cos (src/solver.f90:202-317) with the y rule bc_in_y
  subroutine solver2d_cos(pz)
    real, intent(inout), dimension(1:, 1:, 1:) :: pz
    real, allocatable :: tmp(:, :, :)
    integer(c_int32_t) :: rule
    select case (trim(bc_in_y))
    case ('Walls')
       rule = ICLS_BCY_WALLS
    case ('InOut')
       rule = ICLS_BCY_INOUT
    case ('OutOut')
       rule = ICLS_BCY_OUTOUT
    case default
       write(6, *) 'solver2d_cos: BC_in_y must be Walls, InOut or OutOut, got ', bc_in_y
       error stop 1
    end select
    call icls_check(icls_solver2d_cos(rule), 'solver2d_cos')
    if (sync_host) then
       allocate(tmp(0:i1, 0:j1, 0:k1))
       call pull_c(F_P, tmp)
       pz = tmp(1:imax, 1:jmax, 1:kmax)
       deallocate(tmp)
    end if
  end subroutine solver2d_cos
end module mod_solver

! boundp, bounduvw (src/bound.f90:15-287) for the arrays main.f90 passes
module mod_bound
  use mod_icls_state
  implicit none
  private
  public boundp, bounduvw
contains
  subroutine boundp(p)
    real, dimension(0:, 0:, 0:) :: p
    call icls_check(icls_boundp(F_P), 'boundp')
    if (sync_host) call pull_c(F_P, p)
  end subroutine boundp
  subroutine bounduvw(u, v, w)
    real, dimension(0:, 0:, 0:) :: u, v, w
    call icls_check(icls_bounduvw(0_c_int32_t), 'bounduvw')
    if (sync_host) then
       call pull_c(F_UNEW, u); call pull_c(F_VNEW, v); call pull_c(F_WNEW, w)
    end if
  end subroutine bounduvw
end module mod_bound

! correc (src/correc.f90:15-44)
module mod_correc
  use mod_icls_state
  implicit none
  private
  public correc
contains
  subroutine correc(p)
    real, intent(in), dimension(0:, 0:, 0:) :: p
    call icls_check(icls_correc(), 'correc')
  end subroutine correc
end module mod_correc

! chkdiv (src/chkdiv.f90:10-90): prints like the reference and stops on > 1e-12
module mod_chkdiv
  use mod_icls_state
  implicit none
  private
  public chkdiv
contains
  subroutine chkdiv(istep)
    integer, intent(in) :: istep
    real :: divtot, divmax
    integer(c_int32_t) :: ijk(3), rc
    rc = icls_chkdiv(int(istep, c_int32_t), divtot, divmax, ijk)
    write(6, '(A,I5,A,I5,A,I5)') '! Maximal divergence at i = ', ijk(1), ' j = ', ijk(2), ' k = ', ijk(3)
    write(6, '(A,ES13.6,A,ES13.6,A,I5)') '! Divergence: Tot = ', divtot, ' L_infty = ', divmax, ' Rank = ', 0
    if (rc == ICLS_ERR_DIVERGENCE) then
       write(6, *) 'Error: Magnitude of the maximal divergence > 1e-12.'
       write(6, *) 'Program aborted.'
       error stop 4
    end if
    call icls_check(rc, 'chkdiv')
  end subroutine chkdiv
end module mod_chkdiv

! chkdt (src/chkdt.f90:11-105)
module mod_chkdt
  use mod_icls_state
  implicit none
  private
  public chkdt
contains
  subroutine chkdt(dtmax)
    real, intent(inout) :: dtmax
    call icls_check(icls_chkdt(dtmax), 'chkdt')
  end subroutine chkdt
end module mod_chkdt

! main.f90:142 and main.f90:164-185 as calls (they are array statements in the
! reference's main program)
module mod_icls_main_ops
  use mod_icls_state
  implicit none
  private
  public add_pjump, update_pressure, check_nan
contains
  subroutine add_pjump(p, p_jump)
    real, dimension(0:, 0:, 0:) :: p, p_jump
    call icls_check(icls_add_pjump(), 'p = p + p_jump')
  end subroutine add_pjump
  subroutine update_pressure(first_step)
    logical, intent(in) :: first_step
    integer(c_int32_t) :: f
    f = 0
    if (first_step) f = 1
    call icls_check(icls_update_pressure(f), 'update_pressure')
  end subroutine update_pressure
  subroutine check_nan()
    integer(c_int32_t) :: rc
    rc = icls_check_nan()
    if (rc == ICLS_ERR_NAN) then
       write(6, *) 'Got NaN... Terminate brutally. :('
       error stop 3
    end if
    call icls_check(rc, 'check_nan')
  end subroutine check_nan
end module mod_icls_main_ops

! write_vtk (src/write_vtk.f90:10-39): datadir//'para'//istep(i7.7)//'.vtk'
module mod_write_vtk
  use mod_icls_state
  implicit none
  private
  public write_vtk
contains
  subroutine write_vtk(istep)
    integer, intent(in) :: istep
    character(len=7) :: istepchar
    write(istepchar, '(i7.7)') istep
    call icls_check(icls_write_vtk('data/para'//istepchar//'.vtk'//c_null_char), 'write_vtk')
  end subroutine write_vtk
end module mod_write_vtk

! post2d (src/output.f90:179-427): datadir//'yz_mid'//istep(i7.7)//'.dat'
module mod_output
  use mod_icls_state
  implicit none
  private
  public post2d
contains
  subroutine post2d(istep)
    integer, intent(in) :: istep
    character(len=7) :: filenumber
    write(filenumber, '(i7.7)') istep
    call icls_check(icls_post2d('data/yz_mid'//filenumber//'.dat'//c_null_char), 'post2d')
  end subroutine post2d
end module mod_output

! loadflds (src/loadd.f90:16-108), writing direction (in = 1): datadir//'fld'//nr(i7.7).
! Reading back is init_program's restart branch (icls_init_program_restart).
module mod_loadd
  use mod_icls_state
  implicit none
  private
  public loadflds
contains
  subroutine loadflds(in, nr)
    integer, intent(in) :: in, nr
    character(len=7) :: istepchar
    if (in /= 1) then
       write(6, *) 'loadflds: reading is done by init_program (begin > 0)'
       error stop 2
    end if
    write(istepchar, '(i7.7)') nr
    call icls_check(icls_write_fld('data/fld'//istepchar//c_null_char, int(nr, c_int32_t)), 'loadflds')
  end subroutine loadflds
end module mod_loadd
