! icls_main.f90 -- Fortran main program driving the GPU step in the order of the
! reference's time loop (src/main.f90:103-291).  The run is configured by the
! namelist /icls/ in the file named as the first command-line argument (the
! reference compiles these in, src/param.f90).  Optional: out_file names a raw
! stream file receiving unew, vnew, wnew, pnew (0:i1,0:j1,0:k1) and lvset(:,:,:,1)
! (-2:i1+2,...) in Fortran order after the last step; iout2d, iout3d and ioutfld
! (param.f90:86-87) write yz_mid, para .vtk and fld files into data/ (0 = off).
program ICLS
  use iso_c_binding
  use icls_capi
  use mod_icls_state
  use mod_init
  use mod_levelset
  use mod_misc
  use mod_ab
  use mod_fillps
  use mod_gfm
  use mod_solver
  use mod_bound
  use mod_correc
  use mod_chkdiv
  use mod_chkdt
  use mod_icls_main_ops
  use mod_write_vtk
  use mod_output
  use mod_loadd
  implicit none

  integer :: itot = 64, jtot = 64, ktot = 64, nstep = 100, begin = 0, ioutchk = 20, ndrops = 1
  real :: lx = 1., ly = 1., lz = 1., timestep = 1e-4, rho1 = 1., rho2 = 10., miu1 = 1., miu2 = 10.
  real :: Re = 10., We = 1., iFr = 0., g_unit(3) = (/0., 0., 0./), wall_speed = 1.5
  integer :: ls_adv_space = 0, re_initialization = 1, reinit_step = 1, reinit_iter_max = 1
  integer :: mass_correction = 1, inflate_step = 1, ns_np = 1, constPoi = 0, iniu = 0, mode = 0
  real :: drop_x(ICLS_MAX_DROPS) = 0.5, drop_y(ICLS_MAX_DROPS) = 0.5, drop_z(ICLS_MAX_DROPS) = 0.5
  real :: drop_r(ICLS_MAX_DROPS) = 0.2
  integer :: iout2d = 0, iout3d = 0, ioutfld = 0
  character(len=256) :: out_file = ''
  namelist /icls/ itot, jtot, ktot, nstep, begin, ioutchk, lx, ly, lz, timestep, rho1, rho2, miu1, miu2, &
       Re, We, iFr, g_unit, wall_speed, ls_adv_space, re_initialization, reinit_step, reinit_iter_max, &
       mass_correction, inflate_step, ns_np, constPoi, iniu, mode, ndrops, drop_x, drop_y, drop_z, drop_r, &
       iout2d, iout3d, ioutfld, out_file

  type(icls_params) :: p0
  integer :: istep, u, l
  real :: time, dtmax, v_bulk, vel_max, massloss, vol, vol0
  real, allocatable :: duold(:, :, :), dvold(:, :, :), dwold(:, :, :)
  real, allocatable :: pold(:, :, :), p_jump(:, :, :), curv_restart(:, :, :)
  real, allocatable :: unew(:, :, :), vnew(:, :, :), wnew(:, :, :), ls(:, :, :)
  real, allocatable, target :: p(:, :, :)
  real, pointer :: pp2(:, :, :)
  character(len=256) :: cfg

  if (command_argument_count() >= 1) then
     call get_command_argument(1, cfg)
     open(newunit=u, file=trim(cfg), status='old', action='read')
     read(u, nml=icls)
     close(u)
  end if

  p0%itot = itot; p0%jtot = jtot; p0%ktot = ktot; p0%lmax = ndrops
  p0%lx = lx; p0%ly = ly; p0%lz = lz; p0%timestep = timestep
  p0%rho1 = rho1; p0%rho2 = rho2; p0%miu1 = miu1; p0%miu2 = miu2
  p0%Re = Re; p0%We = We; p0%iFr = iFr; p0%g_unit = g_unit
  p0%ls_adv_space = ls_adv_space; p0%re_initialization = re_initialization
  p0%reinit_step = reinit_step; p0%reinit_iter_max = reinit_iter_max
  p0%mass_correction = mass_correction; p0%inflate_step = inflate_step; p0%ns_np = ns_np
  p0%wall_speed = wall_speed; p0%constPoi = constPoi; p0%iniu = iniu; p0%mode = mode; p0%ioutchk = ioutchk
  p0%cfl = 0.; p0%ls_full = 0; p0%mass_trigger_exp = 0; p0%reserved = 0
  p0%drop_xc = drop_x; p0%drop_yc = drop_y; p0%drop_zc = drop_z; p0%drop_r = drop_r
  call icls_setup(p0)

  allocate(duold(imax, jmax, kmax), dvold(imax, jmax, kmax), dwold(imax, jmax, kmax))
  allocate(pold(0:i1, 0:j1, 0:k1), p_jump(0:i1, 0:j1, 0:k1), curv_restart(0:i1, 0:j1, 0:k1))
  allocate(unew(0:i1, 0:j1, 0:k1), vnew(0:i1, 0:j1, 0:k1), wnew(0:i1, 0:j1, 0:k1), p(0:i1, 0:j1, 0:k1))

  !---- initialisation (main.f90:80-82)
  call init_program(begin, duold, dvold, dwold, pold, curv_restart)
  pp2 => p(1:imax, 1:jmax, 1:kmax)
  ! 0 on a fresh start; on a restart the time read from data/fld<begin> (loadd.f90:61)
  call icls_check(icls_get_scalar(S_TIME, time), 'time')

  !---- main loop (main.f90:103-291)
  do istep = begin + 1, nstep
     time = time + timestep
     call icls_check(icls_set_scalar(S_TIME, time), 'time')   ! common time (main.f90:107)

     call ls_main(istep)

     call average_stream_vel(v_bulk)
     call ab2(duold, dvold, dwold, pold)

     call fillps(p)
     call gfm4p(p_jump)
     call add_pjump(p, p_jump)

     call solver2d(pp2)
     call boundp(p)

     call correc(p)
     call bounduvw(unew, vnew, wnew)

     call update_pressure(istep == begin + 1)
     call average_vel_at_center

     call check_nan()

     if (mod(istep, ioutchk) == 0) then
        call chkdiv(istep)
        call chkdt(dtmax)
        write(6, '(A,1ES12.4)') '! dt calculated by chkdt =', dtmax
        call max_vel_at_center(vel_max)
        massloss = 0.
        do l = 1, ndrops
           call icls_check(icls_get_volume(int(l - 1, c_int32_t), vol, vol0), 'volume')
           massloss = massloss + (1. - vol / vol0)
        end do
        write(6, '(A,I8,A,1ES12.4,A,1ES12.4,A,1ES12.4)') '! istep', istep, '  time =', time, &
             '  massloss =', massloss, '  max |u_c| =', vel_max
     end if

     !---- output files (main.f90:262-277); 0 switches a kind off
     if (iout2d > 0) then
        if (mod(istep, iout2d) == 0) call post2d(istep)
     end if
     if (iout3d > 0) then
        if (mod(istep, iout3d) == 0) call write_vtk(istep)
     end if
     if (ioutfld > 0) then
        if (mod(istep, ioutfld) == 0) call loadflds(1, istep)
     end if
  end do

  if (len_trim(out_file) > 0) then
     allocate(ls(-2:i1 + 2, -2:j1 + 2, -2:k1 + 2))
     call pull_c(F_UNEW, unew); call pull_c(F_VNEW, vnew); call pull_c(F_WNEW, wnew); call pull_c(F_PNEW, p)
     call icls_check(icls_download(F_LVSET, 0_c_int32_t, ls), 'download lvset')
     open(newunit=u, file=trim(out_file), access='stream', form='unformatted', status='replace')
     write(u) unew, vnew, wnew, p, ls
     close(u)
  end if

  call icls_check(icls_finalize(), 'icls_finalize')
  write(6, '(A)') '************** Simulation completed. **************'
end program ICLS
