!
! boxzy_driver -- the BOXZYHYDRO main program with the hydro step on the GPU.
!
! Keeps the reference's namelist (the six groups of input.f90:7-16, same keys
! and defaults as parameters.f90) and the call sequence of the main loop
! (main.f90:196-555); every hydro subroutine is the boxzy_hip wrapper of the
! C-ABI entry point that replaces it.
!
! Usage:  boxzy_driver <namelist> [ic_file|-] [out_prefix] [max_steps] [mode]
!   ic_file  stream-binary float64 cons(5,ngrid) u(3,ngrid) p(ngrid)
!            adindx(ngrid); "-" or absent: the reference's shock-tube set-up
!            (init_conditions.f90:26-77, split at y=0)
!   mode     calls (default): one C-ABI call per reference subroutine
!            fused: whole steps through bx_step (two fused kernels per step)
! Writes <out_prefix>.dt (dt, time per step) and <out_prefix>.final
! (cons, u, p), the layout of the reference driver in oracle/.
!
module boxzy_namelist
  implicit none
  integer, parameter :: pre = 8
  ! &io_control
  logical :: write_ascii = .true.
  ! &grid_params
  integer :: nx = 50, ny = 50, nz = 50
  real(pre) :: dx = 0.25d0, dy = 0.25d0, dz = 0.25d0, yoffset = 0d0
  logical :: fluxangmom = .false.
  logical :: no_outflow_xl = .true., no_outflow_xr = .true., no_outflow_yl = .true.
  logical :: no_outflow_yr = .true., no_outflow_zl = .true., no_outflow_zr = .true.
  ! &hydro
  real(pre) :: small_rho = 1d-10, small_eps = 1d-40, vlimit = 1d1, endtime = 0d0
  real(pre) :: den_change_tol = 0.1d0, avmagx = 0d0, avmagy = 0d0, avmagz = 0d0
  real(pre) :: dtout = 10d0, starttime = 0d0, cfl = 0.1d0
  integer :: irestart = 0
  ! &gravity
  integer :: yml_max = 10, nrad_yml = 100, anchor_space = 10, miniter = 10
  real(pre) :: grav_err_tol = 1e-5, grav_err_tol_low = 1e-5
  ! &eos_input
  real(pre) :: mu_z = 16.78d0, brot = 85.4d0, vib = 5987d0, dTk_eos = 5d0, tk_bgrnd = 1d-6
  real(pre) :: ac = 1d0, bc = 3d0, gammafix = 1.66666666667d0
  real(pre) :: xabun = 0.73d0, yabun = 0.27d0, zabun = 0.0d0
  integer :: NEOS_RHO = 1000, H2STAT = 0
  ! &particle_input
  integer :: npart = 100000, npart_direct = 0
  real(pre) :: tg_immediate_couple = 1450d0, a_sublimate_limit = 6.67e-20
  logical :: use_pic = .true., initialize_particles_now = .true.

  namelist /io_control/ write_ascii
  namelist /grid_params/ nx, ny, nz, dx, dy, dz, yoffset, fluxangmom, &
       no_outflow_xl, no_outflow_xr, no_outflow_yl, no_outflow_yr, &
       no_outflow_zl, no_outflow_zr
  namelist /hydro/ small_rho, small_eps, vlimit, endtime, den_change_tol, &
       avmagx, avmagy, avmagz, dtout, starttime, irestart, cfl
  namelist /gravity/ yml_max, nrad_yml, anchor_space, grav_err_tol, miniter, grav_err_tol_low
  namelist /eos_input/ mu_z, brot, vib, dTk_eos, tk_bgrnd, ac, bc, gammafix, &
       xabun, yabun, zabun, NEOS_RHO, H2STAT
  namelist /particle_input/ npart, npart_direct, tg_immediate_couple, a_sublimate_limit, &
       use_pic, initialize_particles_now

contains

  ! input.f90:21-36: every group is mandatory, each read after a rewind
  subroutine read_params(filename)
    character(*), intent(in) :: filename
    open(102, file=trim(filename), status='old')
    read(102, nml=io_control)
    rewind(102)
    read(102, nml=grid_params)
    rewind(102)
    read(102, nml=hydro)
    rewind(102)
    read(102, nml=gravity)
    rewind(102)
    read(102, nml=eos_input)
    rewind(102)
    read(102, nml=particle_input)
    close(102)
  end subroutine

end module boxzy_namelist

program boxzy_driver
  use iso_c_binding
  use boxzy_namelist
  use boxzy_hip
  implicit none

  character(len=256) :: nmlfile, icfile, prefix, arg, mode
  integer :: maxsteps, step, n, ix, iy, iz, igrid
  real(c_double) :: f1, dtold, tphase, yc, nextstop
  real(c_double), allocatable :: cons(:,:), u(:,:), p(:), adindx(:), dts(:)
  type(bx_params) :: prm
  type(bx_eos_params) :: eosprm
  integer(c_int32_t) :: st, taken
  logical :: frame

  call get_command_argument(1, nmlfile)
  icfile = '-'
  prefix = 'boxzy'
  maxsteps = 1000000
  mode = 'calls'
  if (command_argument_count() >= 2) call get_command_argument(2, icfile)
  if (command_argument_count() >= 3) call get_command_argument(3, prefix)
  if (command_argument_count() >= 4) then
     call get_command_argument(4, arg)
     read(arg, *) maxsteps
  end if
  if (command_argument_count() >= 5) call get_command_argument(5, mode)

  call read_params(trim(nmlfile))

  ! init_grid.f90 counterpart: the context owns the device arrays and tables
  prm%nx = nx; prm%ny = ny; prm%nz = nz
  prm%dx = dx; prm%dy = dy; prm%dz = dz
  prm%gammafix = gammafix
  prm%small_rho = small_rho; prm%small_eps = small_eps; prm%vlimit = vlimit; prm%cfl = cfl
  prm%xabun = xabun; prm%yabun = yabun; prm%zabun = zabun; prm%mu_z = mu_z
  prm%avmag = [avmagx, avmagy, avmagz]
  prm%no_outflow = [merge(1, 0, no_outflow_xl), merge(1, 0, no_outflow_xr), &
       merge(1, 0, no_outflow_yl), merge(1, 0, no_outflow_yr), &
       merge(1, 0, no_outflow_zl), merge(1, 0, no_outflow_zr)]
  prm%precision = BX_PREC_FP64
  prm%fluxangmom = merge(1, 0, fluxangmom)       ! main.f90:343-369 flux_ang, source.f90:73-90
  prm%yoffset = yoffset
  call bx_check(bx_create(prm, 0_c_int32_t, bxctx), 'create')
  if (H2STAT >= 0) then
     ! main.f90:70-71: initialize_eos + calc_eos_table (tables kept on the device)
     call bx_eos_defaults(eosprm)
     eosprm%H2STAT = H2STAT; eosprm%NEOS_RHO = NEOS_RHO
     eosprm%dTk_eos = dTk_eos; eosprm%tk_bgrnd = tk_bgrnd; eosprm%brot = brot; eosprm%vib = vib
     eosprm%ac = ac; eosprm%bc = bc
     call bx_check(bx_eos_init(bxctx, eosprm), 'eos_init')
  end if

  n = nx * ny * nz
  allocate(cons(5, n), u(3, n), p(n), adindx(n))
  if (irestart > 0) then
     ! main.f90:83-84: restart from celldump.<irestart> (read_hydro.f90)
     call read_hydro(irestart, step)
     starttime = time
  else if (trim(icfile) == '-') then
     ! init_conditions.f90:26-77: shock tube split at y=0, then the outer
     ! ghost layer gets small_rho/small_eps with p and u left at the IC
     igrid = 0
     do iz = 1, nz
        do iy = 1, ny
           yc = dy * (dble(iy) - (dble(ny) + 1d0) / 2d0)
           do ix = 1, nx
              igrid = igrid + 1
              if (yc < 0d0) then
                 cons(1, igrid) = 1d0
                 p(igrid) = 1d0
              else
                 cons(1, igrid) = 0.125
                 p(igrid) = 1d0 / 10d0
              end if
              adindx(igrid) = gammafix
              cons(5, igrid) = p(igrid) / (adindx(igrid) - 1d0)
              cons(2:4, igrid) = 0d0
              u(:, igrid) = 0d0
              frame = ix == 1 .or. ix == nx .or. iy == 1 .or. iy == ny
              if (nz >= 6) frame = frame .or. iz == 1 .or. iz == nz
              if (frame) then
                 cons(1, igrid) = small_rho
                 cons(2:4, igrid) = 0d0
                 cons(5, igrid) = small_eps
              end if
           end do
        end do
     end do
  else
     open(unit=201, file=trim(icfile), access='stream', form='unformatted', status='old')
     read(201) cons
     read(201) u
     read(201) p
     read(201) adindx
     close(201)
  end if
  if (irestart <= 0) call bx_check(bx_upload(bxctx, cons, u, p, adindx), 'upload')

  open(unit=202, file=trim(prefix)//'.dt', access='stream', form='unformatted', status='replace')
  time = starttime
  nextstop = starttime                            ! main.f90:102-104
  if (irestart > 0) nextstop = nextstop + dtout
  f1 = 1d0
  dtold = 0d0
  step = 0
  if (irestart > 0) step = irestart               ! main.f90:199

  if (trim(mode) == 'fused') then
     ! the whole step in one call per step (bx_step), output on main.f90:548-553's cadence
     allocate(dts(1))
     st = step
     do
        if (time >= endtime .or. st >= maxsteps .or. st >= 1000000) exit
        call bx_check(bx_step_c(bxctx, BX_STEP_FUSED, 1_c_int32_t, endtime, time, dtold, st, dts, taken), 'step')
        if (taken /= 1) exit
        write(202) dts(1), time
        if (time >= nextstop .or. time == endtime) then
           call write_files(int(st), write_ascii)
           nextstop = nextstop + dtout
        end if
     end do
  else
     do  ! main.f90:211-555
        if (time >= endtime .or. step >= maxsteps) exit
        step = step + 1
        call state()
        call courant()
        dt = dt * f1
        if (step > 1 .and. dtold > 0d0) dt = min(dt, dtold * 1.1d0)
        dtold = dt
        if (dt + time > endtime) dt = endtime - time
        dt = dt * 0.5d0
        call source(time, 0)
        call set_ghost_cells()
        call state()
        call velocity()
        call copy_old()
        dt = dt * 2d0
        call flux(0)
        call set_ghost_cells()
        call state()
        call velocity()
        call flux(1)
        call set_ghost_cells()
        call state()
        call velocity()
        dt = 0.5d0 * dt
        tphase = time + dt
        call source(tphase, 1)
        call set_ghost_cells()
        call state()
        call velocity()
        dt = dt * 2d0
        time = time + dt
        write(202) dt, time
        if (time >= nextstop .or. time == endtime) then  ! main.f90:548-553
           call write_files(step, write_ascii)
           nextstop = nextstop + dtout
        end if
     end do
  end if
  close(202)

  call bx_check(bx_download(bxctx, cons, u, p, adindx), 'download')
  open(unit=205, file=trim(prefix)//'.final', access='stream', form='unformatted', status='replace')
  write(205) cons
  write(205) u
  write(205) p
  close(205)
  call bx_check(bx_destroy(bxctx), 'destroy')
end program boxzy_driver
