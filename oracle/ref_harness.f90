!
! Test-infrastructure driver for the reference BOXZYHYDRO hydro step.
!
! This program is part of the oracle: it links the reference's own Fortran
! subroutines (state, velocity, courant, source, flux, set_ghost_cells,
! init_grid, the eos and units setup) compiled from their sources, and runs the
! main-loop body of main.f90:211-555 with every optional-physics flag off.
! It is never shipped and never called by the product path; its outputs are
! golden vectors for the C restatement (boxzy_oracle.c) and the HIP kernels.
!
! Usage:  ref_harness <namelist> <ic_file> <out_prefix> <max_steps> <mode> [diag] [dump] [grav]
!   ic_file   stream-binary float64: cons(5,ngrid) u(3,ngrid) p(ngrid) adindx(ngrid)
!             (Fortran order, exactly the reference's grid_commons arrays)
!   mode 0    run steps; write <prefix>.dt (dt, time per step),
!             <prefix>.final (cons, u, p after the last step) and
!             <prefix>.timing (wall seconds of the step loop, steps, threads)
!   mode 1    run steps; additionally write <prefix>.snap with the state after
!             every sub-call of the first step (tagged records)
!   mode 2    restart: read_hydro() reads celldump.<max_steps> (irestart =
!             max_steps) instead of ic_file; write <prefix>.final and stop
!   mode 3    write the EOS tables calc_eos_table() leaves to <prefix>.eos and
!             stop: NEOS_T, NEOS_RHO (int32); log_rho_eos_low, drho_eos (f64);
!             rho_table, deng_eos_array (NEOS_RHO); tk_table (NEOS_T);
!             eng_table, gamma_table, muc_table, eng_table2, tk_table2,
!             muc_table2, gamma_table2 (NEOS_T x NEOS_RHO, Fortran order)
!   mode 4    one state() call on ic_file; write <prefix>.final and stop
!   every run also writes <prefix>.eos_state: adindx, muc_array (ngrid each)
!   diag 1    also run the per-step conservation diagnostics (main.f90:222-223)
!   dump 1|2  after the run, write_files(step) with write_ascii .false.|.true.
!             (celldump.<step> in the working directory)
!   grav 1    the potential phi(ngrid) is a fifth record of ic_file; gforce
!             from calc_gforce() before the loop and before source(tphase,1)
!             of every step (main.f90:146-148, 479-481, USEPOTENTIAL with a
!             fixed potential: no Poisson solve)
!   grav 2    as 1, and external_grav() after each calc_gforce()
!             (main.f90:152-154, 485-487, EXTERNALGRAV)
!
! avisc (avisc.f90:4-26) is not driven here: for outer-layer ghosts it reads
! grid(0) and u(:,0) (ineigh(2:6)=0 from init_grid.f90:178), outside the
! arrays, so its output there is undefined and the read can fault.
!
! The step sequence below restates main.f90:211-555 with RADTRAN, COOL,
! SELFGRAVITY, PARTICLE, EXTERNALGRAV, ENERGYCOR off; flux or flux_ang as
! the namelist's fluxangmom selects (main.f90:343-369).
!
program ref_harness
 use parameters
 use derived_types
 use grid_commons
 use eos
 use input
 use utils
 use selfgravity, only: external_grav
 use omp_lib
 implicit none

 character(len=256)::nmlfile,icfile,prefix,arg
 character(len=80)::nml80
 integer::maxsteps,mode,diag,dump,grav,step,igrid,idim,ios
 real(pre)::f1,dtold,tphase
 real(pre)::etot_loc0,mass_loc0,f_loc0
 real(8)::wt0,wt1
 type(units)::scale

 call get_command_argument(1,nmlfile)
 call get_command_argument(2,icfile)
 call get_command_argument(3,prefix)
 call get_command_argument(4,arg); read(arg,*)maxsteps
 call get_command_argument(5,arg); read(arg,*)mode
 diag=0
 if(command_argument_count()>=6)then
   call get_command_argument(6,arg); read(arg,*)diag
 endif
 dump=0
 if(command_argument_count()>=7)then
   call get_command_argument(7,arg); read(arg,*)dump
 endif
 grav=0
 if(command_argument_count()>=8)then
   call get_command_argument(8,arg); read(arg,*)grav
 endif

 nml80=nmlfile(1:80)
 call read_params(nml80)
 call init_grid()
 call initialize_eos
 call calc_eos_table

! initial conditions from file, then the bookkeeping that init_conditions.f90
! performs for every set-up: phi=0, gforce from phi (-> -0.0), cons_old=cons
 if(mode==3)then
   open(unit=207,file=trim(prefix)//'.eos',access='stream',form='unformatted',status='replace')
   write(207)NEOS_T,NEOS_RHO
   write(207)log_rho_eos_low,drho_eos
   write(207)rho_table,deng_eos_array,tk_table
   write(207)eng_table,gamma_table,muc_table
   write(207)eng_table2,tk_table2,muc_table2,gamma_table2
   close(207)
   stop
 endif
 if(mode==2)then
   irestart=maxsteps
   phi=zero
   call read_hydro()
   call write_final()
   stop
 endif
 open(unit=201,file=trim(icfile),access='stream',form='unformatted',status='old',iostat=ios)
 if(ios/=0)stop "cannot open ic file"
 read(201)cons
 read(201)u
 read(201)p
 read(201)adindx
 phi=zero
 if(grav>0)read(201)phi
 close(201)
!$OMP PARALLEL DEFAULT(SHARED)
 call calc_gforce()
 if(grav==2)then
!$OMP BARRIER
   call external_grav(zero,0)
 endif
!$OMP END PARALLEL
 cons_old=cons

 if(mode==4)then
!$OMP PARALLEL DEFAULT(SHARED)
   call state()
!$OMP END PARALLEL
   call write_final()
   stop
 endif

 if(mode==1)then
   open(unit=203,file=trim(prefix)//'.snap',access='stream',form='unformatted',status='replace')
   call snap(0)
 endif

 open(unit=202,file=trim(prefix)//'.dt',access='stream',form='unformatted',status='replace')

 call get_units(scale)
 time=starttime
 f1=one
 dtold=zero
 step=0
 wt0=omp_get_wtime()

 do
  if(time>=endtime.or.step>=maxsteps)exit
  step=step+1
  etot_loc0=zero;mass_loc0=zero;f_loc0=zero

!$OMP PARALLEL DEFAULT(SHARED)
  if(diag==1)then
    call conservation_diagnostic()
    call get_etot(etot_loc0,mass_loc0,f_loc0)
  endif
  call state()
!$OMP BARRIER
  call courant()
!$OMP BARRIER
!$OMP MASTER
  if(mode==1.and.step==1)call snap(1)
  dt=dt*f1
  if(step>1.and.dtold>zero)dt=min(dt,dtold*1.1d0)
  dtold=dt
  if(dt+time>endtime)dt=endtime-time
  dt=dt*half
!$OMP END MASTER
!$OMP BARRIER
  call source(time,0)
!$OMP BARRIER
!$OMP MASTER
  if(mode==1.and.step==1)call snap(2)
!$OMP END MASTER
!$OMP BARRIER
  call set_ghost_cells()
!$OMP BARRIER
  call state()
!$OMP BARRIER
!$OMP MASTER
  if(mode==1.and.step==1)call snap(3)
!$OMP END MASTER
!$OMP BARRIER
  call velocity()
!$OMP BARRIER
!$OMP DO SCHEDULE(STATIC)
  do igrid=1,ngrid
   do idim=1,5
    cons_old(idim,igrid)=cons(idim,igrid)
   enddo
  enddo
!$OMP ENDDO
!$OMP END PARALLEL
  if(mode==1.and.step==1)call snap(4)

  dt=dt*two
  if(fluxangmom)then   ! main.f90:343-347
    call flux_ang(0)
  else
    call flux(0)
  endif
  if(mode==1.and.step==1)call snap(5)

!$OMP PARALLEL DEFAULT(SHARED)
  call set_ghost_cells()
!$OMP BARRIER
  call state()
!$OMP BARRIER
  call velocity()
!$OMP END PARALLEL
  if(mode==1.and.step==1)call snap(6)

  if(fluxangmom)then   ! main.f90:365-369
    call flux_ang(1)
  else
    call flux(1)
  endif
  if(mode==1.and.step==1)call snap(7)

!$OMP PARALLEL DEFAULT(SHARED)
  call set_ghost_cells()
!$OMP BARRIER
  call state()
!$OMP BARRIER
  call velocity()
!$OMP END PARALLEL
  if(mode==1.and.step==1)call snap(8)

  dt=half*dt
!$OMP PARALLEL DEFAULT(SHARED)
  if(grav>0)call calc_gforce()
  tphase=time+dt
  if(grav==2)then
!$OMP BARRIER
    call external_grav(tphase,0)
  endif
!$OMP BARRIER
  call source(tphase,1)
  call set_ghost_cells()
!$OMP BARRIER
  call state()
!$OMP BARRIER
  call velocity()
!$OMP END PARALLEL
  if(mode==1.and.step==1)call snap(9)

  dt=dt*two
  time=time+dt
  write(202)dt,time
 enddo
 wt1=omp_get_wtime()
 close(202)
 if(mode==1)close(203)
 open(unit=206,file=trim(prefix)//'.timing',form='formatted',status='replace')
 write(206,*)wt1-wt0,step,omp_get_max_threads()
 close(206)

 call write_final()

 if(dump>0)then
   write_ascii=(dump==2)
   call write_files(step)
 endif

contains

 subroutine write_final()
  open(unit=205,file=trim(prefix)//'.final',access='stream',form='unformatted',status='replace')
  write(205)cons
  write(205)u
  write(205)p
  close(205)
  open(unit=208,file=trim(prefix)//'.eos_state',access='stream',form='unformatted',status='replace')
  write(208)adindx
  write(208)muc_array
  close(208)
 end subroutine

 subroutine snap(tag)
  integer,intent(in)::tag
  write(203)tag
  write(203)dt
  write(203)cons
  write(203)u
  write(203)p
  write(203)cons_old
 end subroutine

end program
