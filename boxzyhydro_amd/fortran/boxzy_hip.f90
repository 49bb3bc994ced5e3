!
! boxzy_hip -- ISO_C_BINDING interface to libboxzy_hip.so (include/boxzy.h),
! plus the reference-shaped wrappers a Fortran driver calls.
!
! The reference's hot path is argument-less subroutines acting on module
! globals (grid_commons.f90:7-36).  This module keeps that shape: the device
! context and the scalars dt and time are module variables, and state(),
! velocity(), courant(), source(tphase,siter), set_ghost_cells(), flux(avg)
! and avisc() have the reference's names and argument meaning.  Each wrapper
! must be called outside OpenMP regions; it stops with the library's message
! on an error.
!
module boxzy_hip
  use iso_c_binding
  implicit none

  integer, parameter :: BX_PREC_FP64 = 0, BX_PREC_MIXED = 1
  integer, parameter :: BX_STEP_EXACT = 0, BX_STEP_FUSED = 1, BX_STEP_GRAPH = 2

  type, bind(C) :: bx_params
     integer(c_int32_t) :: nx, ny, nz
     real(c_double) :: dx, dy, dz
     real(c_double) :: gammafix
     real(c_double) :: small_rho, small_eps, vlimit, cfl
     real(c_double) :: xabun, yabun, zabun, mu_z
     real(c_double) :: avmag(3)
     integer(c_int32_t) :: no_outflow(6)
     integer(c_int32_t) :: precision
     integer(c_int32_t) :: fluxangmom
     real(c_double) :: yoffset
  end type bx_params

  ! &eos_input keys and the table bounds of parameters.f90:44-58
  type, bind(C) :: bx_eos_params
     integer(c_int32_t) :: H2STAT, NEOS_RHO
     real(c_double) :: dTk_eos, tk_bgrnd, brot, vib, ac, bc
     real(c_double) :: diss, ionize, rho_eos_low, rho_eos_high, tk_eos_cutoff
  end type bx_eos_params

  interface
     function bx_last_error() bind(C, name='bx_last_error') result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function
     integer(c_int) function bx_create(prm, device, ctx) bind(C, name='bx_create')
       import :: c_int, c_int32_t, c_ptr, bx_params
       type(bx_params), intent(in) :: prm
       integer(c_int32_t), value :: device
       type(c_ptr), intent(out) :: ctx
     end function
     integer(c_int) function bx_destroy(ctx) bind(C, name='bx_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function bx_upload(ctx, cons, u, p, adindx) bind(C, name='bx_upload')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: cons(*), u(*), p(*), adindx(*)
     end function
     integer(c_int) function bx_download(ctx, cons, u, p, adindx) bind(C, name='bx_download')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: cons(*), u(*), p(*), adindx(*)
     end function
     integer(c_int) function bx_state_c(ctx) bind(C, name='bx_state')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function bx_velocity_c(ctx, maxv) bind(C, name='bx_velocity')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: maxv
     end function
     integer(c_int) function bx_courant_c(ctx, dt) bind(C, name='bx_courant')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: dt
     end function
     integer(c_int) function bx_source_c(ctx, tphase, siter, dt) bind(C, name='bx_source')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: tphase, dt
       integer(c_int32_t), value :: siter
     end function
     integer(c_int) function bx_set_ghost_cells_c(ctx) bind(C, name='bx_set_ghost_cells')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function bx_write_celldump_c(ctx, path, time, step, ascii) bind(C, name='bx_write_celldump')
       import :: c_int, c_int32_t, c_ptr, c_double, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
       real(c_double), value :: time
       integer(c_int32_t), value :: step, ascii
     end function
     subroutine bx_eos_defaults(e) bind(C, name='bx_eos_defaults')
       import :: bx_eos_params
       type(bx_eos_params), intent(out) :: e
     end subroutine
     integer(c_int) function bx_eos_init(ctx, e) bind(C, name='bx_eos_init')
       import :: c_int, c_ptr, bx_eos_params
       type(c_ptr), value :: ctx
       type(bx_eos_params), intent(in) :: e
     end function
     integer(c_int) function bx_download_muc(ctx, muc_array) bind(C, name='bx_download_muc')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: muc_array(*)
     end function
     integer(c_int) function bx_read_celldump_c(ctx, path, time, step) bind(C, name='bx_read_celldump')
       import :: c_int, c_int32_t, c_ptr, c_double, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
       real(c_double), intent(out) :: time
       integer(c_int32_t), intent(out) :: step
     end function
     integer(c_int) function bx_flux_c(ctx, avg, dt) bind(C, name='bx_flux')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: avg
       real(c_double), value :: dt
     end function
     integer(c_int) function bx_copy_old_c(ctx) bind(C, name='bx_copy_old')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function bx_avisc_c(ctx, qq) bind(C, name='bx_avisc')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: qq(*)
     end function
     integer(c_int) function bx_step_c(ctx, mode, nsteps, endtime, time, dtold, step, dts, taken) &
          bind(C, name='bx_step')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: mode, nsteps
       real(c_double), value :: endtime
       real(c_double), intent(inout) :: time, dtold
       integer(c_int32_t), intent(inout) :: step
       real(c_double), intent(out) :: dts(*)
       integer(c_int32_t), intent(out) :: taken
     end function
     integer(c_int) function bx_sync(ctx) bind(C, name='bx_sync')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function bx_gravity_config(ctx, flags, object_mass, object_radius) &
          bind(C, name='bx_gravity_config')
       import :: c_int, c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: flags
       real(c_double), value :: object_mass, object_radius
     end function
     integer(c_int) function bx_set_phi(ctx, phi) bind(C, name='bx_set_phi')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: phi(*)
     end function
     integer(c_int) function bx_calc_gforce_c(ctx) bind(C, name='bx_calc_gforce')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function bx_external_grav_c(ctx, com) bind(C, name='bx_external_grav')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: com(3)
     end function
  end interface

  integer, parameter :: BX_GRAV_POTENTIAL = 1, BX_GRAV_EXTERNAL = 2, BX_GRAV_COM_SEQUENTIAL = 4

  ! the device counterpart of grid_commons: context plus the step scalars
  type(c_ptr), save :: bxctx = c_null_ptr
  real(c_double), save :: dt = 0d0, time = 0d0
  real(c_double), save :: maxv = 0d0

contains

  subroutine bx_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == 0) return
    call c_f_pointer(bx_last_error(), msg, [512])
    text = ' '
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       text(i:i) = msg(i)
    end do
    write(*,'(A,A,A,I0,A,A)') 'boxzy_hip: ', what, ' failed (', rc, '): ', trim(text)
    stop 2
  end subroutine

  subroutine state()             ! state.f90:6-89
    call bx_check(bx_state_c(bxctx), 'state')
  end subroutine

  subroutine velocity()          ! velocity.f90:4-64
    call bx_check(bx_velocity_c(bxctx, maxv), 'velocity')
  end subroutine

  subroutine courant()           ! courant.f90:4-39: sets dt = cfl/max rate
    call bx_check(bx_courant_c(bxctx, dt), 'courant')
  end subroutine

  subroutine source(tphase, siter)   ! source.f90:5-205 (uses dt)
    real(c_double), intent(in) :: tphase
    integer, intent(in) :: siter
    call bx_check(bx_source_c(bxctx, tphase, int(siter, c_int32_t), dt), 'source')
  end subroutine

  subroutine set_ghost_cells()   ! utils.f90:305-480
    call bx_check(bx_set_ghost_cells_c(bxctx), 'set_ghost_cells')
  end subroutine

  subroutine flux(avg)           ! flux.f90:6-235 (uses dt)
    integer, intent(in) :: avg
    call bx_check(bx_flux_c(bxctx, int(avg, c_int32_t), dt), 'flux')
  end subroutine

  subroutine copy_old()          ! main.f90:327-333
    call bx_check(bx_copy_old_c(bxctx), 'copy_old')
  end subroutine

  ! write_files.f90:5-71 (hydro part): state(), then celldump.<step> (binary,
  ! or formatted when ascii); time is the module variable
  ! utils.f90:549-571 on the device (gravity configured with bx_gravity_config
  ! and the potential given with bx_set_phi)
  subroutine calc_gforce()
    call bx_check(bx_calc_gforce_c(bxctx), 'calc_gforce')
  end subroutine

  ! gravity.f90:1477-1549; object_mass as given to bx_gravity_config
  subroutine external_grav(tphase, printcom, object_mass)
    real(c_double), intent(in) :: tphase, object_mass
    integer, intent(in) :: printcom
    real(c_double) :: com(3)
    call bx_check(bx_external_grav_c(bxctx, com), 'external_grav')
    if (printcom == 1) print *, "STAR ", tphase, -com(1) / object_mass, -com(2) / object_mass, &
         -com(3) / object_mass
  end subroutine

  subroutine write_files(step, ascii)
    integer, intent(in) :: step
    logical, intent(in) :: ascii
    character(len=17) :: name
    write(name, '(A,I8.8)') 'celldump.', step
    call bx_check(bx_write_celldump_c(bxctx, trim(name)//c_null_char, time, int(step, c_int32_t), &
         merge(1_c_int32_t, 0_c_int32_t, ascii)), 'write_files')
  end subroutine

  ! read_hydro.f90:4-183: restart from celldump.<irestart>; sets time and returns the file's step
  subroutine read_hydro(irestart, step)
    integer, intent(in) :: irestart
    integer, intent(out) :: step
    character(len=17) :: name
    integer(c_int32_t) :: s
    write(name, '(A,I8.8)') 'celldump.', irestart
    call bx_check(bx_read_celldump_c(bxctx, trim(name)//c_null_char, time, s), 'read_hydro')
    step = s
  end subroutine

end module boxzy_hip
