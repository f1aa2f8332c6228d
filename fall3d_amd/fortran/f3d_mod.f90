!-----------------------------------------------------------------------
! ISO_C_BINDING interfaces to libfall3d_amd (include/fall3d_amd.h).
!
! Two groups:
!  * f3d_* : the device-resident context.  A FALL3D driver keeps its
!    control flow (Sources/Fall3d.f90:79-121) and replaces the calls
!    meteo -> f3d_set_meteo, setsrc+sizloop -> f3d_step,
!    getc0+cmass3d (endstep/writim) -> f3d_budget.
!  * the per-routine externals (setbcc, advctx, ...) need no interface:
!    the library exports them with the reference's own implicit-interface
!    ABI, so the reference's sizloop.f90 links against them unchanged.
!-----------------------------------------------------------------------
module fall3d_amd
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: F3D_FP64 = 0, F3D_FP32 = 1
  integer(c_int), parameter, public :: F3D_OK = 0

  type, bind(C), public :: f3d_grid
     integer(c_int) :: nx, ny, nz, nc
     real(c_double) :: dx, dy
     type(c_ptr)    :: zlayer = c_null_ptr
     type(c_ptr)    :: hm1 = c_null_ptr
     integer(c_int) :: lonlat = 0
  end type f3d_grid

  public :: f3d_create, f3d_destroy, f3d_set_classes, f3d_set_models, f3d_set_meteo, &
            f3d_get_dt, f3d_set_source, f3d_step, f3d_budget, f3d_get_state, f3d_sync, &
            f3d_get_clock, f3d_error_message
  ! batching and tuning (no counterpart in the reference)
  public :: f3d_batch_create, f3d_batch_destroy, f3d_batch_step, f3d_batch_sync, &
            f3d_batch_autotune, f3d_autotune

  interface
     integer(c_int) function f3d_create(ctx, grid, precision) bind(C, name='f3d_create')
       import :: c_int, c_ptr, f3d_grid
       type(c_ptr), intent(out) :: ctx
       type(f3d_grid), intent(in) :: grid
       integer(c_int), value :: precision
     end function f3d_create

     subroutine f3d_destroy(ctx) bind(C, name='f3d_destroy')
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine f3d_destroy

     integer(c_int) function f3d_set_classes(ctx, diam, rhop, sphe, modv) &
          bind(C, name='f3d_set_classes')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: diam(*), rhop(*), sphe(*)
       integer(c_int), value :: modv
     end function f3d_set_classes

     integer(c_int) function f3d_set_models(ctx, modkv, modkh, rkv0, rkh0, csh) &
          bind(C, name='f3d_set_models')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: modkv, modkh
       real(c_double), value :: rkv0, rkh0, csh
     end function f3d_set_models

     integer(c_int) function f3d_set_meteo(ctx, vx, vy, vz, tempe, vptemp, rho, ustar, zi, rl, &
          vdep) bind(C, name='f3d_set_meteo')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: vx(*), vy(*), vz(*), tempe(*), vptemp(*), rho(*)
       real(c_double), intent(in) :: ustar(*), zi(*), rl(*), vdep(*)
     end function f3d_set_meteo

     ! LDU and znt (Master.f90:96,108): later meteo reads run setvdrydep
     ! (meteo.f90:99-100) on the device
     integer(c_int) function f3d_set_landuse(ctx, ldu, znt) bind(C, name='f3d_set_landuse')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: ldu(*), znt(*)
     end function f3d_set_landuse

     ! topg (Master): LON-LAT grids run corvver (meteo.f90:116-124)
     integer(c_int) function f3d_set_topography(ctx, topg) bind(C, name='f3d_set_topography')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: topg(*)
     end function f3d_set_topography

     ! prate (Master, mm/h): wet deposition in later f3d_step calls (not in batches)
     integer(c_int) function f3d_set_precipitation(ctx, prate) bind(C, name='f3d_set_precipitation')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: prate(*)
     end function f3d_set_precipitation

     ! printres_nc products: load(nx,ny,0:nc), thick(nx,ny), colmass(nx,ny)
     integer(c_int) function f3d_output_products(ctx, member, load, thick, colmass) &
          bind(C, name='f3d_output_products')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: member
       real(c_double), intent(out) :: load(*), thick(*), colmass(*)
     end function f3d_output_products

     integer(c_int) function f3d_get_vdep(ctx, vdep) bind(C, name='f3d_get_vdep')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: vdep(*)
     end function f3d_get_vdep

     integer(c_int) function f3d_get_dt(ctx, dt, dtc) bind(C, name='f3d_get_dt')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: dt, dtc
     end function f3d_get_dt

     integer(c_int) function f3d_set_source(ctx, nsou, isrc, jsrc, ksrc, tmrat, flowrate) &
          bind(C, name='f3d_set_source')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsou
       integer(c_int), intent(in) :: isrc(*), jsrc(*), ksrc(*)
       real(c_double), intent(in) :: tmrat(*)
       real(c_double), value :: flowrate
     end function f3d_set_source

     integer(c_int) function f3d_step(ctx, nsteps) bind(C, name='f3d_step')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
     end function f3d_step

     integer(c_int) function f3d_budget(ctx, out) bind(C, name='f3d_budget')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out(6)
     end function f3d_budget

     integer(c_int) function f3d_get_state(ctx, c, cdry) bind(C, name='f3d_get_state')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: c(*), cdry(*)
     end function f3d_get_state

     integer(c_int) function f3d_sync(ctx) bind(C, name='f3d_sync')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function f3d_sync

     integer(c_int) function f3d_get_clock(ctx, time, iiter, erumass, outmass) &
          bind(C, name='f3d_get_clock')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: time, erumass, outmass
       integer(c_int), intent(out) :: iiter
     end function f3d_get_clock

     ! several contexts stepped together (ctxs: array of n context handles)
     integer(c_int) function f3d_batch_create(batch, ctxs, n) bind(C, name='f3d_batch_create')
       import :: c_int, c_ptr
       type(c_ptr), intent(out) :: batch
       type(c_ptr), intent(in) :: ctxs(*)
       integer(c_int), value :: n
     end function f3d_batch_create

     subroutine f3d_batch_destroy(batch) bind(C, name='f3d_batch_destroy')
       import :: c_ptr
       type(c_ptr), value :: batch
     end subroutine f3d_batch_destroy

     integer(c_int) function f3d_batch_step(batch, nsteps) bind(C, name='f3d_batch_step')
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       integer(c_int), value :: nsteps
     end function f3d_batch_step

     integer(c_int) function f3d_batch_sync(batch) bind(C, name='f3d_batch_sync')
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
     end function f3d_batch_sync

     integer(c_int) function f3d_batch_autotune(batch, trials, tx, ty, slabs, best_ms) &
          bind(C, name='f3d_batch_autotune')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: batch
       integer(c_int), value :: trials
       integer(c_int), intent(out) :: tx, ty, slabs
       real(c_double), intent(out) :: best_ms
     end function f3d_batch_autotune

     integer(c_int) function f3d_autotune(ctx, trials, tx, ty, slabs, best_ms) &
          bind(C, name='f3d_autotune')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: trials
       integer(c_int), intent(out) :: tx, ty, slabs
       real(c_double), intent(out) :: best_ms
     end function f3d_autotune

     type(c_ptr) function f3d_last_error_c() bind(C, name='f3d_last_error')
       import :: c_ptr
     end function f3d_last_error_c
  end interface

contains

  ! The library's last error message as a Fortran string
  function f3d_error_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: i
    msg = ' '
    call c_f_pointer(f3d_last_error_c(), p, [512])
    do i = 1, 512
       if (p(i) == c_null_char) exit
       msg(i:i) = p(i)
    end do
  end function f3d_error_message

end module fall3d_amd
