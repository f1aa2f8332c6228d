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
     real(c_double) :: dlon = 0.0_c_double, dlat = 0.0_c_double
  end type f3d_grid

  ! run date, extent and clock of a restart file (wrirst.f90:120-172)
  type, bind(C), public :: f3d_run_info
     integer(c_int) :: year = 0, month = 0, day = 0, run_start = 0, run_end = 0
     real(c_double) :: xmin = 0.0_c_double, ymin = 0.0_c_double
     real(c_double) :: xmax = 0.0_c_double, ymax = 0.0_c_double
     real(c_double) :: erumass = 0.0_c_double, outmass = 0.0_c_double
  end type f3d_run_info

  ! the GRAVITY_CURRENT block (readat.f90:149-200), in the grid's coordinates
  type, bind(C), public :: f3d_gravity
     real(c_double) :: c_flow_rate = 0.0_c_double, lambda_grav = 0.0_c_double
     real(c_double) :: k_entrain = 0.0_c_double, brunt_vaisala = 0.0_c_double
     real(c_double) :: center_x = 0.0_c_double, center_y = 0.0_c_double
     real(c_double) :: xorigr = 0.0_c_double, yorigr = 0.0_c_double
     real(c_double) :: dx = 0.0_c_double, dy = 0.0_c_double, beg_time = 0.0_c_double
  end type f3d_gravity

  public :: f3d_create, f3d_destroy, f3d_set_classes, f3d_set_models, f3d_set_meteo, &
            f3d_get_dt, f3d_set_dt, f3d_set_source, f3d_step, f3d_budget, f3d_get_state, &
            f3d_set_state, f3d_get_coef, f3d_sync, f3d_get_clock, f3d_set_clock, &
            f3d_get_cd_state, f3d_set_cd_state, f3d_error_message
  ! meteo-stage extensions and output products
  public :: f3d_set_landuse, f3d_set_topography, f3d_set_precipitation, &
            f3d_output_products, f3d_get_vdep, f3d_set_pressure, f3d_output_concentration, &
            f3d_write_restart, f3d_read_restart, f3d_restart_write, f3d_restart_read, &
            f3d_set_gravity_current, f3d_get_gravity_current, f3d_get_divu
  ! ensembles and resident meteorology (no counterpart in the reference)
  public :: f3d_create_ensemble, f3d_set_member_source, f3d_ensemble_budget, &
            f3d_get_member_state, f3d_set_meteo_window
  ! batching, tuning, timing and memory (no counterpart in the reference)
  public :: f3d_batch_create, f3d_batch_destroy, f3d_batch_step, f3d_batch_sync, &
            f3d_batch_autotune, f3d_autotune, f3d_batch_set_tile, f3d_batch_get_tile, &
            f3d_batch_use_graph, f3d_set_tile, f3d_set_tile_config, f3d_timing, &
            f3d_kernel_time, f3d_device_bytes, f3d_mem_info, f3d_hbm_copy, &
            f3d_cmass3d_noalloc, f3d_batch_timing, f3d_batch_kernel_time

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
     ! gc: the parameters, or c_null_ptr to turn the gravity current off
     integer(c_int) function f3d_set_gravity_current(ctx, gc, nsou, zs) &
          bind(C, name='f3d_set_gravity_current')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, gc
       integer(c_int), value :: nsou
       real(c_double), intent(in) :: zs(*)
     end function f3d_set_gravity_current
     integer(c_int) function f3d_get_gravity_current(ctx, diag, corrdiv, active) &
          bind(C, name='f3d_get_gravity_current')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: diag(9), corrdiv
       integer(c_int), intent(out) :: active
     end function f3d_get_gravity_current
     integer(c_int) function f3d_get_divu(ctx, divu) bind(C, name='f3d_get_divu')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: divu(*)
     end function f3d_get_divu
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

     ! pressure p(nx,ny,nz) in Pa for the flight-level cuts
     integer(c_int) function f3d_set_pressure(ctx, p) bind(C, name='f3d_set_pressure')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: p(*)
     end function f3d_set_pressure

     ! printres_nc concentrations: cgrnd(nx,ny), con(nx,ny,nz), cfl(nx,ny,nfl)
     integer(c_int) function f3d_output_concentration(ctx, member, cgrnd, con, asl, nfl, fl, cfl) &
          bind(C, name='f3d_output_concentration')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: member, asl, nfl
       real(c_double), intent(out) :: cgrnd(*), con(*), cfl(*)
       real(c_double), intent(in) :: fl(*)
     end function f3d_output_concentration

     ! wrirst / rearst (path: a NUL-terminated character string)
     integer(c_int) function f3d_write_restart(ctx, member, path, info) &
          bind(C, name='f3d_write_restart')
       import :: c_int, c_ptr, c_char, f3d_run_info
       type(c_ptr), value :: ctx
       integer(c_int), value :: member
       character(kind=c_char), intent(in) :: path(*)
       type(f3d_run_info), intent(in) :: info
     end function f3d_write_restart

     integer(c_int) function f3d_read_restart(ctx, path, cload, info) &
          bind(C, name='f3d_read_restart')
       import :: c_int, c_ptr, c_char, c_double, f3d_run_info
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: path(*)
       real(c_double), intent(out) :: cload(*)
       type(f3d_run_info), intent(out) :: info
     end function f3d_read_restart

     ! the restart file on host arrays c(nx+2,ny+2,nz+2,nc), cload(nx,ny,nc+1)
     integer(c_int) function f3d_restart_write(path, nx, ny, nz, nc, lonlat, c, cload, info) &
          bind(C, name='f3d_restart_write')
       import :: c_int, c_char, c_double, f3d_run_info
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int), value :: nx, ny, nz, nc, lonlat
       real(c_double), intent(in) :: c(*), cload(*)
       type(f3d_run_info), intent(in) :: info
     end function f3d_restart_write

     integer(c_int) function f3d_restart_read(path, nx, ny, nz, nc, c, cload, info) &
          bind(C, name='f3d_restart_read')
       import :: c_int, c_char, c_double, f3d_run_info
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int), value :: nx, ny, nz, nc
       real(c_double), intent(out) :: c(*), cload(*)
       type(f3d_run_info), intent(out) :: info
     end function f3d_restart_read

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

     integer(c_int) function f3d_create_ensemble(ctx, grid, nmembers) &
          bind(C, name='f3d_create_ensemble')
       import :: c_int, c_ptr, f3d_grid
       type(c_ptr), intent(out) :: ctx
       type(f3d_grid), intent(in) :: grid
       integer(c_int), value :: nmembers
     end function f3d_create_ensemble

     ! meteorology from the window (i0+1.., j0+1.., 1..nz) of a region
     ! context's fields in HBM
     integer(c_int) function f3d_set_meteo_window(ctx, region, i0, j0) &
          bind(C, name='f3d_set_meteo_window')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, region
       integer(c_int), value :: i0, j0
     end function f3d_set_meteo_window

     integer(c_int) function f3d_set_dt(ctx, dt) bind(C, name='f3d_set_dt')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt
     end function f3d_set_dt

     integer(c_int) function f3d_set_member_source(ctx, member, nsou, isrc, jsrc, ksrc, &
          tmrat, flowrate) bind(C, name='f3d_set_member_source')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: member, nsou
       integer(c_int), intent(in) :: isrc(*), jsrc(*), ksrc(*)
       real(c_double), intent(in) :: tmrat(*)
       real(c_double), value :: flowrate
     end function f3d_set_member_source

     integer(c_int) function f3d_ensemble_budget(ctx, out) bind(C, name='f3d_ensemble_budget')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out(*)
     end function f3d_ensemble_budget

     integer(c_int) function f3d_get_member_state(ctx, member, c, cdry) &
          bind(C, name='f3d_get_member_state')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: member
       real(c_double), intent(out) :: c(*), cdry(*)
     end function f3d_get_member_state

     ! restart: c (0:nx+1,0:ny+1,0:nz+1,nc) and cdry (nx,ny,0:nc)
     integer(c_int) function f3d_set_state(ctx, c, cdry) bind(C, name='f3d_set_state')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: c(*), cdry(*)
     end function f3d_set_state

     integer(c_int) function f3d_get_coef(ctx, vx, rho, rkhor, rkh1, rkver, vset) &
          bind(C, name='f3d_get_coef')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: vx(*), rho(*), rkhor(*), rkh1(*), rkver(*), vset(*)
     end function f3d_get_coef

     integer(c_int) function f3d_set_clock(ctx, time, iiter, erumass, outmass) &
          bind(C, name='f3d_set_clock')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: time, erumass, outmass
       integer(c_int), value :: iiter
     end function f3d_set_clock

     ! the SAVEd vsettl drag coefficient (setvset.f90:89)
     integer(c_int) function f3d_get_cd_state(ctx, cd) bind(C, name='f3d_get_cd_state')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: cd
     end function f3d_get_cd_state

     integer(c_int) function f3d_set_cd_state(ctx, cd) bind(C, name='f3d_set_cd_state')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: cd
     end function f3d_set_cd_state

     integer(c_int) function f3d_set_tile(ctx, tx, ty) bind(C, name='f3d_set_tile')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: tx, ty
     end function f3d_set_tile

     integer(c_int) function f3d_set_tile_config(ctx, tx, ty, waves, slabs) &
          bind(C, name='f3d_set_tile_config')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: tx, ty, waves, slabs
     end function f3d_set_tile_config

     integer(c_int) function f3d_timing(ctx, enable) bind(C, name='f3d_timing')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: enable
     end function f3d_timing

     integer(c_int) function f3d_kernel_time(ctx, reset, avg_ms, count) &
          bind(C, name='f3d_kernel_time')
       import :: c_int, c_ptr, c_double, c_long
       type(c_ptr), value :: ctx
       integer(c_int), value :: reset
       real(c_double), intent(out) :: avg_ms
       integer(c_long), intent(out) :: count
     end function f3d_kernel_time

     integer(c_long_long) function f3d_device_bytes(ctx) bind(C, name='f3d_device_bytes')
       import :: c_long_long, c_ptr
       type(c_ptr), value :: ctx
     end function f3d_device_bytes

     integer(c_int) function f3d_mem_info(free_bytes, total_bytes) bind(C, name='f3d_mem_info')
       import :: c_int, c_long_long
       integer(c_long_long), intent(out) :: free_bytes, total_bytes
     end function f3d_mem_info

     integer(c_int) function f3d_hbm_copy(nbytes, reps, avg_ms) bind(C, name='f3d_hbm_copy')
       import :: c_int, c_long_long, c_double
       integer(c_long_long), value :: nbytes
       integer(c_int), value :: reps
       real(c_double), intent(out) :: avg_ms
     end function f3d_hbm_copy

     integer(c_int) function f3d_batch_set_tile(batch, tx, ty, slabs) &
          bind(C, name='f3d_batch_set_tile')
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       integer(c_int), value :: tx, ty, slabs
     end function f3d_batch_set_tile

     integer(c_int) function f3d_batch_get_tile(batch, tx, ty, slabs) &
          bind(C, name='f3d_batch_get_tile')
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       integer(c_int), intent(out) :: tx, ty, slabs
     end function f3d_batch_get_tile

     integer(c_int) function f3d_batch_timing(batch, enable) bind(C, name='f3d_batch_timing')
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       integer(c_int), value :: enable
     end function f3d_batch_timing

     integer(c_int) function f3d_batch_kernel_time(batch, reset, avg_ms, count) &
          bind(C, name='f3d_batch_kernel_time')
       import :: c_int, c_ptr, c_double, c_long
       type(c_ptr), value :: batch
       integer(c_int), value :: reset
       real(c_double), intent(out) :: avg_ms
       integer(c_long), intent(out) :: count
     end function f3d_batch_kernel_time

     integer(c_int) function f3d_batch_use_graph(batch, on) bind(C, name='f3d_batch_use_graph')
       import :: c_int, c_ptr
       type(c_ptr), value :: batch
       integer(c_int), value :: on
     end function f3d_batch_use_graph

     ! cmass3d over device arrays, allocating nothing (device pointers;
     ! stream: a hipStream_t or c_null_ptr)
     integer(c_int) function f3d_cmass3d_noalloc(c, cload, cdry, cwet, nx, ny, nz, nc, &
          nmembers, dx, dy, dz, out, stream) bind(C, name='f3d_cmass3d_noalloc')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: c, cload, cdry, cwet, dz, out, stream
       integer(c_int), value :: nx, ny, nz, nc, nmembers
       real(c_double), value :: dx, dy
     end function f3d_cmass3d_noalloc

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
