!-----------------------------------------------------------------------
! Fortran host driver over the GPU transport core.
!
! Control flow of Sources/Fall3d.f90:73-121 (meteo and source cadence,
! time += dt and iiter += 1 before each step, stop when time >= end_time,
! the 600 s mass-budget cadence of endstep.f90:51), with the transport
! core replaced by calls through the ISO_C_BINDING module fall3d_amd.
!
! Inputs come from a scenario file written by fall3d_amd.scenario
! (Scenario.write_binary); outputs are the final concentration, deposit and
! budget as a raw little-endian stream.
!
! usage: f3d_driver scenario.bin out.bin [fp32|fp64]
!-----------------------------------------------------------------------
program f3d_driver
  use iso_c_binding
  use fall3d_amd
  implicit none

  character(len=1024) :: fin, fout, fprec
  integer(c_int) :: nx, ny, nz, nc, nsou, modv, modkv, modkh, prec, rc, iiter, it_dev
  real(c_double) :: dx, dy, rkv0, rkh0, csh, flowrate, end_time
  real(c_double), allocatable, target :: zlayer(:)
  real(c_double), allocatable :: diam(:), rhop(:), sphe(:)
  real(c_double), allocatable :: vx(:), vy(:), vz(:), tempe(:), vptemp(:), rho(:)
  real(c_double), allocatable :: ustar(:), zi(:), rl(:), vdep(:), tmrat(:)
  integer(c_int), allocatable :: isrc(:), jsrc(:), ksrc(:)
  real(c_double), allocatable :: c(:), cdry(:)
  real(c_double) :: time, dt, dtc, meteo_time, out(6), t_dev, erumass, outmass
  type(f3d_grid) :: grid
  type(c_ptr) :: ctx
  logical :: go_on, meteotime
  integer :: u, n3, n2, nlog

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, fprec)
  prec = F3D_FP32
  if (trim(fprec) == 'fp64') prec = F3D_FP64

  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) nx, ny, nz, nc, nsou, modv, modkv, modkh
  read(u) dx, dy, rkv0, rkh0, csh, flowrate, end_time
  n3 = nx*ny*nz
  n2 = nx*ny
  allocate(zlayer(nz), diam(nc), rhop(nc), sphe(nc))
  allocate(vx(n3), vy(n3), vz(n3), tempe(n3), vptemp(n3), rho(n3))
  allocate(ustar(n2), zi(n2), rl(n2), vdep(n2*nc))
  allocate(isrc(max(nsou,1)), jsrc(max(nsou,1)), ksrc(max(nsou,1)), tmrat(max(nsou,1)*nc))
  read(u) zlayer, diam, rhop, sphe
  read(u) vx, vy, vz, tempe, vptemp, rho, ustar, zi, rl, vdep
  if (nsou > 0) read(u) isrc(1:nsou), jsrc(1:nsou), ksrc(1:nsou), tmrat(1:nsou*nc)
  close(u)

  grid%nx = nx; grid%ny = ny; grid%nz = nz; grid%nc = nc
  grid%dx = dx; grid%dy = dy
  grid%zlayer = c_loc(zlayer)
  call check(f3d_create(ctx, grid, prec), 'f3d_create')
  call check(f3d_set_classes(ctx, diam, rhop, sphe, modv), 'f3d_set_classes')
  call check(f3d_set_models(ctx, modkv, modkh, rkv0, rkh0, csh), 'f3d_set_models')
  call check(f3d_set_source(ctx, nsou, isrc, jsrc, ksrc, tmrat, flowrate), 'f3d_set_source')

  ! Fall3d.f90:73-121
  time = 0.0_c_double
  meteo_time = 0.0_c_double
  meteotime = .true.
  iiter = 0
  nlog = 0
  go_on = .true.
  do while (go_on)
     if (meteo_time <= time) then
        if (meteotime) then
           call check(f3d_set_meteo(ctx, vx, vy, vz, tempe, vptemp, rho, ustar, zi, rl, vdep), &
                'f3d_set_meteo')
           call check(f3d_get_dt(ctx, dt, dtc), 'f3d_get_dt')
           meteo_time = end_time          ! one meteo interval covers the run
        end if
        if (meteo_time >= end_time) meteotime = .false.
     end if
     time  = time + dt
     iiter = iiter + 1
     call check(f3d_step(ctx, 1), 'f3d_step')   ! setsrc + sizloop on the device
     ! endstep.f90:51: mass budget every 600 s of model time
     if (iiter == 1 .or. mod(time, 600.0_c_double) <= dt) then
        call check(f3d_budget(ctx, out), 'f3d_budget')
        nlog = nlog + 1
        write(*,'(a,i7,a,f10.3,a,f12.1,a,es13.6,a,es13.6,a,es13.6,a,es13.6)') &
             'step ', iiter, '  dt ', dt, '  time ', time, '  massvol ', out(1), &
             '  massdep ', out(2), '  outmass ', out(5), '  erupted ', out(6)
     end if
     if (time >= end_time) go_on = .false.
  end do

  call check(f3d_budget(ctx, out), 'f3d_budget')
  call check(f3d_get_clock(ctx, t_dev, it_dev, erumass, outmass), 'f3d_get_clock')
  if (it_dev /= iiter .or. t_dev /= time) then
     write(*,*) 'clock mismatch between driver and device: ', iiter, it_dev, time, t_dev
     stop 1
  end if
  allocate(c((nx+2)*(ny+2)*(nz+2)*nc), cdry(n2*(nc+1)))
  call check(f3d_get_state(ctx, c, cdry), 'f3d_get_state')
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) iiter, time, out, c, cdry
  close(u)
  call f3d_destroy(ctx)
  write(*,'(a,i7,a,i5,a)') 'done: ', iiter, ' steps, ', nlog, ' budget records'

contains

  subroutine check(rc_, what)
    integer(c_int), intent(in) :: rc_
    character(len=*), intent(in) :: what
    if (rc_ /= F3D_OK) then
       write(*,*) trim(what), ' failed: ', trim(f3d_error_message())
       stop 2    ! runend (runend.f90:42-67)
    end if
  end subroutine check

end program f3d_driver
