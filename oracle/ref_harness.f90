!-----------------------------------------------------------------------
! Test harness around the FALL3D-7.0 reference routines (test
! infrastructure only: it is linked with the reference's own Fortran
! sources, compiled in place by oracle/Makefile into oracle/_ref/, and is
! only ever loaded by tests/ and tests/golden/ to produce or check golden
! vectors).  Nothing in the product links this file.
!
! The harness owns no algorithm.  It fills the reference's module state
! (Master, Numeric, Domain, Parallel) from C pointers and calls the
! reference routines in the order the reference driver does:
!   Fall3d.f90:112-118   time += dt; iiter += 1; erumass += flowrate*dt;
!                        setsrc; sizloop
!   meteo.f90:27,81,100  kappa3, setvset, setvdrydep
!   meteo.f90:108-112    Hm1 scaling of vx, rho and rkh1 (three array
!                        statements; meteo itself cannot be linked because
!                        it calls the netCDF reader reamet)
!   meteo.f90:140,145    addradialwind (gravity current on), setdivu
!   meteo.f90:149        finddt
!   endstep.f90:71       divcorr (gravity current on, at the log times)
!   endstep.f90:80-106   getc0 (restated: endstep.f90 cannot be linked
!                        because it calls the netCDF writer printres_nc)
!   writim.f90:18        cmass3d
!-----------------------------------------------------------------------
module ref_harness
  use iso_c_binding
  use KindType
  use Numeric
  use Master
  use Parallel, only: parallel_startup, parallel_setgroups
  use Domain,   only: ic_loc, nc_loc, iz_loc, nz_loc
  use InpOut,   only: nlst
  implicit none
  logical, save :: started = .false.
contains

  subroutine ref_alloc(nx_, ny_, nz_, nc_, nsou_) bind(C, name='ref_alloc')
    integer(c_int), value :: nx_, ny_, nz_, nc_, nsou_
    integer :: ic
    if (.not. started) then
       call parallel_startup
       call parallel_setgroups(1)
       started = .true.
    end if
    nx = nx_; ny = ny_; nz = nz_; nc = nc_; np = nc_; ng = 0; nsou = nsou_
    ic_loc = 1; nc_loc = nc; iz_loc = 1; nz_loc = nz
    safety_factor = 0.9_rp          ! inidat.f90:22
    lmeth = 1
    wetdeposition = .false.
    aggregation = .false.
    gravity_current = .false.
    coord_sys = 'UTM'
    do ic = 1, ncmax
       write(classname(ic), '(a,i2.2)') 'class-', ic
    end do
    if (allocated(c)) call ref_free()
    allocate(Hm1(nx,ny), zlayer(nz), dz(0:nz), topg(nx,ny))
    allocate(vx(nx,ny,nz), vy(nx,ny,nz), vz(nx,ny,nz), tempe(nx,ny,nz), vptemp(nx,ny,nz))
    allocate(rho(nx,ny,nz), rkhor(nx,ny,nz), rkh1(nx,ny,nz), rkver(nx,ny,nz))
    allocate(ustar(nx,ny), zi(nx,ny), rl(nx,ny))
    allocate(vset(nx,ny,nz,nc), vdep(nx,ny,nc))
    allocate(c(0:nx+1,0:ny+1,0:nz+1,nc))
    allocate(cdry(nx,ny,0:nc), cwet(nx,ny,0:nc), cload(nx,ny,0:nc))
    allocate(work(max(nx+1,ny+1),2))
    allocate(isrc(max(nsou,1)), jsrc(max(nsou,1)), ksrc(max(nsou,1)), tmrat(max(nsou,1),nc))
    allocate(zs(max(nsou,1)), divu(nx,ny,nz))
    zs = 0.0_rp; divu = 0.0_rp
    c = 0.0_rp; cdry = 0.0_rp; cwet = 0.0_rp; cload = 0.0_rp
    vset = 0.0_rp; vdep = 0.0_rp; work = 0.0_rp
    rkhor = 0.0_rp; rkh1 = 0.0_rp; rkver = 0.0_rp
    Hm1 = 1.0_rp; topg = 0.0_rp
    time = 0.0_rp; iiter = 0; erumass = 0.0_rp; outmass = 0.0_rp; flowrate = 0.0_rp
    dt = 0.0_rp; dtc = 0.0_rp
  end subroutine ref_alloc

  subroutine ref_free() bind(C, name='ref_free')
    deallocate(Hm1, zlayer, dz, topg, vx, vy, vz, tempe, vptemp, rho, rkhor, rkh1, rkver)
    deallocate(ustar, zi, rl, vset, vdep, c, cdry, cwet, cload, work, isrc, jsrc, ksrc, tmrat)
    deallocate(zs, divu)
  end subroutine ref_free

  ! Grid: uniform horizontal spacing (UTM, Hm1 = 1), zlayer(1:nz); dz follows
  ! reagrd.f90:216-220 (the harness fills dz the same way the reader does).
  subroutine ref_set_grid(dx_, dy_, zl) bind(C, name='ref_set_grid')
    real(c_double), value :: dx_, dy_
    real(c_double) :: zl(nz)
    integer :: k
    dx = dx_; dy = dy_; dX1 = dx_; dX2 = dy_
    zlayer(1:nz) = zl(1:nz)
    do k = 1, nz-1
       dz(k) = zlayer(k+1) - zlayer(k)
    end do
    dz(0) = dz(1)
    dz(nz) = dz(nz-1)
  end subroutine ref_set_grid

  ! Scale factor and coordinate system.  lonlat_ = 1: a LON-LAT grid whose
  ! spacing is dlon_ x dlat_ degrees, dX1 = Rearth*(dx*pi/180) as
  ! reagrd.f90:155-158 computes it (kappa3's CMAQ Khn takes dx, dy in
  ! degrees, everything else dX1, dX2); lonlat_ = 0 keeps UTM and the
  ! spacing of ref_set_grid.  Hm1 is given per column (reagrd.f90:247-270
  ! reads it from the grid's latitudes).
  subroutine ref_set_hm1(lonlat_, dlon_, dlat_, hm1_) bind(C, name='ref_set_hm1')
    integer(c_int), value :: lonlat_
    real(c_double), value :: dlon_, dlat_
    real(c_double) :: hm1_(nx,ny)
    if (lonlat_ /= 0) then
       coord_sys = 'LON-LAT'
       dx = dlon_; dy = dlat_
       dX1 = Rearth*(dx*pi/180.0_rp)
       dX2 = Rearth*(dy*pi/180.0_rp)
    else
       coord_sys = 'UTM'
    end if
    Hm1 = hm1_
  end subroutine ref_set_hm1

  subroutine ref_get_scaled(dx1_, dx2_) bind(C, name='ref_get_scaled')
    real(c_double) :: dx1_, dx2_
    dx1_ = dX1; dx2_ = dX2
  end subroutine ref_get_scaled

  subroutine ref_set_models(modkv_, modkh_, modv_, rkv0_, rkh0_) bind(C, name='ref_set_models')
    integer(c_int), value :: modkv_, modkh_, modv_
    real(c_double), value :: rkv0_, rkh0_
    modkv = modkv_; modkh = modkh_; modv = modv_; rkv0 = rkv0_; rkh0 = rkh0_
  end subroutine ref_set_models

  subroutine ref_set_classes(diam_, rhop_, sphe_) bind(C, name='ref_set_classes')
    real(c_double) :: diam_(nc), rhop_(nc), sphe_(nc)
    diam(1:nc) = diam_(1:nc); rhop(1:nc) = rhop_(1:nc); sphe(1:nc) = sphe_(1:nc)
    call setpsi(psi, sphe, diam, modv, nc)       ! readat.f90:248
  end subroutine ref_set_classes

  subroutine ref_get_psi(out) bind(C, name='ref_get_psi')
    real(c_double) :: out(nc)
    out(1:nc) = psi(1:nc)
  end subroutine ref_get_psi

  subroutine ref_set_meteo(vx_, vy_, vz_, tempe_, vptemp_, rho_, ustar_, zi_, rl_, vdep_, topg_) &
       bind(C, name='ref_set_meteo')
    real(c_double) :: vx_(nx,ny,nz), vy_(nx,ny,nz), vz_(nx,ny,nz), tempe_(nx,ny,nz)
    real(c_double) :: vptemp_(nx,ny,nz), rho_(nx,ny,nz)
    real(c_double) :: ustar_(nx,ny), zi_(nx,ny), rl_(nx,ny), vdep_(nx,ny,nc), topg_(nx,ny)
    vx = vx_; vy = vy_; vz = vz_; tempe = tempe_; vptemp = vptemp_; rho = rho_
    ustar = ustar_; zi = zi_; rl = rl_; vdep = vdep_; topg = topg_
  end subroutine ref_set_meteo

  ! meteo.f90:99-100: setvdrydep from the land use LDU and the roughness
  ! length znt; its other inputs (rl, ustar, tempe, rho, zlayer, diam) are
  ! the module state set above.  vdep_ is in/out: classes coarser than
  ! 100 um keep their value.
  subroutine ref_setvdrydep(ldu_, znt_, vdep_) bind(C, name='ref_setvdrydep')
    real(c_double) :: ldu_(nx,ny), znt_(nx,ny), vdep_(nx,ny,nc)
    if (allocated(LDU)) deallocate(LDU)
    if (allocated(znt)) deallocate(znt)
    allocate(LDU(nx,ny), znt(nx,ny))
    LDU = ldu_; znt = znt_; vdep = vdep_
    call setvdrydep
    vdep_ = vdep
  end subroutine ref_setvdrydep

  ! meteo.f90:116-124: corvver on given Hm1, topg and (scaled) winds; vz_
  ! in/out
  subroutine ref_corvver(hm1_, topg_, vx_, vy_, vz_) bind(C, name='ref_corvver')
    real(c_double) :: hm1_(nx,ny), topg_(nx,ny), vx_(nx,ny,nz), vy_(nx,ny,nz), vz_(nx,ny,nz)
    Hm1 = hm1_; topg = topg_; vx = vx_; vy = vy_; vz = vz_
    call corvver
    vz_ = vz
  end subroutine ref_corvver

  ! wet deposition on (sizloop.f90:65-67) with the precipitation rate
  ! prate(nx,ny) in mm/h; zi is the module state set above
  subroutine ref_set_wetdep(prate_) bind(C, name='ref_set_wetdep')
    real(c_double) :: prate_(nx,ny)
    if (allocated(prate)) deallocate(prate)
    allocate(prate(nx,ny))
    prate = prate_
    wetdeposition = .true.
  end subroutine ref_set_wetdep

  subroutine ref_get_cwet(cwet_) bind(C, name='ref_get_cwet')
    real(c_double) :: cwet_(nx,ny,0:nc)
    cwet_ = cwet
  end subroutine ref_get_cwet

  subroutine ref_kappa3() bind(C, name='ref_kappa3')
    call kappa3
  end subroutine ref_kappa3

  subroutine ref_setvset() bind(C, name='ref_setvset')
    call setvset
  end subroutine ref_setvset

  ! meteo.f90:108-112 (the three scaling statements), corvver on LON-LAT
  ! grids (meteo.f90:116-124), then finddt (meteo.f90:149)
  subroutine ref_scale_and_finddt(dt_, dtc_) bind(C, name='ref_scale_and_finddt')
    real(c_double) :: dt_, dtc_
    integer :: k
    do k = 1, nz
       vx  (1:nx,1:ny,k) = vx   (1:nx,1:ny,k)/ Hm1(1:nx,1:ny)
       rho (1:nx,1:ny,k) = rho  (1:nx,1:ny,k)* Hm1(1:nx,1:ny)
       rkh1(1:nx,1:ny,k) = rkhor(1:nx,1:ny,k)/(Hm1(1:nx,1:ny)*Hm1(1:nx,1:ny))
    end do
    if (coord_sys == 'LON-LAT') call corvver
    if (gravity_current) call addradialwind
    call setdivu
    call finddt
    dt_ = dt; dtc_ = dtc
  end subroutine ref_scale_and_finddt

  ! The GRAVITY_CURRENT block (readat.f90:149-200): prm = {c_flow_rate,
  ! lambda_grav, k_entrain, brunt_vaisala, centre x, centre y, xorigr,
  ! yorigr} (the centre is RADIAL_WIND_LON/LAT on LON-LAT grids, else
  ! RADIAL_WIND_X/Y), beg_time and the source heights zs(nsou).  The
  ! routine's log lines (addradialwind.f90:88-96) go to /dev/null.
  subroutine ref_set_gravity(prm, beg_time_, zs_) bind(C, name='ref_set_gravity')
    real(c_double) :: prm(8), zs_(nsou)
    real(c_double), value :: beg_time_
    c_flow_rate = prm(1); lambda_grav = prm(2); k_entrain = prm(3); brunt_vaisala = prm(4)
    if (coord_sys == 'LON-LAT') then
       radial_wind_lon = prm(5); radial_wind_lat = prm(6)
    else
       radial_wind_x = prm(5); radial_wind_y = prm(6)
    end if
    xorigr = prm(7); yorigr = prm(8)
    beg_time = beg_time_
    zs(1:nsou) = zs_(1:nsou)
    open(unit=nlst, file='/dev/null', status='old', action='write')
    gravity_current = .true.
  end subroutine ref_set_gravity

  ! endstep.f90:71 after writim's cmass3d (ref_budget)
  subroutine ref_divcorr() bind(C, name='ref_divcorr')
    call divcorr
  end subroutine ref_divcorr

  subroutine ref_get_wind(vx_, vy_, divu_) bind(C, name='ref_get_wind')
    real(c_double) :: vx_(nx,ny,nz), vy_(nx,ny,nz), divu_(nx,ny,nz)
    vx_ = vx; vy_ = vy; divu_ = divu
  end subroutine ref_get_wind

  subroutine ref_set_dt(dt_) bind(C, name='ref_set_dt')
    real(c_double), value :: dt_
    dt = dt_
  end subroutine ref_set_dt

  subroutine ref_set_source(is_, js_, ks_, tm_, flow_) bind(C, name='ref_set_source')
    integer(c_int) :: is_(nsou), js_(nsou), ks_(nsou)
    real(c_double) :: tm_(nsou,nc)
    real(c_double), value :: flow_
    isrc(1:nsou) = is_(1:nsou); jsrc(1:nsou) = js_(1:nsou); ksrc(1:nsou) = ks_(1:nsou)
    tmrat(1:nsou,1:nc) = tm_(1:nsou,1:nc)
    flowrate = flow_
  end subroutine ref_set_source

  ! One pass of the driver loop body, Fall3d.f90:112-117
  subroutine ref_step() bind(C, name='ref_step')
    time    = time  + dt
    iiter   = iiter + 1
    erumass = erumass + flowrate*dt
    call setsrc
    call sizloop
  end subroutine ref_step

  ! getc0 (endstep.f90:80-106), cload (endstep.f90:35), cmass3d (writim.f90:18)
  subroutine ref_budget(mv, md, mdry, mwet, om, em) bind(C, name='ref_budget')
    real(c_double) :: mv, md, mdry, mwet, om, em
    integer :: i, j, ic
    do i = 1, nx
       do j = 1, ny
          cdry(i,j,0) = 0.0_rp
          cwet(i,j,0) = 0.0_rp
          do ic = 1, np
             cdry(i,j,0) = cdry(i,j,0) + cdry(i,j,ic)
             cwet(i,j,0) = cwet(i,j,0) + cwet(i,j,ic)
          end do
       end do
    end do
    cload(1:nx,1:ny,0:nc) = cdry(1:nx,1:ny,0:nc) + cwet(1:nx,1:ny,0:nc)
    call cmass3d(c,cload,cdry,cwet,nx,ny,nz,nc,dX1,dX2,dz,massvol,massdep,massdry,masswet)
    mv = massvol; md = massdep; mdry = massdry; mwet = masswet; om = outmass; em = erumass
  end subroutine ref_budget

  subroutine ref_get_state(c_, cdry_, t_, it_) bind(C, name='ref_get_state')
    real(c_double) :: c_(0:nx+1,0:ny+1,0:nz+1,nc), cdry_(nx,ny,0:nc), t_
    integer(c_int) :: it_
    c_ = c; cdry_ = cdry; t_ = time; it_ = iiter
  end subroutine ref_get_state

  subroutine ref_set_state(c_, cdry_) bind(C, name='ref_set_state')
    real(c_double) :: c_(0:nx+1,0:ny+1,0:nz+1,nc), cdry_(nx,ny,0:nc)
    c = c_; cdry = cdry_
  end subroutine ref_set_state

  subroutine ref_get_vz(vz_) bind(C, name='ref_get_vz')
    real(c_double) :: vz_(nx,ny,nz)
    vz_ = vz
  end subroutine ref_get_vz

  subroutine ref_get_coef(vx_, rho_, rkhor_, rkh1_, rkver_, vset_) bind(C, name='ref_get_coef')
    real(c_double) :: vx_(nx,ny,nz), rho_(nx,ny,nz), rkhor_(nx,ny,nz), rkh1_(nx,ny,nz)
    real(c_double) :: rkver_(nx,ny,nz), vset_(nx,ny,nz,nc)
    vx_ = vx; rho_ = rho; rkhor_ = rkhor; rkh1_ = rkh1; rkver_ = rkver; vset_ = vset
  end subroutine ref_get_coef

  ! Direct setters for the coefficient arrays (to drive sizloop with chosen fields)
  subroutine ref_set_coef(vx_, vy_, vz_, rho_, rkhor_, rkh1_, rkver_, vset_, vdep_) &
       bind(C, name='ref_set_coef')
    real(c_double) :: vx_(nx,ny,nz), vy_(nx,ny,nz), vz_(nx,ny,nz), rho_(nx,ny,nz)
    real(c_double) :: rkhor_(nx,ny,nz), rkh1_(nx,ny,nz), rkver_(nx,ny,nz), vset_(nx,ny,nz,nc)
    real(c_double) :: vdep_(nx,ny,nc)
    vx = vx_; vy = vy_; vz = vz_; rho = rho_; rkhor = rkhor_; rkh1 = rkh1_; rkver = rkver_
    vset = vset_; vdep = vdep_
  end subroutine ref_set_coef

  ! Output-side routines of printres_nc.f90 (which itself needs netCDF):
  ! the flight-level cut setcut_FL (printres_nc.f90:556, setcut.f90:1-57) of
  ! var3d on the pressure p at flight level fl (m), the ICAO standard
  ! atmosphere it takes the nominal pressure from (setcut.f90:61-107), and
  ! ter2asl (printres_nc.f90:586, ter2asl.f90) on the module's zlayer
  subroutine ref_setcut_fl(var3d_, p_, topg_, fl_, var2d_) bind(C, name='ref_setcut_fl')
    real(c_double) :: var3d_(nx,ny,nz), p_(nx,ny,nz), topg_(nx,ny), var2d_(nx,ny)
    real(c_double), value :: fl_
    real(rp) :: flm
    flm = fl_
    call setcut_FL(nx,ny,nz,zlayer,var2d_,var3d_,topg_,p_,flm)
  end subroutine ref_setcut_fl

  subroutine ref_standard_atm(z_, p_, t_) bind(C, name='ref_standard_atm')
    real(c_double), value :: z_
    real(c_double) :: p_, t_
    real(rp) :: z
    z = z_
    call standard_atm(p_, t_, z)
  end subroutine ref_standard_atm

  subroutine ref_ter2asl(conc_, topg_) bind(C, name='ref_ter2asl')
    real(c_double) :: conc_(nx,ny,nz), topg_(nx,ny)
    call ter2asl(conc_, topg_, zlayer, nx, ny, nz)
  end subroutine ref_ter2asl

  subroutine ref_set_clock(time_, iiter_, erumass_, outmass_) bind(C, name='ref_set_clock')
    real(c_double), value :: time_, erumass_, outmass_
    integer(c_int), value :: iiter_
    time = time_; iiter = iiter_; erumass = erumass_; outmass = outmass_
  end subroutine ref_set_clock

end module ref_harness
