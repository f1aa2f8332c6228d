/*
 * fall3d_amd: MI355X-native transport core for FALL3D-7.0 (C ABI).
 *
 * Two layers, both plain C: pointers, sizes, no device or C++ types.
 *
 * 1. Per-routine drop-ins with the reference's own external-procedure ABI
 *    (lower case + one trailing underscore, every argument by reference,
 *    Fortran column-major arrays, no hidden lengths).  A reference build
 *    that links libfall3d_amd.so ahead of its own objects gets the GPU
 *    versions of these routines without a source change.  They stage host
 *    arrays to the device, run the plain FP64 kernel, and copy back.  Only
 *    the serial decomposition (nz_loc == nz, iz_loc == 1; ic_loc = 1,
 *    nc_loc = nc) is accepted; anything else leaves the outputs untouched
 *    and sets f3d_last_error().
 *
 * 2. A device-resident context (f3d_*) that holds grid, meteorology,
 *    coefficients, concentration and budgets in HBM and runs the whole
 *    per-step transport core there: the calls the driver loop of
 *    Sources/Fall3d.f90:79-121 makes (meteo, setsrc, sizloop, endstep
 *    budgets).  Precision F3D_FP64 replays the reference routine by
 *    routine in FP64; F3D_FP32 runs one fused FP32 kernel per step over
 *    all classes with FP64 deposit, budgets, CFL step and settling.
 *
 * Errors: every int-returning call returns 0 on success and a negative
 * code otherwise (F3D_E*); f3d_last_error() gives the message.  All calls
 * on one context must come from one host thread at a time.
 */
#ifndef FALL3D_AMD_H
#define FALL3D_AMD_H
#ifdef __cplusplus
extern "C" {
#endif

enum {
    F3D_OK = 0,
    F3D_EINVAL = -1,   /* bad argument, size or unsupported decomposition */
    F3D_EHIP = -2,     /* HIP runtime error (message carries the HIP error) */
    F3D_ENOMEM = -3,   /* device allocation failed */
    F3D_ECONV = -4,    /* vsettl / get_gama did not converge (runend in the reference) */
    F3D_ESTATE = -5    /* call out of order (e.g. step before meteo) */
};

enum { F3D_FP64 = 0, F3D_FP32 = 1 };

/* ---------------- layer 1: reference-ABI drop-ins (FP64) ---------------- */

/* Sources/setbcc.f90:1-13 */
void setbcc_(double *c, int *nx, int *ny, int *nz, int *nz_loc, int *iz_loc, const double *vx,
             const double *vy, const double *vz, const double *vset, const double *vdep,
             const double *rkver, const double *dz);
/* Sources/advctzc.f90:1-17 (lmeth unused, as in the reference) */
void advctzc_(double *c, int *nx, int *ny, int *nz, int *nz_loc, int *iz_loc, const double *vz,
              const double *vset, const double *dz, double *dt, int *lmeth);
/* Sources/advctx.f90:1-19 (work(nx+1,2) is not touched) */
void advctx_(double *c, int *nx, int *ny, int *nz, int *nz_loc, int *iz_loc, const double *vx,
             double *dx, double *dt, double *work, int *lmeth);
/* Sources/advcty.f90:1-19 */
void advcty_(double *c, int *nx, int *ny, int *nz, int *nz_loc, int *iz_loc, const double *vy,
             double *dy, double *dt, double *work, int *lmeth);
/* Sources/diffz.f90:1-13 */
void diffz_(double *c, int *nx, int *ny, int *nz, int *nz_loc, int *iz_loc, const double *dz,
            double *dt, const double *rkver, const double *rho, double *work);
/* Sources/diffx.f90:1-13 (called with rkh1 by sizloop.f90:56,60) */
void diffx_(double *c, int *nx, int *ny, int *nz, int *nz_loc, int *iz_loc, double *dx,
            double *dt, const double *rkhor, const double *rho, double *work);
/* Sources/diffy.f90:1-13 (called with rkhor by sizloop.f90:57,59) */
void diffy_(double *c, int *nx, int *ny, int *nz, int *nz_loc, int *iz_loc, double *dy,
            double *dt, const double *rkhor, const double *rho, double *work);
/* Sources/accum.f90:1-15; cdry is (nx,ny,0:*) and ic selects the slot */
void accum_(const double *c, double *cdry, int *ic, int *nx, int *ny, int *nz, int *nz_loc,
            int *iz_loc, const double *vz, double *dt, const double *vset, const double *vdep);
/* Sources/wetdep.f90:1 (called at sizloop.f90:65-67): wet deposition of
 * class ic (1-based) into cwet(nx,ny,0:*); classes above 100 um return at
 * once, as the reference does. */
void wetdep_(double *c, double *cwet, int *ic, int *nx, int *ny, int *nz, int *nz_loc,
             int *iz_loc, double *dt, const double *diam, const double *zlayer, const double *dz,
             const double *zi, const double *prate);
/* Sources/cmass2d.f90:1-30; outmass accumulates */
void cmass2d_(const double *c, const double *vx, const double *vy, const double *vz,
              const double *vset, int *nx, int *ny, int *nz, int *nc, int *ic_loc, int *nc_loc,
              int *iz_loc, int *nz_loc, double *dx, double *dy, const double *dz, double *dt,
              double *outmass);
/* Sources/cmass3d.f90:1-24 */
void cmass3d_(const double *c, const double *cload, const double *cdry, const double *cwet,
              int *nx, int *ny, int *nz, int *nc, double *dx, double *dy, const double *dz,
              double *massvol, double *massdep, double *massdry, double *masswet);

/* No counterpart in the reference: cmass3d_ (Sources/cmass3d.f90:1-24) over
 * arrays already in device memory, for callers that cannot allocate (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * It calls no hipMalloc, hipFree or allocating routine: its partial sums
 * live in one fixed device array of 512 KiB (4 x 256 partials per member,
 * at most 64 members).  c: nmembers slabs of (0:nx+1,0:ny+1,0:nz+1,nc),
 * member after member; cload, cdry, cwet: nmembers arrays (nx,ny,0:nc), of
 * which slot 0 is read; dz: (0:nz); out: 4*nmembers doubles, massvol,
 * massdep, massdry, masswet of each member (times dx*dy, as cmass3d_).  All
 * pointers are device pointers; the work is queued on `stream` (a
 * hipStream_t, NULL: the null stream) and the call returns at once.  Every
 * call on one GPU shares the partials: calls that overlap on one GPU race.
 * More than 64 members is refused (F3D_EINVAL). */
int f3d_cmass3d_noalloc(const double *c, const double *cload, const double *cdry,
                        const double *cwet, int nx, int ny, int nz, int nc, int nmembers,
                        double dx, double dy, const double *dz, double *out, void *stream);

/* ---------------- layer 2: device-resident context ---------------- */

typedef struct f3d_ctx f3d_ctx;

typedef struct {
    int nx, ny, nz, nc;          /* grid (Numeric.f90:18) and classes (Master.f90:44) */
    double dx, dy;               /* dX1, dX2 in m (reagrd.f90:194-200) */
    const double *zlayer;        /* (nz) layer heights; dz follows reagrd.f90:216-220 */
    const double *hm1;           /* (nx,ny) scale factor, NULL for UTM (Hm1 = 1) */
    int lonlat;                  /* 0 UTM, 1 LON-LAT: kappa3's CMAQ constant and
                                    spacing, corvver (meteo.f90:116-124) */
    double dlon, dlat;           /* LON-LAT: the spacing dx, dy in degrees of the
                                    reference (reagrd.f90:155-158), which kappa3's
                                    CMAQ Khn takes (kappa3.f90:226); 0 derives it
                                    from dX1, dX2 and Rearth = 6356 km (Numeric.f90:14).
                                    Ignored on UTM grids. */
} f3d_grid;

/* Allocates every device buffer (Master.f90:getmem1 equivalent). */
int f3d_create(f3d_ctx **ctx, const f3d_grid *grid, int precision);
/* Ensemble of `nmembers` runs on one grid, one meteorology and one dt that
 * differ only in their source (FP32 fused path): every member is a FALL3D
 * run of its own (Fall3d.f90:79-121) and one step advances all of them in
 * one launch.  No counterpart in the reference, which runs one scenario per
 * process. */
int f3d_create_ensemble(f3d_ctx **ctx, const f3d_grid *grid, int nmembers);
void f3d_destroy(f3d_ctx *ctx);

/* Classes and settling model (reagrn.f90:20-87, readat.f90:218-248).
 * modv: 1 Arastoopour, 2 Ganser, 3 Wilson, 4 Dellino. */
int f3d_set_classes(f3d_ctx *ctx, const double *diam, const double *rhop, const double *sphe,
                    int modv);
/* Turbulence models (readat.f90:251-324): modkv 0 CONSTANT, 1 SIMILARITY,
 * 2 SURFACE_LAYER; modkh 0 CONSTANT, 1 RAMS, 2 CMAQ. */
int f3d_set_models(f3d_ctx *ctx, int modkv, int modkh, double rkv0, double rkh0, double csh);

/* Meteo read (meteo.f90:19-149): uploads the fields (Fortran order, host)
 * and runs kappa3, setvset, the Hm1 scaling and finddt on the device.
 * 3-D fields are (nx,ny,nz); ustar, zi, rl are (nx,ny); vdep is (nx,ny,nc)
 * (the setvdrydep output).  vz(:,:,1) must already be 0 (reamet.f90:153). */
int f3d_set_meteo(f3d_ctx *ctx, const double *vx, const double *vy, const double *vz,
                  const double *tempe, const double *vptemp, const double *rho,
                  const double *ustar, const double *zi, const double *rl, const double *vdep);
/* The same from meteorology already in HBM: the window (i0+1..i0+nx,
 * j0+1..j0+ny, 1..nz) of a larger context's fields (same dx, dy and
 * levels; Hm1 = 1).  No counterpart in the reference: it lets many
 * forecasts on one resident regional meteorology start without a host copy.
 * Deposition velocities come from the region when the class counts agree,
 * else they are zero. */
int f3d_set_meteo_window(f3d_ctx *ctx, const f3d_ctx *region, int i0, int j0);
int f3d_get_dt(f3d_ctx *ctx, double *dt, double *dtc);
/* Override dt (the reference keeps dt until the next meteo read). */
int f3d_set_dt(f3d_ctx *ctx, double dt);

/* Source nodes (source.f90, 1-based indices) with tmrat(nsou,nc) in kg/s,
 * already multiplied by Hm1 (source.f90:75-77); flowrate for erumass. */
int f3d_set_source(f3d_ctx *ctx, int nsou, const int *isrc, const int *jsrc, const int *ksrc,
                   const double *tmrat, double flowrate);
/* Source of one ensemble member (0-based); f3d_set_source sets all members. */
int f3d_set_member_source(f3d_ctx *ctx, int member, int nsou, const int *isrc, const int *jsrc,
                          const int *ksrc, const double *tmrat, double flowrate);

/* nsteps passes of the loop body Fall3d.f90:112-117: time += dt,
 * iiter += 1, erumass += flowrate*dt, setsrc, sizloop (with accum, clip and
 * cmass2d).  Asynchronous on the context's stream. */
int f3d_step(f3d_ctx *ctx, int nsteps);

/* getc0 + cload + cmass3d (endstep.f90:35,80-106; writim.f90:18).
 * out[6] = massvol, massdep, massdry, masswet, outmass, erumass. */
int f3d_budget(f3d_ctx *ctx, double *out);
/* f3d_budget for every member: out[6*m + 0..5]. */
int f3d_ensemble_budget(f3d_ctx *ctx, double *out);

/* State transfer, FP64 host arrays in reference layout:
 * c (0:nx+1,0:ny+1,0:nz+1,nc), cdry (nx,ny,0:nc) (slot 0 as last set by
 * f3d_budget).  Either pointer may be NULL.  f3d_set_state: single runs only. */
int f3d_get_state(f3d_ctx *ctx, double *c, double *cdry);
/* State of one ensemble member; f3d_get_state is member 0. */
int f3d_get_member_state(f3d_ctx *ctx, int member, double *c, double *cdry);
int f3d_set_state(f3d_ctx *ctx, const double *c, const double *cdry);
/* Coefficients after the meteo update (scaled as the reference holds them):
 * vx/Hm1, rho*Hm1, rkhor, rkh1, rkver (nx,ny,nz) and vset (nx,ny,nz,nc). */
int f3d_get_coef(f3d_ctx *ctx, double *vx, double *rho, double *rkhor, double *rkh1,
                 double *rkver, double *vset);
/* Land use and roughness length per column (nx,ny): LDU (USGS category) and
 * znt of Master.f90:96,108, the fields setvdrydep reads (setvdrydep.f90,
 * called at meteo.f90:99-100).  Once set, every meteo read computes vdep on
 * the device for the classes up to 100 um, on the density before the Hm1
 * scaling; classes above 100 um keep the vdep f3d_set_meteo was given.
 * Without it vdep is the caller's. */
int f3d_set_landuse(f3d_ctx *ctx, const double *ldu, const double *znt);
/* Terrain height per column (nx,ny), Master topg.  On LON-LAT grids
 * (f3d_grid.lonlat = 1) every later meteo read runs corvver
 * (corvver.f90, called at meteo.f90:116-124): vz minus the terrain-slope
 * projection of the scaled horizontal wind.  Without it the terrain is
 * taken as flat (no correction); UTM grids never correct. */
int f3d_set_topography(f3d_ctx *ctx, const double *topg);
/* Precipitation rate per column (nx,ny) in mm/h, Master prate: turns on
 * the wet deposition of sizloop.f90:65-67 (wetdep.f90, Jung and Shao) in
 * every later f3d_step, into cwet and the budgets' masswet, on the FP64 and
 * the FP32 path.  A context with precipitation steps through f3d_step only:
 * f3d_batch_create refuses it.  Needs f3d_set_classes first. */
int f3d_set_precipitation(f3d_ctx *ctx, const double *prate);
/* Output products of one member from the device state, as printres_nc.f90
 * (:150-241) forms them before writing: load (nx,ny,0:nc) = cload/Hm1 in
 * kg/m2 with cload = cdry + cwet (endstep.f90:35) and slot 0 the class sum;
 * thick (nx,ny) = 1e2*sum(cload/rhop)/Hm1, the deposit thickness in cm;
 * colmass (nx,ny) = 1e3*trapezoid over the levels of sum_ic(c)/Hm1, the
 * z-cumulative concentration in g/m2.  Any output may be NULL.  The deposit
 * slots as the steps left them (slot 0 is formed here, as getc0 would). */
int f3d_output_products(f3d_ctx *ctx, int member, double *load, double *thick, double *colmass);
/* vdep (nx,ny,nc) as the last meteo read left it (Master vdep). */
int f3d_get_vdep(f3d_ctx *ctx, double *vdep);
/* Pressure p(nx,ny,nz) in Pa (Master.f90:103, read with the meteo in
 * reamet), which the flight-level cuts interpolate in. */
int f3d_set_pressure(f3d_ctx *ctx, const double *p);
/* Concentration outputs of printres_nc.f90 for one member, in g/m3 (every
 * class is a particle class, np = nc), each in the reference's operation
 * order; null outputs are skipped:
 *   cgrnd (nx,ny)     C_GRND = 1e3*sum_ic c(:,:,1,ic)/Hm1 (printres_nc.f90:391-396)
 *   con   (nx,ny,nz)  CON = 1e3*sum_ic c(:,:,1:nz,ic)/Hm1 (printres_nc.f90:574-584);
 *                     asl = 1 interpolates it to the heights zlayer above sea
 *                     level on the topography (ter2asl.f90, the reference's
 *                     terrain_following = .false., printres_nc.f90:586;
 *                     needs f3d_set_topography)
 *   cfl   (nx,ny,nfl) C_FLxxx: sum_ic c/Hm1 cut at flight level fl[l]
 *                     (hundreds of feet) by the pressure of f3d_set_pressure
 *                     against the ICAO standard atmosphere, times 1e3
 *                     (printres_nc.f90:545-560, setcut.f90:1-107) */
int f3d_output_concentration(f3d_ctx *ctx, int member, double *cgrnd, double *con, int asl,
                             int nfl, const double *fl, double *cfl);

/* Gravity current of the umbrella cloud: the GRAVITY_CURRENT block of the
 * input file (readat.f90:149-200, Master.f90:69-78).  Coordinates are the
 * grid's own, as XORIGR/YORIGR and the spacing dx, dy are (degrees on
 * LON-LAT grids, m on UTM grids); center is RADIAL_WIND_LON/LAT (LON-LAT) or
 * RADIAL_WIND_X/Y (UTM); beg_time is the eruption start in s, on the clock
 * of f3d_set_clock. */
typedef struct {
    double c_flow_rate, lambda_grav, k_entrain, brunt_vaisala;
    double center_x, center_y;
    double xorigr, yorigr, dx, dy;
    double beg_time;
} f3d_gravity;
/* Turns the gravity current on (gc non-null) or off (gc NULL) for a
 * single-member context; zs (nsou) are the source heights of the current
 * source (Master zs), of which the reference takes maxval(zs) as the column
 * top: call again when the source changes.  Every later meteo read adds the
 * radial wind after the Hm1 scaling and corvver (addradialwind.f90, called
 * at meteo.f90:140; not while the source's total rate sum(tmrat) is at most
 * 0.1) and forms the wind divergence (setdivu.f90, meteo.f90:145), before
 * finddt; every f3d_step that ends at a log time (the first step, and
 * whenever mod(time, 600) <= dt, endstep.f90:53) applies divcorr
 * (divcorr.f90, endstep.f90:71) on that step's budgets.  The reference
 * clamps the crater node (i_wind, j_wind) on the low side only; the
 * library also clamps it to nx, ny.  Replaces (has no counterpart for)
 * nothing: the reference's three routines take module state, so they have
 * no argument-level drop-in. */
int f3d_set_gravity_current(f3d_ctx *ctx, const f3d_gravity *gc, int nsou, const double *zs);
/* The radial wind's scalars of the last meteo read, diag[9] = {min_radius,
 * max_radius, radius_grav, c_grav_factor, kmin, kmax, i_wind, j_wind,
 * k_wind} (addradialwind.f90:52-155), the factor corrdiv of the last
 * divcorr, and whether the last meteo read added the wind; any may be NULL. */
int f3d_get_gravity_current(f3d_ctx *ctx, double *diag, double *corrdiv, int *active);
/* The wind divergence divu (nx,ny,nz) of setdivu.f90:16-30 (Master divu)
 * on the current scaled wind: formed at every meteo read while the gravity
 * current is on, else on this call. */
int f3d_get_divu(f3d_ctx *ctx, double *divu);

/* Run date, extent and clock recorded in a restart file (the global
 * attributes of wrirst.f90:120-172). */
typedef struct {
    int year, month, day;           /* YEAR, MONTH, DAY */
    int run_start, run_end;         /* RUN_START, RUN_END (s) */
    double xmin, ymin, xmax, ymax;  /* XMIN..YMAX (UTM, m) or LONMIN..LATMAX (LON-LAT, deg) */
    double erumass, outmass;        /* ERUPTED_MASS, OUT_MASS: read back only; written
                                       from the member's clock */
} f3d_run_info;
/* Restart file of wrirst.f90 in the netCDF classic format (64-bit offsets;
 * written without the netCDF library): dimensions nx, ny, nc, nc1, nx2,
 * ny2, nz2; CON(nx2,ny2,nz2,nc) = the state c with its ghost frame and
 * LOAD(nx,ny,nc1) = cload(nx,ny,0:nc) = cdry + cwet with the class sums in
 * slot 0 (endstep.f90:35,80-106), both 32-bit floats as NF90_FLOAT makes
 * them; the global attributes of info and the member's erupted mass and
 * outflow. */
int f3d_write_restart(f3d_ctx *ctx, int member, const char *path, const f3d_run_info *info);
/* rearst.f90: checks the file's grid against the context's (nx, ny, nz2 - 2,
 * nc; F3D_EINVAL on a mismatch), sets the state c from CON and the clock's
 * erupted mass and outflow from ERUPTED_MASS and OUT_MASS, and zeroes the
 * dry and wet deposits, which the reference also starts again from zero
 * (rearst.f90:186 restores cload only, which the next endstep overwrites
 * with cdry + cwet, endstep.f90:35).  cload (nx,ny,nc+1) receives LOAD and
 * info the attributes; both may be NULL.  Single-member contexts. */
int f3d_read_restart(f3d_ctx *ctx, const char *path, double *cload, f3d_run_info *info);
/* The file side of the two above on host arrays (no device work): c
 * (nx+2,ny+2,nz+2,nc) and cload (nx,ny,nc+1) in, the attributes of info
 * (erumass and outmass included) and the coordinate system (lonlat) with
 * them; the reader checks the grid and fills c, cload and info (each may
 * be NULL). */
int f3d_restart_write(const char *path, int nx, int ny, int nz, int nc, int lonlat,
                      const double *c, const double *cload, const f3d_run_info *info);
int f3d_restart_read(const char *path, int nx, int ny, int nz, int nc, double *c, double *cload,
                     f3d_run_info *info);
int f3d_get_clock(f3d_ctx *ctx, double *time, int *iiter, double *erumass, double *outmass);
int f3d_set_clock(f3d_ctx *ctx, double time, int iiter, double erumass, double outmass);
/* The SAVEd vsettl drag coefficient carried between meteo reads
 * (setvset.f90:89); starts at 1.0. */
int f3d_get_cd_state(f3d_ctx *ctx, double *cd);
int f3d_set_cd_state(f3d_ctx *ctx, double cd);

int f3d_sync(f3d_ctx *ctx);
/* Tile shape (columns x rows) of the fused FP32 step kernel.  No
 * counterpart in the reference: a tuning knob set by the autotuner.
 * Supported: 58x10, 58x18, 58x26 (the flux kernel: 64-lane region rows,
 * 58 output columns).  Default: 58x18 with two member*class slabs per
 * workgroup (f3d_set_tile_config). */
int f3d_set_tile(f3d_ctx *ctx, int tx, int ty);
/* The same with the 58-column kernel's options: waves per workgroup (0 or
 * (ty+6)/2: two region rows per wave; or 8 with ty = 18: three rows per
 * wave) and member*class slabs per workgroup (0 or 1: one; 2: two slabs
 * share the coefficient loads and barriers).  No counterpart in the
 * reference. */
int f3d_set_tile_config(f3d_ctx *ctx, int tx, int ty, int waves, int slabs);
/* HIP-event timing of the step kernels (enabled by f3d_timing(ctx,1)):
 * average duration in ms and launch count of the dominant step kernel
 * since the last reset. */
int f3d_timing(f3d_ctx *ctx, int enable);
int f3d_kernel_time(f3d_ctx *ctx, int reset, double *avg_ms, long *count);
/* Device bytes held by the context. */
long long f3d_device_bytes(const f3d_ctx *ctx);

const char *f3d_last_error(void);

/* ---------------- Layer 3: batched steps and tile autotuning ----------------
 * No counterpart in the reference (its parallel form is one scenario split
 * over MPI ranks).  A batch steps a set of FP32 contexts -- any grid sizes,
 * class and member counts, each with its own meteorology, sources and dt --
 * together: per step one persistent launch draws every (scenario, tile,
 * slab group) work item from a queue, and the step's device work is
 * captured in a HIP graph and replayed.  Each context keeps its own clock,
 * sources and budgets (f3d_budget, f3d_get_state work on it as usual); a
 * context's result in a batch equals the one it gets from f3d_step with the
 * same tile shape.  The contexts must stay alive while the batch exists;
 * they may be on different iiter parities (each keeps its own sweep order).
 * Contexts with wet deposition or the gravity current step through f3d_step
 * only: f3d_batch_create refuses them, and f3d_batch_step refuses a context
 * that turned either on after joining. */
typedef struct f3d_batch f3d_batch;
int f3d_batch_create(f3d_batch **batch, f3d_ctx *const *ctxs, int n);
void f3d_batch_destroy(f3d_batch *batch);
/* tile shape of the batched flux kernel: 58x10, 58x18 or 58x26, 1 or 2
 * slabs per workgroup (default 58x18, 2) */
int f3d_batch_set_tile(f3d_batch *batch, int tx, int ty, int slabs);
int f3d_batch_get_tile(const f3d_batch *batch, int *tx, int *ty, int *slabs);
/* capture steps in HIP graphs (default on) or launch them directly */
int f3d_batch_use_graph(f3d_batch *batch, int on);
int f3d_batch_step(f3d_batch *batch, int nsteps);
int f3d_batch_sync(f3d_batch *batch);
/* HIP-event timing of the batch's persistent flux launches, as f3d_timing /
 * f3d_kernel_time for a context: while it is on, steps launch directly
 * instead of from graphs (the same kernels, one launch each) */
int f3d_batch_timing(f3d_batch *batch, int enable);
int f3d_batch_kernel_time(f3d_batch *batch, int reset, double *avg_ms, long *count);
/* time each candidate tile shape over `trials` flux steps of the batch (no
 * state changes) and keep the fastest; outputs may be null */
int f3d_batch_autotune(f3d_batch *batch, int trials, int *tx, int *ty, int *slabs,
                       double *best_ms);
/* the same for one context and its f3d_step */
int f3d_autotune(f3d_ctx *ctx, int trials, int *tx, int *ty, int *slabs, double *best_ms);

/* Free and total device memory (admission control of the serving loop).  No
 * counterpart in the reference. */
int f3d_mem_info(long long *free_bytes, long long *total_bytes);

/* No counterpart in the reference: a plain device copy of nbytes (read) +
 * nbytes (write), repeated reps times; *avg_ms is the average launch time.
 * It is the known-size kernel against which HBM counters (rocprofv3
 * FETCH_SIZE / WRITE_SIZE) and the achievable bandwidth are calibrated. */
int f3d_hbm_copy(long long nbytes, int reps, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif
