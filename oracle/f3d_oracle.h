/*
 * FALL3D-7.0 transport core: FP64 CPU restatement (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker
 * or the timed CPU baseline.  The product (fall3d_amd/) never links it.
 *
 * Array conventions follow the reference (Sources/Master.f90:92-136):
 *   c    (0:nx+1,0:ny+1,0:nz+1,nc)  Fortran column-major, one ghost cell
 *   met  (nx,ny,nz)                 vx vy vz rho rkhor rkh1 rkver tempe vptemp
 *   vset (nx,ny,nz,nc)   vdep (nx,ny,nc)   cdry/cwet/cload (nx,ny,0:nc)
 *   dz   (0:nz)          zlayer (nz)       2-D fields (nx,ny)
 * The serial build is restated (nz_loc = nz, iz_loc = 1, ic_loc = 1).
 */
#ifndef F3D_ORACLE_H
#define F3D_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

void or_setbcc(double *c, int nx, int ny, int nz, const double *vx, const double *vy,
               const double *vz, const double *vset, const double *vdep,
               const double *rkver, const double *dz);
void or_advctzc(double *c, int nx, int ny, int nz, const double *vz, const double *vset,
                const double *dz, double dt);
void or_advctx(double *c, int nx, int ny, int nz, const double *vx, double dx, double dt);
void or_advcty(double *c, int nx, int ny, int nz, const double *vy, double dy, double dt);
void or_diffz(double *c, int nx, int ny, int nz, const double *dz, double dt,
              const double *rkver, const double *rho);
void or_diffx(double *c, int nx, int ny, int nz, double dx, double dt, const double *rkh,
              const double *rho);
void or_diffy(double *c, int nx, int ny, int nz, double dy, double dt, const double *rkh,
              const double *rho);
void or_accum(const double *c, double *cdry, int ic, int nx, int ny, int nz, double dt,
              const double *vset, const double *vdep);
void or_clip(double *c, long n);
void or_cmass2d(const double *c, const double *vx, const double *vy, const double *vz,
                const double *vset, int nx, int ny, int nz, int nc, double dx, double dy,
                const double *dz, double dt, double *outmass);
void or_cmass3d(const double *c, const double *cload, const double *cdry, const double *cwet,
                int nx, int ny, int nz, int nc, double dx, double dy, const double *dz,
                double *massvol, double *massdep, double *massdry, double *masswet);
void or_getc0(double *cdry, double *cwet, int nx, int ny, int np);

double or_finddt(int nx, int ny, int nz, int nc, const double *dz, double dX1, double dX2,
                 const double *vx, const double *vy, const double *vz, const double *rkh1,
                 const double *rkhor, const double *rkver, const double *vset, double *dtc);

/* modkv: 0 CONSTANT, 1 SIMILARITY, 2 SURFACE_LAYER; modkh: 0 CONSTANT, 1 RAMS, 2 CMAQ.
 * lonlat: 0 UTM, 1 LON-LAT (only selects Khn0 in CMAQ). */
void or_kappa3(int modkv, int modkh, int nx, int ny, int nz, const double *dz,
               const double *zlayer, const double *vx, const double *vy, const double *vptemp,
               const double *zi, const double *rl, const double *ustar, const double *Hm1,
               double dX1, double dX2, double dx, double dy, double rkv0, double rkh0,
               double Csh, int lonlat, double *rkver, double *rkhor);

int or_get_gama(double diam, double sphe, double *gama);
void or_setpsi(double *psi, const double *sphe, const double *diam, int modv, int nc);
/* Per-class constants of the Ganser model that depend only on (diam, sphe, psi). */
int or_ganser_consts(double diam, double sphe, double psi, double *rk1, double *rk2);
int or_vsettl(double diam, double rhop, double rhoa, double visc, double *vset, int model,
              double psi, double sphe, double *cd_state);
/* Returns 0, or -(1+cell index) of the first non-converged cell. */
long or_setvset(int nx, int ny, int nz, int np, const double *tempe, const double *rho,
                const double *diam, const double *rhop, const double *psi, const double *sphe,
                int modv, double *vset, double *cd_state);
void or_meteo_scale(int nx, int ny, int nz, const double *Hm1, double *vx, double *rho,
                    const double *rkhor, double *rkh1);

/* setvdrydep.f90:54-264 (tempe1, rho1: first level, before the Hm1
 * scaling; vdep(nx,ny,nc) in/out: classes above 100 um keep theirs) */
void or_setvdrydep(int nx, int ny, int nc, double zlayer1, const double *diam,
                   const double *ldu, const double *znt, const double *rl, const double *ustar,
                   const double *tempe1, const double *rho1, double *vdep);

/* corvver.f90:16-36 on the scaled vx (LON-LAT grids) */
void or_corvver(int nx, int ny, int nz, const double *Hm1, const double *topg, double dX1,
                double dX2, const double *vx, const double *vy, double *vz);

void or_setsrc(double *c, int nx, int ny, int nz, int nc, int nsou, const int *isrc,
               const int *jsrc, const int *ksrc, const double *tmrat, const double *dz,
               const double *zlayer, const double *Hm1, double dX1, double dX2, double dt);

/* printres_nc.f90:150-241 output products: load (nx,ny,0:nc) kg/m2,
 * thick (nx,ny) cm, colmass (nx,ny) g/m2; c (0:nx+1,0:ny+1,0:nz+1,nc) */
void or_standard_atm(double z, double *p, double *T);
void or_output_concentration(int nx, int ny, int nz, int nc, const double *c, const double *Hm1,
                             const double *zlayer, const double *topg, const double *p, int nfl,
                             const double *fl, double *cgrnd, double *con, int asl, double *cfl);
void or_output_products(int nx, int ny, int nz, int nc, const double *c, const double *cdry,
                        const double *cwet, const double *Hm1, const double *rhop,
                        const double *dz, double *load, double *thick, double *colmass);

/* wetdep.f90:28-50 (ic 1-based; cwet(nx,ny,0:nc)) */
void or_wetdep(double *c, double *cwet, int ic, int nx, int ny, int nz, double dt,
               const double *diam, const double *zlayer, const double *dz, const double *zi,
               const double *prate);

/* sizloop.f90:31-83 for all classes; nthreads > 1 splits the class loop
 * over POSIX threads (classes are independent until the clip).  prate
 * non-null: wet deposition into cwet (sizloop.f90:65-67). */
void or_sizloop(double *c, double *cdry, int nx, int ny, int nz, int nc, int iiter, double dt,
                const double *vx, const double *vy, const double *vz, const double *vset,
                const double *vdep, const double *rkver, const double *rkhor,
                const double *rkh1, const double *rho, const double *dz, double dX1,
                double dX2, double *outmass, int nthreads, double *cwet, const double *diam,
                const double *zlayer, const double *zi, const double *prate);

/* Gravity current: setdivu.f90:16-30; addradialwind.f90:47-173 (prm =
 * {c_flow_rate, lambda_grav, k_entrain, brunt_vaisala, centre x, centre y,
 * xorigr, yorigr, dx, dy}; returns 0 when the source is off, out: 9
 * diagnostics or null); divcorr.f90:22-38 (returns corrdiv). */
void or_setdivu(int nx, int ny, int nz, const double *vx, const double *vy, const double *vz,
                const double *Hm1, double dX1, double dX2, const double *dz, double *divu);
int or_addradialwind(int nx, int ny, int nz, double *vx, double *vy, const double *Hm1,
                     double dX1, double dX2, const double *zlayer, const double *prm,
                     double time_eruption, double mass_flow_rate, double h_tot, double *out);
double or_divcorr(double *c, int nx, int ny, int nz, int nc, const double *divu, double dt,
                  double erumass, double massvol, double massdep, double outmass);

#ifdef __cplusplus
}
#endif
#endif
