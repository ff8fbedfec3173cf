/*
 * FP64 CPU oracle for MagLua's per-timestep spin-dynamics hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or as the timed CPU baseline.  The product (maglua_amd/, libmsd.so) never
 * links, loads or calls it.
 *
 * Each function restates the reference algorithm it names (paths relative to
 * the MagLua source tree).  Pinning of each function (golden vectors, closed
 * forms, brute force) is listed in the header comment of msd_oracle.cpp.
 */
#ifndef MSD_ORACLE_H
#define MSD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based RNG (Philox4x32-10, Salmon et al., SC'11) ---- */
void ora_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* 3 standard normals for (seed, step, site): the build's thermal-noise stream */
void ora_thermal_normals(uint64_t seed, uint64_t step, int64_t site, double out[3]);
/* unit spins uniform on the sphere for sites [0,n), keyed by seed */
void ora_random_unit_spins(uint64_t seed, int64_t n, double* x, double* y, double* z);

/* ---- FFT (unnormalised, sign -1 = forward, as FFTW) ---- */
int  ora_fft1d(double* data, int n, int sign);                  /* interleaved re,im, in place */
void ora_dft1d_naive(const double* in, double* out, int n, int sign);
int  ora_fft2d_layers(double* data, int nx, int ny, int nz, int sign);

/* ---- dipole tensor (periodic in-plane lattice sum) ---- */
int ora_dipole_tensor(int nx, int ny, int nz, const int gmax[4], const double ABC[9],
                      double* XX, double* XY, double* XZ, double* YY, double* YZ, double* ZZ);

/* ---- object model ---- */
typedef struct ora_sys ora_sys;
typedef struct ora_ex  ora_ex;
typedef struct ora_ani ora_ani;
typedef struct ora_th  ora_th;
typedef struct ora_lr  ora_lr;
typedef struct ora_llg ora_llg;

ora_sys* ora_sys_new(int nx, int ny, int nz);
void     ora_sys_free(ora_sys*);
ora_sys* ora_sys_copy(const ora_sys*);
void     ora_sys_copy_spins_to(const ora_sys* src, ora_sys* dst);
void     ora_sys_set_spins(ora_sys*, const double* x, const double* y, const double* z);
void     ora_sys_set_spin(ora_sys*, int site, double sx, double sy, double sz, double m);
void     ora_sys_get_spins(const ora_sys*, double* x, double* y, double* z, double* ms);
void     ora_sys_set_params(ora_sys*, double alpha, double gamma, double dt, double time);
void     ora_sys_get_params(const ora_sys*, double out[4]);  /* alpha, gamma, dt, time */
void     ora_sys_set_site_alpha(ora_sys*, const double* a);  /* NULL clears */
void     ora_sys_set_site_gamma(ora_sys*, const double* g);
int      ora_sys_slot(const ora_sys*, const char* name);      /* -1 if unregistered */
int      ora_sys_get_field(const ora_sys*, const char* name, double* hx, double* hy, double* hz);
int      ora_sys_set_field(ora_sys*, const char* name, const double* hx, const double* hy, const double* hz);
void     ora_sys_reset_fields(ora_sys*);
void     ora_sys_sum_fields(ora_sys*);
void     ora_sys_net_moment(const ora_sys*, double out[8]);
double   ora_sys_field_dot(const ora_sys*, const char* name);

ora_ex*  ora_ex_new(int nx, int ny, int nz);
void     ora_ex_free(ora_ex*);
void     ora_ex_set_pbc(ora_ex*, int px, int py, int pz);
void     ora_ex_set_normalize(ora_ex*, int yes);
void     ora_ex_set_scale(ora_ex*, double s);
int      ora_ex_add(ora_ex*, int ax, int ay, int az, int bx, int by, int bz, double J); /* 0-based coords */
void     ora_ex_add_path(ora_ex*, int from_site, int to_site, double J);
void     ora_ex_add_paths(ora_ex*, const int* from, const int* to, const double* J, int count);
int      ora_ex_apply(ora_ex*, ora_sys*);

ora_ani* ora_ani_new(int nx, int ny, int nz);
void     ora_ani_free(ora_ani*);
void     ora_ani_set_scale(ora_ani*, double s);
int      ora_ani_add(ora_ani*, int site, double ax, double ay, double az, double K);
void     ora_ani_add_many(ora_ani*, const int* site, const double* axes, const double* K, int count);
int      ora_ani_apply(ora_ani*, ora_sys*);

ora_th*  ora_th_new(int nx, int ny, int nz);
void     ora_th_free(ora_th*);
void     ora_th_set_temperature(ora_th*, double T);
void     ora_th_set_scale(ora_th*, double s);
void     ora_th_scale_site(ora_th*, int site, double v);
/* explicit normals (3n, SoA x|y|z) -> field; the oracle's own stream if normals==NULL */
int      ora_th_apply(ora_th*, ora_sys*, const double* normals, uint64_t seed, uint64_t step);

ora_lr*  ora_lr_new(int nx, int ny, int nz);
void     ora_lr_free(ora_lr*);
void     ora_lr_set_strength(ora_lr*, double g);
void     ora_lr_set_scale(ora_lr*, double s);
void     ora_lr_set_truncation(ora_lr*, const int gmax[4]);
void     ora_lr_set_unit_cell(ora_lr*, const double ABC[9]);
int      ora_lr_set_tensor(ora_lr*, int which, const double* t);   /* [nx*ny*(2nz-1)] */
int      ora_lr_get_tensor(ora_lr*, int which, double* t);
int      ora_lr_apply(ora_lr*, ora_sys*);

/* 3D FFT and LongRange3D (periodic, or open on the doubled grid) */
int      ora_fft3d(double* data, int nx, int ny, int nz, int sign);
int      ora_dipole_tensor3d_periodic(int nx, int ny, int nz, int R,
                                      double* XX, double* XY, double* XZ, double* YY, double* YZ, double* ZZ);
int      ora_dipole_tensor3d_open(int nx, int ny, int nz,
                                  double* XX, double* XY, double* XZ, double* YY, double* YZ, double* ZZ);
/* open-boundary dipole field at chosen sites by a direct real-space sum (no FFT):
 * out[3k..3k+2] = g sum_{s != t} Gamma(r_t - r_s) m(s), t = sites[k] */
int      ora_dipole_direct_open(int nx, int ny, int nz, const double* x, const double* y, const double* z,
                                int nsites, const int* sites, double g, double* out);
typedef struct ora_lr3 ora_lr3;
ora_lr3* ora_lr3_new(int nx, int ny, int nz, int open);
void     ora_lr3_free(ora_lr3*);
void     ora_lr3_set_strength(ora_lr3*, double g);
int      ora_lr3_set_tensor(ora_lr3*, int which, const double* t);   /* 0..8 = XX XY XZ YX YY YZ ZX ZY ZZ */
int      ora_lr3_apply(ora_lr3*, ora_sys*);

/* LongRange2D: T [9][nz][nz][ny][nx] (XX..ZZ, labels [dest][src]); apply reads [src][dest] */
typedef struct ora_lr2d ora_lr2d;
ora_lr2d* ora_lr2d_new(int nx, int ny, int nz);
void      ora_lr2d_free(ora_lr2d*);
void      ora_lr2d_set_scale(ora_lr2d*, double s);
void      ora_lr2d_set_layer_scales(ora_lr2d*, const double* src_moment, const double* dest_field);
int       ora_lr2d_set_tensor(ora_lr2d*, int dest, int src, int ab, const double* a);
int       ora_lr2d_get_tensor(ora_lr2d*, int dest, int src, int ab, double* a);
int       ora_lr2d_apply(ora_lr2d*, ora_sys*, const char* slot);
/* Magnetostatics2D: prism-prism demag element N_ab (ab = 0..8), and makeData's tensors
 * T [9][source layer][destination layer][ny][nx] (-N summed over images) */
double    ora_mag2d_element(int ab, double X, double Y, double Z, const double d1[3], const double d2[3]);
int       ora_mag2d_tensor(int nx, int ny, int nz, const double* ABC, const double* grain, const int trunc[4], double* T);
int       ora_dipole2d_tensor(int nx, int ny, int nz, const double* ABC, const double* off, const int trunc[4],
                              int physical, double* T);

ora_llg* ora_llg_new(void);
void     ora_llg_free(ora_llg*);
void     ora_llg_set_flags(ora_llg*, int thermal_only_first_term, int disable_renorm);
void     ora_llg_set_quaternion(ora_llg*, int yes);   /* LLG.Quaternion (llgquat.cpp) */
int      ora_llg_apply(ora_llg*, const ora_sys* from, double scale, const ora_sys* dmdt, ora_sys* to, int advance);

/* ---- integrators (PredictorCorrector.lua / RungeKutta.lua) over a fixed operator set ---- */
typedef struct ora_sim ora_sim;
ora_sim* ora_sim_new(ora_sys* ss, ora_ex* ex, ora_ani* ani, ora_lr* lr, ora_th* th, ora_llg* llg);
void     ora_sim_free(ora_sim*);
void     ora_sim_set_seed(ora_sim*, uint64_t seed);
void     ora_sim_set_lr3(ora_sim*, ora_lr3* lr3);
void     ora_sim_set_lr2d(ora_sim*, ora_lr2d* lr2d, const char* slot);
int      ora_sim_calc_fields(ora_sim*, ora_sys* target);
int      ora_sim_heun(ora_sim*, int nsteps);                  /* one PC corrector pass per step */
int      ora_sim_rk(ora_sim*, const char* name, int nsteps);  /* RK1, RK2, RK3, K3, RK4, RK4_38, RK6, BS3 */
void     ora_sim_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
