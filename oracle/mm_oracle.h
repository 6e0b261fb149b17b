/*
 * mm_oracle.h -- CPU oracle for the finite-difference field + LLG hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the CPU path of
 * MicroMagnetic.jl (mesh, interaction kernels, FFT demag, LLG right-hand side,
 * RK4 / Heun / Dormand-Prince / Bogacki-Shampine steppers).  It exists to check
 * the HIP library and to time a CPU baseline.  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it; the product never does.
 *
 * Conventions follow the reference: spins are AoS [mx,my,mz] per cell, x
 * fastest, then y, then z (src/micro/mesh.jl:24-29); mu0*Ms is stored
 * (src/sim.jl:138); every array is double precision.
 */
#ifndef MM_ORACLE_H
#define MM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

/* pbc bit mask: 1 = x, 2 = y, 4 = z */
orc_sim *orc_create(int nx, int ny, int nz, double dx, double dy, double dz, int pbc);
void orc_destroy(orc_sim *s);
int orc_num_threads(int n); /* set OpenMP threads (n>0); returns current count */

void orc_neighbours(const orc_sim *s, int32_t *ngbs_out /* 6*N, 1-based, -1 none */);

void orc_set_Ms(orc_sim *s, const double *Ms);          /* A/m; stored as mu0*Ms */
void orc_set_mu0Ms(orc_sim *s, const double *mu0Ms);    /* raw (tests that poke mu0_Ms) */
void orc_set_pins(orc_sim *s, const uint8_t *pins);
int orc_init_m0(orc_sim *s, const double *m, int norm); /* -1 on NaN */
void orc_set_m(orc_sim *s, const double *m);            /* raw copy, no normalisation */
void orc_get_m(const orc_sim *s, double *m);

/* interaction terms; each returns its index in insertion order */
int orc_add_exch(orc_sim *s, double Ax, double Ay, double Az);
int orc_add_exch_spatial(orc_sim *s, const double *A);
int orc_add_dmi_bulk(orc_sim *s, double Dx, double Dy, double Dz);
int orc_add_dmi_bulk_spatial(orc_sim *s, const double *Dx, const double *Dy, const double *Dz);
int orc_add_dmi_interfacial(orc_sim *s, const double *D);
int orc_add_anis(orc_sim *s, const double *Ku, double ax, double ay, double az);
int orc_add_anis_spatial(orc_sim *s, const double *Ku, const double *axes /* 3N, normalised */);
int orc_add_cubic_anis(orc_sim *s, const double *Kc, const double a1[3], const double a2[3]);
int orc_add_zeeman(orc_sim *s, const double *H /* 3N */);
int orc_add_demag(orc_sim *s, int Nx, int Ny, int Nz);
int orc_add_demag_direct(orc_sim *s, int Nx, int Ny, int Nz);
int orc_add_thermal(orc_sim *s, const double *Temp, double T0, uint64_t seed);
int orc_add_sot(orc_sim *s, const double *aj, double bj, double px, double py, double pz);
int orc_add_stt_slonczewski(orc_sim *s, const double *J, double beta, double Lambda, double xi, double P,
                            double px, double py, double pz);
int orc_add_stt_zhangli(orc_sim *s, const double *J /* 3N */, double b, const double *xi, int xi_factor);
int orc_set_term_time_fn(orc_sim *s, int term, double (*fn)(double, void *), void *user);
void orc_set_alpha(orc_sim *s, const double *alpha /* N or NULL */);
int orc_add_torque(orc_sim *s, void (*fn)(double, const double *, double *, void *), void *user);
int orc_add_magnetoelastic(orc_sim *s, int cubic, const double *data /* 6N */, double lambda_s, double B1,
                           double B2);
int orc_add_sahe_torque(orc_sim *s, const double *sigma_s, const double *sigma_sa1, const double *sigma_sa2,
                        const double *a1, const double *a2, const double *a3, double beta /* 3N AoS each */);
int orc_add_twin_mono_anis(orc_sim *s, const double K[6], const double *axis_a, const double *axis_b,
                           const double *axis_u /* 3N each */);
int orc_add_exch_int(orc_sim *s, const double *J /* nx*ny */, int k1, int k2);
int orc_add_dmi_int(orc_sim *s, double Dx, double Dy, double Dz, int k1, int k2);
int orc_add_hex_anis(orc_sim *s, double K1, double K2, double K3);
int orc_add_zeeman_time(orc_sim *s, const double *H /* 3N */, void (*fn)(double, double *, void *), void *user);
int orc_num_terms(const orc_sim *s);

/* the compact real-space tensor of a demag term: 6*N doubles, order xx,yy,zz,xy,xz,yz */
int orc_demag_tensor(const orc_sim *s, int term, double *out);

/* field evaluation */
void orc_effective_field(orc_sim *s, const double *spin, double t, double *field_out);
void orc_term_field(orc_sim *s, int term, const double *spin, double t,
                    double *field_out, double *energy_out);
void orc_term_energies(orc_sim *s, const double *spin, double t, double *E_out);

/* LLG driver */
void orc_set_llg(orc_sim *s, double alpha, double gamma, int precession);
void orc_llg_rhs(orc_sim *s, const double *spin, double t, double *dmdt_out);

/* integrators (src/integrator/ rk.jl, heun.jl, adaptive_rk*.jl); kind: 0 RK4, 1 Heun, 2 DOPRI54, 3 BS23 */
void orc_set_integrator(orc_sim *s, int kind, double step, double tol);
int orc_advance_step(orc_sim *s);       /* one (accepted) step; 0 ok, <0 failure */
int orc_run_until(orc_sim *s, double t_end);
void orc_integrator_state(const orc_sim *s, double *t, double *step, double *step_next,
                          int64_t *nsteps, int64_t *nfevals);
int orc_relax(orc_sim *s, int max_steps, double stopping_dmdt, int *steps_done,
              double *max_dmdt_out);

/* standalone complex FFT used by the demag (exposed for its own test) */
void orc_fft(double *data /* interleaved re,im */, int n, int sign);

/* counter-based normal variates used for the thermal field */
void orc_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out);
void orc_philox_normal(uint64_t seed, uint64_t step, uint64_t n, double *out);

#ifdef __cplusplus
}
#endif
#endif
