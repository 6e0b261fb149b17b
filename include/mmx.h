/*
 * mmx.h -- C ABI of libmmx, the MI355X (gfx950) finite-difference field + LLG
 * hot path that sits behind MicroMagnetic.jl's API (Sim / FDMesh / add_exch /
 * add_dmi / add_anis / add_cubic_anis / add_zeeman / add_demag /
 * effective_field / run_until / run_sim / relax).
 *
 * Conventions (the reference's, src/micro/mesh.jl:24-29):
 *   - cell I = (k-1) nx ny + (j-1) nx + i, x fastest; vectors are AoS
 *     [mx,my,mz] per cell, so a vector array holds 3N doubles;
 *   - host pointers are borrowed for the duration of the call only;
 *   - every call is synchronous at the boundary and returns 0 (or a term index
 *     >= 0 for the add_* calls) on success and a negative MMX_E* status on
 *     failure; mmx_last_error() then describes it (thread-local);
 *   - a context owns its device memory and one HIP stream; it is not thread-
 *     safe, distinct contexts are independent.
 */
#ifndef MMX_H
#define MMX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MMX_OK 0
#define MMX_EINVAL (-1)  /* invalid argument (error(...) / ArgumentError in the reference) */
#define MMX_EHIP (-2)    /* HIP runtime error */
#define MMX_ENAN (-3)    /* NaN error norm: src/integrator/adaptive_rk.jl:112-117 */
#define MMX_ETRIES (-4)  /* more than 50 tries: src/integrator/adaptive_rk.jl:119-122 */
#define MMX_ESTATE (-5)  /* call not valid in the current state (e.g. run_until with RK4) */
#define MMX_ENOMEM (-6)  /* device allocation failed */

#define MMX_F64 64
#define MMX_F32 32

#define MMX_DMI_BULK 0
#define MMX_DMI_INTERFACIAL 1

#define MMX_RK4 0     /* "RungeKutta"   src/integrator/rk.jl */
#define MMX_HEUN 1    /* "Heun"         src/integrator/heun.jl */
#define MMX_DOPRI54 2 /* "DormandPrince" src/integrator/adaptive_rk_steps.jl:49 */
#define MMX_BS23 3    /* "BS23"         src/integrator/adaptive_rk_steps.jl:11 */
#define MMX_SD 4      /* driver "SD": steepest descent, src/sd.jl (step and tol unused) */

typedef struct mmx_ctx mmx_ctx;

/* FDMesh(; dx, dy, dz, nx, ny, nz, pbc): src/micro/mesh.jl:69-98 */
typedef struct mmx_mesh_desc {
    int32_t nx, ny, nz;
    int32_t pbc;       /* bit 0: x periodic, bit 1: y, bit 2: z */
    double dx, dy, dz; /* m */
    int32_t precision; /* MMX_F64 or MMX_F32 (set_precision, src/MicroMagnetic.jl:35) */
    int32_t device;    /* HIP device ordinal */
} mmx_mesh_desc;

/* Sim(mesh): src/sim.jl:48-107 (LLG driver, DormandPrince integrator, alpha 0.1, gamma 2.21e5) */
int mmx_create(mmx_ctx **out, const mmx_mesh_desc *mesh);
int mmx_destroy(mmx_ctx *ctx);
const char *mmx_last_error(void);

/* set_Ms(sim, Ms): src/sim.jl:129-143 -- Ms in A/m, N values, stored as mu0*Ms */
int mmx_set_Ms(mmx_ctx *ctx, const double *Ms);
/* set_pinning(sim, ids): src/sim.jl:172-177 -- N bytes, nonzero = pinned */
int mmx_set_pinning(mmx_ctx *ctx, const uint8_t *pins);
/* init_m0(sim, m0; norm): src/sim.jl:278-309 -- 3N values; normalised if norm,
 * zeroed where Ms == 0; MMX_EINVAL on NaN */
int mmx_init_m0(mmx_ctx *ctx, const double *m, int norm);
/* Array(sim.spin) */
int mmx_get_m(mmx_ctx *ctx, double *m);

/* add_exch(sim, A): src/micro/add_field.jl:233-263.  A != NULL: per-cell A
 * (SpatialExchange, harmonic mean); else uniform (Ax, Ay, Az) (UniformExchange). */
int mmx_add_exch(mmx_ctx *ctx, const double *A, double Ax, double Ay, double Az);
/* add_dmi(sim, D; type): src/micro/add_field.jl:293-345.
 * bulk: Dx/Dy/Dz == NULL -> BulkDMI(Dx0, Dy0, Dz0); else SpatialVectorBulkDMI (all three arrays).
 * interfacial: Dx = per-cell D, or NULL with the uniform value Dx0. */
int mmx_add_dmi(mmx_ctx *ctx, int type, const double *Dx, const double *Dy, const double *Dz, double Dx0,
                double Dy0, double Dz0);
/* add_anis(sim, Ku; axis): src/micro/add_field.jl:430-465.  Ku per cell (or NULL
 * with the uniform Ku0); axes = 3N per-cell axis (normalised here) or NULL with (ax, ay, az). */
int mmx_add_anis(mmx_ctx *ctx, const double *Ku, double Ku0, const double *axes, double ax, double ay,
                 double az);
/* add_cubic_anis(sim, Kc; axis1, axis2): src/micro/add_field.jl:541-574 */
int mmx_add_cubic_anis(mmx_ctx *ctx, const double *Kc, double Kc0, const double axis1[3],
                       const double axis2[3]);
/* add_zeeman(sim, H0): src/micro/add_field.jl:11-48 -- 3N per-cell field or NULL with (Hx,Hy,Hz) */
int mmx_add_zeeman(mmx_ctx *ctx, const double *H, double Hx, double Hy, double Hz);
/* update_zeeman(sim, H0): src/micro/add_field.jl:60-100 */
int mmx_update_zeeman(mmx_ctx *ctx, int term, const double *H, double Hx, double Hy, double Hz);
/* add_hex_anis(sim; K1, K2, K3): src/micro/add_field.jl:590-606, hexagonal_anisotropy_kernel!
 * src/micro/kernels.jl:142-167 -- E = K1 (1-mz^2) + K2 (1-mz^2)^2 + K3 Re((mx + i my)^6). */
int mmx_add_hex_anis(mmx_ctx *ctx, double K1, double K2, double K3);
/* add_torque(sim, fun): src/micro/add_field.jl:998-1006, TorqueField field.jl:282-293,
 * torque_kernel! src/micro/kernels.jl:639-651 -- at every evaluation fn(t, m, y, user) gets
 * the state m (3N AoS, the stage vector inside the steppers) and writes the torque y
 * (3N AoS, zeroed before the call); the term is H = (1/gamma) m x y.  Runs on the host:
 * every evaluation then copies the state out and the torque in. */
int mmx_add_torque(mmx_ctx *ctx, void (*fn)(double t, const double *m, double *y, void *user), void *user);
/* add_magnetoelastic(sim; model, lambda_s, B1, B2, sigma / strain) (add_mel):
 * src/micro/add_field.jl:1367-1405, kernels src/micro/kernels.jl:742-845 -- data holds 6N
 * values, cell by cell in Voigt order (xx, yy, zz, xy, xz, yz) as the reference's (6, nx, ny, nz)
 * array; MMX_MEL_TENSOR: stress and lambda_s; MMX_MEL_CUBIC: strain and B1, B2. */
#define MMX_MEL_TENSOR 0
#define MMX_MEL_CUBIC 1
int mmx_add_magnetoelastic(mmx_ctx *ctx, int model, const double *data, double lambda_s, double B1, double B2);
/* add_sahe_torque(sim, sigma_s; sigma_sa1, sigma_sa2, a1, a2, a3, beta): src/micro/add_field.jl:892-936,
 * sahe_torque_kernel! src/micro/kernels.jl:589-613 -- six 3N AoS vector fields; the torque enters as
 * H = (beta sigma + m x sigma)/gamma with sigma = sigma_s x a1 - (m.a2)^2 sigma_sa1 x a1
 * + (m.a3) sigma_sa2 .* m (no energy). */
int mmx_add_sahe_torque(mmx_ctx *ctx, const double *sigma_s, const double *sigma_sa1, const double *sigma_sa2,
                        const double *a1, const double *a2, const double *a3, double beta);
/* add_twin_mono_anis(sim; Ka, Kb, Kaa, Kbb, Kab, Ku, axis_a, axis_b, axis_u111):
 * src/micro/add_field.jl:629-660, twin_monoclinic_anisotropy_kernel! src/micro/kernels.jl:169-225
 * -- K = {Ka, Kb, Kaa, Kbb, Kab, Ku}; three per-cell axes, 3N values each (AoS),
 * normalised here as the reference normalises them. */
int mmx_add_twin_mono_anis(mmx_ctx *ctx, const double K[6], const double *axis_a, const double *axis_b,
                           const double *axis_u111);
/* add_exch_int(sim, J; k1, k2): src/micro/add_field.jl:789-816, interlayer_exch_kernel!
 * src/micro/kernels.jl:496-536 -- J holds nx*ny values (one per column), layers 1-based,
 * k2 = -1 for the top layer; H(k1) = J/(mu0Ms dz) m(k2) and the other way round. */
int mmx_add_exch_int(mmx_ctx *ctx, const double *J, int k1, int k2);
/* add_dmi_int(sim, D; k1, k2): src/micro/add_field.jl:835-860, interlayer_dmi_kernel!
 * src/micro/kernels.jl:539-584 -- H(k1) = D x m(k2)/(mu0Ms dz), H(k2) = -D x m(k1)/(mu0Ms dz). */
int mmx_add_dmi_int(mmx_ctx *ctx, double Dx, double Dy, double Dz, int k1, int k2);
/* add_zeeman(sim, H0, ft): src/micro/add_field.jl:161-200, TimeZeeman field.jl:12-33,
 * time_zeeman_kernel! src/micro/kernels.jl:39-49 -- H = H0 .* (fx, fy, fz)(t); fn writes the
 * three factors (a scalar time_fun gives three equal ones) and runs on the host at every
 * evaluation time.  3N per-cell H0 or NULL with (Hx,Hy,Hz). */
int mmx_add_zeeman_time(mmx_ctx *ctx, const double *H, double Hx, double Hy, double Hz,
                        void (*fn)(double t, double *f, void *user), void *user);
/* add_demag(sim; Nx, Ny, Nz): src/micro/add_field.jl:403-419 -> init_demag, src/micro/demag.jl:30-100.
 * The FP64 Newell tensor is built on the host; its transform runs on the GPU in FP64. */
int mmx_add_demag(mmx_ctx *ctx, int Nx, int Ny, int Nz);
/* add_thermal_noise(sim, T; T0): src/micro/add_field.jl:734-771.  T per cell (or NULL
 * with T_uniform).  The noise is Philox4x32-10 keyed by (seed, integrator step, index). */
int mmx_add_thermal_noise(mmx_ctx *ctx, const double *T, double T_uniform, double T0, uint64_t seed);
/* Spin torques enter H_eff as further terms in insertion order and carry no
 * energy (compute_system_energy skips them, src/micro/field.jl:520-528).
 * add_sot(sim, aj, bj, p): src/micro/add_field.jl:951-966, df_torque_kernel!
 * src/micro/kernels.jl:619-633 -- aj per cell (or NULL with aj0), p not normalised. */
int mmx_add_sot(mmx_ctx *ctx, const double *aj, double aj0, double bj, double px, double py, double pz);
/* add_stt(sim; model=:slonczewski, ...): src/micro/add_field.jl:1195-1227,
 * slonczewski_torque_kernel! src/micro/kernels.jl:657-674.  J per cell (or NULL with
 * J0); beta = hbar/(mu0 Ms e tf) as the reference computes it. */
int mmx_add_stt_slonczewski(mmx_ctx *ctx, const double *J, double J0, double beta, double Lambda, double xi,
                            double P, double px, double py, double pz);
/* add_stt(sim; model=:zhang_li, ...): src/micro/add_field.jl:1150-1192,
 * zhangli_torque_kernel! src/micro/kernels.jl:680-739.  J = 3N AoS current density (or
 * NULL with (Jx,Jy,Jz)), b the coefficient (b, or P mu_B/(e Ms)); xi per cell (or NULL
 * with xi0); xi_factor = 1 applies 1/(1 + xi^2) as the (P, Ms) form does. */
int mmx_add_stt_zhangli(mmx_ctx *ctx, const double *J, double Jx, double Jy, double Jz, double b,
                        const double *xi, double xi0, int xi_factor);
/* The ft / ufun time profile of a Slonczewski or Zhang-Li term (field.jl:335,351),
 * or time_D of a bulk DMI term, which makes it add_dmi(sim, D, ft) (add_field.jl:360-393,
 * TimeBulkDMI / TimeSpatialBulkDMI field.jl:145-193: D scaled by time_D(t)):
 * called on the host at every evaluation time (stage times inside the steppers);
 * fn = NULL restores the constant 1. */
int mmx_set_term_time_fn(mmx_ctx *ctx, int term, double (*fn)(double t, void *user), void *user);
int mmx_num_terms(mmx_ctx *ctx);

/* effective_field(sim, spin, t): src/micro/field.jl:496-505 -- sim.field (3N) */
int mmx_effective_field(mmx_ctx *ctx, double t, double *H);
/* effective_field(term, sim, spin, t): term.field (3N) and term.energy (N), either nullable */
int mmx_term_field(mmx_ctx *ctx, int term, double t, double *H, double *E);
/* compute_system_energy: src/micro/field.jl:520-536 -- sum(term.energy) per term */
int mmx_compute_energies(mmx_ctx *ctx, double t, double *E);

/* No counterpart in the reference: mmx_compute_energies for callers that cannot
 * allocate (code under HIP graph capture, serving loops).  Writes the per-term
 * energies of the current state to E_dev, a device array of mmx_num_terms
 * doubles, on the context's stream without a host synchronisation (call
 * mmx_synchronize before reading).  Calls no hipMalloc: the per-block partial
 * sums live in a static device array of 65536 doubles per GPU; a mesh whose
 * evaluation needs more blocks than that is refused with MMX_EINVAL.  Calls that
 * overlap on one GPU share that array and race.  Contexts with an add_torque
 * term are refused (its host callback allocates). */
int mmx_compute_energies_noalloc(mmx_ctx *ctx, double t, double *E_dev);

/* LLG driver fields alpha, gamma, precession: src/driver.jl:15-21 */
int mmx_set_llg(mmx_ctx *ctx, double alpha, double gamma, int precession);
/* Sim(mesh; driver="SpatialLLG") + set_alpha(sim, alpha): src/driver.jl:32-38,149-152,
 * src/sim.jl:213-222, spatial_llg_rhs_kernel! src/llg/llg.jl:62-97 -- alpha per cell
 * (N values); NULL returns to the uniform alpha of mmx_set_llg.  The stochastic
 * field and the adjoint need the uniform alpha. */
int mmx_set_alpha(mmx_ctx *ctx, const double *alpha);
/* llg_call_back(sim, dm_dt, spin, t): src/llg/llg.jl:45-58; m == NULL uses the current state */
int mmx_llg_rhs(mmx_ctx *ctx, double t, const double *m, double *dmdt);

/* create_driver("LLG", integrator): src/driver.jl:126-145.  step: fixed step of
 * RK4/Heun; tol: tolerance of DOPRI54/BS23.  Resets t, nsteps and nfevals.
 * kind MMX_SD selects create_driver("SD") (src/driver.jl:93-100) instead:
 * advance_step is then run_step(sim, ::SD) and relax uses tau/(gamma/2) as its
 * step size; mmx_get_state reports tau as `step`. */
int mmx_set_integrator(mmx_ctx *ctx, int kind, double step, double tol);
/* advance_step(sim) n times: src/integrator/rk.jl:23, heun.jl:28, adaptive_rk.jl:93 */
int mmx_advance_step(mmx_ctx *ctx, int64_t n);
/* run_until(sim, t_end): src/sim.jl:547-579 + integrate_to_time, adaptive_rk.jl:168-186 */
int mmx_run_until(mmx_ctx *ctx, double t_end);
/* relax(sim; max_steps, stopping_dmdt) with the LLG driver: src/sim.jl:392-490 */
int mmx_relax(mmx_ctx *ctx, int64_t max_steps, double stopping_dmdt, int64_t *steps, double *max_dmdt);
/* integrator.t, .step, .step_next, .nsteps, .nfevals: src/integrator/adaptive_rk.jl:37-55 */
int mmx_get_state(mmx_ctx *ctx, double *t, double *step, double *step_next, int64_t *nsteps,
                  int64_t *nfevals);

/* Replace the per-cell scalar parameter of a term (N values): A of a spatial
 * exchange, D of an interfacial DMI, Ku of an anisotropy, Kc of a cubic
 * anisotropy, T of a thermal term.  A uniform term becomes spatial.  The
 * reference changes these by rebuilding the term (add_anis(sim, Ku) again,
 * src/micro/add_field.jl:430-465); this keeps the demag set-up. */
int mmx_set_term_param(mmx_ctx *ctx, int term, const double *values);

/* Differentiation form.  The reference's only derivative of this path is the
 * Enzyme forward-mode Jacobian of gamma m x H_eff (ext/EnzymeExt.jl:6-59);
 * these entry points give reverse-mode gradients of a loss on the state after
 * an RK4 trajectory instead.
 * mmx_adjoint_record: advance nsteps RK4 steps as mmx_advance_step does,
 * keeping a checkpoint of the state every `every` steps (RK4 only). */
int mmx_adjoint_record(mmx_ctx *ctx, int64_t nsteps, int64_t every);
/* Pull the cotangent dL/dm(T) (3N AoS) of the recorded trajectory's final
 * state back to dL/dm(0) (grad_m0, 3N AoS, nullable), recomputing each segment
 * from its checkpoint; also forms the per-cell parameter gradients below. */
int mmx_adjoint_backward(mmx_ctx *ctx, const double *lambda_T, double *grad_m0);
/* dL/dp per cell (N values) of the last backward pass for the term's scalar
 * parameter field: Ku (anisotropy), Kc (cubic), D (interfacial DMI), A
 * (spatial exchange).  Bulk DMI (kernels.jl:306-422) gives 3N values, AoS
 * (dL/dDx, dL/dDy, dL/dDz per cell): for a spatial vector D(x) the gradient of
 * each cell's components; for a uniform D each cell's share of the gradient of
 * the uniform components (their sum over cells). */
int mmx_adjoint_param_grad(mmx_ctx *ctx, int term, double *grad);
/* dL/dMs per cell (N values) of the last backward pass: Ms enters the local and
 * stencil terms as 1/Ms and the demag term through M = Ms m (set_Ms,
 * src/sim.jl:129-143); zero at cells with Ms = 0. */
int mmx_adjoint_ms_grad(mmx_ctx *ctx, double *grad);
/* The Jacobian of one llg_call_back evaluation (src/llg/llg.jl:45-58) at the
 * state m (3N AoS; NULL: the current state) and time t, applied to v (JVP:
 * out = J v) or transposed to a (VJP: out = J^T a), 3N AoS each.  The JVP is
 * what ext/EnzymeExt.jl:32-59 (dynamic_matrix) builds column by column with
 * Enzyme's forward mode; the field's Jacobian covers exchange, DMI, uniaxial
 * and cubic anisotropy, static Zeeman and demag terms, uniform alpha.  The VJP,
 * like the adjoint, takes m = 0 at cells with Ms = 0 (as init_m0 leaves them). */
int mmx_llg_rhs_jvp(mmx_ctx *ctx, double t, const double *m, const double *v, double *out);
int mmx_llg_rhs_vjp(mmx_ctx *ctx, double t, const double *m, const double *a, double *out);

/* Batching form.  Many independent systems (thermal-ensemble members,
 * parameter sweeps) of any mesh sizes advanced together: every stage of a
 * step is one launch over all members without demag, plus, for each group of
 * members with demag that share a geometry (mesh, cell size, periodicity and
 * demag images, hence one tensor), one launch per demag pass over the group;
 * two steps are replayed from a HIP graph.  The reference runs such systems
 * one Sim at a time (one run_sim / advance_step per Sim, src/sim.jl:813-857);
 * each member keeps its own terms, parameters, state and counters, exactly as
 * if advanced alone.  Members: same precision and device, all RungeKutta or
 * all Heun with a step set, constant terms.  The members stay usable on their
 * own. */
typedef struct mmx_batch mmx_batch;
int mmx_batch_create(mmx_batch **out, mmx_ctx *const *members, int n);
int mmx_batch_destroy(mmx_batch *batch);
/* advance_step(sim) nsteps times for every member */
int mmx_batch_advance(mmx_batch *batch, int64_t nsteps);
/* the launch shape the set-up picked by timing (cells per thread and grid of
 * the stage launch, number of work items) and the candidates' times in ms */
int mmx_batch_info(mmx_batch *batch, int32_t *cells_per_thread, int32_t *grid, int32_t *work_items,
                   float *tune_ms, int32_t n_tune);
/* The whole launch shape the set-up picked (threads per block, cells per
 * thread, grid of the members without demag) and the number of geometry
 * groups of members with demag.  No counterpart in the reference (it has no
 * batched launch); for callers that report or pin the tuned shape. */
int mmx_batch_shape(mmx_batch *batch, int32_t *block_threads, int32_t *cells_per_thread, int32_t *grid,
                    int32_t *n_groups);

/* Nudged elastic band with the LLG driver: NEB(sim, init_images,
 * frames_between_images; spring_constant, clib_image), src/neb/neb.jl:84-153.
 * images: (n_free + 2) x 3N AoS -- the fixed left image, the n_free free
 * images (already interpolated, e.g. by slerp as init_images does) and the
 * fixed right image.  The field and energy of each image come from `sim` (its
 * terms, demag included); sim must outlive the band.  Defaults as the
 * reference sets them: BS23, tol 1e-5, alpha 0.2, gamma 2.21e5, precession
 * off.  climbing_image is 1-based among the free images, -1 for none. */
typedef struct mmx_neb mmx_neb;
int mmx_neb_create(mmx_neb **out, mmx_ctx *sim, const double *images, int n_free, double spring_constant,
                   int climbing_image);
int mmx_neb_destroy(mmx_neb *neb);
/* neb.driver.alpha / gamma / precession / tol */
int mmx_neb_set_driver(mmx_neb *neb, double alpha, double gamma, int precession, double tol);
/* advance_step(neb, integrator) n times (adaptive_rk.jl:93-148) */
int mmx_neb_advance(mmx_neb *neb, int64_t nsteps);
/* relax(neb; max_steps, stopping_dmdt), src/neb/neb.jl:231-296 (no file output) */
int mmx_neb_relax(mmx_neb *neb, int64_t max_steps, double stopping_dmdt, int64_t *steps, double *max_dmdt);
/* free images (n_free x 3N AoS), energies E_total of all n_free+2 images (J),
 * the n_free+1 geodesic distances, and {t, step, step_next, nsteps, nfevals};
 * each output nullable */
int mmx_neb_get(mmx_neb *neb, double *images, double *energies, double *distances, double *state);

/* diagnostics: the compact FP64 real-space tensor of a demag term (6N values,
 * xx,yy,zz,xy,xz,yz), the padded FFT geometry and per-pass kernel timers */
int mmx_demag_tensor(mmx_ctx *ctx, int term, double *out);
int mmx_demag_geometry(mmx_ctx *ctx, int32_t *Px, int32_t *Py, int32_t *Pz, int32_t *KS);
int mmx_synchronize(mmx_ctx *ctx);
#define MMX_PASS_P1 0    /* x real->half-spectrum */
#define MMX_PASS_P2 1    /* y forward */
#define MMX_PASS_P3 2    /* z forward, tensor multiply, z inverse */
#define MMX_PASS_P4 3    /* y inverse */
#define MMX_PASS_P5 4    /* x inverse + fused local field, LLG and stage update */
#define MMX_PASS_LOCAL 5 /* local field + LLG + stage update (no demag) */
#define MMX_NPASS 6
/* diagnostic (no counterpart in the reference): a streaming copy of `bytes` between two
 * fresh device buffers on `device`, timed with HIP events over `reps` launches of `grid`
 * blocks after one warm-up launch; the practical HBM ceiling the passes are compared with
 * (read + write = 2 * bytes per copy). */
int mmx_diag_copy(int device, int64_t bytes, int reps, int grid, double *ms_per_copy);
int mmx_profile(mmx_ctx *ctx, int enable);
int mmx_profile_read(mmx_ctx *ctx, double *ms /* MMX_NPASS */, int64_t *launches /* MMX_NPASS */);

#ifdef __cplusplus
}
#endif
#endif
