/*
 * mm_oracle.c -- CPU oracle (test infrastructure only; see mm_oracle.h).
 *
 * A plain-C restatement of MicroMagnetic.jl's CPU path for the finite-
 * difference effective field, the LLG right-hand side and the steppers.  Every
 * function cites the reference file:line whose behaviour it restates.  The
 * operation order of each formula follows the reference so that results agree
 * to rounding with the HIP library, which is checked against this file.
 */
#include "mm_oracle.h"

#include <complex.h>
#undef I /* the imaginary unit is spelled CI below; I names cell indices */
#define CI _Complex_I
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MU0 (4.0 * M_PI * 1e-7) /* src/const.jl:1 */
#define K_B 1.3806505e-23       /* src/const.jl:3 */

typedef double complex cplx;

enum {
    T_ZEEMAN = 1,
    T_EXCH,
    T_EXCH_S,
    T_DMI_BULK,
    T_DMI_BULK_S,
    T_DMI_INT,
    T_ANIS,
    T_ANIS_S,
    T_CUBIC,
    T_DEMAG,
    T_DEMAG_DIRECT,
    T_THERMAL,
    T_SOT,
    T_STT_SL,
    T_STT_ZL,
    T_ZEEMAN_T,
    T_HEX,
    T_EXCH_IL,
    T_DMI_IL,
    T_TWIN,
    T_SAHE,
    T_MEL_T,
    T_MEL_C,
    T_TORQUE_FN
};

typedef struct {
    int kind;
    double *field;  /* 3N */
    double *energy; /* N */
    double p[12];
    double *a, *b, *c; /* per-cell parameter arrays */
    double *ext[2];    /* two more (spin anomalous Hall torque: a2, a3) */
    /* demag */
    int Px, Py, Pz;
    double *ktensor[6]; /* real k-space tensors, Px*Py*Pz each */
    cplx *buf[3];
    double *rtensor; /* compact real-space tensor, 6*N (xx,yy,zz,xy,xz,yz) */
    /* thermal */
    uint64_t seed;
    int64_t eta_step;
    double *eta;
    /* torque amplitude ufun(t) (NULL: 1) */
    double (*ufun)(double, void *);
    void *uuser;
    /* time Zeeman factors (fx, fy, fz)(t) */
    void (*ft3)(double, double *, void *);
    /* add_torque: fn(t, m, y, user) writes the torque y for the state m (3N AoS) */
    void (*tfn)(double, const double *, double *, void *);
} term_t;

enum { I_RK4 = 0, I_HEUN = 1, I_DOPRI54 = 2, I_BS23 = 3 };

struct orc_sim {
    int nx, ny, nz, N;
    double dx, dy, dz, volume;
    int px, py, pz;
    int32_t *ngbs; /* 6*N, 1-based, -1 = none (src/micro/mesh.jl:71-85) */
    double *mu0_Ms;
    uint8_t *pins;
    double *spin, *prespin, *field;
    int nterms;
    term_t terms[32];
    /* LLG driver (src/driver.jl:148 defaults) */
    double alpha, gamma;
    int precession;
    /* integrator */
    int ikind;
    double t, step, step_next, tol;
    int64_t nsteps, nfevals;
    int succeed;
    double *k[7];
    double *y_current, *y_next, *y_temp;
    double t_eval; /* time of the evaluation in progress (torque ufun) */
    double *alphas; /* SpatialLLG per-cell damping, NULL: uniform alpha */
};

/* ------------------------------------------------------------------------ */
/* mesh: src/micro/mesh.jl:24-43 (index / indexpbc), :71-85 (ngbs)          */
/* ------------------------------------------------------------------------ */
static int64_t mesh_index(int64_t i, int64_t j, int64_t k, int nx, int ny, int nz)
{
    if (i < 1 || j < 1 || k < 1 || k > nz || j > ny || i > nx) return -1;
    return (k - 1) * (int64_t)nx * ny + (j - 1) * (int64_t)nx + i;
}

static int64_t mesh_indexpbc(int64_t i, int64_t j, int64_t k, int nx, int ny, int nz, int px,
                             int py, int pz)
{
    if (px && i < 1) i += nx;
    if (px && i > nx) i -= nx;
    if (py && j < 1) j += ny;
    if (py && j > ny) j -= ny;
    if (pz && k < 1) k += nz;
    if (pz && k > nz) k -= nz;
    return mesh_index(i, j, k, nx, ny, nz);
}

orc_sim *orc_create(int nx, int ny, int nz, double dx, double dy, double dz, int pbc)
{
    orc_sim *s = calloc(1, sizeof(orc_sim));
    s->nx = nx;
    s->ny = ny;
    s->nz = nz;
    s->N = nx * ny * nz;
    s->dx = dx;
    s->dy = dy;
    s->dz = dz;
    s->volume = dx * dy * dz; /* mesh.jl:86 */
    s->px = (pbc & 1) != 0;
    s->py = (pbc & 2) != 0;
    s->pz = (pbc & 4) != 0;
    int N = s->N;
    s->ngbs = malloc(sizeof(int32_t) * 6 * (size_t)N);
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                int64_t id = mesh_index(i, j, k, nx, ny, nz) - 1;
                int32_t *g = s->ngbs + 6 * id;
                g[0] = (int32_t)mesh_indexpbc(i - 1, j, k, nx, ny, nz, s->px, s->py, s->pz);
                g[1] = (int32_t)mesh_indexpbc(i + 1, j, k, nx, ny, nz, s->px, s->py, s->pz);
                g[2] = (int32_t)mesh_indexpbc(i, j - 1, k, nx, ny, nz, s->px, s->py, s->pz);
                g[3] = (int32_t)mesh_indexpbc(i, j + 1, k, nx, ny, nz, s->px, s->py, s->pz);
                g[4] = (int32_t)mesh_indexpbc(i, j, k - 1, nx, ny, nz, s->px, s->py, s->pz);
                g[5] = (int32_t)mesh_indexpbc(i, j, k + 1, nx, ny, nz, s->px, s->py, s->pz);
            }
    s->mu0_Ms = calloc(N, sizeof(double));
    s->pins = calloc(N, 1);
    s->spin = calloc(3 * (size_t)N, sizeof(double));
    s->prespin = calloc(3 * (size_t)N, sizeof(double));
    s->field = calloc(3 * (size_t)N, sizeof(double));
    s->alpha = 0.1;
    s->gamma = 2.21e5;
    s->precession = 1;
    s->ikind = I_DOPRI54;
    s->tol = 1e-6;
    s->step = 0;
    return s;
}

static void free_term(term_t *t)
{
    free(t->field);
    free(t->energy);
    free(t->a);
    free(t->b);
    free(t->c);
    free(t->ext[0]);
    free(t->ext[1]);
    for (int i = 0; i < 6; i++) free(t->ktensor[i]);
    for (int i = 0; i < 3; i++) free(t->buf[i]);
    free(t->rtensor);
    free(t->eta);
}

void orc_destroy(orc_sim *s)
{
    if (!s) return;
    for (int i = 0; i < s->nterms; i++) free_term(&s->terms[i]);
    for (int i = 0; i < 7; i++) free(s->k[i]);
    free(s->y_current);
    free(s->y_next);
    free(s->y_temp);
    free(s->ngbs);
    free(s->mu0_Ms);
    free(s->pins);
    free(s->spin);
    free(s->prespin);
    free(s->alphas);
    free(s->field);
    free(s);
}

int orc_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

void orc_neighbours(const orc_sim *s, int32_t *out)
{
    memcpy(out, s->ngbs, sizeof(int32_t) * 6 * (size_t)s->N);
}

/* set_Ms: src/sim.jl:129-143 (Ms .*= mu_0) */
void orc_set_Ms(orc_sim *s, const double *Ms)
{
    for (int i = 0; i < s->N; i++) s->mu0_Ms[i] = Ms[i] * MU0;
}

void orc_set_mu0Ms(orc_sim *s, const double *v) { memcpy(s->mu0_Ms, v, sizeof(double) * s->N); }

void orc_set_pins(orc_sim *s, const uint8_t *p) { memcpy(s->pins, p, s->N); }

/* normalise: src/helper.jl:255-271 */
static void normalise(double *a, int N)
{
#pragma omp parallel for schedule(static)
    for (int id = 0; id < N; id++) {
        double *v = a + 3 * (size_t)id;
        double m2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if (m2 > 0) {
            double len = 1 / sqrt(m2);
            v[0] *= len;
            v[1] *= len;
            v[2] *= len;
        }
    }
}

static void integrator_set_initial(orc_sim *s);

/* init_m0: src/sim.jl:278-309 */
int orc_init_m0(orc_sim *s, const double *m, int norm)
{
    size_t n3 = 3 * (size_t)s->N;
    memcpy(s->spin, m, n3 * sizeof(double));
    if (norm) normalise(s->spin, s->N);
    for (int i = 0; i < s->N; i++)
        if (s->mu0_Ms[i] == 0) s->spin[3 * i] = s->spin[3 * i + 1] = s->spin[3 * i + 2] = 0;
    for (size_t i = 0; i < n3; i++)
        if (isnan(s->spin[i])) return -1;
    memcpy(s->prespin, s->spin, n3 * sizeof(double));
    integrator_set_initial(s);
    return 0;
}

void orc_set_m(orc_sim *s, const double *m)
{
    memcpy(s->spin, m, 3 * (size_t)s->N * sizeof(double));
    integrator_set_initial(s);
}

void orc_get_m(const orc_sim *s, double *m) { memcpy(m, s->spin, 3 * (size_t)s->N * sizeof(double)); }

/* ------------------------------------------------------------------------ */
/* cross products: src/util.jl:6-16                                         */
/* ------------------------------------------------------------------------ */
static inline double cross_x(double x1, double x2, double x3, double y1, double y2, double y3)
{
    (void)x1;
    (void)y1;
    return -x3 * y2 + x2 * y3;
}
static inline double cross_y(double x1, double x2, double x3, double y1, double y2, double y3)
{
    (void)x2;
    (void)y2;
    return x3 * y1 - x1 * y3;
}
static inline double cross_z(double x1, double x2, double x3, double y1, double y2, double y3)
{
    (void)x3;
    (void)y3;
    return -x2 * y1 + x1 * y2;
}

/* ------------------------------------------------------------------------ */
/* terms                                                                    */
/* ------------------------------------------------------------------------ */
static term_t *new_term(orc_sim *s, int kind)
{
    term_t *t = &s->terms[s->nterms++];
    memset(t, 0, sizeof(*t));
    t->kind = kind;
    t->field = calloc(3 * (size_t)s->N, sizeof(double));
    t->energy = calloc(s->N, sizeof(double));
    return t;
}

static double *dup_array(const double *a, size_t n)
{
    double *r = malloc(n * sizeof(double));
    memcpy(r, a, n * sizeof(double));
    return r;
}

/* add_exch with a number / 3-tuple: src/micro/add_field.jl:233-247 */
int orc_add_exch(orc_sim *s, double Ax, double Ay, double Az)
{
    term_t *t = new_term(s, T_EXCH);
    t->p[0] = Ax;
    t->p[1] = Ay;
    t->p[2] = Az;
    return s->nterms - 1;
}

/* add_exch with an array / function: add_field.jl:248-253 */
int orc_add_exch_spatial(orc_sim *s, const double *A)
{
    term_t *t = new_term(s, T_EXCH_S);
    t->a = dup_array(A, s->N);
    return s->nterms - 1;
}

/* add_dmi type="bulk" with a number / numeric 3-tuple: add_field.jl:299-303 */
int orc_add_dmi_bulk(orc_sim *s, double Dx, double Dy, double Dz)
{
    term_t *t = new_term(s, T_DMI_BULK);
    t->p[0] = Dx;
    t->p[1] = Dy;
    t->p[2] = Dz;
    return s->nterms - 1;
}

/* add_dmi type="bulk" with arrays / functions: add_field.jl:304-325 */
int orc_add_dmi_bulk_spatial(orc_sim *s, const double *Dx, const double *Dy, const double *Dz)
{
    term_t *t = new_term(s, T_DMI_BULK_S);
    t->a = dup_array(Dx, s->N);
    t->b = dup_array(Dy, s->N);
    t->c = dup_array(Dz, s->N);
    return s->nterms - 1;
}

/* add_dmi type="interfacial": add_field.jl:327-335 (D always stored per cell) */
int orc_add_dmi_interfacial(orc_sim *s, const double *D)
{
    term_t *t = new_term(s, T_DMI_INT);
    t->a = dup_array(D, s->N);
    return s->nterms - 1;
}

/* add_anis with a tuple axis: add_field.jl:430-446 (axis normalised) */
int orc_add_anis(orc_sim *s, const double *Ku, double ax, double ay, double az)
{
    term_t *t = new_term(s, T_ANIS);
    t->a = dup_array(Ku, s->N);
    double lt = sqrt(ax * ax + ay * ay + az * az);
    t->p[0] = ax / lt;
    t->p[1] = ay / lt;
    t->p[2] = az / lt;
    return s->nterms - 1;
}

/* add_anis with a per-cell axis: add_field.jl:447-453 */
int orc_add_anis_spatial(orc_sim *s, const double *Ku, const double *axes)
{
    term_t *t = new_term(s, T_ANIS_S);
    t->a = dup_array(Ku, s->N);
    t->b = dup_array(axes, 3 * (size_t)s->N);
    normalise(t->b, s->N);
    return s->nterms - 1;
}

/* add_cubic_anis: add_field.jl:541-574.  axis3 = axis1 x axis2 of the inputs
 * as given (not normalised), as the reference does. */
int orc_add_cubic_anis(orc_sim *s, const double *Kc, const double a1[3], const double a2[3])
{
    double n1 = sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
    double n2 = sqrt(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
    double na1[3] = {a1[0] / n1, a1[1] / n1, a1[2] / n1};
    double na2[3] = {a2[0] / n2, a2[1] / n2, a2[2] / n2};
    if (fabs(na1[0] * na2[0] + na1[1] * na2[1] + na1[2] * na2[2]) > 1e-10) return -1;
    term_t *t = new_term(s, T_CUBIC);
    t->a = dup_array(Kc, s->N);
    t->p[0] = na1[0];
    t->p[1] = na1[1];
    t->p[2] = na1[2];
    t->p[3] = na2[0];
    t->p[4] = na2[1];
    t->p[5] = na2[2];
    t->p[6] = a1[1] * a2[2] - a1[2] * a2[1];
    t->p[7] = a1[2] * a2[0] - a1[0] * a2[2];
    t->p[8] = a1[0] * a2[1] - a1[1] * a2[0];
    return s->nterms - 1;
}

/* add_zeeman: add_field.jl:11-48 (static field array) */
int orc_add_zeeman(orc_sim *s, const double *H)
{
    term_t *t = new_term(s, T_ZEEMAN);
    memcpy(t->field, H, 3 * (size_t)s->N * sizeof(double));
    return s->nterms - 1;
}

int orc_num_terms(const orc_sim *s) { return s->nterms; }

/* uniform_exchange_kernel!: src/micro/kernels.jl:269-304 */
static void field_exch(orc_sim *s, term_t *t, const double *m)
{
    const double ax = 2 * t->p[0] / (s->dx * s->dx);
    const double ay = 2 * t->p[1] / (s->dy * s->dy);
    const double az = 2 * t->p[2] / (s->dz * s->dz);
    const double nabla[6] = {ax, ax, ay, ay, az, az};
    const double *Ms = s->mu0_Ms;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < s->N; I++) {
        size_t i = 3 * (size_t)I;
        double *h = t->field;
        if (Ms[I] == 0) {
            t->energy[I] = 0;
            h[i] = h[i + 1] = h[i + 2] = 0;
            continue;
        }
        double fx = 0, fy = 0, fz = 0;
        for (int j = 0; j < 6; j++) {
            int32_t id = s->ngbs[6 * (size_t)I + j];
            if (id > 0 && Ms[id - 1] > 0) {
                size_t k = 3 * (size_t)(id - 1);
                fx += nabla[j] * (m[k] - m[i]);
                fy += nabla[j] * (m[k + 1] - m[i + 1]);
                fz += nabla[j] * (m[k + 2] - m[i + 2]);
            }
        }
        double Ms_inv = 1.0 / Ms[I];
        t->energy[I] = -0.5 * (fx * m[i] + fy * m[i + 1] + fz * m[i + 2]) * s->volume;
        h[i] = fx * Ms_inv;
        h[i + 1] = fy * Ms_inv;
        h[i + 2] = fz * Ms_inv;
    }
}

/* exchange_kernel! (spatial A, harmonic mean): src/micro/kernels.jl:226-266 */
static void field_exch_spatial(orc_sim *s, term_t *t, const double *m)
{
    const double ax = 2.0 / (s->dx * s->dx);
    const double ay = 2.0 / (s->dy * s->dy);
    const double az = 2.0 / (s->dz * s->dz);
    const double nabla[6] = {ax, ax, ay, ay, az, az};
    const double *Ms = s->mu0_Ms, *A = t->a;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < s->N; I++) {
        size_t i = 3 * (size_t)I;
        double *h = t->field;
        double A_I = A[I];
        if (Ms[I] == 0 || A_I == 0) {
            t->energy[I] = 0;
            h[i] = h[i + 1] = h[i + 2] = 0;
            continue;
        }
        double fx = 0, fy = 0, fz = 0;
        for (int j = 0; j < 6; j++) {
            int32_t id = s->ngbs[6 * (size_t)I + j];
            if (id > 0 && Ms[id - 1] > 0) {
                double A_nb = A[id - 1];
                if (A_nb != 0) {
                    size_t k = 3 * (size_t)(id - 1);
                    double A_eff = 2 * A_I * A_nb / (A_I + A_nb);
                    double coeff = A_eff * nabla[j];
                    fx += coeff * (m[k] - m[i]);
                    fy += coeff * (m[k + 1] - m[i + 1]);
                    fz += coeff * (m[k + 2] - m[i + 2]);
                }
            }
        }
        double Ms_inv = 1.0 / Ms[I];
        t->energy[I] = -0.5 * (fx * m[i] + fy * m[i + 1] + fz * m[i + 2]) * s->volume;
        h[i] = fx * Ms_inv;
        h[i + 1] = fy * Ms_inv;
        h[i + 2] = fz * Ms_inv;
    }
}

/* bulkdmi_kernel!: src/micro/kernels.jl:306-344 */
static void field_dmi_bulk(orc_sim *s, term_t *t, const double *m)
{
    /* TimeBulkDMI (field.jl:145-163): D = time_D(t) * D before the kernel */
    const double f = t->ufun ? t->ufun(s->t_eval, t->uuser) : 1.0;
    const double Dx = f * t->p[0], Dy = f * t->p[1], Dz = f * t->p[2];
    const double Ds[6] = {Dx / s->dx, Dx / s->dx, Dy / s->dy, Dy / s->dy, Dz / s->dz, Dz / s->dz};
    const double ax[6] = {1, -1, 0, 0, 0, 0};
    const double ay[6] = {0, 0, 1, -1, 0, 0};
    const double az[6] = {0, 0, 0, 0, 1, -1};
    const double *Ms = s->mu0_Ms;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < s->N; I++) {
        size_t i = 3 * (size_t)I;
        double *h = t->field;
        if (Ms[I] == 0) {
            t->energy[I] = 0;
            h[i] = h[i + 1] = h[i + 2] = 0;
            continue;
        }
        double fx = 0, fy = 0, fz = 0;
        for (int j = 0; j < 6; j++) {
            int32_t id = s->ngbs[6 * (size_t)I + j];
            if (id > 0 && Ms[id - 1] > 0) {
                size_t k = 3 * (size_t)(id - 1);
                fx += Ds[j] * cross_x(ax[j], ay[j], az[j], m[k], m[k + 1], m[k + 2]);
                fy += Ds[j] * cross_y(ax[j], ay[j], az[j], m[k], m[k + 1], m[k + 2]);
                fz += Ds[j] * cross_z(ax[j], ay[j], az[j], m[k], m[k + 1], m[k + 2]);
            }
        }
        double Ms_inv = 1.0 / Ms[I];
        t->energy[I] = -0.5 * (fx * m[i] + fy * m[i + 1] + fz * m[i + 2]) * s->volume;
        h[i] = fx * Ms_inv;
        h[i + 1] = fy * Ms_inv;
        h[i + 2] = fz * Ms_inv;
    }
}

/* spatial_vector_bulkdmi_kernel!: src/micro/kernels.jl:347-422 */
static void field_dmi_bulk_spatial(orc_sim *s, term_t *t, const double *m)
{
    const double axes[6] = {1 / s->dx, -1 / s->dx, 1 / s->dy, -1 / s->dy, 1 / s->dz, -1 / s->dz};
    const double *Ms = s->mu0_Ms, *Dxs = t->a, *Dys = t->b, *Dzs = t->c;
    /* TimeSpatialBulkDMI (field.jl:180-193): the kernel gets time_D(t) .* D */
    const double f = t->ufun ? t->ufun(s->t_eval, t->uuser) : 1.0;
#define DSC(v) (t->ufun ? f * (v) : (v))
#pragma omp parallel for schedule(static)
    for (int I = 0; I < s->N; I++) {
        size_t i = 3 * (size_t)I;
        double *h = t->field;
        if (Ms[I] == 0) {
            t->energy[I] = 0;
            h[i] = h[i + 1] = h[i + 2] = 0;
            continue;
        }
        double Dx_I = DSC(Dxs[I]), Dy_I = DSC(Dys[I]), Dz_I = DSC(Dzs[I]);
        double fx = 0, fy = 0, fz = 0;
        const int32_t *g = s->ngbs + 6 * (size_t)I;
        for (int j = 0; j < 2; j++) {
            int32_t id = g[j];
            if (id > 0 && Ms[id - 1] > 0) {
                double Dx_n = DSC(Dxs[id - 1]);
                if (Dx_I != 0 && Dx_n != 0) {
                    size_t k = 3 * (size_t)(id - 1);
                    double D_eff = 2 * Dx_I * Dx_n / (Dx_I + Dx_n);
                    double a = axes[j];
                    fy += D_eff * (-a * m[k + 2]);
                    fz += D_eff * (a * m[k + 1]);
                }
            }
        }
        for (int j = 2; j < 4; j++) {
            int32_t id = g[j];
            if (id > 0 && Ms[id - 1] > 0) {
                double Dy_n = DSC(Dys[id - 1]);
                if (Dy_I != 0 && Dy_n != 0) {
                    size_t k = 3 * (size_t)(id - 1);
                    double D_eff = 2 * Dy_I * Dy_n / (Dy_I + Dy_n);
                    double a = axes[j];
                    fx += D_eff * (a * m[k + 2]);
                    fz += D_eff * (-a * m[k]);
                }
            }
        }
        for (int j = 4; j < 6; j++) {
            int32_t id = g[j];
            if (id > 0 && Ms[id - 1] > 0) {
                double Dz_n = DSC(Dzs[id - 1]);
                if (Dz_I != 0 && Dz_n != 0) {
                    size_t k = 3 * (size_t)(id - 1);
                    double D_eff = 2 * Dz_I * Dz_n / (Dz_I + Dz_n);
                    double a = axes[j];
                    fx += D_eff * (-a * m[k + 1]);
                    fy += D_eff * (a * m[k]);
                }
            }
        }
        double Ms_inv = 1.0 / Ms[I];
        t->energy[I] = -0.5 * (fx * m[i] + fy * m[i + 1] + fz * m[i + 2]) * s->volume;
        h[i] = fx * Ms_inv;
        h[i + 1] = fy * Ms_inv;
        h[i + 2] = fz * Ms_inv;
    }
}
#undef DSC

/* interfacial_dmi_kernel!: src/micro/kernels.jl:426-466 */
static void field_dmi_int(orc_sim *s, term_t *t, const double *m)
{
    const double Dd[4] = {1 / s->dx, 1 / s->dx, 1 / s->dy, 1 / s->dy};
    const double ax[4] = {0, 0, -1, 1};
    const double ay[4] = {1, -1, 0, 0};
    const double az[4] = {0, 0, 0, 0};
    const double *Ms = s->mu0_Ms, *Ds = t->a;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < s->N; I++) {
        size_t i = 3 * (size_t)I;
        double *h = t->field;
        double D_I = Ds[I];
        if (Ms[I] == 0 || D_I == 0) {
            t->energy[I] = 0;
            h[i] = h[i + 1] = h[i + 2] = 0;
            continue;
        }
        double fx = 0, fy = 0, fz = 0;
        for (int j = 0; j < 4; j++) {
            int32_t id = s->ngbs[6 * (size_t)I + j];
            if (id > 0 && Ms[id - 1] > 0) {
                double D_nb = Ds[id - 1];
                if (D_nb != 0) {
                    size_t k = 3 * (size_t)(id - 1);
                    double D_eff = 2 * D_I * D_nb / (D_I + D_nb);
                    double coeff = D_eff * Dd[j];
                    fx += coeff * cross_x(ax[j], ay[j], az[j], m[k], m[k + 1], m[k + 2]);
                    fy += coeff * cross_y(ax[j], ay[j], az[j], m[k], m[k + 1], m[k + 2]);
                    fz += coeff * cross_z(ax[j], ay[j], az[j], m[k], m[k + 1], m[k + 2]);
                }
            }
        }
        double Ms_inv = 1.0 / Ms[I];
        t->energy[I] = -0.5 * (fx * m[i] + fy * m[i + 1] + fz * m[i + 2]) * s->volume;
        h[i] = fx * Ms_inv;
        h[i + 1] = fy * Ms_inv;
        h[i + 2] = fz * Ms_inv;
    }
}

/* anisotropy_kernel! / spatial_anisotropy_kernel!: src/micro/kernels.jl:54-105 */
static void field_anis(orc_sim *s, term_t *t, const double *m, int spatial)
{
    const double *Ms = s->mu0_Ms, *Ku = t->a;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < s->N; I++) {
        size_t j = 3 * (size_t)I;
        double *h = t->field;
        double ax = t->p[0], ay = t->p[1], az = t->p[2];
        if (spatial) {
            ax = t->b[j];
            ay = t->b[j + 1];
            az = t->b[j + 2];
        }
        double Ms_local = Ms[I];
        if (Ms_local == 0.0) {
            t->energy[I] = 0;
            h[j] = h[j + 1] = h[j + 2] = 0;
        } else {
            double Ms_inv = 2.0 / Ms_local;
            double sa = m[j] * ax + m[j + 1] * ay + m[j + 2] * az;
            h[j] = Ku[I] * sa * ax * Ms_inv;
            h[j + 1] = Ku[I] * sa * ay * Ms_inv;
            h[j + 2] = Ku[I] * sa * az * Ms_inv;
            t->energy[I] = Ku[I] * (1.0 - sa * sa) * s->volume;
        }
    }
}

/* cubic_anisotropy_kernel!: src/micro/kernels.jl:110-137 */
static void field_cubic(orc_sim *s, term_t *t, const double *m)
{
    const double *Ms = s->mu0_Ms, *Kc = t->a;
    const double a1x = t->p[0], a1y = t->p[1], a1z = t->p[2];
    const double a2x = t->p[3], a2y = t->p[4], a2z = t->p[5];
    const double a3x = t->p[6], a3y = t->p[7], a3z = t->p[8];
#pragma omp parallel for schedule(static)
    for (int I = 0; I < s->N; I++) {
        size_t j = 3 * (size_t)I;
        double *h = t->field;
        double Ms_local = Ms[I];
        if (Ms_local == 0.0) {
            t->energy[I] = 0;
            h[j] = h[j + 1] = h[j + 2] = 0;
        } else {
            double Ms_inv = 4.0 * Kc[I] / Ms_local;
            double mxp = a1x * m[j] + a1y * m[j + 1] + a1z * m[j + 2];
            double myp = a2x * m[j] + a2y * m[j + 1] + a2z * m[j + 2];
            double mzp = a3x * m[j] + a3y * m[j + 1] + a3z * m[j + 2];
            double mxp3 = mxp * mxp * mxp;
            double myp3 = myp * myp * myp;
            double mzp3 = mzp * mzp * mzp;
            h[j] = Ms_inv * (mxp3 * a1x + myp3 * a2x + mzp3 * a3x);
            h[j + 1] = Ms_inv * (mxp3 * a1y + myp3 * a2y + mzp3 * a3y);
            h[j + 2] = Ms_inv * (mxp3 * a1z + myp3 * a2z + mzp3 * a3z);
            t->energy[I] = -Kc[I] * (mxp * mxp3 + myp * myp3 + mzp * mzp3) * s->volume;
        }
    }
}

/* zeeman_kernel!: src/micro/kernels.jl:14-20 (energy only; the field is static) */
static void energy_zeeman(orc_sim *s, term_t *t, const double *m)
{
#pragma omp parallel for schedule(static)
    for (int id = 0; id < s->N; id++) {
        size_t j = 3 * (size_t)id;
        double mh = m[j] * t->field[j] + m[j + 1] * t->field[j + 1] + m[j + 2] * t->field[j + 2];
        t->energy[id] = -s->volume * s->mu0_Ms[id] * mh;
    }
}

/* ------------------------------------------------------------------------ */
/* FFT: a recursive mixed-radix decimation-in-time transform (radix 4, 2, 3, */
/* 5 and a direct DFT for any other prime factor).  Forward uses            */
/* exp(-2 pi i jk/n) (sign = -1), like FFTW's and numpy's forward transform. */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n, sign;
    int nf;
    int fac[64];
    cplx *tw; /* exp(sign 2 pi i t / n), t < n */
} fft_plan;

static void fft_plan_init(fft_plan *p, int n, int sign)
{
    p->n = n;
    p->sign = sign;
    p->nf = 0;
    int r = n;
    while (r % 4 == 0) {
        p->fac[p->nf++] = 4;
        r /= 4;
    }
    while (r % 2 == 0) {
        p->fac[p->nf++] = 2;
        r /= 2;
    }
    for (int q = 3; r > 1;) {
        if (r % q == 0) {
            p->fac[p->nf++] = q;
            r /= q;
        } else {
            q += 2;
            if (q * q > r) q = r;
        }
    }
    p->tw = malloc(sizeof(cplx) * (n > 0 ? n : 1));
    for (int t = 0; t < n; t++) {
        double ang = 2.0 * M_PI * (double)t / (double)n;
        p->tw[t] = cos(ang) + sign * sin(ang) * CI;
    }
}

static void fft_plan_free(fft_plan *p) { free(p->tw); }

static void fft_rec(const fft_plan *pl, const cplx *in, cplx *out, int n, size_t istride,
                    const int *fac, int twstride)
{
    int p = fac[0], m = n / p;
    if (m == 1) {
        for (int q = 0; q < p; q++) out[q] = in[q * istride];
    } else {
        for (int q = 0; q < p; q++)
            fft_rec(pl, in + q * istride, out + q * m, m, istride * p, fac + 1, twstride * p);
    }
    /* butterflies: out[k + s m] = sum_q W_n^{qk} out[q m + k] W_p^{qs} */
    cplx tmp[64];
    cplx *t = tmp;
    cplx *big = NULL;
    if (p > 64) t = big = malloc(sizeof(cplx) * p);
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < p; q++) {
            cplx v = out[q * m + k];
            t[q] = (q == 0 || k == 0) ? v : v * pl->tw[(size_t)q * k * twstride];
        }
        if (p == 2) {
            out[k] = t[0] + t[1];
            out[k + m] = t[0] - t[1];
        } else if (p == 4) {
            cplx a0 = t[0] + t[2], a1 = t[0] - t[2];
            cplx a2 = t[1] + t[3], a3 = (t[1] - t[3]) * (pl->sign * CI);
            out[k] = a0 + a2;
            out[k + 2 * m] = a0 - a2;
            out[k + m] = a1 + a3;
            out[k + 3 * m] = a1 - a3;
        } else {
            /* direct DFT of length p with roots W_p = W_n^{m} */
            for (int s2 = 0; s2 < p; s2++) {
                cplx acc = t[0];
                for (int q = 1; q < p; q++) {
                    size_t e = ((size_t)q * s2) % p;
                    acc += t[q] * pl->tw[e * m * twstride];
                }
                out[k + s2 * m] = acc;
            }
        }
    }
    free(big);
}

static void fft_exec(const fft_plan *pl, const cplx *in, size_t istride, cplx *out)
{
    if (pl->n == 1) {
        out[0] = in[0];
        return;
    }
    fft_rec(pl, in, out, pl->n, istride, pl->fac, 1);
}

void orc_fft(double *data, int n, int sign)
{
    fft_plan pl;
    fft_plan_init(&pl, n, sign);
    cplx *in = malloc(sizeof(cplx) * n), *out = malloc(sizeof(cplx) * n);
    for (int i = 0; i < n; i++) in[i] = data[2 * i] + data[2 * i + 1] * CI;
    fft_exec(&pl, in, 1, out);
    for (int i = 0; i < n; i++) {
        data[2 * i] = creal(out[i]);
        data[2 * i + 1] = cimag(out[i]);
    }
    free(in);
    free(out);
    fft_plan_free(&pl);
}

/* unnormalised 3-D transform of a (Px,Py,Pz) complex array, x fastest */
static void fft3d(cplx *a, int Px, int Py, int Pz, int sign)
{
    fft_plan px, py, pz;
    fft_plan_init(&px, Px, sign);
    fft_plan_init(&py, Py, sign);
    fft_plan_init(&pz, Pz, sign);
    int lmax = Px > Py ? Px : Py;
    lmax = lmax > Pz ? lmax : Pz;
#pragma omp parallel
    {
        cplx *tmp = malloc(sizeof(cplx) * lmax);
#pragma omp for schedule(static) collapse(2)
        for (int z = 0; z < Pz; z++)
            for (int y = 0; y < Py; y++) {
                cplx *row = a + ((size_t)z * Py + y) * Px;
                fft_exec(&px, row, 1, tmp);
                memcpy(row, tmp, sizeof(cplx) * Px);
            }
#pragma omp for schedule(static) collapse(2)
        for (int z = 0; z < Pz; z++)
            for (int x = 0; x < Px; x++) {
                cplx *col = a + (size_t)z * Py * Px + x;
                fft_exec(&py, col, Px, tmp);
                for (int y = 0; y < Py; y++) col[(size_t)y * Px] = tmp[y];
            }
#pragma omp for schedule(static) collapse(2)
        for (int y = 0; y < Py; y++)
            for (int x = 0; x < Px; x++) {
                cplx *col = a + (size_t)y * Px + x;
                fft_exec(&pz, col, (size_t)Px * Py, tmp);
                for (int z = 0; z < Pz; z++) col[(size_t)z * Px * Py] = tmp[z];
            }
        free(tmp);
    }
    fft_plan_free(&px);
    fft_plan_free(&py);
    fft_plan_free(&pz);
}

/* ------------------------------------------------------------------------ */
/* Demag tensor: src/micro/demag.jl:133-311, src/atomistic/demag.jl:106-140 */
/* ------------------------------------------------------------------------ */
static double newell_f(double x, double y, double z) /* demag.jl:133-155 */
{
    double x2 = x * x, y2 = y * y, z2 = z * z;
    double R = sqrt(x2 + y2 + z2);
    if (R == 0.0) return 0.0;
    double f = 1.0 / 6 * (2 * x2 - y2 - z2) * R;
    if (x2 > 0) f -= x * y * z * atan(y * z / (x * R));
    if (x2 + z2 > 0) f += 0.5 * y * (z2 - x2) * asinh(y / (sqrt(x2 + z2)));
    if (x2 + y2 > 0) f += 0.5 * z * (y2 - x2) * asinh(z / (sqrt(x2 + y2)));
    return f;
}

static double newell_g(double x, double y, double z) /* demag.jl:157-192 */
{
    double x2 = x * x, y2 = y * y, z2 = z * z;
    double R = sqrt(x2 + y2 + z2);
    if (R == 0.0) return 0.0;
    double g = -1.0 / 3 * x * y * R;
    if (z2 > 0) g -= 1.0 / 6 * z2 * z * atan(x * y / (z * R));
    if (y2 > 0) g -= 0.5 * y2 * z * atan(x * z / (y * R));
    if (x2 > 0) g -= 0.5 * x2 * z * atan(y * z / (x * R));
    if (x2 + y2 > 0) g += x * y * z * asinh(z / (sqrt(x2 + y2)));
    if (y2 + z2 > 0) g += 1.0 / 6 * y * (3 * z2 - y2) * asinh(x / (sqrt(y2 + z2)));
    if (x2 + z2 > 0) g += 1.0 / 6 * x * (3 * z2 - x2) * asinh(y / (sqrt(x2 + z2)));
    return g;
}

static double dipolar_xx(double x, double y, double z) /* atomistic/demag.jl:106-113 */
{
    double R = x * x + y * y + z * z;
    if (R == 0) return 0.0;
    return -(2 * x * x - y * y - z * z) / (R * R * sqrt(R));
}

static double dipolar_xy(double x, double y, double z) /* atomistic/demag.jl:115-122 */
{
    double R = x * x + y * y + z * z;
    if (R == 0) return 0.0;
    return -3 * x * y / (R * R * sqrt(R));
}

static double fmax3(double a, double b, double c)
{
    double m = a > b ? a : b;
    return m > c ? m : c;
}

typedef double (*newell_fn)(double, double, double);

/* the 27-point Newell combination of demag.jl:203-234 (f) and :254-285 (g) */
static double newell_27(newell_fn F, double x, double y, double z, double dx, double dy, double dz)
{
    double tensor = 8.0 * F(x, y, z);
    tensor -= 4.0 * F(x + dx, y, z);
    tensor -= 4.0 * F(x - dx, y, z);
    tensor -= 4.0 * F(x, y - dy, z);
    tensor -= 4.0 * F(x, y + dy, z);
    tensor -= 4.0 * F(x, y, z - dz);
    tensor -= 4.0 * F(x, y, z + dz);
    tensor += 2.0 * F(x + dx, y + dy, z);
    tensor += 2.0 * F(x + dx, y - dy, z);
    tensor += 2.0 * F(x - dx, y - dy, z);
    tensor += 2.0 * F(x - dx, y + dy, z);
    tensor += 2.0 * F(x + dx, y, z + dz);
    tensor += 2.0 * F(x + dx, y, z - dz);
    tensor += 2.0 * F(x - dx, y, z + dz);
    tensor += 2.0 * F(x - dx, y, z - dz);
    tensor += 2.0 * F(x, y - dy, z - dz);
    tensor += 2.0 * F(x, y - dy, z + dz);
    tensor += 2.0 * F(x, y + dy, z + dz);
    tensor += 2.0 * F(x, y + dy, z - dz);
    tensor -= F(x + dx, y + dy, z + dz);
    tensor -= F(x + dx, y + dy, z - dz);
    tensor -= F(x + dx, y - dy, z + dz);
    tensor -= F(x + dx, y - dy, z - dz);
    tensor -= F(x - dx, y + dy, z + dz);
    tensor -= F(x - dx, y + dy, z - dz);
    tensor -= F(x - dx, y - dy, z + dz);
    tensor -= F(x - dx, y - dy, z - dz);
    return tensor / (4.0 * M_PI * dx * dy * dz);
}

static double tensor_xx(double x, double y, double z, double dx, double dy, double dz)
{ /* demag.jl:196-235 */
    double R = sqrt(x * x + y * y + z * z) / fmax3(dx, dy, dz);
    if (R > 60) return dipolar_xx(x, y, z) * dx * dy * dz / (4 * M_PI);
    return newell_27(newell_f, x, y, z, dx, dy, dz);
}

static double tensor_xy(double x, double y, double z, double dx, double dy, double dz)
{ /* demag.jl:247-286 */
    double R = sqrt(x * x + y * y + z * z) / fmax3(dx, dy, dz);
    if (R > 60) return dipolar_xy(x, y, z) * dx * dy * dz / (4 * M_PI);
    return newell_27(newell_g, x, y, z, dx, dy, dz);
}

/* component c of the tensor at displacement (x,y,z): demag.jl:237-296 */
static double tensor_comp(int c, double x, double y, double z, double dx, double dy, double dz)
{
    switch (c) {
    case 0: return tensor_xx(x, y, z, dx, dy, dz);
    case 1: return tensor_xx(y, x, z, dy, dx, dz);
    case 2: return tensor_xx(z, y, x, dz, dy, dx);
    case 3: return tensor_xy(x, y, z, dx, dy, dz);
    case 4: return tensor_xy(x, z, y, dx, dz, dy);
    default: return tensor_xy(y, z, x, dy, dz, dx);
    }
}

/* tensors_kernel_*!: demag.jl:298-386 -- image sum over (2Nx+1)(2Ny+1)(2Nz+1) */
static void compute_rtensor(orc_sim *s, term_t *t, int Nx, int Ny, int Nz)
{
    double ms = fmax3(s->dx, s->dy, s->dz); /* demag.jl:32-35 */
    double dx = s->dx / ms, dy = s->dy / ms, dz = s->dz / ms;
    int nx = s->nx, ny = s->ny, nz = s->nz;
    t->rtensor = malloc(sizeof(double) * 6 * (size_t)s->N);
    for (int c = 0; c < 6; c++) {
        double *out = t->rtensor + (size_t)c * s->N;
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    double sum = 0.0;
                    for (int p = -Nx; p <= Nx; p++)
                        for (int q = -Ny; q <= Ny; q++)
                            for (int r = -Nz; r <= Nz; r++) {
                                double x = (i + (double)p * nx) * dx;
                                double y = (j + (double)q * ny) * dy;
                                double z = (k + (double)r * nz) * dz;
                                sum += tensor_comp(c, x, y, z, dx, dy, dz);
                            }
                    out[((size_t)k * ny + j) * nx + i] = sum;
                }
    }
}

/* padded length: demag.jl:41-44 */
static int pad_len(int n) { return n > 3 ? 2 * n : 2 * n - 1; }

/* init_demag: demag.jl:30-100 (fill_tensors :394-417, real(rfft) of each) */
int orc_add_demag(orc_sim *s, int Nx, int Ny, int Nz)
{
    term_t *t = new_term(s, T_DEMAG);
    compute_rtensor(s, t, Nx, Ny, Nz);
    int nx = s->nx, ny = s->ny, nz = s->nz;
    int Px = pad_len(nx), Py = pad_len(ny), Pz = pad_len(nz);
    t->Px = Px;
    t->Py = Py;
    t->Pz = Pz;
    size_t P = (size_t)Px * Py * Pz;
    cplx *work = malloc(sizeof(cplx) * P);
    /* odd components: xy odd in x,y; xz in x,z; yz in y,z (demag.jl:69-80) */
    const int odd[6][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    for (int c = 0; c < 6; c++) {
        const double *tc = t->rtensor + (size_t)c * s->N;
#pragma omp parallel for schedule(static) collapse(2)
        for (int k = 1; k <= Pz; k++)
            for (int j = 1; j <= Py; j++)
                for (int i = 1; i <= Px; i++) {
                    size_t o = ((size_t)(k - 1) * Py + (j - 1)) * Px + (i - 1);
                    if ((Px % 2 == 0 && i == nx + 1) || (Py % 2 == 0 && j == ny + 1) ||
                        (Pz % 2 == 0 && k == nz + 1)) {
                        work[o] = 0;
                        continue;
                    }
                    int x = (i <= nx) ? i : Px - i + 2;
                    int y = (j <= ny) ? j : Py - j + 2;
                    int z = (k <= nz) ? k : Pz - k + 2;
                    int sx = (odd[c][0] && i > nx) ? -1 : 1;
                    int sy = (odd[c][1] && j > ny) ? -1 : 1;
                    int sz = (odd[c][2] && k > nz) ? -1 : 1;
                    work[o] = sx * sy * sz * tc[((size_t)(z - 1) * ny + (y - 1)) * nx + (x - 1)];
                }
        fft3d(work, Px, Py, Pz, -1);
        t->ktensor[c] = malloc(sizeof(double) * P);
        for (size_t i = 0; i < P; i++) t->ktensor[c][i] = creal(work[i]);
    }
    free(work);
    for (int c = 0; c < 3; c++) t->buf[c] = malloc(sizeof(cplx) * P);
    return s->nterms - 1;
}

int orc_demag_tensor(const orc_sim *s, int term, double *out)
{
    if (term < 0 || term >= s->nterms) return -1;
    const term_t *t = &s->terms[term];
    if (!t->rtensor) return -1;
    memcpy(out, t->rtensor, sizeof(double) * 6 * (size_t)s->N);
    return 0;
}

/* effective_field(::Demag): demag.jl:102-130, distribute_m :419-427,
 * add_tensor_M :463-481, collect_h :435-449.  irfft is normalised by P. */
static void field_demag(orc_sim *s, term_t *t, const double *m)
{
    int nx = s->nx, ny = s->ny, nz = s->nz, Px = t->Px, Py = t->Py, Pz = t->Pz;
    size_t P = (size_t)Px * Py * Pz;
    for (int c = 0; c < 3; c++) {
        cplx *b = t->buf[c];
        memset(b, 0, sizeof(cplx) * P);
#pragma omp parallel for schedule(static) collapse(2)
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) {
                    size_t I = ((size_t)k * ny + j) * nx + i;
                    b[((size_t)k * Py + j) * Px + i] = m[3 * I + c] * s->mu0_Ms[I];
                }
        fft3d(b, Px, Py, Pz, -1);
    }
    cplx *Mx = t->buf[0], *My = t->buf[1], *Mz = t->buf[2];
    double **T = t->ktensor;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < P; i++) {
        double xx = T[0][i], yy = T[1][i], zz = T[2][i], xy = T[3][i], xz = T[4][i], yz = T[5][i];
        cplx Hx = xx * Mx[i] + xy * My[i] + xz * Mz[i];
        cplx Hy = xy * Mx[i] + yy * My[i] + yz * Mz[i];
        cplx Hz = xz * Mx[i] + yz * My[i] + zz * Mz[i];
        Mx[i] = Hx;
        My[i] = Hy;
        Mz[i] = Hz;
    }
    for (int c = 0; c < 3; c++) fft3d(t->buf[c], Px, Py, Pz, +1);
    const double factor = -1.0 / MU0;
    const double invP = 1.0 / (double)P;
#pragma omp parallel for schedule(static) collapse(2)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t I = ((size_t)k * ny + j) * nx + i;
                size_t o = ((size_t)k * Py + j) * Px + i;
                size_t p = 3 * I;
                double *h = t->field;
                h[p] = factor * (creal(Mx[o]) * invP);
                h[p + 1] = factor * (creal(My[o]) * invP);
                h[p + 2] = factor * (creal(Mz[o]) * invP);
                double mh = m[p] * h[p] + m[p + 1] * h[p + 1] + m[p + 2] * h[p + 2];
                t->energy[I] = -0.5 * s->mu0_Ms[I] * s->volume * mh;
            }
}

/* init_direct_demag / direct evaluation: src/micro/demag_direct.jl:13-101.
 * H_i = -(1/mu0) sum_j N(r_i - r_j) mu0Ms_j m_j with the same tensor. */
int orc_add_demag_direct(orc_sim *s, int Nx, int Ny, int Nz)
{
    term_t *t = new_term(s, T_DEMAG_DIRECT);
    compute_rtensor(s, t, Nx, Ny, Nz);
    return s->nterms - 1;
}

static void field_demag_direct(orc_sim *s, term_t *t, const double *m)
{
    int nx = s->nx, ny = s->ny, N = s->N;
    const double *R = t->rtensor;
    const double factor = -1.0 / MU0;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < N; I++) {
        int i = I % nx, j = (I / nx) % ny, k = I / (nx * ny);
        double hx = 0, hy = 0, hz = 0;
        for (int J = 0; J < N; J++) {
            int i2 = J % nx, j2 = (J / nx) % ny, k2 = J / (nx * ny);
            int di = i - i2, dj = j - j2, dk = k - k2;
            double sx = di < 0 ? -1 : 1, sy = dj < 0 ? -1 : 1, sz = dk < 0 ? -1 : 1;
            size_t o = ((size_t)abs(dk) * ny + abs(dj)) * nx + abs(di);
            double xx = R[o], yy = R[(size_t)N + o], zz = R[2 * (size_t)N + o];
            double xy = sx * sy * R[3 * (size_t)N + o];
            double xz = sx * sz * R[4 * (size_t)N + o];
            double yz = sy * sz * R[5 * (size_t)N + o];
            double Mx = s->mu0_Ms[J] * m[3 * J], My = s->mu0_Ms[J] * m[3 * J + 1],
                   Mz = s->mu0_Ms[J] * m[3 * J + 2];
            hx += xx * Mx + xy * My + xz * Mz;
            hy += xy * Mx + yy * My + yz * Mz;
            hz += xz * Mx + yz * My + zz * Mz;
        }
        double *h = t->field + 3 * (size_t)I;
        h[0] = factor * hx;
        h[1] = factor * hy;
        h[2] = factor * hz;
        double mh = m[3 * I] * h[0] + m[3 * I + 1] * h[1] + m[3 * I + 2] * h[2];
        t->energy[I] = -0.5 * s->mu0_Ms[I] * s->volume * mh;
    }
}

/* ------------------------------------------------------------------------ */
/* Thermal field: StochasticField (src/micro/field.jl:244-280,              */
/* kernels.jl:469-494).  The reference draws eta with the backend randn!,   */
/* which no fixture pins; here eta comes from a counter-based Philox4x32-10 */
/* generator keyed by (seed, integrator step, component index), so the GPU */
/* library can draw the identical stream.                                   */
/* ------------------------------------------------------------------------ */
static void philox4x32_10(uint32_t ctr[4], const uint32_t key_in[2])
{
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * ctr[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * ctr[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ ctr[1] ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ ctr[3] ^ k1;
        uint32_t n3 = (uint32_t)p0;
        ctr[0] = n0;
        ctr[1] = n1;
        ctr[2] = n2;
        ctr[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

/* normal variate number `idx` of the stream (seed, step): Box-Muller over
 * two 53-bit uniforms; even idx takes the cosine branch, odd the sine. */
static double philox_normal(uint64_t seed, uint64_t step, uint64_t idx)
{
    uint64_t pair = idx >> 1;
    uint32_t ctr[4] = {(uint32_t)pair, (uint32_t)(pair >> 32), (uint32_t)step,
                       (uint32_t)(step >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    philox4x32_10(ctr, key);
    double u1 = ((double)(((uint64_t)(ctr[0] >> 6) << 27) | (ctr[1] >> 5)) + 0.5) *
                (1.0 / 9007199254740992.0);
    double u2 = ((double)(((uint64_t)(ctr[2] >> 6) << 27) | (ctr[3] >> 5)) + 0.5) *
                (1.0 / 9007199254740992.0);
    double r = sqrt(-2.0 * log(u1));
    return (idx & 1) ? r * sin(2.0 * M_PI * u2) : r * cos(2.0 * M_PI * u2);
}

/* the raw block, for the known-answer checks of the round function */
void orc_philox4x32_10(const uint32_t *ctr_in, const uint32_t *key, uint32_t *out)
{
    uint32_t c[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
    philox4x32_10(c, key);
    for (int q = 0; q < 4; q++) out[q] = c[q];
}

void orc_philox_normal(uint64_t seed, uint64_t step, uint64_t n, double *out)
{
    for (uint64_t i = 0; i < n; i++) out[i] = philox_normal(seed, step, i);
}

/* add_thermal_noise: src/micro/add_field.jl:734-771 (T0 = offset temperature) */
int orc_add_thermal(orc_sim *s, const double *Temp, double T0, uint64_t seed)
{
    term_t *t = new_term(s, T_THERMAL);
    t->a = dup_array(Temp, s->N);
    t->p[0] = T0;
    t->seed = seed;
    t->eta_step = -1;
    t->eta = calloc(3 * (size_t)s->N, sizeof(double));
    return s->nterms - 1;
}

static void field_thermal(orc_sim *s, term_t *t, const double *m)
{
    /* eta refreshed only when the integrator step count advances (field.jl:250-253) */
    if (s->nsteps > t->eta_step) {
        orc_philox_normal(t->seed, (uint64_t)s->nsteps, 3 * (uint64_t)s->N, t->eta);
        t->eta_step = s->nsteps;
    }
    double dt = s->step;
    double factor = 2 * s->alpha * K_B / (s->volume * s->gamma * dt); /* field.jl:261 */
    const double *Ms = s->mu0_Ms;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < s->N; I++) {
        size_t j = 3 * (size_t)I;
        double *h = t->field;
        double T_local = t->a[I] + t->p[0];
        if (Ms[I] > 0) {
            double scale = sqrt(factor * T_local / Ms[I]);
            h[j] = t->eta[j] * scale;
            h[j + 1] = t->eta[j + 1] * scale;
            h[j + 2] = t->eta[j + 2] * scale;
            t->energy[I] = -Ms[I] * s->volume * (m[j] * h[j] + m[j + 1] * h[j + 1] + m[j + 2] * h[j + 2]);
        } else {
            t->energy[I] = 0;
            h[j] = h[j + 1] = h[j + 2] = 0;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* spin torques as effective fields (no energy): src/micro/field.jl:317-356 */
/* ------------------------------------------------------------------------ */
/* cross_x/y/z: src/util.jl:6-16 */
static inline double crx(double x1, double x2, double x3, double y1, double y2, double y3) { return -x3 * y2 + x2 * y3; }
static inline double cry(double x1, double x2, double x3, double y1, double y2, double y3) { return x3 * y1 - x1 * y3; }
static inline double crz(double x1, double x2, double x3, double y1, double y2, double y3) { return -x2 * y1 + x1 * y2; }

/* add_sot: src/micro/add_field.jl:951-966 (aj per cell, p as given, not normalised) */
int orc_add_sot(orc_sim *s, const double *aj, double bj, double px, double py, double pz)
{
    term_t *t = new_term(s, T_SOT);
    t->a = dup_array(aj, s->N);
    t->p[0] = px;
    t->p[1] = py;
    t->p[2] = pz;
    t->p[4] = bj;
    return s->nterms - 1;
}

/* df_torque_kernel!: src/micro/kernels.jl:619-633 -- H = (1/gamma)(aj m x p + bj p) */
static void field_sot(orc_sim *s, term_t *t, const double *m)
{
    const double gamma = s->gamma, px = t->p[0], py = t->p[1], pz = t->p[2];
#pragma omp parallel for schedule(static)
    for (int I = 0; I < s->N; I++) {
        size_t j = 3 * (size_t)I;
        double *h = t->field;
        double a = t->a[I] / gamma;
        double b = t->p[4] / gamma;
        double mx = m[j], my = m[j + 1], mz = m[j + 2];
        h[j] = a * crx(mx, my, mz, px, py, pz) + b * px;
        h[j + 1] = a * cry(mx, my, mz, px, py, pz) + b * py;
        h[j + 2] = a * crz(mx, my, mz, px, py, pz) + b * pz;
        t->energy[I] = 0;
    }
}

/* add_stt(model=:slonczewski): src/micro/add_field.jl:1195-1227; beta as computed
 * there (h_bar/(mu_0*Ms*c_e*tf)) is an argument */
int orc_add_stt_slonczewski(orc_sim *s, const double *J, double beta, double Lambda, double xi, double P,
                            double px, double py, double pz)
{
    term_t *t = new_term(s, T_STT_SL);
    t->a = dup_array(J, s->N);
    t->p[0] = px;
    t->p[1] = py;
    t->p[2] = pz;
    t->p[4] = Lambda * Lambda; /* lambda_sq = T(torque.Lambda^2), field.jl:349 */
    t->p[5] = P;
    t->p[6] = xi;
    t->p[7] = beta;
    return s->nterms - 1;
}

/* slonczewski_torque_kernel!: src/micro/kernels.jl:657-674; ft = ufun(t)*beta (field.jl:350) */
static void field_stt_sl(orc_sim *s, term_t *t, const double *m)
{
    const double px = t->p[0], py = t->p[1], pz = t->p[2];
    const double lambda_sq = t->p[4], P = t->p[5], xi = t->p[6];
    const double ft = (t->ufun ? t->ufun(s->t_eval, t->uuser) : 1.0) * t->p[7];
#pragma omp parallel for schedule(static)
    for (int I = 0; I < s->N; I++) {
        size_t j = 3 * (size_t)I;
        double *h = t->field;
        double mx = m[j], my = m[j + 1], mz = m[j + 2];
        double mp = mx * px + my * py + mz * pz;
        double epsilon = P * lambda_sq / (lambda_sq + 1 + (lambda_sq - 1) * mp);
        double a = t->a[I] * ft;
        h[j] = a * (epsilon * crx(mx, my, mz, px, py, pz) + xi * px);
        h[j + 1] = a * (epsilon * cry(mx, my, mz, px, py, pz) + xi * py);
        h[j + 2] = a * (epsilon * crz(mx, my, mz, px, py, pz) + xi * pz);
        t->energy[I] = 0;
    }
}

/* add_stt(model=:zhang_li): src/micro/add_field.jl:1150-1192.  bJ = J .* b, and for
 * the (P, Ms) form (xi_factor = 1) each cell's vector times 1/(1 + xi^2) */
int orc_add_stt_zhangli(orc_sim *s, const double *J /* 3N */, double b, const double *xi, int xi_factor)
{
    term_t *t = new_term(s, T_STT_ZL);
    t->a = dup_array(xi, s->N);
    t->b = malloc(3 * (size_t)s->N * sizeof(double));
    for (int I = 0; I < s->N; I++) {
        double f = 1 / (1 + xi[I] * xi[I]);
        for (int q = 0; q < 3; q++) {
            double v = J[3 * (size_t)I + q] * b;
            t->b[3 * (size_t)I + q] = xi_factor ? v * f : v;
        }
    }
    return s->nterms - 1;
}

/* zhangli_torque_kernel!: src/micro/kernels.jl:680-739; ut = ufun(t)/gamma (field.jl:334).
 * Mesh neighbours (ngbs) only -- no Ms test; a missing side falls back to the cell
 * itself with a one-sided 1/d factor. */
static void field_stt_zl(orc_sim *s, term_t *t, const double *m)
{
    const double ut = (t->ufun ? t->ufun(s->t_eval, t->uuser) : 1.0) / s->gamma;
    const double d[3] = {s->dx, s->dy, s->dz};
#pragma omp parallel for schedule(static)
    for (int I = 0; I < s->N; I++) {
        size_t j = 3 * (size_t)I;
        double *h = t->field;
        double fx = 0, fy = 0, fz = 0;
        for (int q = 0; q < 3; q++) {
            int64_t i1 = s->ngbs[6 * (size_t)I + 2 * q], i2 = s->ngbs[6 * (size_t)I + 2 * q + 1];
            double factor = (i1 > 0 && i2 > 0) ? 1 / (2 * d[q]) : 1 / d[q];
            if (i1 < 0) i1 = I + 1;
            if (i2 < 0) i2 = I + 1;
            size_t j1 = 3 * (size_t)(i1 - 1), j2 = 3 * (size_t)(i2 - 1);
            double u = t->b[j + q] * factor;
            fx += u * (m[j2] - m[j1]);
            fy += u * (m[j2 + 1] - m[j1 + 1]);
            fz += u * (m[j2 + 2] - m[j1 + 2]);
        }
        fx = ut * fx;
        fy = ut * fy;
        fz = ut * fz;
        double mx = m[j], my = m[j + 1], mz = m[j + 2];
        h[j] = crx(mx, my, mz, fx, fy, fz) + t->a[I] * fx;
        h[j + 1] = cry(mx, my, mz, fx, fy, fz) + t->a[I] * fy;
        h[j + 2] = crz(mx, my, mz, fx, fy, fz) + t->a[I] * fz;
        t->energy[I] = 0;
    }
}

/* add_hex_anis(sim; K1, K2, K3): src/micro/add_field.jl:590-606 */
int orc_add_hex_anis(orc_sim *s, double K1, double K2, double K3)
{
    term_t *t = new_term(s, T_HEX);
    t->p[0] = K1;
    t->p[1] = K2;
    t->p[2] = K3;
    return s->nterms - 1;
}

/* hexagonal_anisotropy_kernel!: src/micro/kernels.jl:142-167.  The integer
 * powers are products (mx^5 = mx^4 * mx); the reference's Float64 power
 * differs from them in the last bits only. */
static void field_hex(orc_sim *s, term_t *t, const double *m)
{
    const double K1 = t->p[0], K2 = t->p[1], K3 = t->p[2];
    const double *Ms = s->mu0_Ms;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < s->N; I++) {
        size_t j = 3 * (size_t)I;
        double *h = t->field;
        if (Ms[I] == 0) {
            t->energy[I] = 0;
            h[j] = h[j + 1] = h[j + 2] = 0;
            continue;
        }
        const double Ms_inv = 1.0 / Ms[I];
        const double mx = m[j], my = m[j + 1], mz = m[j + 2];
        const double x2 = mx * mx, y2 = my * my, x4 = x2 * x2, y4 = y2 * y2;
        h[j] = -6 * K3 * Ms_inv * (x4 * mx - 10 * (x2 * mx) * y2 + 5 * mx * y4);
        h[j + 1] = -6 * K3 * Ms_inv * (-5 * x4 * my + 10 * x2 * (y2 * my) - y4 * my);
        h[j + 2] = 2 * mz * Ms_inv * (K1 + 2 * K2 * (1 - mz * mz));
        const double u = 1 - mz * mz;
        t->energy[I] = (K1 * u + K2 * u * u + K3 * (x4 * x2 - 15 * x4 * y2 + 15 * x2 * y4 - y4 * y2)) * s->volume;
    }
}

/* add_torque(sim, fun): src/micro/add_field.jl:998-1006 */
int orc_add_torque(orc_sim *s, void (*fn)(double, const double *, double *, void *), void *user)
{
    term_t *t = new_term(s, T_TORQUE_FN);
    t->tfn = fn;
    t->uuser = user;
    t->a = (double *)calloc(3 * (size_t)s->N, sizeof(double));
    return s->nterms - 1;
}

/* TorqueField: field.jl:282-293 then torque_kernel! src/micro/kernels.jl:639-651 */
static void field_torque_fn(orc_sim *s, term_t *t, const double *m)
{
    /* torque.field: zero when the term is added, then handed to the function
     * holding the previous evaluation's field (torque_kernel! writes H into it) */
    double *y = t->a;
    t->tfn(s->t_eval, m, y, t->uuser);
    const double a = 1 / s->gamma;
    for (int I = 0; I < s->N; I++) {
        size_t j = 3 * (size_t)I;
        const double mx = m[j], my = m[j + 1], mz = m[j + 2], hx = y[j], hy = y[j + 1], hz = y[j + 2];
        t->field[j] = a * cross_x(mx, my, mz, hx, hy, hz);
        t->field[j + 1] = a * cross_y(mx, my, mz, hx, hy, hz);
        t->field[j + 2] = a * cross_z(mx, my, mz, hx, hy, hz);
        t->energy[I] = 0;
    }
    memcpy(y, t->field, 3 * (size_t)s->N * sizeof(double));
}

/* add_magnetoelastic(sim; model, lambda_s, B1, B2, sigma / strain): src/micro/add_field.jl:1367-1405;
 * data: 6N values in Voigt order per cell (the reference's (6, nx, ny, nz) array) */
int orc_add_magnetoelastic(orc_sim *s, int cubic, const double *data, double lambda_s, double B1, double B2)
{
    term_t *t = new_term(s, cubic ? T_MEL_C : T_MEL_T);
    t->a = dup_array(data, 6 * (size_t)s->N);
    t->p[0] = lambda_s;
    t->p[1] = B1;
    t->p[2] = B2;
    return s->nterms - 1;
}

/* magnetoelastic_tensor_kernel! / magnetoelastic_cubic_kernel!: src/micro/kernels.jl:742-845 */
static void field_mel(orc_sim *s, term_t *t, const double *m)
{
    const double *Ms = s->mu0_Ms, V = s->volume;
#pragma omp parallel for schedule(static)
    for (int id = 0; id < s->N; id++) {
        size_t j = 3 * (size_t)id;
        const double *q = t->a + 6 * (size_t)id;
        double *h = t->field;
        if (Ms[id] == 0) {
            t->energy[id] = 0;
            h[j] = h[j + 1] = h[j + 2] = 0;
            continue;
        }
        const double mx = m[j], my = m[j + 1], mz = m[j + 2];
        const double xx = q[0], yy = q[1], zz = q[2], xy = q[3], xz = q[4], yz = q[5];
        if (t->kind == T_MEL_T) {
            const double ls = t->p[0];
            t->energy[id] = -1.5 * ls *
                            (xx * mx * mx + yy * my * my + zz * mz * mz + 2 * xy * mx * my + 2 * xz * mx * mz +
                             2 * yz * my * mz) * V;
            const double f = 3 * ls / Ms[id];
            h[j] = f * (xx * mx + xy * my + xz * mz);
            h[j + 1] = f * (xy * mx + yy * my + yz * mz);
            h[j + 2] = f * (xz * mx + yz * my + zz * mz);
        } else {
            const double B1 = t->p[1], B2 = t->p[2];
            t->energy[id] = (B1 * (xx * mx * mx + yy * my * my + zz * mz * mz) +
                             2 * B2 * (xy * mx * my + xz * mx * mz + yz * my * mz)) * V;
            const double f = 2 / Ms[id];
            h[j] = -f * (B1 * xx * mx + B2 * (xy * my + xz * mz));
            h[j + 1] = -f * (B1 * yy * my + B2 * (xy * mx + yz * mz));
            h[j + 2] = -f * (B1 * zz * mz + B2 * (xz * mx + yz * my));
        }
    }
}

/* add_sahe_torque: src/micro/add_field.jl:892-936 (six 3N AoS fields).  Kept as
 * a = sigma_s x a1, b = sigma_sa1 x a1 (cross_product, src/util.jl:35-46),
 * c = sigma_sa2, ext[0] = a2, ext[1] = a3; p[0] = beta */
int orc_add_sahe_torque(orc_sim *s, const double *sigma_s, const double *sigma_sa1, const double *sigma_sa2,
                        const double *a1, const double *a2, const double *a3, double beta)
{
    term_t *t = new_term(s, T_SAHE);
    const size_t n3 = 3 * (size_t)s->N;
    t->a = dup_array(sigma_s, n3);
    t->b = dup_array(sigma_sa1, n3);
    for (int i = 0; i < s->N; i++) {
        const double *y = a1 + 3 * (size_t)i;
        double *x = t->a + 3 * (size_t)i, *z = t->b + 3 * (size_t)i;
        const double r0 = -x[2] * y[1] + x[1] * y[2], r1 = x[2] * y[0] - x[0] * y[2], r2 = -x[1] * y[0] + x[0] * y[1];
        x[0] = r0;
        x[1] = r1;
        x[2] = r2;
        const double q0 = -z[2] * y[1] + z[1] * y[2], q1 = z[2] * y[0] - z[0] * y[2], q2 = -z[1] * y[0] + z[0] * y[1];
        z[0] = q0;
        z[1] = q1;
        z[2] = q2;
    }
    t->c = dup_array(sigma_sa2, n3);
    t->ext[0] = dup_array(a2, n3);
    t->ext[1] = dup_array(a3, n3);
    t->p[0] = beta;
    return s->nterms - 1;
}

/* sahe_torque_kernel!: src/micro/kernels.jl:589-613 (no energy) */
static void field_sahe(orc_sim *s, term_t *t, const double *m)
{
    const double beta = t->p[0], gamma = s->gamma, *Ms = s->mu0_Ms;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < s->N; I++) {
        size_t j = 3 * (size_t)I;
        double *h = t->field;
        t->energy[I] = 0;
        if (Ms[I] == 0) {
            h[j] = h[j + 1] = h[j + 2] = 0;
            continue;
        }
        const double mx = m[j], my = m[j + 1], mz = m[j + 2];
        const double *A2 = t->ext[0] + j, *A3 = t->ext[1] + j, *S1 = t->a + j, *S2 = t->b + j, *S3 = t->c + j;
        double ma2 = mx * A2[0] + my * A2[1] + mz * A2[2];
        ma2 = ma2 * ma2;
        const double ma3 = mx * A3[0] + my * A3[1] + mz * A3[2];
        const double sx = S1[0] - ma2 * S2[0] + ma3 * S3[0] * mx;
        const double sy = S1[1] - ma2 * S2[1] + ma3 * S3[1] * my;
        const double sz = S1[2] - ma2 * S2[2] + ma3 * S3[2] * mz;
        h[j] = (beta * sx + cross_x(mx, my, mz, sx, sy, sz)) / gamma;
        h[j + 1] = (beta * sy + cross_y(mx, my, mz, sx, sy, sz)) / gamma;
        h[j + 2] = (beta * sz + cross_z(mx, my, mz, sx, sy, sz)) / gamma;
    }
}

/* add_twin_mono_anis: src/micro/add_field.jl:629-660; axes 3N AoS, normalised here
 * (normalise, src/helper.jl:255-271); p = {Ka, Kb, Kaa, Kbb, Kab, Ku} */
int orc_add_twin_mono_anis(orc_sim *s, const double K[6], const double *axis_a, const double *axis_b,
                           const double *axis_u)
{
    term_t *t = new_term(s, T_TWIN);
    for (int q = 0; q < 6; q++) t->p[q] = K[q];
    double **dst[3] = {&t->a, &t->b, &t->c};
    const double *src[3] = {axis_a, axis_b, axis_u};
    for (int q = 0; q < 3; q++) {
        double *v = dup_array(src[q], 3 * (size_t)s->N);
        for (int i = 0; i < s->N; i++) {
            double *w = v + 3 * (size_t)i;
            double m2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
            if (m2 > 0) {
                double f = 1.0 / sqrt(m2);
                w[0] *= f;
                w[1] *= f;
                w[2] *= f;
            }
        }
        *dst[q] = v;
    }
    return s->nterms - 1;
}

/* twin_monoclinic_anisotropy_kernel!: src/micro/kernels.jl:169-225 */
static void field_twin(orc_sim *s, term_t *t, const double *m)
{
    const double Ka = t->p[0], Kb = t->p[1], Kaa = t->p[2], Kbb = t->p[3], Kab = t->p[4], Ku = t->p[5];
    const double *Ms = s->mu0_Ms;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < s->N; I++) {
        size_t j = 3 * (size_t)I;
        double *h = t->field;
        if (Ms[I] == 0) {
            t->energy[I] = 0;
            h[j] = h[j + 1] = h[j + 2] = 0;
            continue;
        }
        const double *A = t->a + j, *B = t->b + j, *U = t->c + j;
        const double mx = m[j], my = m[j + 1], mz = m[j + 2];
        const double al_a = mx * A[0] + my * A[1] + mz * A[2];
        const double al_b = mx * B[0] + my * B[1] + mz * B[2];
        const double al_u = mx * U[0] + my * U[1] + mz * U[2];
        const double a2 = al_a * al_a, b2 = al_b * al_b;
        const double dEa = 2 * Ka * al_a + 4 * Kaa * a2 * al_a + 2 * Kab * al_a * b2;
        const double dEb = 2 * Kb * al_b + 4 * Kbb * b2 * al_b + 2 * Kab * al_b * a2;
        const double dEu = -2 * Ku * al_u;
        const double Ms_inv = 1 / Ms[I];
        for (int c = 0; c < 3; c++) h[j + c] = -Ms_inv * (dEa * A[c] + dEb * B[c] + dEu * U[c]);
        t->energy[I] = (Ka * a2 + Kb * b2 + Kaa * a2 * a2 + Kbb * b2 * b2 + Kab * a2 * b2 - Ku * al_u * al_u) *
                       s->volume;
    }
}

/* add_exch_int(sim, J; k1, k2): src/micro/add_field.jl:789-816 (J: nx*ny values, 1-based layers) */
int orc_add_exch_int(orc_sim *s, const double *J, int k1, int k2)
{
    term_t *t = new_term(s, T_EXCH_IL);
    t->a = dup_array(J, (size_t)s->nx * s->ny);
    t->p[4] = k1;
    t->p[5] = k2 == -1 ? s->nz : k2;
    return s->nterms - 1;
}

/* add_dmi_int(sim, D; k1, k2): src/micro/add_field.jl:835-860 */
int orc_add_dmi_int(orc_sim *s, double Dx, double Dy, double Dz, int k1, int k2)
{
    term_t *t = new_term(s, T_DMI_IL);
    t->p[0] = Dx;
    t->p[1] = Dy;
    t->p[2] = Dz;
    t->p[4] = k1;
    t->p[5] = k2 == -1 ? s->nz : k2;
    return s->nterms - 1;
}

/* interlayer_exch_kernel! / interlayer_dmi_kernel!: src/micro/kernels.jl:496-584, one
 * (i, j) column at a time: layer k1 is written first, then layer k2; every other cell
 * keeps a zero field. */
static void field_interlayer(orc_sim *s, term_t *t, const double *m)
{
    const int nx = s->nx, ny = s->ny, K1 = (int)t->p[4], K2 = (int)t->p[5];
    const double dz = s->dz, V = s->volume, *Ms = s->mu0_Ms;
    const double Dx = t->p[0], Dy = t->p[1], Dz = t->p[2];
    double *h = t->field;
    memset(h, 0, 3 * (size_t)s->N * sizeof(double));
    memset(t->energy, 0, (size_t)s->N * sizeof(double));
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const size_t id1 = (size_t)(K1 - 1) * nx * ny + (size_t)j * nx + i;
            const size_t id2 = (size_t)(K2 - 1) * nx * ny + (size_t)j * nx + i;
            const double mbx = m[3 * id1], mby = m[3 * id1 + 1], mbz = m[3 * id1 + 2];
            const double mtx = m[3 * id2], mty = m[3 * id2 + 1], mtz = m[3 * id2 + 2];
            const double Ms1 = Ms[id1], Ms2 = Ms[id2];
            if (!(Ms1 > 0 && Ms2 > 0)) continue;
            double *h1 = h + 3 * id1, *h2 = h + 3 * id2;
            if (t->kind == T_EXCH_IL) {
                const double J = t->a[(size_t)j * nx + i];
                double c = J / (Ms1 * dz);
                h1[0] = c * mtx;
                h1[1] = c * mty;
                h1[2] = c * mtz;
                t->energy[id1] = -0.5 * Ms1 * (h1[0] * mbx + h1[1] * mby + h1[2] * mbz) * V;
                c = J / (Ms2 * dz);
                h2[0] = c * mbx;
                h2[1] = c * mby;
                h2[2] = c * mbz;
                t->energy[id2] = -0.5 * Ms2 * (h2[0] * mtx + h2[1] * mty + h2[2] * mtz) * V;
            } else {
                double c = 1.0 / (Ms1 * dz);
                h1[0] = c * cross_x(Dx, Dy, Dz, mtx, mty, mtz);
                h1[1] = c * cross_y(Dx, Dy, Dz, mtx, mty, mtz);
                h1[2] = c * cross_z(Dx, Dy, Dz, mtx, mty, mtz);
                t->energy[id1] = -0.5 * Ms1 * (h1[0] * mbx + h1[1] * mby + h1[2] * mbz) * V;
                c = -1.0 / (Ms2 * dz);
                h2[0] = c * cross_x(Dx, Dy, Dz, mbx, mby, mbz);
                h2[1] = c * cross_y(Dx, Dy, Dz, mbx, mby, mbz);
                h2[2] = c * cross_z(Dx, Dy, Dz, mbx, mby, mbz);
                t->energy[id2] = -0.5 * Ms2 * (h2[0] * mtx + h2[1] * mty + h2[2] * mtz) * V;
            }
        }
}

/* add_zeeman(sim, H0, ft): src/micro/add_field.jl:161-200 (H0 3N) */
int orc_add_zeeman_time(orc_sim *s, const double *H, void (*fn)(double, double *, void *), void *user)
{
    term_t *t = new_term(s, T_ZEEMAN_T);
    t->a = dup_array(H, 3 * (size_t)s->N);
    t->ft3 = fn;
    t->uuser = user;
    return s->nterms - 1;
}

/* TimeZeeman: src/micro/field.jl:12-33, time_zeeman_kernel! src/micro/kernels.jl:39-49 */
static void field_zeeman_time(orc_sim *s, term_t *t, const double *m)
{
    double f[3] = {1, 1, 1};
    if (t->ft3) t->ft3(s->t_eval, f, t->uuser);
    const double V = s->volume;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < s->N; I++) {
        size_t j = 3 * (size_t)I;
        double *h = t->field;
        h[j] = t->a[j] * f[0];
        h[j + 1] = t->a[j + 1] * f[1];
        h[j + 2] = t->a[j + 2] * f[2];
        double mh = m[j] * h[j] + m[j + 1] * h[j + 1] + m[j + 2] * h[j + 2];
        t->energy[I] = -V * s->mu0_Ms[I] * mh;
    }
}

int orc_set_term_time_fn(orc_sim *s, int term, double (*fn)(double, void *), void *user)
{
    if (term < 0 || term >= s->nterms) return -1;
    s->terms[term].ufun = fn;
    s->terms[term].uuser = user;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* dispatch and aggregation: src/micro/field.jl:496-505, :520-536           */
/* ------------------------------------------------------------------------ */
static void term_eval(orc_sim *s, term_t *t, const double *m)
{
    switch (t->kind) {
    case T_ZEEMAN: energy_zeeman(s, t, m); break;
    case T_EXCH: field_exch(s, t, m); break;
    case T_EXCH_S: field_exch_spatial(s, t, m); break;
    case T_DMI_BULK: field_dmi_bulk(s, t, m); break;
    case T_DMI_BULK_S: field_dmi_bulk_spatial(s, t, m); break;
    case T_DMI_INT: field_dmi_int(s, t, m); break;
    case T_ANIS: field_anis(s, t, m, 0); break;
    case T_ANIS_S: field_anis(s, t, m, 1); break;
    case T_CUBIC: field_cubic(s, t, m); break;
    case T_DEMAG: field_demag(s, t, m); break;
    case T_DEMAG_DIRECT: field_demag_direct(s, t, m); break;
    case T_THERMAL: field_thermal(s, t, m); break;
    case T_SOT: field_sot(s, t, m); break;
    case T_STT_SL: field_stt_sl(s, t, m); break;
    case T_STT_ZL: field_stt_zl(s, t, m); break;
    case T_ZEEMAN_T: field_zeeman_time(s, t, m); break;
    case T_HEX: field_hex(s, t, m); break;
    case T_EXCH_IL:
    case T_DMI_IL: field_interlayer(s, t, m); break;
    case T_TWIN: field_twin(s, t, m); break;
    case T_SAHE: field_sahe(s, t, m); break;
    case T_MEL_T:
    case T_MEL_C: field_mel(s, t, m); break;
    case T_TORQUE_FN: field_torque_fn(s, t, m); break;
    }
}

void orc_effective_field(orc_sim *s, const double *spin, double t, double *field_out)
{
    s->t_eval = t;
    size_t n3 = 3 * (size_t)s->N;
    memset(s->field, 0, n3 * sizeof(double));
    for (int i = 0; i < s->nterms; i++) {
        term_t *tm = &s->terms[i];
        if (tm->kind != T_ZEEMAN) term_eval(s, tm, spin);
        double *f = s->field;
        const double *g = tm->field;
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < n3; q++) f[q] = f[q] + g[q];
    }
    if (field_out) memcpy(field_out, s->field, n3 * sizeof(double));
}

void orc_term_field(orc_sim *s, int term, const double *spin, double t, double *field_out,
                    double *energy_out)
{
    s->t_eval = t;
    term_t *tm = &s->terms[term];
    term_eval(s, tm, spin);
    if (field_out) memcpy(field_out, tm->field, 3 * (size_t)s->N * sizeof(double));
    if (energy_out) memcpy(energy_out, tm->energy, (size_t)s->N * sizeof(double));
}

/* pairwise sum (Julia's sum over Float64 arrays is pairwise) */
static double psum(const double *a, size_t n)
{
    if (n <= 128) {
        double r = 0;
        for (size_t i = 0; i < n; i++) r += a[i];
        return r;
    }
    size_t h = n / 2;
    return psum(a, h) + psum(a + h, n - h);
}

/* compute_system_energy: field.jl:520-536; returns the sum of each term's energy */
void orc_term_energies(orc_sim *s, const double *spin, double t, double *E_out)
{
    s->t_eval = t;
    for (int i = 0; i < s->nterms; i++) {
        term_eval(s, &s->terms[i], spin);
        E_out[i] = psum(s->terms[i].energy, s->N);
    }
}

/* ------------------------------------------------------------------------ */
/* LLG: src/llg/llg.jl:6-58                                                 */
/* ------------------------------------------------------------------------ */
void orc_set_llg(orc_sim *s, double alpha, double gamma, int precession)
{
    s->alpha = alpha;
    s->gamma = gamma;
    s->precession = precession;
}

/* set_alpha with the SpatialLLG driver: src/sim.jl:213-222 (NULL: uniform again) */
void orc_set_alpha(orc_sim *s, const double *alpha)
{
    free(s->alphas);
    s->alphas = alpha ? dup_array(alpha, s->N) : NULL;
}

void orc_llg_rhs(orc_sim *s, const double *m, double t, double *dm_dt)
{
    orc_effective_field(s, m, t, NULL);
    const double *h = s->field;
    const double alpha = s->alpha, gamma = s->gamma;
    const int precession = s->precession;
#pragma omp parallel for schedule(static)
    for (int I = 0; I < s->N; I++) {
        size_t j = 3 * (size_t)I;
        if (s->pins[I]) {
            dm_dt[j] = dm_dt[j + 1] = dm_dt[j + 2] = 0;
            continue;
        }
        /* spatial_llg_rhs_kernel! (llg.jl:62-97) takes alpha per cell */
        const double al = s->alphas ? s->alphas[I] : alpha;
        double a = -gamma / (1 + al * al);
        double mx = m[j], my = m[j + 1], mz = m[j + 2];
        double mm = mx * mx + my * my + mz * mz;
        double mh = mx * h[j] + my * h[j + 1] + mz * h[j + 2];
        double h1 = mm * h[j] - mh * mx;
        double h2 = mm * h[j + 1] - mh * my;
        double h3 = mm * h[j + 2] - mh * mz;
        double f1 = 0, f2 = 0, f3 = 0;
        if (precession) {
            f1 = cross_x(mx, my, mz, h1, h2, h3);
            f2 = cross_y(mx, my, mz, h1, h2, h3);
            f3 = cross_z(mx, my, mz, h1, h2, h3);
        }
        dm_dt[j] = a * (f1 - h1 * al);
        dm_dt[j + 1] = a * (f2 - h2 * al);
        dm_dt[j + 2] = a * (f3 - h3 * al);
    }
}

/* ------------------------------------------------------------------------ */
/* integrators                                                              */
/* ------------------------------------------------------------------------ */
static void ensure_integrator_arrays(orc_sim *s)
{
    size_t n3 = 3 * (size_t)s->N;
    for (int i = 0; i < 7; i++)
        if (!s->k[i]) s->k[i] = calloc(n3, sizeof(double));
    if (!s->y_current) s->y_current = calloc(n3, sizeof(double));
    if (!s->y_next) s->y_next = calloc(n3, sizeof(double));
    if (!s->y_temp) s->y_temp = calloc(n3, sizeof(double));
}

/* create_driver defaults (src/driver.jl:126-145): Heun 1e-14, RK4 5e-14, tol 1e-6 */
void orc_set_integrator(orc_sim *s, int kind, double step, double tol)
{
    ensure_integrator_arrays(s);
    s->ikind = kind;
    s->t = 0;
    s->nsteps = 0;
    s->nfevals = 0;
    s->succeed = 0;
    s->step_next = 0;
    s->tol = tol;
    s->step = (kind == I_RK4 || kind == I_HEUN) ? step : 0.0;
    integrator_set_initial(s);
}

/* set_initial_condition!: src/integrator/adaptive_rk.jl:234-238 */
static void integrator_set_initial(orc_sim *s)
{
    if (s->y_current) memcpy(s->y_current, s->spin, 3 * (size_t)s->N * sizeof(double));
}

/* RungeKutta.advance_step: src/integrator/rk.jl:23-52 */
static int rk4_step(orc_sim *s)
{
    size_t n3 = 3 * (size_t)s->N;
    double h = s->step;
    double *k1 = s->k[0], *k2 = s->k[1], *k3 = s->k[2], *k4 = s->k[3];
    double *y = s->spin, *ys = s->prespin;
    orc_llg_rhs(s, y, s->t, k1);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++) ys[i] = y[i] + 0.5 * h * k1[i];
    orc_llg_rhs(s, ys, s->t + 0.5 * h, k2);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++) ys[i] = y[i] + 0.5 * h * k2[i];
    orc_llg_rhs(s, ys, s->t + 0.5 * h, k3);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++) ys[i] = y[i] + h * k3[i];
    orc_llg_rhs(s, ys, s->t + h, k4);
    const double c = 1.0 / 6 * h;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++) {
        ys[i] = y[i];
        y[i] += c * (k1[i] + 2 * k2[i] + 2 * k3[i] + k4[i]);
    }
    normalise(y, s->N);
    s->nsteps += 1;
    s->t = s->nsteps * h;
    return 0;
}

/* ModifiedEuler.advance_step: src/integrator/heun.jl:28-49 */
static int heun_step(orc_sim *s)
{
    size_t n3 = 3 * (size_t)s->N;
    double h = s->step;
    double *k1 = s->k[0], *k2 = s->k[1];
    double *y = s->spin, *ys = s->prespin;
    orc_llg_rhs(s, y, s->t, k1);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++) ys[i] = y[i] + h * k1[i]; /* vector_add2 util.jl:86 */
    orc_llg_rhs(s, ys, s->t + h, k2);
    const double c = 0.5 * h;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++) {
        ys[i] = y[i];
        y[i] = ys[i] + c * k1[i] + c * k2[i]; /* vector_add3 util.jl:97 */
    }
    normalise(y, s->N);
    s->nsteps += 1;
    s->t = s->nsteps * h;
    return 0;
}

static double max_err_norm(const double *err, const double *yn, size_t n3)
{
    double m = 0;
    int nan = 0;
    for (size_t i = 0; i < n3; i++) {
        double v = fabs(err[i] / (1 + fabs(yn[i])));
        if (isnan(v)) nan = 1;
        if (v > m) m = v;
    }
    return nan ? NAN : m;
}

/* rk_step_dopri54!: src/integrator/adaptive_rk_steps.jl:49-105 (FSAL off) */
static double dopri54_try(orc_sim *s, double step, double t)
{
    static const double c[6] = {1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1.0, 1.0};
    static const double a1 = 1.0 / 5;
    static const double a2[2] = {3.0 / 40, 9.0 / 40};
    static const double a3[3] = {44.0 / 45, -56.0 / 15, 32.0 / 9};
    static const double a4[4] = {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729};
    static const double a5[5] = {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176,
                                 -5103.0 / 18656};
    static const double a6[6] = {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784,
                                 11.0 / 84};
    static const double w[7] = {71.0 / 57600, 0, -71.0 / 16695, 71.0 / 1920, -17253.0 / 339200,
                                22.0 / 525, -1.0 / 40};
    size_t n3 = 3 * (size_t)s->N;
    double **k = s->k;
    double *y = s->y_current, *yn = s->y_next, *yt = s->y_temp;
    orc_llg_rhs(s, y, t, k[0]);
    s->nfevals += 1;
    double b1 = a1 * step;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++) yn[i] = y[i] + b1 * k[0][i];
    orc_llg_rhs(s, yn, t + c[0] * step, k[1]);
    double b21 = a2[0] * step, b22 = a2[1] * step;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++) yn[i] = y[i] + b21 * k[0][i] + b22 * k[1][i];
    orc_llg_rhs(s, yn, t + c[1] * step, k[2]);
    double b31 = a3[0] * step, b32 = a3[1] * step, b33 = a3[2] * step;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++) yn[i] = y[i] + b31 * k[0][i] + b32 * k[1][i] + b33 * k[2][i];
    orc_llg_rhs(s, yn, t + c[2] * step, k[3]);
    double b41 = a4[0] * step, b42 = a4[1] * step, b43 = a4[2] * step, b44 = a4[3] * step;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++)
        yn[i] = y[i] + b41 * k[0][i] + b42 * k[1][i] + b43 * k[2][i] + b44 * k[3][i];
    orc_llg_rhs(s, yn, t + c[3] * step, k[4]);
    double b51 = a5[0] * step, b52 = a5[1] * step, b53 = a5[2] * step, b54 = a5[3] * step,
           b55 = a5[4] * step;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++)
        yn[i] = y[i] + b51 * k[0][i] + b52 * k[1][i] + b53 * k[2][i] + b54 * k[3][i] +
                b55 * k[4][i];
    orc_llg_rhs(s, yn, t + step, k[5]);
    double b61 = a6[0] * step, b63 = a6[2] * step, b64 = a6[3] * step, b65 = a6[4] * step,
           b66 = a6[5] * step;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++)
        yn[i] = y[i] + b61 * k[0][i] + b63 * k[2][i] + b64 * k[3][i] + b65 * k[4][i] +
                b66 * k[5][i];
    normalise(yn, s->N);
    orc_llg_rhs(s, yn, t + step, k[6]);
    s->nfevals += 6;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++)
        yt[i] = w[0] * k[0][i] + w[2] * k[2][i] + w[3] * k[3][i] + w[4] * k[4][i] +
                w[5] * k[5][i] + w[6] * k[6][i];
    return max_err_norm(yt, yn, n3) * step / s->tol;
}

/* rk_step_bs23!: src/integrator/adaptive_rk_steps.jl:11-47 (FSAL on) */
static double bs23_try(orc_sim *s, double step, double t)
{
    size_t n3 = 3 * (size_t)s->N;
    double **k = s->k;
    double *y = s->y_current, *yn = s->y_next, *yt = s->y_temp;
    if (s->succeed && t >= s->t) {
        memcpy(k[0], k[3], n3 * sizeof(double));
    } else {
        orc_llg_rhs(s, y, t, k[0]);
        s->nfevals += 1;
    }
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++) yt[i] = y[i] + step * 0.5 * k[0][i];
    orc_llg_rhs(s, yt, t + (1.0 / 2) * step, k[1]);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++) yt[i] = y[i] + step * 0.75 * k[1][i];
    orc_llg_rhs(s, yt, t + (3.0 / 4) * step, k[2]);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++)
        yn[i] = y[i] + step * (2.0 / 9 * k[0][i] + 1.0 / 3 * k[1][i] + 4.0 / 9 * k[2][i]);
    normalise(yn, s->N);
    orc_llg_rhs(s, yn, t + step, k[3]);
    s->nfevals += 3;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++)
        yt[i] = -(5.0 / 72) * k[0][i] + (1.0 / 12) * k[1][i] + (1.0 / 9) * k[2][i] +
                (-1.0 / 8) * k[3][i];
    return max_err_norm(yt, yn, n3) * step / s->tol;
}

/* compute_initial_step: src/integrator/adaptive_rk.jl:151-165 */
static double compute_initial_step(orc_sim *s, double dt, int order_main)
{
    double abs_step = dt, abs_step_tmp = dt;
    size_t n3 = 3 * (size_t)s->N;
    s->step = 1e-15;
    orc_llg_rhs(s, s->y_current, s->t, s->y_temp);
    double mx = 0;
    for (size_t i = 0; i < n3; i++)
        if (fabs(s->y_temp[i]) > mx) mx = fabs(s->y_temp[i]);
    double r_step = mx / (0.824 * pow(s->tol, 1.0 / order_main));
    s->nfevals += 1;
    if (abs_step * r_step > 0.001) abs_step_tmp = 0.001 / r_step;
    return abs_step < abs_step_tmp ? abs_step : abs_step_tmp;
}

/* AdaptiveRK.advance_step: src/integrator/adaptive_rk.jl:93-148 */
static int adaptive_step(orc_sim *s)
{
    const int max_nsteps = 50;
    double t = s->t;
    int order_main = 5, order = 4;
    if (s->ikind == I_BS23) {
        order_main = 3;
        order = 2;
    }
    if (s->step_next <= 0) s->step_next = compute_initial_step(s, 1e-12, order_main);
    s->step = s->step_next;
    int nstep = 1;
    for (;;) {
        double max_error = s->ikind == I_BS23 ? bs23_try(s, s->step, t) : dopri54_try(s, s->step, t);
        if (isnan(max_error)) {
            s->step_next = 1e-14;
            return -3;
        }
        nstep += 1;
        if (nstep > max_nsteps) return -4;
        s->succeed = (max_error <= 1.0);
        if (s->succeed) {
            memcpy(s->y_current, s->y_next, 3 * (size_t)s->N * sizeof(double));
            s->nsteps += 1;
            s->t += s->step;
            double factor = 0.824 * pow(1.0 / max_error, 1.0 / (order + 1));
            double f = factor < 0.2 ? 0.2 : factor;
            s->step_next = s->step * (5.0 < f ? 5.0 : f);
            break;
        } else {
            double factor = 0.824 * pow(1.0 / max_error, 1.0 / order);
            double f = factor < 0.2 ? 0.2 : factor;
            s->step = s->step * (5.0 < f ? 5.0 : f);
        }
    }
    return 0;
}

int orc_advance_step(orc_sim *s)
{
    ensure_integrator_arrays(s);
    switch (s->ikind) {
    case I_RK4: return rk4_step(s);
    case I_HEUN: return heun_step(s);
    default: {
        int r = adaptive_step(s);
        return r;
    }
    }
}

/* run_until + integrate_to_time: src/sim.jl:547-571, adaptive_rk.jl:168-186 */
int orc_run_until(orc_sim *s, double t_end)
{
    ensure_integrator_arrays(s);
    if (s->ikind == I_RK4 || s->ikind == I_HEUN) return -5;
    if (t_end < s->t - s->step) return 0;
    if (t_end == s->t) return 0;
    size_t n3 = 3 * (size_t)s->N;
    memcpy(s->prespin, s->spin, n3 * sizeof(double));
    int rc = 0;
    while (s->t < t_end) {
        if (s->step_next + s->t > t_end) s->step_next = t_end - s->t;
        rc = adaptive_step(s);
        if (rc != 0) break;
    }
    memcpy(s->spin, s->y_current, n3 * sizeof(double));
    return rc;
}

void orc_integrator_state(const orc_sim *s, double *t, double *step, double *step_next,
                          int64_t *nsteps, int64_t *nfevals)
{
    if (t) *t = s->t;
    if (step) *step = s->step;
    if (step_next) *step_next = s->step_next;
    if (nsteps) *nsteps = s->nsteps;
    if (nfevals) *nfevals = s->nfevals;
}

/* relax with the LLG driver: src/sim.jl:392-490 (compute_dm! helper.jl:273-286) */
int orc_relax(orc_sim *s, int max_steps, double stopping_dmdt, int *steps_done,
              double *max_dmdt_out)
{
    ensure_integrator_arrays(s);
    size_t n3 = 3 * (size_t)s->N;
    const double dmdt_factor = (2 * M_PI / 360) * 1e9;
    double max_dmdt = 0;
    int i, rc = 0;
    for (i = 1; i <= max_steps; i++) {
        if (s->ikind == I_RK4 || s->ikind == I_HEUN) {
            rc = orc_advance_step(s); /* fixed-step steppers copy spin to prespin themselves */
        } else {
            rc = adaptive_step(s);
            memcpy(s->prespin, s->spin, n3 * sizeof(double));
            memcpy(s->spin, s->y_current, n3 * sizeof(double));
        }
        if (rc != 0) break;
        double mx = 0;
        for (int I = 0; I < s->N; I++) {
            double ax = s->prespin[3 * I] - s->spin[3 * I];
            double ay = s->prespin[3 * I + 1] - s->spin[3 * I + 1];
            double az = s->prespin[3 * I + 2] - s->spin[3 * I + 2];
            double d = sqrt(ax * ax + ay * ay + az * az);
            if (d > mx) mx = d;
        }
        max_dmdt = mx / s->step;
        if (max_dmdt < stopping_dmdt * dmdt_factor) break;
    }
    if (steps_done) *steps_done = i > max_steps ? max_steps : i;
    if (max_dmdt_out) *max_dmdt_out = max_dmdt;
    return rc;
}
