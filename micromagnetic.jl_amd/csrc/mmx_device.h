// mmx_device.h -- per-cell device code shared by the local-field kernel and the
// last demag pass: the interaction terms, the LLG right-hand side and the fused
// integrator-stage epilogue.
//
// Formulas restate src/micro/kernels.jl (terms), src/llg/llg.jl:6-40 (LLG),
// src/util.jl:86-183 and src/integrator/*.jl (stage combinations),
// src/helper.jl:255-271 (normalise).  Per term, the arithmetic keeps the
// reference's operation order; the sum over terms follows insertion order
// (src/micro/field.jl:496-505).  Floating-point contraction is off in the
// per-cell code so that every product and sum rounds the way the reference's
// CPU path does (this code is bandwidth-bound; fused multiply-adds buy nothing).
#pragma once
#include "mmx_common.h"

namespace mmx {

// ---- counter-based normal variates for the thermal field ------------------
// Philox4x32-10 (Salmon et al., SC'11) keyed by the stream seed; counter =
// (pair index, integrator step).  Box-Muller over two 53-bit uniforms.
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
        uint32_t lo0 = 0xD2511F53u * c[0];
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
        uint32_t lo1 = 0xCD9E8D57u * c[2];
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ void philox_normal2(uint64_t seed, uint64_t step, uint64_t pair,
                                               double &n_cos, double &n_sin)
{
    uint32_t c[4] = {(uint32_t)pair, (uint32_t)(pair >> 32), (uint32_t)step, (uint32_t)(step >> 32)};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    double u1 = ((double)(((uint64_t)(c[0] >> 6) << 27) | (c[1] >> 5)) + 0.5) * (1.0 / 9007199254740992.0);
    double u2 = ((double)(((uint64_t)(c[2] >> 6) << 27) | (c[3] >> 5)) + 0.5) * (1.0 / 9007199254740992.0);
    double r = sqrt(-2.0 * log(u1));
    double sn, cs;
    sincospi(2.0 * u2, &sn, &cs);
    n_cos = r * cs;
    n_sin = r * sn;
}

__device__ __forceinline__ double philox_normal(uint64_t seed, uint64_t step, uint64_t idx)
{
    double a, b;
    philox_normal2(seed, step, idx >> 1, a, b);
    return (idx & 1) ? b : a;
}

// ---- neighbours: indexpbc of src/micro/mesh.jl:31-43 ---------------------
// order (-x, +x, -y, +y, -z, +z); -1 when there is no neighbour.
template <class T>
__device__ __forceinline__ void neighbours(const FieldParams<T> &P, int i, int j, int k, int64_t nb[6])
{
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    const int64_t sxy = (int64_t)nx * ny;
    const int64_t base = (int64_t)k * sxy + (int64_t)j * nx;
    int t;
    t = i - 1;
    if (P.px && t < 0) t += nx;
    nb[0] = (t >= 0) ? base + t : -1;
    t = i + 1;
    if (P.px && t >= nx) t -= nx;
    nb[1] = (t < nx) ? base + t : -1;
    const int64_t bx = (int64_t)k * sxy + i;
    t = j - 1;
    if (P.py && t < 0) t += ny;
    nb[2] = (t >= 0) ? bx + (int64_t)t * nx : -1;
    t = j + 1;
    if (P.py && t >= ny) t -= ny;
    nb[3] = (t < ny) ? bx + (int64_t)t * nx : -1;
    const int64_t by = (int64_t)j * nx + i;
    t = k - 1;
    if (P.pz && t < 0) t += nz;
    nb[4] = (t >= 0) ? (int64_t)t * sxy + by : -1;
    t = k + 1;
    if (P.pz && t >= nz) t -= nz;
    nb[5] = (t < nz) ? (int64_t)t * sxy + by : -1;
}

// Field (A/m) of every term selected by `mask` at cell I, summed in insertion
// order; hd = demag field of this cell when a demag term is present.
// If esum != nullptr, it receives the summed energy (J) of the selected terms.
template <class T>
__device__ __forceinline__ void cell_field(const FieldParams<T> &P, const Dyn &dyn, uint32_t mask, int64_t I,
                                           int i, int j, int k, const T *__restrict__ m, T hdx, T hdy,
                                           T hdz, T &Hx, T &Hy, T &Hz, double *esum)
{
#pragma clang fp contract(off)
    const int64_t N = P.N;
    const T mx = m[I], my = m[N + I], mz = m[2 * N + I];
    const T Ms = P.mu0Ms[I];
    const T V = P.volume;
    int64_t nb[6];
    neighbours(P, i, j, k, nb);
    T nmx[6], nmy[6], nmz[6];
    bool ok[6];
#pragma unroll
    for (int q = 0; q < 6; q++) {
        int64_t id = nb[q];
        ok[q] = false;
        nmx[q] = nmy[q] = nmz[q] = T(0);
        if (id >= 0) {
            ok[q] = P.mu0Ms[id] > T(0);
            nmx[q] = m[id];
            nmy[q] = m[N + id];
            nmz[q] = m[2 * N + id];
        }
    }
    Hx = Hy = Hz = T(0);
    for (int tt = 0; tt < P.nterms; tt++) {
        if (!((mask >> tt) & 1u)) continue;
        const TermDev<T> &D = P.t[tt];
        T fx = 0, fy = 0, fz = 0;  // "f" of the exchange/DMI kernels before 1/Ms
        T hx = 0, hy = 0, hz = 0;
        double e = 0;
        switch (D.kind) {
        case K_ZEEMAN:
        case K_ZEEMAN_U: {  // static field; energy src/micro/kernels.jl:14-20
            if (D.kind == K_ZEEMAN) {
                hx = D.a[I];
                hy = D.b[I];
                hz = D.c[I];
            } else {
                hx = D.s[0];
                hy = D.s[1];
                hz = D.s[2];
            }
            T mh = mx * hx + my * hy + mz * hz;
            e = (double)(-V * Ms * mh);
        } break;
        case K_EXCH: {  // uniform_exchange_kernel!: kernels.jl:269-304
            if (Ms == T(0)) break;
            const T nab[6] = {D.s[0], D.s[0], D.s[1], D.s[1], D.s[2], D.s[2]};
#pragma unroll
            for (int q = 0; q < 6; q++) {
                if (ok[q]) {
                    fx += nab[q] * (nmx[q] - mx);
                    fy += nab[q] * (nmy[q] - my);
                    fz += nab[q] * (nmz[q] - mz);
                }
            }
            T Ms_inv = T(1.0) / Ms;
            e = (double)(T(-0.5) * (fx * mx + fy * my + fz * mz) * V);
            hx = fx * Ms_inv;
            hy = fy * Ms_inv;
            hz = fz * Ms_inv;
        } break;
        case K_EXCH_S: {  // exchange_kernel!: kernels.jl:226-266
            T A_I = D.a[I];
            if (Ms == T(0) || A_I == T(0)) break;
            const T nab[6] = {D.s[0], D.s[0], D.s[1], D.s[1], D.s[2], D.s[2]};
#pragma unroll
            for (int q = 0; q < 6; q++) {
                if (ok[q]) {
                    T A_nb = D.a[nb[q]];
                    if (A_nb != T(0)) {
                        T A_eff = T(2) * A_I * A_nb / (A_I + A_nb);
                        T coeff = A_eff * nab[q];
                        fx += coeff * (nmx[q] - mx);
                        fy += coeff * (nmy[q] - my);
                        fz += coeff * (nmz[q] - mz);
                    }
                }
            }
            T Ms_inv = T(1.0) / Ms;
            e = (double)(T(-0.5) * (fx * mx + fy * my + fz * mz) * V);
            hx = fx * Ms_inv;
            hy = fy * Ms_inv;
            hz = fz * Ms_inv;
        } break;
        case K_DMI_BULK: {  // bulkdmi_kernel!: kernels.jl:306-344, a_j x m_j expanded
            if (Ms == T(0)) break;
            const T Dsx = D.s[0], Dsy = D.s[1], Dsz = D.s[2];
            if (ok[0]) { fy += Dsx * (-nmz[0]); fz += Dsx * nmy[0]; }
            if (ok[1]) { fy += Dsx * nmz[1];    fz += Dsx * (-nmy[1]); }
            if (ok[2]) { fx += Dsy * nmz[2];    fz += Dsy * (-nmx[2]); }
            if (ok[3]) { fx += Dsy * (-nmz[3]); fz += Dsy * nmx[3]; }
            if (ok[4]) { fx += Dsz * (-nmy[4]); fy += Dsz * nmx[4]; }
            if (ok[5]) { fx += Dsz * nmy[5];    fy += Dsz * (-nmx[5]); }
            T Ms_inv = T(1.0) / Ms;
            e = (double)(T(-0.5) * (fx * mx + fy * my + fz * mz) * V);
            hx = fx * Ms_inv;
            hy = fy * Ms_inv;
            hz = fz * Ms_inv;
        } break;
        case K_DMI_BULK_S: {  // spatial_vector_bulkdmi_kernel!: kernels.jl:347-422
            if (Ms == T(0)) break;
            const T DxI = D.a[I], DyI = D.b[I], DzI = D.c[I];
#pragma unroll
            for (int q = 0; q < 2; q++) {
                if (ok[q]) {
                    T Dn = D.a[nb[q]];
                    if (DxI != T(0) && Dn != T(0)) {
                        T De = T(2) * DxI * Dn / (DxI + Dn);
                        T ax = q == 0 ? D.s[0] : -D.s[0];
                        fy += De * (-ax * nmz[q]);
                        fz += De * (ax * nmy[q]);
                    }
                }
            }
#pragma unroll
            for (int q = 2; q < 4; q++) {
                if (ok[q]) {
                    T Dn = D.b[nb[q]];
                    if (DyI != T(0) && Dn != T(0)) {
                        T De = T(2) * DyI * Dn / (DyI + Dn);
                        T ay = q == 2 ? D.s[1] : -D.s[1];
                        fx += De * (ay * nmz[q]);
                        fz += De * (-ay * nmx[q]);
                    }
                }
            }
#pragma unroll
            for (int q = 4; q < 6; q++) {
                if (ok[q]) {
                    T Dn = D.c[nb[q]];
                    if (DzI != T(0) && Dn != T(0)) {
                        T De = T(2) * DzI * Dn / (DzI + Dn);
                        T az = q == 4 ? D.s[2] : -D.s[2];
                        fx += De * (-az * nmy[q]);
                        fy += De * (az * nmx[q]);
                    }
                }
            }
            T Ms_inv = T(1.0) / Ms;
            e = (double)(T(-0.5) * (fx * mx + fy * my + fz * mz) * V);
            hx = fx * Ms_inv;
            hy = fy * Ms_inv;
            hz = fz * Ms_inv;
        } break;
        case K_DMI_INT: {  // interfacial_dmi_kernel!: kernels.jl:426-466
            T DI = D.uniform ? D.s[3] : D.a[I];
            if (Ms == T(0) || DI == T(0)) break;
            const T Dd[4] = {D.s[0], D.s[0], D.s[1], D.s[1]};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (ok[q]) {
                    T Dn = D.uniform ? D.s[3] : D.a[nb[q]];
                    if (Dn != T(0)) {
                        T De = T(2) * DI * Dn / (DI + Dn);
                        T c = De * Dd[q];
                        // e = (0,1,0), (0,-1,0), (-1,0,0), (1,0,0) for -x,+x,-y,+y
                        if (q == 0) { fx += c * nmz[q];    fz += c * (-nmx[q]); }
                        if (q == 1) { fx += c * (-nmz[q]); fz += c * nmx[q]; }
                        if (q == 2) { fy += c * nmz[q];    fz += c * (-nmy[q]); }
                        if (q == 3) { fy += c * (-nmz[q]); fz += c * nmy[q]; }
                    }
                }
            }
            T Ms_inv = T(1.0) / Ms;
            e = (double)(T(-0.5) * (fx * mx + fy * my + fz * mz) * V);
            hx = fx * Ms_inv;
            hy = fy * Ms_inv;
            hz = fz * Ms_inv;
        } break;
        case K_ANIS:
        case K_ANIS_S: {  // anisotropy_kernel! / spatial: kernels.jl:54-105
            if (Ms == T(0)) break;
            T ax, ay, az;
            if (D.kind == K_ANIS) {
                ax = D.s[0];
                ay = D.s[1];
                az = D.s[2];
            } else {
                ax = D.b[I];
                ay = D.b[N + I];
                az = D.b[2 * N + I];
            }
            T Ku = D.uniform ? D.s[3] : D.a[I];
            T Ms_inv = T(2.0) / Ms;
            T sa = mx * ax + my * ay + mz * az;
            hx = Ku * sa * ax * Ms_inv;
            hy = Ku * sa * ay * Ms_inv;
            hz = Ku * sa * az * Ms_inv;
            e = (double)(Ku * (T(1.0) - sa * sa) * V);
        } break;
        case K_CUBIC: {  // cubic_anisotropy_kernel!: kernels.jl:110-137
            if (Ms == T(0)) break;
            T Kc = D.uniform ? D.s[9] : D.a[I];
            T Ms_inv = T(4.0) * Kc / Ms;
            T mxp = D.s[0] * mx + D.s[1] * my + D.s[2] * mz;
            T myp = D.s[3] * mx + D.s[4] * my + D.s[5] * mz;
            T mzp = D.s[6] * mx + D.s[7] * my + D.s[8] * mz;
            T mxp3 = mxp * mxp * mxp;
            T myp3 = myp * myp * myp;
            T mzp3 = mzp * mzp * mzp;
            hx = Ms_inv * (mxp3 * D.s[0] + myp3 * D.s[3] + mzp3 * D.s[6]);
            hy = Ms_inv * (mxp3 * D.s[1] + myp3 * D.s[4] + mzp3 * D.s[7]);
            hz = Ms_inv * (mxp3 * D.s[2] + myp3 * D.s[5] + mzp3 * D.s[8]);
            e = (double)(-Kc * (mxp * mxp3 + myp * myp3 + mzp * mzp3) * V);
        } break;
        case K_DEMAG: {  // collect_h_kernel!: src/micro/demag.jl:435-449
            hx = hdx;
            hy = hdy;
            hz = hdz;
            T mh = mx * hx + my * hy + mz * hz;
            e = (double)(T(-0.5) * Ms * V * mh);
        } break;
        case K_THERMAL: {  // stochastic_field_kernel!: kernels.jl:469-494
            if (!(Ms > T(0))) break;
            T Tl = (D.uniform ? D.s[1] : D.a[I]) + D.s[0];
            T scale = (T)sqrt((T)dyn.thermal_factor * Tl / Ms);
            uint64_t base = 3 * (uint64_t)I;
            hx = (T)philox_normal(P.seed, dyn.noise_step, base) * scale;
            hy = (T)philox_normal(P.seed, dyn.noise_step, base + 1) * scale;
            hz = (T)philox_normal(P.seed, dyn.noise_step, base + 2) * scale;
            e = (double)(-Ms * V * (mx * hx + my * hy + mz * hz));
        } break;
        default: break;
        }
        Hx = Hx + hx;
        Hy = Hy + hy;
        Hz = Hz + hz;
        if (esum) *esum += e;
    }
}

// llg_rhs_kernel!: src/llg/llg.jl:6-40
template <class T>
__device__ __forceinline__ void llg_rhs(const FieldParams<T> &P, int64_t I, T mx, T my, T mz, T hx,
                                        T hy, T hz, T &dx, T &dy, T &dz)
{
#pragma clang fp contract(off)
    if (P.pins && P.pins[I]) {
        dx = dy = dz = T(0);
        return;
    }
    const T alpha = P.alpha;
    const T a = -P.gamma / (T(1) + alpha * alpha);
    T mm = mx * mx + my * my + mz * mz;
    T mh = mx * hx + my * hy + mz * hz;
    T h1 = mm * hx - mh * mx;
    T h2 = mm * hy - mh * my;
    T h3 = mm * hz - mh * mz;
    T f1 = 0, f2 = 0, f3 = 0;
    if (P.precession) {
        f1 = -mz * h2 + my * h3;
        f2 = mz * h1 - mx * h3;
        f3 = -my * h1 + mx * h2;
    }
    dx = a * (f1 - h1 * alpha);
    dy = a * (f2 - h2 * alpha);
    dz = a * (f3 - h3 * alpha);
}

// NaN-propagating max (a NaN anywhere makes the result NaN, like Julia's maximum)
__device__ __forceinline__ double pmax(double a, double b)
{
    if (a != a) return a;
    if (b != b) return b;
    return b > a ? b : a;
}

// The fused last step of one RHS evaluation for cell I.  `red` accumulates the
// thread's reduction value (ERR: running max; ENERGY: sum of the masked terms).
template <class T>
__device__ __forceinline__ void finish_cell(const FieldParams<T> &P, const Dyn &dyn, const Epi<T> &E, int64_t I,
                                            int i, int j, int k, const T *__restrict__ m, T hdx, T hdy, T hdz,
                                            double &red)
{
#pragma clang fp contract(off)
    const int64_t N = P.N;
    T Hx, Hy, Hz;
    if (E.mode == EPI_ENERGY) {
        double es = 0;
        cell_field(P, dyn, E.term_mask, I, i, j, k, m, hdx, hdy, hdz, Hx, Hy, Hz, &es);
        red += es;
        return;
    }
    if (E.mode == EPI_FIELD) {
        if (E.e_out) {
            double es = 0;
            cell_field(P, dyn, E.term_mask, I, i, j, k, m, hdx, hdy, hdz, Hx, Hy, Hz, &es);
            E.e_out[I] = (T)es;
        } else {
            cell_field(P, dyn, E.term_mask, I, i, j, k, m, hdx, hdy, hdz, Hx, Hy, Hz, (double *)nullptr);
        }
        E.out[I] = Hx;
        E.out[N + I] = Hy;
        E.out[2 * N + I] = Hz;
        return;
    }
    cell_field(P, dyn, E.term_mask, I, i, j, k, m, hdx, hdy, hdz, Hx, Hy, Hz, (double *)nullptr);
    const T mx = m[I], my = m[N + I], mz = m[2 * N + I];
    T kx, ky, kz;
    llg_rhs(P, I, mx, my, mz, Hx, Hy, Hz, kx, ky, kz);
    if (E.mode == EPI_RHS) {
        E.out[I] = kx;
        E.out[N + I] = ky;
        E.out[2 * N + I] = kz;
        return;
    }
    if (E.kstore) {
        E.kstore[I] = kx;
        E.kstore[N + I] = ky;
        E.kstore[2 * N + I] = kz;
    }
    const T kk[3] = {kx, ky, kz};
    if (E.mode == EPI_ERR) {  // vector_add6b / BS23 error combination, then |err|/(1+|y|)
        double mval = 0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int64_t o = c * N + I;
            double err = 0;
#pragma unroll
            for (int q = 0; q < 6; q++)
                if (q < E.np) err = (q == 0) ? E.cp[0] * (double)E.kp[0][o] : err + E.cp[q] * (double)E.kp[q][o];
            err = err + E.cself * (double)kk[c];
            const double v = fabs(err / (1.0 + fabs((double)m[o])));
            mval = pmax(mval, v);
        }
        red = pmax(red, mval);
        return;
    }
    T y[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int64_t o = c * N + I;
        if (E.mode == EPI_RK4_FINAL) {  // src/integrator/rk.jl:46
            T s = E.kp[0][o] + T(2) * E.kp[1][o] + T(2) * E.kp[2][o] + kk[c];
            y[c] = E.y0[o] + (T)E.cself * s;
        } else if (E.mode == EPI_STAGE_SCALED) {  // adaptive_rk_steps.jl:35
            T s = T(0);
#pragma unroll
            for (int q = 0; q < 6; q++)
                if (q < E.np) s = (q == 0) ? (T)E.cp[0] * E.kp[0][o] : s + (T)E.cp[q] * E.kp[q][o];
            s = s + (T)E.cself * kk[c];
            y[c] = E.y0[o] + (T)E.outer * s;
        } else {  // vector_add2..6: a1 + c2 a2 + ... (util.jl:86-166)
            T s = E.y0[o];
#pragma unroll
            for (int q = 0; q < 6; q++)
                if (q < E.np) s = s + (T)E.cp[q] * E.kp[q][o];
            y[c] = s + (T)E.cself * kk[c];
        }
    }
    if (E.normalize) {  // normalise: src/helper.jl:255-271
        T m2 = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
        if (m2 > T(0)) {
            T len = T(1) / sqrt(m2);
            y[0] *= len;
            y[1] *= len;
            y[2] *= len;
        }
    }
    E.yout[I] = y[0];
    E.yout[N + I] = y[1];
    E.yout[2 * N + I] = y[2];
}

// Block-wide reduction of `nval` doubles per thread into
// partial[blockIdx*stride + v]; works for any block size up to 1024.
template <int NV>
__device__ __forceinline__ void block_reduce_store(double *red, int nval, bool is_max, double *partial,
                                                   int stride)
{
    __shared__ double sred[1024];
    const int nt = blockDim.x * blockDim.y;
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    int p2 = 1;
    while (p2 < nt) p2 <<= 1;
    for (int v = 0; v < nval; v++) {
        sred[tid] = red[v];
        __syncthreads();
        for (int s = p2 / 2; s > 0; s >>= 1) {
            if (tid < s && tid + s < nt) {
                double a = sred[tid], b = sred[tid + s];
                sred[tid] = is_max ? pmax(a, b) : a + b;
            }
            __syncthreads();
        }
        if (tid == 0) partial[(size_t)blockIdx.x * stride + v] = sred[0];
        __syncthreads();
    }
}

}  // namespace mmx
