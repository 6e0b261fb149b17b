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

// NaN-propagating max (a NaN anywhere makes the result NaN, like Julia's maximum)
__device__ __forceinline__ double pmax(double a, double b)
{
    if (a != a) return a;
    if (b != b) return b;
    return b > a ? b : a;
}

// ---- neighbours: indexpbc of src/micro/mesh.jl:31-43 ---------------------
// order (-x, +x, -y, +y, -z, +z); -1 when there is no neighbour.
template <class T>
__device__ __forceinline__ void neighbours(const MMX_CONST FieldParams<T> &P, int i, int j, int k, int64_t nb[6])
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
__device__ __forceinline__ void cell_field(const MMX_CONST FieldParams<T> &P, const Dyn &dyn, uint32_t mask, int64_t I,
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
    // a missing neighbour reads the cell itself and is masked: no load sits
    // behind a branch, so the 24 loads issue together
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const bool ex = nb[q] >= 0;
        const int64_t id = ex ? nb[q] : I;
        const T msn = gld(P.mu0Ms, id);
        const T ax = gld(m, id), ay = gld(m, N + id), az = gld(m, 2 * N + id);
        ok[q] = ex && msn > T(0);
        nmx[q] = ex ? ax : T(0);
        nmy[q] = ex ? ay : T(0);
        nmz[q] = ex ? az : T(0);
    }
    Hx = Hy = Hz = T(0);
    for (int tt = 0; tt < P.nterms; tt++) {
        if (!((mask >> tt) & 1u)) continue;
        const MMX_CONST TermDev<T> &D = P.t[tt];
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
            // TimeBulkDMI (field.jl:145-163): Ds = (time_D(t) Dx)/dx, set per evaluation
            const T Dsx = D.tdep ? (T)dyn.tfac[tt] : D.s[0];
            const T Dsy = D.tdep ? (T)dyn.tfy[tt] : D.s[1];
            const T Dsz = D.tdep ? (T)dyn.tfz[tt] : D.s[2];
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
            // TimeSpatialBulkDMI (field.jl:180-193): every D read is time_D(t) * D
            const T ft = D.tdep ? (T)dyn.tfac[tt] : T(1);
            const T DxI = D.a[I] * ft, DyI = D.b[I] * ft, DzI = D.c[I] * ft;
#pragma unroll
            for (int q = 0; q < 2; q++) {
                if (ok[q]) {
                    T Dn = D.a[nb[q]] * ft;
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
                    T Dn = D.b[nb[q]] * ft;
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
                    T Dn = D.c[nb[q]] * ft;
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
        case K_HEX: {  // hexagonal_anisotropy_kernel!: kernels.jl:142-167
            if (Ms == T(0)) break;
            const T K1 = D.s[0], K2 = D.s[1], K3 = D.s[2];
            const T Ms_inv = T(1) / Ms;
            const T x2 = mx * mx, y2 = my * my, z2 = mz * mz;
            const T x3 = x2 * mx, y3 = y2 * my, x4 = x2 * x2, y4 = y2 * y2;
            const T c6 = T(-6) * K3 * Ms_inv;
            hx = c6 * ((x4 * mx - T(10) * x3 * y2) + T(5) * mx * y4);
            hy = c6 * ((T(-5) * x4 * my + T(10) * x2 * y3) - y4 * my);
            const T u = T(1) - z2;
            hz = T(2) * mz * Ms_inv * (K1 + T(2) * K2 * u);
            const T p6 = ((x4 * x2 - T(15) * x4 * y2) + T(15) * x2 * y4) - y4 * y2;
            e = (double)(((K1 * u + K2 * u * u) + K3 * p6) * V);
        } break;
        case K_MEL_T:
        case K_MEL_C: {  // magnetoelastic_tensor_kernel! / _cubic_kernel!: kernels.jl:742-845
            if (Ms == T(0)) break;
            const T *Q = D.a;  // Voigt planes xx, yy, zz, xy, xz, yz
            const T qxx = Q[I], qyy = Q[N + I], qzz = Q[2 * N + I];
            const T qxy = Q[3 * N + I], qxz = Q[4 * N + I], qyz = Q[5 * N + I];
            if (D.kind == K_MEL_T) {
                const T ls = D.s[0];
                const T sum = ((((qxx * mx * mx + qyy * my * my) + qzz * mz * mz) + T(2) * qxy * mx * my) +
                               T(2) * qxz * mx * mz) + T(2) * qyz * my * mz;
                e = (double)(-T(1.5) * ls * sum * V);
                const T f = T(3) * ls / Ms;
                hx = f * ((qxx * mx + qxy * my) + qxz * mz);
                hy = f * ((qxy * mx + qyy * my) + qyz * mz);
                hz = f * ((qxz * mx + qyz * my) + qzz * mz);
            } else {
                const T B1 = D.s[0], B2 = D.s[1];
                e = (double)((B1 * ((qxx * mx * mx + qyy * my * my) + qzz * mz * mz) +
                              T(2) * B2 * ((qxy * mx * my + qxz * mx * mz) + qyz * my * mz)) * V);
                const T nf = -(T(2) / Ms);
                hx = nf * (B1 * qxx * mx + B2 * (qxy * my + qxz * mz));
                hy = nf * (B1 * qyy * my + B2 * (qxy * mx + qyz * mz));
                hz = nf * (B1 * qzz * mz + B2 * (qxz * mx + qyz * my));
            }
        } break;
        case K_TORQUE_FN: {  // torque_kernel!: kernels.jl:639-651, H = (1/gamma) m x T
            const T tx = D.a[I], ty = D.a[N + I], tz = D.a[2 * N + I];
            const T a = T(1) / P.gamma;
            hx = a * (-mz * ty + my * tz);
            hy = a * (mz * tx - mx * tz);
            hz = a * (-my * tx + mx * ty);
        } break;
        case K_SAHE: {  // sahe_torque_kernel!: kernels.jl:589-613 (no energy)
            if (Ms == T(0)) break;
            const T *Q = D.a;  // plane f*3 + c of the packed vector fields
#define MMX_SAHE(f, c) Q[(int64_t)(3 * (f) + (c)) * N + I]
            const T pa2 = (mx * MMX_SAHE(3, 0) + my * MMX_SAHE(3, 1)) + mz * MMX_SAHE(3, 2);
            const T ma2 = pa2 * pa2;
            const T ma3 = (mx * MMX_SAHE(4, 0) + my * MMX_SAHE(4, 1)) + mz * MMX_SAHE(4, 2);
            const T sx = (MMX_SAHE(0, 0) - ma2 * MMX_SAHE(1, 0)) + ma3 * MMX_SAHE(2, 0) * mx;
            const T sy = (MMX_SAHE(0, 1) - ma2 * MMX_SAHE(1, 1)) + ma3 * MMX_SAHE(2, 1) * my;
            const T sz = (MMX_SAHE(0, 2) - ma2 * MMX_SAHE(1, 2)) + ma3 * MMX_SAHE(2, 2) * mz;
#undef MMX_SAHE
            const T beta = D.s[0], g = P.gamma;
            hx = (beta * sx + (-mz * sy + my * sz)) / g;
            hy = (beta * sy + (mz * sx - mx * sz)) / g;
            hz = (beta * sz + (-my * sx + mx * sy)) / g;
        } break;
        case K_TWIN: {  // twin_monoclinic_anisotropy_kernel!: kernels.jl:169-225
            if (Ms == T(0)) break;
            const T ax = D.a[I], ay = D.a[N + I], az = D.a[2 * N + I];
            const T bx = D.b[I], by = D.b[N + I], bz = D.b[2 * N + I];
            const T ux = D.c[I], uy = D.c[N + I], uz = D.c[2 * N + I];
            const T Ka = D.s[0], Kb = D.s[1], Kaa = D.s[2], Kbb = D.s[3], Kab = D.s[4], Ku = D.s[5];
            const T pa = (mx * ax + my * ay) + mz * az;
            const T pb = (mx * bx + my * by) + mz * bz;
            const T pu = (mx * ux + my * uy) + mz * uz;
            const T pa2 = pa * pa, pb2 = pb * pb;
            const T dEa = (T(2) * Ka * pa + T(4) * Kaa * pa2 * pa) + T(2) * Kab * pa * pb2;
            const T dEb = (T(2) * Kb * pb + T(4) * Kbb * pb2 * pb) + T(2) * Kab * pb * pa2;
            const T dEu = T(-2) * Ku * pu;
            const T nMs_inv = -(T(1) / Ms);
            hx = nMs_inv * ((dEa * ax + dEb * bx) + dEu * ux);
            hy = nMs_inv * ((dEa * ay + dEb * by) + dEu * uy);
            hz = nMs_inv * ((dEa * az + dEb * bz) + dEu * uz);
            e = (double)((((((Ka * pa2 + Kb * pb2) + Kaa * pa2 * pa2) + Kbb * pb2 * pb2) + Kab * pa2 * pb2) -
                          Ku * pu * pu) * V);
        } break;
        case K_EXCH_IL:
        case K_DMI_IL: {  // interlayer_exch_kernel! / interlayer_dmi_kernel!: kernels.jl:496-584
            const int k1 = (int)D.s[4], k2 = (int)D.s[5];
            // the kernel writes layer k1's field, then layer k2's (the last wins if k1 == k2)
            const bool top = k == k2, bot = !top && k == k1;
            if (!top && !bot) break;
            const int64_t Ip = I + (int64_t)((top ? k1 : k2) - k) * ((int64_t)P.nx * P.ny);
            const T Msp = P.mu0Ms[Ip];
            if (!(Ms > T(0) && Msp > T(0))) break;
            const T px = m[Ip], py = m[N + Ip], pz = m[2 * N + Ip];
            const T dz = D.s[3];
            if (D.kind == K_EXCH_IL) {
                const T c = D.a[i + (int64_t)j * P.nx] / (Ms * dz);
                hx = c * px;
                hy = c * py;
                hz = c * pz;
            } else {
                const T c = (top ? T(-1) : T(1)) / (Ms * dz);
                const T Dx = D.s[0], Dy = D.s[1], Dz = D.s[2];
                hx = c * (-Dz * py + Dy * pz);
                hy = c * (Dz * px - Dx * pz);
                hz = c * (-Dy * px + Dx * py);
            }
            e = (double)(T(-0.5) * Ms * ((hx * mx + hy * my) + hz * mz) * V);
        } break;
        case K_ZEEMAN_T: {  // time_zeeman_kernel!: kernels.jl:39-49, factors from field.jl:12-33
            const T fx = (T)dyn.tfac[tt], fy = (T)dyn.tfy[tt], fz = (T)dyn.tfz[tt];
            if (D.uniform) {
                hx = D.s[0] * fx;
                hy = D.s[1] * fy;
                hz = D.s[2] * fz;
            } else {
                hx = D.a[I] * fx;
                hy = D.b[I] * fy;
                hz = D.c[I] * fz;
            }
            T mh = mx * hx + my * hy + mz * hz;
            e = (double)(-V * Ms * mh);
        } break;
        case K_SOT: {  // df_torque_kernel!: kernels.jl:619-633 (no energy, no Ms rule)
            const T a = (D.uniform ? D.s[3] : D.a[I]) / P.gamma;
            const T b = D.s[4] / P.gamma;
            const T px = D.s[0], py = D.s[1], pz = D.s[2];
            hx = a * (-mz * py + my * pz) + b * px;
            hy = a * (mz * px - mx * pz) + b * py;
            hz = a * (-my * px + mx * py) + b * pz;
        } break;
        case K_STT_SL: {  // slonczewski_torque_kernel!: kernels.jl:657-674
            const T px = D.s[0], py = D.s[1], pz = D.s[2];
            const T lsq = D.s[4], Pp = D.s[5], xi = D.s[6];
            const T mp = mx * px + my * py + mz * pz;
            const T eps = Pp * lsq / (lsq + T(1) + (lsq - T(1)) * mp);
            const T a = (D.uniform ? D.s[3] : D.a[I]) * (T)dyn.tfac[tt];
            hx = a * (eps * (-mz * py + my * pz) + xi * px);
            hy = a * (eps * (mz * px - mx * pz) + xi * py);
            hz = a * (eps * (-my * px + mx * py) + xi * pz);
        } break;
        case K_STT_ZL: {  // zhangli_torque_kernel!: kernels.jl:680-739 (mesh neighbours only)
            T gx = 0, gy = 0, gz = 0;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const int64_t i1 = nb[2 * d], i2 = nb[2 * d + 1];
                const T dd = D.s[d];
                const T factor = (i1 >= 0 && i2 >= 0) ? T(1) / (T(2) * dd) : T(1) / dd;
                const T m1x = i1 >= 0 ? nmx[2 * d] : mx, m1y = i1 >= 0 ? nmy[2 * d] : my, m1z = i1 >= 0 ? nmz[2 * d] : mz;
                const T m2x = i2 >= 0 ? nmx[2 * d + 1] : mx, m2y = i2 >= 0 ? nmy[2 * d + 1] : my,
                        m2z = i2 >= 0 ? nmz[2 * d + 1] : mz;
                const T u = D.b[d * N + I] * factor;
                gx += u * (m2x - m1x);
                gy += u * (m2y - m1y);
                gz += u * (m2z - m1z);
            }
            const T ut = (T)dyn.tfac[tt];
            gx = ut * gx;
            gy = ut * gy;
            gz = ut * gz;
            const T xi = D.uniform ? D.s[3] : D.a[I];
            hx = (-mz * gy + my * gz) + xi * gx;
            hy = (mz * gx - mx * gz) + xi * gy;
            hz = (-my * gx + mx * gy) + xi * gz;
        } break;
        default: break;
        }
        Hx = Hx + hx;
        Hy = Hy + hy;
        Hz = Hz + hz;
        if (esum) *esum += e;
    }
}

// ---- slot form ---------------------------------------------------------------
// When every kind of term appears at most once, the kernels take this form.
// SlotCfg holds what every cell needs from the parameter block (wave-uniform
// values that the compiler keeps in scalar registers across the cell loop);
// the switches and the insertion order come packed (FieldParams::flags,
// ::ordp), and what only spatial or rarely used terms need is read inside
// their branches.  A cell issues its loads (own values, six neighbours,
// spatial parameters) before its arithmetic, and the field of each slot is
// added in insertion order exactly as cell_field adds the terms.
template <class T>
struct SlotCfg {
    uint32_t fl, ordp;
    int nx, ny, nz, N;
    const T *mu0Ms;
    const uint8_t *nbm;
    int se, sb, si, sa, sc, sz, st;
    T nab0, nab1, nab2;
    T id0, id1, iDe, iDu;
    T ax, ay, az, Ku;
    T zx, zy, zz;
};

template <class T>
__device__ __forceinline__ SlotCfg<T> slot_cfg(const MMX_CONST FieldParams<T> &P)
{
    SlotCfg<T> C;
    C.fl = P.flags;
    C.ordp = P.ordp;
    C.nx = P.nx;
    C.ny = P.ny;
    C.nz = P.nz;
    C.N = (int)P.N;
    C.mu0Ms = P.mu0Ms;
    C.nbm = P.nbm;
    C.se = max(P.slot[SL_EXCH], 0);
    C.sb = max(P.slot[SL_DMIB], 0);
    C.si = max(P.slot[SL_DMII], 0);
    C.sa = max(P.slot[SL_ANIS], 0);
    C.sc = max(P.slot[SL_CUBIC], 0);
    C.sz = max(P.slot[SL_ZEE], 0);
    C.st = max(P.slot[SL_THERM], 0);
    C.nab0 = P.t[C.se].s[0];
    C.nab1 = P.t[C.se].s[1];
    C.nab2 = P.t[C.se].s[2];
    C.id0 = P.t[C.si].s[0];
    C.id1 = P.t[C.si].s[1];
    C.iDu = P.t[C.si].s[3];
    C.iDe = P.t[C.si].s[4];
    C.ax = P.t[C.sa].s[0];
    C.ay = P.t[C.sa].s[1];
    C.az = P.t[C.sa].s[2];
    C.Ku = P.t[C.sa].s[3];
    C.zx = P.t[C.sz].s[0];
    C.zy = P.t[C.sz].s[1];
    C.zz = P.t[C.sz].s[2];
    return C;
}

template <class X>
__device__ __forceinline__ void gst(X *p, int64_t i, X v)
{
    ((__attribute__((address_space(1))) X *)p)[i] = v;
}

// m of a block's x rows staged in LDS by the last demag pass: component c of
// the cell's own row at p[c*cs + x], the block's row r' = r + d at p + d*rs
// for 0 <= r' < nrows (the rows are consecutive z at one y).
template <class T>
struct RowTile {
    const T *p;
    int cs, rs, r, nrows;
};

// Field of cell I in slot form (see SlotCfg); returns m of the cell and its
// neighbour / pin mask.  TILE: m of the cell, of its x neighbours and of the
// z neighbours inside the block come from the LDS row tile tl.
template <class T, uint32_t SIG, bool TILE = false>
__device__ __forceinline__ void cell_field_fast(const MMX_CONST FieldParams<T> &P, const Dyn &dyn, int I, int i, int j,
                                                int k, const T *__restrict__ m, T hdx, T hdy, T hdz, T &mx, T &my,
                                                T &mz, T &Hx, T &Hy, T &Hz, uint32_t &msk,
                                                const RowTile<T> &tl = RowTile<T>{})
{
#pragma clang fp contract(off)
    const SlotCfg<T> C = slot_cfg(P);
    const int N = C.N;
    constexpr uint32_t GEOM = FL_PX | FL_PY | FL_PZ | FL_PREC;
    const uint32_t fl = SIG ? (sig_flags(SIG) | (C.fl & GEOM)) : C.fl;
    const uint32_t ordp = SIG ? sig_ordp(SIG) : C.ordp;
    const int ex = (fl >> FL_EX_SHIFT) & 3, bd = (fl >> FL_BD_SHIFT) & 3, idm = (fl >> FL_ID_SHIFT) & 3;
    const int ze = (fl >> FL_ZE_SHIFT) & 3;
    // ---- loads ----
    msk = gld(C.nbm, I);
    const T Ms = gld(C.mu0Ms, I);
    if constexpr (TILE) {
        mx = lld(tl.p, i);
        my = lld(tl.p, tl.cs + i);
        mz = lld(tl.p, 2 * tl.cs + i);
    } else {
        mx = gld(m, I);
        my = gld(m, N + I);
        mz = gld(m, 2 * N + I);
    }
    const bool stencil = ex || bd || idm;
    int nb[6];
    // TILE: x index of each x neighbour in the row; block-row offset of each z
    // neighbour and whether that row is one of the block's
    int tx[2] = {i, i}, tz[2] = {0, 0};
    bool tzin[2] = {false, false};
    {
        const int nx = C.nx, ny = C.ny, nz = C.nz, sxy = nx * ny;
        int t;
        t = i - 1;
        if ((fl & FL_PX) && t < 0) t += nx;
        nb[0] = (t >= 0) ? I - i + t : I;
        tx[0] = (t >= 0) ? t : i;
        t = i + 1;
        if ((fl & FL_PX) && t >= nx) t -= nx;
        nb[1] = (t < nx) ? I - i + t : I;
        tx[1] = (t < nx) ? t : i;
        t = j - 1;
        if ((fl & FL_PY) && t < 0) t += ny;
        nb[2] = (t >= 0) ? I + (t - j) * nx : I;
        t = j + 1;
        if ((fl & FL_PY) && t >= ny) t -= ny;
        nb[3] = (t < ny) ? I + (t - j) * nx : I;
        t = k - 1;
        if ((fl & FL_PZ) && t < 0) t += nz;
        nb[4] = (t >= 0) ? I + (t - k) * sxy : I;
        if constexpr (TILE) {
            tz[0] = (t >= 0) ? t - k : 0;
            tzin[0] = tl.r + tz[0] >= 0 && tl.r + tz[0] < tl.nrows;
        }
        t = k + 1;
        if ((fl & FL_PZ) && t >= nz) t -= nz;
        nb[5] = (t < nz) ? I + (t - k) * sxy : I;
        if constexpr (TILE) {
            tz[1] = (t < nz) ? t - k : 0;
            tzin[1] = tl.r + tz[1] >= 0 && tl.r + tz[1] < tl.nrows;
        }
    }
    T nmx[6], nmy[6], nmz[6];
#pragma unroll
    for (int q = 0; q < 6; q++) nmx[q] = nmy[q] = nmz[q] = T(0);
    if (stencil) {
#pragma unroll
        for (int q = 0; q < 6; q++) {
            if constexpr (TILE) {
                if (q < 2) {
                    nmx[q] = lld(tl.p, tx[q]);
                    nmy[q] = lld(tl.p, tl.cs + tx[q]);
                    nmz[q] = lld(tl.p, 2 * tl.cs + tx[q]);
                    continue;
                }
                if (q >= 4) {
                    // one generic (flat) access reads the tile or HBM per lane:
                    // no branch, so these loads issue with the others
                    const bool in = tzin[q - 4];
                    const T *px = in ? tl.p + tz[q - 4] * tl.rs + i : m + nb[q];
                    const int64_t st = in ? (int64_t)tl.cs : (int64_t)N;
                    nmx[q] = px[0];
                    nmy[q] = px[st];
                    nmz[q] = px[2 * st];
                    continue;
                }
            }
            nmx[q] = gld(m, nb[q]);
            nmy[q] = gld(m, N + nb[q]);
            nmz[q] = gld(m, 2 * N + nb[q]);
        }
    }
    T exAI = 0, exAn[6] = {0, 0, 0, 0, 0, 0};
    if (ex == 2) {
        const T *A = P.t[C.se].a;
        exAI = gld(A, I);
#pragma unroll
        for (int q = 0; q < 6; q++) exAn[q] = gld(A, nb[q]);
    }
    T bDI[3] = {0, 0, 0}, bDn[6] = {0, 0, 0, 0, 0, 0};
    if (bd == 2) {
        const T *Dx = P.t[C.sb].a, *Dy = P.t[C.sb].b, *Dz = P.t[C.sb].c;
        bDI[0] = gld(Dx, I);
        bDI[1] = gld(Dy, I);
        bDI[2] = gld(Dz, I);
#pragma unroll
        for (int q = 0; q < 6; q++) bDn[q] = gld(q < 2 ? Dx : (q < 4 ? Dy : Dz), nb[q]);
    }
    T iDI = C.iDu, iDn[4] = {C.iDu, C.iDu, C.iDu, C.iDu};
    if (idm == 2) {
        const T *D = P.t[C.si].a;
        iDI = gld(D, I);
#pragma unroll
        for (int q = 0; q < 4; q++) iDn[q] = gld(D, nb[q]);
    }
    T ax = C.ax, ay = C.ay, az = C.az, Ku = C.Ku;
    if (fl & FL_AN_AXIS) {
        const T *A = P.t[C.sa].b;
        ax = gld(A, I);
        ay = gld(A, N + I);
        az = gld(A, 2 * N + I);
    }
    if (fl & FL_AN_KU) Ku = gld(P.t[C.sa].a, I);
    T Kc = 0;
    if (fl & FL_CU) Kc = (fl & FL_CU_KC) ? gld(P.t[C.sc].a, I) : P.t[C.sc].s[9];
    T zhx = 0, zhy = 0, zhz = 0;
    if (ze == 1) {
        zhx = C.zx;
        zhy = C.zy;
        zhz = C.zz;
        if (!SIG) {  // a uniform TimeZeeman takes this form too: H0 .* f(t) (static: factors 1)
            zhx = zhx * (T)dyn.tfac[C.sz];
            zhy = zhy * (T)dyn.tfy[C.sz];
            zhz = zhz * (T)dyn.tfz[C.sz];
        }
    } else if (ze == 2) {
        zhx = gld(P.t[C.sz].a, I);
        zhy = gld(P.t[C.sz].b, I);
        zhz = gld(P.t[C.sz].c, I);
    }
    T Tl = 0;
    if (fl & FL_TH) {
        const MMX_CONST TermDev<T> &D = P.t[C.st];
        Tl = (D.uniform ? D.s[1] : gld(D.a, I)) + D.s[0];
    }
    // ---- arithmetic ----
    const bool ex_on = ex && (Ms != T(0)) && (ex == 1 || exAI != T(0));
    const bool bd_on = bd && (Ms != T(0));
    const bool id_on = idm && (Ms != T(0)) && (iDI != T(0));
    T fex = 0, fey = 0, fez = 0, fbx = 0, fby = 0, fbz = 0, fix = 0, fiy = 0, fiz = 0;
    T bs0 = 0, bs1 = 0, bs2 = 0;
    if (bd) {
        bs0 = P.t[C.sb].s[0];
        bs1 = P.t[C.sb].s[1];
        bs2 = P.t[C.sb].s[2];
    }
    if (stencil) {
#pragma unroll
        for (int q = 0; q < 6; q++) {
            if (!((msk >> q) & 1u)) continue;
            const T ax_ = nmx[q], ay_ = nmy[q], az_ = nmz[q];
            if (ex == 1) {  // kernels.jl:269-304
                const T nab = (q < 2) ? C.nab0 : ((q < 4) ? C.nab1 : C.nab2);
                fex += nab * (ax_ - mx);
                fey += nab * (ay_ - my);
                fez += nab * (az_ - mz);
            } else if (ex == 2) {  // kernels.jl:226-266
                const T nab = (q < 2) ? C.nab0 : ((q < 4) ? C.nab1 : C.nab2);
                const T A_nb = exAn[q];
                if (A_nb != T(0)) {
                    const T coeff = (T(2) * exAI * A_nb / (exAI + A_nb)) * nab;
                    fex += coeff * (ax_ - mx);
                    fey += coeff * (ay_ - my);
                    fez += coeff * (az_ - mz);
                }
            }
            if (bd == 1) {  // bulkdmi_kernel!: kernels.jl:306-344
                if (q == 0) { fby += bs0 * (-az_); fbz += bs0 * ay_; }
                if (q == 1) { fby += bs0 * az_;    fbz += bs0 * (-ay_); }
                if (q == 2) { fbx += bs1 * az_;    fbz += bs1 * (-ax_); }
                if (q == 3) { fbx += bs1 * (-az_); fbz += bs1 * ax_; }
                if (q == 4) { fbx += bs2 * (-ay_); fby += bs2 * ax_; }
                if (q == 5) { fbx += bs2 * ay_;    fby += bs2 * (-ax_); }
            } else if (bd == 2) {  // spatial_vector_bulkdmi_kernel!: kernels.jl:347-422
                const T DI = bDI[q / 2], Dn = bDn[q];
                if (DI != T(0) && Dn != T(0)) {
                    const T De = T(2) * DI * Dn / (DI + Dn);
                    if (q < 2) {
                        const T a = q == 0 ? bs0 : -bs0;
                        fby += De * (-a * az_);
                        fbz += De * (a * ay_);
                    } else if (q < 4) {
                        const T a = q == 2 ? bs1 : -bs1;
                        fbx += De * (a * az_);
                        fbz += De * (-a * ax_);
                    } else {
                        const T a = q == 4 ? bs2 : -bs2;
                        fbx += De * (-a * ay_);
                        fby += De * (a * ax_);
                    }
                }
            }
            if (idm && q < 4) {  // interfacial_dmi_kernel!: kernels.jl:426-466
                const T Dn = iDn[q];
                if (Dn != T(0)) {
                    const T De = idm == 1 ? C.iDe : T(2) * iDI * Dn / (iDI + Dn);
                    const T c = De * (q < 2 ? C.id0 : C.id1);
                    if (q == 0) { fix += c * az_;    fiz += c * (-ax_); }
                    if (q == 1) { fix += c * (-az_); fiz += c * ax_; }
                    if (q == 2) { fiy += c * az_;    fiz += c * (-ay_); }
                    if (q == 3) { fiy += c * (-az_); fiz += c * ay_; }
                }
            }
        }
    }
    const T Ms_inv = (Ms != T(0)) ? T(1.0) / Ms : T(0);
    T ahx = 0, ahy = 0, ahz = 0;  // uniaxial, kernels.jl:54-105 (2/Ms == 2 * (1/Ms) exactly)
    if ((fl & FL_AN) && Ms != T(0)) {
        const T Ms2 = T(2.0) * Ms_inv;
        const T sa = mx * ax + my * ay + mz * az;
        ahx = Ku * sa * ax * Ms2;
        ahy = Ku * sa * ay * Ms2;
        ahz = Ku * sa * az * Ms2;
    }
    T chx = 0, chy = 0, chz = 0;  // cubic, kernels.jl:110-137
    if ((fl & FL_CU) && Ms != T(0)) {
        const MMX_CONST T *ca = P.t[C.sc].s;
        const T Ms4 = T(4.0) * Kc / Ms;
        const T mxp = ca[0] * mx + ca[1] * my + ca[2] * mz;
        const T myp = ca[3] * mx + ca[4] * my + ca[5] * mz;
        const T mzp = ca[6] * mx + ca[7] * my + ca[8] * mz;
        const T mxp3 = mxp * mxp * mxp, myp3 = myp * myp * myp, mzp3 = mzp * mzp * mzp;
        chx = Ms4 * (mxp3 * ca[0] + myp3 * ca[3] + mzp3 * ca[6]);
        chy = Ms4 * (mxp3 * ca[1] + myp3 * ca[4] + mzp3 * ca[7]);
        chz = Ms4 * (mxp3 * ca[2] + myp3 * ca[5] + mzp3 * ca[8]);
    }
    T thx = 0, thy = 0, thz = 0;  // thermal, kernels.jl:469-494
    if ((fl & FL_TH) && Ms > T(0)) {
        const T scale = (T)sqrt((T)dyn.thermal_factor * Tl / Ms);
        const uint64_t base = 3 * (uint64_t)I;
        thx = (T)philox_normal(P.seed, dyn.noise_step, base) * scale;
        thy = (T)philox_normal(P.seed, dyn.noise_step, base + 1) * scale;
        thz = (T)philox_normal(P.seed, dyn.noise_step, base + 2) * scale;
    }
    Hx = Hy = Hz = T(0);
    const int nterms = (fl >> FL_NT_SHIFT) & 15;
#pragma unroll
    for (int q = 0; q < MAX_TERMS; q++) {
        if (q >= nterms) break;
        const int sk = (ordp >> (3 * q)) & 7;
        T hx, hy, hz;
        if (sk == SL_EXCH) { hx = ex_on ? fex * Ms_inv : T(0); hy = ex_on ? fey * Ms_inv : T(0); hz = ex_on ? fez * Ms_inv : T(0); }
        else if (sk == SL_DMIB) { hx = bd_on ? fbx * Ms_inv : T(0); hy = bd_on ? fby * Ms_inv : T(0); hz = bd_on ? fbz * Ms_inv : T(0); }
        else if (sk == SL_DMII) { hx = id_on ? fix * Ms_inv : T(0); hy = id_on ? fiy * Ms_inv : T(0); hz = id_on ? fiz * Ms_inv : T(0); }
        else if (sk == SL_ANIS) { hx = ahx; hy = ahy; hz = ahz; }
        else if (sk == SL_CUBIC) { hx = chx; hy = chy; hz = chz; }
        else if (sk == SL_ZEE) { hx = zhx; hy = zhy; hz = zhz; }
        else if (sk == SL_DEMAG) { hx = hdx; hy = hdy; hz = hdz; }
        else { hx = thx; hy = thy; hz = thz; }
        Hx = Hx + hx;
        Hy = Hy + hy;
        Hz = Hz + hz;
    }
}

// Epilogue of the specialised kernels: MODE is EPI_STAGE, EPI_RK4_FINAL or EPI_ERR.
// ynext (nullable): receives the stored new state of the cell (STAGE / RK4_FINAL).
template <class T, int MODE, bool NOKP, uint32_t SIG, bool TILE = false>
__device__ __forceinline__ void finish_fast(const MMX_CONST FieldParams<T> &P, const Dyn &dyn, const Epi<T> &E, int I,
                                            int i, int j, int k, const T *__restrict__ m, T hdx, T hdy, T hdz,
                                            double &red, T *ynext = nullptr, const RowTile<T> &tl = RowTile<T>{})
{
#pragma clang fp contract(off)
    const int N = (int)P.N;
    const T alpha_ = P.alpha, llg_a_ = P.llg_a;
    const bool prec_ = (P.flags & FL_PREC) != 0;
    // operands of the stage combination, loaded with the field's loads; the
    // first stage of a step evaluates at its base state (y0 == m), whose
    // values the field reads anyway
    T y0v[3] = {0, 0, 0}, kpv[3][3];
    const bool y0m = E.y0 == m;
    if constexpr (MODE != EPI_ERR) {
        if (!y0m)
#pragma unroll
            for (int c = 0; c < 3; c++) y0v[c] = gld(E.y0, c * N + I);
    }
    if constexpr (MODE == EPI_RK4_FINAL) {
#pragma unroll
        for (int q = 0; q < 3; q++)
#pragma unroll
            for (int c = 0; c < 3; c++) kpv[q][c] = gld(E.kp[q], c * N + I);
    }
    const T alpha_c = P.alphas ? gld(P.alphas, I) : alpha_;
    T mx, my, mz, Hx, Hy, Hz;
    uint32_t msk;
    cell_field_fast<T, SIG, TILE>(P, dyn, I, i, j, k, m, hdx, hdy, hdz, mx, my, mz, Hx, Hy, Hz, msk, tl);
    if constexpr (MODE != EPI_ERR) {
        if (y0m) {
            y0v[0] = mx;
            y0v[1] = my;
            y0v[2] = mz;
        }
    }
    // the earlier stages' k of a DOPRI54 / BS23 / Heun combination or error
    // estimate: all issued here together (uniform guards, no use inside),
    // after the field so that they do not add to its register peak
    constexpr bool KPL = (MODE == EPI_STAGE && !NOKP) || MODE == EPI_ERR;
    T kq[KPL ? 6 : 1][3];
    if constexpr (KPL) {
#pragma unroll
        for (int q = 0; q < 6; q++)
            if (q < E.np)
#pragma unroll
                for (int c = 0; c < 3; c++) kq[q][c] = gld(E.kp[q], c * N + I);
    }
    // llg_rhs_kernel!: src/llg/llg.jl:6-40 (a = -gamma/(1+alpha^2) precomputed)
    T kk[3];
    if ((msk >> 6) & 1u) {
        kk[0] = kk[1] = kk[2] = T(0);
    } else {
        T alpha = alpha_, a = llg_a_;
        if (P.alphas) {  // spatial_llg_rhs_kernel! (llg.jl:62-97): alpha per cell
            alpha = alpha_c;
            a = -P.gamma / (T(1) + alpha * alpha);
        }
        const T mm = mx * mx + my * my + mz * mz;
        const T mh = mx * Hx + my * Hy + mz * Hz;
        const T h1 = mm * Hx - mh * mx, h2 = mm * Hy - mh * my, h3 = mm * Hz - mh * mz;
        T f1 = 0, f2 = 0, f3 = 0;
        if (prec_) {
            f1 = -mz * h2 + my * h3;
            f2 = mz * h1 - mx * h3;
            f3 = -my * h1 + mx * h2;
        }
        kk[0] = a * (f1 - h1 * alpha);
        kk[1] = a * (f2 - h2 * alpha);
        kk[2] = a * (f3 - h3 * alpha);
    }
    if (E.kstore) {
        gst(E.kstore, I, kk[0]);
        gst(E.kstore, N + I, kk[1]);
        gst(E.kstore, 2 * N + I, kk[2]);
    }
    if constexpr (MODE == EPI_ERR) {
        const T mv[3] = {mx, my, mz};
        double mval = 0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double err = 0;
#pragma unroll
            for (int q = 0; q < 6; q++)
                if (q < E.np) {
                    const double kd = (double)kq[q][c];
                    err = (q == 0) ? E.cp[0] * kd : err + E.cp[q] * kd;
                }
            err = err + E.cself * (double)kk[c];
            mval = pmax(mval, fabs(err / (1.0 + fabs((double)mv[c]))));
        }
        red = pmax(red, mval);
    } else {
        T y[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            if constexpr (MODE == EPI_RK4_FINAL) {
                const T s = kpv[0][c] + T(2) * kpv[1][c] + T(2) * kpv[2][c] + kk[c];
                y[c] = y0v[c] + (T)E.cself * s;
            } else {
                T s = y0v[c];
                if constexpr (!NOKP) {
#pragma unroll
                    for (int q = 0; q < 6; q++)
                        if (q < E.np) s = s + (T)E.cp[q] * kq[q][c];
                }
                y[c] = s + (T)E.cself * kk[c];
            }
        }
        if (E.normalize) {
            const T m2 = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
            if (m2 > T(0)) {
                const T len = T(1) / sqrt(m2);
                y[0] *= len;
                y[1] *= len;
                y[2] *= len;
            }
        }
        gst(E.yout, I, y[0]);
        gst(E.yout, N + I, y[1]);
        gst(E.yout, 2 * N + I, y[2]);
        if (ynext) {
            ynext[0] = y[0];
            ynext[1] = y[1];
            ynext[2] = y[2];
        }
    }
}

// llg_rhs_kernel!: src/llg/llg.jl:6-40
template <class T>
__device__ __forceinline__ void llg_rhs(const MMX_CONST FieldParams<T> &P, int64_t I, T mx, T my, T mz, T hx,
                                        T hy, T hz, T &dx, T &dy, T &dz)
{
#pragma clang fp contract(off)
    if (P.pins && P.pins[I]) {
        dx = dy = dz = T(0);
        return;
    }
    // spatial_llg_rhs_kernel! (llg.jl:62-97) reads alpha per cell
    const T alpha = P.alphas ? P.alphas[I] : P.alpha;
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

// Pull-backs through the LLG right-hand side of llg_rhs,
//   F(m, h) = a (m x p - alpha p),  p = |m|^2 h - (m.h) m,  a = -gamma/(1+alpha^2):
//   (dF/dh)^T l = -a Pi q,              Pi v = |m|^2 v - (m.v) m
//   (dF/dm)^T l = a (p x l - Pm^T q),   Pm^T v = 2 m (h.v) - h (m.v) - (m.h) v
// with q = m x l + alpha l; the m x terms drop when precession is off.
template <class T>
__device__ __forceinline__ void llg_q(const MMX_CONST FieldParams<T> &P, T mx, T my, T mz, T lx, T ly, T lz, T &qx, T &qy,
                                      T &qz)
{
    const T alpha = P.alpha;
    qx = alpha * lx;
    qy = alpha * ly;
    qz = alpha * lz;
    if (P.precession) {
        qx += my * lz - mz * ly;
        qy += mz * lx - mx * lz;
        qz += mx * ly - my * lx;
    }
}

template <class T>
__device__ __forceinline__ void llg_vjp_h(const MMX_CONST FieldParams<T> &P, T mx, T my, T mz, T lx, T ly, T lz, T &ux,
                                          T &uy, T &uz)
{
    const T alpha = P.alpha;
    const T a = -P.gamma / (T(1) + alpha * alpha);
    T qx, qy, qz;
    llg_q(P, mx, my, mz, lx, ly, lz, qx, qy, qz);
    const T mm = mx * mx + my * my + mz * mz;
    const T mq = mx * qx + my * qy + mz * qz;
    ux = -a * (mm * qx - mq * mx);
    uy = -a * (mm * qy - mq * my);
    uz = -a * (mm * qz - mq * mz);
}

template <class T>
__device__ __forceinline__ void llg_vjp_m(const MMX_CONST FieldParams<T> &P, T mx, T my, T mz, T hx, T hy, T hz, T lx, T ly,
                                          T lz, T &rx, T &ry, T &rz)
{
    const T alpha = P.alpha;
    const T a = -P.gamma / (T(1) + alpha * alpha);
    T qx, qy, qz;
    llg_q(P, mx, my, mz, lx, ly, lz, qx, qy, qz);
    const T mm = mx * mx + my * my + mz * mz;
    const T mh = mx * hx + my * hy + mz * hz;
    const T hq = hx * qx + hy * qy + hz * qz;
    const T mq = mx * qx + my * qy + mz * qz;
    rx = -(T(2) * mx * hq - hx * mq - mh * qx);
    ry = -(T(2) * my * hq - hy * mq - mh * qy);
    rz = -(T(2) * mz * hq - hz * mq - mh * qz);
    if (P.precession) {
        const T px = mm * hx - mh * mx, py = mm * hy - mh * my, pz = mm * hz - mh * mz;
        rx += py * lz - pz * ly;
        ry += pz * lx - px * lz;
        rz += px * ly - py * lx;
    }
    rx *= a;
    ry *= a;
    rz *= a;
}

// Jacobian-vector product of llg_rhs at (m, h) along (v, w = dh):
//   p = |m|^2 h - (m.h) m,  dp = 2 (m.v) h - (v.h) m - (m.h) v + |m|^2 w - (m.w) m,
//   F = a (m x p - alpha p)  ->  dF = a (v x p + m x dp - alpha dp)
// (the m x terms only with precession).
template <class T>
__device__ __forceinline__ void llg_jvp(const MMX_CONST FieldParams<T> &P, T mx, T my, T mz, T hx, T hy, T hz, T vx,
                                        T vy, T vz, T wx, T wy, T wz, T &ox, T &oy, T &oz)
{
    const T alpha = P.alpha;
    const T a = -P.gamma / (T(1) + alpha * alpha);
    const T mm = mx * mx + my * my + mz * mz;
    const T mh = mx * hx + my * hy + mz * hz;
    const T mv = mx * vx + my * vy + mz * vz;
    const T vh = vx * hx + vy * hy + vz * hz;
    const T mw = mx * wx + my * wy + mz * wz;
    const T px = mm * hx - mh * mx, py = mm * hy - mh * my, pz = mm * hz - mh * mz;
    const T dx = T(2) * mv * hx - vh * mx - mh * vx + mm * wx - mw * mx;
    const T dy = T(2) * mv * hy - vh * my - mh * vy + mm * wy - mw * my;
    const T dz = T(2) * mv * hz - vh * mz - mh * vz + mm * wz - mw * mz;
    ox = -alpha * dx;
    oy = -alpha * dy;
    oz = -alpha * dz;
    if (P.precession) {
        ox += (vy * pz - vz * py) + (my * dz - mz * dy);
        oy += (vz * px - vx * pz) + (mz * dx - mx * dz);
        oz += (vx * py - vy * px) + (mx * dy - my * dx);
    }
    ox *= a;
    oy *= a;
    oz *= a;
}

// e_q x v for the interfacial DMI directions e = (0,1,0), (0,-1,0), (-1,0,0),
// (1,0,0) of the -x, +x, -y, +y neighbours (src/micro/kernels.jl:426-466)
template <class T>
__device__ __forceinline__ void idmi_cross(int q, T vx, T vy, T vz, T &wx, T &wy, T &wz)
{
    wx = wy = wz = T(0);
    if (q == 0) { wx = vz; wz = -vx; }
    if (q == 1) { wx = -vz; wz = vx; }
    if (q == 2) { wy = vz; wz = -vy; }
    if (q == 3) { wy = -vz; wz = vy; }
}

// EPI_ADJ for cell I (see AdjIO): ut is this pass's input vector.
template <class T>
__device__ __forceinline__ void adj_cell(const MMX_CONST FieldParams<T> &P, const Dyn &dyn, const Epi<T> &E, int64_t I, int i,
                                         int j, int k, const T *__restrict__ ut, T hdx, T hdy, T hdz)
{
    const int64_t N = P.N;
    const AdjIO<T> &A = E.adj;
    // constant terms have no Jacobian; the cubic one is linearised at Y below
    uint32_t mask = E.term_mask;
    for (int tt = 0; tt < P.nterms; tt++) {
        const int kd = P.t[tt].kind;
        if (kd == K_ZEEMAN || kd == K_ZEEMAN_U || kd == K_THERMAL || kd == K_CUBIC) mask &= ~(1u << tt);
    }
    T Hx, Hy, Hz;
    cell_field(P, dyn, mask, I, i, j, k, ut, hdx, hdy, hdz, Hx, Hy, Hz, (double *)nullptr);
    const T Ms = P.mu0Ms[I];
    const T Yx = A.Y[I], Yy = A.Y[N + I], Yz = A.Y[2 * N + I];
    const T ux = ut[I], uy = ut[N + I], uz = ut[2 * N + I];
    T mux = 0, muy = 0, muz = 0;
    if (Ms > T(0)) {
        for (int tt = 0; tt < P.nterms; tt++) {
            const MMX_CONST TermDev<T> &D = P.t[tt];
            if (D.kind != K_CUBIC || !((E.term_mask >> tt) & 1u)) continue;
            // J = (12 Kc / mu0Ms) sum_r (Y.a_r)^2 a_r a_r^T  (cubic_anisotropy_kernel!, kernels.jl:110-137)
            const T Kc = D.uniform ? D.s[9] : D.a[I];
            const T c = T(12) * Kc / Ms;
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const T ax = D.s[3 * r], ay = D.s[3 * r + 1], az = D.s[3 * r + 2];
                const T yp = Yx * ax + Yy * ay + Yz * az;
                const T up = ux * ax + uy * ay + uz * az;
                const T f = c * yp * yp * up;
                Hx += f * ax;
                Hy += f * ay;
                Hz += f * az;
            }
        }
        mux = Ms * Hx;
        muy = Ms * Hy;
        muz = Ms * Hz;
    }
    if (!(P.pins && P.pins[I])) {
        T rx, ry, rz;
        llg_vjp_m(P, Yx, Yy, Yz, A.hY[I], A.hY[N + I], A.hY[2 * N + I], A.a[I], A.a[N + I], A.a[2 * N + I], rx, ry,
                  rz);
        mux += rx;
        muy += ry;
        muz += rz;
    }
    A.lam[I] += mux;
    A.lam[N + I] += muy;
    A.lam[2 * N + I] += muz;
    if (A.next) {
        const T c = (T)A.cnext;
        A.next[I] += c * mux;
        A.next[N + I] += c * muy;
        A.next[2 * N + I] += c * muz;
    }
    // parameter gradients: g_p += sum_i u_i . dH_i/dp, u = mu0Ms ut
    if (!(Ms > T(0))) return;
    int64_t nb[6];
    bool nbd = false;
    if (A.gms) {
        // mu0Ms enters the stencil and local terms as 1/mu0Ms (dH/dmu = -H/mu at
        // this cell) and the demag term through M = mu0Ms m of every cell: the
        // demag value of this pass is D u (the operator is symmetric), so cell
        // I gains Y_I . (D u)_I
        uint32_t lm = 0;
        for (int tt = 0; tt < P.nterms; tt++) {
            const int kd = P.t[tt].kind;
            if (kd == K_EXCH || kd == K_EXCH_S || kd == K_DMI_BULK || kd == K_DMI_BULK_S || kd == K_DMI_INT ||
                kd == K_ANIS || kd == K_ANIS_S || kd == K_CUBIC)
                lm |= ((E.term_mask >> tt) & 1u) << tt;
        }
        T lx = 0, ly = 0, lz = 0;
        if (lm) cell_field(P, dyn, lm, I, i, j, k, A.Y, T(0), T(0), T(0), lx, ly, lz, (double *)nullptr);
        const T gl = -(ux * lx + uy * ly + uz * lz);  // u . (-h/mu) with u = mu ut
        const T gd = Yx * hdx + Yy * hdy + Yz * hdz;
        A.gms[I] += gl + gd;
    }
    for (int tt = 0; tt < P.nterms; tt++) {
        T *g = A.g[tt];
        if (!g) continue;
        const MMX_CONST TermDev<T> &D = P.t[tt];
        if (D.kind == K_DMI_BULK || D.kind == K_DMI_BULK_S) {
            // bulk DMI (kernels.jl:306-422): f = sum_q De_q s_q (e_q x Y_q) over the
            // neighbours along axis e_q, s_q = +-1/d (-: the + neighbour).  Uniform D:
            // the per-cell share ut . f_c / D_c of dL/dD_c (summed over cells it is
            // the gradient of the uniform component); vector D(x): the harmonic mean
            // De = 2 D_p D_n / (D_p + D_n) of both cells of each pair, differentiated
            // for D_p as this cell's and as its neighbour's value.
            if (!nbd) {
                neighbours(P, i, j, k, nb);
                nbd = true;
            }
            T gc[3] = {0, 0, 0};
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const int64_t id = nb[q];
                if (id < 0 || !(P.mu0Ms[id] > T(0))) continue;
                const int ax = q >> 1;
                const T sg = (q & 1) ? T(-1) : T(1);
                const T jx = A.Y[id], jy = A.Y[N + id], jz = A.Y[2 * N + id];
                // e x Y_j and e x Y_p for axis ax
                T wx, wy, wz, zx, zy, zz;
                if (ax == 0) { wx = 0; wy = -jz; wz = jy; zx = 0; zy = -Yz; zz = Yy; }
                else if (ax == 1) { wx = jz; wy = 0; wz = -jx; zx = Yz; zy = 0; zz = -Yx; }
                else { wx = -jy; wy = jx; wz = 0; zx = -Yy; zy = Yx; zz = 0; }
                if (D.kind == K_DMI_BULK) {
                    const T Dc = D.s[6 + ax];
                    if (Dc != T(0)) gc[ax] += (D.s[ax] * sg) * (ux * wx + uy * wy + uz * wz) / Dc;
                } else {
                    const T *Dp = ax == 0 ? D.a : (ax == 1 ? D.b : D.c);
                    const T DI = Dp[I], Dn = Dp[id];
                    if (DI == T(0) || Dn == T(0)) continue;
                    const T ssum = DI + Dn;
                    const T dDe = T(2) * Dn * Dn / (ssum * ssum);
                    const T vx = ut[id], vy = ut[N + id], vz = ut[2 * N + id];
                    const T t = (ux * wx + uy * wy + uz * wz) - (vx * zx + vy * zy + vz * zz);
                    gc[ax] += dDe * (D.s[ax] * sg) * t;
                }
            }
            g[I] += gc[0];
            g[N + I] += gc[1];
            g[2 * N + I] += gc[2];
            continue;
        }
        T gv = 0;
        switch (D.kind) {
        case K_ANIS:
        case K_ANIS_S: {  // dH/dKu = (2/mu0Ms)(m.u)u
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
            gv = T(2) * (ux * ax + uy * ay + uz * az) * (Yx * ax + Yy * ay + Yz * az);
        } break;
        case K_CUBIC: {  // dH/dKc = (4/mu0Ms) sum_r (m.a_r)^3 a_r
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const T ax = D.s[3 * r], ay = D.s[3 * r + 1], az = D.s[3 * r + 2];
                const T yp = Yx * ax + Yy * ay + Yz * az;
                gv += T(4) * yp * yp * yp * (ux * ax + uy * ay + uz * az);
            }
        } break;
        case K_DMI_INT: {  // D_eff = 2 D_p D_j / (D_p + D_j) enters row p and row j
            const T DI = D.uniform ? D.s[3] : D.a[I];
            if (DI == T(0)) break;
            if (!nbd) {
                neighbours(P, i, j, k, nb);
                nbd = true;
            }
            const T Dd[4] = {D.s[0], D.s[0], D.s[1], D.s[1]};
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int64_t id = nb[q];
                if (id < 0 || !(P.mu0Ms[id] > T(0))) continue;
                const T Dn = D.uniform ? D.s[3] : D.a[id];
                if (Dn == T(0)) continue;
                const T s = DI + Dn;
                const T dDe = T(2) * Dn * Dn / (s * s);
                const T jx = A.Y[id], jy = A.Y[N + id], jz = A.Y[2 * N + id];
                const T vx = ut[id], vy = ut[N + id], vz = ut[2 * N + id];
                T wx, wy, wz, zx, zy, zz;
                idmi_cross(q, jx, jy, jz, wx, wy, wz);  // e_q x Y_j
                idmi_cross(q, Yx, Yy, Yz, zx, zy, zz);  // e_q x Y_p
                const T t = (ux * wx + uy * wy + uz * wz) - (vx * zx + vy * zy + vz * zz);
                gv += dDe * Dd[q] * t;
            }
        } break;
        case K_EXCH_S: {  // A_eff = 2 A_p A_j / (A_p + A_j)
            const T AI = D.a[I];
            if (AI == T(0)) break;
            if (!nbd) {
                neighbours(P, i, j, k, nb);
                nbd = true;
            }
            const T nab[6] = {D.s[0], D.s[0], D.s[1], D.s[1], D.s[2], D.s[2]};
#pragma unroll
            for (int q = 0; q < 6; q++) {
                const int64_t id = nb[q];
                if (id < 0 || !(P.mu0Ms[id] > T(0))) continue;
                const T An = D.a[id];
                if (An == T(0)) continue;
                const T s = AI + An;
                const T dAe = T(2) * An * An / (s * s);
                const T dx = A.Y[id] - Yx, dy = A.Y[N + id] - Yy, dz = A.Y[2 * N + id] - Yz;
                const T t = (ux - ut[id]) * dx + (uy - ut[N + id]) * dy + (uz - ut[2 * N + id]) * dz;
                gv += dAe * nab[q] * t;
            }
        } break;
        default: break;
        }
        g[I] += gv;
    }
}

// The fused last step of one RHS evaluation for cell I.  `red` accumulates the
// thread's reduction value (ERR: running max; ENERGY: sum of the masked terms).
// MM: the modes this instantiation serves (MM_* of mmx_common.h); a launch
// picks the instantiation whose set holds E.mode.
template <class T, uint32_t MM = MM_GENERIC>
__device__ __forceinline__ void finish_cell(const MMX_CONST FieldParams<T> &P, const Dyn &dyn, const Epi<T> &E, int64_t I,
                                            int i, int j, int k, const T *__restrict__ m, T hdx, T hdy, T hdz,
                                            double &red)
{
#pragma clang fp contract(off)
    const int64_t N = P.N;
    T Hx, Hy, Hz;
    if constexpr ((MM & MM_ADJ) != 0) {
        if (E.mode == EPI_ADJ) {
            adj_cell<T>(P, dyn, E, I, i, j, k, m, hdx, hdy, hdz);
            return;
        }
    }
    if constexpr (MM == MM_ADJ) return;
    if ((MM & (1u << EPI_ENERGY)) && E.mode == EPI_ENERGY) {
        double es = 0;
        cell_field(P, dyn, E.term_mask, I, i, j, k, m, hdx, hdy, hdz, Hx, Hy, Hz, &es);
        red += es;
        return;
    }
    if ((MM & MM_FIELD) && E.mode == EPI_FIELD) {
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
    // nothing below for an instantiation without the RHS, stage and error modes
    if constexpr ((MM & ~(MM_FIELD | (1u << EPI_ENERGY) | MM_ADJ)) == 0) return;
    cell_field(P, dyn, E.term_mask, I, i, j, k, m, hdx, hdy, hdz, Hx, Hy, Hz, (double *)nullptr);
    // operands of the stage combination / error estimate, issued together
    // after the field (uniform guards, no use inside)
    const int nkp = E.mode == EPI_RK4_FINAL ? 3 : (E.mode == EPI_RHS ? 0 : E.np);
    const bool uy0 = E.mode != EPI_RHS && E.mode != EPI_ERR;
    T kq[6][3], y0v[3];
#pragma unroll
    for (int q = 0; q < 6; q++)
        if (q < nkp)
#pragma unroll
            for (int c = 0; c < 3; c++) kq[q][c] = gld(E.kp[q], c * N + I);
    if (uy0)
#pragma unroll
        for (int c = 0; c < 3; c++) y0v[c] = gld(E.y0, c * N + I);
    const T mx = m[I], my = m[N + I], mz = m[2 * N + I];
    T kx, ky, kz;
    llg_rhs(P, I, mx, my, mz, Hx, Hy, Hz, kx, ky, kz);
    if ((MM & (1u << EPI_RHS)) && E.mode == EPI_RHS) {
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
    if (E.out) {  // the field at this stage (adjoint recomputation)
        E.out[I] = Hx;
        E.out[N + I] = Hy;
        E.out[2 * N + I] = Hz;
    }
    const T kk[3] = {kx, ky, kz};
    if ((MM & (1u << EPI_ERR)) && E.mode == EPI_ERR) {  // vector_add6b / BS23 error combination, then |err|/(1+|y|)
        double mval = 0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int64_t o = c * N + I;
            double err = 0;
#pragma unroll
            for (int q = 0; q < 6; q++)
                if (q < E.np) err = (q == 0) ? E.cp[0] * (double)kq[0][c] : err + E.cp[q] * (double)kq[q][c];
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
        if (E.mode == EPI_RK4_FINAL) {  // src/integrator/rk.jl:46
            T s = kq[0][c] + T(2) * kq[1][c] + T(2) * kq[2][c] + kk[c];
            y[c] = y0v[c] + (T)E.cself * s;
        } else if (E.mode == EPI_STAGE_SCALED) {  // adaptive_rk_steps.jl:35
            T s = T(0);
#pragma unroll
            for (int q = 0; q < 6; q++)
                if (q < E.np) s = (q == 0) ? (T)E.cp[0] * kq[0][c] : s + (T)E.cp[q] * kq[q][c];
            s = s + (T)E.cself * kk[c];
            y[c] = y0v[c] + (T)E.outer * s;
        } else {  // vector_add2..6: a1 + c2 a2 + ... (util.jl:86-166)
            T s = y0v[c];
#pragma unroll
            for (int q = 0; q < 6; q++)
                if (q < E.np) s = s + (T)E.cp[q] * kq[q][c];
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
// partial[blockIdx*stride + v]: wave reduction with shuffles, then one value
// per wave through a small shared array (any block size up to 1024).
__device__ __forceinline__ double wave_reduce(double v, bool is_max)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        v = is_max ? pmax(v, o) : v + o;
    }
    return v;
}

template <int NV>
__device__ __forceinline__ void block_reduce_store(double *red, int nval, bool is_max, double *partial,
                                                   int stride)
{
    __shared__ double sw[16];
    const int nt = blockDim.x * blockDim.y;
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int nw = (nt + 63) / 64;
    for (int v = 0; v < nval; v++) {
        double x = wave_reduce(red[v], is_max);
        if ((tid & 63) == 0) sw[tid >> 6] = x;
        __syncthreads();
        if (tid == 0) {
            double acc = sw[0];
            for (int w = 1; w < nw; w++) acc = is_max ? pmax(acc, sw[w]) : acc + sw[w];
            partial[(size_t)blockIdx.x * stride + v] = acc;
        }
        __syncthreads();
    }
}

}  // namespace mmx
