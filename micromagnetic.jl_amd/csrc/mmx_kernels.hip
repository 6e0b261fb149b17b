// mmx_kernels.hip -- HIP kernels for gfx950: the local-field/LLG/stage kernel
// and the five passes of the zero-padded FFT demag convolution.
//
// Demag pipeline for one RHS evaluation (the reference does fill + 3 rfft +
// tensor multiply + 3 irfft + collect, src/micro/demag.jl:102-130):
//   P1  x: real -> half-spectrum, only the nx nonzero inputs are read    (A)
//   P2  y: forward, only ny nonzero inputs per column                     (A -> B)
//   P3  z: forward, 3x3 symmetric tensor multiply, inverse, keep z < nz  (B in place)
//   P4  y: inverse, keep y < ny                                           (B -> A)
//   P5  x: half-spectrum -> real, keep x < nx, then the fused epilogue:
//       local terms + demag field + LLG + integrator stage update.
// Each pass stages a batch of 1-D transforms in LDS and runs a radix-4/2
// Stockham FFT over them with FP64-accurate twiddles from a table.
#include "mmx_device.h"
#include "mmx_kernels.h"

namespace mmx {

template <class T>
struct Cap;
template <>
struct Cap<double> {
    static constexpr int E = 4096;  // complex elements per workgroup batch (64 KB)
    static constexpr int Q = 4;     // radix-4 butterflies per thread per stage
};
template <>
struct Cap<float> {
    static constexpr int E = 8192;
    static constexpr int Q = 8;
};

// In-place batched Stockham FFT over LDS: nseq sequences of length 2^lgL,
// sequence s at buf[s*LS ...].  tw[t] = exp(-2 pi i t / 2^lgLmax).
// Stage with radix R and current span Ns: butterfly j reads x[j + r L/R],
// multiplies by W_{Ns R}^{r (j mod Ns)}, and writes y[(j/Ns) Ns R + j mod Ns + r Ns].
template <class T>
__device__ void lds_fft(cplx<T> *buf, int nseq, int lgL, int LS, bool inv, const cplx<T> *__restrict__ tw,
                        int lgLmax)
{
    constexpr int Q = Cap<T>::Q;
    const int tid = threadIdx.x;
    int lgNs = 0;
    int rem = lgL;
    while (rem > 0) {
        const int lgR = (rem >= 2) ? 2 : 1;
        const int lgM = lgL - lgR;
        const int M = 1 << lgM;
        const int NB = nseq << lgM;  // butterflies in this stage
        const int shift = lgLmax - lgNs - lgR;
        const int Nsm = (1 << lgNs) - 1;
        if (lgR == 2) {
            cplx<T> v[Q][4];
#pragma unroll
            for (int q = 0; q < Q; q++) {
                int b = tid + q * BLOCK;
                if (b < NB) {
                    int s = b >> lgM, jj = b & (M - 1);
                    const cplx<T> *p = buf + s * LS + jj;
#pragma unroll
                    for (int r = 0; r < 4; r++) v[q][r] = p[r << lgM];
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < Q; q++) {
                int b = tid + q * BLOCK;
                if (b < NB) {
                    int s = b >> lgM, jj = b & (M - 1);
                    int kk = jj & Nsm;
                    if (kk) {
#pragma unroll
                        for (int r = 1; r < 4; r++) {
                            cplx<T> w = tw[(r * kk) << shift];
                            if (inv) w.y = -w.y;
                            v[q][r] = cmul(v[q][r], w);
                        }
                    }
                    cplx<T> a0 = cadd(v[q][0], v[q][2]);
                    cplx<T> a1 = csub(v[q][0], v[q][2]);
                    cplx<T> a2 = cadd(v[q][1], v[q][3]);
                    cplx<T> d = csub(v[q][1], v[q][3]);
                    cplx<T> a3 = inv ? cplx<T>{-d.y, d.x} : cplx<T>{d.y, -d.x};
                    int base = s * LS + ((jj >> lgNs) << (lgNs + 2)) + kk;
                    buf[base] = cadd(a0, a2);
                    buf[base + (1 << lgNs)] = cadd(a1, a3);
                    buf[base + (2 << lgNs)] = csub(a0, a2);
                    buf[base + (3 << lgNs)] = csub(a1, a3);
                }
            }
        } else {
            cplx<T> v[2 * Q][2];
#pragma unroll
            for (int q = 0; q < 2 * Q; q++) {
                int b = tid + q * BLOCK;
                if (b < NB) {
                    int s = b >> lgM, jj = b & (M - 1);
                    const cplx<T> *p = buf + s * LS + jj;
                    v[q][0] = p[0];
                    v[q][1] = p[M];
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 2 * Q; q++) {
                int b = tid + q * BLOCK;
                if (b < NB) {
                    int s = b >> lgM, jj = b & (M - 1);
                    int kk = jj & Nsm;
                    if (kk) {
                        cplx<T> w = tw[kk << shift];
                        if (inv) w.y = -w.y;
                        v[q][1] = cmul(v[q][1], w);
                    }
                    int base = s * LS + ((jj >> lgNs) << (lgNs + 1)) + kk;
                    buf[base] = cadd(v[q][0], v[q][1]);
                    buf[base + (1 << lgNs)] = csub(v[q][0], v[q][1]);
                }
            }
        }
        __syncthreads();
        lgNs += lgR;
        rem -= lgR;
    }
}

// ---------------------------------------------------------------------------
// P1: rows of n_in reals (zero-padded to Px = 2H) -> Kx = H+1 complex values.
// Row r = z*Ry + y of component c starts at in[c*cstride + r*n_in]; inputs are
// multiplied by scale[r*n_in + x] when scale != null (mu0 Ms, demag.jl:419-427).
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(BLOCK) k_p1(DemagGeom G, const T *__restrict__ in, int64_t cstride,
                                              const T *__restrict__ scale, int n_in, int Ry, int Rz,
                                              int ncomp, int rpb, cplx<T> *__restrict__ A,
                                              const cplx<T> *__restrict__ tw)
{
    __shared__ cplx<T> lds[Cap<T>::E + 64];
    const int H = G.H, HS = H + 1;
    const int nrows_tot = Ry * Rz;
    const int row0 = blockIdx.x * rpb;
    const int nrows = min(rpb, nrows_tot - row0);
    const int nseq = ncomp * nrows;
    for (int idx = threadIdx.x; idx < nseq * H; idx += BLOCK) {
        int s = idx / H, j = idx - s * H;
        int c = s / nrows, r = s - c * nrows;
        int64_t rb = (int64_t)(row0 + r) * n_in;
        int x0 = 2 * j, x1 = 2 * j + 1;
        T v0 = 0, v1 = 0;
        if (x0 < n_in) v0 = scale ? in[c * cstride + rb + x0] * scale[rb + x0] : in[c * cstride + rb + x0];
        if (x1 < n_in) v1 = scale ? in[c * cstride + rb + x1] * scale[rb + x1] : in[c * cstride + rb + x1];
        lds[s * HS + j] = {v0, v1};
    }
    __syncthreads();
    lds_fft<T>(lds, nseq, G.lgH, HS, false, tw, G.lgLmax);
    const int sh = G.lgLmax - (G.lgH + 1);  // W_Px^k = tw[k << sh]
    for (int idx = threadIdx.x; idx < nseq * (H + 1); idx += BLOCK) {
        int s = idx / (H + 1), k = idx - s * (H + 1);
        int c = s / nrows, r = s - c * nrows;
        cplx<T> Zk = lds[s * HS + (k & (H - 1))];
        cplx<T> Zc = lds[s * HS + ((H - k) & (H - 1))];
        Zc.y = -Zc.y;
        cplx<T> Fe = {(Zk.x + Zc.x) * T(0.5), (Zk.y + Zc.y) * T(0.5)};
        cplx<T> d = csub(Zk, Zc);
        cplx<T> Fo = {d.y * T(0.5), -d.x * T(0.5)};
        cplx<T> X = cadd(Fe, cmul(tw[k << sh], Fo));
        A[((int64_t)c * nrows_tot + row0 + r) * G.KS + k] = X;
    }
}

// ---------------------------------------------------------------------------
// Column transforms along y over tiles of TY consecutive kx columns.
// Input column (plane p, comp c): in[((c*nplanes + p)*Rin + y)*KS + kx], y < n_in.
// Output: out[((c*nplanes + p)*Rout + y)*KS + kx] for y < n_out.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(BLOCK) k_coly(DemagGeom G, const cplx<T> *__restrict__ in,
                                                cplx<T> *__restrict__ out, int n_in, int Rin, int n_out,
                                                int Rout, int nplanes, int lgL, int TY, int inverse,
                                                const cplx<T> *__restrict__ tw)
{
    __shared__ cplx<T> lds[Cap<T>::E + 64];
    const int L = 1 << lgL, LS = L + 1;
    const int kx0 = blockIdx.x * TY;
    const int p = blockIdx.y, c = blockIdx.z;
    const int KS = G.KS, Kx = G.Kx;
    const cplx<T> *src = in + ((int64_t)c * nplanes + p) * Rin * KS;
    for (int idx = threadIdx.x; idx < TY * L; idx += BLOCK) {
        int t = idx % TY, y = idx / TY;
        int kx = kx0 + t;
        cplx<T> v = {0, 0};
        if (y < n_in && kx < Kx) v = src[(int64_t)y * KS + kx];
        lds[t * LS + y] = v;
    }
    __syncthreads();
    lds_fft<T>(lds, TY, lgL, LS, inverse != 0, tw, G.lgLmax);
    cplx<T> *dst = out + ((int64_t)c * nplanes + p) * Rout * KS;
    for (int idx = threadIdx.x; idx < TY * n_out; idx += BLOCK) {
        int t = idx % TY, y = idx / TY;
        int kx = kx0 + t;
        if (kx < Kx) dst[(int64_t)y * KS + kx] = lds[t * LS + y];
    }
}

// ---------------------------------------------------------------------------
// P3: z transform, per-k 3x3 symmetric tensor multiply, inverse (conv mode),
// or forward only (tensor set-up).  B element (c, z, ky, kx) at
// ((c*nzb + z)*Py + ky)*KS + kx.  Kt holds the real k-space tensor folded to
// ky <= Py/2, kz <= Pz/2: (comp, kzf, kyf, kx) at ((comp*Pz2 + kzf)*Py2 + kyf)*KS + kx,
// with the sign of the odd components restored on unfolding (xy odd in y,
// xz odd in z, yz odd in y and z; src/micro/demag.jl:69-80).  The reference
// keeps real(rfft(tensor)), whose odd components also carry the cosine
// transform of the tensor's on-axis line (xy: x=y=0, a function of kz; xz:
// x=z=0, of ky; yz: y=z=0, of kx) -- rounding residue of the Newell sums that
// is not odd; Lline holds these three even transforms: [Pz2 | Py2 | Kx].
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(BLOCK) k_convz(DemagGeom G, cplx<T> *__restrict__ B,
                                                 const T *__restrict__ Kt, const T *__restrict__ Lline,
                                                 int nzb, int n_in, int ncomp, int TZ, int fwd_only,
                                                 const cplx<T> *__restrict__ tw)
{
    __shared__ cplx<T> lds[Cap<T>::E + 64];
    const int Pz = G.Pz, LS = Pz + 1;
    const int KS = G.KS, Kx = G.Kx, Py = G.Py;
    const int kx0 = blockIdx.x * TZ;
    const int ky = blockIdx.y;
    const int nseq = ncomp * TZ;
    for (int idx = threadIdx.x; idx < nseq * Pz; idx += BLOCK) {
        int t = idx % TZ, rest = idx / TZ;
        int z = rest % Pz, c = rest / Pz;
        int kx = kx0 + t;
        cplx<T> v = {0, 0};
        if (z < n_in && kx < Kx) v = B[(((int64_t)c * nzb + z) * Py + ky) * KS + kx];
        lds[(c * TZ + t) * LS + z] = v;
    }
    __syncthreads();
    lds_fft<T>(lds, nseq, G.lgPz, LS, false, tw, G.lgLmax);
    if (fwd_only) {
        for (int idx = threadIdx.x; idx < nseq * Pz; idx += BLOCK) {
            int t = idx % TZ, rest = idx / TZ;
            int z = rest % Pz, c = rest / Pz;
            int kx = kx0 + t;
            if (kx < Kx) B[(((int64_t)c * nzb + z) * Py + ky) * KS + kx] = lds[(c * TZ + t) * LS + z];
        }
        return;
    }
    const int kyf = (ky <= Py / 2) ? ky : Py - ky;
    const T sy = (ky <= Py / 2) ? T(1) : T(-1);
    const int64_t cs = (int64_t)G.Pz2 * G.Py2 * KS;  // component stride of Kt
    for (int idx = threadIdx.x; idx < TZ * Pz; idx += BLOCK) {
        int t = idx % TZ, kz = idx / TZ;
        int kx = kx0 + t;
        if (kx >= Kx) continue;
        int kzf = (kz <= Pz / 2) ? kz : Pz - kz;
        T sz = (kz <= Pz / 2) ? T(1) : T(-1);
        const T *kp = Kt + ((int64_t)kzf * G.Py2 + kyf) * KS + kx;
        T xx = kp[0], yy = kp[cs], zz = kp[2 * cs];
        T xy = sy * kp[3 * cs] + Lline[kzf];
        T xz = sz * kp[4 * cs] + Lline[G.Pz2 + kyf];
        T yz = sy * sz * kp[5 * cs] + Lline[G.Pz2 + G.Py2 + kx];
        cplx<T> Mx = lds[(0 * TZ + t) * LS + kz];
        cplx<T> My = lds[(1 * TZ + t) * LS + kz];
        cplx<T> Mz = lds[(2 * TZ + t) * LS + kz];
        lds[(0 * TZ + t) * LS + kz] = {xx * Mx.x + xy * My.x + xz * Mz.x, xx * Mx.y + xy * My.y + xz * Mz.y};
        lds[(1 * TZ + t) * LS + kz] = {xy * Mx.x + yy * My.x + yz * Mz.x, xy * Mx.y + yy * My.y + yz * Mz.y};
        lds[(2 * TZ + t) * LS + kz] = {xz * Mx.x + yz * My.x + zz * Mz.x, xz * Mx.y + yz * My.y + zz * Mz.y};
    }
    __syncthreads();
    lds_fft<T>(lds, nseq, G.lgPz, LS, true, tw, G.lgLmax);
    const int nz = G.nz;
    for (int idx = threadIdx.x; idx < nseq * nz; idx += BLOCK) {
        int t = idx % TZ, rest = idx / TZ;
        int z = rest % nz, c = rest / nz;
        int kx = kx0 + t;
        if (kx < Kx) B[(((int64_t)c * nzb + z) * Py + ky) * KS + kx] = lds[(c * TZ + t) * LS + z];
    }
}

// ---------------------------------------------------------------------------
// P5: half-spectrum -> real along x for rpb rows, then the fused epilogue for
// every cell of those rows.  A element (c, row, kx) at (c*ny*nz + row)*KS + kx.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(BLOCK) k_p5(DemagGeom G, const cplx<T> *__restrict__ A, FieldParams<T> P,
                                              Epi<T> E, const T *__restrict__ m, int rpb,
                                              const cplx<T> *__restrict__ tw)
{
    __shared__ cplx<T> lds[Cap<T>::E + 64];
    const int H = G.H, HS = H + 1, nx = G.nx, ny = G.ny;
    const int nrows_tot = G.ny * G.nz;
    const int row0 = blockIdx.x * rpb;
    const int nrows = min(rpb, nrows_tot - row0);
    const int nseq = 3 * nrows;
    const int sh = G.lgLmax - (G.lgH + 1);
    for (int idx = threadIdx.x; idx < nseq * H; idx += BLOCK) {
        int s = idx / H, k = idx - s * H;
        int c = s / nrows, r = s - c * nrows;
        const cplx<T> *row = A + ((int64_t)c * nrows_tot + row0 + r) * G.KS;
        cplx<T> Xk = row[k];
        cplx<T> Xc = row[H - k];
        Xc.y = -Xc.y;
        cplx<T> Fe = cadd(Xk, Xc);
        cplx<T> w = tw[k << sh];
        w.y = -w.y;  // W_Px^{-k}
        cplx<T> Fo = cmul(csub(Xk, Xc), w);
        lds[s * HS + k] = {Fe.x - Fo.y, Fe.y + Fo.x};  // Fe + i Fo
    }
    __syncthreads();
    lds_fft<T>(lds, nseq, G.lgH, HS, true, tw, G.lgLmax);
    double red[MAX_TERMS];
    for (int v = 0; v < MAX_TERMS; v++) red[v] = 0;
    for (int idx = threadIdx.x; idx < nrows * nx; idx += BLOCK) {
        int r = idx / nx, x = idx - r * nx;
        int row = row0 + r;
        int y = row % ny, z = row / ny;
        cplx<T> h0 = lds[(0 * nrows + r) * HS + (x >> 1)];
        cplx<T> h1 = lds[(1 * nrows + r) * HS + (x >> 1)];
        cplx<T> h2 = lds[(2 * nrows + r) * HS + (x >> 1)];
        T hx = (x & 1) ? h0.y : h0.x;
        T hy = (x & 1) ? h1.y : h1.x;
        T hz = (x & 1) ? h2.y : h2.x;
        int64_t I = (int64_t)row * nx + x;
        finish_cell(P, E, I, x, y, z, m, hx, hy, hz, red);
    }
    if (E.mode == EPI_ERR) block_reduce_store<1>(red, 1, true, E.partial, 1);
    else if (E.mode == EPI_ENERGY) block_reduce_store<MAX_TERMS>(red, P.nterms, false, E.partial, MAX_TERMS);
}

// ---------------------------------------------------------------------------
// Local-field kernel (meshes without demag): one thread per cell.
// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(BLOCK) k_local(FieldParams<T> P, Epi<T> E, const T *__restrict__ m)
{
    double red[MAX_TERMS];
    for (int v = 0; v < MAX_TERMS; v++) red[v] = 0;
    const int nx = P.nx, ny = P.ny;
    for (int64_t I = (int64_t)blockIdx.x * BLOCK + threadIdx.x; I < P.N; I += (int64_t)gridDim.x * BLOCK) {
        int i = (int)(I % nx);
        int64_t q = I / nx;
        int j = (int)(q % ny);
        int k = (int)(q / ny);
        finish_cell(P, E, I, i, j, k, m, T(0), T(0), T(0), red);
    }
    if (E.mode == EPI_ERR) block_reduce_store<1>(red, 1, true, E.partial, 1);
    else if (E.mode == EPI_ENERGY) block_reduce_store<MAX_TERMS>(red, P.nterms, false, E.partial, MAX_TERMS);
}

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_aos_to_soa(const double *__restrict__ aos, T *__restrict__ soa, int64_t N)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < N; i += (int64_t)gridDim.x * BLOCK) {
        soa[i] = (T)aos[3 * i];
        soa[N + i] = (T)aos[3 * i + 1];
        soa[2 * N + i] = (T)aos[3 * i + 2];
    }
}

template <class T>
__global__ void k_soa_to_aos(const T *__restrict__ soa, double *__restrict__ aos, int64_t N)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < N; i += (int64_t)gridDim.x * BLOCK) {
        aos[3 * i] = (double)soa[i];
        aos[3 * i + 1] = (double)soa[N + i];
        aos[3 * i + 2] = (double)soa[2 * N + i];
    }
}

template <class T>
__global__ void k_cast(const double *__restrict__ src, T *__restrict__ dst, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
        dst[i] = (T)src[i];
}

template <class T>
__global__ void k_to_double(const T *__restrict__ src, double *__restrict__ dst, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
        dst[i] = (double)src[i];
}

// padded real-space tensor component (mirror of fill_tensors_kernel!,
// src/micro/demag.jl:394-411) on the power-of-two padded grid
__global__ void k_tensor_pad(DemagGeom G, const double *__restrict__ rt, int comp, double *__restrict__ pad)
{
    const int64_t P = (int64_t)G.Px * G.Py * G.Pz;
    const int64_t N = (int64_t)G.nx * G.ny * G.nz;
    const bool ox = comp == 3 || comp == 4, oy = comp == 3 || comp == 5, oz = comp == 4 || comp == 5;
    for (int64_t o = (int64_t)blockIdx.x * BLOCK + threadIdx.x; o < P; o += (int64_t)gridDim.x * BLOCK) {
        int x = (int)(o % G.Px);
        int64_t q = o / G.Px;
        int y = (int)(q % G.Py);
        int z = (int)(q / G.Py);
        int dx, dy, dz;
        bool nx_ = false, ny_ = false, nz_ = false, gap = false;
        if (x < G.nx) dx = x;
        else if (x > G.Px - G.nx) { dx = G.Px - x; nx_ = true; }
        else { dx = 0; gap = true; }
        if (y < G.ny) dy = y;
        else if (y > G.Py - G.ny) { dy = G.Py - y; ny_ = true; }
        else { dy = 0; gap = true; }
        if (z < G.nz) dz = z;
        else if (z > G.Pz - G.nz) { dz = G.Pz - z; nz_ = true; }
        else { dz = 0; gap = true; }
        double v = 0;
        // odd components: the planes where an odd axis has displacement 0 are
        // left out here; their only contribution to the real part of the
        // transform (the line on both odd axes) is added back as Lline in P3
        const bool plane0 = (ox && dx == 0) || (oy && dy == 0) || (oz && dz == 0);
        if (!gap && !plane0) {
            v = rt[comp * N + ((int64_t)dz * G.ny + dy) * G.nx + dx];
            if (ox && nx_) v = -v;
            if (oy && ny_) v = -v;
            if (oz && nz_) v = -v;
        }
        pad[o] = v;
    }
}

template <class T>
__global__ void k_tensor_extract(DemagGeom G, const cplx<double> *__restrict__ Bt, int comp, double scale,
                                 T *__restrict__ Kt)
{
    const int64_t n = (int64_t)G.Pz2 * G.Py2 * G.Kx;
    for (int64_t o = (int64_t)blockIdx.x * BLOCK + threadIdx.x; o < n; o += (int64_t)gridDim.x * BLOCK) {
        int kx = (int)(o % G.Kx);
        int64_t q = o / G.Kx;
        int kyf = (int)(q % G.Py2);
        int kzf = (int)(q / G.Py2);
        double v = Bt[((int64_t)kzf * G.Py + kyf) * G.KS + kx].x * scale;
        Kt[(((int64_t)comp * G.Pz2 + kzf) * G.Py2 + kyf) * G.KS + kx] = (T)v;
    }
}

__global__ void k_reduce(const double *__restrict__ partial, int nblocks, int stride, int nval, int is_max,
                         double *__restrict__ out)
{
    __shared__ double s[BLOCK];
    for (int v = 0; v < nval; v++) {
        double acc = 0;
        for (int b = threadIdx.x; b < nblocks; b += BLOCK) {
            double x = partial[(size_t)b * stride + v];
            acc = is_max ? pmax(acc, x) : acc + x;
        }
        s[threadIdx.x] = acc;
        __syncthreads();
        for (int h = BLOCK / 2; h > 0; h >>= 1) {
            if ((int)threadIdx.x < h) s[threadIdx.x] = is_max ? pmax(s[threadIdx.x], s[threadIdx.x + h])
                                                              : s[threadIdx.x] + s[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[v] = s[0];
        __syncthreads();
    }
}

template <class T>
__global__ void k_maxdiff(const T *__restrict__ a, const T *__restrict__ b, int64_t N, double *__restrict__ partial)
{
    double red[1] = {0};
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < N; i += (int64_t)gridDim.x * BLOCK) {
        T dx = a[i] - b[i], dy = a[N + i] - b[N + i], dz = a[2 * N + i] - b[2 * N + i];
        red[0] = pmax(red[0], (double)sqrt(dx * dx + dy * dy + dz * dz));
    }
    block_reduce_store<1>(red, 1, true, partial, 1);
}

template <class T>
__global__ void k_axpy(const T *__restrict__ y, double c, const T *__restrict__ k, T *__restrict__ out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
        out[i] = y[i] + (T)c * k[i];
}

template <class T>
__global__ void k_maxabs(const T *__restrict__ v, int64_t n, double *__restrict__ partial)
{
    double r[1] = {0};
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
        r[0] = pmax(r[0], fabs((double)v[i]));
    block_reduce_store<1>(r, 1, true, partial, 1);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int grid_for(int64_t n, int cap = 8192)
{
    int64_t g = (n + BLOCK - 1) / BLOCK;
    if (g < 1) g = 1;
    return (int)(g > cap ? cap : g);
}

int max_local_blocks() { return 8192; }

template <class T>
void launch_aos_to_soa(const double *aos, T *soa, int64_t N, hipStream_t s)
{
    hipLaunchKernelGGL(k_aos_to_soa<T>, dim3(grid_for(N)), dim3(BLOCK), 0, s, aos, soa, N);
}
template <class T>
void launch_soa_to_aos(const T *soa, double *aos, int64_t N, hipStream_t s)
{
    hipLaunchKernelGGL(k_soa_to_aos<T>, dim3(grid_for(N)), dim3(BLOCK), 0, s, soa, aos, N);
}
template <class T>
void launch_cast(const double *src, T *dst, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_cast<T>, dim3(grid_for(n)), dim3(BLOCK), 0, s, src, dst, n);
}
template <class T>
void launch_to_double(const T *src, double *dst, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_double<T>, dim3(grid_for(n)), dim3(BLOCK), 0, s, src, dst, n);
}

template <class T>
int launch_local(const FieldParams<T> &P, const Epi<T> &E, const T *m, hipStream_t s)
{
    int g = grid_for(P.N, max_local_blocks());
    hipLaunchKernelGGL(k_local<T>, dim3(g), dim3(BLOCK), 0, s, P, E, m);
    return g;
}

template <class T>
void launch_p1(const DemagGeom &G, const T *in, int64_t cstride, const T *scale, int n_in, int Ry, int Rz,
               int ncomp, cplx<T> *A, const cplx<T> *tw, hipStream_t s)
{
    int rpb = Cap<T>::E / (ncomp * (G.H + 1));
    if (rpb < 1) rpb = 1;
    int rows = Ry * Rz;
    int g = (rows + rpb - 1) / rpb;
    hipLaunchKernelGGL(k_p1<T>, dim3(g), dim3(BLOCK), 0, s, G, in, cstride, scale, n_in, Ry, Rz, ncomp, rpb, A, tw);
}

template <class T>
void launch_coly(const DemagGeom &G, const cplx<T> *in, cplx<T> *out, int n_in, int Rin, int n_out, int Rout,
                 int nplanes, int ncomp, int inverse, const cplx<T> *tw, hipStream_t s)
{
    int L = G.Py;
    int TY = Cap<T>::E / (L + 1);
    if (TY > 16) TY = 16;
    if (TY < 1) TY = 1;
    dim3 grid((G.Kx + TY - 1) / TY, nplanes, ncomp);
    hipLaunchKernelGGL(k_coly<T>, grid, dim3(BLOCK), 0, s, G, in, out, n_in, Rin, n_out, Rout, nplanes, G.lgPy,
                       TY, inverse, tw);
}

template <class T>
void launch_convz(const DemagGeom &G, cplx<T> *B, const T *Kt, const T *Lline, int nzb, int n_in, int ncomp,
                  int fwd_only, const cplx<T> *tw, hipStream_t s)
{
    int TZ = Cap<T>::E / (ncomp * (G.Pz + 1));
    if (TZ > 16) TZ = 16;
    if (TZ < 1) TZ = 1;
    dim3 grid((G.Kx + TZ - 1) / TZ, G.Py);
    hipLaunchKernelGGL(k_convz<T>, grid, dim3(BLOCK), 0, s, G, B, Kt, Lline, nzb, n_in, ncomp, TZ, fwd_only, tw);
}

template <class T>
int launch_p5(const DemagGeom &G, const cplx<T> *A, const FieldParams<T> &P, const Epi<T> &E, const T *m,
              const cplx<T> *tw, hipStream_t s)
{
    int rpb = Cap<T>::E / (3 * (G.H + 1));
    if (rpb < 1) rpb = 1;
    int rows = G.ny * G.nz;
    int g = (rows + rpb - 1) / rpb;
    hipLaunchKernelGGL(k_p5<T>, dim3(g), dim3(BLOCK), 0, s, G, A, P, E, m, rpb, tw);
    return g;
}

void launch_tensor_pad(const DemagGeom &G, const double *rt, int comp, double *pad, hipStream_t s)
{
    int64_t P = (int64_t)G.Px * G.Py * G.Pz;
    hipLaunchKernelGGL(k_tensor_pad, dim3(grid_for(P)), dim3(BLOCK), 0, s, G, rt, comp, pad);
}

template <class T>
void launch_tensor_extract(const DemagGeom &G, const cplx<double> *Bt, int comp, double scale, T *Kt,
                           hipStream_t s)
{
    int64_t n = (int64_t)G.Pz2 * G.Py2 * G.Kx;
    hipLaunchKernelGGL(k_tensor_extract<T>, dim3(grid_for(n)), dim3(BLOCK), 0, s, G, Bt, comp, scale, Kt);
}

void launch_reduce(const double *partial, int nblocks, int stride, int nval, int is_max, double *out,
                   hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(BLOCK), 0, s, partial, nblocks, stride, nval, is_max, out);
}

template <class T>
int launch_maxdiff(const T *a, const T *b, int64_t N, double *partial, hipStream_t s)
{
    int g = grid_for(N, max_local_blocks());
    hipLaunchKernelGGL(k_maxdiff<T>, dim3(g), dim3(BLOCK), 0, s, a, b, N, partial);
    return g;
}

template <class T>
void launch_axpy(const T *y, double c, const T *k, T *out, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_axpy<T>, dim3(grid_for(n)), dim3(BLOCK), 0, s, y, c, k, out, n);
}

template <class T>
int launch_maxabs(const T *v, int64_t n, double *partial, hipStream_t s)
{
    int g = grid_for(n, max_local_blocks());
    hipLaunchKernelGGL(k_maxabs<T>, dim3(g), dim3(BLOCK), 0, s, v, n, partial);
    return g;
}

#define MMX_INST(T)                                                                                          \
    template void launch_aos_to_soa<T>(const double *, T *, int64_t, hipStream_t);                           \
    template void launch_soa_to_aos<T>(const T *, double *, int64_t, hipStream_t);                           \
    template void launch_cast<T>(const double *, T *, int64_t, hipStream_t);                                 \
    template void launch_to_double<T>(const T *, double *, int64_t, hipStream_t);                            \
    template int launch_local<T>(const FieldParams<T> &, const Epi<T> &, const T *, hipStream_t);            \
    template void launch_p1<T>(const DemagGeom &, const T *, int64_t, const T *, int, int, int, int, cplx<T> *, \
                               const cplx<T> *, hipStream_t);                                                \
    template void launch_coly<T>(const DemagGeom &, const cplx<T> *, cplx<T> *, int, int, int, int, int, int,  \
                                 int, const cplx<T> *, hipStream_t);                                         \
    template void launch_convz<T>(const DemagGeom &, cplx<T> *, const T *, const T *, int, int, int, int,      \
                                  const cplx<T> *, hipStream_t);                                                              \
    template int launch_p5<T>(const DemagGeom &, const cplx<T> *, const FieldParams<T> &, const Epi<T> &,     \
                              const T *, const cplx<T> *, hipStream_t);                                      \
    template void launch_tensor_extract<T>(const DemagGeom &, const cplx<double> *, int, double, T *,         \
                                           hipStream_t);                                                     \
    template int launch_maxdiff<T>(const T *, const T *, int64_t, double *, hipStream_t);          \
    template void launch_axpy<T>(const T *, double, const T *, T *, int64_t, hipStream_t);        \
    template int launch_maxabs<T>(const T *, int64_t, double *, hipStream_t);

MMX_INST(double)
MMX_INST(float)

}  // namespace mmx
