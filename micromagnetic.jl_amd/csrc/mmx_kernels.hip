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
#include "mmx_fft.h"
#include "mmx_kernels.h"
#include "mmx_pass.h"

namespace mmx {

// Block shapes.  A pass works on batches of sequences; each sequence of length
// L uses TPS(L) threads holding EPT(L) elements (mmx_fft.h).

// ---------------------------------------------------------------------------
// P1: rows of n_in reals (zero-padded to Px = 2H) -> Kx = H+1 complex values
// per row, via the half-length complex transform of z[j] = x[2j] + i x[2j+1].
// Sequence s = row*ncomp + c; row r of component c starts at in[c*cstride +
// r*n_in]; inputs are multiplied by scale[r*n_in + x] when scale != null
// (mu0 Ms, demag.jl:419-427).  Output A[(c*nrows + row)*KS + k].
// ---------------------------------------------------------------------------
template <class T, int H>
__global__ void __launch_bounds__(Shape<H>::BX) k_p1(DemagGeom G, const T *__restrict__ in, int64_t cstride,
                                                      const T *__restrict__ scale, int n_in, int nrows, int ncomp,
                                                      cplx<T> *__restrict__ A, const cplx<T> *__restrict__ tw)
{
    using S = Shape<H>;
    constexpr int EPT = S::EPT, TPS = S::TPS, SX = S::SX;
    constexpr int HS = FftShape<H>::XS + 1;
    __shared__ cplx<T> lds[SX * HS];
    const int ls = threadIdx.x / TPS, j = threadIdx.x % TPS;
    const int64_t seq = (int64_t)blockIdx.x * SX + ls;
    const bool active = seq < (int64_t)ncomp * nrows;
    // the components of one row are neighbouring sequences, so the row's
    // scale values (mu0 Ms) are fetched once and then hit the cache
    const int c = active ? (int)(seq % ncomp) : 0;
    const int64_t row = active ? seq / ncomp : 0;
    const int64_t rb = row * n_in;
    const T *src = in + c * cstride + rb;
    cplx<T> v[EPT];
    // even rows whose start is a multiple of 2 elements: one 2-element load per
    // pair (x[2k], x[2k+1]) for the input and for the scale
    const bool pairs = (n_in % 2 == 0) && (cstride % 2 == 0);
    // indices past the row are clamped to its last element and the value is
    // replaced by zero: no load waits behind a branch
    const int nh = (n_in + 1) >> 1;
#pragma unroll
    for (int m = 0; m < EPT; m++) {
        const int k = j + m * TPS;
        if (m * TPS >= nh) {  // uniform: the whole element is padding
            v[m] = {0, 0};
            continue;
        }
        const bool in0 = active && 2 * k < n_in, in1 = active && 2 * k + 1 < n_in;
        T a, b;
        if (pairs) {
            const int kc = min(k, nh - 1);
            gld2(src, kc, a, b);
            if (scale) {
                T sa, sb;
                gld2(scale + rb, kc, sa, sb);
                a = a * sa;
                b = b * sb;
            }
        } else {
            const int x0 = min(2 * k, n_in - 1), x1 = min(2 * k + 1, n_in - 1);
            a = scale ? src[x0] * scale[rb + x0] : src[x0];
            b = scale ? src[x1] * scale[rb + x1] : src[x1];
        }
        v[m] = {in0 ? a : T(0), in1 ? b : T(0)};
    }
    cplx<T> *buf = lds + ls * HS;
    fft_reg<T, H, false, 8, RowWave<H>::on>(v, j, buf, tw, G.lgLmax);
    r2c_split_store<T, H, RowWave<H>::on>(v, j, buf, active, A + ((int64_t)c * nrows + row) * G.KS, tw,
                                          G.lgLmax);
}

// ---------------------------------------------------------------------------
// Column transforms along y over tiles of TY consecutive kx columns.
// Input column (plane p, comp c): in[((c*nplanes + p)*Rin + y)*KS + kx], y < n_in.
// Output: out[((c*nplanes + p)*Rout + y)*KS + kx] for y < n_out.
// ---------------------------------------------------------------------------
// FP32 columns are half as wide in bytes: MMX_COLY_F32X times more of them
// per block (up to 1024 threads) keep whole cache lines per row
#ifndef MMX_COLY_F32X
#define MMX_COLY_F32X 2
#endif
#ifndef MMX_COLY_F64X
#define MMX_COLY_F64X 1
#endif
template <class T, int L>
struct ColyShape {
    static constexpr int TPS = Shape<L>::TPS;
    static constexpr int TY0 = Shape<L>::TY * (sizeof(T) == 4 ? MMX_COLY_F32X : MMX_COLY_F64X);
    static constexpr int TY = TY0 * TPS > 1024 ? (1024 / TPS < 1 ? 1 : 1024 / TPS) : TY0;
    static constexpr int BY = TY * TPS;
};

template <class T, int L, bool INV>
__global__ void __launch_bounds__((ColyShape<T, L>::BY)) k_coly(DemagGeom G, const cplx<T> *__restrict__ in,
                                                               cplx<T> *__restrict__ out, int n_in, int Rin,
                                                               int n_out, int Rout, int nplanes,
                                                               const cplx<T> *__restrict__ tw)
{
    using S = Shape<L>;
    constexpr int EPT = S::EPT, TPS = S::TPS, TY = ColyShape<T, L>::TY;
    constexpr int LS = FftShape<L>::XS + 1;
    __shared__ cplx<T> lds[TY * LS];
    const int t = threadIdx.x % TY, j = threadIdx.x / TY;
    const int kx = blockIdx.x * TY + t;
    const int p = blockIdx.y, c = blockIdx.z;
    const int KS = G.KS;
    const bool ok = kx < G.Kx;
    const cplx<T> *src = in + ((int64_t)c * nplanes + p) * Rin * KS + kx;
    cplx<T> v[EPT];
#pragma unroll
    for (int m = 0; m < EPT; m++) {
        const int y = j + m * TPS;
        v[m] = (ok && y < n_in) ? src[(int64_t)y * KS] : cplx<T>{0, 0};
    }
    fft_reg<T, L, INV>(v, j, lds + t * LS, tw, G.lgLmax);
    if (!ok) return;
    cplx<T> *dst = out + ((int64_t)c * nplanes + p) * Rout * KS + kx;
#pragma unroll
    for (int m = 0; m < EPT; m++) {
        const int y = j + m * TPS;
        if (y < n_out) dst[(int64_t)y * KS] = v[m];
    }
}

// ---------------------------------------------------------------------------
// P3: z transform, per-k 3x3 symmetric tensor multiply, inverse, keep z < nz
// (conv mode); or forward only over one component (tensor set-up).  B element
// (c, z, ky, kx) at ((c*nzb + z)*Py + ky)*KS + kx.  Kt holds the real k-space
// tensor folded to ky <= Py/2, kz <= Pz/2: (comp, kzf, kyf, kx) at
// ((comp*Pz2 + kzf)*Py2 + kyf)*KS + kx, with the sign of the odd components
// restored on unfolding (xy odd in y, xz odd in z, yz odd in y and z;
// src/micro/demag.jl:69-80).  The reference keeps real(rfft(tensor)), whose
// odd components also carry the cosine transform of the tensor's on-axis line
// (xy: x=y=0, a function of kz; xz: x=z=0, of ky; yz: y=z=0, of kx) -- the
// rounding residue of the Newell sums that is not odd; Lline holds these three
// even transforms: [Pz2 | Py2 | Kx].
// ---------------------------------------------------------------------------
// MMX_P3_MC: the three components' transforms share each exchange's barriers
// (one LDS buffer per component); MMX_P3_PF: the tensor entries are read
// before the forward transform so that their latency overlaps it.
#ifndef MMX_P3_MC
#define MMX_P3_MC 1
#endif
#ifndef MMX_P3_PF
#define MMX_P3_PF 1
#endif
template <class T, int L, bool FWD_ONLY>
__global__ void __launch_bounds__(Shape<L>::BZ) k_convz(DemagGeom G, cplx<T> *__restrict__ B,
                                                         const T *__restrict__ Kt, const T *__restrict__ Lline,
                                                         int nzb, int n_in, const cplx<T> *__restrict__ tw)
{
    using S = Shape<L>;
    constexpr int EPT = FftShape<L, MMX_P3_E>::EPT, TPS = FftShape<L, MMX_P3_E>::TPS, TZ = S::TZ;
    constexpr int LS = FftShape<L, MMX_P3_E>::XS + 1;
    constexpr int NC = FWD_ONLY ? 1 : 3;
    // one exchange buffer per component when the three transforms share barriers
    constexpr int CS = 2 * TZ * LS;
    constexpr int NB = (MMX_P3_MC && !FWD_ONLY && 3 * CS * sizeof(cplx<T>) <= 65536) ? 3 : 1;
    __shared__ cplx<T> lds[NB * CS];
    const int half = threadIdx.x / (TZ * TPS);
    const int tl = threadIdx.x % (TZ * TPS);
    const int t = tl % TZ, j = tl / TZ;
    const int kx = blockIdx.x * TZ + t;
    const int KS = G.KS, Py = G.Py;
    const int ky0 = blockIdx.y;
    // half 1 takes the mirror Py-ky0 when it is a distinct row (0 < ky0 < Py/2)
    const int ky = half ? Py - ky0 : ky0;
    const bool ok = kx < G.Kx && (half == 0 || (ky0 >= 1 && ky0 < Py / 2));
    const int64_t cs_b = (int64_t)nzb * Py * KS;  // component stride in B
    const int64_t zs_b = (int64_t)Py * KS;        // z stride in B
    cplx<T> *col = B + (int64_t)blockIdx.z * 3 * cs_b + (int64_t)(ok ? ky : 0) * KS + (ok ? kx : 0);
    cplx<T> v[NC][EPT];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int m = 0; m < EPT; m++) {
            const int z = j + m * TPS;
            v[c][m] = (ok && z < n_in) ? col[c * cs_b + z * zs_b] : cplx<T>{0, 0};
        }
    cplx<T> *buf = lds + (half * TZ + t) * LS;
    if constexpr (FWD_ONLY) {
        fft_reg<T, L, false, MMX_P3_E>(v[0], j, buf, tw, G.lgLmax);
        if (!ok) return;
#pragma unroll
        for (int m = 0; m < EPT; m++) col[(int64_t)(j + m * TPS) * zs_b] = v[0][m];
        return;
    } else {
        const int kyf = (ky <= Py / 2) ? ky : Py - ky;
        const T sy = (ky <= Py / 2) ? T(1) : T(-1);
        const int64_t cs = (int64_t)G.Pz2 * G.Py2 * KS;  // component stride of Kt
        const int kxs = ok ? kx : 0;
        // tensor entries of this thread's kz (read before the transform when
        // MMX_P3_PF so that their latency overlaps it)
        T kxx[EPT], kyy[EPT], kzz[EPT], kxy[EPT], kxz[EPT], kyz[EPT];
        auto load_tensor = [&]() {
            const T lyz = Lline[G.Pz2 + G.Py2 + kxs];
            const T lxz = Lline[G.Pz2 + kyf];
#pragma unroll
            for (int m = 0; m < EPT; m++) {
                const int kz = j + m * TPS;
                const int kzf = (kz <= L / 2) ? kz : L - kz;
                const T sz = (kz <= L / 2) ? T(1) : T(-1);
                const T *kp = Kt + ((int64_t)kzf * G.Py2 + kyf) * KS + kxs;
                kxx[m] = kp[0];
                kyy[m] = kp[cs];
                kzz[m] = kp[2 * cs];
                kxy[m] = sy * kp[3 * cs] + Lline[kzf];
                kxz[m] = sz * kp[4 * cs] + lxz;
                kyz[m] = sy * sz * kp[5 * cs] + lyz;
            }
        };
        if constexpr (MMX_P3_PF) load_tensor();
        if constexpr (NB == 3) {
            fft_reg_multi<T, L, false, MMX_P3_E, 3>(&v[0][0], j, buf, CS, tw, G.lgLmax);
        } else {
#pragma unroll
            for (int c = 0; c < NC; c++) fft_reg<T, L, false, MMX_P3_E>(v[c], j, buf, tw, G.lgLmax);
        }
        if constexpr (!MMX_P3_PF) load_tensor();
#pragma unroll
        for (int m = 0; m < EPT; m++) {
            const T xx = kxx[m], yy = kyy[m], zz = kzz[m], xy = kxy[m], xz = kxz[m], yz = kyz[m];
            const cplx<T> Mx = v[0][m], My = v[1][m], Mz = v[2][m];
            v[0][m] = {xx * Mx.x + xy * My.x + xz * Mz.x, xx * Mx.y + xy * My.y + xz * Mz.y};
            v[1][m] = {xy * Mx.x + yy * My.x + yz * Mz.x, xy * Mx.y + yy * My.y + yz * Mz.y};
            v[2][m] = {xz * Mx.x + yz * My.x + zz * Mz.x, xz * Mx.y + yz * My.y + zz * Mz.y};
        }
        if constexpr (NB == 3) {
            fft_reg_multi<T, L, true, MMX_P3_E, 3>(&v[0][0], j, buf, CS, tw, G.lgLmax);
        } else {
#pragma unroll
            for (int c = 0; c < NC; c++) fft_reg<T, L, true, MMX_P3_E>(v[c], j, buf, tw, G.lgLmax);
        }
        if (!ok) return;
        const int nz = G.nz;
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int m = 0; m < EPT; m++) {
                const int z = j + m * TPS;
                if (z < nz) col[c * cs_b + z * zs_b] = v[c][m];
            }
    }
}

// ---------------------------------------------------------------------------
// P3, wave form (conv mode, TPS = Pz/4 <= 64): every z sequence lies in one
// wave, so the exchanges need no block barrier and the waves of a block run
// independently.  A block is 4 waves: waves 0 and 1 take TW = 64/TPS
// consecutive kx each at row ky0, waves 2 and 3 the same kx at the mirror row
// Py-ky0 (same folded tensor entries); lane = j*TW + t.  The folded tensor is
// stored in groups of 4 kx, (comp, kyf, kx/4, kzf, kx%4), so that the entries
// of one wave for all kz form one contiguous run.
// ---------------------------------------------------------------------------
#ifndef MMX_P3_WAVE
#define MMX_P3_WAVE 1
#endif
// Transforms of zero-padded data: the nonzero inputs and the wanted outputs
// of a length-L convolution transform lie in its first half (n <= L/2, since
// L >= 2n - 1), i.e. in elements m < EPT/2 of every thread; with MMX_PRUNE the
// others are compile-time zeros / dead values, which prunes the first forward
// and the last inverse stage.
#ifndef MMX_PRUNE
#define MMX_PRUNE 1
#endif
// elements per thread of the wave-form z transform for FP32 (FP64: MMX_P3_E)
#ifndef MMX_P3_E32
#define MMX_P3_E32 8
#endif
// Waves of a block along kx: 2 (the other two waves take the mirror row
// Py-ky, which reads the same folded tensor entries) or 4 (the block covers 4
// wave widths of one row and the grid every ky).  Long z transforms leave a
// wave few kx columns (TW = 64/TPS); from TW <= MMX_P3W_WX4 on the wider
// tile measured faster (Pz = 128: 1634 -> 1462 us), below it slower.
#ifndef MMX_P3W_WX4
#define MMX_P3W_WX4 2
#endif
template <class T>
struct P3W {
    static constexpr int E = sizeof(T) == 8 ? MMX_P3_E : MMX_P3_E32;
    template <int L>
    static constexpr int wx()
    {
        return 64 / FftShape<L, E>::TPS <= MMX_P3W_WX4 ? 4 : 2;
    }
};
#ifndef MMX_P3_NBW
#define MMX_P3_NBW 3
#endif
// exchange-buffer padding per sequence, in elements (LDS banks: 16-byte
// elements are conflict-free on the reads at +2, 8-byte ones at +4)
#ifndef MMX_P3_PAD
#define MMX_P3_PAD (sizeof(T) == 8 ? 2 : 4)
#endif
#ifndef MMX_P3W_PF
#define MMX_P3W_PF 1
#endif
// Tensor loads in kx pairs: the two lanes of a kx pair (t even/odd) load
// both kx of three components each (one 2-value load per component) and swap
// halves with a DPP move, and a quad of lanes (same j) loads four Lline[kzf]
// at once and broadcasts them: per element 3 + 1/4 vector loads per lane
// instead of 6 + 1
#ifndef MMX_P3W_PAIR
#define MMX_P3W_PAIR 1
#endif
template <class T, int L>
__global__ void __launch_bounds__(256) k_convw(DemagGeom G, cplx<T> *__restrict__ B, const T *__restrict__ Kt,
                                               const T *__restrict__ Lline, int n_in, int nzb,
                                               const cplx<T> *__restrict__ tw)
{
    using F = FftShape<L, P3W<T>::E>;
    constexpr int EPT = F::EPT, TPS = F::TPS;
    static_assert(TPS <= 64, "one sequence per wave");
    constexpr bool PR = MMX_PRUNE && L >= 2;  // Pz = 1 (nz = 1) is the one unpadded length
    constexpr int TW = 64 / TPS;
    constexpr int LS = F::XS + MMX_P3_PAD;
    constexpr int NB = MMX_P3_NBW;
    constexpr int WS = TW * LS;
    __shared__ cplx<T> lds[4 * NB * WS];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int t = lane % TW, j = lane / TW;
    constexpr int WX = P3W<T>::template wx<L>();
    const int half = WX == 2 ? wave >> 1 : 0;
    const int Py = G.Py, KS = G.KS;
    const int ky0 = blockIdx.y;
    // the mirror row exists only for 0 < ky0 < Py/2 (a whole wave's decision)
    if (half && !(ky0 >= 1 && ky0 < Py / 2)) return;
    const int ky = half ? Py - ky0 : ky0;
    const int kx = blockIdx.x * (WX * TW) + (wave % WX) * TW + t;
    const bool ok = kx < G.Kx;
    const int kxs = ok ? kx : 0;
    const int64_t cs_b = (int64_t)nzb * Py * KS;
    const int64_t zs_b = (int64_t)Py * KS;
    cplx<T> *col = B + (int64_t)blockIdx.z * 3 * cs_b + (int64_t)ky * KS + kxs;  // grid z: image
    cplx<T> v[3][EPT];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int m = 0; m < EPT; m++) {
            const int z = j + m * TPS;
            v[c][m] = (ok && z < n_in && (!PR || m < EPT / 2)) ? col[c * cs_b + z * zs_b] : cplx<T>{0, 0};
        }
    const int kyf = (ky <= Py / 2) ? ky : Py - ky;
    const T sy = (ky <= Py / 2) ? T(1) : T(-1);
    const int Pz2 = G.Pz2;
    const int64_t cs = (int64_t)G.Py2 * KS * Pz2;
    const T *kb = Kt + (((int64_t)kyf * (KS >> 2) + (kxs >> 2)) * Pz2) * 4 + (kxs & 3);
    T kxx[EPT], kyy[EPT], kzz[EPT], kxy[EPT], kxz[EPT], kyz[EPT];
    constexpr bool PAIR = MMX_P3W_PAIR && TW % 2 == 0;
    // the kz line table by lane quads needs 4 kx lanes per j and 4 | EPT
    constexpr bool LQ = TW % 4 == 0 && EPT % 4 == 0;
    // raw 16-byte loads (issued before the forward transform) and the lane
    // exchanges that complete them (after it, so the transform does not wait
    // for the loads: the wave barriers of its exchanges keep the DPP moves
    // from being scheduled above it)
    cplx<T> pa[PAIR ? EPT : 1], pb[PAIR ? EPT : 1], pc[PAIR ? EPT : 1];
    T lq[PAIR ? (LQ ? EPT / 4 : EPT) : 1], lyz = 0, lxz = 0;
    const int e = t & 1;
    auto load_tensor_pairs = [&]() {
        lyz = Lline[Pz2 + G.Py2 + kxs];
        lxz = Lline[Pz2 + kyf];
        if constexpr (LQ) {
            // lane (j, t) loads Lline at kz = j + (4h + (t&3))*TPS for h < EPT/4
#pragma unroll
            for (int h = 0; h < EPT / 4; h++) {
                const int kzq = j + (4 * h + (t & 3)) * TPS;
                lq[h] = Lline[kzq <= L / 2 ? kzq : L - kzq];
            }
        } else {
#pragma unroll
            for (int m = 0; m < EPT; m++) {
                const int kz = j + m * TPS;
                lq[m] = Lline[kz <= L / 2 ? kz : L - kz];
            }
        }
        // kx pair (2pk, 2pk+1), clamped into the padded rows (lanes past them hold no data)
        const int pk = min(kx >> 1, (KS >> 1) - 1);
        const T *kbp = Kt + (((int64_t)kyf * (KS >> 2) + (pk >> 1)) * Pz2) * 4 + 2 * (pk & 1) + e * cs;
#pragma unroll
        for (int m = 0; m < EPT; m++) {
            const int kz = j + m * TPS;
            const int kzf = (kz <= L / 2) ? kz : L - kz;
            const T *kp = kbp + kzf * 4;
            // even lanes load components 0, 2, 4 (xx, zz, xz), odd lanes 1, 3, 5 (yy, xy, yz)
            gld2(kp, 0, pa[m].x, pa[m].y);
            gld2(kp + 2 * cs, 0, pb[m].x, pb[m].y);
            gld2(kp + 4 * cs, 0, pc[m].x, pc[m].y);
        }
    };
    auto finish_tensor_pairs = [&]() {
        T lz[EPT];
        if constexpr (LQ) {
            // element m is taken from lane m%4 of the quad (load m/4)
#pragma unroll
            for (int h = 0; h < EPT / 4; h++) {
                lz[4 * h + 0] = quad_perm<QP_BCAST<0>>(lq[h]);
                lz[4 * h + 1] = quad_perm<QP_BCAST<1>>(lq[h]);
                lz[4 * h + 2] = quad_perm<QP_BCAST<2>>(lq[h]);
                lz[4 * h + 3] = quad_perm<QP_BCAST<3>>(lq[h]);
            }
        } else {
#pragma unroll
            for (int m = 0; m < EPT; m++) lz[m] = lq[m];
        }
#pragma unroll
        for (int m = 0; m < EPT; m++) {
            const int kz = j + m * TPS;
            const T sz = (kz <= L / 2) ? T(1) : T(-1);
            const cplx<T> a = pa[m], b = pb[m], c = pc[m];
            const T ra = quad_perm<QP_XOR1>(e ? a.x : a.y);
            const T rb = quad_perm<QP_XOR1>(e ? b.x : b.y);
            const T rc = quad_perm<QP_XOR1>(e ? c.x : c.y);
            const T oa = e ? a.y : a.x, ob = e ? b.y : b.x, oc = e ? c.y : c.x;
            kxx[m] = e ? ra : oa;
            kyy[m] = e ? oa : ra;
            kzz[m] = e ? rb : ob;
            kxy[m] = sy * (e ? ob : rb) + lz[m];
            kxz[m] = sz * (e ? rc : oc) + lxz;
            kyz[m] = sy * sz * (e ? oc : rc) + lyz;
        }
    };
    auto load_tensor = [&]() {
        if constexpr (PAIR) {
            load_tensor_pairs();
        } else {
            const T lyz = Lline[Pz2 + G.Py2 + kxs];
            const T lxz = Lline[Pz2 + kyf];
#pragma unroll
            for (int m = 0; m < EPT; m++) {
                const int kz = j + m * TPS;
                const int kzf = (kz <= L / 2) ? kz : L - kz;
                const T sz = (kz <= L / 2) ? T(1) : T(-1);
                const T *kp = kb + kzf * 4;
                kxx[m] = kp[0];
                kyy[m] = kp[cs];
                kzz[m] = kp[2 * cs];
                kxy[m] = sy * kp[3 * cs] + Lline[kzf];
                kxz[m] = sz * kp[4 * cs] + lxz;
                kyz[m] = sy * sz * kp[5 * cs] + lyz;
            }
        }
    };
    // one set of twiddle bases serves the forward and the inverse transform;
    // loaded before the tensor (loads complete in order, and the forward
    // transform needs the twiddles, not the tensor)
    const TwSet<T, L, P3W<T>::E> ts = fft_twiddles<T, L, P3W<T>::E>(j, tw, G.lgLmax);
    if constexpr (MMX_P3W_PF) load_tensor();
    cplx<T> *buf = lds + wave * NB * WS + t * LS;
    if constexpr (NB == 3) {
        fft_wave_multi<T, L, false, P3W<T>::E, 3>(&v[0][0], j, buf, WS, ts);
    } else {
#pragma unroll
        for (int c = 0; c < 3; c++) fft_wave_multi<T, L, false, P3W<T>::E, 1>(v[c], j, buf, 0, ts);
    }
    if constexpr (!MMX_P3W_PF) load_tensor();
    if constexpr (PAIR) finish_tensor_pairs();
#pragma unroll
    for (int m = 0; m < EPT; m++) {
        const T xx = kxx[m], yy = kyy[m], zz = kzz[m], xy = kxy[m], xz = kxz[m], yz = kyz[m];
        const cplx<T> Mx = v[0][m], My = v[1][m], Mz = v[2][m];
        v[0][m] = {xx * Mx.x + xy * My.x + xz * Mz.x, xx * Mx.y + xy * My.y + xz * Mz.y};
        v[1][m] = {xy * Mx.x + yy * My.x + yz * Mz.x, xy * Mx.y + yy * My.y + yz * Mz.y};
        v[2][m] = {xz * Mx.x + yz * My.x + zz * Mz.x, xz * Mx.y + yz * My.y + zz * Mz.y};
    }
    if constexpr (NB == 3) {
        fft_wave_multi<T, L, true, P3W<T>::E, 3>(&v[0][0], j, buf, WS, ts);
    } else {
#pragma unroll
        for (int c = 0; c < 3; c++) fft_wave_multi<T, L, true, P3W<T>::E, 1>(v[c], j, buf, 0, ts);
    }
    if (!ok) return;
    const int nz = G.nz;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int m = 0; m < EPT; m++) {
            const int z = j + m * TPS;
            if ((!PR || m < EPT / 2) && z < nz) col[c * cs_b + z * zs_b] = v[c][m];
        }
}

// ---------------------------------------------------------------------------
// P5: half-spectrum -> real along x for RX rows x 3 components, then the
// fused epilogue for every cell of those rows.  A element (c, row, kx) at
// (c*ny*nz + row)*KS + kx.
// ---------------------------------------------------------------------------
// V selects the epilogue: 0 = any mode with the generic field (cell_field),
// or the slot-form specialisations 1 = EPI_STAGE, 2 = EPI_RK4_FINAL, 3 = EPI_ERR.



// ---------------------------------------------------------------------------
// Batches of systems without demag: one persistent launch per stage covers
// every member.  Work item w = (member, first cell) spans BLOCK*CPT cells of
// one member; blocks b with b % 8 == x (one XCD) take the contiguous item
// range x*ceil(n/8) ... so that the y/z neighbour rows of a member's item
// were fetched by the same XCD's L2 shortly before.
// ---------------------------------------------------------------------------
template <class T, int V, int CPT, int BS>
__global__ void __launch_bounds__(BS) k_batch_local(const int2 *__restrict__ items, int nitems,
                                                        const FieldParams<T> *const *__restrict__ Ps,
                                                        const BatchStage<T> *__restrict__ st,
                                                        const Dyn *__restrict__ dyns)
{
    const int G = gridDim.x;
    const int xcd = blockIdx.x & 7, r = blockIdx.x >> 3;
    const int Gx = (G - xcd + 7) >> 3;
    const int per = (nitems + 7) >> 3;
    const int lo = xcd * per, hi = min(nitems, lo + per);
    for (int it = lo + r; it < hi; it += Gx) {
        const int2 w = items[it];
        const MMX_CONST FieldParams<T> &P = *(const MMX_CONST FieldParams<T> *)Ps[w.x];
        // the member's epilogue and per-launch values are read in place (a private
        // copy of these structs would live in scratch: their arrays are indexed
        // by term)
        const Epi<T> &E = st[w.x].E;
        const T *m = st[w.x].m;
        const Dyn &d = dyns[w.x];
        const unsigned nx = P.nx, ny = P.ny;
        const int N = (int)P.N;
        double red = 0;
        for (int c = 0; c < CPT; c++) {
            const int I = w.y + c * BS + threadIdx.x;
            if (I < N) {
                const unsigned q = (unsigned)I / nx;
                const int i = (int)((unsigned)I - q * nx);
                const unsigned k = q / ny;
                const int j = (int)(q - k * ny);
                epilogue<T, V, 0u, MM_STAGES>(P, d, E, I, i, j, (int)k, m, T(0), T(0), T(0), red);
            }
        }
    }
}

// the thermal noise of every member moves to its next integrator step
static __global__ void k_batch_tick(Dyn *__restrict__ dyns, int M)
{
    for (int m = blockIdx.x * BLOCK + threadIdx.x; m < M; m += gridDim.x * BLOCK) dyns[m].noise_step += 1;
}


// ---------------------------------------------------------------------------
// Nudged elastic band (src/neb/neb.jl, neb_kernels.jl).  Images are stored
// image-major, each as three planes of N values.  Grids are (blocks per
// image, images); per-image block partials go to partial[(q*gridDim.x+b)*S+v].
// ---------------------------------------------------------------------------
// effective_field_energy(neb, spin) for meshes without demag (neb.jl:220-230):
// the field and energy of every free image in one launch.  Image q reads
// m + q*3N and writes its field to out + q*3N; each block leaves the sum of its
// cells' energies, each rounded to T as the per-cell energy array holds it, in
// partial[q*gridDim.x + b].
template <class T>
__global__ void __launch_bounds__(BLOCK) k_neb_field(const FieldParams<T> *__restrict__ Pp, Dyn dyn, uint32_t mask,
                                                     const T *__restrict__ Y, T *__restrict__ F,
                                                     double *__restrict__ partial)
{
    const MMX_CONST FieldParams<T> &P = *(const MMX_CONST FieldParams<T> *)Pp;
    const unsigned nx = P.nx, ny = P.ny;
    const int N = (int)P.N;
    const int64_t off = (int64_t)blockIdx.y * 3 * N;
    const T *m = Y + off;
    T *out = F + off;
    double r[1] = {0};
    for (int I = blockIdx.x * BLOCK + threadIdx.x; I < N; I += gridDim.x * BLOCK) {
        const unsigned q = (unsigned)I / nx;
        const int i = (int)((unsigned)I - q * nx);
        const unsigned k = q / ny;
        const int j = (int)(q - k * ny);
        T Hx, Hy, Hz;
        double es = 0;
        cell_field(P, dyn, mask, I, i, j, (int)k, m, T(0), T(0), T(0), Hx, Hy, Hz, &es);
        r[0] += (double)(T)es;
        out[I] = Hx;
        out[N + I] = Hy;
        out[2 * N + I] = Hz;
    }
    block_reduce_store<1>(r, 1, false, partial + (size_t)blockIdx.y * gridDim.x, 1);
}

// per-segment sums: out[q] = sum_b partial[q*nb + b]
static __global__ void k_segment_sum(const double *__restrict__ partial, int nb, double *__restrict__ out)
{
    __shared__ double s[BLOCK];
    const int q = blockIdx.x;
    double a = 0;
    for (int i = threadIdx.x; i < nb; i += BLOCK) a += partial[(size_t)q * nb + i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int h = BLOCK / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = s[0];
}

// max over every image and cell of |a - b| (relax(::NEB), compute_dm!), one launch
template <class T>
__global__ void k_neb_maxdiff(const T *__restrict__ a, const T *__restrict__ b, int64_t N,
                              double *__restrict__ partial)
{
    const int64_t off = (int64_t)blockIdx.y * 3 * N;
    a += off;
    b += off;
    double red[1] = {0};
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < N; i += (int64_t)gridDim.x * BLOCK) {
        T dx = a[i] - b[i], dy = a[N + i] - b[N + i], dz = a[2 * N + i] - b[2 * N + i];
        red[0] = pmax(red[0], (double)sqrt(dx * dx + dy * dy + dz * dz));
    }
    block_reduce_store<1>(red, 1, true, partial + (size_t)blockIdx.y * gridDim.x, 1);
}

// sum of n values (per-cell energies -> image energy)
template <class T>
__global__ void k_sum(const T *__restrict__ v, int64_t n, double *__restrict__ partial)
{
    double r[1] = {0};
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * BLOCK)
        r[0] += (double)v[i];
    block_reduce_store<1>(r, 1, false, partial, 1);
}

// compute_distance_kernel! (neb_kernels.jl:66-81): squared geodesic angles of
// image pair q = (q-1, q) (q = 0: left end, q = n: right end), summed per block
template <class T>
__global__ void k_neb_dist(int64_t N, int n, const T *__restrict__ Y, const T *__restrict__ L,
                           const T *__restrict__ R, double *__restrict__ partial)
{
    const int q = blockIdx.y;
    const T *a = q == 0 ? L : Y + (int64_t)(q - 1) * 3 * N;
    const T *b = q == n ? R : Y + (int64_t)q * 3 * N;
    double r[1] = {0};
    for (int64_t I = (int64_t)blockIdx.x * BLOCK + threadIdx.x; I < N; I += (int64_t)gridDim.x * BLOCK) {
        const T mx1 = a[I], my1 = a[N + I], mz1 = a[2 * N + I];
        const T mx2 = b[I], my2 = b[N + I], mz2 = b[2 * N + I];
        const T cx = -my2 * mz1 + my1 * mz2;
        const T cy = mx2 * mz1 - mx1 * mz2;
        const T cz = -mx2 * my1 + mx1 * my2;
        const T mm = mx1 * mx2 + my1 * my2 + mz1 * mz2;
        const T d = sqrt(cx * cx + cy * cy + cz * cz);
        const double ang = (double)atan2(d, mm);
        r[0] += ang * ang;
    }
    block_reduce_store<1>(r, 1, false, partial + (size_t)q * gridDim.x, 1);
}

// compute_tangents_kernel! + reduce_tangent_kernel! (neb_kernels.jl:84-109)
// at the stage vector Ys; per-block partials of |t|^2 and f.t
template <class T>
__global__ void k_neb_tangent(int64_t N, int n, const T *__restrict__ Ys, const T *__restrict__ L,
                              const T *__restrict__ R, const double *__restrict__ energy, const T *__restrict__ F,
                              T *__restrict__ tan, double *__restrict__ partial)
{
#pragma clang fp contract(off)
    const int q = blockIdx.y;
    const int64_t V = 3 * N;
    const double E1 = energy[q], E2 = energy[q + 1], E3 = energy[q + 2];
    const double dEmax = fmax(fabs(E3 - E2), fabs(E2 - E1));
    const double dEmin = fmin(fabs(E3 - E2), fabs(E2 - E1));
    int br;
    if (E1 > E2 && E2 > E3) br = 0;
    else if (E3 > E2 && E2 > E1) br = 1;
    else if (E3 > E1) br = 2;
    else if (E3 < E1) br = 3;
    else br = 4;
    const T *m = Ys + q * V;
    const T *mp = q == n - 1 ? R : Ys + (q + 1) * V;
    const T *mm = q == 0 ? L : Ys + (q - 1) * V;
    const T *f = F + q * V;
    T *t = tan + q * V;
    double r[2] = {0, 0};
    for (int64_t I = (int64_t)blockIdx.x * BLOCK + threadIdx.x; I < N; I += (int64_t)gridDim.x * BLOCK) {
        T tv[3], mv[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int64_t o = c * N + I;
            const T mk = m[o];
            const T tip = mp[o] - mk;
            const T tim = mk - mm[o];
            T x;
            if (br == 0) x = tim;
            else if (br == 1) x = tip;
            else if (br == 2) x = (T)dEmax * tip + (T)dEmin * tim;
            else if (br == 3) x = (T)dEmin * tip + (T)dEmax * tim;
            else x = tim + tip;
            tv[c] = x;
            mv[c] = mk;
        }
        const T d = tv[0] * mv[0] + tv[1] * mv[1] + tv[2] * mv[2];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int64_t o = c * N + I;
            tv[c] -= d * mv[c];
            t[o] = tv[c];
            r[0] += (double)tv[c] * (double)tv[c];
            r[1] += (double)f[o] * (double)tv[c];
        }
    }
    block_reduce_store<2>(r, 2, false, partial + (size_t)q * gridDim.x * 2, 2);
}

// per image: reduce the block partials; coef[q] = {1/|t|, s_q}, where the NEB
// force is f + s_q t/|t| with s_q = k (d[q+1] - d[q]) - f.t/|t| (or -2 f.t/|t|
// for the climbing image), src/neb/neb.jl:196-216
static __global__ void k_neb_coef(const double *__restrict__ partial, int nb, const double *__restrict__ dist, double k,
                           int climb, double *__restrict__ coef)
{
    __shared__ double s0[BLOCK], s1[BLOCK];
    const int q = blockIdx.x;
    double a = 0, b = 0;
    for (int i = threadIdx.x; i < nb; i += BLOCK) {
        a += partial[((size_t)q * nb + i) * 2];
        b += partial[((size_t)q * nb + i) * 2 + 1];
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    __syncthreads();
    for (int h = BLOCK / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s0[threadIdx.x] += s0[threadIdx.x + h];
            s1[threadIdx.x] += s1[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double nrm = sqrt(s0[0]);
        const double ft = s1[0] / nrm;
        coef[2 * q] = nrm;
        coef[2 * q + 1] = (q + 1 == climb) ? -2.0 * ft : k * (dist[q + 1] - dist[q]) - ft;
    }
}

// per image pair: distance = sqrt(sum of block partials)
static __global__ void k_neb_dist_final(const double *__restrict__ partial, int nb, double *__restrict__ dist)
{
    __shared__ double s[BLOCK];
    const int q = blockIdx.x;
    double a = 0;
    for (int i = threadIdx.x; i < nb; i += BLOCK) a += partial[(size_t)q * nb + i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int h = BLOCK / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) dist[q] = sqrt(s[0]);
}

// NEB force + LLG (precession as set) at the stage vector, then the BS23
// stage combination over the whole band (adaptive_rk_steps.jl:11-47):
//   mode 0: kstore = k; yout = y0 + c0 k
//   mode 1: kstore = k; yout = normalise(y0 + outer ((c0 k1 + c1 k2) + c2 k))
//   mode 2: kstore = k; err = ((e0 k1 + e1 k2) + e2 k3) + e3 k; max |err|/(1+|Ys|)
//   mode 3: kstore = k; max |k|
struct NebStage {
    int mode;
    double c[4];
    double outer;
};

template <class T>
__global__ void k_neb_stage(const FieldParams<T> *__restrict__ Pp, int64_t N, const T *__restrict__ Ys,
                            const T *__restrict__ F, const T *__restrict__ tan, const double *__restrict__ coef,
                            NebStage S, T *__restrict__ kstore, const T *__restrict__ y0,
                            const T *__restrict__ k1, const T *__restrict__ k2, const T *__restrict__ k3,
                            T *__restrict__ yout, double *__restrict__ partial)
{
#pragma clang fp contract(off)
    const MMX_CONST FieldParams<T> &P = *(const MMX_CONST FieldParams<T> *)Pp;
    const int q = blockIdx.y;
    const int64_t V = 3 * N, base = q * V;
    const double nrm = coef[2 * q], sq = coef[2 * q + 1];
    double red = 0;
    for (int64_t I = (int64_t)blockIdx.x * BLOCK + threadIdx.x; I < N; I += (int64_t)gridDim.x * BLOCK) {
        T mv[3], hv[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int64_t o = base + c * N + I;
            const T th = tan[o] / (T)nrm;
            hv[c] = F[o] + (T)sq * th;
            mv[c] = Ys[o];
        }
        T kk[3];
        llg_rhs(P, I, mv[0], mv[1], mv[2], hv[0], hv[1], hv[2], kk[0], kk[1], kk[2]);
        if (S.mode == 1) {
            T y[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const int64_t o = base + c * N + I;
                const T s = ((T)S.c[0] * k1[o] + (T)S.c[1] * k2[o]) + (T)S.c[2] * kk[c];
                y[c] = y0[o] + (T)S.outer * s;
            }
            const T m2 = y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
            if (m2 > T(0)) {
                const T len = T(1) / sqrt(m2);
                y[0] *= len;
                y[1] *= len;
                y[2] *= len;
            }
#pragma unroll
            for (int c = 0; c < 3; c++) yout[base + c * N + I] = y[c];
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int64_t o = base + c * N + I;
            kstore[o] = kk[c];
            if (S.mode == 0) {
                yout[o] = y0[o] + (T)S.c[0] * kk[c];
            } else if (S.mode == 2) {
                const double e = ((S.c[0] * (double)k1[o] + S.c[1] * (double)k2[o]) + S.c[2] * (double)k3[o]) +
                                 S.c[3] * (double)kk[c];
                red = pmax(red, fabs(e / (1.0 + fabs((double)mv[c]))));
            } else if (S.mode == 3) {
                red = pmax(red, fabs((double)kk[c]));
            }
        }
    }
    if (S.mode >= 2) {
        double r[1] = {red};
        block_reduce_store<1>(r, 1, true, partial + (size_t)q * gridDim.x, 1);
    }
}


// neighbour / pin mask of every cell (FieldParams::nbm), from the same rules
// as neighbours() and the Ms > 0 test of the stencil kernels
template <class T>
__global__ void k_nbmask(const FieldParams<T> *__restrict__ Pp, uint8_t *__restrict__ out)
{
    const MMX_CONST FieldParams<T> &P = *(const MMX_CONST FieldParams<T> *)Pp;
    for (int64_t I = (int64_t)blockIdx.x * BLOCK + threadIdx.x; I < P.N; I += (int64_t)gridDim.x * BLOCK) {
        const int i = (int)(I % P.nx);
        const int64_t q = I / P.nx;
        const int j = (int)(q % P.ny);
        const int k = (int)(q / P.ny);
        int64_t nb[6];
        neighbours(P, i, j, k, nb);
        uint32_t msk = 0;
#pragma unroll
        for (int d = 0; d < 6; d++)
            if (nb[d] >= 0 && P.mu0Ms[nb[d]] > T(0)) msk |= 1u << d;
        if (P.pins && P.pins[I]) msk |= 1u << 6;
        out[I] = (uint8_t)msk;
    }
}


// ---------------------------------------------------------------------------
// Steepest descent with Barzilai-Borwein steps (src/sd.jl)
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void sd_cross(T x1, T x2, T x3, T y1, T y2, T y3, T &a, T &b, T &c)
{
    // cross_product: src/util.jl:30-33
    a = -x3 * y2 + x2 * y3;
    b = x3 * y1 - x1 * y3;
    c = -x2 * y1 + x1 * y2;
}

// compute_gk_kernel_1! / _2! (sd.jl:9-37): g = m x (m x h); with prev != null
// also the block sums of s.s, s.f, f.f (s = m - m_pre, f = g - g_old)
template <class T>
__global__ void k_sd_g(int64_t N, const T *__restrict__ m, const T *__restrict__ mpre, const T *__restrict__ h,
                       T *__restrict__ g, double *__restrict__ partial)
{
#pragma clang fp contract(off)
    double r[3] = {0, 0, 0};
    for (int64_t I = (int64_t)blockIdx.x * BLOCK + threadIdx.x; I < N; I += (int64_t)gridDim.x * BLOCK) {
        const T mx = m[I], my = m[N + I], mz = m[2 * N + I];
        T fx, fy, fz, gx, gy, gz;
        sd_cross(mx, my, mz, h[I], h[N + I], h[2 * N + I], fx, fy, fz);
        sd_cross(mx, my, mz, fx, fy, fz, gx, gy, gz);
        if (mpre) {
            const T sx = mx - mpre[I], sy = my - mpre[N + I], sz = mz - mpre[2 * N + I];
            const T dx = gx - g[I], dy = gy - g[N + I], dz = gz - g[2 * N + I];
            r[0] += (double)(sx * sx + sy * sy + sz * sz);
            r[1] += (double)(sx * dx + sy * dy + sz * dz);
            r[2] += (double)(dx * dx + dy * dy + dz * dz);
        }
        g[I] = gx;
        g[N + I] = gy;
        g[2 * N + I] = gz;
    }
    if (mpre) block_reduce_store<3>(r, 3, false, partial, 3);
}

// compute_tau (sd.jl:39-66) from the three sums; one thread
static __global__ void k_sd_tau(const double *__restrict__ sums, int odd, double min_tau, double max_tau,
                         double *__restrict__ tau)
{
    const double s1 = sums[0], s2 = sums[1], s3 = sums[2];
    const double tau1 = s2 != 0.0 ? s1 / s2 : min_tau;
    const double tau2 = s3 != 0.0 ? s2 / s3 : min_tau;
    double t = odd ? fabs(tau1) : fabs(tau2);
    if (t > max_tau) t = max_tau;
    tau[0] = t;
}

// run_step_kernel! (sd.jl:39-55): prespin = m, then the Cayley-like update
template <class T>
__global__ void k_sd_update(const FieldParams<T> *__restrict__ Pp, const T *__restrict__ g, T *__restrict__ m,
                            T *__restrict__ mpre, const T *__restrict__ h, const double *__restrict__ taup)
{
#pragma clang fp contract(off)
    const MMX_CONST FieldParams<T> &P = *(const MMX_CONST FieldParams<T> *)Pp;
    const int64_t N = P.N;
    const T tau = (T)taup[0];
    for (int64_t I = (int64_t)blockIdx.x * BLOCK + threadIdx.x; I < N; I += (int64_t)gridDim.x * BLOCK) {
        const T mx = m[I], my = m[N + I], mz = m[2 * N + I];
        mpre[I] = mx;
        mpre[N + I] = my;
        mpre[2 * N + I] = mz;
        if (P.pins && P.pins[I]) continue;
        T fx, fy, fz;
        sd_cross(mx, my, mz, h[I], h[N + I], h[2 * N + I], fx, fy, fz);
        const T factor = T(0.25) * (fx * fx + fy * fy + fz * fz) * (tau * tau);
        const T ax = (T(1) - factor) * mx - tau * g[I];
        const T ay = (T(1) - factor) * my - tau * g[N + I];
        const T az = (T(1) - factor) * mz - tau * g[2 * N + I];
        m[I] = ax / (T(1) + factor);
        m[N + I] = ay / (T(1) + factor);
        m[2 * N + I] = az / (T(1) + factor);
    }
}

// normalise: src/helper.jl:255-271
template <class T>
__global__ void k_normalise(int64_t N, T *__restrict__ m)
{
#pragma clang fp contract(off)
    for (int64_t I = (int64_t)blockIdx.x * BLOCK + threadIdx.x; I < N; I += (int64_t)gridDim.x * BLOCK) {
        const T x = m[I], y = m[N + I], z = m[2 * N + I];
        const T m2 = x * x + y * y + z * z;
        if (m2 > T(0)) {
            const T len = T(1) / sqrt(m2);
            m[I] = x * len;
            m[N + I] = y * len;
            m[2 * N + I] = z * len;
        }
    }
}

// ---------------------------------------------------------------------------
// adjoint helpers (RK4 step pull-back)
// ---------------------------------------------------------------------------
// ut = (dF/dh)^T a / mu0Ms at state Y (zero where Ms = 0 or pinned): the
// input vector of the EPI_ADJ pass.
template <class T>
__global__ void k_adj_pre(const FieldParams<T> *__restrict__ Pp, const T *__restrict__ Y, const T *__restrict__ a,
                          T *__restrict__ ut)
{
    const MMX_CONST FieldParams<T> &P = *(const MMX_CONST FieldParams<T> *)Pp;
    const int64_t N = P.N;
    for (int64_t I = (int64_t)blockIdx.x * BLOCK + threadIdx.x; I < N; I += (int64_t)gridDim.x * BLOCK) {
        const T Ms = P.mu0Ms[I];
        T ux = 0, uy = 0, uz = 0;
        if (Ms > T(0) && !(P.pins && P.pins[I])) {
            llg_vjp_h(P, Y[I], Y[N + I], Y[2 * N + I], a[I], a[N + I], a[2 * N + I], ux, uy, uz);
            ux /= Ms;
            uy /= Ms;
            uz /= Ms;
        }
        ut[I] = ux;
        ut[N + I] = uy;
        ut[2 * N + I] = uz;
    }
}

// Jacobian-vector product of one RHS evaluation: out = (dF/dm) v + (dF/dh) w at
// (Y, hY), where w = hv (the field of v under the terms that are linear in m)
// plus the cubic anisotropy's Jacobian at Y applied to v; zero at pinned cells
// (llg_rhs_kernel! leaves them at rest).  out may alias hv.
template <class T>
__global__ void k_llg_jvp(const FieldParams<T> *__restrict__ Pp, const T *__restrict__ Y, const T *__restrict__ hY,
                          const T *__restrict__ v, const T *hv, T *out)
{
#pragma clang fp contract(off)
    const MMX_CONST FieldParams<T> &P = *(const MMX_CONST FieldParams<T> *)Pp;
    const int64_t N = P.N;
    for (int64_t I = (int64_t)blockIdx.x * BLOCK + threadIdx.x; I < N; I += (int64_t)gridDim.x * BLOCK) {
        const T mx = Y[I], my = Y[N + I], mz = Y[2 * N + I];
        const T vx = v[I], vy = v[N + I], vz = v[2 * N + I];
        T wx = hv[I], wy = hv[N + I], wz = hv[2 * N + I];
        const T Ms = P.mu0Ms[I];
        if (Ms > T(0)) {
            for (int tt = 0; tt < P.nterms; tt++) {
                const MMX_CONST TermDev<T> &D = P.t[tt];
                if (D.kind != K_CUBIC) continue;
                // (12 Kc / mu0Ms) sum_r (Y.a_r)^2 (v.a_r) a_r  (cubic_anisotropy_kernel!, kernels.jl:110-137)
                const T Kc = D.uniform ? D.s[9] : D.a[I];
                const T c = T(12) * Kc / Ms;
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    const T ax = D.s[3 * r], ay = D.s[3 * r + 1], az = D.s[3 * r + 2];
                    const T yp = mx * ax + my * ay + mz * az;
                    const T f = c * yp * yp * (vx * ax + vy * ay + vz * az);
                    wx += f * ax;
                    wy += f * ay;
                    wz += f * az;
                }
            }
        }
        T ox = 0, oy = 0, oz = 0;
        if (!(P.pins && P.pins[I]))
            llg_jvp(P, mx, my, mz, hY[I], hY[N + I], hY[2 * N + I], vx, vy, vz, wx, wy, wz, ox, oy, oz);
        out[I] = ox;
        out[N + I] = oy;
        out[2 * N + I] = oz;
    }
}

// Pull-back of y1 = normalise(y0 + c (((k1 + 2 k2) + 2 k3) + k4)) (rk.jl:46,
// helper.jl:255-271): lam_z = (lam - (lam.y1) y1) / |z|; writes lam0 = lam_z
// and the stage cotangents a1 = a4 = c lam_z, a2 = a3 = 2 c lam_z.
template <class T>
__global__ void k_adj_final(int64_t N, const T *__restrict__ y0, const T *__restrict__ k1, const T *__restrict__ k2,
                            const T *__restrict__ k3, const T *__restrict__ k4, double c, const T *__restrict__ lam,
                            T *__restrict__ lam0, T *__restrict__ a1, T *__restrict__ a2, T *__restrict__ a3,
                            T *__restrict__ a4)
{
#pragma clang fp contract(off)
    for (int64_t I = (int64_t)blockIdx.x * BLOCK + threadIdx.x; I < N; I += (int64_t)gridDim.x * BLOCK) {
        T z[3], l[3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int64_t o = q * N + I;
            T s = k1[o] + T(2) * k2[o] + T(2) * k3[o] + k4[o];
            z[q] = y0[o] + (T)c * s;
            l[q] = lam[o];
        }
        const T m2 = z[0] * z[0] + z[1] * z[1] + z[2] * z[2];
        if (m2 > T(0)) {
            const T len = T(1) / sqrt(m2);
            const T y1[3] = {z[0] * len, z[1] * len, z[2] * len};
            const T ly = l[0] * y1[0] + l[1] * y1[1] + l[2] * y1[2];
#pragma unroll
            for (int q = 0; q < 3; q++) l[q] = (l[q] - ly * y1[q]) * len;
        }
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int64_t o = q * N + I;
            lam0[o] = l[q];
            a1[o] = (T)c * l[q];
            a2[o] = T(2) * (T)c * l[q];
            a3[o] = T(2) * (T)c * l[q];
            a4[o] = (T)c * l[q];
        }
    }
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
static __global__ void k_tensor_pad(DemagGeom G, const double *__restrict__ rt, int comp, double *__restrict__ pad)
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
        const int64_t o_kt = G.ktw ? ((((int64_t)comp * G.Py2 + kyf) * (G.KS >> 2) + (kx >> 2)) * G.Pz2 + kzf) * 4 + (kx & 3)
                                   : (((int64_t)comp * G.Pz2 + kzf) * G.Py2 + kyf) * G.KS + kx;
        Kt[o_kt] = (T)v;
    }
}

static __global__ void k_reduce(const double *__restrict__ partial, int nblocks, int stride, int nval, int is_max,
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

#ifdef MMX_KT_MAIN
int max_local_blocks() { return 8192; }
#endif

#ifdef MMX_KT_MAIN
// Streaming copy of n 16-byte words (the practical HBM ceiling the passes are
// compared with): grid-stride loop, four independent 16-byte loads in flight
// per thread before their stores.
static __global__ void __launch_bounds__(BLOCK) k_stream_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                               int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * BLOCK;
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

void launch_stream_copy(const void *src, void *dst, int64_t n16, int grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(BLOCK), 0, s, (const uint4 *)src, (uint4 *)dst, n16);
}
#endif


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
int launch_local(int64_t N, const FieldParams<T> *P, const Dyn &dyn, const Epi<T> &E, const T *m, int fast,
                 uint32_t sig, hipStream_t s)
{
    int g = grid_for(N, max_local_blocks());
    if (fast && sig && launch_local_sig<T>(N, P, dyn, E, m, sig, g, s) >= 0) return g;
    const int v = variant(E.mode, fast, E.np);
    if (v == 4) hipLaunchKernelGGL((k_local<T, 4>), dim3(g), dim3(BLOCK), 0, s, P, dyn, E, m);
    else if (v == 1) hipLaunchKernelGGL((k_local<T, 1>), dim3(g), dim3(BLOCK), 0, s, P, dyn, E, m);
    else if (v == 2) hipLaunchKernelGGL((k_local<T, 2>), dim3(g), dim3(BLOCK), 0, s, P, dyn, E, m);
    else if (v == 3) hipLaunchKernelGGL((k_local<T, 3>), dim3(g), dim3(BLOCK), 0, s, P, dyn, E, m);
    else if (v == 5) hipLaunchKernelGGL((k_local<T, 5>), dim3(g), dim3(BLOCK), 0, s, P, dyn, E, m);
    else hipLaunchKernelGGL((k_local<T, 0>), dim3(g), dim3(BLOCK), 0, s, P, dyn, E, m);
    return g;
}


template <class T>
void launch_p1(const DemagGeom &G, const T *in, int64_t cstride, const T *scale, int n_in, int Ry, int Rz,
               int ncomp, cplx<T> *A, const cplx<T> *tw, hipStream_t s)
{
    const int nrows = Ry * Rz;
#define P1_CALL(H)                                                                                   \
    {                                                                                                \
        int g = (int)(((int64_t)ncomp * nrows + Shape<H>::SX - 1) / Shape<H>::SX);                   \
        hipLaunchKernelGGL((k_p1<T, H>), dim3(g), dim3(Shape<H>::BX), 0, s, G, in, cstride, scale, n_in, nrows, \
                           ncomp, A, tw);                                                            \
    }
    MMX_LG_SWITCH(G.lgH, P1_CALL)
#undef P1_CALL
}

template <class T>
void launch_coly(const DemagGeom &G, const cplx<T> *in, cplx<T> *out, int n_in, int Rin, int n_out, int Rout,
                 int nplanes, int ncomp, int inverse, const cplx<T> *tw, hipStream_t s)
{
#define CY_CALL(L)                                                                                   \
    {                                                                                                \
        constexpr int TY = ColyShape<T, L>::TY, BY = ColyShape<T, L>::BY;                           \
        dim3 grid((G.Kx + TY - 1) / TY, nplanes, ncomp);                                             \
        if (inverse)                                                                                 \
            hipLaunchKernelGGL((k_coly<T, L, true>), grid, dim3(BY), 0, s, G, in, out, n_in, Rin, n_out,     \
                               Rout, nplanes, tw);                                                   \
        else                                                                                         \
            hipLaunchKernelGGL((k_coly<T, L, false>), grid, dim3(BY), 0, s, G, in, out, n_in, Rin, n_out,    \
                               Rout, nplanes, tw);                                                   \
    }
    MMX_LG_SWITCH(G.lgPy, CY_CALL)
#undef CY_CALL
}

#ifdef MMX_KT_MAIN
int convz_grouped_tensor(const DemagGeom &G, int elsize)
{
    const int e = elsize == 8 ? P3W<double>::E : P3W<float>::E;
    const int ept = G.Pz < e ? G.Pz : e;
    return (MMX_P3_WAVE && G.Pz / ept <= 64) ? 1 : 0;
}
#endif


template <class T>
void launch_convz(const DemagGeom &G, cplx<T> *B, const T *Kt, const T *Lline, int nzb, int n_in, int ncomp,
                  int fwd_only, const cplx<T> *tw, hipStream_t s, int nimg)
{
    (void)ncomp;
#define CZ_CALL(L)                                                                                   \
    {                                                                                                \
        dim3 grid((G.Kx + Shape<L>::TZ - 1) / Shape<L>::TZ, G.Py / 2 + 1, nimg);                     \
        if constexpr (FftShape<L, P3W<T>::E>::TPS <= 64) {                                           \
            if (!fwd_only && G.ktw) {                                                                \
                constexpr int TW = 64 / FftShape<L, P3W<T>::E>::TPS, WX = P3W<T>::template wx<L>();  \
                dim3 gw((G.Kx + WX * TW - 1) / (WX * TW), WX == 2 ? G.Py / 2 + 1 : G.Py, nimg);     \
                hipLaunchKernelGGL((k_convw<T, L>), gw, dim3(256), 0, s, G, B, Kt, Lline, n_in, nzb, tw); \
                break;                                                                               \
            }                                                                                        \
        }                                                                                            \
        if (fwd_only)                                                                                \
            hipLaunchKernelGGL((k_convz<T, L, true>), grid, dim3(Shape<L>::BZ), 0, s, G, B, Kt, Lline, nzb, n_in, tw); \
        else                                                                                         \
            hipLaunchKernelGGL((k_convz<T, L, false>), grid, dim3(Shape<L>::BZ), 0, s, G, B, Kt, Lline, nzb, n_in, tw); \
    }
    MMX_LG_SWITCH(G.lgPz, CZ_CALL)
#undef CZ_CALL
}

// sums of n stacked vectors of N values: partial[q*gridDim.x + b], then out[q]
template <class T>
__global__ void k_sum_images(const T *__restrict__ v, int64_t N, double *__restrict__ partial)
{
    const T *x = v + (int64_t)blockIdx.y * N;
    double r[1] = {0};
    for (int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x; i < N; i += (int64_t)gridDim.x * BLOCK)
        r[0] += (double)x[i];
    block_reduce_store<1>(r, 1, false, partial + (size_t)blockIdx.y * gridDim.x, 1);
}

template <class T>
void launch_sum_images(const T *v, int64_t N, int n, int nb, double *partial, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_sum_images<T>, dim3(nb, n), dim3(BLOCK), 0, s, v, N, partial);
    hipLaunchKernelGGL(k_segment_sum, dim3(n), dim3(BLOCK), 0, s, partial, nb, out);
}



#ifdef MMX_KT_MAIN
void launch_tensor_pad(const DemagGeom &G, const double *rt, int comp, double *pad, hipStream_t s)
{
    int64_t P = (int64_t)G.Px * G.Py * G.Pz;
    hipLaunchKernelGGL(k_tensor_pad, dim3(grid_for(P)), dim3(BLOCK), 0, s, G, rt, comp, pad);
}
#endif


template <class T>
void launch_tensor_extract(const DemagGeom &G, const cplx<double> *Bt, int comp, double scale, T *Kt,
                           hipStream_t s)
{
    int64_t n = (int64_t)G.Pz2 * G.Py2 * G.Kx;
    hipLaunchKernelGGL(k_tensor_extract<T>, dim3(grid_for(n)), dim3(BLOCK), 0, s, G, Bt, comp, scale, Kt);
}

#ifdef MMX_KT_MAIN
void launch_reduce(const double *partial, int nblocks, int stride, int nval, int is_max, double *out,
                   hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(BLOCK), 0, s, partial, nblocks, stride, nval, is_max, out);
}
#endif


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

template <class T>
void launch_adj_pre(int64_t N, const FieldParams<T> *P, const T *Y, const T *a, T *ut, hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_pre<T>, dim3(grid_for(N)), dim3(BLOCK), 0, s, P, Y, a, ut);
}

template <class T>
void launch_llg_jvp(int64_t N, const FieldParams<T> *P, const T *Y, const T *hY, const T *v, const T *hv, T *out,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_llg_jvp<T>, dim3(grid_for(N)), dim3(BLOCK), 0, s, P, Y, hY, v, hv, out);
}

template <class T>
void launch_adj_final(int64_t N, const T *y0, const T *const k[4], double c, const T *lam, T *lam0, T *const a[4],
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_adj_final<T>, dim3(grid_for(N)), dim3(BLOCK), 0, s, N, y0, k[0], k[1], k[2], k[3], c, lam,
                       lam0, a[0], a[1], a[2], a[3]);
}

template <class T>
void launch_batch_local(const int2 *items, int nitems, const FieldParams<T> *const *Ps, const BatchStage<T> *st,
                        const Dyn *dyns, int V, int cpt, int bs, int grid, hipStream_t s)
{
#define BL_CALL(VV, CC, BB) \
    hipLaunchKernelGGL((k_batch_local<T, VV, CC, BB>), dim3(grid), dim3(BB), 0, s, items, nitems, Ps, st, dyns)
#define BL_CPT(VV, BB)                      \
    if (cpt == 1) BL_CALL(VV, 1, BB);       \
    else if (cpt == 2) BL_CALL(VV, 2, BB);  \
    else BL_CALL(VV, 4, BB);
#define BL_BS(VV)                                  \
    if (bs == 128) { BL_CPT(VV, 128) }             \
    else if (bs == 512) { BL_CPT(VV, 512) }        \
    else { BL_CPT(VV, 256) }
    if (V == 4) { BL_BS(4) }
    else if (V == 1) { BL_BS(1) }
    else if (V == 2) { BL_BS(2) }
    else { BL_CPT(0, 256) }  // the term-list variant: 256 threads only
#undef BL_BS
#undef BL_CPT
#undef BL_CALL
}

#ifdef MMX_KT_MAIN
void launch_batch_tick(Dyn *dyns, int M, hipStream_t s)
{
    hipLaunchKernelGGL(k_batch_tick, dim3((M + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, dyns, M);
}
#endif


#ifdef MMX_KT_MAIN
int batch_variant(int mode, int fast, int np) { return variant(mode, fast, np); }
#endif


template <class T>
int launch_sum(const T *v, int64_t n, double *partial, hipStream_t s)
{
    int g = grid_for(n, max_local_blocks());
    hipLaunchKernelGGL(k_sum<T>, dim3(g), dim3(BLOCK), 0, s, v, n, partial);
    return g;
}

template <class T>
void launch_neb_dist(int64_t N, int n, const T *Y, const T *L, const T *R, int nb, double *partial, double *dist,
                     hipStream_t s)
{
    hipLaunchKernelGGL(k_neb_dist<T>, dim3(nb, n + 1), dim3(BLOCK), 0, s, N, n, Y, L, R, partial);
    hipLaunchKernelGGL(k_neb_dist_final, dim3(n + 1), dim3(BLOCK), 0, s, partial, nb, dist);
}

template <class T>
void launch_neb_field(const FieldParams<T> *P, const Dyn &dyn, uint32_t mask, int64_t N, int n, int nb, const T *Y,
                      T *F, double *partial, double *energy, hipStream_t s)
{
    hipLaunchKernelGGL(k_neb_field<T>, dim3(nb, n), dim3(BLOCK), 0, s, P, dyn, mask, Y, F, partial);
    hipLaunchKernelGGL(k_segment_sum, dim3(n), dim3(BLOCK), 0, s, partial, nb, energy);
}

template <class T>
void launch_neb_maxdiff(const T *a, const T *b, int64_t N, int n, int nb, double *partial, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_neb_maxdiff<T>, dim3(nb, n), dim3(BLOCK), 0, s, a, b, N, partial);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(BLOCK), 0, s, partial, nb * n, 1, 1, 1, out);
}

template <class T>
void launch_neb_tangent(int64_t N, int n, const T *Ys, const T *L, const T *R, const double *energy, const T *F,
                        T *tan, int nb, double *partial, const double *dist, double k, int climb, double *coef,
                        hipStream_t s)
{
    hipLaunchKernelGGL(k_neb_tangent<T>, dim3(nb, n), dim3(BLOCK), 0, s, N, n, Ys, L, R, energy, F, tan, partial);
    hipLaunchKernelGGL(k_neb_coef, dim3(n), dim3(BLOCK), 0, s, partial, nb, dist, k, climb, coef);
}

template <class T>
void launch_neb_stage(const FieldParams<T> *P, int64_t N, int n, int nb, const T *Ys, const T *F, const T *tan,
                      const double *coef, int mode, const double c[4], double outer, T *kstore, const T *y0,
                      const T *k1, const T *k2, const T *k3, T *yout, double *partial, hipStream_t s)
{
    NebStage S{};
    S.mode = mode;
    for (int i = 0; i < 4; i++) S.c[i] = c[i];
    S.outer = outer;
    hipLaunchKernelGGL(k_neb_stage<T>, dim3(nb, n), dim3(BLOCK), 0, s, P, N, Ys, F, tan, coef, S, kstore, y0, k1,
                       k2, k3, yout, partial);
}

template <class T>
void launch_nbmask(int64_t N, const FieldParams<T> *P, uint8_t *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_nbmask<T>, dim3(grid_for(N)), dim3(BLOCK), 0, s, P, out);
}

template <class T>
int launch_sd_g(int64_t N, const T *m, const T *mpre, const T *h, T *g, double *partial, hipStream_t s)
{
    int gr = grid_for(N, max_local_blocks());
    hipLaunchKernelGGL(k_sd_g<T>, dim3(gr), dim3(BLOCK), 0, s, N, m, mpre, h, g, partial);
    return gr;
}

#ifdef MMX_KT_MAIN
void launch_sd_tau(const double *sums, int odd, double min_tau, double max_tau, double *tau, hipStream_t s)
{
    hipLaunchKernelGGL(k_sd_tau, dim3(1), dim3(1), 0, s, sums, odd, min_tau, max_tau, tau);
}
#endif


template <class T>
void launch_sd_update(int64_t N, const FieldParams<T> *P, const T *g, T *m, T *mpre, const T *h, const double *tau,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_sd_update<T>, dim3(grid_for(N)), dim3(BLOCK), 0, s, P, g, m, mpre, h, tau);
}

template <class T>
void launch_normalise(int64_t N, T *m, hipStream_t s)
{
    hipLaunchKernelGGL(k_normalise<T>, dim3(grid_for(N)), dim3(BLOCK), 0, s, N, m);
}

#define MMX_INST(T)                                                                                          \
    template void launch_aos_to_soa<T>(const double *, T *, int64_t, hipStream_t);                           \
    template void launch_soa_to_aos<T>(const T *, double *, int64_t, hipStream_t);                           \
    template void launch_cast<T>(const double *, T *, int64_t, hipStream_t);                                 \
    template void launch_to_double<T>(const T *, double *, int64_t, hipStream_t);                            \
    template int launch_local<T>(int64_t, const FieldParams<T> *, const Dyn &, const Epi<T> &, const T *, int,  \
                                 uint32_t, hipStream_t);                                                               \
    template void launch_p1<T>(const DemagGeom &, const T *, int64_t, const T *, int, int, int, int, cplx<T> *, \
                               const cplx<T> *, hipStream_t);                                                \
    template void launch_coly<T>(const DemagGeom &, const cplx<T> *, cplx<T> *, int, int, int, int, int, int,  \
                                 int, const cplx<T> *, hipStream_t);                                         \
    template void launch_convz<T>(const DemagGeom &, cplx<T> *, const T *, const T *, int, int, int, int,      \
                                  const cplx<T> *, hipStream_t, int);                                                         \
    template void launch_sum_images<T>(const T *, int64_t, int, int, double *, double *, hipStream_t);        \
    template void launch_tensor_extract<T>(const DemagGeom &, const cplx<double> *, int, double, T *,         \
                                           hipStream_t);                                                     \
    template int launch_maxdiff<T>(const T *, const T *, int64_t, double *, hipStream_t);          \
    template void launch_axpy<T>(const T *, double, const T *, T *, int64_t, hipStream_t);        \
    template int launch_maxabs<T>(const T *, int64_t, double *, hipStream_t);                    \
    template void launch_batch_local<T>(const int2 *, int, const FieldParams<T> *const *, const BatchStage<T> *, \
                                        const Dyn *, int, int, int, int, hipStream_t);                              \
    template int launch_sd_g<T>(int64_t, const T *, const T *, const T *, T *, double *, hipStream_t); \
    template void launch_sd_update<T>(int64_t, const FieldParams<T> *, const T *, T *, T *, const T *,        \
                                      const double *, hipStream_t);                                         \
    template void launch_normalise<T>(int64_t, T *, hipStream_t);                                         \
    template void launch_nbmask<T>(int64_t, const FieldParams<T> *, uint8_t *, hipStream_t);      \
    template int launch_sum<T>(const T *, int64_t, double *, hipStream_t);                          \
    template void launch_neb_dist<T>(int64_t, int, const T *, const T *, const T *, int, double *, double *,  \
                                     hipStream_t);                                                          \
    template void launch_neb_field<T>(const FieldParams<T> *, const Dyn &, uint32_t, int64_t, int, int, const T *, \
                                      T *, double *, double *, hipStream_t);                                \
    template void launch_neb_maxdiff<T>(const T *, const T *, int64_t, int, int, double *, double *, hipStream_t); \
    template void launch_neb_tangent<T>(int64_t, int, const T *, const T *, const T *, const double *,        \
                                        const T *, T *, int, double *, const double *, double, int, double *, \
                                        hipStream_t);                                                       \
    template void launch_neb_stage<T>(const FieldParams<T> *, int64_t, int, int, const T *, const T *,        \
                                      const T *, const double *, int, const double[4], double, T *,          \
                                      const T *, const T *, const T *, const T *, T *, double *, hipStream_t); \
    template void launch_adj_pre<T>(int64_t, const FieldParams<T> *, const T *, const T *, T *, hipStream_t); \
    template void launch_llg_jvp<T>(int64_t, const FieldParams<T> *, const T *, const T *, const T *, const T *,   \
                                    T *, hipStream_t);                                                          \
    template void launch_adj_final<T>(int64_t, const T *, const T *const[4], double, const T *, T *,          \
                                      T *const[4], hipStream_t);

#ifndef MMX_KT
#error "compile once per precision with -DMMX_KT=double or -DMMX_KT=float"
#endif
MMX_INST(MMX_KT)

}  // namespace mmx
