// mmx_fft.h -- register-resident Stockham FFT for power-of-two lengths.
//
// One sequence of length L is spread over TPS(L) = L / EPT(L) threads; thread
// j in [0, TPS) holds EPT(L) = min(L, E) elements (E = 8 or 4) v[m] = x[j + m*TPS] on entry
// and X[j + m*TPS] on exit, so a column can be loaded from and stored to
// global memory in the same pattern.  Stages are radix E (a final smaller
// radix when log2 L is not a multiple of log2 E); values move between stages through an
// LDS exchange buffer of L elements per sequence.  Twiddles come from an FP64-
// accurate table tw[t] = exp(-2 pi i t / 2^lgLmax) indexed at t << (lgLmax - lgL):
// each butterfly loads W^k (all stages' loads issue before the first stage)
// and forms W^(2k) .. W^((R-1)k) by multiplication.
#pragma once
#include "mmx_common.h"

namespace mmx {

// Exchange-buffer index map: element i of a sequence sits at i + (i >> SW)
// (one gap every 2^SW elements), which spreads the strided writes of the
// early stages over the LDS banks; SW = 0 disables the gaps.  SW8 applies to
// the 8-element-per-thread transforms, SW4 to the 4-element ones.
#ifndef MMX_FFT_SW8
#define MMX_FFT_SW8 0
#endif
#ifndef MMX_FFT_SW4
#define MMX_FFT_SW4 0
#endif

template <int L, int E = 8>
struct FftShape {
    static constexpr int EPT = L >= E ? E : L;
    static constexpr int TPS = L / EPT;
    static constexpr int LG = L <= 1 ? 0 : (L <= 2 ? 1 : (L <= 4 ? 2 : (L <= 8 ? 3 : (L <= 16 ? 4 : (L <= 32 ? 5 : (L <= 64 ? 6 : (L <= 128 ? 7 : (L <= 256 ? 8 : (L <= 512 ? 9 : (L <= 1024 ? 10 : (L <= 2048 ? 11 : 12)))))))))));
    static constexpr int SW = E >= 8 ? MMX_FFT_SW8 : MMX_FFT_SW4;
    // elements one sequence's exchange buffer spans
    static constexpr int XS = SW ? L + (L >> SW) : L;
};

template <int SW>
__device__ __forceinline__ int xsw(int i)
{
    if constexpr (SW > 0) return i + (i >> SW);
    else return i;
}

template <class T>
__device__ __forceinline__ cplx<T> mul_mi(cplx<T> a, bool inv)  // a * (-i) forward, a * (+i) inverse
{
    return inv ? cplx<T>{-a.y, a.x} : cplx<T>{a.y, -a.x};
}

template <class T, bool INV>
__device__ __forceinline__ void dft2(cplx<T> *a)
{
    cplx<T> t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
}

template <class T, bool INV>
__device__ __forceinline__ void dft4(cplx<T> *a)
{
    cplx<T> b0 = cadd(a[0], a[2]), b1 = csub(a[0], a[2]);
    cplx<T> b2 = cadd(a[1], a[3]), b3 = mul_mi(csub(a[1], a[3]), INV);
    a[0] = cadd(b0, b2);
    a[2] = csub(b0, b2);
    a[1] = cadd(b1, b3);
    a[3] = csub(b1, b3);
}

// 8-point DFT as two 4-point DFTs (even/odd inputs) and the W8 combination
template <class T, bool INV>
__device__ __forceinline__ void dft8(cplx<T> *a)
{
    const T r = T(0.70710678118654752440084436210485);
    cplx<T> b0 = cadd(a[0], a[4]), b1 = csub(a[0], a[4]);
    cplx<T> b2 = cadd(a[2], a[6]), b3 = mul_mi(csub(a[2], a[6]), INV);
    cplx<T> b4 = cadd(a[1], a[5]), b5 = csub(a[1], a[5]);
    cplx<T> b6 = cadd(a[3], a[7]), b7 = mul_mi(csub(a[3], a[7]), INV);
    cplx<T> c0 = cadd(b0, b2), c2 = csub(b0, b2), c1 = cadd(b1, b3), c3 = csub(b1, b3);
    cplx<T> c4 = cadd(b4, b6), c6 = csub(b4, b6), c5 = cadd(b5, b7), c7 = csub(b5, b7);
    // W8 c5, -i c6, W8^3 c7 (forward); conjugate roots for the inverse
    cplx<T> d5, d7;
    if (!INV) {
        d5 = {(c5.x + c5.y) * r, (c5.y - c5.x) * r};
        d7 = {(c7.y - c7.x) * r, -(c7.x + c7.y) * r};
    } else {
        d5 = {(c5.x - c5.y) * r, (c5.y + c5.x) * r};
        d7 = {-(c7.x + c7.y) * r, (c7.x - c7.y) * r};
    }
    cplx<T> d6 = mul_mi(c6, INV);
    a[0] = cadd(c0, c4);
    a[4] = csub(c0, c4);
    a[1] = cadd(c1, d5);
    a[5] = csub(c1, d5);
    a[2] = cadd(c2, d6);
    a[6] = csub(c2, d6);
    a[3] = cadd(c3, d7);
    a[7] = csub(c3, d7);
}

template <class T, int R, bool INV>
__device__ __forceinline__ void dftR(cplx<T> *a)
{
    if constexpr (R == 8) dft8<T, INV>(a);
    else if constexpr (R == 4) dft4<T, INV>(a);
    else if constexpr (R == 2) dft2<T, INV>(a);
}

// Orders the LDS accesses of one wave before and after this point.  A wave's
// LDS instructions execute in issue order, so when every thread of a sequence
// sits in the same wave the exchange needs no block barrier, only this
// compiler-level fence (no instruction is emitted for wavefront scope).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool WAVE>
__device__ __forceinline__ void exchange_sync()
{
    if constexpr (WAVE) wave_lds_sync();
    else __syncthreads();
}

// Twiddle bases of every stage of a length-L transform for thread j, loaded
// together before the first stage so that their latency overlaps it instead
// of stalling each later stage.  The stage at span Ns keeps W^(kk*L/(Ns*Rs))
// (kk = jb mod Ns) for each of its NQ butterflies; the other twiddles of a
// butterfly are products of its powers (w^r = w^(r/2) w^(r - r/2), at most
// 3 roundings deep), so a stage costs one table load per butterfly.
template <class T, int L, int E, int Ns = 1>
struct TwSet {
    static constexpr int EPT = FftShape<L, E>::EPT, TPS = FftShape<L, E>::TPS;
    static constexpr int Rs = (L / Ns) >= EPT ? EPT : L / Ns;
    static constexpr int NQ = EPT / Rs;
    static constexpr int TWSTEP = L / (Ns * Rs);
    cplx<T> w[NQ];
    TwSet<T, L, E, Ns * Rs> next;
    __device__ __forceinline__ void load(int j, const cplx<T> *__restrict__ tw, int twsh)
    {
        if constexpr (Ns > 1) {
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const int kk = (j + q * TPS) & (Ns - 1);
                gld2(reinterpret_cast<const T *>(tw), (int64_t)((kk * TWSTEP) << twsh), w[q].x, w[q].y);
            }
        }
        next.load(j, tw, twsh);
    }
};
template <class T, int L, int E>
struct TwSet<T, L, E, L> {
    __device__ __forceinline__ void load(int, const cplx<T> *__restrict__, int) {}
};

// One Stockham stage of radix Rs at span Ns over NC sequences held by the same
// threads (v[c*EPT + m]; sequence c exchanges through lds + c*cs).  Butterfly
// jb = j + q*TPS (q < EPT/Rs) reads v[q + r*EPT/Rs] = x[jb + r*L/Rs], applies
// W_{Ns Rs}^{r (jb mod Ns)}, and writes y[(jb/Ns) Ns Rs + jb mod Ns + r Ns].  The
// last stage (Ns Rs = L) leaves its outputs in v at the input positions.  The
// NC sequences share each exchange's barriers and twiddles.  WAVE: all threads
// of each sequence are lanes of one wave (exchange without block barriers).
template <class T, int L, int E, int Rs, int Ns, bool INV, int NC, bool WAVE = false>
__device__ __forceinline__ void fft_stage(cplx<T> *v, int j, cplx<T> *lds, int cs, const cplx<T> *wb)
{
    constexpr int EPT = FftShape<L, E>::EPT, TPS = FftShape<L, E>::TPS;
    constexpr int NQ = EPT / Rs;
    constexpr bool LAST = (Ns * Rs == L);
    cplx<T> o[LAST ? NC * EPT : 1];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int jb = j + q * TPS;
        const int kk = jb & (Ns - 1);
        cplx<T> w[Rs];
        if constexpr (Ns > 1) {
            cplx<T> w1 = wb[q];
            if (INV) w1.y = -w1.y;
            w[1] = w1;
#pragma unroll
            for (int r = 2; r < Rs; r++) w[r] = cmul(w[r >> 1], w[r - (r >> 1)]);
        }
#pragma unroll
        for (int c = 0; c < NC; c++) {
            cplx<T> a[Rs];
#pragma unroll
            for (int r = 0; r < Rs; r++) a[r] = v[c * EPT + q + r * NQ];
            if constexpr (Ns > 1) {
#pragma unroll
                for (int r = 1; r < Rs; r++) a[r] = cmul(a[r], w[r]);
            }
            dftR<T, Rs, INV>(a);
            if constexpr (LAST) {
#pragma unroll
                for (int r = 0; r < Rs; r++) o[c * EPT + q + r * NQ] = a[r];
            } else {
                const int base = (jb / Ns) * Ns * Rs + kk;
#pragma unroll
                for (int r = 0; r < Rs; r++) lds[c * cs + xsw<FftShape<L, E>::SW>(base + r * Ns)] = a[r];
            }
        }
    }
    if constexpr (LAST) {
#pragma unroll
        for (int m = 0; m < NC * EPT; m++) v[m] = o[m];
    } else {
        exchange_sync<WAVE>();
#pragma unroll
        for (int c = 0; c < NC; c++)
#pragma unroll
            for (int m = 0; m < EPT; m++) v[c * EPT + m] = lds[c * cs + xsw<FftShape<L, E>::SW>(j + m * TPS)];
        exchange_sync<WAVE>();
    }
}

template <class T, int L, int E, int Ns, bool INV, int NC, bool WAVE = false>
__device__ __forceinline__ void fft_stages(cplx<T> *v, int j, cplx<T> *lds, int cs, const TwSet<T, L, E, Ns> &ts)
{
    if constexpr (Ns < L) {
        constexpr int Rs = TwSet<T, L, E, Ns>::Rs;
        fft_stage<T, L, E, Rs, Ns, INV, NC, WAVE>(v, j, lds, cs, ts.w);
        fft_stages<T, L, E, Ns * Rs, INV, NC, WAVE>(v, j, lds, cs, ts.next);
    }
}

// Twiddle bases of thread j for a length-L transform (shared by the forward
// and the inverse transform of the same length).
template <class T, int L, int E = 8>
__device__ __forceinline__ TwSet<T, L, E> fft_twiddles(int j, const cplx<T> *__restrict__ tw, int lgLmax)
{
    TwSet<T, L, E> ts;
    ts.load(j, tw, lgLmax - FftShape<L, E>::LG);
    return ts;
}

// Transform of one sequence.  Threads of the same block must all call this
// together (the exchanges use __syncthreads); `lds` is this sequence's buffer
// of L elements.  `active` threads hold data; inactive ones only join barriers.
template <class T, int L, bool INV, int E = 8, bool WAVE = false>
__device__ __forceinline__ void fft_reg(cplx<T> *v, int j, cplx<T> *lds, const cplx<T> *__restrict__ tw, int lgLmax)
{
    if constexpr (L > 1) fft_stages<T, L, E, 1, INV, 1, WAVE>(v, j, lds, 0, fft_twiddles<T, L, E>(j, tw, lgLmax));
}

// NC sequences per thread (v[c*EPT + m]), exchanging through lds + c*cs.
template <class T, int L, bool INV, int E, int NC>
__device__ __forceinline__ void fft_reg_multi(cplx<T> *v, int j, cplx<T> *lds, int cs, const cplx<T> *__restrict__ tw,
                                              int lgLmax)
{
    if constexpr (L > 1) fft_stages<T, L, E, 1, INV, NC>(v, j, lds, cs, fft_twiddles<T, L, E>(j, tw, lgLmax));
}

// NC sequences per thread whose threads all sit in one wave: the exchanges
// synchronise the wave only.  ts: the twiddle bases from fft_twiddles.
template <class T, int L, bool INV, int E, int NC>
__device__ __forceinline__ void fft_wave_multi(cplx<T> *v, int j, cplx<T> *lds, int cs, const TwSet<T, L, E> &ts)
{
    static_assert(FftShape<L, E>::TPS <= 64, "a sequence must fit in one wave");
    if constexpr (L > 1) fft_stages<T, L, E, 1, INV, NC, true>(v, j, lds, cs, ts);
}

}  // namespace mmx
