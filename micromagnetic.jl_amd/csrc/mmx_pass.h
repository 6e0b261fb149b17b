// mmx_pass.h -- the tile shapes of the demag passes and the kernels that end an
// RHS evaluation (the last demag pass k_p5 and the local-field kernel k_local),
// shared by the translation units that instantiate them.
#pragma once
#include "mmx_common.h"
#include "mmx_device.h"
#include "mmx_fft.h"

namespace mmx {

// z pass: elements per thread of the length-Pz transform, and the number of
// threads per block the column tile aims at
#ifndef MMX_P3_E
#define MMX_P3_E 4
#endif
#ifndef MMX_P3_THREADS
#define MMX_P3_THREADS 256
#endif

template <int L>
struct Shape : FftShape<L> {
    using FftShape<L>::TPS;
    // y columns: TY consecutive kx columns per block (>= 8 for whole cache lines)
    static constexpr int TY = (256 / TPS) >= 8 ? (256 / TPS) : ((1024 / TPS) < 8 ? (1024 / TPS) : 8);
    static constexpr int BY = TY * TPS;
    // z columns: TZ columns at ky and TZ at Py-ky (which share the folded
    // tensor), all three components in one thread, 4 elements each
    static constexpr int TPSZ = FftShape<L, MMX_P3_E>::TPS;
    static constexpr int TZ0 = MMX_P3_THREADS / 2 / TPSZ;
    static constexpr int TZ = TZ0 > 64 ? 64 : (TZ0 < 1 ? 1 : TZ0);
    static constexpr int BZ = 2 * TZ * TPSZ;
    // x rows: SX sequences per block
    static constexpr int SX = (256 / TPS) < 1 ? 1 : (256 / TPS);
    static constexpr int BX = SX * TPS;
    // last pass: RX rows x 3 components per block; 3*RX*TPS is a whole number
    // of 64-wide waves (the block reductions shuffle across full waves)
    static constexpr int GX = TPS >= 64 ? 1 : 64 / TPS;
    static constexpr int RX = (256 / (3 * TPS)) / GX * GX < GX ? GX : (256 / (3 * TPS)) / GX * GX;
    static constexpr int B5 = 3 * RX * TPS;
    static_assert(B5 % 64 == 0, "last-pass blocks must be whole waves");
};

// V: 1 EPI_STAGE, 2 EPI_RK4_FINAL, 3 EPI_ERR, 4 EPI_STAGE without earlier stages
// in the combination (np == 0: RK4 stages, Heun's first)
template <int V>
struct VMode {
    static constexpr int mode = (V == 1 || V == 4) ? EPI_STAGE : (V == 2 ? EPI_RK4_FINAL : EPI_ERR);
    static constexpr bool nokp = V == 4;
};

// V: 0 the term list with the modes of MM, 5 the adjoint (term list), 1-4 the
// slot-form integrator stages (VMode)
// TILE (slot form only): m of the cell's row and of the block's other rows
// comes from the LDS row tile tl.
template <class T, int V, uint32_t SIG = 0, uint32_t MM = MM_GENERIC, bool TILE = false>
__device__ __forceinline__ void epilogue(const MMX_CONST FieldParams<T> &P, const Dyn &dyn, const Epi<T> &E, int64_t I,
                                         int i, int j, int k, const T *__restrict__ m, T hx, T hy, T hz, double &red,
                                         T *ynext = nullptr, const RowTile<T> &tl = RowTile<T>{})
{
    static_assert(!TILE || (V >= 1 && V <= 4), "the row tile serves the slot-form epilogues");
    if constexpr (V == 0) finish_cell<T, MM>(P, dyn, E, I, i, j, k, m, hx, hy, hz, red);
    else if constexpr (V == 5) finish_cell<T, MM_ADJ>(P, dyn, E, I, i, j, k, m, hx, hy, hz, red);
    else
        finish_fast<T, VMode<V>::mode, VMode<V>::nokp, SIG, TILE>(P, dyn, E, (int)I, i, j, k, m, hx, hy, hz, red,
                                                                  ynext, tl);
}

// Split-step twiddles W_Px^(j + q*TPS), q < EPT, of thread j: one per-lane
// table load of W^j times the wave-uniform W^(q*TPS) (scalar loads), so the
// split step costs one vector load per thread instead of EPT.
template <class T, int EPT, int TPS>
__device__ __forceinline__ void split_twiddles(cplx<T> *w, int j, const cplx<T> *__restrict__ tw, int sh)
{
    cplx<T> wj;
    gld2(reinterpret_cast<const T *>(tw), (int64_t)(j << sh), wj.x, wj.y);
    w[0] = wj;
#pragma unroll
    for (int q = 1; q < EPT; q++) w[q] = cmul(wj, tw[(q * TPS) << sh]);
}

// Real-to-complex x transform tail shared by the first and the last pass: v
// holds this thread's outputs Z[j + q*TPS] of the half-length complex FFT of
// one row packed as z[n] = (x[2n], x[2n+1]); buf is that sequence's exchange
// buffer (all of its threads arrive here together).  Writes the Px/2+1
// spectrum values X[k] = Fe + W_Px^k Fo (X[H] = Fe(0) - Fo(0)) to dst.
template <class T, int H, bool WAVE = false>
__device__ __forceinline__ void r2c_split_store(const cplx<T> *v, int j, cplx<T> *buf, bool active,
                                                cplx<T> *__restrict__ dst, const cplx<T> *__restrict__ tw, int lgLmax)
{
    constexpr int EPT = FftShape<H>::EPT, TPS = FftShape<H>::TPS;
#pragma unroll
    for (int m = 0; m < EPT; m++) buf[j + m * TPS] = v[m];
    exchange_sync<WAVE>();
    if (!active) return;
    // the twiddle pointer passes through an empty asm statement after the
    // barrier so that these loads are not hoisted over the transform (where
    // their registers would add to its peak)
    asm volatile("" : "+s"(tw));
    const int sh = lgLmax - (FftShape<H>::LG + 1);  // W_Px^k = tw[k << sh]
    cplx<T> ws[EPT];
    split_twiddles<T, EPT, TPS>(ws, j, tw, sh);
#pragma unroll
    for (int m = 0; m < EPT; m++) {
        const int k = j + m * TPS;
        cplx<T> Zk = v[m];
        cplx<T> Zc = buf[(H - k) & (H - 1)];
        Zc.y = -Zc.y;
        cplx<T> Fe = {(Zk.x + Zc.x) * T(0.5), (Zk.y + Zc.y) * T(0.5)};
        cplx<T> d = csub(Zk, Zc);
        cplx<T> Fo = {d.y * T(0.5), -d.x * T(0.5)};
        dst[k] = cadd(Fe, cmul(ws[m], Fo));
        if (k == 0) dst[H] = csub(Fe, Fo);  // W_Px^H = -1
    }
}

// Row transforms (first and last pass): thread j of a sequence is lane
// (tid % TPS), so a sequence of TPS <= 64 threads lies inside one wave and
// its exchanges need only a wave-level fence, not a block barrier.
#ifndef MMX_ROWWAVE
#define MMX_ROWWAVE 1
#endif
#ifndef MMX_P5_MIRROR_LDS
#define MMX_P5_MIRROR_LDS 1
#endif
template <int H>
struct RowWave {
    static constexpr bool on = MMX_ROWWAVE && FftShape<H>::TPS <= 64;
};

// Slot-form last pass: the threads that transformed a row also stage m of
// that row in the unused upper half of the row's LDS buffer, so the cell loop
// reads m of the cell, of its x neighbours and of its in-block z neighbours
// from LDS (one 16-byte load per two cells instead of up to 12 loads per cell)
#ifndef MMX_P5_TILE
#define MMX_P5_TILE 1
#endif
// 1: the tile's loads issue with the spectrum's, before the row transform
#ifndef MMX_P5_TILE_EARLY
#define MMX_P5_TILE_EARLY 0
#endif
// at least MMX_P5_WPE waves per SIMD for the last pass: its cell loop is
// latency-bound, and the register budget this sets costs no spills
#ifndef MMX_P5_WPE
#define MMX_P5_WPE 3
#endif
// launches of the last pass with fewer blocks take fewer rows per block (p5_rows)
#ifndef MMX_P5_MIN_BLOCKS
#define MMX_P5_MIN_BLOCKS 512
#endif
// the term-list variants (V = 0, 5) hold more live values: 2 waves per SIMD
// leave them their registers without scratch spills
#define MMX_P5_ATTR __attribute__((amdgpu_waves_per_eu((V == 0 || V == 5) ? 2 : MMX_P5_WPE)))
// IMG: the launch covers a stack of images (grid y = image), each with its own
// spectra, spin and field planes; fields and per-cell energies only.
// The last pass of one system (one block's rows); the kernels below supply
// its operands: the launch's own arguments (k_p5) or a batch member's entries
// of the batch tables (k_p5_batch).
template <class T, int H, int V, uint32_t SIG, bool IMG>
__device__ __forceinline__ void p5_body(const DemagGeom &G, const cplx<T> *__restrict__ A,
                                        const FieldParams<T> *__restrict__ Pp, const Dyn &dyn, const Epi<T> &E,
                                        const T *__restrict__ m, const cplx<T> *__restrict__ tw)
{
    using S = Shape<H>;
    constexpr int EPT = S::EPT, TPS = S::TPS, RX = S::RX;
    constexpr int HS = FftShape<H>::XS + 1;
    __shared__ cplx<T> lds[3 * RX * HS];
    const int nx = G.nx, ny = G.ny, nz = G.nz;
    const int nrows_tot = ny * nz;
    // A block takes rx <= RX rows (x lines) at one y and consecutive z (rows
    // r >= rx of its layout stay idle in the transforms).  Groups are
    // ordered y-major, z-minor, and each XCD (launch indices go round-robin
    // over the 8 XCDs) gets one contiguous run of groups: the z+-1 rows a cell
    // reads were just fetched by this or the previous block, the y+-1 rows a
    // few blocks earlier, so both come from that XCD's L2.
    const int rx = G.p5rx;
    const int gz = (nz + rx - 1) / rx;
    const int ngroups = ny * gz;
    const int per = (ngroups + 7) / 8;
    const int lb = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (lb >= ngroups) {
        if (threadIdx.x == 0 && (E.mode == EPI_ERR || E.mode == EPI_ENERGY)) E.partial[blockIdx.x] = 0.0;
        return;
    }
    const int y0 = lb / gz, z0 = (lb % gz) * rx;
    const int nrows = min(rx, nz - z0);
    const int ls = threadIdx.x / TPS, j = threadIdx.x % TPS;
    const int c = ls / RX, r = ls % RX;
    const bool active = r < nrows;
    const cplx<T> *row = A + ((int64_t)c * nrows_tot + (int64_t)(z0 + (active ? r : 0)) * ny + y0) * G.KS;
    const int sh = G.lgLmax - (FftShape<H>::LG + 1);
    cplx<T> v[EPT];
    cplx<T> *buf = lds + (c * RX + r) * HS;
    // with wave-local rows the mirror values X[H-k] come from LDS: each
    // spectrum value is read from HBM once
    constexpr bool MIR = RowWave<H>::on && MMX_P5_MIRROR_LDS;
    // Row buffer (c, r) as 2*HS reals: the field of cell x at x (x < nx) and,
    // with the tile, m of the row at mo + x (mo + nx <= 2H + 2 since nx <= H).
    // The row's threads write the tile after they end the row's transform,
    // so no barrier is needed before they overwrite its exchange buffer.
    // Pairs past the row are clamped to its last pair (stored again, same
    // value); the loop stops at a uniform bound, so no load waits on a branch.
    constexpr bool TL = MMX_P5_TILE && V >= 1 && V <= 4 && RowWave<H>::on;
    const int mo = (nx + 1) & ~1;
    const T *mrow = m + (int64_t)c * nx * ny * nz + ((int64_t)(z0 + (active ? r : 0)) * ny + y0) * nx;
    const bool tpair = (nx & 1) == 0 && ((uintptr_t)mrow & (2 * sizeof(T) - 1)) == 0;
    cplx<T> tv[TL ? EPT : 1];
    auto tile_load = [&]() {
        const int nh = (nx + 1) >> 1;
#pragma unroll
        for (int q = 0; q < EPT; q++) {
            if (q * TPS >= nh) break;
            const int k = min(j + q * TPS, nh - 1);
            if (tpair) {
                gld2(mrow, k, tv[q].x, tv[q].y);
            } else {
                tv[q].x = gld(mrow, 2 * k);
                tv[q].y = gld(mrow, min(2 * k + 1, nx - 1));
            }
        }
    };
    auto tile_store = [&]() {
        if (!active) return;
        const int nh = (nx + 1) >> 1;
        T *dst = reinterpret_cast<T *>(buf) + mo;
#pragma unroll
        for (int q = 0; q < EPT; q++) {
            if (q * TPS >= nh) break;
            const int k = min(j + q * TPS, nh - 1);
            dst[2 * k] = tv[q].x;
            if (2 * k + 1 < nx) dst[2 * k + 1] = tv[q].y;
        }
    };
    if constexpr (TL && MMX_P5_TILE_EARLY) tile_load();
    // (row points at a valid row also for inactive threads: their loads are
    // harmless and their results are replaced by zeros, so no load sits
    // behind a branch and they all issue together)
    cplx<T> wq[EPT];
    split_twiddles<T, EPT, TPS>(wq, j, tw, sh);
    if constexpr (MIR) {
        cplx<T> xr[EPT];
#pragma unroll
        for (int q = 0; q < EPT; q++) xr[q] = row[j + q * TPS];
        const cplx<T> xh = row[H];
        if (active) {
#pragma unroll
            for (int q = 0; q < EPT; q++) buf[j + q * TPS] = xr[q];
            if (j == 0) buf[H] = xh;
        }
        exchange_sync<true>();
    }
#pragma unroll
    for (int q = 0; q < EPT; q++) {
        const int k = j + q * TPS;
        cplx<T> Xk = MIR ? buf[k] : row[k];
        cplx<T> Xc = MIR ? buf[H - k] : row[H - k];
        Xc.y = -Xc.y;
        cplx<T> Fe = cadd(Xk, Xc);
        cplx<T> w = wq[q];
        w.y = -w.y;  // W_Px^{-k}
        cplx<T> Fo = cmul(csub(Xk, Xc), w);
        v[q] = active ? cplx<T>{Fe.x - Fo.y, Fe.y + Fo.x} : cplx<T>{0, 0};  // Fe + i Fo
    }
    if constexpr (MIR) exchange_sync<true>();
    fft_reg<T, H, true, 8, RowWave<H>::on>(v, j, buf, tw, G.lgLmax);
#pragma unroll
    for (int q = 0; q < EPT; q++) {
        const int k = j + q * TPS;
        if (!TL || 2 * k < nx) buf[k] = v[q];
    }
    if constexpr (TL) {
        if constexpr (!MMX_P5_TILE_EARLY) tile_load();
        tile_store();
    }
    __syncthreads();
    const MMX_CONST FieldParams<T> &P = *(const MMX_CONST FieldParams<T> *)Pp;
    double red = 0;
    // next spectrum (STAGE / RK4_FINAL of the specialised epilogues): the cell
    // replaces its demag value in LDS by mu0Ms * (new state), and the rows are
    // transformed as the first pass would transform them for the next stage
    constexpr bool NXOK = V == 1 || V == 2 || V == 4;
    const bool nx_on = NXOK && E.anext != nullptr;
    T *ldsr = reinterpret_cast<T *>(lds);
    // (row, x) of this thread's cell, advanced without dividing
    int rr = (int)threadIdx.x / nx, x = (int)threadIdx.x - rr * nx;
    for (int idx = threadIdx.x; idx < nrows * nx; idx += blockDim.x) {
        const int y = y0, z = z0 + rr;
        const int rw = z * ny + y;
        const cplx<T> h0 = lds[(0 * RX + rr) * HS + (x >> 1)];
        const cplx<T> h1 = lds[(1 * RX + rr) * HS + (x >> 1)];
        const cplx<T> h2 = lds[(2 * RX + rr) * HS + (x >> 1)];
        const T hx = (x & 1) ? h0.y : h0.x;
        const T hy = (x & 1) ? h1.y : h1.x;
        const T hz = (x & 1) ? h2.y : h2.x;
        RowTile<T> tl{};
        if constexpr (TL) tl = RowTile<T>{ldsr + 2 * rr * HS + mo, 2 * RX * HS, 2 * HS, rr, nrows};
        if constexpr (NXOK) {
            T yn[3];
            const int I = rw * nx + x;
            // read with the cell's other loads (the field reads the same value)
            const T ms = gld(P.mu0Ms, I);
            epilogue<T, V, SIG, MM_GENERIC, TL>(P, dyn, E, (int64_t)I, x, y, z, m, hx, hy, hz, red, yn, tl);
            if (nx_on) {
#pragma unroll
                for (int c = 0; c < 3; c++) ldsr[2 * (c * RX + rr) * HS + x] = yn[c] * ms;
            }
        } else {
            epilogue<T, V, SIG, IMG ? MM_FIELD : MM_GENERIC, TL>(P, dyn, E, (int64_t)rw * nx + x, x, y, z, m, hx, hy,
                                                               hz, red, nullptr, tl);
        }
        x += blockDim.x;
        while (x >= nx) {
            x -= nx;
            rr++;
        }
    }
    if constexpr (!NXOK) {
        if (E.mode == EPI_ERR || E.mode == EPI_ENERGY) block_reduce_store<1>(&red, 1, E.mode == EPI_ERR, E.partial, 1);
    }
    if (NXOK && nx_on) {
        __syncthreads();
        // the thread's sequence index and buffer pass through an empty asm
        // statement: otherwise the compiler keeps the first transform's LDS
        // addresses live through the cell loop to reuse them here, at the
        // cost of occupancy
        int tid = (int)threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int jt = tid % TPS;
        cplx<T> *bt = lds + (tid / TPS) * HS;  // (c*RX + r)*HS
        // row r of component c: z[n] = (M[2n], M[2n+1]) for 2n < nx, zero beyond
        // k < H lies inside the buffer: read unconditionally, select the zeros
#pragma unroll
        for (int q = 0; q < EPT; q++) {
            const int k = jt + q * TPS;
            const cplx<T> Z = bt[k];
            const bool in = active && 2 * k < nx;
            v[q] = {in ? Z.x : T(0), (in && 2 * k + 1 < nx) ? Z.y : T(0)};
        }
        exchange_sync<RowWave<H>::on>();
        fft_reg<T, H, false, 8, RowWave<H>::on>(v, jt, bt, tw, G.lgLmax);
        cplx<T> *dst = E.anext + ((int64_t)c * nrows_tot + (int64_t)(z0 + (active ? r : 0)) * ny + y0) * G.KS;
        r2c_split_store<T, H, RowWave<H>::on>(v, jt, bt, active, dst, tw, G.lgLmax);
    }
}

template <class T, int H, int V, uint32_t SIG = 0, bool IMG = false>
__global__ void __launch_bounds__(Shape<H>::B5) MMX_P5_ATTR k_p5(DemagGeom G, const cplx<T> *__restrict__ A,
                                                      const FieldParams<T> *__restrict__ Pp, Dyn dyn, Epi<T> E,
                                                      const T *__restrict__ m, const cplx<T> *__restrict__ tw)
{
    if constexpr (IMG) {
        const int64_t q = blockIdx.y, Nc = (int64_t)G.nx * G.ny * G.nz;
        A += q * 3 * (int64_t)G.ny * G.nz * G.KS;
        m += q * 3 * Nc;
        E.out += q * 3 * Nc;
        if (E.e_out) E.e_out += q * Nc;
    }
    p5_body<T, H, V, SIG, IMG>(G, A, Pp, dyn, E, m, tw);
}

// Batch of systems of one geometry (one demag tensor): grid y = member q of
// the group, whose spectra lie at A + q*astride; its parameters, dynamic
// scalars and stage descriptor (epilogue and input vector) are entries gidx[q]
// of the batch tables.  Every member takes the same stage mode.
template <class T, int H, int V>
__global__ void __launch_bounds__(Shape<H>::B5) MMX_P5_ATTR
    k_p5_batch(DemagGeom G, const cplx<T> *__restrict__ A, int64_t astride, const int *__restrict__ gidx,
               const FieldParams<T> *const *__restrict__ Ps, const BatchStage<T> *__restrict__ st,
               const Dyn *__restrict__ dyns, const cplx<T> *__restrict__ tw)
{
    const int q = blockIdx.y;
    const int mi = gidx[q];
    const BatchStage<T> &bs = st[mi];
    p5_body<T, H, V, 0u, false>(G, A + q * astride, Ps[mi], dyns[mi], bs.E, bs.m, tw);
}

// ---------------------------------------------------------------------------
// Local-field kernel (meshes without demag): one thread per cell.
// ---------------------------------------------------------------------------
template <class T, int V, uint32_t SIG = 0>
__global__ void __launch_bounds__(BLOCK) k_local(const FieldParams<T> *__restrict__ Pp, Dyn dyn, Epi<T> E,
                                                 const T *__restrict__ m)
{
    const MMX_CONST FieldParams<T> &P = *(const MMX_CONST FieldParams<T> *)Pp;
    double red = 0;
    const unsigned nx = P.nx, ny = P.ny;
    const int N = (int)P.N;  // meshes hold fewer than 2^31 cells
    for (int I = blockIdx.x * BLOCK + threadIdx.x; I < N; I += gridDim.x * BLOCK) {
        const unsigned q = (unsigned)I / nx;
        const int i = (int)((unsigned)I - q * nx);
        const unsigned k = q / ny;
        const int j = (int)(q - k * ny);
        epilogue<T, V, SIG>(P, dyn, E, I, i, j, (int)k, m, T(0), T(0), T(0), red);
    }
    if (E.mode == EPI_ERR || E.mode == EPI_ENERGY) block_reduce_store<1>(&red, 1, E.mode == EPI_ERR, E.partial, 1);
}

// dispatch of a run-time log2 length to a compile-time length
#define MMX_LG_SWITCH(lg, CALL)                                                 \
    switch (lg) {                                                              \
    case 0: CALL(1); break;                                                    \
    case 1: CALL(2); break;                                                    \
    case 2: CALL(4); break;                                                    \
    case 3: CALL(8); break;                                                    \
    case 4: CALL(16); break;                                                   \
    case 5: CALL(32); break;                                                   \
    case 6: CALL(64); break;                                                   \
    case 7: CALL(128); break;                                                  \
    case 8: CALL(256); break;                                                  \
    case 9: CALL(512); break;                                                  \
    case 10: CALL(1024); break;                                                \
    case 11: CALL(2048); break;                                                \
    default: break;                                                            \
    }

// epilogue variant for a mode: the slot-form kernels take the integrator stages,
// the adjoint its own term-list instantiation
inline int variant(int mode, int fast, int np)
{
    if (mode == EPI_ADJ) return 5;
    if (!fast) return 0;
    if (mode == EPI_STAGE) return np == 0 ? 4 : 1;
    if (mode == EPI_RK4_FINAL) return 2;
    if (mode == EPI_ERR) return 3;
    return 0;
}

}  // namespace mmx
