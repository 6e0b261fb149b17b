// mmx_kernels_p5.hip -- launchers of the last demag pass (k_p5, mmx_pass.h) for
// every transform length and epilogue variant.  Compiled once per precision
// (-DMMX_KT=double / float) so the instantiations build in parallel.
#include <hip/hip_runtime.h>

#include "mmx_common.h"
#include "mmx_kernels.h"
#include "mmx_pass.h"

#ifndef MMX_KT
#error "compile once per precision with -DMMX_KT=double or -DMMX_KT=float"
#endif

namespace mmx {

#ifdef MMX_KT_MAIN_P5
// blocks of one last-pass launch: G.p5rx rows per block, groups rounded up to
// the 8 XCDs
int p5_grid(const DemagGeom &G)
{
    return (G.ny * ((G.nz + G.p5rx - 1) / G.p5rx) + 7) / 8 * 8;
}

// Rows per block of the last pass for a launch over `copies` systems of one
// geometry: the largest of RX, RX/2, ..., 1 that gives the launch at least
// MMX_P5_MIN_BLOCKS blocks (2 per CU), else the one that gives the most.  A
// small mesh (128x128x8: 256 blocks at RX = 4) otherwise leaves each CU one
// block of three waves, and the pass waits on memory latency.
int p5_rows(const DemagGeom &G, int copies)
{
    int rx = 1;
#define P5R_CALL(H) rx = Shape<H>::RX;
    MMX_LG_SWITCH(G.lgH, P5R_CALL)
#undef P5R_CALL
    auto blocks = [&](int r) { return (int64_t)copies * G.ny * ((G.nz + r - 1) / r); };
    int best = rx;
    for (int r = rx / 2; r >= 1 && blocks(best) < MMX_P5_MIN_BLOCKS; r /= 2)
        if (blocks(r) > blocks(best)) best = r;
    return best;
}
#endif

template <class T>
int launch_p5(const DemagGeom &G, const cplx<T> *A, const FieldParams<T> *P, const Dyn &dyn, const Epi<T> &E,
              const T *m, int fast, uint32_t sig, const cplx<T> *tw, hipStream_t s)
{
    if (fast && sig) {
        const int gs = launch_p5_sig<T>(G, A, P, dyn, E, m, sig, tw, s);
        if (gs >= 0) return gs;
    }
    int g = 0;
#define P5_CALL(H)                                                                                   \
    {                                                                                                \
        g = p5_grid(G);                                                                              \
        const int v = variant(E.mode, fast, E.np);                                                   \
        if (v == 4)                                                                                  \
            hipLaunchKernelGGL((k_p5<T, H, 4>), dim3(g), dim3(Shape<H>::B5), 0, s, G, A, P, dyn, E, m, tw); \
        else if (v == 1)                                                                             \
            hipLaunchKernelGGL((k_p5<T, H, 1>), dim3(g), dim3(Shape<H>::B5), 0, s, G, A, P, dyn, E, m, tw); \
        else if (v == 2)                                                                             \
            hipLaunchKernelGGL((k_p5<T, H, 2>), dim3(g), dim3(Shape<H>::B5), 0, s, G, A, P, dyn, E, m, tw); \
        else if (v == 3)                                                                             \
            hipLaunchKernelGGL((k_p5<T, H, 3>), dim3(g), dim3(Shape<H>::B5), 0, s, G, A, P, dyn, E, m, tw); \
        else if (v == 5)                                                                             \
            hipLaunchKernelGGL((k_p5<T, H, 5>), dim3(g), dim3(Shape<H>::B5), 0, s, G, A, P, dyn, E, m, tw); \
        else                                                                                         \
            hipLaunchKernelGGL((k_p5<T, H, 0>), dim3(g), dim3(Shape<H>::B5), 0, s, G, A, P, dyn, E, m, tw); \
    }
    MMX_LG_SWITCH(G.lgH, P5_CALL)
#undef P5_CALL
    return g;
}

template <class T>
void launch_p5_images(const DemagGeom &G, const cplx<T> *A, const FieldParams<T> *P, const Dyn &dyn,
                      const Epi<T> &E, const T *m, int nimg, const cplx<T> *tw, hipStream_t s)
{
#define P5I_CALL(H)                                                                                  \
    {                                                                                                \
        const int g = p5_grid(Gi);                                                                   \
        hipLaunchKernelGGL((k_p5<T, H, 0, 0u, true>), dim3(g, nimg), dim3(Shape<H>::B5), 0, s, Gi, A, P, dyn, E, m, \
                           tw);                                                                      \
    }
    DemagGeom Gi = G;
    Gi.p5rx = p5_rows(G, nimg);
    MMX_LG_SWITCH(Gi.lgH, P5I_CALL)
#undef P5I_CALL
}

template <class T>
void launch_p5_batch(const DemagGeom &G, const cplx<T> *A, int64_t astride, const int *gidx, int nmem,
                     const FieldParams<T> *const *Ps, const BatchStage<T> *st, const Dyn *dyns, int v,
                     const cplx<T> *tw, hipStream_t s)
{
#define P5B_CALL(H)                                                                                  \
    {                                                                                                \
        const dim3 g(p5_grid(Gb), nmem);                                                             \
        const dim3 b(Shape<H>::B5);                                                                  \
        if (v == 4)                                                                                  \
            hipLaunchKernelGGL((k_p5_batch<T, H, 4>), g, b, 0, s, Gb, A, astride, gidx, Ps, st, dyns, tw); \
        else if (v == 1)                                                                             \
            hipLaunchKernelGGL((k_p5_batch<T, H, 1>), g, b, 0, s, Gb, A, astride, gidx, Ps, st, dyns, tw); \
        else if (v == 2)                                                                             \
            hipLaunchKernelGGL((k_p5_batch<T, H, 2>), g, b, 0, s, Gb, A, astride, gidx, Ps, st, dyns, tw); \
        else                                                                                         \
            hipLaunchKernelGGL((k_p5_batch<T, H, 0>), g, b, 0, s, Gb, A, astride, gidx, Ps, st, dyns, tw); \
    }
    DemagGeom Gb = G;
    Gb.p5rx = p5_rows(G, nmem);
    MMX_LG_SWITCH(Gb.lgH, P5B_CALL)
#undef P5B_CALL
}

template int launch_p5<MMX_KT>(const DemagGeom &, const cplx<MMX_KT> *, const FieldParams<MMX_KT> *, const Dyn &,
                               const Epi<MMX_KT> &, const MMX_KT *, int, uint32_t, const cplx<MMX_KT> *, hipStream_t);
template void launch_p5_batch<MMX_KT>(const DemagGeom &, const cplx<MMX_KT> *, int64_t, const int *, int,
                                      const FieldParams<MMX_KT> *const *, const BatchStage<MMX_KT> *, const Dyn *,
                                      int, const cplx<MMX_KT> *, hipStream_t);
template void launch_p5_images<MMX_KT>(const DemagGeom &, const cplx<MMX_KT> *, const FieldParams<MMX_KT> *,
                                       const Dyn &, const Epi<MMX_KT> &, const MMX_KT *, int, const cplx<MMX_KT> *,
                                       hipStream_t);

}  // namespace mmx
