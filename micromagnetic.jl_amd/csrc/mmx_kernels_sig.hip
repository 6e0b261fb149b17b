// mmx_kernels_sig.hip -- the last demag pass and the local-field kernel
// compiled for fixed term signatures (mmx_common.h): with the switches and
// the insertion order known at compile time the cell code is straight-line.
// The launchers return -1 when no kernel matches, and the caller then takes
// the signature-free kernels.
#include <hip/hip_runtime.h>

#include "mmx_common.h"
#include "mmx_kernels.h"
#include "mmx_pass.h"

namespace mmx {

template <class T, int H, uint32_t SIG>
static int p5_v(int v, int g, const DemagGeom &G, const cplx<T> *A, const FieldParams<T> *P, const Dyn &dyn,
                const Epi<T> &E, const T *m, const cplx<T> *tw, hipStream_t s)
{
    const dim3 b(Shape<H>::B5);
    if (v == 4) hipLaunchKernelGGL((k_p5<T, H, 4, SIG>), dim3(g), b, 0, s, G, A, P, dyn, E, m, tw);
    else if (v == 1) hipLaunchKernelGGL((k_p5<T, H, 1, SIG>), dim3(g), b, 0, s, G, A, P, dyn, E, m, tw);
    else if (v == 2) hipLaunchKernelGGL((k_p5<T, H, 2, SIG>), dim3(g), b, 0, s, G, A, P, dyn, E, m, tw);
    else hipLaunchKernelGGL((k_p5<T, H, 3, SIG>), dim3(g), b, 0, s, G, A, P, dyn, E, m, tw);
    return g;
}

template <class T, uint32_t SIG>
static int p5_h(int v, const DemagGeom &G, const cplx<T> *A, const FieldParams<T> *P, const Dyn &dyn,
                const Epi<T> &E, const T *m, const cplx<T> *tw, hipStream_t s)
{
#define P5S(HH)                                                                                      \
    case HH: {                                                                                       \
        const int g = p5_grid(G);                                                                    \
        return p5_v<T, HH, SIG>(v, g, G, A, P, dyn, E, m, tw, s);                                     \
    }
    switch (G.H) {
        P5S(16)
        P5S(32)
        P5S(64)
        P5S(128)
        P5S(256)
        P5S(512)
        P5S(1024)
    default: return -1;
    }
#undef P5S
}

template <class T>
int launch_p5_sig(const DemagGeom &G, const cplx<T> *A, const FieldParams<T> *P, const Dyn &dyn, const Epi<T> &E,
                  const T *m, uint32_t sig, const cplx<T> *tw, hipStream_t s)
{
    const int v = variant(E.mode, 1, E.np);
    if (v == 0 || v == 5) return -1;
    switch (sig) {
    case SIG_CHIRAL: return p5_h<T, SIG_CHIRAL>(v, G, A, P, dyn, E, m, tw, s);
    case SIG_SOFT: return p5_h<T, SIG_SOFT>(v, G, A, P, dyn, E, m, tw, s);
    case SIG_BULK: return p5_h<T, SIG_BULK>(v, G, A, P, dyn, E, m, tw, s);
    case SIG_CHIRAL_CELL: return p5_h<T, SIG_CHIRAL_CELL>(v, G, A, P, dyn, E, m, tw, s);
    default: return -1;
    }
}

template <class T>
int launch_local_sig(int64_t N, const FieldParams<T> *P, const Dyn &dyn, const Epi<T> &E, const T *m, uint32_t sig,
                     int grid, hipStream_t s)
{
    const int v = variant(E.mode, 1, E.np);
    if (v == 0 || v == 5 || sig != SIG_CHIRAL_LOCAL) return -1;
    const dim3 g(grid), b(BLOCK);
    if (v == 4) hipLaunchKernelGGL((k_local<T, 4, SIG_CHIRAL_LOCAL>), g, b, 0, s, P, dyn, E, m);
    else if (v == 1) hipLaunchKernelGGL((k_local<T, 1, SIG_CHIRAL_LOCAL>), g, b, 0, s, P, dyn, E, m);
    else if (v == 2) hipLaunchKernelGGL((k_local<T, 2, SIG_CHIRAL_LOCAL>), g, b, 0, s, P, dyn, E, m);
    else hipLaunchKernelGGL((k_local<T, 3, SIG_CHIRAL_LOCAL>), g, b, 0, s, P, dyn, E, m);
    return grid;
}

#define MMX_INST_SIG(T)                                                                                   \
    template int launch_p5_sig<T>(const DemagGeom &, const cplx<T> *, const FieldParams<T> *, const Dyn &,    \
                                  const Epi<T> &, const T *, uint32_t, const cplx<T> *, hipStream_t);        \
    template int launch_local_sig<T>(int64_t, const FieldParams<T> *, const Dyn &, const Epi<T> &, const T *, \
                                     uint32_t, int, hipStream_t);

#ifndef MMX_KT
#error "compile once per precision with -DMMX_KT=double or -DMMX_KT=float"
#endif
MMX_INST_SIG(MMX_KT)

}  // namespace mmx
