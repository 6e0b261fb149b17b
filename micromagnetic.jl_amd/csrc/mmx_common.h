// mmx_common.h -- types shared by the HIP kernels and the host library.
//
// Device data layout (all per-cell arrays are x-fastest, then y, then z, the
// cell order of src/micro/mesh.jl:24-29):
//   m, k_s, y  : structure-of-arrays, three planes of N values (mx | my | mz)
//   mu0Ms      : N values (mu0 * Ms, as src/sim.jl:138 stores it)
//   pins       : N bytes
// Precision T is double (FP64 path) or float (FP32 path).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mmx {

constexpr int MAX_TERMS = 10;

// The parameter block is read through the constant address space: the
// compiler then knows that no store of a kernel can change it, keeps its
// fields in scalar registers across the cell loops and loads them with
// scalar loads (the block stays in ordinary device memory).
#define MMX_CONST __attribute__((address_space(4)))

// Loads from per-cell arrays through the global address space (plain
// pointers would otherwise become flat accesses, which also count against
// the scalar/LDS wait counter).
template <class X>
__device__ __forceinline__ X gld(const X *p, int64_t i)
{
    return ((const __attribute__((address_space(1))) X *)p)[i];
}
// Load of the two values p[2i], p[2i+1] (p 2*sizeof(X)-aligned) as one vector
// access through the global address space.
template <class X>
__device__ __forceinline__ void gld2(const X *p, int64_t i, X &a, X &b)
{
    typedef X v2 __attribute__((vector_size(2 * sizeof(X))));
    const v2 v = ((const __attribute__((address_space(1))) v2 *)p)[i];
    a = v[0];
    b = v[1];
}
// Loads from LDS through a pointer the compiler cannot trace to a __shared__
// array (LDS address space: ds_read, not flat).
template <class X>
__device__ __forceinline__ X lld(const X *p, int i)
{
    return ((const __attribute__((address_space(3))) X *)p)[i];
}
constexpr int BLOCK = 256;

enum TermKind : int {
    K_NONE = 0,
    K_ZEEMAN,       // static per-cell field (three planes in a/b/c)
    K_ZEEMAN_U,     // static uniform field s[0..2]
    K_EXCH,         // uniform exchange, s = {ax, ay, az} = 2A/d^2
    K_EXCH_S,       // spatial A (a), s = {2/dx^2, 2/dy^2, 2/dz^2}
    K_DMI_BULK,     // uniform bulk DMI, s = {Dx/dx, Dy/dy, Dz/dz}
    K_DMI_BULK_S,   // spatial bulk DMI (a,b,c = Dx,Dy,Dz), s = {1/dx, 1/dy, 1/dz}
    K_DMI_INT,      // interfacial DMI, per-cell D (a) or uniform s[3] when a == null
    K_ANIS,         // uniaxial, Ku (a) or uniform s[3]; axis s[0..2]
    K_ANIS_S,       // uniaxial with per-cell axis planes (b), Ku (a) or s[3]
    K_CUBIC,        // cubic, Kc (a) or uniform s[9]; axes s[0..8]
    K_DEMAG,        // FFT demag; its field comes from the convolution passes
    K_THERMAL,      // stochastic field, T (a) or uniform s[1]; s[0] = T0
    K_SOT,          // damping/field-like torque, aj (a) or uniform s[3]; p s[0..2], bj s[4]
    K_STT_SL,       // Slonczewski torque, J (a) or uniform s[3]; p s[0..2], Lambda^2 s[4], P s[5], xi s[6]
    K_STT_ZL,       // Zhang-Li torque, b*J planes (b), xi (a) or uniform s[3]; dx, dy, dz s[0..2]
    K_ZEEMAN_T,     // time-dependent Zeeman: static planes (a,b,c) or uniform s[0..2], times ft(t)
    K_HEX,          // hexagonal anisotropy, s = {K1, K2, K3}
    K_EXCH_IL,      // interlayer exchange: J per (i,j) in a (nx*ny), s = {-, -, -, dz, k1, k2}
    K_DMI_IL,       // interlayer DMI: s = {Dx, Dy, Dz, dz, k1, k2} (layers 0-based)
    K_MEL_T,        // magnetoelastic, tensor model: a = 6 planes of stress (xx,yy,zz,xy,xz,yz), s[0] = lambda_s
    K_MEL_C,        // magnetoelastic, cubic model: a = 6 planes of strain, s = {B1, B2}
    K_TORQUE_FN,    // add_torque: the torque T (3 planes in a) written by a host function per evaluation
    K_SAHE,         // spin anomalous Hall torque: a = 15 planes (sigma_s x a1, sigma_sa1 x a1, sigma_sa2, a2, a3), s[0] = beta
    K_TWIN,         // twin monoclinic anisotropy: s = {Ka, Kb, Kaa, Kbb, Kab, Ku}; axes a, b, u111 (3 planes each)
};

template <class T>
struct TermDev {
    int kind;
    int uniform;     // parameter given as a scalar instead of an array
    int tdep;        // bulk DMI scaled by time_D(t): values in Dyn::tfac/tfy/tfz
    T s[12];
    const T *a, *b, *c;
};

template <class T>
struct FieldParams {
    int nx, ny, nz;
    int px, py, pz;  // periodic flags
    int64_t N;
    T volume;
    const T *mu0Ms;
    const uint8_t *pins;
    // per cell: bit q (q = 0..5, order -x,+x,-y,+y,-z,+z) set when neighbour q
    // exists and has Ms > 0 (the stencils' "use this neighbour" test,
    // src/micro/kernels.jl:291); bit 6 set when the cell is pinned
    const uint8_t *nbm;
    int nterms;
    TermDev<T> t[MAX_TERMS];
    // LLG driver (src/llg/llg.jl:6-40); alphas: per-cell damping of the
    // SpatialLLG driver (src/llg/llg.jl:62-97), null for the uniform alpha
    T alpha, gamma;
    const T *alphas;
    int precession;
    uint64_t seed;  // thermal noise stream
    // Slot form used by the specialised kernels when every kind of term
    // appears at most once: slot[k] = term index of kind-slot k (or -1), and
    // ord[q] = kind-slot of the q-th term in insertion order.
    int fast;
    int slot[8];
    int ord[MAX_TERMS];
    T llg_a;  // -gamma / (1 + alpha^2), src/llg/llg.jl:18
    // the slot form's switches and insertion order packed into two words
    // (FL_* bits below; ordp holds ord[q] in bits 3q..3q+2)
    uint32_t flags, ordp;
    uint32_t sig;  // term signature (0 when none applies)
};

enum SlotFlags : uint32_t {
    FL_PX = 1u << 0, FL_PY = 1u << 1, FL_PZ = 1u << 2,
    FL_EX_SHIFT = 3,   // 2 bits: 0 none, 1 uniform, 2 spatial A
    FL_BD_SHIFT = 5,   // 2 bits: bulk DMI
    FL_ID_SHIFT = 7,   // 2 bits: interfacial DMI
    FL_AN = 1u << 9, FL_AN_AXIS = 1u << 10, FL_AN_KU = 1u << 11,
    FL_CU = 1u << 12, FL_CU_KC = 1u << 13,
    FL_ZE_SHIFT = 14,  // 2 bits: 0 none, 1 uniform, 2 per cell
    FL_TH = 1u << 16,
    FL_PREC = 1u << 18,
    FL_NT_SHIFT = 20,  // 4 bits: number of terms
};



enum SlotKind : int { SL_EXCH = 0, SL_DMIB, SL_DMII, SL_ANIS, SL_CUBIC, SL_ZEE, SL_DEMAG, SL_THERM };

// A term signature: the set of slots present (bit = SlotKind) when there is
// no thermal term, the terms were added in slot order and every term has
// uniform parameters (no per-cell arrays) except that interfacial DMI and the
// uniaxial Ku may be given per cell (bits SG_ID_CELL, SG_KU_CELL).  Kernels compiled for a signature know
// the switches and the insertion order at compile time; signature 0 reads
// them from the parameter block.
constexpr uint32_t SG_EX = 1u << SL_EXCH, SG_BD = 1u << SL_DMIB, SG_ID = 1u << SL_DMII, SG_AN = 1u << SL_ANIS,
                   SG_CU = 1u << SL_CUBIC, SG_ZE = 1u << SL_ZEE, SG_DM = 1u << SL_DEMAG;
constexpr uint32_t SG_ID_CELL = 1u << 8, SG_KU_CELL = 1u << 9;
// the signatures the kernels are compiled for
constexpr uint32_t SIG_CHIRAL = SG_EX | SG_ID | SG_AN | SG_ZE | SG_DM;    // film with iDMI, Ku, field, demag
constexpr uint32_t SIG_SOFT = SG_EX | SG_DM;                              // exchange + demag
constexpr uint32_t SIG_BULK = SG_EX | SG_BD | SG_CU | SG_DM;              // bulk DMI, cubic, demag
constexpr uint32_t SIG_CHIRAL_LOCAL = SG_EX | SG_ID | SG_AN | SG_ZE;      // as SIG_CHIRAL without demag
// film with D(x,y) and Ku(x,y) maps and demag, no field
constexpr uint32_t SIG_CHIRAL_CELL = SG_EX | SG_ID | SG_AN | SG_DM | SG_ID_CELL | SG_KU_CELL;

constexpr uint32_t sig_flags(uint32_t sig)
{
    uint32_t f = 0, n = 0;
    for (int b = 0; b < 8; b++) n += (sig >> b) & 1u;
    if (sig & SG_EX) f |= 1u << FL_EX_SHIFT;
    if (sig & SG_BD) f |= 1u << FL_BD_SHIFT;
    if (sig & SG_ID) f |= ((sig & SG_ID_CELL) ? 2u : 1u) << FL_ID_SHIFT;
    if (sig & SG_AN) f |= FL_AN | ((sig & SG_KU_CELL) ? (uint32_t)FL_AN_KU : 0u);
    if (sig & SG_CU) f |= FL_CU;
    if (sig & SG_ZE) f |= 1u << FL_ZE_SHIFT;
    return f | (n << FL_NT_SHIFT);
}

constexpr uint32_t sig_ordp(uint32_t sig)
{
    uint32_t o = 0, q = 0;
    for (uint32_t b = 0; b < 8; b++)
        if ((sig >> b) & 1u) {
            o |= b << (3 * q);
            q++;
        }
    return o;
}

// Per-launch values (by value; the parameter block above lives in device
// memory and is read with wave-uniform scalar loads).
struct Dyn {
    double thermal_factor;  // 2 alpha kB / (V gamma dt), src/micro/field.jl:261
    uint64_t noise_step;    // integrator step count keying the noise stream
    double tfac[MAX_TERMS]; // time factor of a term at the evaluation time (Slonczewski:
                            // ufun(t)*beta; Zhang-Li: ufun(t)/gamma; time Zeeman: fx(t))
    double tfy[MAX_TERMS], tfz[MAX_TERMS];  // time Zeeman: fy(t), fz(t)
};

// What the last pass does with the field of a cell.
enum EpiMode : int {
    EPI_FIELD = 0,   // write H (3 planes) and, if e != null, per-cell energy of the masked terms
    EPI_RHS,         // write dm/dt
    EPI_STAGE,       // k = dm/dt; optionally store k (kstore) and H (out); y_out = y0 + sum cp_j kp_j + cself k
    EPI_STAGE_SCALED,// y_out = y0 + outer * (sum cp_j kp_j + cself k)    (BS23 stage 3)
    EPI_RK4_FINAL,   // y_out = y0 + cself * (((k1 + 2 k2) + 2 k3) + k)  (src/integrator/rk.jl:46)
    EPI_ERR,         // store k; err = sum cp_j kp_j + cself k; block max of |err|/(1+|y_in|)
    EPI_ENERGY,      // per-term energy block sums (FP64)
    EPI_ADJ,         // adjoint of one RHS evaluation (AdjIO below)
};

// Sets of epilogue modes a term-list kernel is compiled for (bit = EpiMode):
// each instantiation carries only the code of the modes it serves, which keeps
// its registers within the file and off the scratch segment.
constexpr uint32_t MM_ADJ = 1u << EPI_ADJ;
constexpr uint32_t MM_FIELD = 1u << EPI_FIELD;
constexpr uint32_t MM_STAGES = (1u << EPI_STAGE) | (1u << EPI_STAGE_SCALED) | (1u << EPI_RK4_FINAL);
constexpr uint32_t MM_GENERIC = ~MM_ADJ;  // every mode but the adjoint

// Adjoint of k = f(Y) = F(Y, H(Y)) applied to a cotangent a (EPI_ADJ).  The
// pass's input vector is ut = (dF/dh)^T a / mu0Ms (zero where Ms = 0 or the
// cell is pinned); the epilogue forms
//   mu = mu0Ms * H_lin(ut) + (dF/dm)^T a
// (H_lin: the field's Jacobian at Y, which for the energy-derived terms is
// diag(mu0Ms)^-1-symmetric, so J_H^T u = mu0Ms * J_H(u / mu0Ms)), adds mu to
// lam and cnext * mu to next, and adds the terms' parameter gradients
// u . dH/dp to g[term] and u . dH/d(mu0Ms) to gms.
template <class T>
struct AdjIO {
    const T *Y;   // state the RHS was evaluated at
    const T *hY;  // total field at Y
    const T *a;   // cotangent of k
    T *lam;       // += mu
    T *next;      // += cnext * mu (nullable)
    double cnext;
    T *g[MAX_TERMS];  // per-cell parameter gradient of each term (nullable; 3 planes for bulk DMI)
    T *gms;           // per-cell gradient with respect to mu0Ms (nullable)
};

template <class T>
struct cplx;

template <class T>
struct Epi {
    int mode;
    uint32_t term_mask;  // which terms enter the field (bit per term index)
    T *out;              // FIELD: H planes; RHS: dm/dt planes
    T *e_out;            // FIELD: per-cell energy (nullable)
    T *kstore;           // store k (nullable)
    const T *y0;         // base state of the stage combination
    const T *kp[6];
    double cp[6];
    int np;
    double cself;
    double outer;
    T *yout;
    int normalize;
    double *partial;     // one slot per block (ERR: max; ENERGY: nterms slots per block)
    AdjIO<T> adj;
    cplx<T> *anext;      // last demag pass, STAGE / RK4_FINAL: x-spectrum of mu0Ms*yout (nullable)
};

// One stage of one member of a batch (many independent systems advanced by
// the same launches): the member's epilogue and the vector the stage reads.
template <class T>
struct BatchStage {
    Epi<T> E;
    const T *m;
};

template <class T>
struct cplx {
    T x, y;
};

template <class T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

template <class T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b)
{
    return {a.x + b.x, a.y + b.y};
}

template <class T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b)
{
    return {a.x - b.x, a.y - b.y};
}

// Quad permute across lanes (DPP, a VALU move: no LDS or memory access):
// lane 4a+q returns the value of lane 4a+((CTRL >> 2q) & 3).  Every lane of
// the quad must be active.
template <int CTRL>
__device__ __forceinline__ float quad_perm(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double quad_perm(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
constexpr int QP_XOR1 = 0xB1;                                        // [1,0,3,2]
template <int Q>
constexpr int QP_BCAST = Q | (Q << 2) | (Q << 4) | (Q << 6);         // [Q,Q,Q,Q]

// Demag convolution geometry.  Padded lengths are powers of two >= 2n-1
// (any such length gives the same linear convolution as the reference's
// 2n / 2n-1 padding of src/micro/demag.jl:41-44).  x is transformed real-to-
// complex: H = Px/2 complex points per row, Kx = H+1 output columns, rows of
// the intermediate arrays are KS >= Kx complex values apart.
struct DemagGeom {
    int nx, ny, nz;
    int Px, Py, Pz;
    int H, Kx, KS;
    int lgH, lgPy, lgPz;
    int Lmax, lgLmax;   // twiddle table length
    int Py2, Pz2;       // Py/2+1, Pz/2+1 (folded tensor extents)
    int ktw;            // 1: folded tensor grouped by 4 kx per (kyf, kzf run), for the wave-form z pass
    int p5rx;           // rows per block of the last pass (at most its shape's RX; p5_rows)
};

}  // namespace mmx
