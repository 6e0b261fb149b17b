// mmx_api.cpp -- the C ABI of libmmx (include/mmx.h): context and parameter
// management, demag set-up, and the integrators that drive the fused kernels.
//
// The steppers restate src/integrator/rk.jl (RK4), heun.jl (Heun),
// adaptive_rk.jl + adaptive_rk_steps.jl (DOPRI54, BS23) and the run loops of
// src/sim.jl (run_until, relax).  Every stage combination is fused into the
// last kernel of the RHS evaluation that produces its k (mmx_device.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/mmx.h"
#include "mmx_common.h"
#include "mmx_kernels.h"

namespace mmx {
void newell_tensor(int nx, int ny, int nz, double Dx, double Dy, double Dz, int Nx, int Ny, int Nz,
                   double *out, int nthreads);
}

using namespace mmx;

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

struct HipError {
    int code;
};

#define HIPCHECK(x)                                                                              \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) {                                                                  \
            fail(MMX_EHIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
            throw HipError{MMX_EHIP};                                                            \
        }                                                                                        \
    } while (0)

// Device memory of the allocation-free energy entry point
// (mmx_compute_energies_noalloc): the per-block partial sums of one term's
// evaluation, reused term by term.  Each GPU the library runs on holds its own
// copy until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for that device (the array's name is a host-side
// stand-in).  Calls that overlap on one GPU share it and race.
static constexpr int NOALLOC_SLOTS = 65536;
static __device__ __align__(256) double scratch[NOALLOC_SLOTS];
static constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex scratch_mu;
static double *scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// once per device, with that device current (called from mmx_create)
static int scratch_init(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return fail(MMX_EINVAL, "device %d out of range", device);
    std::lock_guard<std::mutex> l(scratch_mu);
    if (scratch_ptr[device]) return MMX_OK;
    double *p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(MMX_EHIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device, hipGetErrorString(e));
    scratch_ptr[device] = p;
    return MMX_OK;
}

static constexpr double MU0 = 4.0 * M_PI * 1e-7;  // src/const.jl:1
static constexpr double KB = 1.3806505e-23;       // src/const.jl:3

template <class X>
struct DevBuf {
    X *p = nullptr;
    size_t n = 0;
    void alloc(size_t count)
    {
        release();
        if (count == 0) return;
        hipError_t e = hipMalloc((void **)&p, count * sizeof(X));
        if (e != hipSuccess) {
            p = nullptr;
            fail(MMX_ENOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(X), hipGetErrorString(e));
            throw HipError{MMX_ENOMEM};
        }
        n = count;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

// host-side description of one interaction term (insertion order = index)
struct HostTerm {
    int kind = K_NONE;
    int uniform = 0;
    double s[12] = {0};
    int narrays = 0;  // number of device arrays owned
    double thermal_T0 = 0;
    uint64_t seed = 0;
    std::vector<double> rtensor;  // demag: compact real-space tensor (6N)
    double (*ufun)(double, void *) = nullptr;  // torque amplitude ufun(t) (null: 1)
    void *uuser = nullptr;
    void (*ft3)(double, double *, void *) = nullptr;  // time Zeeman (fx, fy, fz)(t)
    // add_torque: fn(t, m, y, user) writes the torque y (3N AoS) for the state m (3N AoS)
    void (*tfn)(double, const double *, double *, void *) = nullptr;
    // the host torque buffer handed to fn: zero when the term is added, then kept
    // from call to call like the reference's torque.field (field.jl:282-293)
    std::vector<double> tbuf;
};

struct CtxBase {
    virtual ~CtxBase() = default;
};

struct mmx_ctx {
    mmx_mesh_desc mesh;
    int64_t N;
    double volume;
    hipStream_t stream = nullptr;
    std::unique_ptr<CtxBase> impl;
    std::vector<HostTerm> terms;
    std::vector<double> mu0Ms;  // host copy (FP64)
    // LLG driver and integrator state (src/driver.jl:148, adaptive_rk.jl:37-55)
    double alpha = 0.1, gamma = 2.21e5;
    int precession = 1;
    int ikind = MMX_DOPRI54;
    double t = 0, step = 0, step_next = 0, tol = 1e-6;
    int64_t nsteps = 0, nfevals = 0;
    int succeed = 0;
    int demag_term = -1;
    int demag_img[3] = {0, 0, 0};  // macro-PBC image counts of the demag term
    bool profile = false;
    // SD driver (src/driver.jl:3-12): tau, min/max tau, step count
    double sd_tau = 0, sd_min_tau = 1e-10, sd_max_tau = 1.0;
    int64_t sd_steps = 0;
};

struct mmx_neb {
    mmx_ctx *sim = nullptr;
    std::unique_ptr<CtxBase> impl;
};

struct mmx_batch {
    std::vector<mmx_ctx *> members;
    int precision = MMX_F64;
    int device = 0;
    hipStream_t stream = nullptr;
    std::unique_ptr<CtxBase> impl;
};

namespace {

template <class T>
struct Engine : CtxBase {
    using value_type = T;
    mmx_ctx *C;
    int64_t N;
    DevBuf<T> mu0Ms;
    DevBuf<T> alphas;  // SpatialLLG damping per cell (empty: uniform alpha)
    DevBuf<uint8_t> pins;
    bool have_pins = false;
    bool any_tfn = false;  // an add_torque term is present
    DevBuf<uint8_t> nbm;  // neighbour / pin mask, refreshed when Ms or the pins change
    // state: current spin, three stage buffers, seven k buffers (SoA 3N each)
    DevBuf<T> ybuf[4];
    T *y = nullptr;
    T *stg[3] = {nullptr, nullptr, nullptr};
    DevBuf<T> kbuf[7];
    T *k[7] = {nullptr};
    // per-term device arrays
    std::vector<std::unique_ptr<DevBuf<T>>> arrays;
    std::vector<const T *> tA, tB, tC;
    // demag
    bool demag = false;
    DemagGeom G{};
    DevBuf<cplx<T>> Abuf, Bbuf;
    DevBuf<T> Kt, Lline;
    DevBuf<cplx<T>> tw;
    // kernel parameter block (device copy, re-uploaded when it changes)
    DevBuf<FieldParams<T>> Pdev;
    FieldParams<T> Plast{};
    bool Pvalid = false;
    // reductions and scratch
    DevBuf<double> partial, red;
    DevBuf<double> dscratch;  // 3N doubles for AoS transfers
    // profiling
    struct Ev {
        int pass;
        hipEvent_t a, b;
    };
    std::vector<Ev> evs;
    double prof_ms[MMX_NPASS] = {0};
    int64_t prof_n[MMX_NPASS] = {0};

    explicit Engine(mmx_ctx *c) : C(c), N(c->N)
    {
        mu0Ms.alloc(N);
        HIPCHECK(hipMemsetAsync(mu0Ms.p, 0, N * sizeof(T), C->stream));
        for (auto &b : ybuf) {
            b.alloc(3 * N);
            HIPCHECK(hipMemsetAsync(b.p, 0, 3 * N * sizeof(T), C->stream));
        }
        y = ybuf[0].p;
        for (int i = 0; i < 3; i++) stg[i] = ybuf[i + 1].p;
        partial.alloc((size_t)max_local_blocks() + 64);
        red.alloc(64);
        dscratch.alloc(3 * N);
        nbm.alloc(N);
        update_mask();
    }

    void update_mask()
    {
        const FieldParams<T> *P = params_dev();
        launch_nbmask<T>(N, P, nbm.p, C->stream);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(C->stream));
    }
    ~Engine() override
    {
        for (auto &e : evs) {
            (void)hipEventDestroy(e.a);
            (void)hipEventDestroy(e.b);
        }
    }

    void ensure_k()
    {
        for (int i = 0; i < 7; i++)
            if (!kbuf[i].p) {
                kbuf[i].alloc(3 * N);
                HIPCHECK(hipMemsetAsync(kbuf[i].p, 0, 3 * N * sizeof(T), C->stream));
                k[i] = kbuf[i].p;
            }
    }

    // ---- transfers ---------------------------------------------------------
    void upload_soa(const double *aos, T *dst)
    {
        HIPCHECK(hipMemcpyAsync(dscratch.p, aos, 3 * N * sizeof(double), hipMemcpyHostToDevice, C->stream));
        launch_aos_to_soa<T>(dscratch.p, dst, N, C->stream);
        HIPCHECK(hipGetLastError());
    }
    void download_aos(const T *src, double *aos)
    {
        launch_soa_to_aos<T>(src, dscratch.p, N, C->stream);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(aos, dscratch.p, 3 * N * sizeof(double), hipMemcpyDeviceToHost, C->stream));
        HIPCHECK(hipStreamSynchronize(C->stream));
    }
    void upload_scalar(const double *h, T *dst, int64_t n)
    {
        HIPCHECK(hipMemcpyAsync(dscratch.p, h, n * sizeof(double), hipMemcpyHostToDevice, C->stream));
        launch_cast<T>(dscratch.p, dst, n, C->stream);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(C->stream));
    }
    void download_scalar(const T *src, double *h, int64_t n)
    {
        launch_to_double<T>(src, dscratch.p, n, C->stream);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(h, dscratch.p, n * sizeof(double), hipMemcpyDeviceToHost, C->stream));
        HIPCHECK(hipStreamSynchronize(C->stream));
    }
    // a device copy of a per-cell parameter array (n = N or 3N, AoS -> SoA for 3N)
    const T *new_array(const double *h, int64_t n, bool vec)
    {
        auto b = std::make_unique<DevBuf<T>>();
        b->alloc(n);
        if (vec) {
            upload_soa(h, b->p);
            HIPCHECK(hipStreamSynchronize(C->stream));
        } else {  // through the 3N staging buffer, in pieces for longer arrays
            for (int64_t off = 0; off < n; off += 3 * N) upload_scalar(h + off, b->p + off, std::min<int64_t>(3 * N, n - off));
        }
        const T *p = b->p;
        arrays.push_back(std::move(b));
        return p;
    }

    // ---- parameters of the kernels ----------------------------------------
    FieldParams<T> params() const
    {
        FieldParams<T> P{};
        P.nx = C->mesh.nx;
        P.ny = C->mesh.ny;
        P.nz = C->mesh.nz;
        P.px = (C->mesh.pbc & 1) != 0;
        P.py = (C->mesh.pbc & 2) != 0;
        P.pz = (C->mesh.pbc & 4) != 0;
        P.N = N;
        P.volume = (T)C->volume;
        P.mu0Ms = mu0Ms.p;
        P.pins = have_pins ? pins.p : nullptr;
        P.nbm = nbm.p;
        P.nterms = (int)C->terms.size();
        for (int i = 0; i < P.nterms; i++) {
            const HostTerm &h = C->terms[i];
            TermDev<T> &d = P.t[i];
            d.kind = h.kind;
            d.uniform = h.uniform;
            d.tdep = h.ufun != nullptr && (h.kind == K_DMI_BULK || h.kind == K_DMI_BULK_S);
            for (int q = 0; q < 12; q++) d.s[q] = (T)h.s[q];
            d.a = tA[i];
            d.b = tB[i];
            d.c = tC[i];
        }
        P.alpha = (T)C->alpha;
        P.gamma = (T)C->gamma;
        P.alphas = alphas.p;
        P.precession = C->precession;
        P.seed = 0;
        for (const auto &h : C->terms)
            if (h.kind == K_THERMAL) P.seed = h.seed;
        // slot form: every kind of term at most once
        for (int q = 0; q < 8; q++) P.slot[q] = -1;
        P.fast = 1;
        for (int i = 0; i < P.nterms; i++) {
            int sk;
            switch (C->terms[i].ufun ? K_NONE : C->terms[i].kind) {  // time-dependent: generic
            case K_EXCH: case K_EXCH_S: sk = SL_EXCH; break;
            case K_DMI_BULK: case K_DMI_BULK_S: sk = SL_DMIB; break;
            case K_DMI_INT: sk = SL_DMII; break;
            case K_ANIS: case K_ANIS_S: sk = SL_ANIS; break;
            case K_CUBIC: sk = SL_CUBIC; break;
            case K_ZEEMAN: case K_ZEEMAN_U: sk = SL_ZEE; break;
            case K_ZEEMAN_T: sk = C->terms[i].uniform ? SL_ZEE : -1; break;  // factors in Dyn
            case K_DEMAG: sk = SL_DEMAG; break;
            case K_THERMAL: sk = SL_THERM; break;
            default: sk = -1; break;  // torque terms: generic kernels only
            }
            if (sk < 0) {
                P.fast = 0;
                P.ord[i] = 0;
                continue;
            }
            if (P.slot[sk] >= 0) P.fast = 0;
            P.slot[sk] = i;
            P.ord[i] = sk;
        }
        // packed switches of the slot form
        {
            uint32_t f = (P.px ? FL_PX : 0u) | (P.py ? FL_PY : 0u) | (P.pz ? FL_PZ : 0u);
            auto kind_at = [&](int sk) { return P.slot[sk] >= 0 ? C->terms[P.slot[sk]].kind : (int)K_NONE; };
            auto uni_at = [&](int sk) { return P.slot[sk] >= 0 ? C->terms[P.slot[sk]].uniform : 1; };
            const int ke = kind_at(SL_EXCH), kb = kind_at(SL_DMIB), ki = kind_at(SL_DMII), ka = kind_at(SL_ANIS);
            const int kc = kind_at(SL_CUBIC), kz = kind_at(SL_ZEE);
            f |= (uint32_t)(ke == K_NONE ? 0 : (ke == K_EXCH_S ? 2 : 1)) << FL_EX_SHIFT;
            f |= (uint32_t)(kb == K_NONE ? 0 : (kb == K_DMI_BULK_S ? 2 : 1)) << FL_BD_SHIFT;
            f |= (uint32_t)(ki == K_NONE ? 0 : (uni_at(SL_DMII) ? 1 : 2)) << FL_ID_SHIFT;
            if (ka != K_NONE) {
                f |= FL_AN;
                if (ka == K_ANIS_S) f |= FL_AN_AXIS;
                if (!uni_at(SL_ANIS)) f |= FL_AN_KU;
            }
            if (kc != K_NONE) {
                f |= FL_CU;
                if (!uni_at(SL_CUBIC)) f |= FL_CU_KC;
            }
            f |= (uint32_t)(kz == K_NONE ? 0 : (kz == K_ZEEMAN ? 2 : 1)) << FL_ZE_SHIFT;
            if (P.slot[SL_THERM] >= 0) f |= FL_TH;
            if (C->precession) f |= FL_PREC;
            f |= (uint32_t)P.nterms << FL_NT_SHIFT;
            P.flags = f;
            uint32_t o = 0;
            for (int i = 0; i < P.nterms; i++) o |= (uint32_t)P.ord[i] << (3 * i);
            P.ordp = o;
        }
        // term signature: no thermal term, added in slot order, every term uniform
        // but interfacial DMI and uniaxial Ku, which may be per cell
        {
            bool ok = P.fast != 0;
            uint32_t sig = 0;
            int last = -1;
            for (int i = 0; i < P.nterms && ok; i++) {
                const HostTerm &h = C->terms[i];
                const bool uni = h.kind == K_EXCH || h.kind == K_DMI_BULK || h.kind == K_ZEEMAN_U ||
                                 h.kind == K_DEMAG || h.kind == K_DMI_INT || h.kind == K_ANIS ||
                                 (h.kind == K_CUBIC && h.uniform);
                ok = uni && P.ord[i] > last;
                last = P.ord[i];
                sig |= 1u << P.ord[i];
                if (h.kind == K_DMI_INT && !h.uniform) sig |= SG_ID_CELL;
                if (h.kind == K_ANIS && !h.uniform) sig |= SG_KU_CELL;
            }
            P.sig = ok ? sig : 0u;
        }
                // -gamma/(1+alpha^2) rounded as the kernel computes it (llg.jl:18)
        volatile T al = (T)C->alpha;
        T den = T(1) + al * al;
        P.llg_a = -(T)C->gamma / den;
        return P;
    }

    // time of the next evaluation (a stage time inside the steppers); enters
    // only through the torque terms' ufun(t)
    double t_eval = 0;

    Dyn dyn() const
    {
        Dyn d{};
        // stochastic field factor, src/micro/field.jl:261 (dt = integrator.step)
        d.thermal_factor = (C->step > 0) ? 2 * C->alpha * KB / (C->volume * C->gamma * C->step) : 0.0;
        d.noise_step = (uint64_t)C->nsteps;
        for (size_t i = 0; i < C->terms.size(); i++) {
            const HostTerm &h = C->terms[i];
            const double u = h.ufun ? h.ufun(t_eval, h.uuser) : 1.0;
            if (h.kind == K_STT_SL) d.tfac[i] = u * h.s[7];    // ft = ufun(t)*beta, field.jl:351
            if (h.kind == K_STT_ZL) d.tfac[i] = u / C->gamma;  // ut = ufun(t)/gamma, field.jl:335
            if (h.kind == K_DMI_BULK && h.ufun) {              // Ds = (time_D(t) D)/d, field.jl:150-152
                d.tfac[i] = (double)((T)(u * h.s[6]) / (T)h.s[9]);
                d.tfy[i] = (double)((T)(u * h.s[7]) / (T)h.s[10]);
                d.tfz[i] = (double)((T)(u * h.s[8]) / (T)h.s[11]);
            }
            if (h.kind == K_DMI_BULK_S && h.ufun) d.tfac[i] = u;  // time_D(t) .* D, field.jl:185
            if (h.kind == K_ZEEMAN_U) d.tfac[i] = d.tfy[i] = d.tfz[i] = 1.0;  // the slot form scales by these
            if (h.kind == K_ZEEMAN_T) {                         // time_fun(t), field.jl:16-26
                double f[3] = {1, 1, 1};
                if (h.ft3) h.ft3(t_eval, f, h.uuser);
                d.tfac[i] = f[0];
                d.tfy[i] = f[1];
                d.tfz[i] = f[2];
            }
        }
        return d;
    }

    // device copy of the parameter block, refreshed when it changed
    const FieldParams<T> *params_dev()
    {
        FieldParams<T> P = params();
        if (!Pvalid || std::memcmp(&P, &Plast, sizeof P) != 0) {
            if (!Pdev.p) Pdev.alloc(1);
            HIPCHECK(hipStreamSynchronize(C->stream));
            HIPCHECK(hipMemcpy(Pdev.p, &P, sizeof P, hipMemcpyHostToDevice));
            Plast = P;
            Pvalid = true;
        }
        return Pdev.p;
    }

    uint32_t all_mask() const { return (uint32_t)((1ull << C->terms.size()) - 1); }

    // ---- profiling ---------------------------------------------------------
    void mark(int pass, bool begin)
    {
        if (!C->profile) return;
        if (begin) {
            Ev e;
            e.pass = pass;
            HIPCHECK(hipEventCreate(&e.a));
            HIPCHECK(hipEventCreate(&e.b));
            HIPCHECK(hipEventRecord(e.a, C->stream));
            evs.push_back(e);
        } else {
            HIPCHECK(hipEventRecord(evs.back().b, C->stream));
        }
    }
    void collect_profile()
    {
        HIPCHECK(hipStreamSynchronize(C->stream));
        for (auto &e : evs) {
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, e.a, e.b));
            prof_ms[e.pass] += ms;
            prof_n[e.pass] += 1;
            HIPCHECK(hipEventDestroy(e.a));
            HIPCHECK(hipEventDestroy(e.b));
        }
        evs.clear();
    }

    // Stage chaining: a caller sets nx_req right before an evaluation whose
    // yout is the input of the very next evaluation; the last pass then also
    // leaves the x-spectrum of mu0Ms*yout in Abuf (a_ready = yout), and that
    // next evaluation skips its first pass.  Both are cleared by every eval
    // and at every entry into the engine.
    bool nx_req = false;
    const T *a_ready = nullptr;

    // ---- one RHS evaluation with a fused epilogue; returns partial blocks --
    // add_torque terms: the state of this evaluation goes to the host, each
    // function writes its torque, which returns to the term's planes before the
    // passes read it (field.jl:282-293 calls torque_fun on the same stream order)
    std::vector<double> tfn_m;
    void run_torque_fns(const T *m)
    {
        for (size_t q = 0; q < C->terms.size(); q++) {
            HostTerm &h = C->terms[q];
            if (!h.tfn) continue;
            if (tfn_m.empty()) {
                tfn_m.resize(3 * N);
                download_aos(m, tfn_m.data());
            }
            h.tfn(t_eval, tfn_m.data(), h.tbuf.data(), h.uuser);
            upload_soa(h.tbuf.data(), const_cast<T *>(tA[q]));
            // torque_kernel! (kernels.jl:639-651) overwrites torque.field with
            // H = (1/gamma) m x T in place, so the reference's next call of the
            // function starts from that field: the host buffer follows it
            const T a = T(1) / (T)C->gamma;
            double *y = h.tbuf.data();
            const double *mm = tfn_m.data();
            for (int64_t i = 0; i < 3 * N; i += 3) {
                const T mx = (T)mm[i], my = (T)mm[i + 1], mz = (T)mm[i + 2];
                const T hx = (T)y[i], hy = (T)y[i + 1], hz = (T)y[i + 2];
                y[i] = (double)(a * (-mz * hy + my * hz));
                y[i + 1] = (double)(a * (mz * hx - mx * hz));
                y[i + 2] = (double)(a * (-my * hx + mx * hy));
            }
            HIPCHECK(hipStreamSynchronize(C->stream));
        }
        tfn_m.clear();
    }

    int eval(const T *m, const Epi<T> &E, bool generic = false)
    {
        if (any_tfn) run_torque_fns(m);
        const FieldParams<T> *P = params_dev();
        const Dyn d = dyn();
        const int fast = generic ? 0 : Plast.fast;
        const bool have_a = demag && a_ready == m;
        const bool want_nx = nx_req && demag && fast && (E.mode == EPI_STAGE || E.mode == EPI_RK4_FINAL);
        a_ready = nullptr;
        nx_req = false;
        int nb;
        if (!demag) {
            mark(MMX_PASS_LOCAL, true);
            nb = launch_local<T>(N, P, d, E, m, fast, Plast.sig, C->stream);
            mark(MMX_PASS_LOCAL, false);
        } else {
            const int nx = G.nx, ny = G.ny, nz = G.nz;
            if (!have_a) {
                mark(MMX_PASS_P1, true);
                launch_p1<T>(G, m, N, mu0Ms.p, nx, ny, nz, 3, Abuf.p, tw.p, C->stream);
                mark(MMX_PASS_P1, false);
            }
            mark(MMX_PASS_P2, true);
            launch_coly<T>(G, Abuf.p, Bbuf.p, ny, ny, G.Py, G.Py, nz, 3, 0, tw.p, C->stream);
            mark(MMX_PASS_P2, false);
            mark(MMX_PASS_P3, true);
            launch_convz<T>(G, Bbuf.p, Kt.p, Lline.p, nz, nz, 3, 0, tw.p, C->stream);
            mark(MMX_PASS_P3, false);
            mark(MMX_PASS_P4, true);
            launch_coly<T>(G, Bbuf.p, Abuf.p, G.Py, G.Py, ny, ny, nz, 3, 1, tw.p, C->stream);
            mark(MMX_PASS_P4, false);
            mark(MMX_PASS_P5, true);
            if (want_nx) {
                Epi<T> En = E;
                En.anext = Abuf.p;
                nb = launch_p5<T>(G, Abuf.p, P, d, En, m, fast, Plast.sig, tw.p, C->stream);
                a_ready = E.yout;
            } else {
                nb = launch_p5<T>(G, Abuf.p, P, d, E, m, fast, Plast.sig, tw.p, C->stream);
            }
            mark(MMX_PASS_P5, false);
        }
        HIPCHECK(hipGetLastError());
        return nb;
    }

    // fields and per-cell energies of nimg stacked images (three planes of N
    // each) through the five passes at once: the images' spectra lie side by
    // side in A and B (3*nimg components), so each pass is one launch.
    void eval_images(const T *Y, int nimg, T *F, T *ecell, cplx<T> *A, cplx<T> *B)
    {
        const int nx = G.nx, ny = G.ny, nz = G.nz;
        launch_p1<T>(G, Y, N, mu0Ms.p, nx, ny, nz, 3 * nimg, A, tw.p, C->stream);
        launch_coly<T>(G, A, B, ny, ny, G.Py, G.Py, nz, 3 * nimg, 0, tw.p, C->stream);
        launch_convz<T>(G, B, Kt.p, Lline.p, nz, nz, 3, 0, tw.p, C->stream, nimg);
        launch_coly<T>(G, B, A, G.Py, G.Py, ny, ny, nz, 3 * nimg, 1, tw.p, C->stream);
        Epi<T> x = epi(EPI_FIELD);
        x.out = F;
        x.e_out = ecell;
        launch_p5_images<T>(G, A, params_dev(), dyn(), x, Y, nimg, tw.p, C->stream);
        HIPCHECK(hipGetLastError());
    }

    double reduce_scalar(int nblocks, bool is_max)
    {
        launch_reduce(partial.p, nblocks, 1, 1, is_max ? 1 : 0, red.p, C->stream);
        HIPCHECK(hipGetLastError());
        double v = 0;
        HIPCHECK(hipMemcpyAsync(&v, red.p, sizeof(double), hipMemcpyDeviceToHost, C->stream));
        HIPCHECK(hipStreamSynchronize(C->stream));
        return v;
    }

    Epi<T> epi(int mode) const
    {
        Epi<T> E{};
        E.mode = mode;
        E.term_mask = all_mask();
        E.partial = partial.p;
        return E;
    }

    // ---- demag set-up -------------------------------------------------------
    static int next_pow2(int v)
    {
        int p = 1;
        while (p < v) p <<= 1;
        return p;
    }
    static int ilog2(int v)
    {
        int l = 0;
        while ((1 << l) < v) l++;
        return l;
    }

    void setup_demag(const std::vector<double> &rt)
    {
        const int nx = C->mesh.nx, ny = C->mesh.ny, nz = C->mesh.nz;
        G.nx = nx;
        G.ny = ny;
        G.nz = nz;
        G.Px = std::max(2, next_pow2(2 * nx - 1));
        G.Py = ny == 1 ? 1 : next_pow2(2 * ny - 1);
        G.Pz = nz == 1 ? 1 : next_pow2(2 * nz - 1);
        G.H = G.Px / 2;
        G.Kx = G.H + 1;
        G.KS = (G.Kx + 15) / 16 * 16;
        G.lgH = ilog2(G.H);
        G.lgPy = ilog2(G.Py);
        G.lgPz = ilog2(G.Pz);
        G.Lmax = std::max(G.Px, std::max(G.Py, G.Pz));
        G.lgLmax = ilog2(G.Lmax);
        G.Py2 = G.Py / 2 + 1;
        G.Pz2 = G.Pz / 2 + 1;
        G.ktw = convz_grouped_tensor(G, (int)sizeof(T));
        G.p5rx = p5_rows(G, 1);
        if (G.H > 2048 || G.Py > 2048 || G.Pz > 2048) {
            fail(MMX_EINVAL, "mesh %dx%dx%d exceeds the demag transform limits of this build", nx, ny, nz);
            throw HipError{MMX_EINVAL};
        }
        // FP64 twiddle table W_Lmax^t (long double sincos, rounded once)
        std::vector<cplx<double>> twd(G.Lmax);
        for (int t = 0; t < G.Lmax; t++) {
            long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)G.Lmax;
            twd[t] = {(double)cosl(a), (double)sinl(a)};
        }
        DevBuf<cplx<double>> twD;
        twD.alloc(G.Lmax);
        HIPCHECK(hipMemcpyAsync(twD.p, twd.data(), G.Lmax * sizeof(cplx<double>), hipMemcpyHostToDevice, C->stream));
        tw.alloc(G.Lmax);
        if (sizeof(T) == 8) {
            HIPCHECK(hipMemcpyAsync(tw.p, twD.p, G.Lmax * sizeof(cplx<double>), hipMemcpyDeviceToDevice, C->stream));
        } else {
            std::vector<cplx<T>> twf(G.Lmax);
            for (int t = 0; t < G.Lmax; t++) twf[t] = {(T)twd[t].x, (T)twd[t].y};
            HIPCHECK(hipMemcpyAsync(tw.p, twf.data(), G.Lmax * sizeof(cplx<T>), hipMemcpyHostToDevice, C->stream));
            HIPCHECK(hipStreamSynchronize(C->stream));
        }
        // k-space tensor: pad + 3-D r2c transform in FP64 on the device, keep the
        // real part folded to ky <= Py/2, kz <= Pz/2, scaled by -1/(mu0 Px Py Pz)
        // (irfft normalisation and the -1/mu0 of collect_h, demag.jl:441).
        const int64_t Ptot = (int64_t)G.Px * G.Py * G.Pz;
        const int64_t nA = (int64_t)G.Pz * G.Py * G.KS;
        DevBuf<double> rtD, pad;
        DevBuf<cplx<double>> At, Bt;
        rtD.alloc(6 * N);
        pad.alloc(Ptot);
        At.alloc(nA);
        Bt.alloc(nA);
        HIPCHECK(hipMemcpyAsync(rtD.p, rt.data(), 6 * N * sizeof(double), hipMemcpyHostToDevice, C->stream));
        Kt.alloc((int64_t)6 * G.Pz2 * G.Py2 * G.KS);
        HIPCHECK(hipMemsetAsync(Kt.p, 0, Kt.n * sizeof(T), C->stream));
        const double scale = -1.0 / (MU0 * (double)Ptot);
        for (int c = 0; c < 6; c++) {
            launch_tensor_pad(G, rtD.p, c, pad.p, C->stream);
            launch_p1<double>(G, pad.p, 0, nullptr, G.Px, G.Py, G.Pz, 1, At.p, twD.p, C->stream);
            launch_coly<double>(G, At.p, Bt.p, G.Py, G.Py, G.Py, G.Py, G.Pz, 1, 0, twD.p, C->stream);
            launch_convz<double>(G, Bt.p, nullptr, nullptr, G.Pz, G.Pz, 1, 1, twD.p, C->stream);
            launch_tensor_extract<T>(G, Bt.p, c, scale, Kt.p, C->stream);
            HIPCHECK(hipGetLastError());
        }
        HIPCHECK(hipStreamSynchronize(C->stream));
        // even cosine transforms of the odd components' on-axis lines (see
        // k_convz): L(k) = T(0) + 2 sum_{d>=1} T(d) cos(2 pi k d / P), scaled
        auto line_dft = [&](int comp, int n, int P, int nk, int64_t stride, double *out) {
            const double *t0 = rt.data() + (int64_t)comp * N;
            for (int kk = 0; kk < nk; kk++) {
                long double acc = 0;
                for (int d = 1; d < n; d++) {
                    long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)((int64_t)kk * d % P) / P;
                    acc += (long double)t0[d * stride] * cosl(a);
                }
                out[kk] = (double)((long double)t0[0] + 2.0L * acc) * scale;
            }
        };
        {
            std::vector<double> L(G.Pz2 + G.Py2 + G.Kx);
            line_dft(3, nz, G.Pz, G.Pz2, (int64_t)nx * ny, L.data());                // xy line along z
            line_dft(4, ny, G.Py, G.Py2, nx, L.data() + G.Pz2);                     // xz line along y
            line_dft(5, nx, G.Px, G.Kx, 1, L.data() + G.Pz2 + G.Py2);               // yz line along x
            DevBuf<double> Ld;  // staging sized for L (may exceed the 3N scratch on tiny meshes)
            Ld.alloc(L.size());
            HIPCHECK(hipMemcpyAsync(Ld.p, L.data(), L.size() * sizeof(double), hipMemcpyHostToDevice, C->stream));
            Lline.alloc(L.size());
            launch_cast<T>(Ld.p, Lline.p, (int64_t)L.size(), C->stream);
            HIPCHECK(hipGetLastError());
            HIPCHECK(hipStreamSynchronize(C->stream));
        }
        // the last pass writes one reduction slot group per block
        const int64_t slots = std::max<int64_t>(max_local_blocks(), p5_grid(G)) + 64;
        if ((int64_t)partial.n < slots) partial.alloc(slots);
        Abuf.alloc((int64_t)3 * nz * ny * G.KS);
        Bbuf.alloc((int64_t)3 * nz * G.Py * G.KS);
        HIPCHECK(hipMemsetAsync(Abuf.p, 0, Abuf.n * sizeof(cplx<T>), C->stream));
        HIPCHECK(hipMemsetAsync(Bbuf.p, 0, Bbuf.n * sizeof(cplx<T>), C->stream));
        HIPCHECK(hipStreamSynchronize(C->stream));
        demag = true;
    }

    // ---- integrators --------------------------------------------------------
    void swap_y(T *&other)
    {
        T *t = y;
        y = other;
        other = t;
    }

    // RungeKutta.advance_step: src/integrator/rk.jl:23-52.  rk4_into writes the
    // new state of y0 to out (out may be stg[1] but not stg[0] or y0).
    // chain_final: the next evaluation after this step is at `out`.
    void rk4_into(const T *y0, T *out, double h, bool chain_final = false)
    {
        ensure_k();
        T *a = stg[0], *b = stg[1];
        const double t0 = C->t;  // rk.jl:31-43: t, t + h/2, t + h/2, t + h
        Epi<T> E = epi(EPI_STAGE);
        E.y0 = y0;
        E.kstore = k[0];
        E.cself = 0.5 * h;
        E.yout = a;
        nx_req = true;
        t_eval = t0;
        eval(y0, E);
        E.kstore = k[1];
        E.yout = b;
        nx_req = true;
        t_eval = t0 + 0.5 * h;
        eval(a, E);
        E.kstore = k[2];
        E.cself = h;
        E.yout = a;
        nx_req = true;
        eval(b, E);
        Epi<T> F = epi(EPI_RK4_FINAL);
        F.y0 = y0;
        F.kp[0] = k[0];
        F.kp[1] = k[1];
        F.kp[2] = k[2];
        F.np = 3;
        F.cself = 1.0 / 6 * h;
        F.yout = out;
        F.normalize = 1;
        nx_req = chain_final;
        t_eval = t0 + h;
        eval(a, F);
    }

    void rk4_step(bool more = false)
    {
        const double h = C->step;
        rk4_into(y, stg[1], h, more);
        swap_y(stg[1]);  // new state in stg[1]'s buffer; the previous state stays in stg[1] (prespin)
        C->nfevals += 4;
        C->nsteps += 1;
        C->t = C->nsteps * h;
    }

    // ModifiedEuler.advance_step: src/integrator/heun.jl:28-49
    void heun_step(bool more = false)
    {
        ensure_k();
        const double h = C->step;
        T *a = stg[0], *b = stg[1];
        Epi<T> E = epi(EPI_STAGE);
        E.y0 = y;
        E.kstore = k[0];
        E.cself = h;
        E.yout = a;
        nx_req = true;
        t_eval = C->t;  // heun.jl:34-38: t, t + h
        eval(y, E);
        Epi<T> F = epi(EPI_STAGE);
        F.y0 = y;
        F.kp[0] = k[0];
        F.cp[0] = 0.5 * h;
        F.np = 1;
        F.cself = 0.5 * h;
        F.yout = b;
        F.normalize = 1;
        nx_req = more;
        t_eval = C->t + h;
        eval(a, F);
        swap_y(stg[1]);
        C->nfevals += 2;
        C->nsteps += 1;
        C->t = C->nsteps * h;
    }

    // rk_step_dopri54!: src/integrator/adaptive_rk_steps.jl:49-105; the new
    // state is left in stg[1]; returns the scaled error norm
    double dopri54_try(double step)
    {
        static const double a1 = 1.0 / 5;
        static const double a2[2] = {3.0 / 40, 9.0 / 40};
        static const double a3[3] = {44.0 / 45, -56.0 / 15, 32.0 / 9};
        static const double a4[4] = {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729};
        static const double a5[5] = {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656};
        static const double a6[6] = {35.0 / 384, 0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784, 11.0 / 84};
        static const double w[7] = {71.0 / 57600, 0, -71.0 / 16695, 71.0 / 1920, -17253.0 / 339200, 22.0 / 525,
                                    -1.0 / 40};
        // stage times t + c_i h (adaptive_rk_steps.jl:50,67-96)
        static const double c[6] = {1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1.0, 1.0};
        const double t0 = C->t;
        ensure_k();
        T *A = stg[0], *B = stg[1];
        Epi<T> E = epi(EPI_STAGE);
        E.y0 = y;
        // k1 -> y + a1 h k1
        E.kstore = k[0];
        E.np = 0;
        E.cself = a1 * step;
        E.yout = A;
        nx_req = true;
        t_eval = t0;
        eval(y, E);
        // k2
        E.kstore = k[1];
        E.kp[0] = k[0];
        E.cp[0] = a2[0] * step;
        E.np = 1;
        E.cself = a2[1] * step;
        E.yout = B;
        nx_req = true;
        t_eval = t0 + c[0] * step;
        eval(A, E);
        // k3
        E.kstore = k[2];
        E.kp[1] = k[1];
        E.cp[0] = a3[0] * step;
        E.cp[1] = a3[1] * step;
        E.np = 2;
        E.cself = a3[2] * step;
        E.yout = A;
        nx_req = true;
        t_eval = t0 + c[1] * step;
        eval(B, E);
        // k4
        E.kstore = k[3];
        E.kp[2] = k[2];
        for (int q = 0; q < 3; q++) E.cp[q] = a4[q] * step;
        E.np = 3;
        E.cself = a4[3] * step;
        E.yout = B;
        nx_req = true;
        t_eval = t0 + c[2] * step;
        eval(A, E);
        // k5
        E.kstore = k[4];
        E.kp[3] = k[3];
        for (int q = 0; q < 4; q++) E.cp[q] = a5[q] * step;
        E.np = 4;
        E.cself = a5[4] * step;
        E.yout = A;
        nx_req = true;
        t_eval = t0 + c[3] * step;
        eval(B, E);
        // k6 -> y_next = y + a61 k1 + a63 k3 + a64 k4 + a65 k5 + a66 k6, normalised
        E.kstore = k[5];
        E.kp[0] = k[0];
        E.kp[1] = k[2];
        E.kp[2] = k[3];
        E.kp[3] = k[4];
        E.cp[0] = a6[0] * step;
        E.cp[1] = a6[2] * step;
        E.cp[2] = a6[3] * step;
        E.cp[3] = a6[4] * step;
        E.np = 4;
        E.cself = a6[5] * step;
        E.yout = B;
        E.normalize = 1;
        nx_req = true;
        t_eval = t0 + step;
        eval(A, E);
        // k7 at y_next and the error combination (vector_add6b)
        Epi<T> R = epi(EPI_ERR);
        R.kstore = k[6];
        R.kp[0] = k[0];
        R.kp[1] = k[2];
        R.kp[2] = k[3];
        R.kp[3] = k[4];
        R.kp[4] = k[5];
        R.cp[0] = w[0];
        R.cp[1] = w[2];
        R.cp[2] = w[3];
        R.cp[3] = w[4];
        R.cp[4] = w[5];
        R.np = 5;
        R.cself = w[6];
        t_eval = t0 + step;
        int nb = eval(B, R);
        C->nfevals += 7;
        return reduce_scalar(nb, true) * step / C->tol;
    }

    // rk_step_bs23!: src/integrator/adaptive_rk_steps.jl:11-47 (FSAL: k1 <- k4)
    double bs23_try(double step, double t)
    {
        ensure_k();
        T *A = stg[0], *B = stg[1];
        Epi<T> E = epi(EPI_STAGE);
        E.y0 = y;
        if (C->succeed && t >= C->t) {
            std::swap(k[0], k[3]);
            // y_temp = y + (step*0.5) k1 without a new evaluation
            axpy(y, step * 0.5, k[0], A);
        } else {
            E.kstore = k[0];
            E.cself = step * 0.5;
            E.yout = A;
            nx_req = true;
            t_eval = t;  // adaptive_rk_steps.jl:22-37: t, t + step/2, t + 3 step/4, t + step
            eval(y, E);
            C->nfevals += 1;
        }
        // k2 -> y + (step*0.75) k2
        E.kstore = k[1];
        E.np = 0;
        E.cself = step * 0.75;
        E.yout = B;
        nx_req = true;
        t_eval = t + (1.0 / 2) * step;
        eval(A, E);
        // k3 -> y_next = y + step (2/9 k1 + 1/3 k2 + 4/9 k3), normalised
        Epi<T> S = epi(EPI_STAGE_SCALED);
        S.y0 = y;
        S.kstore = k[2];
        S.kp[0] = k[0];
        S.kp[1] = k[1];
        S.cp[0] = 2.0 / 9;
        S.cp[1] = 1.0 / 3;
        S.np = 2;
        S.cself = 4.0 / 9;
        S.outer = step;
        S.yout = A;
        S.normalize = 1;
        t_eval = t + (3.0 / 4) * step;
        eval(B, S);
        // k4 at y_next and the error combination
        Epi<T> R = epi(EPI_ERR);
        R.kstore = k[3];
        R.kp[0] = k[0];
        R.kp[1] = k[1];
        R.kp[2] = k[2];
        R.cp[0] = -(5.0 / 72);
        R.cp[1] = 1.0 / 12;
        R.cp[2] = 1.0 / 9;
        R.np = 3;
        R.cself = -1.0 / 8;
        t_eval = t + step;
        int nb = eval(A, R);
        C->nfevals += 3;
        // keep the candidate in stg[1] as the DOPRI path does
        std::swap(stg[0], stg[1]);
        return reduce_scalar(nb, true) * step / C->tol;
    }

    void axpy(const T *yv, double c, const T *kv, T *out);

    // compute_initial_step: src/integrator/adaptive_rk.jl:151-165
    double initial_step(double dt, int order_main)
    {
        ensure_k();
        double abs_step = dt, abs_step_tmp = dt;
        C->step = 1e-15;
        Epi<T> E = epi(EPI_RHS);
        E.out = k[6];
        t_eval = C->t;
        eval(y, E);
        double mx = maxabs(k[6]);
        double r_step = mx / (0.824 * std::pow(C->tol, 1.0 / order_main));
        C->nfevals += 1;
        if (abs_step * r_step > 0.001) abs_step_tmp = 0.001 / r_step;
        return std::min(abs_step, abs_step_tmp);
    }

    double maxabs(const T *v);

    // AdaptiveRK.advance_step: src/integrator/adaptive_rk.jl:93-148
    int adaptive_step()
    {
        const int max_nsteps = 50;
        const double t = C->t;
        int order_main = 5, order = 4;
        if (C->ikind == MMX_BS23) {
            order_main = 3;
            order = 2;
        }
        if (C->step_next <= 0) C->step_next = initial_step(1e-12, order_main);
        C->step = C->step_next;
        int nstep = 1;
        for (;;) {
            double max_error = C->ikind == MMX_BS23 ? bs23_try(C->step, t) : dopri54_try(C->step);
            if (std::isnan(max_error)) {
                C->step_next = 1e-14;
                return fail(MMX_ENAN, "NaN error detected, reducing step size");
            }
            nstep += 1;
            if (nstep > max_nsteps) return fail(MMX_ETRIES, "Too many inner steps (%d), system may not converge!", nstep);
            C->succeed = (max_error <= 1.0);
            if (C->succeed) {
                swap_y(stg[1]);
                C->nsteps += 1;
                C->t += C->step;
                double factor = 0.824 * std::pow(1.0 / max_error, 1.0 / (order + 1));
                C->step_next = C->step * std::min(5.0, std::max(0.2, factor));
                break;
            } else {
                double factor = 0.824 * std::pow(1.0 / max_error, 1.0 / order);
                C->step = C->step * std::min(5.0, std::max(0.2, factor));
            }
        }
        return MMX_OK;
    }


    // run_step(sim, ::SD): src/sd.jl:68-81 (field at t = 0, Barzilai-Borwein tau,
    // prespin = spin, update, normalise every 10 steps)
    DevBuf<T> sd_g;
    DevBuf<double> sd_dev;  // [sums(3), tau]
    void sd_step()
    {
        ensure_k();
        if (!sd_g.p) {
            sd_g.alloc(3 * N);
            HIPCHECK(hipMemsetAsync(sd_g.p, 0, 3 * N * sizeof(T), C->stream));
            sd_dev.alloc(8);
        }
        T *h = k[6];
        Epi<T> E = epi(EPI_FIELD);
        E.out = h;
        const double save_step = C->step;
        C->step = 0;  // the stochastic field factor needs a time step; SD has none
        t_eval = 0;
        eval(y, E);
        C->step = save_step;
        if (C->sd_steps == 0) {  // compute_gk_kernel_1!, tau = min_tau
            launch_sd_g<T>(N, y, nullptr, h, sd_g.p, partial.p, C->stream);
            const double t0 = C->sd_min_tau;
            HIPCHECK(hipMemcpyAsync(sd_dev.p + 3, &t0, sizeof(double), hipMemcpyHostToDevice, C->stream));
        } else {
            const int g = launch_sd_g<T>(N, y, stg[1], h, sd_g.p, partial.p, C->stream);
            launch_reduce(partial.p, g, 3, 3, 0, sd_dev.p, C->stream);
            launch_sd_tau(sd_dev.p, (int)(C->sd_steps % 2), C->sd_min_tau, C->sd_max_tau, sd_dev.p + 3, C->stream);
        }
        launch_sd_update<T>(N, params_dev(), sd_g.p, y, stg[1], h, sd_dev.p + 3, C->stream);
        C->sd_steps += 1;
        if (C->sd_steps % 10 == 0) launch_normalise<T>(N, y, C->stream);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpyAsync(&C->sd_tau, sd_dev.p + 3, sizeof(double), hipMemcpyDeviceToHost, C->stream));
        HIPCHECK(hipStreamSynchronize(C->stream));
        C->nfevals += 1;
        C->nsteps += 1;
    }

    int advance(int64_t n)
    {
        for (int64_t i = 0; i < n; i++) {
            if (C->ikind == MMX_RK4) rk4_step(i + 1 < n);
            else if (C->ikind == MMX_HEUN) heun_step(i + 1 < n);
            else if (C->ikind == MMX_SD) sd_step();
            else {
                int rc = adaptive_step();
                if (rc != MMX_OK) return rc;
            }
        }
        return MMX_OK;
    }


    // ---- differentiation form: RK4 trajectory pull-back ----------------------
    // Checkpoints every `every` steps; the backward pass recomputes each
    // segment's states, then per step the four stages with their fields, and
    // pulls the cotangent back through the final combination, the
    // normalisation and each stage's RHS (EPI_ADJ pass).
    struct AdjState {
        bool recorded = false;
        int64_t nsteps = 0, every = 0, step0 = 0;
        double h = 0, t0 = 0;
        DevBuf<T> ck, seg, work, gms;
        std::vector<std::unique_ptr<DevBuf<T>>> g;
    } adj;

    static bool has_param_grad(int kind)
    {
        return kind == K_ANIS || kind == K_ANIS_S || kind == K_CUBIC || kind == K_DMI_INT || kind == K_EXCH_S ||
               kind == K_DMI_BULK || kind == K_DMI_BULK_S;
    }
    // bulk DMI gradients hold the three components (Dx, Dy, Dz planes)
    static int64_t param_grad_planes(int kind) { return kind == K_DMI_BULK || kind == K_DMI_BULK_S ? 3 : 1; }

    int adj_record(int64_t nsteps, int64_t every)
    {
        if (C->ikind != MMX_RK4) return fail(MMX_ESTATE, "the adjoint needs the RungeKutta (RK4) integrator");
        if (alphas.p) return fail(MMX_ESTATE, "the adjoint takes a uniform alpha");
        if (!(C->step > 0)) return fail(MMX_ESTATE, "the RK4 step is not set");
        for (const auto &h : C->terms)
            if (h.kind == K_SOT || h.kind == K_STT_SL || h.kind == K_STT_ZL || h.kind == K_ZEEMAN_T ||
                h.kind == K_HEX || h.kind == K_EXCH_IL || h.kind == K_DMI_IL || h.kind == K_TWIN || h.kind == K_SAHE || h.kind == K_MEL_T || h.kind == K_MEL_C ||
                h.kind == K_TORQUE_FN || h.ufun)
                return fail(MMX_EINVAL, "the adjoint covers the time-independent exchange, DMI, uniaxial and cubic "
                                        "anisotropy, Zeeman and demag terms");
        if (nsteps < 1 || every < 1) return fail(MMX_EINVAL, "nsteps and every must be >= 1");
        const int64_t nck = (nsteps + every - 1) / every;
        adj.recorded = false;
        adj.ck.alloc((size_t)nck * 3 * N);
        adj.step0 = C->nsteps;
        adj.t0 = C->t;
        adj.h = C->step;
        for (int64_t s = 0; s < nsteps; s++) {
            if (s % every == 0)
                HIPCHECK(hipMemcpyAsync(adj.ck.p + (s / every) * 3 * N, y, 3 * N * sizeof(T), hipMemcpyDeviceToDevice,
                                        C->stream));
            rk4_step();
        }
        HIPCHECK(hipStreamSynchronize(C->stream));
        adj.nsteps = nsteps;
        adj.every = every;
        adj.recorded = true;
        return MMX_OK;
    }

    int adj_backward(const double *lamT, double *grad_m0)
    {
        if (!adj.recorded) return fail(MMX_ESTATE, "no recorded trajectory (call mmx_adjoint_record first)");
        const int64_t V = 3 * N;
        const double h = adj.h;
        // work vectors: lam, lam0, ut, Y2..Y4, h1..h4, k1..k4, a1..a4
        adj.work.alloc((size_t)18 * V);
        HIPCHECK(hipMemsetAsync(adj.work.p, 0, adj.work.n * sizeof(T), C->stream));
        T *w = adj.work.p;
        T *lam = w, *lam0 = w + V, *U = w + 2 * V;
        T *Ys[3] = {w + 3 * V, w + 4 * V, w + 5 * V};
        T *hs[4], *ks[4], *as[4];
        for (int q = 0; q < 4; q++) {
            hs[q] = w + (6 + q) * V;
            ks[q] = w + (10 + q) * V;
            as[q] = w + (14 + q) * V;
        }
        adj.seg.alloc((size_t)adj.every * V);
        adj.g.clear();
        adj.g.resize(C->terms.size());
        for (size_t tt = 0; tt < C->terms.size(); tt++)
            if (has_param_grad(C->terms[tt].kind)) {
                const int64_t n = param_grad_planes(C->terms[tt].kind) * N;
                adj.g[tt] = std::make_unique<DevBuf<T>>();
                adj.g[tt]->alloc(n);
                HIPCHECK(hipMemsetAsync(adj.g[tt]->p, 0, n * sizeof(T), C->stream));
            }
        adj.gms.alloc(N);
        HIPCHECK(hipMemsetAsync(adj.gms.p, 0, N * sizeof(T), C->stream));
        upload_soa(lamT, lam);
        const int64_t save_n = C->nsteps;
        const double save_t = C->t;
        const int64_t nseg = (adj.nsteps + adj.every - 1) / adj.every;
        for (int64_t sg = nseg - 1; sg >= 0; sg--) {
            const int64_t s0 = sg * adj.every;
            const int64_t len = std::min(adj.every, adj.nsteps - s0);
            T *seg = adj.seg.p;
            HIPCHECK(hipMemcpyAsync(seg, adj.ck.p + sg * V, V * sizeof(T), hipMemcpyDeviceToDevice, C->stream));
            for (int64_t n = 0; n + 1 < len; n++) {
                C->nsteps = adj.step0 + s0 + n;
                rk4_into(seg + n * V, seg + (n + 1) * V, h);
            }
            for (int64_t n = len - 1; n >= 0; n--) {
                const T *y0 = seg + n * V;
                C->nsteps = adj.step0 + s0 + n;
                // stages with their fields
                Epi<T> E = epi(EPI_STAGE);
                E.y0 = y0;
                E.cself = 0.5 * h;
                E.kstore = ks[0];
                E.out = hs[0];
                E.yout = Ys[0];
                eval(y0, E, true);
                E.kstore = ks[1];
                E.out = hs[1];
                E.yout = Ys[1];
                eval(Ys[0], E, true);
                E.cself = h;
                E.kstore = ks[2];
                E.out = hs[2];
                E.yout = Ys[2];
                eval(Ys[1], E, true);
                E.cself = 0;
                E.kstore = ks[3];
                E.out = hs[3];
                E.yout = U;
                eval(Ys[2], E, true);
                launch_adj_final<T>(N, y0, ks, 1.0 / 6 * h, lam, lam0, as, C->stream);
                HIPCHECK(hipGetLastError());
                const T *Yst[4] = {y0, Ys[0], Ys[1], Ys[2]};
                const FieldParams<T> *P = params_dev();
                for (int s = 3; s >= 0; s--) {
                    launch_adj_pre<T>(N, P, Yst[s], as[s], U, C->stream);
                    HIPCHECK(hipGetLastError());
                    Epi<T> A = epi(EPI_ADJ);
                    A.adj.Y = Yst[s];
                    A.adj.hY = hs[s];
                    A.adj.a = as[s];
                    A.adj.lam = lam0;
                    A.adj.next = s > 0 ? as[s - 1] : nullptr;
                    A.adj.cnext = s == 3 ? h : 0.5 * h;
                    for (size_t tt = 0; tt < adj.g.size(); tt++) A.adj.g[tt] = adj.g[tt] ? adj.g[tt]->p : nullptr;
                    A.adj.gms = adj.gms.p;
                    eval(U, A, true);
                }
                std::swap(lam, lam0);
            }
        }
        C->nsteps = save_n;
        C->t = save_t;
        if (grad_m0) download_aos(lam, grad_m0);
        HIPCHECK(hipStreamSynchronize(C->stream));
        return MMX_OK;
    }

    int adj_param_grad(int term, double *out)
    {
        if (term < 0 || term >= (int)adj.g.size() || !adj.g[term])
            return fail(MMX_EINVAL, "no parameter gradient for term %d (run mmx_adjoint_backward; supported: "
                                    "anisotropy Ku, cubic Kc, interfacial and bulk DMI D, spatial exchange A)", term);
        if (param_grad_planes(C->terms[term].kind) == 3) download_aos(adj.g[term]->p, out);
        else download_scalar(adj.g[term]->p, out, N);
        return MMX_OK;
    }

    // dL/dMs per cell: the mu0Ms gradient times mu0 (mu0Ms = mu0 Ms, src/sim.jl:138)
    int adj_ms_grad(double *out)
    {
        if (!adj.gms.p) return fail(MMX_ESTATE, "no gradient yet (run mmx_adjoint_backward)");
        download_scalar(adj.gms.p, out, N);
        for (int64_t i = 0; i < N; i++) out[i] *= MU0;
        return MMX_OK;
    }

    // ---- one RHS evaluation's Jacobian products -----------------------------
    // JVP: J v = (dF/dm) v + (dF/dh) H'(v), H' the field's Jacobian at m (the
    // terms linear in m evaluated on v, plus the cubic term linearised at m);
    // VJP: J^T a through the adjoint pass.  Both at the state m (3N AoS, or the
    // current state when null) and time t; the terms the adjoint covers only.
    DevBuf<T> jbuf[5];
    int jac_check() const
    {
        if (alphas.p) return fail(MMX_ESTATE, "the Jacobian products take a uniform alpha");
        for (const auto &h : C->terms)
            if (h.kind == K_SOT || h.kind == K_STT_SL || h.kind == K_STT_ZL || h.kind == K_ZEEMAN_T || h.kind == K_HEX ||
                h.kind == K_EXCH_IL || h.kind == K_DMI_IL || h.kind == K_TWIN || h.kind == K_SAHE || h.kind == K_MEL_T ||
                h.kind == K_MEL_C || h.kind == K_TORQUE_FN || h.kind == K_THERMAL || h.ufun)
                return fail(MMX_EINVAL, "the Jacobian products cover the exchange, DMI, uniaxial and cubic anisotropy, "
                                        "static Zeeman and demag terms");
        return MMX_OK;
    }
    void jac_setup(const double *m)
    {
        for (auto &b : jbuf)
            if (!b.p) b.alloc(3 * N);
        if (m) upload_soa(m, jbuf[0].p);
        else HIPCHECK(hipMemcpyAsync(jbuf[0].p, y, 3 * N * sizeof(T), hipMemcpyDeviceToDevice, C->stream));
    }
    void field_into(const T *src, T *out, uint32_t mask)
    {
        Epi<T> E = epi(EPI_FIELD);
        E.term_mask = mask;
        E.out = out;
        eval(src, E, true);
    }
    int rhs_jvp(double t, const double *m, const double *v, double *out)
    {
        if (jac_check() != MMX_OK) return MMX_EINVAL;
        jac_setup(m);
        T *Y = jbuf[0].p, *V = jbuf[1].p, *hY = jbuf[2].p, *hv = jbuf[3].p;
        upload_soa(v, V);
        t_eval = t;
        field_into(Y, hY, all_mask());
        uint32_t lin = 0;  // terms linear in m (Zeeman is constant, the cubic term is linearised in k_llg_jvp)
        for (size_t q = 0; q < C->terms.size(); q++) {
            const int kd = C->terms[q].kind;
            if (kd != K_ZEEMAN && kd != K_ZEEMAN_U && kd != K_CUBIC) lin |= 1u << q;
        }
        field_into(V, hv, lin);
        launch_llg_jvp<T>(N, params_dev(), Y, hY, V, hv, hv, C->stream);
        HIPCHECK(hipGetLastError());
        download_aos(hv, out);
        return MMX_OK;
    }
    int rhs_vjp(double t, const double *m, const double *a, double *out)
    {
        if (jac_check() != MMX_OK) return MMX_EINVAL;
        jac_setup(m);
        T *Y = jbuf[0].p, *Ab = jbuf[1].p, *hY = jbuf[2].p, *U = jbuf[3].p, *lam = jbuf[4].p;
        upload_soa(a, Ab);
        t_eval = t;
        field_into(Y, hY, all_mask());
        const FieldParams<T> *P = params_dev();
        launch_adj_pre<T>(N, P, Y, Ab, U, C->stream);
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemsetAsync(lam, 0, 3 * N * sizeof(T), C->stream));
        Epi<T> A = epi(EPI_ADJ);
        A.adj.Y = Y;
        A.adj.hY = hY;
        A.adj.a = Ab;
        A.adj.lam = lam;
        eval(U, A, true);
        download_aos(lam, out);
        return MMX_OK;
    }

    // per-term energies of the current state into out_dev (device, nterms
    // values) without allocating: block partials in `part` (slots device
    // doubles), one term after the other, each reduced on the stream
    int energies_noalloc(double t, double *out_dev, double *part, int slots)
    {
        if (any_tfn) return fail(MMX_EINVAL, "the allocation-free energies take no add_torque term");
        const int nb = demag ? p5_grid(G) : std::min<int64_t>(std::max<int64_t>((N + BLOCK - 1) / BLOCK, 1),
                                                                max_local_blocks());
        if (nb > slots)
            return fail(MMX_EINVAL, "this mesh needs %d partial sums per term; the allocation-free path holds %d",
                        nb, slots);
        t_eval = t;
        for (int q = 0; q < (int)C->terms.size(); q++) {
            Epi<T> E = epi(EPI_ENERGY);
            E.term_mask = 1u << q;
            E.partial = part;
            const int got = eval(y, E);
            if (got != nb) return fail(MMX_ESTATE, "internal: %d partial sums written, %d expected", got, nb);
            launch_reduce(part, nb, 1, 1, 0, out_dev + q, C->stream);
            HIPCHECK(hipGetLastError());
        }
        return MMX_OK;
    }

    double max_dm()
    {
        int nb = launch_maxdiff<T>(stg[1], y, N, partial.p, C->stream);
        HIPCHECK(hipGetLastError());
        return reduce_scalar(nb, true);
    }
};

template <class T>
void Engine<T>::axpy(const T *yv, double c, const T *kv, T *out)
{
    launch_axpy<T>(yv, c, kv, out, 3 * N, C->stream);
    HIPCHECK(hipGetLastError());
}

template <class T>
double Engine<T>::maxabs(const T *v)
{
    int g = launch_maxabs<T>(v, 3 * N, partial.p, C->stream);
    HIPCHECK(hipGetLastError());
    return reduce_scalar(g, true);
}



// ---------------------------------------------------------------------------
// Nudged elastic band with the LLG driver (src/neb/neb.jl).  The band's free
// images live image-major on the device; the field and energy of an image
// come from the simulation's own passes (any terms, demag included).  As in
// the reference, the field, energies and distances used by every stage of a
// step are those of the accepted images (neb.jl:196-230 read neb.spin), while
// tangents and the LLG right-hand side use the stage vector; they are
// computed once per accepted step.  Integrators: BS23 (default) or DOPRI54
// is not offered; BS23 as neb.jl:146-152 sets it (tol 1e-5, alpha 0.2,
// precession off).
// ---------------------------------------------------------------------------
template <class T>
struct NebEngine : CtxBase {
    mmx_neb *Nb;
    mmx_ctx *C;
    Engine<T> &e;
    int n;          // free images
    int64_t N, V;   // cells, 3N
    int nb;         // blocks per image of the band kernels
    DevBuf<T> L, R, Y, Ynext, Ytmp, F, tan, ecell, prespin;
    DevBuf<T> kb[4];
    T *k[4];
    DevBuf<double> energy, dist, coef, part;
    DevBuf<FieldParams<T>> Pneb;  // the sim's block with the band's LLG parameters
    DevBuf<cplx<T>> An, Bn;       // demag spectra of every free image (eval_images)
    DevBuf<T> ecell_n;            // per-cell energies of every free image
    double spring = 1e5;
    int climb = -1;
    double alpha = 0.2, gamma = 2.21e5;
    int precession = 0;
    double tol = 1e-5;
    // AdaptiveRK state (adaptive_rk.jl:37-55)
    double t = 0, step = 0, step_next = 0;
    int64_t nsteps = 0, nfevals = 0;
    bool succeed = false, have_k4 = false, fresh = false;

    NebEngine(mmx_neb *nb_, mmx_ctx *c, int nfree)
        : Nb(nb_), C(c), e(*static_cast<Engine<T> *>(c->impl.get())), n(nfree), N(c->N), V(3 * c->N)
    {
        nb = (int)std::min<int64_t>((N + BLOCK - 1) / BLOCK, 512);
        L.alloc(V);
        R.alloc(V);
        for (DevBuf<T> *b : {&Y, &Ynext, &Ytmp, &F, &tan, &prespin}) b->alloc((size_t)n * V);
        for (int i = 0; i < 4; i++) {
            kb[i].alloc((size_t)n * V);
            k[i] = kb[i].p;
        }
        ecell.alloc(N);
        energy.alloc(n + 2);
        dist.alloc(n + 1);
        coef.alloc(2 * n);
        part.alloc((size_t)std::max<int64_t>((int64_t)(n + 1) * nb * 2, max_local_blocks()) + 64);
    }

    hipStream_t st() const { return C->stream; }

    // field and energy of one image (effective_field_energy(sim, m), field.jl:538-547)
    void field_energy(const T *m, T *out, double *E)
    {
        e.t_eval = 0.0;
        Epi<T> x = e.epi(EPI_FIELD);
        x.out = out ? out : Ytmp.p;  // the field of an end image is not kept
        x.e_out = ecell.p;
        e.eval(m, x);
        int g = launch_sum<T>(ecell.p, N, part.p, st());
        launch_reduce(part.p, g, 1, 1, 0, E, st());
        HIPCHECK(hipGetLastError());
    }

    void set_images(const double *aos)  // (n+2) x 3N AoS: left, free images, right
    {
        e.upload_soa(aos, L.p);
        for (int q = 0; q < n; q++) e.upload_soa(aos + (size_t)(q + 1) * V, Y.p + (size_t)q * V);
        e.upload_soa(aos + (size_t)(n + 1) * V, R.p);
        HIPCHECK(hipMemcpyAsync(prespin.p, Y.p, (size_t)n * V * sizeof(T), hipMemcpyDeviceToDevice, st()));
        field_energy(L.p, nullptr, energy.p);
        field_energy(R.p, nullptr, energy.p + n + 1);
        HIPCHECK(hipStreamSynchronize(st()));
        fresh = false;
    }

    // fields, energies and distances of the accepted images
    void refresh()
    {
        if (fresh) return;
        e.t_eval = 0.0;  // effective_field_energy(sim, m, 0.0), neb.jl:227
        if (!e.demag) {  // every image in one launch
            launch_neb_field<T>(e.params_dev(), e.dyn(), e.all_mask(), N, n, nb, Y.p, F.p, part.p, energy.p + 1,
                                st());
            HIPCHECK(hipGetLastError());
        } else {
            if (!An.p) {
                An.alloc((size_t)n * e.Abuf.n);
                Bn.alloc((size_t)n * e.Bbuf.n);
                ecell_n.alloc((size_t)n * N);
                HIPCHECK(hipMemsetAsync(An.p, 0, An.n * sizeof(cplx<T>), st()));
                HIPCHECK(hipMemsetAsync(Bn.p, 0, Bn.n * sizeof(cplx<T>), st()));
            }
            e.eval_images(Y.p, n, F.p, ecell_n.p, An.p, Bn.p);
            launch_sum_images<T>(ecell_n.p, N, n, nb, part.p, energy.p + 1, st());
            HIPCHECK(hipGetLastError());
        }
        launch_neb_dist<T>(N, n, Y.p, L.p, R.p, nb, part.p, dist.p, st());
        HIPCHECK(hipGetLastError());
        fresh = true;
    }

    // one right-hand side at the stage vector Ys with the stage combination
    void rhs(const T *Ys, int mode, const double c[4], double outer, T *kout, const T *y0, const T *a1,
             const T *a2, const T *a3, T *yout)
    {
        const FieldParams<T> *P = Pneb.p;
        launch_neb_tangent<T>(N, n, Ys, L.p, R.p, energy.p, F.p, tan.p, nb, part.p, dist.p, spring, climb,
                              coef.p, st());
        launch_neb_stage<T>(P, N, n, nb, Ys, F.p, tan.p, coef.p, mode, c, outer, kout, y0, a1, a2, a3, yout,
                            part.p, st());
        HIPCHECK(hipGetLastError());
        nfevals += 1;
    }

    double reduce_max()
    {
        launch_reduce(part.p, n * nb, 1, 1, 1, e.red.p, st());
        HIPCHECK(hipGetLastError());
        double v = 0;
        HIPCHECK(hipMemcpyAsync(&v, e.red.p, sizeof(double), hipMemcpyDeviceToHost, st()));
        HIPCHECK(hipStreamSynchronize(st()));
        return v;
    }

    // rk_step_bs23!: src/integrator/adaptive_rk_steps.jl:11-47
    double bs23_try(double h, double t0)
    {
        const double z[4] = {0, 0, 0, 0};
        if (succeed && t0 >= t && have_k4) {
            std::swap(k[0], k[3]);  // k1 .= k4
            launch_axpy<T>(Y.p, h * 0.5, k[0], Ytmp.p, (int64_t)n * V, st());
        } else {
            double c[4] = {h * 0.5, 0, 0, 0};
            rhs(Y.p, 0, c, 0, k[0], Y.p, nullptr, nullptr, nullptr, Ytmp.p);
        }
        {
            double c[4] = {h * 0.75, 0, 0, 0};
            rhs(Ytmp.p, 0, c, 0, k[1], Y.p, nullptr, nullptr, nullptr, Ynext.p);
        }
        std::swap(Ytmp.p, Ynext.p);  // y_temp now holds y + 3/4 h k2
        {
            double c[4] = {2.0 / 9, 1.0 / 3, 4.0 / 9, 0};
            rhs(Ytmp.p, 1, c, h, k[2], Y.p, k[0], k[1], nullptr, Ynext.p);
        }
        {
            double c[4] = {-(5.0 / 72), 1.0 / 12, 1.0 / 9, -1.0 / 8};
            rhs(Ynext.p, 2, c, 0, k[3], nullptr, k[0], k[1], k[2], nullptr);
        }
        (void)z;
        return reduce_max() * h / tol;
    }

    double initial_step(double dt)  // compute_initial_step, adaptive_rk.jl:151-165 (BS23: order 3)
    {
        double abs_step = dt, tmp = dt;
        const double c[4] = {0, 0, 0, 0};
        rhs(Y.p, 3, c, 0, Ytmp.p, nullptr, nullptr, nullptr, nullptr, nullptr);
        const double r = reduce_max() / (0.824 * std::pow(tol, 1.0 / 3));
        if (abs_step * r > 0.001) tmp = 0.001 / r;
        return std::min(abs_step, tmp);
    }

    // the LLG parameters of the band (neb.jl:146-152) in a parameter block of its own
    void upload_params()
    {
        FieldParams<T> P = e.params();
        P.alpha = (T)alpha;
        P.gamma = (T)gamma;
        P.precession = precession;
        if (!Pneb.p) Pneb.alloc(1);
        HIPCHECK(hipStreamSynchronize(st()));
        HIPCHECK(hipMemcpy(Pneb.p, &P, sizeof P, hipMemcpyHostToDevice));
    }

    int advance_step()  // adaptive_rk.jl:93-148
    {
        refresh();
        const double t0 = t;
        if (step_next <= 0) step_next = initial_step(1e-12);
        step = step_next;
        int nstep = 1;
        for (;;) {
            const double err = bs23_try(step, t0);
            if (std::isnan(err)) {
                step_next = 1e-14;
                return fail(MMX_ENAN, "NaN error detected, reducing step size");
            }
            nstep += 1;
            if (nstep > 50) return fail(MMX_ETRIES, "Too many inner steps (%d), system may not converge!", nstep);
            succeed = err <= 1.0;
            if (succeed) {
                std::swap(Y.p, Ynext.p);
                have_k4 = true;
                nsteps += 1;
                t += step;
                const double factor = 0.824 * std::pow(1.0 / err, 1.0 / 3);
                step_next = step * std::min(5.0, std::max(0.2, factor));
                fresh = false;
                break;
            }
            const double factor = 0.824 * std::pow(1.0 / err, 1.0 / 2);
            step = step * std::min(5.0, std::max(0.2, factor));
        }
        return MMX_OK;
    }

    double max_dm_vs_prespin()
    {
        launch_neb_maxdiff<T>(prespin.p, Y.p, N, n, nb, part.p, e.red.p, st());
        HIPCHECK(hipGetLastError());
        double mx = 0;
        HIPCHECK(hipMemcpyAsync(&mx, e.red.p, sizeof(double), hipMemcpyDeviceToHost, st()));
        HIPCHECK(hipStreamSynchronize(st()));
        return mx;
    }

    // relax(::NEB) with the LLG driver: src/neb/neb.jl:231-296 (no output)
    int relax(int64_t max_steps, double stopping_dmdt, int64_t *steps_out, double *max_dmdt_out)
    {
        const double dmdt_factor = (2 * M_PI / 360) * 1e9;
        double md = 0;
        int64_t i = 0;
        for (i = 0; i <= max_steps; i++) {
            int rc = advance_step();
            if (rc != MMX_OK) return rc;
            md = max_dm_vs_prespin() / step;
            if (md < stopping_dmdt * dmdt_factor) break;
        }
        if (steps_out) *steps_out = std::min(i, max_steps);
        if (max_dmdt_out) *max_dmdt_out = md;
        return MMX_OK;
    }
};

// ---------------------------------------------------------------------------
// Batches: many independent systems (ensemble members, parameter sweeps) of
// any sizes advanced together.  Each stage of a step is ONE launch over every
// member without demag (k_batch_local); members with demag are grouped by
// geometry (mesh, cells, periodicity, demag images: one tensor), and each
// group takes one launch per demag pass and stage, its members' spectra
// stacked side by side (the y and z passes over 3 x members components, the
// last pass k_p5_batch with each member's own parameters and epilogue).  Two
// steps are captured in a HIP graph and replayed.
// The state of member m alternates between its y buffer (A) and its stg[1]
// buffer (B); an odd remainder step is copied back so the state always ends
// in A.  The launch shape (cells per thread, grid) is picked at creation by
// timing the candidates on the batch itself.
// ---------------------------------------------------------------------------
template <class T>
struct BatchEngine : CtxBase {
    mmx_batch *B;
    std::vector<Engine<T> *> E;
    int M = 0, kind = MMX_RK4, S = 4;
    DevBuf<int2> items;
    int nitems = 0;
    DevBuf<const FieldParams<T> *> Ps;
    DevBuf<BatchStage<T>> st;  // [parity][stage][member]
    DevBuf<Dyn> dyns;
    int V[4] = {0, 0, 0, 0};
    int cpt = 1, bs = 256, grid = 2048;
    std::vector<float> tune_ms;
    std::vector<T *> bufA, bufB;  // the two state buffers of each member
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    // members with demag, one group per geometry
    struct Group {
        std::vector<int> mem;  // member indices
        DevBuf<int> gidx;
        DevBuf<cplx<T>> A, Bs;  // stacked spectra (A: P1/P4 layout, Bs: P2/P3 layout)
        int64_t astride = 0;
        bool fast = true;       // slot-form last pass: chains the next spectrum, P1 once per advance
        int vg[4] = {0, 0, 0, 0};
    };
    std::vector<std::unique_ptr<Group>> groups;
    std::vector<int> local;                   // members without demag
    std::vector<BatchStage<T>> hst;           // host copy of the stage descriptors

    explicit BatchEngine(mmx_batch *b) : B(b) {}
    ~BatchEngine() override
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
    }

    static int stages(int k) { return k == MMX_HEUN ? 2 : 4; }

    void build()
    {
        M = (int)B->members.size();
        kind = B->members[0]->ikind;
        S = stages(kind);
        for (auto *c : B->members) E.push_back(static_cast<Engine<T> *>(c->impl.get()));
        for (int m = 0; m < M; m++) {
            bufA.push_back(E[m]->y);
            bufB.push_back(E[m]->stg[1]);
        }
        // group the members with demag by geometry
        for (int m = 0; m < M; m++) {
            mmx_ctx *c = B->members[m];
            if (!E[m]->demag) {
                local.push_back(m);
                continue;
            }
            Group *g = nullptr;
            for (auto &q : groups) {
                const mmx_ctx *c0 = B->members[q->mem[0]];
                const mmx_mesh_desc &a = c0->mesh, &b = c->mesh;
                if (a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.dx == b.dx && a.dy == b.dy && a.dz == b.dz &&
                    a.pbc == b.pbc &&
                    c0->demag_img[0] == c->demag_img[0] && c0->demag_img[1] == c->demag_img[1] &&
                    c0->demag_img[2] == c->demag_img[2]) {
                    g = q.get();
                    break;
                }
            }
            if (!g) {
                groups.push_back(std::make_unique<Group>());
                g = groups.back().get();
            }
            g->mem.push_back(m);
        }
        std::vector<int2> it;
        items_for(256, 4, it);
        std::vector<const FieldParams<T> *> ps(M);
        bool fast = true;
        for (int m = 0; m < M; m++) {
            E[m]->ensure_k();
            ps[m] = E[m]->params_dev();
        }
        for (int m : local) fast = fast && E[m]->Plast.fast;
        for (auto &g : groups) {
            const Engine<T> &e0 = *E[g->mem[0]];
            const int n = (int)g->mem.size();
            for (int m : g->mem) g->fast = g->fast && E[m]->Plast.fast;
            g->astride = e0.Abuf.n;
            g->A.alloc((size_t)n * e0.Abuf.n);
            g->Bs.alloc((size_t)n * e0.Bbuf.n);
            // padded entries that the passes read but never write stay zero
            HIPCHECK(hipMemsetAsync(g->A.p, 0, g->A.n * sizeof(cplx<T>), B->stream));
            HIPCHECK(hipMemsetAsync(g->Bs.p, 0, g->Bs.n * sizeof(cplx<T>), B->stream));
            g->gidx.alloc(n);
            HIPCHECK(hipMemcpy(g->gidx.p, g->mem.data(), n * sizeof(int), hipMemcpyHostToDevice));
        }
        Ps.alloc(M);
        HIPCHECK(hipMemcpy(Ps.p, ps.data(), M * sizeof(ps[0]), hipMemcpyHostToDevice));
        // stage descriptors for both parities
        std::vector<BatchStage<T>> h(2 * S * M);
        for (int par = 0; par < 2; par++)
            for (int m = 0; m < M; m++) {
                Engine<T> &e = *E[m];
                const double dt = e.C->step;
                T *A = e.y, *Bb = e.stg[1];
                const T *yin = par ? Bb : A;
                T *out = par ? A : Bb;
                T *a = e.stg[0];
                BatchStage<T> *d = &h[(size_t)par * S * M];
                if (kind == MMX_RK4) {  // src/integrator/rk.jl:23-52, as rk4_into
                    Epi<T> x = e.epi(EPI_STAGE);
                    x.y0 = yin;
                    x.kstore = e.k[0];
                    x.cself = 0.5 * dt;
                    x.yout = a;
                    d[0 * M + m] = {x, yin};
                    x.kstore = e.k[1];
                    x.yout = out;
                    d[1 * M + m] = {x, a};
                    x.kstore = e.k[2];
                    x.cself = dt;
                    x.yout = a;
                    d[2 * M + m] = {x, out};
                    Epi<T> f = e.epi(EPI_RK4_FINAL);
                    f.y0 = yin;
                    f.kp[0] = e.k[0];
                    f.kp[1] = e.k[1];
                    f.kp[2] = e.k[2];
                    f.np = 3;
                    f.cself = 1.0 / 6 * dt;
                    f.yout = out;
                    f.normalize = 1;
                    d[3 * M + m] = {f, a};
                } else {  // src/integrator/heun.jl:28-49, as heun_step
                    Epi<T> x = e.epi(EPI_STAGE);
                    x.y0 = yin;
                    x.kstore = e.k[0];
                    x.cself = dt;
                    x.yout = a;
                    d[0 * M + m] = {x, yin};
                    Epi<T> f = e.epi(EPI_STAGE);
                    f.y0 = yin;
                    f.kp[0] = e.k[0];
                    f.cp[0] = 0.5 * dt;
                    f.np = 1;
                    f.cself = 0.5 * dt;
                    f.yout = out;
                    f.normalize = 1;
                    d[1 * M + m] = {f, a};
                }
            }
        // group members' last passes also write the next stage's x-spectrum
        // (every stage's output is the next stage's input)
        for (auto &g : groups) {
            if (!g->fast) continue;
            for (size_t q = 0; q < g->mem.size(); q++)
                for (int par = 0; par < 2; par++)
                    for (int s = 0; s < S; s++) h[((size_t)par * S + s) * M + g->mem[q]].E.anext = g->A.p + q * g->astride;
        }
        hst = h;
        st.alloc(h.size());
        HIPCHECK(hipMemcpy(st.p, h.data(), h.size() * sizeof(h[0]), hipMemcpyHostToDevice));
        for (int s = 0; s < S; s++) {
            const Epi<T> &x = h[(size_t)s * M].E;  // every member's stage s has the same mode and np
            V[s] = batch_variant(x.mode, fast ? 1 : 0, x.np);
            for (auto &g : groups) g->vg[s] = batch_variant(x.mode, g->fast ? 1 : 0, x.np);
        }
        dyns.alloc(M);
        upload_dyn();
        tune();
    }

    void items_for(int b, int c, std::vector<int2> &it)
    {
        it.clear();
        for (int m : local)
            for (int64_t c0 = 0; c0 < E[m]->N; c0 += (int64_t)b * c) it.push_back(make_int2(m, (int)c0));
        nitems = (int)it.size();
        if (it.empty()) return;
        items.alloc(it.size());
        HIPCHECK(hipMemcpy(items.p, it.data(), it.size() * sizeof(int2), hipMemcpyHostToDevice));
    }

    // x-spectra of the groups' members at the stage-0 input of parity par
    // (the first pass; later stages take them from the previous last pass)
    void prime(int par)
    {
        for (auto &g : groups)
            for (size_t q = 0; q < g->mem.size(); q++) {
                Engine<T> &e = *E[g->mem[q]];
                const T *in = hst[(size_t)par * S * M + g->mem[q]].m;
                launch_p1<T>(e.G, in, e.N, e.mu0Ms.p, e.G.nx, e.G.ny, e.G.nz, 3, g->A.p + q * g->astride, e.tw.p,
                             B->stream);
            }
        HIPCHECK(hipGetLastError());
    }

    void stage_groups(int par, int s)
    {
        const BatchStage<T> *sts = st.p + ((size_t)par * S + s) * M;
        for (auto &g : groups) {
            const Engine<T> &e0 = *E[g->mem[0]];
            const DemagGeom &G = e0.G;
            const int n = (int)g->mem.size(), ny = G.ny, nz = G.nz;
            if (!g->fast)  // term-list last pass: no chained spectrum, the first pass runs every stage
                for (int q = 0; q < n; q++) {
                    const Engine<T> &e = *E[g->mem[q]];
                    launch_p1<T>(G, hst[((size_t)par * S + s) * M + g->mem[q]].m, e.N, e.mu0Ms.p, G.nx, ny, nz, 3,
                                 g->A.p + q * g->astride, e0.tw.p, B->stream);
                }
            launch_coly<T>(G, g->A.p, g->Bs.p, ny, ny, G.Py, G.Py, nz, 3 * n, 0, e0.tw.p, B->stream);
            launch_convz<T>(G, g->Bs.p, e0.Kt.p, e0.Lline.p, nz, nz, 3, 0, e0.tw.p, B->stream, n);
            launch_coly<T>(G, g->Bs.p, g->A.p, G.Py, G.Py, ny, ny, nz, 3 * n, 1, e0.tw.p, B->stream);
            launch_p5_batch<T>(G, g->A.p, g->astride, g->gidx.p, n, Ps.p, sts, dyns.p, g->vg[s], e0.tw.p, B->stream);
        }
    }

    void upload_dyn()
    {
        std::vector<Dyn> d(M);
        for (int m = 0; m < M; m++) d[m] = E[m]->dyn();
        HIPCHECK(hipMemcpyAsync(dyns.p, d.data(), M * sizeof(Dyn), hipMemcpyHostToDevice, B->stream));
    }

    void stage_local(int par, int s)
    {
        if (nitems)
            launch_batch_local<T>(items.p, nitems, Ps.p, st.p + ((size_t)par * S + s) * M, dyns.p, V[s], cpt,
                                  V[s] == 0 ? 256 : bs, grid, B->stream);
    }

    void stage(int par, int s)
    {
        stage_local(par, s);
        stage_groups(par, s);
    }

    void step(int par)
    {
        for (int s = 0; s < S; s++) stage(par, s);
        launch_batch_tick(dyns.p, M, B->stream);
    }

    // time the first stage of the local members (it writes only stage
    // buffers) for each launch shape: threads per block (the term-list
    // variant runs 256 only), cells per thread of a work item, grid
    void tune()
    {
        if (local.empty()) return;
        bool any_list = false;
        for (int s = 0; s < S; s++) any_list = any_list || V[s] == 0;
        const std::vector<int> cands_b = any_list ? std::vector<int>{256} : std::vector<int>{128, 256, 512};
        const int cands_c[3] = {1, 2, 4};
        const int cands_g[4] = {1024, 2048, 4096, 8192};
        hipEvent_t a, b;
        HIPCHECK(hipEventCreate(&a));
        HIPCHECK(hipEventCreate(&b));
        float best = 1e30f;
        int bb = 256, bc = 1, bg = 2048;
        std::vector<int2> it;
        tune_ms.clear();
        for (int t : cands_b)
            for (int c : cands_c) {
                items_for(t, c, it);
                for (int g : cands_g) {
                    bs = t;
                    cpt = c;
                    grid = g;
                    stage_local(0, 0);
                    HIPCHECK(hipEventRecord(a, B->stream));
                    for (int r = 0; r < 3; r++) stage_local(0, 0);
                    HIPCHECK(hipEventRecord(b, B->stream));
                    HIPCHECK(hipEventSynchronize(b));
                    float ms = 0;
                    HIPCHECK(hipEventElapsedTime(&ms, a, b));
                    tune_ms.push_back(ms / 3);
                    if (ms < best) {
                        best = ms;
                        bb = t;
                        bc = c;
                        bg = g;
                    }
                }
            }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        bs = bb;
        cpt = bc;
        grid = bg;
        items_for(bs, cpt, it);
    }

    void capture()
    {
        if (gexec) return;
        HIPCHECK(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
        step(0);
        step(1);
        hipError_t e = hipStreamEndCapture(B->stream, &graph);
        HIPCHECK(e);
        HIPCHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    }

    int advance(int64_t n)
    {
        for (auto *c : B->members) HIPCHECK(hipStreamSynchronize(c->stream));
        for (int m = 0; m < M; m++) {
            Engine<T> &e = *E[m];
            e.params_dev();
            if (e.y != bufA[m]) {  // the member stepped on its own since: its state is in B
                HIPCHECK(hipMemcpyAsync(bufA[m], e.y, 3 * e.N * sizeof(T), hipMemcpyDeviceToDevice, B->stream));
                e.y = bufA[m];
                e.stg[1] = bufB[m];
            }
        }
        upload_dyn();
        prime(0);
        capture();
        for (int64_t i = 0; i + 1 < n; i += 2) HIPCHECK(hipGraphLaunch(gexec, B->stream));
        if (n & 1) {
            step(0);
            for (int m = 0; m < M; m++)
                HIPCHECK(hipMemcpyAsync(E[m]->y, E[m]->stg[1], 3 * E[m]->N * sizeof(T), hipMemcpyDeviceToDevice,
                                        B->stream));
        }
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipStreamSynchronize(B->stream));
        for (auto *c : B->members) {
            c->nsteps += n;
            c->nfevals += (int64_t)S * n;
            c->t = c->nsteps * c->step;
        }
        return MMX_OK;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// dispatch helpers
// ---------------------------------------------------------------------------
template <class F>
static int with_engine(mmx_ctx *c, F &&f)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    try {
        if (c->mesh.precision == MMX_F64) {
            auto &e = *static_cast<Engine<double> *>(c->impl.get());
            e.nx_req = false;
            e.a_ready = nullptr;
            return f(e);
        }
        auto &e = *static_cast<Engine<float> *>(c->impl.get());
        e.nx_req = false;
        e.a_ready = nullptr;
        return f(e);
    } catch (const HipError &e) {
        return e.code;
    } catch (const std::bad_alloc &) {
        return fail(MMX_ENOMEM, "host allocation failed");
    }
}

template <class F>
static int with_batch(mmx_batch *b, F &&f)
{
    if (!b) return fail(MMX_EINVAL, "null batch");
    try {
        if (b->precision == MMX_F64) return f(*static_cast<BatchEngine<double> *>(b->impl.get()));
        return f(*static_cast<BatchEngine<float> *>(b->impl.get()));
    } catch (const HipError &e) {
        return e.code;
    } catch (const std::bad_alloc &) {
        return fail(MMX_ENOMEM, "host allocation failed");
    }
}


template <class F>
static int with_neb(mmx_neb *b, F &&f)
{
    if (!b || !b->sim) return fail(MMX_EINVAL, "null band");
    try {
        if (b->sim->mesh.precision == MMX_F64) return f(*static_cast<NebEngine<double> *>(b->impl.get()));
        return f(*static_cast<NebEngine<float> *>(b->impl.get()));
    } catch (const HipError &e) {
        return e.code;
    } catch (const std::bad_alloc &) {
        return fail(MMX_ENOMEM, "host allocation failed");
    }
}

static bool finite_all(const double *a, int64_t n)
{
    for (int64_t i = 0; i < n; i++)
        if (std::isnan(a[i])) return false;
    return true;
}

template <class T>
static int add_term(mmx_ctx *c, Engine<T> &eng, HostTerm h, const typename Engine<T>::value_type *a,
                    const typename Engine<T>::value_type *b, const typename Engine<T>::value_type *cc)
{
    if ((int)c->terms.size() >= MAX_TERMS) return fail(MMX_EINVAL, "at most %d interaction terms", MAX_TERMS);
    c->terms.push_back(h);
    eng.tA.push_back(a);
    eng.tB.push_back(b);
    eng.tC.push_back(cc);
    return (int)c->terms.size() - 1;
}

template <class T>
static void ensure_out(Engine<T> &e)
{
    e.ensure_k();
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *mmx_last_error(void) { return g_err.c_str(); }

int mmx_create(mmx_ctx **out, const mmx_mesh_desc *mesh)
{
    if (!out || !mesh) return fail(MMX_EINVAL, "null argument");
    *out = nullptr;
    if (mesh->nx < 1 || mesh->ny < 1 || mesh->nz < 1) return fail(MMX_EINVAL, "mesh sizes must be >= 1");
    if (!(mesh->dx > 0 && mesh->dy > 0 && mesh->dz > 0)) return fail(MMX_EINVAL, "cell sizes must be > 0");
    if (mesh->precision != MMX_F64 && mesh->precision != MMX_F32)
        return fail(MMX_EINVAL, "precision must be 64 or 32");
    int64_t N = (int64_t)mesh->nx * mesh->ny * mesh->nz;
    if (N > (int64_t)1 << 31) return fail(MMX_EINVAL, "mesh too large");
    auto c = std::make_unique<mmx_ctx>();
    c->mesh = *mesh;
    c->N = N;
    c->volume = mesh->dx * mesh->dy * mesh->dz;  // src/micro/mesh.jl:86
    c->mu0Ms.assign(N, 0.0);
    try {
        HIPCHECK(hipSetDevice(mesh->device));
        if (scratch_init(mesh->device) != MMX_OK) return MMX_EHIP;
        HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        if (mesh->precision == MMX_F64) c->impl.reset(new Engine<double>(c.get()));
        else c->impl.reset(new Engine<float>(c.get()));
        HIPCHECK(hipStreamSynchronize(c->stream));
    } catch (const HipError &e) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        return e.code;
    }
    *out = c.release();
    return MMX_OK;
}

int mmx_destroy(mmx_ctx *c)
{
    if (!c) return MMX_OK;
    (void)hipSetDevice(c->mesh.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->impl.reset();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MMX_OK;
}

int mmx_set_Ms(mmx_ctx *c, const double *Ms)
{
    if (!c || !Ms) return fail(MMX_EINVAL, "null argument");
    if (!finite_all(Ms, c->N)) return fail(MMX_EINVAL, "NaN is given by the input Ms!");
    return with_engine(c, [&](auto &e) {
        for (int64_t i = 0; i < c->N; i++) c->mu0Ms[i] = Ms[i] * MU0;
        e.upload_scalar(c->mu0Ms.data(), e.mu0Ms.p, c->N);
        e.update_mask();
        return MMX_OK;
    });
}

int mmx_set_pinning(mmx_ctx *c, const uint8_t *pins)
{
    if (!c || !pins) return fail(MMX_EINVAL, "null argument");
    return with_engine(c, [&](auto &e) {
        if (!e.pins.p) e.pins.alloc(c->N);
        HIPCHECK(hipMemcpyAsync(e.pins.p, pins, c->N, hipMemcpyHostToDevice, c->stream));
        HIPCHECK(hipStreamSynchronize(c->stream));
        e.have_pins = true;
        e.update_mask();
        return MMX_OK;
    });
}

int mmx_init_m0(mmx_ctx *c, const double *m, int norm)
{
    if (!c || !m) return fail(MMX_EINVAL, "null argument");
    const int64_t N = c->N;
    std::vector<double> s(m, m + 3 * N);
    if (norm) {  // normalise: src/helper.jl:255-271
        for (int64_t i = 0; i < N; i++) {
            double *v = &s[3 * i];
            double m2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            if (m2 > 0) {
                double len = 1 / std::sqrt(m2);
                v[0] *= len;
                v[1] *= len;
                v[2] *= len;
            }
        }
    }
    for (int64_t i = 0; i < N; i++)
        if (c->mu0Ms[i] == 0) s[3 * i] = s[3 * i + 1] = s[3 * i + 2] = 0;
    if (!finite_all(s.data(), 3 * N)) return fail(MMX_EINVAL, "NaN is given by the input m0!");
    return with_engine(c, [&](auto &e) {
        e.upload_soa(s.data(), e.y);
        HIPCHECK(hipStreamSynchronize(c->stream));
        c->succeed = 0;
        return MMX_OK;
    });
}

int mmx_get_m(mmx_ctx *c, double *m)
{
    if (!c || !m) return fail(MMX_EINVAL, "null argument");
    return with_engine(c, [&](auto &e) {
        e.download_aos(e.y, m);
        return MMX_OK;
    });
}

int mmx_add_exch(mmx_ctx *c, const double *A, double Ax, double Ay, double Az)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        HostTerm h;
        const double dx = c->mesh.dx, dy = c->mesh.dy, dz = c->mesh.dz;
        if (A) {  // SpatialExchange: nabla = 2/d^2 (kernels.jl:231-233)
            h.kind = K_EXCH_S;
            h.s[0] = (double)(TT(2) / ((TT)dx * (TT)dx));
            h.s[1] = (double)(TT(2) / ((TT)dy * (TT)dy));
            h.s[2] = (double)(TT(2) / ((TT)dz * (TT)dz));
            const TT *a = e.new_array(A, c->N, false);
            return add_term(c, e, h, a, nullptr, nullptr);
        }
        h.kind = K_EXCH;  // UniformExchange: 2A/d^2 (kernels.jl:277-279)
        h.s[0] = (double)(TT(2) * (TT)Ax / ((TT)dx * (TT)dx));
        h.s[1] = (double)(TT(2) * (TT)Ay / ((TT)dy * (TT)dy));
        h.s[2] = (double)(TT(2) * (TT)Az / ((TT)dz * (TT)dz));
        return add_term(c, e, h, nullptr, nullptr, nullptr);
    });
}

int mmx_add_dmi(mmx_ctx *c, int type, const double *Dx, const double *Dy, const double *Dz, double Dx0,
                double Dy0, double Dz0)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    if (type != MMX_DMI_BULK && type != MMX_DMI_INTERFACIAL)
        return fail(MMX_EINVAL, "Supported DMI type: interfacial bulk");
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        HostTerm h;
        const TT dx = (TT)c->mesh.dx, dy = (TT)c->mesh.dy, dz = (TT)c->mesh.dz;
        if (type == MMX_DMI_INTERFACIAL) {  // kernels.jl:426-466, Dd = 1/d
            h.kind = K_DMI_INT;
            h.s[0] = (double)(TT(1) / dx);
            h.s[1] = (double)(TT(1) / dy);
            if (!Dx) {
                h.uniform = 1;
                h.s[3] = (double)(TT)Dx0;
                // harmonic mean of two equal D, rounded as the kernel would (kernels.jl:452)
                volatile TT D = (TT)Dx0;
                TT num = TT(2) * D;
                num = num * D;
                TT den = D + D;
                h.s[4] = (double)(TT)(num / den);
                return add_term(c, e, h, nullptr, nullptr, nullptr);
            }
            const TT *a = e.new_array(Dx, c->N, false);
            return add_term(c, e, h, a, nullptr, nullptr);
        }
        if (!Dx && !Dy && !Dz) {  // BulkDMI: Ds = D/d (kernels.jl:313)
            h.kind = K_DMI_BULK;
            h.s[0] = (double)((TT)Dx0 / dx);
            h.s[1] = (double)((TT)Dy0 / dy);
            h.s[2] = (double)((TT)Dz0 / dz);
            // D and d as stored, for a time profile (TimeBulkDMI)
            h.s[6] = (double)(TT)Dx0;
            h.s[7] = (double)(TT)Dy0;
            h.s[8] = (double)(TT)Dz0;
            h.s[9] = (double)dx;
            h.s[10] = (double)dy;
            h.s[11] = (double)dz;
            return add_term(c, e, h, nullptr, nullptr, nullptr);
        }
        if (!Dx || !Dy || !Dz) return fail(MMX_EINVAL, "spatial bulk DMI needs all three arrays");
        h.kind = K_DMI_BULK_S;  // axes = 1/d (kernels.jl:356)
        h.s[0] = (double)(TT(1) / dx);
        h.s[1] = (double)(TT(1) / dy);
        h.s[2] = (double)(TT(1) / dz);
        const TT *a = e.new_array(Dx, c->N, false);
        const TT *b = e.new_array(Dy, c->N, false);
        const TT *cc = e.new_array(Dz, c->N, false);
        return add_term(c, e, h, a, b, cc);
    });
}

int mmx_add_anis(mmx_ctx *c, const double *Ku, double Ku0, const double *axes, double ax, double ay, double az)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        HostTerm h;
        const TT *ka = nullptr;
        if (Ku) ka = e.new_array(Ku, c->N, false);
        else {
            h.uniform = 1;
            h.s[3] = (double)(TT)Ku0;
        }
        if (axes) {  // spatial axis, normalised (add_field.jl:447-453)
            std::vector<double> ax3(axes, axes + 3 * c->N);
            for (int64_t i = 0; i < c->N; i++) {
                double *v = &ax3[3 * i];
                double m2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                if (m2 > 0) {
                    double len = 1 / std::sqrt(m2);
                    v[0] *= len;
                    v[1] *= len;
                    v[2] *= len;
                }
            }
            h.kind = K_ANIS_S;
            const TT *b = e.new_array(ax3.data(), 3 * c->N, true);
            return add_term(c, e, h, ka, b, nullptr);
        }
        double lt = std::sqrt(ax * ax + ay * ay + az * az);  // add_field.jl:441-443
        if (!(lt > 0)) return fail(MMX_EINVAL, "anisotropy axis must be nonzero");
        h.kind = K_ANIS;
        h.s[0] = (double)(TT)(ax / lt);
        h.s[1] = (double)(TT)(ay / lt);
        h.s[2] = (double)(TT)(az / lt);
        return add_term(c, e, h, ka, nullptr, nullptr);
    });
}

int mmx_add_cubic_anis(mmx_ctx *c, const double *Kc, double Kc0, const double a1[3], const double a2[3])
{
    if (!c || !a1 || !a2) return fail(MMX_EINVAL, "null argument");
    double n1 = std::sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
    double n2 = std::sqrt(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
    double na1[3] = {a1[0] / n1, a1[1] / n1, a1[2] / n1};
    double na2[3] = {a2[0] / n2, a2[1] / n2, a2[2] / n2};
    if (std::fabs(na1[0] * na2[0] + na1[1] * na2[1] + na1[2] * na2[2]) > 1e-10)
        return fail(MMX_EINVAL, "The axis1 and axis2 are not perpendicular to each other!");
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        HostTerm h;
        h.kind = K_CUBIC;
        // axis3 = axis1 x axis2 of the inputs as given (add_field.jl:555)
        double a3[3] = {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2], a1[0] * a2[1] - a1[1] * a2[0]};
        for (int q = 0; q < 3; q++) {
            h.s[q] = (double)(TT)na1[q];
            h.s[3 + q] = (double)(TT)na2[q];
            h.s[6 + q] = (double)(TT)a3[q];
        }
        const TT *ka = nullptr;
        if (Kc) ka = e.new_array(Kc, c->N, false);
        else {
            h.uniform = 1;
            h.s[9] = (double)(TT)Kc0;
        }
        return add_term(c, e, h, ka, nullptr, nullptr);
    });
}

int mmx_add_zeeman(mmx_ctx *c, const double *H, double Hx, double Hy, double Hz)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        HostTerm h;
        if (!H) {
            h.kind = K_ZEEMAN_U;
            h.uniform = 1;
            h.s[0] = (double)(TT)Hx;
            h.s[1] = (double)(TT)Hy;
            h.s[2] = (double)(TT)Hz;
            return add_term(c, e, h, nullptr, nullptr, nullptr);
        }
        h.kind = K_ZEEMAN;
        const TT *p = e.new_array(H, 3 * c->N, true);
        return add_term(c, e, h, p, p + c->N, p + 2 * c->N);
    });
}

int mmx_update_zeeman(mmx_ctx *c, int term, const double *H, double Hx, double Hy, double Hz)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    if (term < 0 || term >= (int)c->terms.size()) return fail(MMX_EINVAL, "no such term");
    HostTerm &h = c->terms[term];
    if (h.kind != K_ZEEMAN && h.kind != K_ZEEMAN_U) return fail(MMX_EINVAL, "term %d is not a Zeeman term", term);
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        if (!H) {
            h.kind = K_ZEEMAN_U;
            h.uniform = 1;
            h.s[0] = (double)(TT)Hx;
            h.s[1] = (double)(TT)Hy;
            h.s[2] = (double)(TT)Hz;
            return MMX_OK;
        }
        const TT *p = e.new_array(H, 3 * c->N, true);
        h.kind = K_ZEEMAN;
        h.uniform = 0;
        e.tA[term] = p;
        e.tB[term] = p + c->N;
        e.tC[term] = p + 2 * c->N;
        return MMX_OK;
    });
}

int mmx_add_demag(mmx_ctx *c, int Nx, int Ny, int Nz)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    if (c->demag_term >= 0) return fail(MMX_EINVAL, "only one demag term per simulation");
    if (Nx < 0 || Ny < 0 || Nz < 0) return fail(MMX_EINVAL, "image counts must be >= 0");
    return with_engine(c, [&](auto &e) {
        HostTerm h;
        h.kind = K_DEMAG;
        h.rtensor.resize(6 * c->N);
        newell_tensor(c->mesh.nx, c->mesh.ny, c->mesh.nz, c->mesh.dx, c->mesh.dy, c->mesh.dz, Nx, Ny, Nz,
                      h.rtensor.data(), 0);
        e.setup_demag(h.rtensor);
        int r = add_term(c, e, h, nullptr, nullptr, nullptr);
        if (r >= 0) {
            c->demag_term = r;
            c->demag_img[0] = Nx;
            c->demag_img[1] = Ny;
            c->demag_img[2] = Nz;
        }
        return r;
    });
}

int mmx_add_thermal_noise(mmx_ctx *c, const double *T, double T_uniform, double T0, uint64_t seed)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    for (const auto &h : c->terms)
        if (h.kind == K_THERMAL) return fail(MMX_EINVAL, "only one thermal term per simulation");
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        if (e.alphas.p) return fail(MMX_EINVAL, "the stochastic field needs a uniform alpha (field.jl:257-261)");
        HostTerm h;
        h.kind = K_THERMAL;
        h.s[0] = (double)(TT)T0;
        h.seed = seed;
        const TT *a = nullptr;
        if (T) a = e.new_array(T, c->N, false);
        else {
            h.uniform = 1;
            h.s[1] = (double)(TT)T_uniform;
        }
        return add_term(c, e, h, a, nullptr, nullptr);
    });
}

int mmx_add_sot(mmx_ctx *c, const double *aj, double aj0, double bj, double px, double py, double pz)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    if (aj && !finite_all(aj, c->N)) return fail(MMX_EINVAL, "NaN in aj");
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        HostTerm h;  // DFTorqueField(T(p), aj array, T(bj)): add_field.jl:951-966
        h.kind = K_SOT;
        h.s[0] = (double)(TT)px;
        h.s[1] = (double)(TT)py;
        h.s[2] = (double)(TT)pz;
        h.s[4] = (double)(TT)bj;
        const TT *a = nullptr;
        if (aj) a = e.new_array(aj, c->N, false);
        else {
            h.uniform = 1;
            h.s[3] = (double)(TT)aj0;
        }
        return add_term(c, e, h, a, nullptr, nullptr);
    });
}

int mmx_add_stt_slonczewski(mmx_ctx *c, const double *J, double J0, double beta, double Lambda, double xi,
                            double P, double px, double py, double pz)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    if (J && !finite_all(J, c->N)) return fail(MMX_EINVAL, "NaN in J");
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        HostTerm h;  // SlonczewskiTorque: add_field.jl:1195-1227; lambda_sq = T(Lambda^2), field.jl:349
        h.kind = K_STT_SL;
        h.s[0] = (double)(TT)px;
        h.s[1] = (double)(TT)py;
        h.s[2] = (double)(TT)pz;
        h.s[4] = (double)(TT)((double)(TT)Lambda * (double)(TT)Lambda);
        h.s[5] = (double)(TT)P;
        h.s[6] = (double)(TT)xi;
        h.s[7] = (double)(TT)beta;
        const TT *a = nullptr;
        if (J) a = e.new_array(J, c->N, false);
        else {
            h.uniform = 1;
            h.s[3] = (double)(TT)J0;
        }
        return add_term(c, e, h, a, nullptr, nullptr);
    });
}

int mmx_add_stt_zhangli(mmx_ctx *c, const double *J, double Jx, double Jy, double Jz, double b, const double *xi,
                        double xi0, int xi_factor)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    const int64_t N = c->N;
    if (J && !finite_all(J, 3 * N)) return fail(MMX_EINVAL, "NaN in J");
    if (xi && !finite_all(xi, N)) return fail(MMX_EINVAL, "NaN in xi");
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        // ZhangLiTorque: add_field.jl:1150-1192.  bJ = T(J) .* b in T, then, for
        // (P, Ms) input, times 1/(1 + xi^2) evaluated in T
        std::vector<double> bJ(3 * N);
        for (int64_t i = 0; i < N; i++) {
            const TT x = (TT)(xi ? xi[i] : xi0);
            const TT f = TT(1) / (TT(1) + x * x);
            for (int q = 0; q < 3; q++) {
                const TT Jv = (TT)(J ? J[3 * i + q] : (q == 0 ? Jx : (q == 1 ? Jy : Jz)));
                TT v = (TT)((double)Jv * b);
                if (xi_factor) v = v * f;
                bJ[3 * i + q] = (double)v;
            }
        }
        HostTerm h;
        h.kind = K_STT_ZL;
        h.s[0] = (double)(TT)c->mesh.dx;
        h.s[1] = (double)(TT)c->mesh.dy;
        h.s[2] = (double)(TT)c->mesh.dz;
        const TT *xa = nullptr;
        if (xi) xa = e.new_array(xi, N, false);
        else {
            h.uniform = 1;
            h.s[3] = (double)(TT)xi0;
        }
        const TT *bb = e.new_array(bJ.data(), 3 * N, true);
        return add_term(c, e, h, xa, bb, nullptr);
    });
}

int mmx_add_hex_anis(mmx_ctx *c, double K1, double K2, double K3)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        HostTerm h;  // HexagonalAnisotropy(T(K1), T(K2), T(K3)): add_field.jl:590-606
        h.kind = K_HEX;
        h.uniform = 1;
        h.s[0] = (double)(TT)K1;
        h.s[1] = (double)(TT)K2;
        h.s[2] = (double)(TT)K3;
        return add_term(c, e, h, nullptr, nullptr, nullptr);
    });
}

int mmx_add_torque(mmx_ctx *c, void (*fn)(double t, const double *m, double *y, void *user), void *user)
{
    if (!c || !fn) return fail(MMX_EINVAL, "null argument");
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        HostTerm h;  // TorqueField(fun): add_field.jl:998-1006
        h.kind = K_TORQUE_FN;
        h.tfn = fn;
        h.uuser = user;
        h.tbuf.assign(3 * c->N, 0.0);
        std::vector<double> zero(3 * c->N, 0.0);
        const TT *a = e.new_array(zero.data(), 3 * c->N, true);
        e.any_tfn = true;
        return add_term(c, e, h, a, nullptr, nullptr);
    });
}

int mmx_add_magnetoelastic(mmx_ctx *c, int model, const double *data, double lambda_s, double B1, double B2)
{
    if (!c || !data) return fail(MMX_EINVAL, "null argument");
    if (model != MMX_MEL_TENSOR && model != MMX_MEL_CUBIC)
        return fail(MMX_EINVAL, "Unknown model: use :tensor or :cubic");
    const int64_t N = c->N;
    if (!finite_all(data, 6 * N)) return fail(MMX_EINVAL, "NaN in the stress / strain field");
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        // Magnetoelastic(model, ...): add_field.jl:1367-1405; the (6, N) Voigt field as 6 planes
        std::vector<double> q((size_t)6 * N);
        for (int64_t i = 0; i < N; i++)
            for (int comp = 0; comp < 6; comp++) q[(size_t)comp * N + i] = data[6 * i + comp];
        HostTerm h;
        h.kind = model == MMX_MEL_TENSOR ? K_MEL_T : K_MEL_C;
        h.s[0] = (double)(TT)(model == MMX_MEL_TENSOR ? lambda_s : B1);
        h.s[1] = (double)(TT)(model == MMX_MEL_TENSOR ? 0.0 : B2);
        const TT *a = e.new_array(q.data(), 6 * N, false);
        return add_term(c, e, h, a, nullptr, nullptr);
    });
}

int mmx_add_sahe_torque(mmx_ctx *c, const double *sigma_s, const double *sigma_sa1, const double *sigma_sa2,
                        const double *a1, const double *a2, const double *a3, double beta)
{
    if (!c || !sigma_s || !sigma_sa1 || !sigma_sa2 || !a1 || !a2 || !a3) return fail(MMX_EINVAL, "null argument");
    const int64_t N = c->N;
    for (const double *v : {sigma_s, sigma_sa1, sigma_sa2, a1, a2, a3})
        if (!finite_all(v, 3 * N)) return fail(MMX_EINVAL, "NaN in a vector field");
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        // SAHETorqueField(sigma_s x a1, sigma_sa1 x a1, sigma_sa2, a2, a3, beta): add_field.jl:892-936,
        // the cross products in T (cross_product, src/util.jl:35-46), as 15 planes of N
        std::vector<double> q((size_t)15 * N);
        auto P = [&](int f, int comp, int64_t i) -> double & { return q[(size_t)(3 * f + comp) * N + i]; };
        for (int64_t i = 0; i < N; i++) {
            const int64_t k = 3 * i;
            const TT x1 = (TT)a1[k], x2 = (TT)a1[k + 1], x3 = (TT)a1[k + 2];
            for (int f = 0; f < 2; f++) {
                const double *sv = f == 0 ? sigma_s : sigma_sa1;
                const TT s1 = (TT)sv[k], s2 = (TT)sv[k + 1], s3 = (TT)sv[k + 2];
                P(f, 0, i) = (double)(TT)(-s3 * x2 + s2 * x3);
                P(f, 1, i) = (double)(TT)(s3 * x1 - s1 * x3);
                P(f, 2, i) = (double)(TT)(-s2 * x1 + s1 * x2);
            }
            for (int comp = 0; comp < 3; comp++) {
                P(2, comp, i) = sigma_sa2[k + comp];
                P(3, comp, i) = a2[k + comp];
                P(4, comp, i) = a3[k + comp];
            }
        }
        HostTerm h;
        h.kind = K_SAHE;
        h.s[0] = (double)(TT)beta;
        const TT *a = e.new_array(q.data(), 15 * N, false);
        return add_term(c, e, h, a, nullptr, nullptr);
    });
}

int mmx_add_twin_mono_anis(mmx_ctx *c, const double K[6], const double *axis_a, const double *axis_b,
                           const double *axis_u111)
{
    if (!c || !K || !axis_a || !axis_b || !axis_u111) return fail(MMX_EINVAL, "null argument");
    const int64_t N = c->N;
    // normalise(axes, n_total): src/helper.jl:255-271 (zero vectors stay zero)
    std::vector<double> ax[3];
    const double *in[3] = {axis_a, axis_b, axis_u111};
    for (int q = 0; q < 3; q++) {
        if (!finite_all(in[q], 3 * N)) return fail(MMX_EINVAL, "NaN in an axis");
        ax[q].assign(in[q], in[q] + 3 * N);
        for (int64_t i = 0; i < N; i++) {
            double *v = ax[q].data() + 3 * i;
            const double m2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            if (m2 > 0) {
                const double f = 1.0 / std::sqrt(m2);
                v[0] *= f;
                v[1] *= f;
                v[2] *= f;
            }
        }
    }
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        HostTerm h;  // TwinMonoclinicAnisotropy: add_field.jl:629-660
        h.kind = K_TWIN;
        for (int q = 0; q < 6; q++) h.s[q] = (double)(TT)K[q];
        const TT *a = e.new_array(ax[0].data(), 3 * N, true);
        const TT *b = e.new_array(ax[1].data(), 3 * N, true);
        const TT *u = e.new_array(ax[2].data(), 3 * N, true);
        return add_term(c, e, h, a, b, u);
    });
}

int mmx_add_exch_int(mmx_ctx *c, const double *J, int k1, int k2)
{
    if (!c || !J) return fail(MMX_EINVAL, "null argument");
    const int nz = c->mesh.nz;
    if (k2 == -1) k2 = nz;  // add_field.jl:802-804
    if (k1 < 1 || k1 > nz || k2 < 1 || k2 > nz) return fail(MMX_EINVAL, "layers k1=%d, k2=%d outside 1..%d", k1, k2, nz);
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        HostTerm h;  // InterlayerExchange(Js, k1, k2): add_field.jl:789-816
        h.kind = K_EXCH_IL;
        h.s[3] = (double)(TT)c->mesh.dz;
        h.s[4] = k1 - 1;
        h.s[5] = k2 - 1;
        const TT *a = e.new_array(J, (int64_t)c->mesh.nx * c->mesh.ny, false);
        return add_term(c, e, h, a, nullptr, nullptr);
    });
}

int mmx_add_dmi_int(mmx_ctx *c, double Dx, double Dy, double Dz, int k1, int k2)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    const int nz = c->mesh.nz;
    if (k2 == -1) k2 = nz;  // add_field.jl:840-842
    if (k1 < 1 || k1 > nz || k2 < 1 || k2 > nz) return fail(MMX_EINVAL, "layers k1=%d, k2=%d outside 1..%d", k1, k2, nz);
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        HostTerm h;  // InterlayerDMI(T(Dx), T(Dy), T(Dz), k1, k2): add_field.jl:835-860
        h.kind = K_DMI_IL;
        h.uniform = 1;
        h.s[0] = (double)(TT)Dx;
        h.s[1] = (double)(TT)Dy;
        h.s[2] = (double)(TT)Dz;
        h.s[3] = (double)(TT)c->mesh.dz;
        h.s[4] = k1 - 1;
        h.s[5] = k2 - 1;
        return add_term(c, e, h, nullptr, nullptr, nullptr);
    });
}

int mmx_add_zeeman_time(mmx_ctx *c, const double *H, double Hx, double Hy, double Hz,
                        void (*fn)(double t, double *f, void *user), void *user)
{
    if (!c || !fn) return fail(MMX_EINVAL, "null argument");
    if (H && !finite_all(H, 3 * c->N)) return fail(MMX_EINVAL, "NaN in H");
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        HostTerm h;  // TimeZeeman: add_field.jl:161-200 (init_field in T)
        h.kind = K_ZEEMAN_T;
        h.ft3 = fn;
        h.uuser = user;
        if (!H) {
            h.uniform = 1;
            h.s[0] = (double)(TT)Hx;
            h.s[1] = (double)(TT)Hy;
            h.s[2] = (double)(TT)Hz;
            return add_term(c, e, h, nullptr, nullptr, nullptr);
        }
        const TT *p = e.new_array(H, 3 * c->N, true);
        return add_term(c, e, h, p, p + c->N, p + 2 * c->N);
    });
}

int mmx_set_term_time_fn(mmx_ctx *c, int term, double (*fn)(double t, void *user), void *user)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    if (term < 0 || term >= (int)c->terms.size()) return fail(MMX_EINVAL, "no term %d", term);
    HostTerm &h = c->terms[term];
    if (h.kind != K_STT_SL && h.kind != K_STT_ZL && h.kind != K_DMI_BULK && h.kind != K_DMI_BULK_S)
        return fail(MMX_EINVAL, "term %d has no time-dependent amplitude", term);
    h.ufun = fn;
    h.uuser = user;
    return MMX_OK;
}

int mmx_num_terms(mmx_ctx *c) { return c ? (int)c->terms.size() : fail(MMX_EINVAL, "null context"); }

int mmx_effective_field(mmx_ctx *c, double t, double *H)
{
    return with_engine(c, [&](auto &e) {
        ensure_out(e);
        auto E = e.epi(EPI_FIELD);
        E.out = e.k[6];
        e.t_eval = t;
        e.eval(e.y, E);
        if (H) e.download_aos(e.k[6], H);
        else HIPCHECK(hipStreamSynchronize(c->stream));
        return MMX_OK;
    });
}

int mmx_term_field(mmx_ctx *c, int term, double t, double *H, double *Eo)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    if (term < 0 || term >= (int)c->terms.size()) return fail(MMX_EINVAL, "no such term");
    return with_engine(c, [&](auto &e) {
        ensure_out(e);
        auto E = e.epi(EPI_FIELD);
        E.term_mask = 1u << term;
        E.out = e.k[6];
        E.e_out = e.k[5];
        e.t_eval = t;
        e.eval(e.y, E);
        if (H) e.download_aos(e.k[6], H);
        if (Eo) e.download_scalar(e.k[5], Eo, c->N);
        HIPCHECK(hipStreamSynchronize(c->stream));
        return MMX_OK;
    });
}

int mmx_compute_energies(mmx_ctx *c, double t, double *Eo)
{
    if (!c || !Eo) return fail(MMX_EINVAL, "null argument");
    return with_engine(c, [&](auto &e) {
        // one masked evaluation per term; each block writes its FP64 partial sum
        e.t_eval = t;
        for (int q = 0; q < (int)c->terms.size(); q++) {
            auto E = e.epi(EPI_ENERGY);
            E.term_mask = 1u << q;
            Eo[q] = e.reduce_scalar(e.eval(e.y, E), false);
        }
        return MMX_OK;
    });
}

int mmx_compute_energies_noalloc(mmx_ctx *c, double t, double *E_dev)
{
    if (!c || !E_dev) return fail(MMX_EINVAL, "null argument");
    double *part;
    {
        std::lock_guard<std::mutex> l(scratch_mu);
        part = scratch_ptr[c->mesh.device];
    }
    if (!part) return fail(MMX_ESTATE, "device %d has no scratch address (context not created by mmx_create)",
                           c->mesh.device);
    return with_engine(c, [&](auto &e) { return e.energies_noalloc(t, E_dev, part, NOALLOC_SLOTS); });
}

int mmx_set_alpha(mmx_ctx *c, const double *alpha)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    if (alpha && !finite_all(alpha, c->N)) return fail(MMX_EINVAL, "NaN in alpha");
    if (alpha)
        for (const auto &h : c->terms)
            if (h.kind == K_THERMAL)
                return fail(MMX_EINVAL, "the stochastic field needs a uniform alpha (field.jl:257-261)");
    return with_engine(c, [&](auto &e) {
        if (!alpha) {
            HIPCHECK(hipStreamSynchronize(c->stream));
            e.alphas.release();
            return MMX_OK;
        }
        if (!e.alphas.p) e.alphas.alloc(c->N);
        e.upload_scalar(alpha, e.alphas.p, c->N);
        HIPCHECK(hipStreamSynchronize(c->stream));
        return MMX_OK;
    });
}

int mmx_set_llg(mmx_ctx *c, double alpha, double gamma, int precession)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    c->alpha = alpha;
    c->gamma = gamma;
    c->precession = precession ? 1 : 0;
    return MMX_OK;
}

int mmx_llg_rhs(mmx_ctx *c, double t, const double *m, double *dmdt)
{
    if (!c || !dmdt) return fail(MMX_EINVAL, "null argument");
    return with_engine(c, [&](auto &e) {
        ensure_out(e);
        const auto *src = e.y;
        if (m) {
            e.upload_soa(m, e.stg[2]);
            src = e.stg[2];
        }
        auto E = e.epi(EPI_RHS);
        E.out = e.k[6];
        e.t_eval = t;
        e.eval(src, E);
        e.download_aos(e.k[6], dmdt);
        return MMX_OK;
    });
}

int mmx_set_integrator(mmx_ctx *c, int kind, double step, double tol)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    if (kind < MMX_RK4 || kind > MMX_SD) return fail(MMX_EINVAL, "unknown integrator %d", kind);
    if ((kind == MMX_RK4 || kind == MMX_HEUN) && !(step > 0)) return fail(MMX_EINVAL, "fixed-step integrators need step > 0");
    if ((kind == MMX_DOPRI54 || kind == MMX_BS23) && !(tol > 0)) return fail(MMX_EINVAL, "tol must be > 0");
    c->ikind = kind;
    c->t = 0;
    c->nsteps = 0;
    c->nfevals = 0;
    c->succeed = 0;
    c->step_next = 0;
    c->tol = tol;
    c->step = (kind == MMX_RK4 || kind == MMX_HEUN) ? step : 0.0;
    if (kind == MMX_SD) {  // create_driver("SD"): src/driver.jl:93-100
        c->sd_tau = 0;
        c->sd_steps = 0;
    }
    return MMX_OK;
}

int mmx_advance_step(mmx_ctx *c, int64_t n)
{
    return with_engine(c, [&](auto &e) {
        int rc = e.advance(n);
        HIPCHECK(hipStreamSynchronize(c->stream));
        return rc;
    });
}

int mmx_run_until(mmx_ctx *c, double t_end)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    if (c->ikind == MMX_RK4 || c->ikind == MMX_HEUN)
        return fail(MMX_ESTATE, "run_until needs an adaptive integrator (integrate_to_time is defined for AdaptiveRK only)");
    return with_engine(c, [&](auto &e) {
        if (t_end < c->t - c->step) return MMX_OK;  // sim.jl:549-551
        if (t_end == c->t) return MMX_OK;
        while (c->t < t_end) {  // integrate_to_time: adaptive_rk.jl:168-186
            if (c->step_next + c->t > t_end) c->step_next = t_end - c->t;
            int rc = e.adaptive_step();
            if (rc != MMX_OK) return rc;
        }
        HIPCHECK(hipStreamSynchronize(c->stream));
        return MMX_OK;
    });
}

int mmx_relax(mmx_ctx *c, int64_t max_steps, double stopping_dmdt, int64_t *steps, double *max_dmdt_out)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    return with_engine(c, [&](auto &e) {
        const double dmdt_factor = (2 * M_PI / 360) * 1e9;  // sim.jl:406
        double max_dmdt = 0;
        int64_t i;
        int rc = MMX_OK;
        for (i = 1; i <= max_steps; i++) {
            rc = e.advance(1);
            if (rc != MMX_OK) break;
            // step_size: the integrator's step, or tau / (gamma/2) for SD (sim.jl:404-405, :430)
            const double step_size = c->ikind == MMX_SD ? c->sd_tau / (2.21e5 / 2) : c->step;
            max_dmdt = e.max_dm() / step_size;  // compute_dm!: helper.jl:273-286
            if (max_dmdt < stopping_dmdt * dmdt_factor) break;
        }
        if (steps) *steps = i > max_steps ? max_steps : i;
        if (max_dmdt_out) *max_dmdt_out = max_dmdt;
        HIPCHECK(hipStreamSynchronize(c->stream));
        return rc;
    });
}

int mmx_get_state(mmx_ctx *c, double *t, double *step, double *step_next, int64_t *nsteps, int64_t *nfevals)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    if (t) *t = c->t;
    if (step) *step = c->ikind == MMX_SD ? c->sd_tau : c->step;
    if (step_next) *step_next = c->step_next;
    if (nsteps) *nsteps = c->nsteps;
    if (nfevals) *nfevals = c->nfevals;
    return MMX_OK;
}

int mmx_demag_tensor(mmx_ctx *c, int term, double *out)
{
    if (!c || !out) return fail(MMX_EINVAL, "null argument");
    if (term < 0 || term >= (int)c->terms.size() || c->terms[term].kind != K_DEMAG)
        return fail(MMX_EINVAL, "term %d is not a demag term", term);
    std::memcpy(out, c->terms[term].rtensor.data(), 6 * c->N * sizeof(double));
    return MMX_OK;
}

int mmx_demag_geometry(mmx_ctx *c, int32_t *Px, int32_t *Py, int32_t *Pz, int32_t *KS)
{
    return with_engine(c, [&](auto &e) {
        if (!e.demag) return fail(MMX_ESTATE, "no demag term");
        if (Px) *Px = e.G.Px;
        if (Py) *Py = e.G.Py;
        if (Pz) *Pz = e.G.Pz;
        if (KS) *KS = e.G.KS;
        return MMX_OK;
    });
}

int mmx_synchronize(mmx_ctx *c)
{
    if (!c) return fail(MMX_EINVAL, "null context");
    try {
        HIPCHECK(hipStreamSynchronize(c->stream));
    } catch (const HipError &e) {
        return e.code;
    }
    return MMX_OK;
}

int mmx_diag_copy(int device, int64_t bytes, int reps, int grid, double *ms_per_copy)
{
    if (bytes < 16 || reps < 1 || grid < 1 || !ms_per_copy) return fail(MMX_EINVAL, "invalid copy probe arguments");
    const int64_t n16 = bytes / 16;
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MMX_OK;
    try {
        HIPCHECK(hipSetDevice(device));
        HIPCHECK(hipStreamCreate(&s));
        HIPCHECK(hipMalloc(&a, n16 * 16));
        HIPCHECK(hipMalloc(&b, n16 * 16));
        HIPCHECK(hipMemsetAsync(a, 0, n16 * 16, s));
        HIPCHECK(hipEventCreate(&e0));
        HIPCHECK(hipEventCreate(&e1));
        launch_stream_copy(a, b, n16, grid, s);  // warm-up
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventRecord(e0, s));
        for (int r = 0; r < reps; r++) launch_stream_copy(a, b, n16, grid, s);
        HIPCHECK(hipEventRecord(e1, s));
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
        *ms_per_copy = ms / reps;
    } catch (const HipError &e) {
        rc = e.code;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (s) (void)hipStreamDestroy(s);
    return rc;
}

int mmx_profile(mmx_ctx *c, int enable)
{
    return with_engine(c, [&](auto &e) {
        e.collect_profile();
        for (int i = 0; i < MMX_NPASS; i++) {
            e.prof_ms[i] = 0;
            e.prof_n[i] = 0;
        }
        c->profile = enable != 0;
        return MMX_OK;
    });
}

int mmx_profile_read(mmx_ctx *c, double *ms, int64_t *launches)
{
    return with_engine(c, [&](auto &e) {
        e.collect_profile();
        for (int i = 0; i < MMX_NPASS; i++) {
            if (ms) ms[i] = e.prof_ms[i];
            if (launches) launches[i] = e.prof_n[i];
        }
        return MMX_OK;
    });
}

int mmx_set_term_param(mmx_ctx *c, int term, const double *values)
{
    if (!c || !values) return fail(MMX_EINVAL, "null argument");
    if (term < 0 || term >= (int)c->terms.size()) return fail(MMX_EINVAL, "no term %d", term);
    const int kind = c->terms[term].kind;
    if (kind != K_EXCH_S && kind != K_DMI_INT && kind != K_ANIS && kind != K_ANIS_S && kind != K_CUBIC &&
        kind != K_THERMAL)
        return fail(MMX_EINVAL, "term %d has no scalar parameter field", term);
    if (!finite_all(values, c->N)) return fail(MMX_EINVAL, "NaN in the parameter values");
    return with_engine(c, [&](auto &e) {
        using TT = std::remove_pointer_t<decltype(e.y)>;
        HostTerm &h = c->terms[term];
        if (h.uniform || !e.tA[term]) {
            e.tA[term] = e.new_array(values, c->N, false);
            h.uniform = 0;
        } else {
            e.upload_scalar(values, const_cast<TT *>(e.tA[term]), c->N);
        }
        return MMX_OK;
    });
}

int mmx_adjoint_record(mmx_ctx *c, int64_t nsteps, int64_t every)
{
    return with_engine(c, [&](auto &e) { return e.adj_record(nsteps, every); });
}

int mmx_adjoint_backward(mmx_ctx *c, const double *lambda_T, double *grad_m0)
{
    if (!lambda_T) return fail(MMX_EINVAL, "null cotangent");
    if (c && !finite_all(lambda_T, 3 * c->N)) return fail(MMX_EINVAL, "NaN in the cotangent");
    return with_engine(c, [&](auto &e) { return e.adj_backward(lambda_T, grad_m0); });
}

int mmx_adjoint_param_grad(mmx_ctx *c, int term, double *grad)
{
    if (!grad) return fail(MMX_EINVAL, "null argument");
    return with_engine(c, [&](auto &e) { return e.adj_param_grad(term, grad); });
}

int mmx_adjoint_ms_grad(mmx_ctx *c, double *grad)
{
    if (!grad) return fail(MMX_EINVAL, "null argument");
    return with_engine(c, [&](auto &e) { return e.adj_ms_grad(grad); });
}

int mmx_llg_rhs_jvp(mmx_ctx *c, double t, const double *m, const double *v, double *out)
{
    if (!c || !v || !out) return fail(MMX_EINVAL, "null argument");
    if (!finite_all(v, 3 * c->N) || (m && !finite_all(m, 3 * c->N))) return fail(MMX_EINVAL, "NaN in the input");
    return with_engine(c, [&](auto &e) { return e.rhs_jvp(t, m, v, out); });
}

int mmx_llg_rhs_vjp(mmx_ctx *c, double t, const double *m, const double *a, double *out)
{
    if (!c || !a || !out) return fail(MMX_EINVAL, "null argument");
    if (!finite_all(a, 3 * c->N) || (m && !finite_all(m, 3 * c->N))) return fail(MMX_EINVAL, "NaN in the input");
    return with_engine(c, [&](auto &e) { return e.rhs_vjp(t, m, a, out); });
}

int mmx_batch_create(mmx_batch **out, mmx_ctx *const *members, int n)
{
    if (!out || !members || n < 1) return fail(MMX_EINVAL, "need at least one member");
    *out = nullptr;
    const int prec = members[0] ? members[0]->mesh.precision : 0;
    for (int i = 0; i < n; i++) {
        mmx_ctx *c = members[i];
        if (!c) return fail(MMX_EINVAL, "member %d is null", i);
        if (c->mesh.precision != prec || c->mesh.device != members[0]->mesh.device)
            return fail(MMX_EINVAL, "members must share precision and device");
        if (c->ikind != MMX_RK4 && c->ikind != MMX_HEUN)
            return fail(MMX_EINVAL, "member %d: batches run the fixed-step RungeKutta or Heun integrators", i);
        if (c->ikind != members[0]->ikind) return fail(MMX_EINVAL, "members must use the same integrator");
        if (!(c->step > 0)) return fail(MMX_EINVAL, "member %d has no step size", i);
        for (const auto &h : c->terms)
            if (h.ufun || h.ft3 || h.tfn)
                return fail(MMX_EINVAL, "member %d has a time-dependent or host-function term; batches take "
                                        "constant ones", i);
        for (int j = 0; j < i; j++)
            if (members[j] == c) return fail(MMX_EINVAL, "member %d appears twice", i);
    }
    auto b = std::make_unique<mmx_batch>();
    b->members.assign(members, members + n);
    b->precision = prec;
    b->device = members[0]->mesh.device;
    try {
        HIPCHECK(hipSetDevice(b->device));
        HIPCHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        for (auto *c : b->members) HIPCHECK(hipStreamSynchronize(c->stream));
        if (prec == MMX_F64) {
            auto *e = new BatchEngine<double>(b.get());
            b->impl.reset(e);
            e->build();
        } else {
            auto *e = new BatchEngine<float>(b.get());
            b->impl.reset(e);
            e->build();
        }
        HIPCHECK(hipStreamSynchronize(b->stream));
    } catch (const HipError &e) {
        b->impl.reset();
        if (b->stream) (void)hipStreamDestroy(b->stream);
        return e.code;
    } catch (const std::bad_alloc &) {
        b->impl.reset();
        if (b->stream) (void)hipStreamDestroy(b->stream);
        return fail(MMX_ENOMEM, "host allocation failed");
    }
    *out = b.release();
    return MMX_OK;
}

int mmx_batch_destroy(mmx_batch *b)
{
    if (!b) return MMX_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->impl.reset();
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return MMX_OK;
}

int mmx_batch_advance(mmx_batch *b, int64_t nsteps)
{
    if (nsteps < 0) return fail(MMX_EINVAL, "nsteps must be >= 0");
    return with_batch(b, [&](auto &e) { return e.advance(nsteps); });
}

int mmx_batch_shape(mmx_batch *b, int32_t *block_threads, int32_t *cells_per_thread, int32_t *grid,
                    int32_t *n_groups)
{
    return with_batch(b, [&](auto &e) {
        if (block_threads) *block_threads = e.bs;
        if (cells_per_thread) *cells_per_thread = e.cpt;
        if (grid) *grid = e.grid;
        if (n_groups) *n_groups = (int32_t)e.groups.size();
        return MMX_OK;
    });
}

int mmx_batch_info(mmx_batch *b, int32_t *cells_per_thread, int32_t *grid, int32_t *work_items, float *tune_ms,
                   int32_t n_tune)
{
    return with_batch(b, [&](auto &e) {
        if (cells_per_thread) *cells_per_thread = e.cpt;
        if (grid) *grid = e.grid;
        if (work_items) *work_items = e.nitems;
        if (tune_ms)
            for (int i = 0; i < n_tune && i < (int)e.tune_ms.size(); i++) tune_ms[i] = e.tune_ms[i];
        return MMX_OK;
    });
}

int mmx_neb_create(mmx_neb **out, mmx_ctx *sim, const double *images, int n_free, double spring_constant,
                   int climbing_image)
{
    if (!out || !sim || !images) return fail(MMX_EINVAL, "null argument");
    *out = nullptr;
    if (n_free < 1) return fail(MMX_EINVAL, "The number of free images should larger than 1 (typically 10~20).");
    for (const auto &h : sim->terms)
        if (h.tfn) return fail(MMX_EINVAL, "the NEB band takes no add_torque term");
    if (!finite_all(images, (int64_t)(n_free + 2) * 3 * sim->N)) return fail(MMX_EINVAL, "NaN in the images");
    auto b = std::make_unique<mmx_neb>();
    b->sim = sim;
    try {
        HIPCHECK(hipSetDevice(sim->mesh.device));
        if (sim->mesh.precision == MMX_F64) {
            auto *x = new NebEngine<double>(b.get(), sim, n_free);
            b->impl.reset(x);
            x->spring = spring_constant;
            x->climb = climbing_image;
            x->set_images(images);
        } else {
            auto *x = new NebEngine<float>(b.get(), sim, n_free);
            b->impl.reset(x);
            x->spring = spring_constant;
            x->climb = climbing_image;
            x->set_images(images);
        }
    } catch (const HipError &e) {
        return e.code;
    } catch (const std::bad_alloc &) {
        return fail(MMX_ENOMEM, "host allocation failed");
    }
    *out = b.release();
    return MMX_OK;
}

int mmx_neb_destroy(mmx_neb *b)
{
    if (!b) return MMX_OK;
    if (b->sim && b->sim->stream) (void)hipStreamSynchronize(b->sim->stream);
    b->impl.reset();
    delete b;
    return MMX_OK;
}

int mmx_neb_set_driver(mmx_neb *b, double alpha, double gamma, int precession, double tol)
{
    if (!(tol > 0)) return fail(MMX_EINVAL, "tol must be > 0");
    return with_neb(b, [&](auto &x) {
        x.alpha = alpha;
        x.gamma = gamma;
        x.precession = precession;
        x.tol = tol;
        x.step_next = 0;
        x.succeed = false;
        x.have_k4 = false;
        return MMX_OK;
    });
}

int mmx_neb_advance(mmx_neb *b, int64_t nsteps)
{
    return with_neb(b, [&](auto &x) {
        x.upload_params();
        for (int64_t i = 0; i < nsteps; i++) {
            int rc = x.advance_step();
            if (rc != MMX_OK) return rc;
        }
        HIPCHECK(hipStreamSynchronize(x.st()));
        return MMX_OK;
    });
}

int mmx_neb_relax(mmx_neb *b, int64_t max_steps, double stopping_dmdt, int64_t *steps, double *max_dmdt)
{
    return with_neb(b, [&](auto &x) {
        x.upload_params();
        return x.relax(max_steps, stopping_dmdt, steps, max_dmdt);
    });
}

int mmx_neb_get(mmx_neb *b, double *images, double *energies, double *distances, double *state)
{
    return with_neb(b, [&](auto &x) {
        x.refresh();
        HIPCHECK(hipStreamSynchronize(x.st()));
        if (images)
            for (int q = 0; q < x.n; q++) x.e.download_aos(x.Y.p + (size_t)q * x.V, images + (size_t)q * x.V);
        if (energies)
            HIPCHECK(hipMemcpy(energies, x.energy.p, (x.n + 2) * sizeof(double), hipMemcpyDeviceToHost));
        if (distances)
            HIPCHECK(hipMemcpy(distances, x.dist.p, (x.n + 1) * sizeof(double), hipMemcpyDeviceToHost));
        if (state) {
            state[0] = x.t;
            state[1] = x.step;
            state[2] = x.step_next;
            state[3] = (double)x.nsteps;
            state[4] = (double)x.nfevals;
        }
        return MMX_OK;
    });
}

}  // extern "C"
