// mmx_kernels.h -- host-side launchers of the HIP kernels (mmx_kernels.hip).
#pragma once
#include "mmx_common.h"

namespace mmx {

// conversions between the caller's AoS layout and the device SoA planes
template <class T>
void launch_aos_to_soa(const double *aos_dev, T *soa, int64_t N, hipStream_t s);
template <class T>
void launch_soa_to_aos(const T *soa, double *aos_dev, int64_t N, hipStream_t s);
template <class T>
void launch_cast(const double *src, T *dst, int64_t n, hipStream_t s);
template <class T>
void launch_to_double(const T *src, double *dst, int64_t n, hipStream_t s);

// local field (no demag) + epilogue over all cells; returns number of blocks
// used (the number of partial-reduction slots written).
template <class T>
int launch_local(int64_t N, const FieldParams<T> *P, const Dyn &dyn, const Epi<T> &E, const T *m, int fast,
                 uint32_t sig, hipStream_t s);

// demag passes
template <class T>
void launch_p1(const DemagGeom &G, const T *in, int64_t cstride, const T *scale, int n_in, int Ry,
               int Rz, int ncomp, cplx<T> *A, const cplx<T> *tw, hipStream_t s);
template <class T>
void launch_coly(const DemagGeom &G, const cplx<T> *in, cplx<T> *out, int n_in, int Rin, int n_out,
                 int Rout, int nplanes, int ncomp, int inverse, const cplx<T> *tw, hipStream_t s);
// 1 when the z pass takes the wave form, whose folded tensor layout groups 4 kx
int convz_grouped_tensor(const DemagGeom &G, int elsize);
template <class T>
void launch_convz(const DemagGeom &G, cplx<T> *B, const T *Kt, const T *Lline, int nzb, int n_in,
                  int ncomp, int fwd_only, const cplx<T> *tw, hipStream_t s, int nimg = 1);
template <class T>
int launch_p5(const DemagGeom &G, const cplx<T> *A, const FieldParams<T> *P, const Dyn &dyn, const Epi<T> &E,
              const T *m, int fast, uint32_t sig, const cplx<T> *tw, hipStream_t s);
// fields and per-cell energies (EPI_FIELD) of nimg stacked images whose spectra
// lie side by side in A (grid y = image)
template <class T>
void launch_p5_images(const DemagGeom &G, const cplx<T> *A, const FieldParams<T> *P, const Dyn &dyn,
                      const Epi<T> &E, const T *m, int nimg, const cplx<T> *tw, hipStream_t s);
// last pass of a group of batch members sharing one geometry: member q's
// spectra at A + q*astride, its tables entries gidx[q]; v is the variant of
// the group's stage mode (4/1/2 slot form, 0 term list)
template <class T>
void launch_p5_batch(const DemagGeom &G, const cplx<T> *A, int64_t astride, const int *gidx, int nmem,
                     const FieldParams<T> *const *Ps, const BatchStage<T> *st, const Dyn *dyns, int v,
                     const cplx<T> *tw, hipStream_t s);
template <class T>
void launch_sum_images(const T *v, int64_t N, int n, int nb, double *partial, double *out, hipStream_t s);
// kernels compiled for a term signature (mmx_kernels_sig.hip); -1 when none matches
template <class T>
int launch_p5_sig(const DemagGeom &G, const cplx<T> *A, const FieldParams<T> *P, const Dyn &dyn, const Epi<T> &E,
                  const T *m, uint32_t sig, const cplx<T> *tw, hipStream_t s);
template <class T>
int launch_local_sig(int64_t N, const FieldParams<T> *P, const Dyn &dyn, const Epi<T> &E, const T *m, uint32_t sig,
                     int grid, hipStream_t s);

// tensor set-up (FP64)
void launch_tensor_pad(const DemagGeom &G, const double *rt, int comp, double *pad, hipStream_t s);
template <class T>
void launch_tensor_extract(const DemagGeom &G, const cplx<double> *Bt, int comp, double scale, T *Kt,
                           hipStream_t s);

// reductions of per-block partials: out[v] = reduce over blocks of partial[b*stride+v]
void launch_reduce(const double *partial, int nblocks, int stride, int nval, int is_max, double *out,
                   hipStream_t s);

// max_i |a_i - b_i| over cells (norm of the 3-vector difference), for relax
template <class T>
int launch_maxdiff(const T *a, const T *b, int64_t N, double *partial, hipStream_t s);

// out = y + c k over n values (BS23 first-same-as-last stage)
template <class T>
void launch_axpy(const T *y, double c, const T *k, T *out, int64_t n, hipStream_t s);
// per-block max |v_i| over n values; returns the number of blocks
template <class T>
int launch_maxabs(const T *v, int64_t n, double *partial, hipStream_t s);

// adjoint of the RK4 step: the EPI_ADJ input vector, and the pull-back of
// the final combination + normalisation
template <class T>
void launch_llg_jvp(int64_t N, const FieldParams<T> *P, const T *Y, const T *hY, const T *v, const T *hv, T *out,
                    hipStream_t s);
template <class T>
void launch_adj_pre(int64_t N, const FieldParams<T> *P, const T *Y, const T *a, T *ut, hipStream_t s);
template <class T>
void launch_adj_final(int64_t N, const T *y0, const T *const k[4], double c, const T *lam, T *lam0, T *const a[4],
                      hipStream_t s);

// batch members without demag: one launch per stage over all of them; V is
// the epilogue variant (batch_variant), cpt the cells per thread of a work
// item, bs the threads per block (128, 256 or 512; 256 for the term list)
template <class T>
void launch_batch_local(const int2 *items, int nitems, const FieldParams<T> *const *Ps, const BatchStage<T> *st,
                        const Dyn *dyns, int V, int cpt, int bs, int grid, hipStream_t s);
void launch_batch_tick(Dyn *dyns, int M, hipStream_t s);
int batch_variant(int mode, int fast, int np);

// nudged elastic band (per-image reductions over grids of nb x images)
template <class T>
int launch_sum(const T *v, int64_t n, double *partial, hipStream_t s);
template <class T>
void launch_neb_dist(int64_t N, int n, const T *Y, const T *L, const T *R, int nb, double *partial, double *dist,
                     hipStream_t s);
template <class T>
void launch_neb_field(const FieldParams<T> *P, const Dyn &dyn, uint32_t mask, int64_t N, int n, int nb, const T *Y,
                      T *F, double *partial, double *energy, hipStream_t s);
template <class T>
void launch_neb_maxdiff(const T *a, const T *b, int64_t N, int n, int nb, double *partial, double *out, hipStream_t s);
template <class T>
void launch_neb_tangent(int64_t N, int n, const T *Ys, const T *L, const T *R, const double *energy, const T *F,
                        T *tan, int nb, double *partial, const double *dist, double k, int climb, double *coef,
                        hipStream_t s);
template <class T>
void launch_neb_stage(const FieldParams<T> *P, int64_t N, int n, int nb, const T *Ys, const T *F, const T *tan,
                      const double *coef, int mode, const double c[4], double outer, T *kstore, const T *y0,
                      const T *k1, const T *k2, const T *k3, T *yout, double *partial, hipStream_t s);

// per-cell neighbour / pin mask (FieldParams::nbm)
template <class T>
void launch_nbmask(int64_t N, const FieldParams<T> *P, uint8_t *out, hipStream_t s);

// steepest descent (src/sd.jl)
template <class T>
int launch_sd_g(int64_t N, const T *m, const T *mpre, const T *h, T *g, double *partial, hipStream_t s);
void launch_sd_tau(const double *sums, int odd, double min_tau, double max_tau, double *tau, hipStream_t s);
template <class T>
void launch_sd_update(int64_t N, const FieldParams<T> *P, const T *g, T *m, T *mpre, const T *h, const double *tau,
                      hipStream_t s);
template <class T>
void launch_normalise(int64_t N, T *m, hipStream_t s);

int max_local_blocks();
// blocks of one launch of the last demag pass for this geometry (G.p5rx rows each)
int p5_grid(const DemagGeom &G);
// rows per block of the last pass for a launch over `copies` systems of G
int p5_rows(const DemagGeom &G, int copies);
// streaming copy of n16 16-byte words (HBM ceiling probe)
void launch_stream_copy(const void *src, void *dst, int64_t n16, int grid, hipStream_t s);

}  // namespace mmx
