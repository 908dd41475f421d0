// Auxiliary device kernels of the Born-sampling path (declarations).
#pragma once
#include <hip/hip_runtime.h>
#include "z2gemm.h"

namespace spp {

// A dense batched tensor seen through one parity-enumerated group covering all
// of its legs (parity 0 tuples = the neutral entries that may be nonzero).
struct FullView {
  Group g;
  long long stride[kMaxLegs];
  void* base;
  long long sstride;
  void* const* ptrs;
};

// out[s] = sum over neutral entries of x*y (x == y gives the squared norm).
// With partials (device scratch of nsamples * kDotChunks doubles) each sample
// is split over kDotChunks workgroups and the partial sums are added in a
// fixed order.
constexpr int kDotChunks = 32;
template <typename T>
hipError_t launch_dot(const FullView& x, const FullView& y, double* out, int nsamples, hipStream_t st,
                      double* partials = nullptr);

// x[entry] *= f[s] when f[s] != 0 (f given as 1/norm or 0 to leave untouched).
template <typename T>
hipError_t launch_scale(const FullView& x, const double* f, int nsamples, hipStream_t st);

// out = (choice[s] == 0 ? a : b) * f[s], neutral entries only (f may be null: 1).
template <typename T>
hipError_t launch_select_copy(const FullView& out, const FullView& a, const FullView& b, const int* choice,
                              const double* f, int nsamples, hipStream_t st);

// Fill a dense buffer with a value.
template <typename T>
hipError_t launch_fill(T* p, long long n, T v, hipStream_t st);

// factor[s] = 1/sqrt(nrm2[s]) (0 when the norm is 0, which leaves the tensor
// untouched); a nonfinite norm sets status[s] = 3 and fail_site[s] = site.
hipError_t launch_safe_factor(const double* nrm2, double* factor, int* status, int* fail_site, int site, int nb,
                              hipStream_t st);

// out[s] = sum_i cand_lognorm[2 i + labels[s][i]].
hipError_t launch_gather_lognorm(const double* cand_lognorm, const int* labels, int nsite, double* out, int nb,
                                 hipStream_t st);

// out[s] = prod_x tensors[x][s * sstrides[x]]: a closed boundary with unit bonds.
template <typename T>
hipError_t launch_close_unit(T* const* tensors, const long long* sstrides, int L, double* out, int nb,
                             hipStream_t st);

// Convert between precisions.
template <typename TO, typename TI>
hipError_t launch_convert(TO* out, const TI* in, long long n, hipStream_t st);

// Double layer of site tensors (W,N,P,A,E,S) -> E[W,N,E,S] with Delta legs in
// the parity-sorted pair order; optional physical operator O[p',p] (2x2) and
// auxiliary projector u[a] (projects the aux leg onto u, the same vector on
// ket and bra). A ket leg with flipk set is the fusion of a virtual leg with a
// unit leg of odd parity: its parity-sorted order lists the original odd
// indices first. ``ket``/``bra`` point to FP64 dense sites with leg dims and
// even counts given per site. One launch builds ``nsites`` layers.
struct LayerJob {
  const double* ket;
  const double* bra;
  void* out;
  int dk[6], ek[6];   // ket leg dims / even counts (W,N,P,A,E,S)
  int db[6], eb[6];   // bra leg dims / even counts
  double op[4];       // physical operator (row-major O[p_out][p_in]); identity if use_op == 0
  int use_op;
  double aux[2];      // aux projector weights; if use_aux == 0: trace over aux
  int use_aux;
  int flipk[6];       // ket virtual leg carries an extra odd unit charge: its parity order is reversed
};
template <typename T>
hipError_t launch_double_layers(const LayerJob* jobs_dev, int njobs, long long maxout, hipStream_t st);

// Batched symmetric eigensolver (cyclic Jacobi in LDS) for problems of size
// n <= 64 stored compactly: A_p at a + p*astride (row-major, leading dim lda).
// Eigenvalues are written in descending order to w + p*wstride and the
// corresponding eigenvectors as columns of V_p (v + p*vstride, leading dim ldv).
// n_of[p] gives each problem's size (problems with n == 0 or pmask[p] == 0 are skipped).
hipError_t launch_jacobi(const double* a, long long astride, int lda, double* w, long long wstride, double* v,
                         long long vstride, int ldv, const int* n_of, const int* pmask, int nprob,
                         hipStream_t st);

}  // namespace spp
