// Host-side engine of the MI355X Born-sampling path.
//
// Device data model: every tensor is dense per sample with parity-sorted legs
// (even indices first). Double-layer legs pair (ket, bra) in the parity-sorted
// pair order of split_pair(). Boundary bonds carry a per-sample even count.
// South boundaries are stored rotated by 180 degrees so that one forward
// compression routine serves both directions.
#pragma once
#include <hip/hip_runtime.h>
#include <map>
#include <string>
#include <vector>
#include <memory>
#include <stdexcept>
#include "z2gemm.h"
#include "kernels.h"
#include "kron.h"

namespace spp {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define SPP_CHECK(expr)                                                                       \
  do {                                                                                        \
    hipError_t e__ = (expr);                                                                  \
    if (e__ != hipSuccess)                                                                    \
      throw ::spp::HipError(std::string(#expr) + ": " + hipGetErrorString(e__) + " at " + \
                            __FILE__ + ":" + std::to_string(__LINE__));                       \
  } while (0)

// Host-to-device copy on the stream that returns once the copy is complete:
// the host source is often a temporary (a job table, a pointer list) released
// right after the call, and a pageable source must outlive its transfer.
inline void upload(void* dst, const void* src, size_t bytes, hipStream_t st) {
  SPP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
  SPP_CHECK(hipStreamSynchronize(st));
}

// Device arena with stack-like release.
class Arena {
 public:
  ~Arena();
  void reserve(size_t bytes);
  void* take(size_t bytes);
  size_t mark() const { return top_; }
  void release(size_t m) { top_ = m; }
  size_t capacity() const { return cap_; }
  size_t peak() const { return peak_; }
  const char* base() const { return base_; }
  // Serve take() from caller-provided device memory [base, base + bytes)
  // until restore(); reserve() then only checks the size (no hipMalloc or
  // hipFree while borrowed). The owned block is kept and comes back on restore().
  void borrow(char* base, size_t bytes);
  void restore();

 private:
  char* base_ = nullptr;
  size_t cap_ = 0, top_ = 0, peak_ = 0;
  bool borrowed_ = false;
  char* own_base_ = nullptr;
  size_t own_cap_ = 0, own_top_ = 0;
};

struct Leg {
  int dim = 1;
  int even = 1;              // fixed even count (when split == nullptr)
  const int* split = nullptr;  // per-sample even counts (device)
  bool flat = false;         // ungraded leg (subspace columns)
};

inline Leg fixed_leg(int dim, int even) {
  Leg l;
  l.dim = dim;
  l.even = even;
  return l;
}
inline Leg flat_leg(int dim) {
  Leg l;
  l.dim = dim;
  l.even = dim;
  l.flat = true;
  return l;
}

// A batched operand: dense (strides from dims) or compact per-parity block matrix.
struct TRef {
  void* base = nullptr;
  long long sstride = 0;
  void* const* ptrs = nullptr;
  long long pstride = 0;
  std::string letters;
  std::vector<long long> stride;
  std::vector<Leg> leg;
  bool compact = false;
  int nrow = 0;       // compact: number of row letters
  long long ld = 0;   // compact: leading dimension
  // Folded dense storage: the last leg (fixed even count) keeps one parity
  // half, max(even, odd) slots; an entry's position in that half follows from
  // its index, and its parity from the neutrality of the stored entries.
  bool fold = false;
  long long offset = 0;  // elements from each sample's base to this view (sub_leg views)
};

TRef dense_ref(void* base, long long sstride, const std::string& letters, const std::vector<Leg>& legs);
// Dense tensor with its last leg folded (neutral entries only, about half the
// storage of dense_ref). Contractions only: no FullView kernels on it.
TRef folded_ref(void* base, long long sstride, const std::string& letters, const std::vector<Leg>& legs);
// Elements per sample of folded_ref(..., legs).
long long folded_count(const std::vector<Leg>& legs);
TRef compact_ref(void* base, long long sstride, long long pstride, long long ld, const std::string& letters,
                 int nrow, const std::vector<Leg>& legs);
TRef with_letters(const TRef& t, const std::string& letters);
TRef with_ptrs(const TRef& t, void* const* ptrs);
// Permute a dense TRef's letters without moving data (e.g. read a south tensor in reversed order).
TRef relabel(const TRef& t, const std::string& new_letters_in_storage_order);
// View of the index range [r0, r0 + chunk.dim) of one leg of a dense operand:
// same strides, the leg replaced by `chunk` (whose parity split must be that of
// the range; see chunk_leg).
TRef sub_leg(const TRef& t, char letter, int r0, const Leg& chunk);
// Default of Context::gram_chunk_bytes.
constexpr double kGramChunkBytes = 1 << 30;

struct Context;
// The leg of the index range [r0, r0 + c) of a parity-sorted leg: its even
// count is clamp(even - r0, 0, c), per sample (written into `split_out`, nb
// ints of device memory, on ctx.stream) when the leg's split varies per sample.
Leg chunk_leg(Context& ctx, const Leg& l, int r0, int c, int* split_out, int nb);

struct Counters {
  double gemm_flops = 0;  // 2*M*N*K summed over launched (sample, parity) problems
  double gemm_launches = 0;
  double gemm_ms = 0;     // summed contraction kernel time (profiling mode)
  double gemm_flops_profiled = 0;
  // profiled launches split by element type: [0] FP64 kernel, [1] FP32 kernel
  double prec_launches[2] = {0, 0};
  double prec_ms[2] = {0, 0};
  double prec_flops[2] = {0, 0};
  // truncation eigensolver: iterative calls, subspace iterations, filter products
  double eig_calls = 0, eig_iterations = 0, eig_filter_gemms = 0;
  // truncation eigenproblems (sample, parity block) whose Ritz residuals were
  // still above subspace_tol when max_subspace_iter ended the iteration
  double eig_unconverged = 0;
  double fu_repairs = 0;  // bonds repaired by the PSD projection of the hole
  // contraction tile autotuner: launch classes tuned, classes whose fastest
  // shape is not the heuristic's, and the milliseconds the tuning launches took
  double tuned_classes = 0, tuned_changed = 0, tune_ms = 0;
  // draw row sweeps captured as HIP graphs, and graph replays
  double graph_captures = 0, graph_replays = 0;
  double kron_launches = 0;  // factorised Gram transfer launches (their multiply-adds are in gemm_flops)
};

// Event pairs recorded around contraction launches when profiling is on.
struct LaunchTimer {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  std::vector<double> pending_flops;
  std::vector<int> pending_prec;  // 0: FP64 kernel, 1: FP32 kernel
};

// Cut spectra waiting for the batched eigenvalue solve of the boundary
// entropy (compact rho blocks copied aside during absorbs, solved together).
constexpr int kEntropyCuts = 32;
struct EntropyTable {
  double* mat[kEntropyCuts];      // [2 nb][ld][ld] blocks (overwritten)
  double* w[kEntropyCuts];        // [2 nb][wstride] eigenvalues
  const int* cnt[kEntropyCuts];   // [2 nb] block dimensions
  double* out[kEntropyCuts];      // [cuts][nb] entropies of the absorb, or nullptr
  double* run_max[kEntropyCuts];  // [nb] running maximum, or nullptr
  long long wstride[kEntropyCuts];
  int ld[kEntropyCuts];
  int cut[kEntropyCuts];          // cut index (after site cut) in the absorb frame
  int nb[kEntropyCuts];           // samples
  int count;
};
struct EntropyWork {
  char* buf = nullptr;
  size_t cap = 0, used = 0;
  struct Cut {
    size_t mat, w, cnt;
    int ld, cut, nb;
    double* out;
    double* run_max;
  };
  std::vector<Cut> pending;
  EntropyWork() = default;
  EntropyWork(const EntropyWork&) = delete;
  EntropyWork& operator=(const EntropyWork&) = delete;
  ~EntropyWork();
};

// Device buffer owned by a context, grown on demand (stream-ordered reuse).
struct DeviceScratch {
  double* p = nullptr;
  size_t n = 0;
  DeviceScratch() = default;
  DeviceScratch(const DeviceScratch&) = delete;
  DeviceScratch& operator=(const DeviceScratch&) = delete;
  ~DeviceScratch();
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  int precision = 0;        // 0: FP64, 1: FP32 contraction GEMMs (FP64 factorizations)
  // FP32 contraction products: 0 on the FP32 matrix cores, 2 or 3 from that
  // many BF16 pieces per operand on the BF16 matrix cores (3: six piece
  // products, FP32-equivalent; 2: three products, about 2^-16 relative)
  int split = 0;
  // Per-sample byte limit of one Gram transfer intermediate of the boundary
  // absorb (the bra bond is taken in index ranges that keep it below this).
  double gram_chunk_bytes = kGramChunkBytes;
  int ksub = 64;            // subspace size for truncation eigenproblems
  int max_subspace_iter = 40;
  double subspace_tol = 1e-12;
  int cheb_degree = 8;      // largest Chebyshev filter degree per subspace step (1: plain power step)
  // Boundary sector budget (BoundaryMPS(chi, sector_space)): at most bcap[q]
  // kept values of bond parity q per cut, intersected with truncrank(chi);
  // negative: no budget.
  int bcap[2] = {-1, -1};
  Arena arena;
  Counters counters;
  std::string last_error;
  long long chunk = 0;      // samples per internal chunk (0: automatic)
  int last_iterations = 0;
  int profile = 0;          // 1: time every contraction launch with HIP events
  LaunchTimer timer;
  EntropyWork entropy;      // boundary-entropy spectra of the current absorb
  DeviceScratch dot_part;   // partial sums of the chunked dot products
  // No-allocation mode (set by the entry points that must not allocate):
  // the chunked dot products take their partial sums from this window of the
  // library's static scratch array instead of growing dot_part.
  double* noalloc_dot = nullptr;
  size_t noalloc_dot_n = 0;
  // Device address of the library's static scratch array on this context's
  // device (resolved once, when the context is created).
  char* static_scratch = nullptr;
  // Test hook of the full update's failure handling: treat bond update
  // fu_fail_bond (sweep order) of sample fu_fail_sample as failed with
  // fu_fail_code (fu_fail_sample < 0: off).
  long long fu_fail_sample = -1;
  int fu_fail_bond = -1, fu_fail_code = 0;
  // Host callback of the full update's PSD repair (spp_fu_repair_fn of the C
  // ABI) and its user pointer; null: a bond that needs the repair fails its sample.
  void* fu_repair_fn = nullptr;
  void* fu_repair_user = nullptr;
  // Tile shapes of the contraction kernel. autotune = 1: the first launch of
  // each launch class (kernel family, operand load modes, sample count, and
  // the dimension and fixed parity split of every M, N and K leg, i.e. the
  // block-size histogram of the launch) times every tile shape of its family
  // on the real operands and the class keeps the fastest; launches that add
  // into their output are never repeated, so a class first met there takes
  // the heuristic shape until a plain launch tunes it. All shapes give
  // bit-identical results. force_tile >= 0 (self-tests): every launch uses
  // that shape index.
  int autotune = 1;
  int force_tile = -1;
  std::map<std::vector<long long>, int> tile_table;
  // HIP graphs of the draw's row sweeps (1, the default): the second time a
  // row sweep of one shape runs on a sampling cache it is captured, and later
  // runs replay the graph; 0: every sweep is launched kernel by kernel.
  int graphs = 1;
  bool capturing = false;  // a row sweep is being captured: no tuning launches
  // Gram transfers of FP32 absorbs whose rows carry single-layer tables
  // (kron.h): 1 (the default) applies each double layer through its ket and
  // bra factors on the vector ALUs, 2 the same on the matrix cores (the vector
  // kernel where no matrix-core instantiation covers the layer), 0 contracts
  // the dense double layer; kron_rows: rows per block of that kernel (0: its
  // default).
  int kron = 1;
  int kron_rows = 0;
  int kron_groups = 4;      // row groups per block of the matrix-core kernel
  void flush_timer();       // accumulate finished event pairs into counters
};

// Boundary entropy of one cut from eigenvalues already at hand (w [2 nb][wstride],
// cnt [2 nb]); out[cut * nb + s] (if out) and the NaN-propagating running
// maximum run_max [nb] (if run_max).
void entropy_cut(Context& ctx, const double* w, long long wstride, const int* cnt, int nb, int cut, double* out,
                 double* run_max);
// Copy the compact blocks rho [2 nb][ld][ld] of one cut aside for the batched
// solve (solving what is pending first when the workspace is full). The
// outputs are written by entropy_flush, which the caller runs before reading
// them.
void entropy_defer(Context& ctx, const double* rho, long long rps, int ld, const int* cnt, int nb, int cut,
                   double* out, double* run_max);
// Solve every deferred cut and write its entropies.
void entropy_flush(Context& ctx);
// Drop pending cuts (their outputs are being freed) and the workspace.
void entropy_reset(Context& ctx);
// All eigenvalues (ascending) of nprob symmetric blocks [nprob][ld][ld]
// (host), of dimensions n_h[i] <= ld, into eig_h [nprob][ld] (host).
void spectrum_selftest(Context& ctx, int nprob, int ld, const int* n_h, const double* mats_h, double* eig_h);

// Scratch for launch_dot's partial sums of nsamples samples on ctx.stream.
double* dot_partials(Context& ctx, int nsamples);

// Contraction C = alpha * A.B (+ C): letters shared by A and B are summed.
template <typename T>
void contract(Context& ctx, const TRef& C, const TRef& A, const TRef& B, int nsamples, double alpha = 1.0,
              bool accumulate = false, const int* pmask = nullptr);

FullView full_view(const TRef& t);

// A factorised double-layer launch (FP32, kron.h), timed like the contraction
// launches when profiling is on.
template <typename F>
void timed_kron(Context& ctx, double flops, F&& launch) {
  if (!ctx.profile) {
    SPP_CHECK(launch());
    return;
  }
  hipEvent_t a, b;
  SPP_CHECK(hipEventCreate(&a));
  SPP_CHECK(hipEventCreate(&b));
  SPP_CHECK(hipEventRecord(a, ctx.stream));
  SPP_CHECK(launch());
  SPP_CHECK(hipEventRecord(b, ctx.stream));
  ctx.timer.pending.emplace_back(a, b);
  ctx.timer.pending_flops.push_back(flops);
  ctx.timer.pending_prec.push_back(1);
  if (ctx.timer.pending.size() >= 4096) ctx.flush_timer();
}

// ------------------------------ boundary MPS -------------------------------
template <typename T>
struct Boundary {
  int L = 0;
  int nb = 0;                     // samples
  std::vector<int> bdim;          // size L+1
  std::vector<int> beven;         // fixed even counts (for shared boundaries)
  std::vector<int*> bsplit;       // per-sample even counts (device), or nullptr
  std::vector<int> pdim, peven;   // physical leg per site
  std::vector<T*> t;              // device tensors [nb][bdim[x]][pdim[x]][bdim[x+1]]
  double* log_scale = nullptr;    // device [nb]
  double* max_trunc = nullptr;    // device [nb]
  // Boundary-entropy recording of absorbs into this boundary (not owned):
  // entropy [L-1][nb] per cut of the last absorb, entropy_max [nb] running
  // maximum; nullptr: not recorded.
  double* entropy = nullptr;
  double* entropy_max = nullptr;
  bool owns = true;
  Leg bond(int b) const {
    Leg l;
    l.dim = bdim[b];
    l.even = beven[b];
    l.split = bsplit[b];
    return l;
  }
  Leg phys(int x) const { return fixed_leg(pdim[x], peven[x]); }
  long long tsize(int x) const { return (long long)bdim[x] * pdim[x] * bdim[x + 1]; }
  TRef ref(int x, const std::string& letters) const {
    return dense_ref(t[x], tsize(x), letters, {bond(x), phys(x), bond(x + 1)});
  }
  void release();
};

// Double layer row of a lattice (shared, batch 1 unless per-sample pointer tables are given).
struct LayerSpec {
  Leg w, n, e, s;
};

template <typename T>
struct RowLayers {
  int L = 0;
  std::vector<LayerSpec> spec;            // per site
  std::vector<void* const*> ptrs;         // per site: per-sample pointer tables (device) or nullptr
  std::vector<T*> shared;                 // per site: shared layer when ptrs == nullptr
  bool rotated = false;                   // read as E_rot[w,n,e,s] = E[e,s,w,n]
  // Single-layer factors of the layers (FP32 rows, unrotated; optional): per
  // site the leg layout and the per-sample table pointers of kron.h. Empty:
  // the Gram transfers contract the dense double layers.
  std::vector<const KronSite*> kron_site;
  std::vector<void* const*> kron_tabs;
  // Reference to site x with the given letters for (W,N,E,S) in the frame of this row.
  TRef ref(int x, const std::string& letters) const;
};

template <typename T>
void trivial_boundary(Context& ctx, Boundary<T>& b, int L, int nb, const std::vector<Leg>& phys);

// Forward (north-type) absorb of a row of layers into env (in place result in out).
template <typename T>
void absorb_forward(Context& ctx, const Boundary<T>& env, const RowLayers<T>& row, int chi, Boundary<T>& out);

}  // namespace spp
