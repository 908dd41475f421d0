// Engine: contraction builder and the batched row absorb with canonical
// compression.
//
// Compression (restates compress_boundary!, src/contraction/boundary_mps.jl:94-134,
// result-equivalently): the reference right-orthogonalises the absorbed MPS by
// LQ and then truncates left to right by SVD. The singular values and left
// singular vectors at cut x are those of rho_x = Y_x G_{x+1} Y_x^T, where Y_x
// is the left-truncated tensor at site x and G_{x+1} the exact Gram matrix of
// the absorbed MPS to the right of the cut (L_{x+1} L_{x+1}^T). So the kernel
// sequence is: right Grams by four contractions per site, then per cut rho_x,
// its top-chi eigenvectors over both parity blocks, U_x and R_x = U_x^T Y_x.
#include "engine.h"
#include <chrono>
#include "boundary_ops.h"
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>

namespace spp {

// ------------------------------- arena --------------------------------------
Arena::~Arena() {
  if (borrowed_) restore();
  if (base_) (void)hipFree(base_);
}

void Arena::borrow(char* base, size_t bytes) {
  if (borrowed_) throw HipError("arena already borrows external memory");
  own_base_ = base_;
  own_cap_ = cap_;
  own_top_ = top_;
  base_ = base;
  cap_ = bytes;
  top_ = 0;
  borrowed_ = true;
}

void Arena::restore() {
  if (!borrowed_) return;
  base_ = own_base_;
  cap_ = own_cap_;
  top_ = own_top_;
  borrowed_ = false;
}

void Arena::reserve(size_t bytes) {
  if (bytes <= cap_) return;
  if (borrowed_)
    throw HipError("device workspace exhausted: need " + std::to_string(bytes) + " bytes, the scratch window holds " +
                   std::to_string(cap_));
  if (top_ != 0) throw HipError("arena resize while in use");
  if (base_) SPP_CHECK(hipFree(base_));
  base_ = nullptr;
  SPP_CHECK(hipMalloc(&base_, bytes));
  cap_ = bytes;
}

DeviceScratch::~DeviceScratch() {
  if (p) (void)hipFree(p);
}

double* dot_partials(Context& ctx, int nsamples) {
  const size_t need = (size_t)std::max(nsamples, 1) * kDotChunks;
  if (ctx.noalloc_dot) {
    if (need > ctx.noalloc_dot_n) throw HipError("dot-product scratch window too small");
    return ctx.noalloc_dot;
  }
  if (ctx.dot_part.n < need) {
    SPP_CHECK(hipStreamSynchronize(ctx.stream));  // earlier dot launches may still read the old buffer
    if (ctx.dot_part.p) SPP_CHECK(hipFree(ctx.dot_part.p));
    ctx.dot_part.p = nullptr;
    ctx.dot_part.n = 0;
    SPP_CHECK(hipMalloc((void**)&ctx.dot_part.p, sizeof(double) * need));
    ctx.dot_part.n = need;
  }
  return ctx.dot_part.p;
}

void* Arena::take(size_t bytes) {
  size_t aligned = (bytes + 255) & ~size_t(255);
  if (top_ + aligned > cap_)
    throw HipError("device workspace exhausted: need " + std::to_string(top_ + aligned) + " bytes, have " +
                   std::to_string(cap_));
  void* p = base_ + top_;
  top_ += aligned;
  peak_ = std::max(peak_, top_);
  return p;
}

// ------------------------------- TRef ---------------------------------------
TRef dense_ref(void* base, long long sstride, const std::string& letters, const std::vector<Leg>& legs) {
  if (letters.size() != legs.size()) throw HipError("dense_ref: letters/legs mismatch for " + letters);
  TRef t;
  t.base = base;
  t.sstride = sstride;
  t.letters = letters;
  t.leg = legs;
  t.stride.assign(legs.size(), 1);
  for (int k = (int)legs.size() - 2; k >= 0; --k) t.stride[k] = t.stride[k + 1] * legs[k + 1].dim;
  return t;
}

static int fold_dim(const Leg& l) { return std::max(l.even, l.dim - l.even); }

TRef folded_ref(void* base, long long sstride, const std::string& letters, const std::vector<Leg>& legs) {
  if (letters.size() != legs.size() || legs.empty()) throw HipError("folded_ref: letters/legs mismatch for " + letters);
  const Leg& last = legs.back();
  if (last.split || last.flat) throw HipError("folded_ref: the last leg needs a fixed even count (" + letters + ")");
  TRef t;
  t.base = base;
  t.sstride = sstride;
  t.letters = letters;
  t.leg = legs;
  t.fold = true;
  t.stride.assign(legs.size(), 1);
  long long d = fold_dim(last);
  for (int k = (int)legs.size() - 2; k >= 0; --k) {
    t.stride[k] = t.stride[k + 1] * d;
    d = legs[k].dim;
  }
  return t;
}

long long folded_count(const std::vector<Leg>& legs) {
  long long n = 1;
  for (size_t k = 0; k + 1 < legs.size(); ++k) n *= legs[k].dim;
  return legs.empty() ? 1 : n * fold_dim(legs.back());
}

TRef compact_ref(void* base, long long sstride, long long pstride, long long ld, const std::string& letters,
                 int nrow, const std::vector<Leg>& legs) {
  TRef t;
  t.base = base;
  t.sstride = sstride;
  t.pstride = pstride;
  t.letters = letters;
  t.leg = legs;
  t.stride.assign(legs.size(), 0);
  t.compact = true;
  t.nrow = nrow;
  t.ld = ld;
  return t;
}

TRef with_letters(const TRef& t, const std::string& letters) {
  if (letters.size() != t.letters.size()) throw HipError("with_letters: rank mismatch");
  TRef r = t;
  r.letters = letters;
  return r;
}

TRef with_ptrs(const TRef& t, void* const* ptrs) {
  TRef r = t;
  r.ptrs = ptrs;
  return r;
}

TRef relabel(const TRef& t, const std::string& new_letters) { return with_letters(t, new_letters); }

TRef sub_leg(const TRef& t, char letter, int r0, const Leg& chunk) {
  const size_t k = t.letters.find(letter);
  if (k == std::string::npos) throw HipError("sub_leg: no leg " + std::string(1, letter) + " in " + t.letters);
  if (t.compact || (t.fold && k + 1 == t.leg.size())) throw HipError("sub_leg: needs a dense (unfolded) leg");
  if (r0 < 0 || chunk.dim < 1 || r0 + chunk.dim > t.leg[k].dim) throw HipError("sub_leg: range outside the leg");
  TRef r = t;
  r.offset += (long long)r0 * t.stride[k];
  r.leg[k] = chunk;
  return r;
}

__global__ void chunk_split_kernel(const int* split, int r0, int c, int* out, int nb) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nb) out[s] = min(max(split[s] - r0, 0), c);
}

Leg chunk_leg(Context& ctx, const Leg& l, int r0, int c, int* split_out, int nb) {
  if (l.flat) throw HipError("chunk_leg: ungraded leg");
  Leg r = fixed_leg(c, std::min(std::max(l.even - r0, 0), c));
  if (l.split) {
    hipLaunchKernelGGL(chunk_split_kernel, dim3((nb + 255) / 256), dim3(256), 0, ctx.stream, l.split, r0, c, split_out,
                       nb);
    SPP_CHECK(hipGetLastError());
    r.split = split_out;
  }
  return r;
}

static long long group_bound(const std::vector<Leg>& legs) {
  long long p = 1;
  bool flat = true, evensplit = false;
  for (auto& l : legs) {
    p *= l.dim;
    if (!l.flat) flat = false;
    if (!l.flat && l.split == nullptr && 2 * l.even == l.dim) evensplit = true;
  }
  if (flat || legs.empty()) return p;
  return evensplit ? p / 2 : p;
}

static double group_estimate(const std::vector<Leg>& legs, int q) {
  // tuple count with per-sample splits replaced by an even split
  double c[2] = {1.0, 0.0};
  bool flat = !legs.empty();
  for (auto& l : legs)
    if (!l.flat) flat = false;
  if (flat) {
    double p = 1;
    for (auto& l : legs) p *= l.dim;
    return p;
  }
  for (auto& l : legs) {
    double e = l.split ? 0.5 * l.dim : l.even;
    double o = l.dim - e;
    double n0 = c[0] * e + c[1] * o, n1 = c[0] * o + c[1] * e;
    c[0] = n0;
    c[1] = n1;
  }
  return c[q];
}

static void fill_group(Group& g, const std::vector<Leg>& legs, const std::string& who) {
  if ((int)legs.size() > kMaxLegs) throw HipError("too many legs in group " + who);
  g.n = (int)legs.size();
  bool anyflat = false, anygraded = false;
  for (int j = 0; j < g.n; ++j) {
    g.dim[j] = legs[j].dim;
    g.even[j] = legs[j].even;
    g.split[j] = legs[j].split;
    if (legs[j].flat) anyflat = true;
    else anygraded = true;
  }
  if (anyflat && anygraded) throw HipError("mixed flat/graded group " + who);
  g.graded = anyflat ? 0 : 1;
}

static void fill_operand_group(OperandGroup& og, const TRef& t, const std::string& letters, int role_group,
                               const std::string& who) {
  og.fold_j = -1;
  og.fold_even = 0;
  if (t.compact) {
    std::string rows = t.letters.substr(0, t.nrow), cols = t.letters.substr(t.nrow);
    og.linear = 1;
    if (letters == rows) og.lstride = t.ld;
    else if (letters == cols) og.lstride = 1;
    else
      throw HipError("compact operand " + t.letters + " cannot serve group '" + letters + "' (" + who + ")");
    (void)role_group;
    return;
  }
  og.linear = 0;
  og.lstride = 0;
  for (size_t j = 0; j < letters.size(); ++j) {
    auto pos = t.letters.find(letters[j]);
    if (pos == std::string::npos) throw HipError("letter missing in operand (" + who + ")");
    og.stride[j] = t.stride[pos];
    if (t.fold && pos + 1 == t.letters.size()) {
      og.fold_j = (int)j;
      og.fold_even = t.leg[pos].even;
    }
  }
}

static void fill_operand(Operand& op, const TRef& t) {
  op.base = t.base;
  op.sstride = t.sstride;
  op.ptrs = t.ptrs;
  op.pstride = t.pstride;
  op.offset = t.offset;
}

static std::vector<Leg> legs_of(const TRef& t, const std::string& letters) {
  std::vector<Leg> out;
  for (char c : letters) out.push_back(t.leg[t.letters.find(c)]);
  return out;
}

// Times every tile shape of the family on this launch (each one writes the
// same C; the first run warms the operands) and returns the fastest, keeping
// the heuristic's shape unless another is faster by more than 2%.
static int tune_tile(Context& ctx, const GemmArgs& g, int family, int pieces, long long mmax, long long nmax) {
  const int def = z2gemm_default_tile(family, mmax, nmax);
  const int count = z2gemm_tile_count(family);
  const auto t0 = std::chrono::steady_clock::now();
  SPP_CHECK(launch_z2gemm_tile(g, family, pieces, def, mmax, nmax, ctx.stream));
  std::vector<hipEvent_t> ev(2 * count);
  for (auto& e : ev) SPP_CHECK(hipEventCreate(&e));
  std::vector<int> ok(count, 0);
  for (int t = 0; t < count; ++t) {
    SPP_CHECK(hipEventRecord(ev[2 * t], ctx.stream));
    ok[t] = launch_z2gemm_tile(g, family, pieces, t, mmax, nmax, ctx.stream) == hipSuccess;
    SPP_CHECK(hipEventRecord(ev[2 * t + 1], ctx.stream));
  }
  SPP_CHECK(hipStreamSynchronize(ctx.stream));
  std::vector<float> ms(count, 0.f);
  for (int t = 0; t < count; ++t) SPP_CHECK(hipEventElapsedTime(&ms[t], ev[2 * t], ev[2 * t + 1]));
  for (auto& e : ev) (void)hipEventDestroy(e);
  int best = def;
  for (int t = 0; t < count; ++t)
    if (ok[t] && ms[t] < 0.98f * ms[best]) best = t;
  ctx.counters.tuned_classes += 1;
  ctx.counters.tuned_changed += best != def ? 1 : 0;
  ctx.counters.tune_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return best;
}

// Tile shape of a launch: forced, tuned for its class (tuning it now when the
// class is new and the launch may be repeated), or the heuristic's. Calls in
// no-allocation mode never tune (the timing waits on the stream).
template <typename T>
static int choose_tile(Context& ctx, const GemmArgs& g, long long mmax, long long nmax,
                       const std::vector<long long>& cls, int& family, int& pieces) {
  family = sizeof(T) == 4 ? (ctx.split ? 2 : 1) : 0;
  pieces = family == 2 ? ctx.split : 0;
  int tile = z2gemm_default_tile(family, mmax, nmax);
  if (ctx.force_tile >= 0) {
    if (ctx.force_tile < z2gemm_tile_count(family)) tile = ctx.force_tile;
  } else if (ctx.autotune && mmax > 0 && nmax > 0 && g.nsamples > 0) {
    auto it = ctx.tile_table.find(cls);
    if (it != ctx.tile_table.end()) {
      tile = it->second;
    } else if (!g.accumulate && !ctx.noalloc_dot && !ctx.capturing) {
      tile = tune_tile(ctx, g, family, pieces, mmax, nmax);
      ctx.tile_table[cls] = tile;
    }
  }
  return tile;
}

template <typename T>
void contract(Context& ctx, const TRef& C, const TRef& A, const TRef& B, int nsamples, double alpha,
              bool accumulate, const int* pmask) {
  std::string M, N, K;
  auto has = [](const std::string& s, char c) { return s.find(c) != std::string::npos; };
  for (char c : C.letters) {
    bool a = has(A.letters, c), b = has(B.letters, c);
    if (a && !b) M += c;
    else if (b && !a) N += c;
    else throw HipError("contract: output letter " + std::string(1, c) + " must appear in exactly one input (" +
                        C.letters + "=" + A.letters + "*" + B.letters + ")");
  }
  for (char c : A.letters)
    if (!has(C.letters, c)) {
      if (!has(B.letters, c)) throw HipError("contract: dangling letter in A");
      K += c;
    }
  for (char c : B.letters)
    if (!has(C.letters, c) && !has(A.letters, c)) throw HipError("contract: dangling letter in B");
  // K order follows A, except that compact A or B dictate it
  if (B.compact && !A.compact) {
    std::string rows = B.letters.substr(0, B.nrow);
    if (rows.size() == K.size()) K = rows;
  }
  std::string who = C.letters + "=" + A.letters + "*" + B.letters;
  GemmArgs g{};
  auto lm = legs_of(A, M), ln = legs_of(B, N), lk = legs_of(A, K);
  fill_group(g.gm, lm, who);
  fill_group(g.gn, ln, who);
  fill_group(g.gk, lk, who);
  fill_operand(g.a, A);
  fill_operand(g.b, B);
  fill_operand(g.c, C);
  fill_operand_group(g.a.g[0], A, M, 0, who);
  fill_operand_group(g.a.g[1], A, K, 1, who);
  fill_operand_group(g.b.g[0], B, K, 0, who);
  fill_operand_group(g.b.g[1], B, N, 1, who);
  fill_operand_group(g.c.g[0], C, M, 0, who);
  fill_operand_group(g.c.g[1], C, N, 1, who);
  // lanes follow the unit-stride index of each operand
  auto last_stride = [](const OperandGroup& og, int nlegs) -> long long {
    if (og.linear) return og.lstride;
    return nlegs > 0 ? og.stride[nlegs - 1] : (1LL << 62);
  };
  // extents from a (sample, parity) base: the kernel addresses operands with
  // 32-bit offsets and reads past-the-end K indices as zeros
  auto extent = [](const TRef& t) -> unsigned long long {
    if (t.compact) return (unsigned long long)t.pstride;
    unsigned long long e = 1;
    for (size_t j = 0; j < t.leg.size(); ++j) {
      const int d = (t.fold && j + 1 == t.leg.size()) ? fold_dim(t.leg[j]) : t.leg[j].dim;
      e += (unsigned long long)(d - 1) * t.stride[j];
    }
    return e;
  };
  const unsigned long long ea = extent(A) * sizeof(T), eb = extent(B) * sizeof(T), ec = extent(C);
  if (ea >= (1ULL << 31) || eb >= (1ULL << 31) || ec >= (1ULL << 31))
    throw HipError("contract: an operand exceeds 2^31 bytes per sample (" + who + ")");
  g.a_bytes = (unsigned)ea;
  g.b_bytes = (unsigned)eb;
  g.a_mfast = last_stride(g.a.g[0], (int)M.size()) < last_stride(g.a.g[1], (int)K.size()) ? 1 : 0;
  g.b_kfast = last_stride(g.b.g[0], (int)K.size()) < last_stride(g.b.g[1], (int)N.size()) ? 1 : 0;
  g.nsamples = nsamples;
  g.sample_map = nullptr;
  g.pmask = pmask;
  g.alpha = alpha;
  g.accumulate = accumulate ? 1 : 0;
  long long mmax = group_bound(lm), nmax = group_bound(ln);
  if (mmax >= (1LL << 31) || nmax >= (1LL << 31) || group_bound(lk) >= (1LL << 31))
    throw HipError("contract: a leg group exceeds 2^31 tuples (" + who + ")");
  double fl = 0;
  for (int q = 0; q < 2; ++q) fl += group_estimate(lm, q) * group_estimate(ln, q) * group_estimate(lk, q);
  ctx.counters.gemm_flops += 2.0 * fl * nsamples;
  ctx.counters.gemm_launches += 1;
  // launch class of the tile autotuner: family, load modes, samples and the
  // block structure of every leg (dimension; fixed even count or -1 per sample)
  std::vector<long long> cls = {(long long)sizeof(T), ctx.split, g.a_mfast, g.b_kfast, nsamples, mmax, nmax};
  for (const auto* grp : {&lm, &ln, &lk}) {
    cls.push_back(-2);
    for (const Leg& l : *grp) {
      cls.push_back(l.dim);
      cls.push_back(l.split ? -1 : l.even);
    }
  }
  int family = 0, pieces = 0;
  const int tile = choose_tile<T>(ctx, g, mmax, nmax, cls, family, pieces);
  if (ctx.profile) {
    hipEvent_t a, b;
    SPP_CHECK(hipEventCreate(&a));
    SPP_CHECK(hipEventCreate(&b));
    SPP_CHECK(hipEventRecord(a, ctx.stream));
    SPP_CHECK(launch_z2gemm_tile(g, family, pieces, tile, mmax, nmax, ctx.stream));
    SPP_CHECK(hipEventRecord(b, ctx.stream));
    ctx.timer.pending.emplace_back(a, b);
    ctx.timer.pending_flops.push_back(2.0 * fl * nsamples);
    ctx.timer.pending_prec.push_back(sizeof(T) == 4 ? 1 : 0);
    if (ctx.timer.pending.size() >= 4096) ctx.flush_timer();
    return;
  }
  SPP_CHECK(launch_z2gemm_tile(g, family, pieces, tile, mmax, nmax, ctx.stream));
}

void Context::flush_timer() {
  if (timer.pending.empty()) return;
  SPP_CHECK(hipStreamSynchronize(stream));
  for (size_t i = 0; i < timer.pending.size(); ++i) {
    float ms = 0.f;
    SPP_CHECK(hipEventElapsedTime(&ms, timer.pending[i].first, timer.pending[i].second));
    counters.gemm_ms += ms;
    counters.gemm_flops_profiled += timer.pending_flops[i];
    const int q = timer.pending_prec[i];
    counters.prec_launches[q] += 1;
    counters.prec_ms[q] += ms;
    counters.prec_flops[q] += timer.pending_flops[i];
    (void)hipEventDestroy(timer.pending[i].first);
    (void)hipEventDestroy(timer.pending[i].second);
  }
  timer.pending.clear();
  timer.pending_flops.clear();
  timer.pending_prec.clear();
}

template void contract<float>(Context&, const TRef&, const TRef&, const TRef&, int, double, bool, const int*);
template void contract<double>(Context&, const TRef&, const TRef&, const TRef&, int, double, bool, const int*);

FullView full_view(const TRef& t) {
  if (t.compact) throw HipError("full_view of a compact operand");
  if (t.fold) throw HipError("full_view of a folded operand");
  if (t.offset) throw HipError("full_view of a sub-range view");
  FullView v{};
  fill_group(v.g, t.leg, "full_view");
  for (size_t j = 0; j < t.leg.size(); ++j) v.stride[j] = t.stride[j];
  v.base = t.base;
  v.sstride = t.sstride;
  v.ptrs = t.ptrs;
  return v;
}

// ------------------------------ boundaries ---------------------------------
template <typename T>
void Boundary<T>::release() {
  if (owns) {
    for (auto p : t)
      if (p) (void)hipFree(p);
    for (auto p : bsplit)
      if (p) (void)hipFree(p);
    if (log_scale) (void)hipFree(log_scale);
    if (max_trunc) (void)hipFree(max_trunc);
  }
  t.clear();
  bsplit.clear();
  log_scale = max_trunc = nullptr;
}
template struct Boundary<float>;
template struct Boundary<double>;

template <typename T>
TRef RowLayers<T>::ref(int x, const std::string& letters) const {
  const LayerSpec& sp = spec[x];
  // spec is given in this row's frame; storage of a rotated row is E[e,s,w,n]
  TRef r;
  if (!rotated) {
    r = dense_ref(shared.empty() ? nullptr : shared[x], 0, letters, {sp.w, sp.n, sp.e, sp.s});
  } else {
    std::string st = {letters[2], letters[3], letters[0], letters[1]};
    r = dense_ref(shared.empty() ? nullptr : shared[x], 0, st, {sp.e, sp.s, sp.w, sp.n});
  }
  if (!ptrs.empty() && ptrs[x]) r.ptrs = ptrs[x];
  return r;
}
template struct RowLayers<float>;
template struct RowLayers<double>;

template <typename T>
void trivial_boundary(Context& ctx, Boundary<T>& b, int L, int nb, const std::vector<Leg>& phys) {
  b.L = L;
  b.nb = nb;
  b.bdim.assign(L + 1, 1);
  b.beven.assign(L + 1, 1);
  b.bsplit.assign(L + 1, nullptr);
  b.pdim.resize(L);
  b.peven.resize(L);
  b.t.assign(L, nullptr);
  for (int x = 0; x < L; ++x) {
    if (phys[x].dim != 1) throw HipError("nontrivial OBC leg");
    b.pdim[x] = 1;
    b.peven[x] = phys[x].even;
    SPP_CHECK(hipMalloc(&b.t[x], sizeof(T) * nb));
    SPP_CHECK(launch_fill<T>(b.t[x], nb, T(1), ctx.stream));
  }
  SPP_CHECK(hipMalloc(&b.log_scale, sizeof(double) * nb));
  SPP_CHECK(hipMalloc(&b.max_trunc, sizeof(double) * nb));
  SPP_CHECK(launch_fill<double>(b.log_scale, nb, 0.0, ctx.stream));
  SPP_CHECK(launch_fill<double>(b.max_trunc, nb, 0.0, ctx.stream));
  b.owns = true;
}
template void trivial_boundary<float>(Context&, Boundary<float>&, int, int, const std::vector<Leg>&);
template void trivial_boundary<double>(Context&, Boundary<double>&, int, int, const std::vector<Leg>&);

// ------------------------- truncation helper kernels ------------------------

// Tuple counts of a group for every (sample, parity) problem.
__global__ void group_counts_kernel(Group g, int* out, int nprob) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nprob) return;
  GroupView v;
  resolve_group(g, p >> 1, v);
  out[p] = (int)group_count(v, p & 1);
}

__device__ inline unsigned long long mix64(unsigned long long v) {
  v += 0x9E3779B97F4A7C15ULL;
  v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ULL;
  v = (v ^ (v >> 27)) * 0x94D049BB133111EBULL;
  return v ^ (v >> 31);
}

// One pass of (shifted) Cholesky QR on the Gram matrices S = Z^T Z of a batch
// of n x n problems (n <= 64): R = D * inv(L)^T where D scales the columns of
// Z to unit norm (scale != 0; identity otherwise) and L L^T = D S D + shift I.
// Then Z R has orthonormal columns to the accuracy the conditioning allows.
// A problem that is inactive or already marked bad is skipped; a dead column
// (norm below 1e-150 of the largest, or not finite) or a non-positive pivot
// marks it bad (bad[p] = 1, ok[p] = 0) for a Householder fallback.
constexpr int CQ = 64;
__global__ __launch_bounds__(CQ) void cholqr_kernel(const double* S, long long sstride, int lds, double* R,
                                                     long long rstride, int ldr, int n, int scale, double shift,
                                                     const int* active, int* ok, int* bad) {
  const int prob = blockIdx.x;
  const int i = threadIdx.x;
  if (!active[prob] || bad[prob]) {
    if (i == 0) ok[prob] = 0;
    return;
  }
  __shared__ double A[CQ * (CQ + 1)];
  __shared__ double X[CQ * (CQ + 1)];
  __shared__ double d[CQ];
  __shared__ double smax;
  __shared__ int fail;
  const double* sp = S + (long long)prob * sstride;
  if (i == 0) {
    fail = 0;
    double m = 0.0;
    for (int k = 0; k < n; ++k) m = fmax(m, sp[(long long)k * lds + k]);
    smax = m;
  }
  __syncthreads();
  if (i < n) {
    const double sii = sp[(long long)i * lds + i];
    double di = 1.0;
    if (scale) {
      if (!(sii > 1e-150 * smax) || !isfinite(sii)) {
        fail = 1;
        di = 0.0;
      } else {
        di = 1.0 / sqrt(sii);
      }
    }
    d[i] = di;
  }
  __syncthreads();
  if (i < n)
    for (int j = 0; j < n; ++j) A[i * (CQ + 1) + j] = d[i] * d[j] * sp[(long long)i * lds + j] + (i == j ? shift : 0.0);
  __syncthreads();
  // right-looking Cholesky, thread i owns row i of the lower triangle
  for (int k = 0; k < n; ++k) {
    if (i == k) {
      const double p = A[k * (CQ + 1) + k];
      if (!(p > 0.0) || !isfinite(p)) fail = 1;
      A[k * (CQ + 1) + k] = p > 0.0 ? sqrt(p) : 1.0;
    }
    __syncthreads();
    if (i > k && i < n) A[i * (CQ + 1) + k] /= A[k * (CQ + 1) + k];
    __syncthreads();
    if (i > k && i < n) {
      const double lik = A[i * (CQ + 1) + k];
      for (int m = k + 1; m <= i; ++m) A[i * (CQ + 1) + m] -= lik * A[m * (CQ + 1) + k];
    }
    __syncthreads();
  }
  // X = inv(L), column i by thread i (forward substitution)
  if (i < n) {
    for (int r = 0; r < n; ++r) {
      double v = (r == i) ? 1.0 : 0.0;
      if (r >= i) {
        for (int m = i; m < r; ++m) v -= A[r * (CQ + 1) + m] * X[m * (CQ + 1) + i];
        v /= A[r * (CQ + 1) + r];
      } else {
        v = 0.0;
      }
      X[r * (CQ + 1) + i] = v;
    }
  }
  __syncthreads();
  double* rp = R + (long long)prob * rstride;
  if (i < n)
    for (int b = 0; b < n; ++b) rp[(long long)i * ldr + b] = (b >= i) ? d[i] * X[b * (CQ + 1) + i] : 0.0;
  if (i == 0) {
    ok[prob] = fail ? 0 : 1;
    if (fail) bad[prob] = 1;
  }
}

// Second orthonormalization pass for a nearly orthonormal block Q (Gram
// S = Q^T Q = I + E): R = (3 I - S) / 2, so that Q R has Gram I + O(E^2).
// A problem with max |E| above 1e-7 (or nonfinite) is marked bad.
__global__ void lowdin_kernel(const double* S, long long sstride, int lds, double* R, long long rstride, int ldr,
                              int n, const int* active, int* ok, int* bad) {
  const int prob = blockIdx.x;
  if (!active[prob] || bad[prob]) {
    if (threadIdx.x == 0) ok[prob] = 0;
    return;
  }
  __shared__ int fail;
  if (threadIdx.x == 0) fail = 0;
  __syncthreads();
  const double* sp = S + (long long)prob * sstride;
  double* rp = R + (long long)prob * rstride;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e - (e / n) * n;
    const double sij = sp[(long long)i * lds + j];
    const double eij = sij - (i == j ? 1.0 : 0.0);
    if (!(fabs(eij) <= 1e-7)) fail = 1;
    rp[(long long)i * ldr + j] = (i == j ? 1.0 : 0.0) - 0.5 * eij;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    ok[prob] = fail ? 0 : 1;
    if (fail) bad[prob] = 1;
  }
}

// out[p] = a[p] && b[p]
__global__ void and_mask_kernel(const int* a, const int* b, int* out, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) out[p] = (a[p] && b[p]) ? 1 : 0;
}

// Deterministic pseudo-random start block (rows >= n zero).
__global__ void random_block_kernel(double* q, long long pstride, int ld, int ncol, const int* n_of, int maxrows) {
  const int prob = blockIdx.y;
  const int n = n_of[prob];
  double* b = q + (long long)prob * pstride;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)maxrows * ncol;
       i += (long long)gridDim.x * blockDim.x) {
    int r = (int)(i / ncol), c = (int)(i % ncol);
    double v = 0.0;
    if (r < n) {
      // seeded by the parity only: a sample's start block (and so its result)
      // does not depend on where it sits in the batch
      unsigned long long h =
          mix64(((unsigned long long)(prob & 1) << 40) ^ ((unsigned long long)r << 20) ^ (unsigned long long)c);
      v = ((double)(h >> 11) * 0x1.0p-53) - 0.5;
    }
    b[(long long)r * ld + c] = v;
  }
}

// Thin Householder QR, explicit Q: Z (n x ncol, ld) -> Q (n x ncol, ld) with
// orthonormal columns spanning range(Z) (LAPACK geqrf/orgqr steps, one
// workgroup per problem, panel kept in global memory). Backward stable, so
// small eigen-directions of the subspace survive (a Gram-matrix
// orthonormalisation would lose those below sqrt(eps)). Z is overwritten by
// the reflectors; tau needs ncol entries per problem.
__global__ __launch_bounds__(256) void householder_q_kernel(double* z, double* q, long long pstride, int ld,
                                                           int ncol, const int* n_of, double* tau_all,
                                                           const int* pmask) {
  const int prob = blockIdx.x;
  if (pmask && !pmask[prob]) return;
  const int n = n_of[prob];
  double* Z = z + (long long)prob * pstride;
  double* Q = q + (long long)prob * pstride;
  double* tau = tau_all + (long long)prob * ncol;
  const int tid = threadIdx.x;
  __shared__ double red[256];
  __shared__ double wcol[64];
  __shared__ double sh_tau, sh_beta, sh_scale;
  const int kmax = min(ncol, n);
  // ---- factorisation ----
  for (int j = 0; j < kmax; ++j) {
    double acc = 0.0;
    for (int r = j + 1 + tid; r < n; r += 256) {
      double v = Z[(long long)r * ld + j];
      acc += v * v;
    }
    red[tid] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (tid < w) red[tid] += red[tid + w];
      __syncthreads();
    }
    if (tid == 0) {
      double alpha = Z[(long long)j * ld + j];
      double sig = red[0];
      if (sig == 0.0) {
        sh_tau = 0.0;
        sh_beta = alpha;
        sh_scale = 0.0;
      } else {
        double nrm = sqrt(alpha * alpha + sig);
        double beta = alpha <= 0.0 ? nrm : -nrm;
        sh_tau = (beta - alpha) / beta;
        sh_scale = 1.0 / (alpha - beta);
        sh_beta = beta;
      }
      tau[j] = sh_tau;
    }
    __syncthreads();
    const double tj = sh_tau, sc = sh_scale;
    for (int r = j + 1 + tid; r < n; r += 256) Z[(long long)r * ld + j] *= sc;
    __syncthreads();
    if (tid == 0) Z[(long long)j * ld + j] = sh_beta;
    // w_c = v^T Z[:, c] for c > j, with v_j = 1
    const int nc = ncol - j - 1;
    if (tj != 0.0 && nc > 0) {
      for (int c0 = 0; c0 < nc; c0 += 64) {
        const int c = c0 + (tid & 63);
        const int part = tid >> 6;
        double a2 = 0.0;
        if (c < nc) {
          const int cc = j + 1 + c;
          if (part == 0) a2 = Z[(long long)j * ld + cc];
          for (int r = j + 1 + part; r < n; r += 4) a2 += Z[(long long)r * ld + j] * Z[(long long)r * ld + cc];
        }
        red[tid] = a2;
        __syncthreads();
        if (part == 0 && c < nc) wcol[tid & 63] = red[tid] + red[tid + 64] + red[tid + 128] + red[tid + 192];
        __syncthreads();
        const int cw = min(64, nc - c0);
        for (int idx = tid; idx < (n - j) * cw; idx += 256) {
          int rr = j + idx / cw, ci = idx % cw;
          int cc = j + 1 + c0 + ci;
          double v = rr == j ? 1.0 : Z[(long long)rr * ld + j];
          Z[(long long)rr * ld + cc] -= tj * v * wcol[ci];
        }
        __syncthreads();
      }
    }
  }
  // ---- explicit thin Q = H_0 ... H_{k-1} [I; 0] ----
  for (int idx = tid; idx < n * ncol; idx += 256) {
    int r = idx / ncol, c = idx % ncol;
    Q[(long long)r * ld + c] = (r == c && c < kmax) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int j = kmax - 1; j >= 0; --j) {
    const double tj = tau[j];
    if (tj == 0.0) continue;
    const int nc = ncol - j;
    for (int c0 = 0; c0 < nc; c0 += 64) {
      const int c = c0 + (tid & 63);
      const int part = tid >> 6;
      double a2 = 0.0;
      if (c < nc) {
        const int cc = j + c;
        if (part == 0) a2 = Q[(long long)j * ld + cc];
        for (int r = j + 1 + part; r < n; r += 4) a2 += Z[(long long)r * ld + j] * Q[(long long)r * ld + cc];
      }
      red[tid] = a2;
      __syncthreads();
      if (part == 0 && c < nc) wcol[tid & 63] = red[tid] + red[tid + 64] + red[tid + 128] + red[tid + 192];
      __syncthreads();
      const int cw = min(64, nc - c0);
      for (int idx = tid; idx < (n - j) * cw; idx += 256) {
        int rr = j + idx / cw, ci = idx % cw;
        int cc = j + c0 + ci;
        double v = rr == j ? 1.0 : Z[(long long)rr * ld + j];
        Q[(long long)rr * ld + cc] -= tj * v * wcol[ci];
      }
      __syncthreads();
    }
  }
}

// Register-resident variant of the thin Householder QR above for n <= NSLAB*RMAX
// rows and ncol <= 64 columns: 64*NSLAB threads, thread (slab, c) = (t / 64, t % 64)
// holds column c of rows slab + NSLAB i. Column reductions go through LDS
// partials; the reflector of step j is broadcast through LDS. Same arithmetic
// as householder_q_kernel.
template <int NSLAB, int RMAX>
__global__ __launch_bounds__(64 * NSLAB) void householder_q_reg_kernel(double* z, double* q, long long pstride, int ld,
                                                                 int ncol, const int* n_of, const int* pmask) {
  const int prob = blockIdx.x;
  if (pmask && !pmask[prob]) return;
  const int n = n_of[prob];
  double* Z = z + (long long)prob * pstride;
  double* Q = q + (long long)prob * pstride;
  const int tid = threadIdx.x;
  const int c = tid & 63, slab = tid >> 6;
  __shared__ double vsh[NSLAB * RMAX];
  __shared__ double part[NSLAB][64];
  __shared__ double wsh[64];
  __shared__ double tau[64];
  __shared__ double sh_scale, sh_beta;
  double zr[RMAX];
#pragma unroll
  for (int i = 0; i < RMAX; ++i) {
    const int r = slab + NSLAB * i;
    zr[i] = (r < n && c < ncol) ? Z[(long long)r * ld + c] : 0.0;
  }
  const int kmax = min(ncol, n);
  for (int j = 0; j < kmax; ++j) {
    // squared norm of column j below the diagonal
    if (c == j) {
      double acc = 0.0;
#pragma unroll
      for (int i = 0; i < RMAX; ++i) {
        const int r = slab + NSLAB * i;
        if (r > j && r < n) acc += zr[i] * zr[i];
      }
      part[slab][0] = acc;
#pragma unroll
      for (int i = 0; i < RMAX; ++i)
        if (slab + NSLAB * i == j) wsh[0] = zr[i];  // alpha = Z[j][j] (static register index)
    }
    __syncthreads();
    if (tid == 0) {
      double sig = 0.0;
      for (int k = 0; k < NSLAB; ++k) sig += part[k][0];
      const double alpha = wsh[0];
      if (sig == 0.0) {
        tau[j] = 0.0;
        sh_beta = alpha;
        sh_scale = 0.0;
      } else {
        const double nrm = sqrt(alpha * alpha + sig);
        const double beta = alpha <= 0.0 ? nrm : -nrm;
        tau[j] = (beta - alpha) / beta;
        sh_scale = 1.0 / (alpha - beta);
        sh_beta = beta;
      }
    }
    __syncthreads();
    const double tj = tau[j];
    if (c == j) {
#pragma unroll
      for (int i = 0; i < RMAX; ++i) {
        const int r = slab + NSLAB * i;
        if (r > j && r < n) {
          zr[i] *= sh_scale;
          vsh[r] = zr[i];
        } else if (r == j) {
          vsh[r] = 1.0;
          zr[i] = sh_beta;
        } else if (r < n) {
          vsh[r] = 0.0;
        }
      }
    }
    __syncthreads();
    if (tj != 0.0) {
      // w_c = v^T Z[:, c] for c > j
      double acc = 0.0;
      if (c > j && c < ncol) {
#pragma unroll
        for (int i = 0; i < RMAX; ++i) {
          const int r = slab + NSLAB * i;
          if (r >= j && r < n) acc += vsh[r] * zr[i];
        }
      }
      part[slab][c] = acc;
      __syncthreads();
      if (slab == 0) {
        double w = 0.0;
        for (int k = 0; k < NSLAB; ++k) w += part[k][c];
        wsh[c] = w;
      }
      __syncthreads();
      if (c > j && c < ncol) {
        const double f = tj * wsh[c];
#pragma unroll
        for (int i = 0; i < RMAX; ++i) {
          const int r = slab + NSLAB * i;
          if (r >= j && r < n) zr[i] -= f * vsh[r];
        }
      }
    }
    __syncthreads();
  }
  // keep the reflectors (below the diagonal) in Z for the Q formation
#pragma unroll
  for (int i = 0; i < RMAX; ++i) {
    const int r = slab + NSLAB * i;
    if (r < n && c < ncol) Z[(long long)r * ld + c] = zr[i];
  }
  // Q = H_0 ... H_{k-1} [I; 0], panel in the same registers
#pragma unroll
  for (int i = 0; i < RMAX; ++i) {
    const int r = slab + NSLAB * i;
    zr[i] = (r == c && c < kmax) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int j = kmax - 1; j >= 0; --j) {
    const double tj = tau[j];
    if (tj == 0.0) continue;
    for (int r = tid; r < n; r += 64 * NSLAB) vsh[r] = r < j ? 0.0 : (r == j ? 1.0 : Z[(long long)r * ld + j]);
    __syncthreads();
    double acc = 0.0;
    if (c >= j && c < ncol) {
#pragma unroll
      for (int i = 0; i < RMAX; ++i) {
        const int r = slab + NSLAB * i;
        if (r >= j && r < n) acc += vsh[r] * zr[i];
      }
    }
    part[slab][c] = acc;
    __syncthreads();
    if (slab == 0) {
      double w = 0.0;
      for (int k = 0; k < NSLAB; ++k) w += part[k][c];
      wsh[c] = w;
    }
    __syncthreads();
    if (c >= j && c < ncol) {
      const double f = tj * wsh[c];
#pragma unroll
      for (int i = 0; i < RMAX; ++i) {
        const int r = slab + NSLAB * i;
        if (r >= j && r < n) zr[i] -= f * vsh[r];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < RMAX; ++i) {
    const int r = slab + NSLAB * i;
    if (r < n && c < ncol) Q[(long long)r * ld + c] = zr[i];
  }
}

// Residuals of Ritz pairs: conv[prob] = max_j<t ||Z_j - theta_j Q_j|| <= tol * |theta_0|.
// Residual test of the Ritz pairs; a problem that converges leaves `active`
// and keeps its Ritz pairs (later iterations skip it).
// log T_k(x) of the Chebyshev polynomial of the first kind, x >= 1
__device__ inline double log_cheb(int k, double x) {
  const double a = k * acosh(x);
  return a + log1p(exp(-2.0 * a)) - 0.69314718055994531;
}

__global__ void residual_kernel(const double* q, const double* z, long long pstride, int ld, const double* th,
                                long long tstride, const int* n_of, int t, double tol, int* conv, double* worst,
                                int* active, int kk, int maxdeg, int* deg) {
  const int prob = blockIdx.x;
  if (!active[prob]) return;
  const int n = n_of[prob];
  const double* qb = q + (long long)prob * pstride;
  const double* zb = z + (long long)prob * pstride;
  const double* tb = th + (long long)prob * tstride;
  __shared__ double part[256];
  __shared__ double mx;
  if (threadIdx.x == 0) mx = 0.0;
  __syncthreads();
  const int col = threadIdx.x % 64, grp = threadIdx.x / 64;
  for (int c0 = 0; c0 < t; c0 += 64) {
    int c = c0 + col;
    double acc = 0.0;
    if (c < t)
      for (int r = grp; r < n; r += 4) {
        double d = zb[(long long)r * ld + c] - tb[c] * qb[(long long)r * ld + c];
        acc += d * d;
      }
    part[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 0; k < 64 && c0 + k < t; ++k) {
        double s = sqrt(part[k] + part[k + 64] + part[k + 128] + part[k + 192]);
        if (s > mx) mx = s;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double scale = n > 0 ? fabs(tb[0]) : 0.0;
    double rel = scale > 0.0 ? mx / scale : 0.0;
    conv[prob] = (rel <= tol) ? 1 : 0;
    worst[prob] = rel;
    active[prob] = (rel <= tol) ? 0 : 1;
    // Chebyshev filter degree for the next step: damp [0, a] with a the
    // smallest Ritz value, the largest degree m <= maxdeg with
    // T_m(x_1) / T_m(x_t) <= 1e8 (x = 2 lambda / a - 1), so that the filtered
    // columns stay within eight decades of one another; 1 = plain power step.
    int m = 1;
    if (n > kk && t > 0 && t < kk) {
      const double a = tb[kk - 1], l1 = tb[0], lt = tb[t - 1];
      if (a > 1e-300 && lt > a * (1.0 + 1e-3) && l1 >= lt) {
        const double x1 = 2.0 * l1 / a - 1.0, xt = 2.0 * lt / a - 1.0;
        for (int k = 2; k <= maxdeg; ++k) {
          if (log_cheb(k, x1) - log_cheb(k, xt) > 18.420680743952367) break;
          m = k;
        }
      }
    }
    deg[prob] = m;
  }
}

// One step of the Chebyshev three-term recurrence on [0, a], a = th[kk-1] of
// each problem (c = e = a / 2):
//   j == 0: out = deg >= 2 ? (z - c y) / e : z                  (Y_1, or the plain step)
//   j >= 1: out = j < deg ? 2 (z - c y) / e - yprev : y         (Y_{j+1}, or carry Y_j)
// z = rho y where needed. out may alias yprev.
__global__ void cheb_step_kernel(const double* z, const double* y, const double* yprev, double* out,
                                 long long pstride, long long count, const double* th, long long tstride, int kk,
                                 const int* deg, const int* active, int j) {
  const int prob = blockIdx.y;
  if (!active[prob]) return;
  const int m = deg[prob];
  const double a = th[(long long)prob * tstride + kk - 1];
  const double c = 0.5 * a, inv_e = 2.0 / a;
  const double* zb = z + (long long)prob * pstride;
  const double* yb = y + (long long)prob * pstride;
  const double* pb = yprev ? yprev + (long long)prob * pstride : nullptr;
  double* ob = out + (long long)prob * pstride;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (long long)gridDim.x * blockDim.x) {
    double v;
    if (j == 0) v = m >= 2 ? (zb[i] - c * yb[i]) * inv_e : zb[i];
    else v = j < m ? 2.0 * (zb[i] - c * yb[i]) * inv_e - pb[i] : yb[i];
    ob[i] = v;
  }
}

// out[p] = active[p] && deg[p] > j
__global__ void degree_mask_kernel(const int* active, const int* deg, int j, int* out, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) out[p] = (active[p] && deg[p] > j) ? 1 : 0;
}


// rho = (A + A^T) / 2 in FP64 from the FP32 compact blocks A (leading n x n
// of each problem, leading dimension ld), through 32 x 32 LDS tiles: block x
// of problem y takes the tile pair (ti, tj), ti <= tj, of the upper triangle,
// reads both tiles and writes both. The same arithmetic as converting to FP64
// first and averaging there.
__global__ void symmetrize_convert_kernel(const float* a, double* rho, long long pstride, int ld, const int* n_of) {
  __shared__ float sa[32][33], sb[32][33];
  const int prob = blockIdx.y;
  const int n = n_of[prob];
  const int T = (ld + 31) / 32;
  int k = blockIdx.x, ti = 0;
  while (ti < T && k >= T - ti) {
    k -= T - ti;
    ++ti;
  }
  const int tj = ti + k;
  if (ti >= T || ti * 32 >= n || tj * 32 >= n) return;
  const float* ab = a + (long long)prob * pstride;
  double* rb = rho + (long long)prob * pstride;
  const int c = threadIdx.x & 31, r0 = threadIdx.x >> 5;
  for (int r = r0; r < 32; r += 8) {
    const int i = ti * 32 + r, j = tj * 32 + c;
    sa[r][c] = (i < n && j < n) ? ab[(long long)i * ld + j] : 0.f;
    const int i2 = tj * 32 + r, j2 = ti * 32 + c;
    sb[r][c] = (i2 < n && j2 < n) ? ab[(long long)i2 * ld + j2] : 0.f;
  }
  __syncthreads();
  for (int r = r0; r < 32; r += 8) {
    const int i = ti * 32 + r, j = tj * 32 + c;
    if (i < n && j < n) rb[(long long)i * ld + j] = 0.5 * ((double)sa[r][c] + (double)sb[c][r]);
    const int i2 = tj * 32 + r, j2 = ti * 32 + c;
    if (ti != tj && i2 < n && j2 < n) rb[(long long)i2 * ld + j2] = 0.5 * ((double)sb[r][c] + (double)sa[c][r]);
  }
}

// Trace of each parity block of rho (compact, ld).
__global__ void trace_kernel(const double* rho, long long pstride, int ld, const int* n_of, double* tr) {
  const int prob = blockIdx.x;
  const int n = n_of[prob];
  const double* b = rho + (long long)prob * pstride;
  __shared__ double part[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) acc += b[(long long)i * ld + i];
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) tr[prob] = part[0];
}

// Select the top chi eigenpairs over both parity blocks (truncrank, ties to
// parity 0 and lower index) intersected with the sector budget caps0/caps1
// (truncspace; negative: none), scatter them into U[k,p,c] (dense, zeroed
// beforehand, bond of capacity bcap) with the new bond parity-sorted, and
// record the new even count, the discarded weight fraction and its running
// maximum. cap bounds the ranked count (the rank of the cut).
template <typename T>
__global__ void truncate_select_kernel(Group gkp, const double* w, long long wstride, const double* v,
                                       long long vstride, int ldv, const int* avail_of, const int* n_of,
                                       const double* tr, int chi, int cap, int caps0, int caps1, int bcap, T* U,
                                       long long ustride, long long ustride_k, long long ustride_p, int* new_even,
                                       double* maxtrunc, double* discarded_out) {
  const int s = blockIdx.x;
  __shared__ int m0s, m1s;
  __shared__ GroupView g;
  if (threadIdx.x == 0) {
    resolve_group(gkp, s, g);
    const double* w0 = w + (long long)(2 * s) * wstride;
    const double* w1 = w + (long long)(2 * s + 1) * wstride;
    int a0 = min(avail_of[2 * s], n_of[2 * s]);
    int a1 = min(avail_of[2 * s + 1], n_of[2 * s + 1]);
    int want = min(chi, min(cap, a0 + a1));
    int i0 = 0, i1 = 0;
    while (i0 + i1 < want) {
      bool take0;
      if (i0 >= a0) take0 = false;
      else if (i1 >= a1) take0 = true;
      else take0 = (w0[i0] >= w1[i1]);
      if (take0) ++i0;
      else ++i1;
    }
    if (caps0 >= 0) i0 = min(i0, caps0);
    if (caps1 >= 0) i1 = min(i1, caps1);
    i1 = min(i1, bcap);  // the bond holds both parities' kept columns
    i0 = min(i0, bcap - i1);
    double kept = 0.0;
    for (int i = 0; i < i0; ++i) kept += fmax(w0[i], 0.0);
    for (int i = 0; i < i1; ++i) kept += fmax(w1[i], 0.0);
    m0s = i0;
    m1s = i1;
    double total = tr[2 * s] + tr[2 * s + 1];
    double disc = total - kept;
    if (disc < 0.0) disc = 0.0;
    double frac = total > 0.0 ? disc / total : 0.0;
    if (frac > maxtrunc[s]) maxtrunc[s] = frac;
    if (discarded_out) discarded_out[s] = frac;
    new_even[s] = bcap - i1;
  }
  __syncthreads();
  const int m0 = m0s, m1 = m1s;
  T* Ub = U + (long long)s * ustride;
  int legs[kMaxLegs];
  for (int q = 0; q < 2; ++q) {
    const int n = n_of[2 * s + q];
    const int m = q == 0 ? m0 : m1;
    const int cbase = q == 0 ? 0 : bcap - m1;
    const double* vb = v + (long long)(2 * s + q) * vstride;
    for (long long i = threadIdx.x; i < (long long)n * m; i += blockDim.x) {
      int r = (int)(i / m), j = (int)(i % m);
      decode(g, q, r, legs);
      Ub[legs[0] * ustride_k + legs[1] * ustride_p + cbase + j] = (T)vb[(long long)r * ldv + j];
    }
  }
}

// log_scale[s] += log(norm), factor[s] = 1/norm (0 when the norm vanishes)
__global__ void norm_finish_kernel(const double* nrm2, double* log_scale, double* factor, int* status, int nb) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nb) return;
  double n = sqrt(nrm2[s]);
  if (!isfinite(n)) {
    if (status) status[s] = 3;  // nonfinite boundary norm
    factor[s] = 0.0;
    return;
  }
  if (n == 0.0) {
    log_scale[s] = -INFINITY;
    factor[s] = 0.0;
  } else {
    log_scale[s] += log(n);
    factor[s] = 1.0 / n;
  }
}

// --------------------------- truncation driver ------------------------------
// Relative Ritz residual at which a truncation eigensolve of an FP32-formed rho stops.
constexpr double kFp32RhoTol = 1e-6;

struct EigResult {
  double* w;        // [nprob][wld]
  long long wstride;
  double* v;        // [nprob] compact rows x cols, ld
  long long vstride;
  int ldv;
  int* avail;       // [nprob]
};

static void eig_topk(Context& ctx, const double* rho, long long rho_pstride, int rho_ld, const Leg& lk,
                     const Leg& lp, const int* n_of, int nmax, int nb, int want, EigResult& res, double tol) {
  const int nprob = 2 * nb;
  hipStream_t st = ctx.stream;
  if (nmax <= ctx.ksub) {
    // direct: full Jacobi on each block
    res.w = (double*)ctx.arena.take(sizeof(double) * nprob * std::max(nmax, 1));
    res.wstride = std::max(nmax, 1);
    res.ldv = std::max(nmax, 1);
    res.vstride = (long long)res.ldv * res.ldv;
    res.v = (double*)ctx.arena.take(sizeof(double) * nprob * res.vstride);
    res.avail = const_cast<int*>(n_of);
    SPP_CHECK(launch_jacobi(rho, rho_pstride, rho_ld, res.w, res.wstride, res.v, res.vstride, res.ldv, n_of,
                            nullptr, nprob, st));
    ctx.last_iterations = 0;
    return;
  }
  const int kk = ctx.ksub;
  const long long bs = (long long)nmax * kk;  // per-problem block size
  double* Qa = (double*)ctx.arena.take(sizeof(double) * nprob * bs);
  double* Za = (double*)ctx.arena.take(sizeof(double) * nprob * bs);
  double* Qb = (double*)ctx.arena.take(sizeof(double) * nprob * bs);
  double* Zb = (double*)ctx.arena.take(sizeof(double) * nprob * bs);
  const long long ss = (long long)kk * kk;
  double* Sm = (double*)ctx.arena.take(sizeof(double) * nprob * ss);
  double* Vm = (double*)ctx.arena.take(sizeof(double) * nprob * ss);
  double* sig = (double*)ctx.arena.take(sizeof(double) * nprob * kk);
  int* kk_of = (int*)ctx.arena.take(sizeof(int) * nprob);
  int* conv = (int*)ctx.arena.take(sizeof(int) * nprob);
  double* worst = (double*)ctx.arena.take(sizeof(double) * nprob);
  SPP_CHECK(launch_fill<int>(kk_of, nprob, kk, st));
  Leg fl = flat_leg(kk);
  // operand references (compact per (sample, parity) blocks; sample stride = 2 blocks)
  auto blockref = [&](double* base, const std::string& letters) {
    return compact_ref(base, 2 * bs, bs, kk, letters, 2, {lk, lp, fl});
  };
  auto smallref = [&](double* base, const std::string& letters) {
    return compact_ref(base, 2 * ss, ss, kk, letters, 1, {fl, fl});
  };
  TRef rhoR = compact_ref(const_cast<double*>(rho), 2 * rho_pstride, rho_pstride, rho_ld, "kpKP", 2,
                          {lk, lp, lk, lp});
  // problems still iterating (sample * 2 + parity); converged ones keep their Ritz pairs
  int* active = (int*)ctx.arena.take(sizeof(int) * nprob);
  SPP_CHECK(launch_fill<int>(active, nprob, 1, st));
  auto matmul_rho = [&](double* Z, double* Q) {  // Z = rho Q
    contract<double>(ctx, blockref(Z, "kpc"), rhoR, blockref(Q, "KPc"), nb, 1.0, false, active);
  };
  auto rmul = [&](double* out, double* Z, double* V) {  // out = Z V
    contract<double>(ctx, blockref(out, "kpb"), blockref(Z, "kpa"), smallref(V, "ab"), nb, 1.0, false, active);
  };
  double* tau = (double*)ctx.arena.take(sizeof(double) * nprob * kk);
  auto orth = [&](double* Q, double* Z, double* tmp) {  // Q <- orth(Z), Householder; Z destroyed
    (void)tmp;
    if (kk <= 64 && nmax <= 8 * 72) {
      hipLaunchKernelGGL((householder_q_reg_kernel<8, 72>), dim3(nprob), dim3(512), 0, st, Z, Q, bs, kk, kk, n_of,
                         active);
    } else if (kk <= 64 && nmax <= 16 * 41) {
      hipLaunchKernelGGL((householder_q_reg_kernel<16, 41>), dim3(nprob), dim3(1024), 0, st, Z, Q, bs, kk, kk, n_of,
                         active);
    } else {
      hipLaunchKernelGGL(householder_q_kernel, dim3(nprob), dim3(256), 0, st, Z, Q, bs, kk, kk, n_of, tau, active);
    }
    SPP_CHECK(hipGetLastError());
  };
  // Q <- orth(Z): a column-scaled, shifted Cholesky QR pass followed by a
  // first-order Lowdin pass (Gram and apply on the matrix cores), Householder
  // for the problems they mark bad; Z is kept, tmp is overwritten. Used once
  // the block is nearly orthogonal up to column scaling (Z = rho Q with Q the
  // previous Ritz vectors).
  int* okm = (int*)ctx.arena.take(sizeof(int) * nprob);
  int* badm = (int*)ctx.arena.take(sizeof(int) * nprob);
  double* Rm = (double*)ctx.arena.take(sizeof(double) * nprob * ss);
  const bool cholqr = kk <= CQ;
  std::vector<int> hbad(nprob);
  auto orth_chol = [&](double* Q, double* Z, double* tmp) {
    SPP_CHECK(launch_fill<int>(badm, nprob, 0, st));
    const double shift = 11.0 * ((double)nmax * kk + (double)kk * (kk + 1)) * 1.1102230246251565e-16;
    // pass 1: column-scaled, shifted Cholesky QR of Z into tmp
    contract<double>(ctx, smallref(Sm, "ab"), compact_ref(Z, 2 * bs, bs, kk, "kpa", 2, {lk, lp, fl}),
                     blockref(Z, "kpb"), nb, 1.0, false, active);
    hipLaunchKernelGGL(cholqr_kernel, dim3(nprob), dim3(CQ), 0, st, Sm, ss, kk, Rm, ss, kk, kk, 1, shift, active,
                       okm, badm);
    SPP_CHECK(hipGetLastError());
    contract<double>(ctx, blockref(tmp, "kpb"), blockref(Z, "kpa"), smallref(Rm, "ab"), nb, 1.0, false, okm);
    // pass 2: first-order Lowdin correction of the nearly orthonormal tmp into Q
    contract<double>(ctx, smallref(Sm, "ab"), compact_ref(tmp, 2 * bs, bs, kk, "kpa", 2, {lk, lp, fl}),
                     blockref(tmp, "kpb"), nb, 1.0, false, okm);
    hipLaunchKernelGGL(lowdin_kernel, dim3(nprob), dim3(256), 0, st, Sm, ss, kk, Rm, ss, kk, kk, active, okm, badm);
    SPP_CHECK(hipGetLastError());
    contract<double>(ctx, blockref(Q, "kpb"), blockref(tmp, "kpa"), smallref(Rm, "ab"), nb, 1.0, false, okm);
    // Householder for the problems marked bad (rare: dead columns, failed pivots)
    SPP_CHECK(hipMemcpyAsync(hbad.data(), badm, sizeof(int) * nprob, hipMemcpyDeviceToHost, st));
    SPP_CHECK(hipStreamSynchronize(st));
    bool any = false;
    for (int p = 0; p < nprob; ++p) any = any || hbad[p];
    if (!any) return;
    hipLaunchKernelGGL(and_mask_kernel, dim3((nprob + 255) / 256), dim3(256), 0, st, active, badm, okm, nprob);
    SPP_CHECK(hipGetLastError());
    if (kk <= 64 && nmax <= 8 * 72) {
      hipLaunchKernelGGL((householder_q_reg_kernel<8, 72>), dim3(nprob), dim3(512), 0, st, Z, Q, bs, kk, kk, n_of,
                         okm);
    } else if (kk <= 64 && nmax <= 16 * 41) {
      hipLaunchKernelGGL((householder_q_reg_kernel<16, 41>), dim3(nprob), dim3(1024), 0, st, Z, Q, bs, kk, kk, n_of,
                         okm);
    } else {
      hipLaunchKernelGGL(householder_q_kernel, dim3(nprob), dim3(256), 0, st, Z, Q, bs, kk, kk, n_of, tau, okm);
    }
    SPP_CHECK(hipGetLastError());
  };
  // start block
  hipLaunchKernelGGL(random_block_kernel, dim3(64, nprob), dim3(256), 0, st, Qb, bs, kk, kk, n_of, nmax);
  SPP_CHECK(hipGetLastError());
  matmul_rho(Za, Qb);
  orth(Qa, Za, Zb);
  const int t = std::min(want, kk);
  std::vector<int> hconv(nprob), hdeg(nprob);
  int* degm = (int*)ctx.arena.take(sizeof(int) * nprob);
  int* stepm = (int*)ctx.arena.take(sizeof(int) * nprob);
  int it = 0;
  for (it = 0; it < ctx.max_subspace_iter; ++it) {
    matmul_rho(Za, Qa);
    // H = Q^T Z
    contract<double>(ctx, smallref(Sm, "ab"), compact_ref(Qa, 2 * bs, bs, kk, "kpa", 2, {lk, lp, fl}),
                     blockref(Za, "kpb"), nb, 1.0, false, active);
    SPP_CHECK(launch_jacobi(Sm, ss, kk, sig, kk, Vm, ss, kk, kk_of, active, nprob, st));
    rmul(Qb, Qa, Vm);
    rmul(Zb, Za, Vm);
    hipLaunchKernelGGL(residual_kernel, dim3(nprob), dim3(256), 0, st, Qb, Zb, bs, kk, sig, (long long)kk, n_of,
                       t, tol, conv, worst, active, kk, cholqr ? ctx.cheb_degree : 1, degm);
    SPP_CHECK(hipGetLastError());
    SPP_CHECK(hipMemcpyAsync(hconv.data(), conv, sizeof(int) * nprob, hipMemcpyDeviceToHost, st));
    SPP_CHECK(hipMemcpyAsync(hdeg.data(), degm, sizeof(int) * nprob, hipMemcpyDeviceToHost, st));
    SPP_CHECK(hipStreamSynchronize(st));
    bool all = true;
    int mdeg = 1;
    for (int p = 0; p < nprob; ++p) {
      all = all && hconv[p];
      if (!hconv[p]) mdeg = std::max(mdeg, hdeg[p]);
    }
    if (all) break;
    if (!cholqr) {
      orth(Qa, Zb, Za);
      continue;
    }
    if (mdeg <= 1) {
      orth_chol(Qa, Zb, Za);
      continue;
    }
    // Chebyshev-filtered step: Y_0 = Ritz vectors (Qb), rho Y_0 = Zb
    const long long cnt = bs;
    const dim3 cg((unsigned)std::min<long long>((cnt + 255) / 256, 64), (unsigned)nprob);
    hipLaunchKernelGGL(cheb_step_kernel, cg, dim3(256), 0, st, Zb, Qb, (const double*)nullptr, Za, bs, cnt, sig,
                       (long long)kk, kk, degm, active, 0);
    SPP_CHECK(hipGetLastError());
    double *yprev = Qb, *ycur = Za, *zfree = Zb;
    for (int j = 1; j < mdeg; ++j) {
      hipLaunchKernelGGL(degree_mask_kernel, dim3((nprob + 255) / 256), dim3(256), 0, st, active, degm, j, stepm,
                         nprob);
      SPP_CHECK(hipGetLastError());
      contract<double>(ctx, blockref(zfree, "kpc"), rhoR, blockref(ycur, "KPc"), nb, 1.0, false, stepm);
      ctx.counters.eig_filter_gemms += 1;
      hipLaunchKernelGGL(cheb_step_kernel, cg, dim3(256), 0, st, zfree, ycur, yprev, yprev, bs, cnt, sig,
                         (long long)kk, kk, degm, active, j);
      SPP_CHECK(hipGetLastError());
      std::swap(yprev, ycur);
    }
    orth_chol(Qa, ycur, yprev);
  }
  ctx.last_iterations = it + 1;
  ctx.counters.eig_calls += 1;
  ctx.counters.eig_iterations += it + 1;
  if (it == ctx.max_subspace_iter)  // the cap ended the iteration: count the blocks short of the tolerance
    for (int p = 0; p < nprob; ++p) ctx.counters.eig_unconverged += hconv[p] ? 0 : 1;
  res.w = sig;
  res.wstride = kk;
  res.v = Qb;
  res.vstride = bs;
  res.ldv = kk;
  res.avail = kk_of;
}

// ------------------------------- absorb -------------------------------------
template <typename T>
void absorb_forward(Context& ctx, const Boundary<T>& env, const RowLayers<T>& row, int chi, Boundary<T>& out) {
  const int L = env.L;
  const int nb = env.nb;
  hipStream_t st = ctx.stream;
  if (row.L != L) throw HipError("absorb: row length mismatch");
  const size_t mark = ctx.arena.mark();
  // ---------------- right Gram matrices G[x], x = L..1 -----------------------
  // G[x] legs (l, w, L, W): bond x of env and the W leg of site x.
  std::vector<T*> G(L + 1, nullptr);
  std::vector<std::vector<Leg>> Gl(L + 1);
  G[L] = (T*)ctx.arena.take(sizeof(T) * nb);
  SPP_CHECK(launch_fill<T>(G[L], nb, T(1), st));
  Gl[L] = {env.bond(L), row.spec[L - 1].e, env.bond(L), row.spec[L - 1].e};
  // The three transfer intermediates carry the bra bond R of the environment
  // as a free leg until the last product contracts it, so they are formed for
  // index ranges of R at a time (G[x] accumulates over the ranges): each
  // range's intermediates stay below ctx.gram_chunk_bytes per sample.
  long long zmax = 0;
  std::vector<int> rchunk(L + 1, 1);
  for (int x = 1; x < L; ++x) {
    const LayerSpec& sp = row.spec[x];
    const long long unit = gram_unit_elems(env.bdim[x], sp.w.dim, sp.n.dim, sp.e.dim, sp.s.dim);
    rchunk[x] = gram_rchunk(unit, env.bdim[x + 1], sizeof(T), ctx.gram_chunk_bytes);
    zmax = std::max(zmax, unit * rchunk[x]);
  }
  T* Z1 = zmax ? (T*)ctx.arena.take(sizeof(T) * nb * zmax) : nullptr;
  T* Z2 = zmax ? (T*)ctx.arena.take(sizeof(T) * nb * zmax) : nullptr;
  int* rsplit = (int*)ctx.arena.take(sizeof(int) * nb);
  bool use_kron = sizeof(T) == sizeof(float) && ctx.kron && (int)row.kron_site.size() == L &&
                  (int)row.kron_tabs.size() == L;
  // role assignments of the two transfers in the row's frame (kron.h)
  const int kst0 = row.rotated ? 3 : 0, kst1 = row.rotated ? 2 : 1;
  for (int x = 1; x < L && use_kron; ++x) use_kron = row.kron_site[x] && row.kron_tabs[x];
  for (int x = L - 1; x >= 1; --x) {
    const LayerSpec& sp = row.spec[x];
    Leg bl = env.bond(x), br = env.bond(x + 1);
    long long gs = (long long)bl.dim * sp.w.dim * bl.dim * sp.w.dim;
    G[x] = (T*)ctx.arena.take(sizeof(T) * nb * gs);
    Gl[x] = {bl, sp.w, bl, sp.w};
    TRef A = env.ref(x, "lnr");
    TRef Gn = dense_ref(G[x + 1], (long long)Gl[x + 1][0].dim * Gl[x + 1][1].dim * Gl[x + 1][2].dim * Gl[x + 1][3].dim,
                        "reRF", Gl[x + 1]);
    TRef g = dense_ref(G[x], gs, "lwLW", Gl[x]);
    const KronSite* ks = use_kron ? row.kron_site[x] : nullptr;
    for (int r0 = 0; r0 < br.dim; r0 += rchunk[x]) {
      const int c = std::min(rchunk[x], br.dim - r0);
      const bool whole = c == br.dim;
      const Leg rc = whole ? br : chunk_leg(ctx, br, r0, c, rsplit, nb);
      if (ks) {
        // rows (l, R, F) of z1 through the ket and bra factors of the layer
        // into z2 [l][R][F][w][p'], then rows (l, R, w) of z2 into z3
        TRef z1 = folded_ref(Z1, zmax, "lRFne", {bl, rc, sp.e, sp.n, sp.e});
        contract<T>(ctx, z1, A, whole ? Gn : sub_leg(Gn, 'R', r0, rc), nb);
        const long long fe = std::max(sp.e.even, sp.e.dim - sp.e.even);
        const long long fs = std::max(sp.s.even, sp.s.dim - sp.s.even);
        KronRows ka;
        ka.nrows = (long long)bl.dim * rc.dim * sp.e.dim;
        ka.nf = sp.e.dim;
        ka.ef = sp.e.even;
        ka.rc = rc.dim;
        ka.lsplit = bl.split;
        ka.leven = bl.even;
        ka.rsplit = rc.split;
        ka.reven = rc.even;
        ka.xs = ka.ys = zmax;
        ka.sx_slab = (long long)sp.e.dim * sp.n.dim * fe;
        ka.sxf = (long long)sp.n.dim * fe;
        ka.sx1 = fe;
        ka.sx2 = 1;
        const double fa =
            2.0 * nb * (ctx.kron == 2 ? kron_mfma_row_macs(*ks, kst0) : kron_row_macs(*ks, kst0)) * ka.nrows;
        timed_kron(ctx, fa, [&] {
          return launch_kron_rows(*ks, kst0, reinterpret_cast<const float*>(Z1), reinterpret_cast<float*>(Z2), ka,
                                  row.kron_tabs[x], nb, ctx.kron_rows, st, ctx.kron_groups, nullptr,
                                  ctx.kron == 2);
        });
        KronRows kb = ka;
        kb.nrows = (long long)bl.dim * rc.dim * sp.w.dim;
        kb.nf = sp.w.dim;
        kb.ef = sp.w.even;
        kb.sx_slab = (long long)sp.e.dim * sp.w.dim * fs;
        kb.sxf = fs;
        kb.sx1 = (long long)sp.w.dim * fs;
        kb.sx2 = 1;
        const double fb =
            2.0 * nb * (ctx.kron == 2 ? kron_mfma_row_macs(*ks, kst1) : kron_row_macs(*ks, kst1)) * kb.nrows;
        timed_kron(ctx, fb, [&] {
          return launch_kron_rows(*ks, kst1, reinterpret_cast<const float*>(Z2), reinterpret_cast<float*>(Z1), kb,
                                  row.kron_tabs[x], nb, ctx.kron_rows, st, ctx.kron_groups, nullptr,
                                  ctx.kron == 2);
        });
        ctx.counters.kron_launches += 2;
        ctx.counters.gemm_flops += fa + fb;
        TRef z3 = folded_ref(Z1, zmax, "lRwWN", {bl, rc, sp.w, sp.w, sp.n});
        TRef Ar = env.ref(x, "LNR");
        contract<T>(ctx, g, z3, whole ? Ar : sub_leg(Ar, 'R', r0, rc), nb, 1.0, r0 > 0);
        continue;
      }
      TRef z1 = folded_ref(Z1, zmax, "lneRF", {bl, sp.n, sp.e, rc, sp.e});
      contract<T>(ctx, z1, A, whole ? Gn : sub_leg(Gn, 'R', r0, rc), nb);
      TRef z2 = folded_ref(Z2, zmax, "lRFwp", {bl, rc, sp.e, sp.w, sp.s});
      contract<T>(ctx, z2, z1, row.ref(x, "wnep"), nb);
      TRef z3 = folded_ref(Z1, zmax, "lRwWN", {bl, rc, sp.w, sp.w, sp.n});
      contract<T>(ctx, z3, z2, row.ref(x, "WNFp"), nb);
      TRef Ar = env.ref(x, "LNR");
      contract<T>(ctx, g, z3, whole ? Ar : sub_leg(Ar, 'R', r0, rc), nb, 1.0, r0 > 0);
    }
  }
  // ---------------- left-to-right truncation sweep --------------------------
  out.L = L;
  out.nb = nb;
  out.bdim.assign(L + 1, 1);
  out.beven.assign(L + 1, 1);
  if (out.bsplit.size() != (size_t)(L + 1)) throw HipError("absorb: output boundary not prepared");
  out.pdim.resize(L);
  out.peven.resize(L);
  for (int x = 0; x < L; ++x) {
    out.pdim[x] = row.spec[x].s.dim;
    out.peven[x] = row.spec[x].s.even;
  }
  SPP_CHECK(hipMemcpyAsync(out.log_scale, env.log_scale, sizeof(double) * nb, hipMemcpyDeviceToDevice, st));
  SPP_CHECK(hipMemcpyAsync(out.max_trunc, env.max_trunc, sizeof(double) * nb, hipMemcpyDeviceToDevice, st));
  // host estimate of LQ ranks kR[x] (columns of L_x), per parity, even-split legs assumed
  std::vector<std::array<long long, 2>> kR(L + 1);
  kR[L] = {1, 0};
  for (int x = L - 1; x >= 1; --x) {
    const LayerSpec& sp = row.spec[x];
    auto cnt2 = [](const Leg& a, long long ae, long long ao, int q) -> long long {
      long long e = a.split ? a.dim : a.even, o = a.split ? a.dim : a.dim - a.even;  // per-sample: bound
      return q == 0 ? e * ae + o * ao : e * ao + o * ae;
    };
    for (int q = 0; q < 2; ++q) {
      long long rows = cnt2(env.bond(x), sp.w.even, sp.w.dim - sp.w.even, q);
      long long cols = cnt2(sp.s, kR[x + 1][0], kR[x + 1][1], q);
      kR[x][q] = std::min(rows, cols);
    }
  }
  T* Rprev = (T*)ctx.arena.take(sizeof(T) * nb);
  SPP_CHECK(launch_fill<T>(Rprev, nb, T(1), st));
  std::vector<Leg> Rl = {fixed_leg(1, 1), env.bond(0), row.spec[0].w};
  Leg kb = fixed_leg(1, 1);
  double* nrm = (double*)ctx.arena.take(sizeof(double) * nb);
  double* fac = (double*)ctx.arena.take(sizeof(double) * nb);
  double* tr = (double*)ctx.arena.take(sizeof(double) * 2 * nb);
  int* ncount = (int*)ctx.arena.take(sizeof(int) * 2 * nb);
  for (int x = 0; x < L; ++x) {
    const LayerSpec& sp = row.spec[x];
    Leg br = env.bond(x + 1);
    // R for the next site is allocated below this site's scratch mark
    const long long rs_cap = (long long)chi * br.dim * sp.e.dim;
    T* Rn = (x < L - 1) ? (T*)ctx.arena.take(sizeof(T) * nb * rs_cap) : nullptr;
    const size_t xm = ctx.arena.mark();
    long long qs = (long long)kb.dim * sp.w.dim * sp.n.dim * br.dim;
    T* Q = (T*)ctx.arena.take(sizeof(T) * nb * qs);
    TRef Rp = dense_ref(Rprev, (long long)Rl[0].dim * Rl[1].dim * Rl[2].dim, "klw", Rl);
    TRef Qr = dense_ref(Q, qs, "kwnr", {kb, sp.w, sp.n, br});
    contract<T>(ctx, Qr, Rp, env.ref(x, "lnr"), nb);
    long long ys = (long long)kb.dim * sp.s.dim * br.dim * sp.e.dim;
    T* Y = (T*)ctx.arena.take(sizeof(T) * nb * ys);
    TRef Yr = dense_ref(Y, ys, "kpre", {kb, sp.s, br, sp.e});
    contract<T>(ctx, Yr, Qr, row.ref(x, "wnep"), nb);
    if (x == L - 1) {
      // last tensor: normalise and store (k, p, 1)
      if (br.dim != 1 || sp.e.dim != 1) throw HipError("absorb: last site must close with unit legs");
      FullView fv = full_view(Yr);
      SPP_CHECK(launch_dot<T>(fv, fv, nrm, nb, st, dot_partials(ctx, nb)));
      hipLaunchKernelGGL(norm_finish_kernel, dim3((nb + 255) / 256), dim3(256), 0, st, nrm, out.log_scale, fac,
                         nullptr, nb);
      SPP_CHECK(hipGetLastError());
      SPP_CHECK(launch_scale<T>(fv, fac, nb, st));
      out.bdim[x] = kb.dim;
      out.beven[x] = kb.even;
      SPP_CHECK(hipMemcpyAsync(out.t[x], Y, sizeof(T) * nb * ys, hipMemcpyDeviceToDevice, st));
      out.bdim[L] = 1;
      out.beven[L] = 1;
      ctx.arena.release(xm);
      break;
    }
    // W = Y G[x+1]; rho = W Y^T (compact per parity block)
    const std::vector<Leg>& gl = Gl[x + 1];
    TRef Gr = dense_ref(G[x + 1], (long long)gl[0].dim * gl[1].dim * gl[2].dim * gl[3].dim, "reRF", gl);
    T* W = (T*)ctx.arena.take(sizeof(T) * nb * ys);
    TRef Wr = dense_ref(W, ys, "kpRF", {kb, sp.s, br, sp.e});
    contract<T>(ctx, Wr, Yr, Gr, nb);
    // rows (k,p) tuple counts per problem and their bound
    long long nmax = group_bound({kb, sp.s});
    Group gkp{};
    fill_group(gkp, {kb, sp.s}, "kp");
    hipLaunchKernelGGL(group_counts_kernel, dim3((2 * nb + 255) / 256), dim3(256), 0, st, gkp, ncount, 2 * nb);
    SPP_CHECK(hipGetLastError());
    const long long rps = nmax * nmax;
    double* rho = (double*)ctx.arena.take(sizeof(double) * 2 * nb * rps);
    if (sizeof(T) == sizeof(double)) {
      contract<T>(ctx, compact_ref(rho, 2 * rps, rps, nmax, "kpKP", 2, {kb, sp.s, kb, sp.s}), Wr,
                  with_letters(Yr, "KPRF"), nb);
    } else {
      T* rho_t = (T*)ctx.arena.take(sizeof(T) * 2 * nb * rps);
      contract<T>(ctx, compact_ref(rho_t, 2 * rps, rps, nmax, "kpKP", 2, {kb, sp.s, kb, sp.s}), Wr,
                  with_letters(Yr, "KPRF"), nb);
      const int tiles = (int)((nmax + 31) / 32);
      hipLaunchKernelGGL(symmetrize_convert_kernel, dim3(tiles * (tiles + 1) / 2, 2 * nb), dim3(256), 0, st,
                         reinterpret_cast<const float*>(rho_t), rho, rps, (int)nmax, ncount);
      SPP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(trace_kernel, dim3(2 * nb), dim3(256), 0, st, rho, rps, (int)nmax, ncount, tr);
    SPP_CHECK(hipGetLastError());
    // bond capacity of this cut
    long long nq[2];
    for (int q = 0; q < 2; ++q) {
      long long e = kb.split ? kb.dim : kb.even, o = kb.split ? kb.dim : kb.dim - kb.even;
      long long pe = sp.s.even, po = sp.s.dim - sp.s.even;
      nq[q] = std::min(nmax, q == 0 ? e * pe + o * po : e * po + o * pe);
    }
    int cap = (int)std::min<long long>(chi, std::min(nq[0], kR[x + 1][0]) + std::min(nq[1], kR[x + 1][1]));
    if (cap < 1) cap = 1;
    // bond capacity under the sector budget (the ranking still runs over the
    // top cap; a parity keeps at most its row bound nq and its budget)
    long long capq[2];
    for (int q = 0; q < 2; ++q) {
      capq[q] = nq[q];
      if (ctx.bcap[q] >= 0) capq[q] = std::min<long long>(capq[q], ctx.bcap[q]);
    }
    int bcap = (int)std::min<long long>(cap, capq[0] + capq[1]);
    if (bcap < 1) bcap = 1;
    EigResult er{};
    // rho formed by FP32 products carries rounding of about 1e-6 of its largest
    // eigenvalue: Ritz residuals below that are not resolved, so the iteration
    // stops there instead of running to its cap
    const double eig_tol = sizeof(T) == sizeof(float) ? std::max(ctx.subspace_tol, kFp32RhoTol) : ctx.subspace_tol;
    eig_topk(ctx, rho, rps, (int)nmax, kb, sp.s, ncount, (int)nmax, nb, cap, er, eig_tol);
    // U[k,p,c] into the output boundary storage
    long long us = (long long)kb.dim * sp.s.dim * bcap;
    T* U = out.t[x];
    SPP_CHECK(launch_fill<T>(U, nb * us, T(0), st));
    hipLaunchKernelGGL(truncate_select_kernel<T>, dim3(nb), dim3(256), 0, st, gkp, er.w, er.wstride, er.v,
                       er.vstride, er.ldv, er.avail, ncount, tr, chi, cap, ctx.bcap[0], ctx.bcap[1], bcap, U, us,
                       (long long)sp.s.dim * bcap, (long long)bcap, out.bsplit[x + 1], out.max_trunc, nullptr);
    SPP_CHECK(hipGetLastError());
    if (out.entropy || out.entropy_max) {
      // pre-truncation spectrum of this cut (eigenvalues of rho)
      if (nmax <= ctx.ksub)
        entropy_cut(ctx, er.w, er.wstride, ncount, nb, x, out.entropy, out.entropy_max);
      else
        entropy_defer(ctx, rho, rps, (int)nmax, ncount, nb, x, out.entropy, out.entropy_max);
    }
    out.bdim[x] = kb.dim;
    out.beven[x] = kb.even;
    Leg nbond;
    nbond.dim = bcap;
    nbond.split = out.bsplit[x + 1];
    nbond.even = bcap;
    // R = U^T Y -> (c, r, e)
    long long rs = (long long)bcap * br.dim * sp.e.dim;
    contract<T>(ctx, dense_ref(Rn, rs, "cre", {nbond, br, sp.e}), dense_ref(U, us, "kpc", {kb, sp.s, nbond}),
                Yr, nb);
    ctx.arena.release(xm);
    Rprev = Rn;
    Rl = {nbond, br, sp.e};
    kb = nbond;
  }
  ctx.arena.release(mark);
}

template void absorb_forward<float>(Context&, const Boundary<float>&, const RowLayers<float>&, int,
                                    Boundary<float>&);
template void absorb_forward<double>(Context&, const Boundary<double>&, const RowLayers<double>&, int,
                                     Boundary<double>&);

}  // namespace spp
