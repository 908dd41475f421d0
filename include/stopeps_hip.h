/*
 * stopeps_hip.h - C ABI of the MI355X Born-sampling path of StoPEPS.jl.
 *
 * The reference has no FFI: its plugin point is Julia dispatch on
 * `RunConfig.contraction::Type{<:ContractionMethod}` (src/driver/config.jl:51),
 * instantiated per chi role by `_boundary(c, role)` (src/driver/particles.jl:27-34)
 * and consumed by `SamplingCache` / `sample_configuration`
 * (src/sampling/born.jl:45-163), `lognorm` (src/contraction/boundary_mps.jl:180-186)
 * and `project_configuration` (born.jl:27-36). A Julia `ccall` layer (or the
 * Python ctypes layer shipped with this library) binds the functions below
 * in place of those methods; the per-sample loops of the callers
 * (src/driver/run.jl:45-66) become one batched call.
 *
 * Conventions
 *  - Sites are indexed row-major, i = x + y*Lx (x fastest), as
 *    `site_order` (src/states/finite.jl:16).
 *  - A site tensor is a dense row-major FP64 array over the legs
 *    (W, N, Pout, Paux, E, S) (src/states/finite.jl:22-26), each leg sorted by
 *    Z2 parity with its `even` first indices of parity 0. Dense
 *    (ComplexSpace) tensors pass even = dim on every leg.
 *  - Host arrays are caller-owned and read/written only during a call; calls
 *    return after the device work finished.
 *  - Return value: 0 on success, negative on an API/HIP error (message via
 *    spp_last_error). Numerical conditions are reported per sample in
 *    `status` (0 ok, 1 invalid Born conditional, 2 zero conditional mass,
 *    3 nonfinite environment), so the caller can raise the reference's
 *    messages (born.jl:137-142, boundary_mps.jl:131,258).
 *  - One context = one HIP stream; a context is not re-entrant. Drive all
 *    contexts of a process from one host thread: calls on separate contexts
 *    from concurrent host threads are not supported (a full-update batch split
 *    over concurrent contexts gave results that varied from run to run).
 *    Contexts used one after the other give the results of a single context.
 */
#ifndef STOPEPS_HIP_H
#define STOPEPS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct spp_ctx spp_ctx;
typedef struct spp_state spp_state;
typedef struct spp_cache spp_cache;

/* Lifecycle. No reference counterpart (Julia owns its process state). */
int spp_create(int device, spp_ctx** out);
int spp_destroy(spp_ctx* ctx);
const char* spp_last_error(const spp_ctx* ctx);
int spp_synchronize(spp_ctx* ctx);

/* Arithmetic of the contraction GEMMs: 0 = FP64 everywhere, 1 = FP32 MFMA
 * contractions with FP64 eigensolvers, log weights and probabilities.
 * 2 and 3 = as 1 (FP32 storage and accumulation), with each product formed on
 * the BF16 matrix cores from BF16 pieces of the FP32 operands: 2 = three
 * pieces and six piece products (FP32-equivalent), 3 = two pieces and three
 * products (about 2^-16 relative). Takes effect for caches built afterwards. */
int spp_set_precision(spp_ctx* ctx, int precision);

/* Tuning: "ksub" (subspace size of the truncation eigensolver, <= 64),
 * "max_subspace_iter", "subspace_tol" (relative Ritz residual at which a
 * truncation eigensolve stops, default 1e-12; where rho is formed by FP32
 * products the stop is at max(subspace_tol, 1e-6), the rounding level of that
 * rho), "cheb_degree" (largest Chebyshev
 * filter degree of a subspace step, 1 = plain power steps; default 8),
 * "chunk" (samples per device batch,
 * 0 = sized from free memory), "profile" (1: time each contraction launch
 * with HIP events on the library stream), "boundary_cap_even" /
 * "boundary_cap_odd" (the sector budget of `BoundaryMPS(chi, sector_space)`,
 * boundary_mps.jl:11-18: every boundary compression keeps truncrank(chi)
 * intersected with at most this many values of each bond parity; negative =
 * no budget, the default; applies to the calls made while it is set);
 * "fu_fail_sample", "fu_fail_bond", "fu_fail_code" (a test hook of the full
 * update's failure handling: bond update fu_fail_bond, in sweep order, of
 * sample fu_fail_sample of the next steps is treated as failed with
 * fu_fail_code; a negative sample turns it off); "gram_chunk_bytes" (per-sample
 * byte limit of one Gram transfer intermediate of a boundary absorb: the bra
 * bond is taken in index ranges that keep it below the limit; default 2^30;
 * results do not depend on it beyond summation order); "autotune" (1, the
 * default: the contraction kernel's tile shape is tuned per launch class — kernel
 * family, operand load modes, sample count and the dimension and fixed parity
 * split of every contracted and free leg — by timing every shape on the first
 * launch of the class; 0: the padded-area heuristic; -1: forget the tuned
 * shapes; every shape gives bit-identical results); "graphs" (1, the default: a
 * sampling cache keeps its draw buffers between calls, and each row's
 * environment-and-draw sweep is captured as a HIP graph the second time it runs
 * with the same shape and replayed afterwards; 0: launched kernel by kernel;
 * off while "profile" is on); "kron" (1, the default: the Gram transfers of
 * the draw's FP32 absorbs apply each drawn layer through its ket and bra
 * single-layer factors in one kernel on the vector ALUs, about 9x fewer
 * multiply-adds than the dense double layer at D=6; 2: the same products on
 * the matrix cores (16x16x4 FP32 tiles; slower than 1 at D=6); 0: the dense
 * double layer; results agree to FP32 rounding); "kron_rows" (rows per block
 * of that kernel, 0 = default); "kron_groups" (row groups per block of the
 * matrix-core kernel, 1..64, default 4);
 * "gemm_tile" (self-tests:
 * force one tile shape index of the kernel family on every launch, -1 = off). */
int spp_set_option(spp_ctx* ctx, const char* name, double value);

/* counters[0] = contraction flops issued (2*M*N*K over parity blocks),
 * counters[1] = contraction launches, counters[2] = subspace iterations of the
 * last truncation, counters[3] = peak workspace bytes, counters[4] = summed
 * contraction kernel milliseconds and counters[5] = the flops of the timed
 * launches (profiling mode); counters[6..8] = launches, milliseconds and flops
 * of the timed FP64-kernel launches, counters[9..11] the same for the FP32
 * kernels (the contraction kernel and the factorised Gram transfer kernel); counters[12..14] = iterative truncation eigensolves, their summed
 * subspace iterations and their Chebyshev filter products; counters[15] = full-
 * update bonds repaired by the PSD projection of their hole; counters[16..18] = launch
 * classes the tile autotuner tuned, those whose fastest shape differs from the
 * heuristic's, and the milliseconds spent tuning; counters[19..20] = draw row sweeps
 * captured as HIP graphs and graph replays; counters[21] = boundary truncation
 * eigenproblems (one per sample and bond parity) whose Ritz residuals were still
 * above "subspace_tol" when "max_subspace_iter" ended the iteration (their kept
 * vectors are those of the last iteration; the reference's boundary SVD has no
 * such cap, boundary_mps.jl:33-43); counters[22] = launches of the factorised
 * Gram transfer kernel (its multiply-adds, register tile included, are added
 * to counters[0]). Returns the count written. */
int spp_get_counters(spp_ctx* ctx, double* counters, int n);
int spp_reset_counters(spp_ctx* ctx);

/* Upload a FinitePEPO (kind 0) or FinitePEPS (kind 1)
 * (`FiniteState`, src/states/finite.jl:27-44). dims/evens: [Lx*Ly][6];
 * data: Lx*Ly pointers to dense site arrays. */
int spp_state_upload(spp_ctx* ctx, int Lx, int Ly, int kind, const int* dims, const int* evens,
                     const double* const* data, double beta, double log_amplitude, spp_state** out);
int spp_state_free(spp_state* state);

/* `project_configuration(checkpoint, configuration)` (src/sampling/born.jl:27-36)
 * of n configurations of an uploaded PEPO, on the device: state s gets site i
 * = sum_a A_i[w,n,p,a,e,s] U[a][k] with k = labels[s][i] (0-based basis index;
 * basis as in spp_sampling_cache_build, NULL = identity) and a one-index
 * auxiliary leg of the label's charge, then `normalize_state!`
 * (src/states/finite.jl:89-100): each site divided by its norm and the logs
 * (normalizations [n][Lx*Ly], may be NULL) added to the log amplitude
 * (log_amplitude [n], may be NULL). capacity[2] = (even, odd) bond space of
 * every interior virtual leg of the new states (NULL: the PEPO's); the PEPO's
 * live sectors are kept (the first live indices of each parity) and the rest
 * of the capacity is zero. out[n] receives new state handles (free each with
 * spp_state_free). Fails with "zero or nonfinite tensor at (x,y)" like the
 * reference. */
int spp_project_batch(spp_ctx* ctx, const spp_state* pepo, int n, const int* labels, const double* basis,
                      const int* capacity, spp_state** out, double* log_amplitude, double* normalizations);

/* Live sectors of every virtual bond (no reference counterpart: the device
 * keeps each leg at its uploaded capacity, and these counts are the bond space
 * the reference would hold after truncation). live: [Lx*Ly][4 legs W,N,E,S][2]
 * = (even, odd) counts in use; entries beyond them must be zero. Set to the
 * full capacity at upload. */
int spp_state_set_live(spp_ctx* ctx, spp_state* state, const int* live);

/* Copy a state back: data[i] receives site i (row-major, x fastest) as dense
 * (W,N,P,A,E,S) in the parity-sorted leg order of the upload; beta,
 * log_amplitude and live (as in spp_state_set_live) may be null. */
int spp_state_download(spp_ctx* ctx, const spp_state* state, double* const* data, double* beta,
                       double* log_amplitude, int* live);

/* `evolution_step!(state, model, delta, D, FullUpdate(...), BoundaryMPS(chi);
 * order, sector_space)` (src/evolution/full_update.jl:563-597) for n states on
 * the same lattice with the same leg spaces (up to the auxiliary charge),
 * updated in place. The bond capacity is the uploaded bond space and doubles
 * as the sector budget (`sector_space`); D is the truncrank of the bond
 * (D <= 0: the capacity's dimension). options = {max_iterations,
 * error_tolerance, diagonal_shift, linear_max_iterations, linear_tolerance}
 * (`FullUpdate`, full_update.jl:11-17). gates_h [Ly*(Lx-1)][d^4] and gates_v
 * [(Ly-1)*Lx][d^4] are `trotter_gate(model, lattice, i, j, delta_gate)`
 * (src/models/tfim.jl:48) as G[p][q][u][v] in the parity-sorted physical
 * basis, bond (x,y)-(x+1,y) at y*(Lx-1)+x and (x,y)-(x,y+1) at y*Lx+x, with
 * delta_gate = delta/2 for order 2. On a one-site lattice gates_h is instead
 * exp(-delta onsite_hamiltonian) as G[p][u] (full_update.jl:568-570) and
 * gates_v is unused. Outputs (may be null): records
 * [n][bond updates in sweep order][8] = {initial_cost, final_cost, residual,
 * iterations, iteration_limit_reached, stop (0 disabled, 1 iteration limit,
 * 2 error tolerance, 3 residual fallback), kept_even, kept_odd};
 * normalizations [n][Lx*Ly] (`normalize_state!` log norms, added to each
 * state's log_amplitude); max_env_trunc [n]; max_boundary_entropy [n] the
 * largest pre-truncation Schmidt entropy (nats) over every boundary cut of the
 * step (`record_entropy=true`, full_update.jl:560-591; NULL: not recorded);
 * status [n] the per-sample outcome: 0 ok, else the failure of the first bond
 * that failed (1 nonfinite norm metric, 2 zero norm metric, 3 zero gated
 * target, 4 negative residual needing the PSD repair, 5 nonfinite
 * residual/target, 6 singular gauge, 7 workspace, 8 eigen/SVD sweep cap
 * reached without convergence), 9 nonfinite row environment, 10 zero or
 * nonfinite site norm, 11 the PSD repair did not remove the residual
 * failure (see spp_fu_set_repair: with a repair callback set, a bond with
 * status 4 is repaired instead of failing).
 * Failure semantics: a sample that fails keeps its sites from the failing
 * bond on (no later bond of the step writes it), the others finish the step;
 * a failed state is marked and every compute entry point refuses it
 * afterwards ("state invalidated by a failed full-update step"). With status
 * NULL the call finishes the step and then returns an error carrying the
 * first failure's message. The reference raises at the failing bond
 * (full_update.jl:491-501, svd.jl:1-11). */
int spp_fu_evolution_step(spp_ctx* ctx, int n, spp_state* const* states, int chi, int order, double delta, int D,
                          const double* options, const double* gates_h, const double* gates_v, double* records,
                          double* normalizations, double* max_env_trunc, double* max_boundary_entropy,
                          int* status);

/* PSD repair of the full-update bond metric (`_positive_reduced_metric` and
 * the repair branch of `update_bond!`, src/evolution/full_update.jl:177-340,
 * 484-508). When the residual check of a bond requests it (status 4,
 * full_update.jl:355-369), spp_fu_evolution_step contracts that sample's
 * complete bond hole on the device and hands it to the host callback set here,
 * together with the Q factors of the reduced pair; the callback writes the
 * reduced metric of the hole's PSD projection, and the bond solve of that
 * sample runs again on it. If the repaired solve still requests the repair,
 * the sample fails with status 11 ("PSD repair did not remove the FU residual
 * failure"). Without a callback the sample fails with status 4.
 *
 * Layouts: a double-layer (pair) leg of a virtual leg with dimension m and e
 * even indices has m*m entries in the pair order (ket i, bra j) ->
 *   both even: i*e + j; both odd: e*e + (i-e)*o + (j-e);
 *   i even, j odd: e*e + o*o + i*o + (j-e); i odd, j even: e*e + o*o + e*o + (i-e)*e + j
 * (o = m - e). hole: H over its six exterior pair legs, row-major; leg k is the
 * pseudo-site slot hole_slot[k] of Q factor hole_q[k] (0: qa, 1: qb) and has
 * virtual dimension hole_dims[k] with hole_evens[k] even indices. qa, qb: the Q
 * factors as dense row-major pseudo sites (W,N,P,A,E,S) of dims *_dims, with
 * the auxiliary leg in the P slot, a unit A slot and the bond r in slot
 * *_core. The metric to write is N[pair(rk, rb)][pair(sk, sb)] over the r legs
 * of qa and qb (dims qa_dims[qa_core] x ..., row-major), the contraction
 * sum H+[(ket legs), (bra legs)] qa[ket.., aux, rk] qa[bra.., aux, rb]
 * qb[ket.., aux', sk] qb[bra.., aux', sb] over the exterior legs and each
 * factor's auxiliary leg, with H+ the PSD projection of H/||H|| as the
 * reference computes it. The callback returns 0 on success. */
typedef struct spp_fu_repair_problem {
  int sample;      /* position of the state in the call's list */
  int bond;        /* bond update in sweep order */
  int horizontal;  /* 1: bond (x,y)-(x+1,y), 0: (x,y)-(x,y+1) */
  int hole_dims[6], hole_evens[6], hole_q[6], hole_slot[6];
  const double* hole;
  int qa_dims[6], qa_evens[6], qb_dims[6], qb_evens[6];
  int qa_core, qb_core;
  const double* qa;
  const double* qb;
  double target_norm; /* <target, N target> of the unrepaired metric */
} spp_fu_repair_problem;
typedef int (*spp_fu_repair_fn)(void* user, const spp_fu_repair_problem* problem, double* metric);
/* Set (fn non-null) or clear the repair callback of a context. */
int spp_fu_set_repair(spp_ctx* ctx, spp_fu_repair_fn fn, void* user);

/* `SamplingCache(state, BoundaryMPS(chi); kind=:auxiliary, basis)`
 * (src/sampling/born.jl:56-73): double layers, aux-projected candidate layers
 * and all south boundary environments. `basis` is the 2x2 row-major matrix
 * U[a][k] whose columns are the local basis vectors (NULL: identity, i.e.
 * `SymmetryBasis()` on Z2Space(0=>1,1=>1); XBasis for dense TFIM). */
int spp_sampling_cache_build(spp_ctx* ctx, const spp_state* pepo, int chi, const double* basis,
                             spp_cache** out);
/* `SamplingCache(state, BoundaryMPS(chi); kind, basis)` (src/sampling/born.jl:56-73)
 * of either kind: kind 0 = :auxiliary (state a PEPO; as spp_sampling_cache_build),
 * kind 1 = :physical (state a PEPS with d = 2 and a one-index auxiliary leg; the
 * resampling proposals of src/resample/event.jl:27-33): unprojected double
 * layers, physically projected candidates, south boundaries, and for a Z2 state
 * (which requires the identity basis, `SymmetryBasis()`) the state's total
 * charge `_state_charge` and reachable charges `_charge_suffix`
 * (src/sampling/basis.jl:89-100). Draws from a physical cache weight label k by
 * ||root v_k||^2 with root the square root of the PSD projection of the site
 * metric `_site_hole` (`_positive_metric(...; square_root=true)`,
 * src/states/tensors.jl:65-94; born.jl:116-134). */
int spp_sampling_cache_build_kind(spp_ctx* ctx, const spp_state* state, int chi, const double* basis, int kind,
                                  spp_cache** out);
int spp_sampling_cache_free(spp_cache* cache);
/* south_max_trunc[Ly+1]: `max_truncation` of every south boundary. */
int spp_sampling_cache_info(const spp_cache* cache, double* south_max_trunc);

/* Batched `sample_configuration(cache; rng)` (src/sampling/born.jl:86-163),
 * auxiliary kind, for n samples at once.
 *  uniforms   [n][Lx*Ly]  the one uniform the reference draws per site
 *                          (`rand(rng)` in _categorical, selection.jl:71),
 *                          or NULL in evaluation mode;
 *  labels_in  [n][Lx*Ly]  0-based basis indices (evaluation mode,
 *                          `configuration=` keyword) or NULL;
 *  labels_out [n][Lx*Ly]  drawn 0-based basis indices;
 *  log_prob   [n]         log probability of the configuration;
 *  conditionals [n][Lx*Ly] drawn conditional probability per site;
 *  max_trunc  [n]         max_environment_truncation;
 *  tensor_log_norm [n]    `lognorm(project_configuration(...)).tensor_log_norm`
 *                          at the sampling chi (may be NULL);
 *  log_amplitude   [n]    log_amplitude of the projected, normalised PEPS
 *                          (`normalize_state!`, finite.jl:89-100; may be NULL);
 *  status [n], fail_site [n] per-sample numerical condition and site index. */
int spp_born_sample_batch(spp_ctx* ctx, spp_cache* cache, int n, const double* uniforms, const int* labels_in,
                          int* labels_out, double* log_prob, double* conditionals, double* max_trunc,
                          double* tensor_log_norm, double* log_amplitude, int* status, int* fail_site);

/* spp_born_sample_batch for either kind of cache, with the per-site `repairs`
 * of the physical kind (born.jl:118-131): repairs [n][Lx*Ly][6] (may be NULL)
 * = {minimum_relative_eigenvalue, negative_eigenvalues, negative_multiplets,
 * negative_spectral_fraction, hermiticity_defect} of the site metric's
 * `_positive_metric` report (tensors.jl:78-85) and, last, the number of
 * allowed outcomes when the proposal weights were all zero and the draw fell
 * back to the uniform distribution over them (0 otherwise); sites not reached
 * (the sample ended earlier) hold NaN. A physical draw has no projected PEPS:
 * tensor_log_norm and log_amplitude are NaN, and the last row is not absorbed
 * (max_trunc covers the absorbs made, born.jl:157-160). In evaluation mode a
 * configuration outside a Z2 state's total charge gets log_prob = -Inf with
 * no conditionals (born.jl:100-104). Status codes of the physical kind add 4
 * nonfinite norm metric (tensors.jl:67), 5 no label fits the total charge
 * (born.jl:116) and 6 the metric's eigensolver reached its iteration cap. */
int spp_born_sample_batch_repairs(spp_ctx* ctx, spp_cache* cache, int n, const double* uniforms,
                                  const int* labels_in, int* labels_out, double* log_prob, double* conditionals,
                                  double* max_trunc, double* tensor_log_norm, double* log_amplitude, int* status,
                                  int* fail_site, double* repairs);

/* Batched `measure_energy(state, model, BoundaryMPS(chi))`
 * (src/measurement/energy.jl:31-78) of the projected samples
 * `project_configuration(pepo, labels)` (src/sampling/born.jl:27-36), as
 * `_observe` runs it for each new particle (src/driver/particles.jl:35-52).
 * `pepo` is the state `cache` was built from; `chi` is the measurement bond.
 *  labels [n][Lx*Ly]   0-based basis indices of the samples;
 *  X[2][2]             onsite operator (row-major), weighted by -h;
 *  F[2][2], G[2][2]    rank-one split of the two-site operator,
 *                      pair[p,q,u,v] = F[p][u] G[q][v] (`_operator_factors`,
 *                      energy.jl:8-12), weighted by -J; factor_odd = 1 when
 *                      the factor leg carries Z2 charge 1 (Z2 TFIM);
 *  energy_total, energy_per_site, log_norm, tensor_log_norm, max_trunc [n]
 *                      as in the reference's NamedTuple;
 *  status [n]          0 ok, 3 nonfinite row environment
 *                      (boundary_mps.jl:258), 4 zero local norm environment
 *                      and 5 nonreal/nonfinite observable (energy.jl:1-6),
 *                      6 norm not finite and positive (boundary_mps.jl:188);
 *  fail_site [n]       row-major site of the failing term, or -1;
 *  boundary_entropy [n][2][Ly-1][Lx-1]  `Se` of `record_entropy=true`
 *                      (boundary_mps.jl:20-57, 193-209): Se.north then
 *                      Se.south [row_cut][bond], pre-truncation Schmidt
 *                      entropy in nats, NaN where undefined (NULL: not
 *                      recorded);
 *  max_boundary_entropy [n]  its maximum over every cut of the north and
 *                      south boundaries (NULL: not recorded). */
int spp_measure_energy_batch(spp_ctx* ctx, const spp_cache* cache, const spp_state* pepo, int chi, int n,
                             const int* labels, const double* X, const double* F, const double* G, int factor_odd,
                             double J, double h, double* energy_total, double* energy_per_site, double* log_norm,
                             double* tensor_log_norm, double* max_trunc, int* status, int* fail_site,
                             double* boundary_entropy, double* max_boundary_entropy);

/* Batched `measure_energy(state, model, BoundaryMPS(chi))`
 * (src/measurement/energy.jl:31-78) of n uploaded states (PEPS or PEPO, all on
 * the same lattice with the same leg spaces; the double layers trace the
 * auxiliary leg), e.g. the particles of `_observe` after `_evolve!`
 * (src/driver/particles.jl:35-40, 86-95). Contractions in the context's
 * precision. Operators, outputs and status codes as in
 * spp_measure_energy_batch; log_norm adds each state's 2 log_amplitude and
 * tensor_log_norm is that of the stored tensors. */
int spp_measure_energy_states(spp_ctx* ctx, int n, const spp_state* const* states, int chi, const double* X,
                              const double* F, const double* G, int factor_odd, double J, double h,
                              double* energy_total, double* energy_per_site, double* log_norm,
                              double* tensor_log_norm, double* max_trunc, int* status, int* fail_site,
                              double* boundary_entropy, double* max_boundary_entropy);

/* Batched `lognorm(state, BoundaryMPS(chi))` (src/contraction/boundary_mps.jl:162-191)
 * of n uploaded states (PEPS or PEPO, all on the same lattice with the same
 * leg spaces), e.g. the particles of `_observe` with energy = false
 * (src/driver/particles.jl:35-40). Contractions in the context's precision.
 *  log_norm [n]         log <psi|psi> + 2 log_amplitude;
 *  tensor_log_norm [n]  log <psi|psi> of the stored tensors;
 *  max_trunc [n]        max_environment_truncation of the north absorbs;
 *  status [n]           0 ok, 6 norm not finite and positive (boundary_mps.jl:188). */
int spp_lognorm_batch(spp_ctx* ctx, int n, const spp_state* const* states, int chi, double* log_norm,
                      double* tensor_log_norm, double* max_trunc, int* status);

/* Device bytes of the static scratch array behind spp_lognorm_batch_noalloc
 * (one per GPU the library runs on, held until the process ends). */
#define SPP_NOALLOC_SCRATCH_BYTES (32u << 20)

/* spp_lognorm_batch without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, outputs and results as spp_lognorm_batch; every device
 * buffer of the call (double layers, boundaries, the absorb workspace) is
 * carved from the library's static scratch array of
 * SPP_NOALLOC_SCRATCH_BYTES on the context's GPU, and states are processed in
 * chunks that fit it. A batch whose single state needs more than the array
 * holds fails with a message before any launch. All calls on one GPU share
 * that GPU's array: calls that overlap on one GPU (other contexts, other
 * threads) race and must be serialised by the caller. Launch timing
 * ("profile") is suspended during the call. */
int spp_lognorm_batch_noalloc(spp_ctx* ctx, int n, const spp_state* const* states, int chi, double* log_norm,
                              double* tensor_log_norm, double* max_trunc, int* status);

/* Gradient of log <psi|psi> with respect to every site tensor of n uploaded
 * states (same lattice and leg spaces), the adjoint of `lognorm`
 * (src/contraction/boundary_mps.jl:162-191). The reference has no adjoint; the
 * object it replaces is the single-site norm matrix `_site_hole`
 * (src/sampling/born.jl:75-83) built from `boundary_environments`
 * (boundary_mps.jl:193-209) and `row_environments` (boundary_mps.jl:261-273):
 *   grad_i = (N_i A_i + N_i^T A_i) / (A_i N_i A_i).
 * This is the environment gradient (truncations held fixed); it equals the
 * derivative of lognorm when chi reaches the exact boundary bond.
 *  grads [n][sum_i size_i]  FP64, site i (row-major, x fastest) dense
 *                           (W,N,P,A,E,S) in the parity-sorted leg order of
 *                           the upload (even indices first on every leg);
 *  log_norm, tensor_log_norm, max_trunc [n] as in spp_lognorm_batch, with
 *                           max_trunc over the north and the south absorbs;
 *  status [n]               0 ok, 3 nonfinite row environment
 *                           (boundary_mps.jl:258), 6 norm or local norm not
 *                           finite and nonzero (boundary_mps.jl:188). */
int spp_lognorm_grad_batch(spp_ctx* ctx, int n, const spp_state* const* states, int chi, double* grads,
                           double* log_norm, double* tensor_log_norm, double* max_trunc, int* status);

/* Host run of the per-bond full-update solve that the device FU kernel runs
 * (selftest; no GPU needed): `_initialize_pair` + `_optimize_pair` +
 * `_fu_balance_bond` (src/evolution/full_update.jl:31-117, 342-482) for one
 * bond. dims = {na, even_a, nb, even_b, d, even_d, D, even_D} (reduced core,
 * physical and bond dimensions, parity-sorted); options = {max_iterations,
 * error_tolerance, diagonal_shift, linear_max_iterations, linear_tolerance};
 * metric N[rb][sb][rk][sk]; target[r][p][q][s]; outputs the new cores
 * left[r][p][k], right[k][q][s] in original coordinates (bond zero-padded to D;
 * sector budget = the bond's even/odd counts) and record = {initial_cost,
 * final_cost, residual, target_norm, iterations, limit_reached, stop (0
 * disabled, 1 iteration limit, 2 error tolerance, 3 residual fallback),
 * status (0 ok, 1 nonfinite metric, 2 zero metric, 3 zero target, 4 PSD repair
 * needed, 5 nonfinite cost, 6 singular gauge, 7 workspace, 8 an eigen/SVD
 * solve reached its sweep cap), kept_even, kept_odd, floored eigenvalues,
 * nonconverged eigen/SVD solves}. Returns 0, or -3 on bad arguments. */
int spp_selftest_fu_bond(const int* dims, const double* options, const double* metric, const double* target,
                         double* left, double* right, double* record);

/* Host run of the bond solve's Jacobi eigensolver (kind 0: symmetric n x n A,
 * m = n; eigenvalues unordered) or one-sided Jacobi SVD (kind 1: m x n A,
 * m >= n; singular values descending) with at most max_sweeps sweeps
 * (selftest; no GPU needed). The device solve caps both at 80 sweeps and
 * reports a solve that reaches the cap (record slot 11, status 8), where the
 * reference retries gesdd -> gesvd and then raises (src/states/svd.jl:1-11).
 * converged: 1 when the sweeps stopped before the cap. Returns 0, or -3 on bad
 * arguments. */
int spp_selftest_jacobi(int kind, int m, int n, const double* A, int max_sweeps, double* values, int* converged);

/* Host run of the reduced-pair stage that the device FU kernel runs
 * (selftest; no GPU needed; `_reduced_pair`, src/evolution/full_update.jl:21-29):
 * QR of both sites of a bond over their live sectors. Sites dense
 * (W,N,P,A,E,S), parity-sorted; live_* [4 legs W,N,E,S][2]; gate[p][q][u][v].
 * Outputs: qa, qb pseudo-site Q tensors (the site's slots with Aux in P, a unit
 * A slot and r in the bond leg's slot; sized from geometry), ca[r][p][k],
 * cb[r][p][k] (k the bond), target[ra][p][q][rb], geometry = {rcap_a,
 * even_a, rcap_b, even_b, live r counts a even, a odd, b even, b odd}. Buffers
 * must hold the capacity sizes. Returns 0. */
int spp_selftest_fu_pair(int horizontal, const int* dims_a, const int* evens_a, const int* dims_b,
                         const int* evens_b, const double* site_a, const double* site_b, const int* live_a,
                         const int* live_b, const double* gate, double* qa, double* qb, double* ca, double* cb,
                         double* target, int* geometry);

/* Self-test of the matrix-core contraction kernel (no reference counterpart):
 * C[M][N] = A[M][K] B[K][N] (row-major, FP64 in/out) with each index a Z2
 * leg whose first evenX entries are even. Only parity-matching (neutral)
 * entries of C are computed; the others are returned as NaN. Pass
 * evenX = X for a plain dense product. layout bit 0: A is given as its
 * transpose [K][M]; bit 1: B is given as its transpose [N][K]. precision as
 * spp_set_precision (for this call only). */
int spp_selftest_gemm(spp_ctx* ctx, int precision, int M, int N, int K, int evenM, int evenN, int evenK,
                      int layout, const double* A, const double* B, double* C);

/* Self-test of the boundary-entropy eigensolver (no reference counterpart;
 * the reference reads the spectrum off its compact SVD, boundary_mps.jl:33-43):
 * all eigenvalues, ascending, of nprob symmetric blocks[nprob][ld][ld]
 * (row-major; only the leading n[i] x n[i] part of block i is used and its
 * two triangles are averaged) into eigenvalues[nprob][ld] (entries past n[i]
 * are 0). */
int spp_selftest_spectrum(spp_ctx* ctx, int nprob, int ld, const int* n, const double* blocks, double* eigenvalues);

/* Self-test of the physical-kind metric eigensolver (no reference
 * counterpart; the reference calls eigh_full, tensors.jl:71): nprob symmetric
 * blocks[nprob][ld][ld] (leading n[i] x n[i] part, triangles averaged) and two
 * vectors[nprob][2][ld] each; returns every eigenvalue (unordered) in
 * eigenvalues[nprob][ld], the projections u_j^T v of each vector on the
 * matching eigenvectors in projections[nprob][2][ld], and failed[nprob] = 1
 * where the QL iteration reached its cap. */
int spp_selftest_psd(spp_ctx* ctx, int nprob, int ld, const int* n, const double* blocks, const double* vectors,
                     double* eigenvalues, double* projections, int* failed);

/* Host run of the implicit QL iteration of the physical-kind metric (selftest;
 * no GPU needed): eigenvalues of the symmetric tridiagonal (d[n], off-diagonal
 * e[0..n-2], e[n-1] scratch) into d, and nvec vectors[nvec][n] replaced by
 * their projections on the eigenvectors. Returns 0, 1 when an eigenvalue
 * needed more than maxit iterations, or -3 on bad arguments. */
int spp_selftest_ql(int n, double* d, double* e, double* vectors, int nvec, int maxit);

/* Self-test of the factorised Gram transfer kernel (no reference
 * counterpart: the reference contracts the double layer, boundary_mps.jl:136-151):
 * single layer S[W][N][PHI][E][S] (d[4] / e[4]: dims and even counts of W, N,
 * E, S; dphi / ephi of PHI; FP64, host, one parity), one sample, bonds l and R
 * of dims nl / nr with even counts el / er. stage 0: x = z1[nl][nr][E][N][fold E]
 * -> y = z2[nl][nr][E][W][fold S] (contracting N and E); stage 1: x = z2 ->
 * y = z3[nl][nr][W][W][fold N] (contracting E and S); stage 2: x =
 * X[W][nl][nr][fold N] -> y = Y[nl][nr][E][fold S] (contracting W and N, the
 * draw's candidate left steps); pair legs (ket, bra) in
 * the parity-sorted order of the double layers, the last leg folded to the
 * parity its neighbours fix. rows_per_block 0: the default. Returns 0, -1 with
 * the error in spp_last_error, -2 when the layer does not fit the kernel. */
int spp_selftest_kron(spp_ctx* ctx, const int* d, const int* e, int dphi, int ephi, const double* S, int stage,
                      int nl, int el, int nr, int er, const float* x, float* y, int rows_per_block);

/* Calibration copy (no reference counterpart): one launch of the library's
 * plain copy kernel moving n FP32 values between two device buffers it
 * allocates, so that HBM traffic counters can be scaled by a known
 * 8*n bytes (4*n read, 4*n written). */
int spp_selftest_copy(spp_ctx* ctx, long long n);

#ifdef __cplusplus
}
#endif

#endif /* STOPEPS_HIP_H */
