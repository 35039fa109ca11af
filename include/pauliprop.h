/*
 * pauliprop.h — C ABI of the MI355X SparsePauliVector engine (libpauliprop.so).
 *
 * The reference (PauliOperators.jl v3.1.0) has no FFI: its interface is Julia
 * dispatch on SparsePauliVector{N,W,T} (src/PauliOperators.jl:42-110).  Each
 * entry point below replaces one of those methods for T = Float64 or
 * T = ComplexF64 (the reference's default, type_SparsePauliVector.jl:138,167);
 * the Julia/Python wrappers validate arguments and raise the reference's
 * exception types before calling in.
 *
 * Conventions
 *  - Keys cross the ABI as W little-endian uint64 words per term for z and
 *    for x (W = 1 for N <= 64, 2 for N <= 128), so a Julia Vector{UInt64} or
 *    Vector{UInt128} passes unchanged (type_SparsePauliVector.jl:23-34).
 *  - Coefficients cross as `double` arrays: one value per term for a Float64
 *    handle, two (re, im) per term for a ComplexF64 handle, so a Julia
 *    Vector{Float64} or Vector{ComplexF64} passes unchanged.  Every `c`
 *    argument of a single-vector call below follows the handle's type.
 *  - Host pointers are borrowed for the duration of a call; device state is
 *    owned by the handle.  One handle per host thread; calls are ordered on
 *    the handle's HIP stream.
 *  - Every function returns PP_OK or an error code; pp_last_error() gives the
 *    message of the last failure on the calling thread.
 */
#ifndef PAULIPROP_H
#define PAULIPROP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PP_OK = 0,
    PP_EARG = 1,         /* invalid argument (reference: ArgumentError / DimensionMismatch) */
    PP_ENOMEM = 2,       /* device allocation failed */
    PP_EHIP = 3,         /* HIP runtime error */
    PP_EOVERFLOW = 4,    /* capacity overflow that growth could not resolve */
    PP_EUNSUPPORTED = 5  /* mode not available in this build */
};

typedef struct pp_handle pp_handle;

/* Byte layout of MergeFilter (src/spv_kernels.jl:33-42). */
typedef struct {
    int64_t wmax, xwmax, mwmax;
    double thresh, alpha_w, dthresh_w, alpha_xw, dthresh_xw;
} pp_filter;

/* WindowCounters (src/spv_evolve.jl:149-161), summed over one pp_evolve call. */
typedef struct {
    int64_t rotations;      /* rotations applied */
    int64_t terms_swept;    /* sum over rotations of the live count entering it */
    int64_t terms_created;  /* children appended (anticommuting, kept by local filter) */
    int64_t merge_in;       /* sum of n + m entering each merge */
    int64_t merge_out;      /* sum of terms leaving each merge */
    int64_t early_merges;   /* capacity-forced merges (spv_evolve.jl:356-362) */
    double t_rotate;        /* seconds in sweep kernels (HIP events) */
    double t_merge;         /* seconds in the merge kernel (HIP events) */
} pp_counters;

/* One step of a boundary strategy that does not compile into a single
 * filter (truncation.jl:37-159; _apply! on SparsePauliVector,
 * spv_evolve.jl:59-119).  A CompositeTruncation becomes its members in order
 * (_apply_tup!), compilable members as PP_OP_FILTER.
 *   PP_OP_FILTER            compaction under `filter` (the clip family)
 *   PP_OP_ADAPTIVE          AdaptiveTruncation(max_terms, min_thresh = value):
 *                           exact top-k threshold (spv_evolve.jl:103-114)
 *   PP_OP_STOCHASTIC_COEFF  StochasticCoeffTruncation(epsilon = value):
 *                           stochastic_clip! (spv_evolve.jl:35-52) with a
 *                           counter-based generator keyed by `seed` (the Julia
 *                           RNG stream is not reproduced)
 *   PP_OP_STOCHASTIC_SAMPLING  StochasticSamplingTruncation(n_keep = max_terms):
 *                           exponential keys u^(1/|c|^2), the n_keep largest
 *                           kept (ties: lower index), rescaled to the original
 *                           Σ|c|^2 (spv_evolve.jl:69-99); same generator */
enum { PP_OP_FILTER = 0, PP_OP_ADAPTIVE = 1, PP_OP_STOCHASTIC_COEFF = 2, PP_OP_STOCHASTIC_SAMPLING = 3 };
typedef struct {
    int32_t kind;
    int32_t reserved;
    pp_filter filter;
    int64_t max_terms;
    double value;
    uint64_t seed;
} pp_strategy_op;

/* SparsePauliVector(N, Float64) with the reference's capacity bookkeeping
 * (type_SparsePauliVector.jl:94-104,167-188): live_cap and append_cap are the
 * reference's buffer lengths, which decide when an early merge happens.  Device
 * memory is sized separately and grows on demand. */
int pp_create(int N, int64_t live_cap, int64_t append_cap, pp_handle** out);
/* SparsePauliVector(N, T) for T = Float64 (complex_coeffs = 0) or ComplexF64
 * (complex_coeffs = 1; type_SparsePauliVector.jl:138-141).  ComplexF64 follows
 * Julia's arithmetic bit for bit: the rotation child is (T(k-2) sinθ) c as a
 * Complex * Complex product (spv_kernels.jl:163), |c| is Julia's hypot
 * (abs(::Complex)) in every filter, sums and real scalings are componentwise.
 * The resident layer launches hold either coefficient type (ComplexF64
 * chunks of at most 3072 terms at N > 64, 16-byte coefficients in LDS). */
int pp_create_typed(int N, int64_t live_cap, int64_t append_cap, int complex_coeffs, pp_handle** out);
/* *complex_coeffs = 1 for a ComplexF64 handle, 0 for Float64. */
int pp_coeff_type(pp_handle* h, int* complex_coeffs);
int pp_destroy(pp_handle* h);

/* Replace the live set by n terms.  sorted = 1: keys already strictly
 * increasing; sorted = 0: sorted here (the constructor's _sort_live!,
 * type_SparsePauliVector.jl:195-204); duplicate keys are an error. */
int pp_upload(pp_handle* h, const uint64_t* z, const uint64_t* x, const double* c, int64_t n, int sorted);
/* Copy the live set out (PauliSum(v), type_SparsePauliVector.jl:212-221). */
int pp_download(pp_handle* h, uint64_t* z, uint64_t* x, double* c, int64_t cap, int64_t* n);
int pp_size(pp_handle* h, int64_t* n);
int pp_capacities(pp_handle* h, int64_t* live_cap, int64_t* append_cap);

/* Pauli-sum products (spv_ops.jl:255-484), computed on the device of h (whose
 * N sets the word count; its live set is not touched):
 *   PP_ALG_MUL             out = op(A) · op(B), op = adjoint when conj != 0
 *                          (_mul_spv and Base.:* on SparsePauliVector and its
 *                          Adjoint, spv_ops.jl:257-298; a single Pauli is a
 *                          one-term operand)
 *   PP_ALG_COMMUTATOR      out = [A, B] (commutator!, spv_ops.jl:418-460)
 *   PP_ALG_ANTICOMMUTATOR  out = {A, B} (anticommutator!, spv_ops.jl:462-472)
 * complex_coeffs = 1: coefficients are ComplexF64 (re, im arrays);
 * complex_coeffs = 0: Float64 (im unused), where a product with a nonzero
 * imaginary part fails as the reference's convert does (InexactError) and the
 * commutator fails ("commutator coefficients are imaginary").  Output: keys
 * strictly increasing, equal keys summed (zero sums kept, as
 * _triples_to_live!), nout terms; cap < nout returns PP_EOVERFLOW with nout
 * set and nothing written (cap = 0 asks for the size).  The device buffers of
 * a call (operands, one coefficient per pair, the result) are kept on h and
 * grow as needed; pp_destroy releases them. */
enum { PP_ALG_MUL = 0, PP_ALG_COMMUTATOR = 1, PP_ALG_ANTICOMMUTATOR = 2 };
typedef struct {
    const uint64_t* z;    /* n x W words */
    const uint64_t* x;
    const double* re;
    const double* im;     /* complex_coeffs only */
    int64_t n;
    int32_t conj;         /* PP_ALG_MUL: use the adjoint (conjugated coefficients) */
} pp_terms;
int pp_algebra(pp_handle* h, int op, int complex_coeffs, const pp_terms* a, const pp_terms* b, uint64_t* oz,
               uint64_t* ox, double* ore, double* oim, int64_t cap, int64_t* nout);

/* Sequence evolution (spv_evolve.jl:312-387): L rotations exp(iθ/2 G) O exp(-iθ/2 G),
 * applied left to right.  gz, gx: L generators of W words each; cosv, sinv:
 * the caller's cos θ and sin θ (spv_evolve.jl:341-342).  strict: the compiled
 * truncation applied at every merge; local: the filter applied to each child
 * as it is created; window >= 1 rotations between merges (window > 1 sorts
 * the pending appends with a stable radix sort, so equal keys are summed in
 * append order where the reference's quicksort leaves the order unspecified).
 * counters may be NULL. */
int pp_evolve(pp_handle* h, const uint64_t* gz, const uint64_t* gx, const double* cosv, const double* sinv,
              int64_t L, int64_t window, const pp_filter* strict, const pp_filter* local, pp_counters* counters);

/* evolve!(v, gens, angs; window, truncation, local_truncation, correction)
 * with any strategy (spv_evolve.jl:312-387): nops > 0 merges with NOFILTER
 * and applies ops[0..nops) after each merge (the `compiled || _apply!` branch
 * of _boundary!, spv_evolve.jl:258-284); nops == 0 uses `strict`.
 * energy_ket != NULL is EnergyCorrection(ψ) (truncation.jl:262-306): the
 * change of ⟨ψ|v|ψ⟩ across every boundary, measured before the merge (live +
 * pending appends) and after the strategy, is added to *energy_acc.  With
 * variance_acc != NULL as well it is EnergyVarianceCorrection(ψ)
 * (truncation.jl:273-310): every boundary merges with NOFILTER, measures
 * ⟨ψ|v|ψ⟩ and variance(v, ψ) on the merged state, applies the strict filter
 * as a compaction (or the strategy), measures again and adds both changes
 * (spv_evolve.jl:246-270). */
int pp_evolve_ex(pp_handle* h, const uint64_t* gz, const uint64_t* gx, const double* cosv, const double* sinv,
                 int64_t L, int64_t window, const pp_filter* strict, const pp_strategy_op* ops, int nops,
                 const pp_filter* local, const uint64_t* energy_ket, double* energy_acc, double* variance_acc,
                 pp_counters* counters);

/* truncate!(v, strategy, corr) (truncation.jl:330-337) for any strategy given
 * as operations; energy_ket / energy_acc / variance_acc as in pp_evolve_ex
 * (may be NULL). */
int pp_truncate(pp_handle* h, const pp_strategy_op* ops, int nops, const uint64_t* energy_ket, double* energy_acc,
                double* variance_acc);

/* variance(v, ψ) (statistics.jl:6-22) for a computational-basis ket (W
 * words): Σ over target kets k' of |⟨k'|v|ψ⟩|^2 minus ⟨ψ|v|ψ⟩^2, with each
 * ⟨k'|v|ψ⟩ summed in key order as the reference accumulates it and the sum
 * over k' in double-double (the reference's order over k' is its Dict's). */
int pp_variance_ket(pp_handle* h, const uint64_t* ket, double* out);

/* get / getindex / haskey (type_SparsePauliVector.jl:286-311): for `count`
 * query keys (W words each in z, x), found[i] = 1 and c[i] = the coefficient
 * when the key is live, else found[i] = 0 and c[i] = 0.  Device binary search. */
int pp_lookup(pp_handle* h, const uint64_t* z, const uint64_t* x, int64_t count, double* c, int32_t* found);
/* pp_lookup without device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * results as pp_lookup; the staged query keys and results live in a fixed
 * per-GPU scratch array of 3.25 MiB, held until the process ends, so at most
 * 65536 queries per call (more: PP_EARG).  Calls that overlap on one GPU (from
 * several host threads or handles) race on that array. */
int pp_lookup_noalloc(pp_handle* h, const uint64_t* z, const uint64_t* x, int64_t count, double* c, int32_t* found);
/* setindex! for `count` keys (type_SparsePauliVector.jl:315-335): present
 * keys get the new coefficient, absent keys are inserted in key order (the
 * reference's live-capacity growth is applied per inserted key); for repeated
 * keys the last value wins. */
int pp_setindex(pp_handle* h, const uint64_t* z, const uint64_t* x, const double* c, int64_t count);
/* delete! for `count` keys (type_SparsePauliVector.jl:345-358); absent keys
 * are ignored; order-preserving. */
int pp_delete(pp_handle* h, const uint64_t* z, const uint64_t* x, int64_t count);
/* sum!(a, b) (spv_ops.jl:182-193): a += b, a merge of b's sorted terms into
 * a with NOFILTER (a and b distinct handles of the same N). */
int pp_sum(pp_handle* a, pp_handle* b);
/* mul!(v, s) (spv_ops.jl:12-17): every coefficient times the real s
 * (componentwise for ComplexF64, as Complex * Real). */
int pp_scale(pp_handle* h, double s);
/* mul!(v, s) for a ComplexF64 handle and a ComplexF64 scalar (re, im):
 * Julia's Complex * Complex.  A Float64 handle returns PP_EARG (InexactError). */
int pp_scale_complex(pp_handle* h, double re, double im);
/* norm(v, p) (spv_ops.jl:57-84): p = 2, 1, Inf or any p > 0. */
int pp_norm(pp_handle* h, double p, double* out);
/* inner_product(a, b) (spv_ops.jl:378-396): Σ over shared keys of
 * conj(c_a) c_b; out[0] (and out[1], the imaginary part, for ComplexF64).
 * a and b have the same coefficient type. */
int pp_inner_product(pp_handle* a, pp_handle* b, double* out);
/* The distance isapprox(a, b; atol, rtol) tests (spv_ops.jl:88-107):
 * sqrt(Σ over the union of keys of (c_a - c_b)^2), absent = 0. */
int pp_diff_norm(pp_handle* a, pp_handle* b, double* out);

/* Single rotation without truncation (spv_evolve.jl:228-236). */
int pp_evolve_single(pp_handle* h, const uint64_t* gz, const uint64_t* gx, double cosv, double sinv);

/* Order-preserving compaction under a filter (spv_kernels.jl:395-408): the
 * SPV clip family (spv_evolve.jl:13-26) and truncate! for compilable
 * strategies. */
int pp_compact(pp_handle* h, const pp_filter* f);

/* Diagonal Pauli channel (channels.jl:203-265).  kind: 0 depolarizing
 * popcount((z|x)&M), 1 dephasing popcount(x&M), 2 bit flip popcount(z&M),
 * 3 bit-phase flip popcount((x^z)&M), 4 general: t0[nX]*t1[nY]*t2[nZ].
 * Tables hold λ^k for k = 0..N as the caller computes them (Julia's
 * Float64^Int gives bit parity); t1, t2 are used by kind 4 only. */
int pp_scale_by_popcount(pp_handle* h, int kind, const uint64_t* mask, const double* t0, const double* t1,
                         const double* t2);

/* ⟨ψ|v|ψ⟩ for a computational-basis ket (spv_evolve.jl:133-134 ->
 * spv_kernels.jl:441-452); ket is W words; out[0] (and out[1], the imaginary
 * part, for ComplexF64).  Synchronises. */
int pp_expectation_ket(pp_handle* h, const uint64_t* ket, double* out);

/* copy(v) (type_SparsePauliVector.jl:249-254): dst takes src's live set and
 * the reference's virtual capacities, device to device (same N and T). */
int pp_copy(pp_handle* dst, pp_handle* src);
/* otimes(v1, v2) = v1 ⊗ v2 (spv_ops.jl:342-358): out (N = N1 + N2, same T,
 * a third handle) becomes the tensor product, keys (z1 | z2 << N1,
 * x1 | x2 << N1) (type_PauliBasis.jl:130-132), coefficients c1 * c2, sorted;
 * out's virtual capacities become the reference's (live max(16, n), append 16).
 * osum (spv_ops.jl:360-366) is v1 ⊗ I + I ⊗ v2: two otimes and a pp_sum. */
int pp_otimes(pp_handle* out, pp_handle* a, pp_handle* b);
/* largest(v) / largest_diag(v) (spv_ops.jl:520-553): the first term with the
 * largest |c| (diag_only: among the x == 0 terms): its key (W words each),
 * coefficient (cw doubles) and 0-based index.  An empty vector (or no
 * diagonal term) is PP_EARG ("collection must be non-empty"). */
int pp_largest(pp_handle* h, int diag_only, uint64_t* z, uint64_t* x, double* c, int64_t* index);
/* diag(v) (offdiag = 0) / offdiag(v) (offdiag = 1) in place (spv_ops.jl:555-587):
 * keep the terms with x == 0 (x != 0), order preserved.  The reference's
 * functions copy first (pp_copy). */
int pp_select_diag(pp_handle* h, int offdiag);
/* Matrix(v) (spv_ops.jl:109-115) for N <= 13: the 2^N x 2^N matrix,
 * column-major (Julia's layout), one double (Float64) or two (re, im) per
 * entry, Σ over terms in key order of Matrix(op) .* c with site 1 the least
 * significant index bit (type_PauliBasis.jl:71-95).  A Float64 vector with an
 * imaginary entry (a term with an odd number of Y) is PP_EARG (InexactError). */
int pp_matrix(pp_handle* h, double* out);
/* check_spv key-order invariant (type_SparsePauliVector.jl:263-278): *ok = 1
 * if the live keys are strictly increasing. */
int pp_check(pp_handle* h, int* ok);

/* ------------------------------------------------------------------------
 * Batches: B independent SparsePauliVectors (qubit counts N[b] <= 128, one
 * word count for the batch) evolved together, one kernel sequence per
 * rotation step for all of them.  This has no counterpart in the reference,
 * which evolves one vector per call; each vector keeps exactly the semantics
 * of the single-vector calls above (one evolve! call per vector per
 * pp_batch_evolve, with the reference's capacity bookkeeping per vector).
 * live_cap / append_cap as in pp_create, per vector.  Batch vectors hold
 * Float64 coefficients.
 */
typedef struct pp_batch pp_batch;
int pp_batch_create(int B, const int* N, const int64_t* live_cap, const int64_t* append_cap, pp_batch** out);
int pp_batch_destroy(pp_batch* p);
/* Words per key of this batch (1 or 2). */
int pp_batch_words(pp_batch* p, int* W);
int pp_batch_upload(pp_batch* p, int b, const uint64_t* z, const uint64_t* x, const double* c, int64_t n, int sorted);
int pp_batch_download(pp_batch* p, int b, uint64_t* z, uint64_t* x, double* c, int64_t cap, int64_t* n);
/* n[b] = live terms of vector b. */
int pp_batch_sizes(pp_batch* p, int64_t* n);
/* nsteps rotation steps; step s of vector b uses generator gz/gx[(s*B+b)*W ..]
 * and cosv/sinv[s*B+b]; active (may be NULL = all) switches a vector off for
 * a step.  Per vector: strict[b] and local[b] (local may be NULL) as in
 * pp_evolve; counters (may be NULL) receive B records. */
int pp_batch_evolve(pp_batch* p, int64_t nsteps, const uint64_t* gz, const uint64_t* gx, const double* cosv,
                    const double* sinv, const uint8_t* active, const pp_filter* strict, const pp_filter* local,
                    pp_counters* counters);
/* Per-vector channel as pp_scale_by_popcount: kind[b] in -1 (none) .. 4,
 * mask[b*W ..], tables[(b*3 + t)*129 + k] = table t, entry k. */
int pp_batch_scale_by_popcount(pp_batch* p, const int* kind, const uint64_t* mask, const double* tables);
/* out[b] = ⟨ψ_b|v_b|ψ_b⟩ for kets[b*W ..]. */
int pp_batch_expectation_ket(pp_batch* p, const uint64_t* kets, double* out);
/* Pre-size device buffers: vector b holds cap[b] live terms without regrowth
 * (buffers grow on demand regardless; this only avoids stall-and-replay). */
int pp_batch_reserve(pp_batch* p, const int64_t* cap);
/* Tile shapes: sweep tile 512 or 2048 terms, merge tile 512, 1024 or 2048
 * outputs (2048 only for N <= 64); 0 = chosen from the live counts. */
int pp_batch_set_tiles(pp_batch* p, int tile, int merge_tile);
/* Capture each evolve call's launch sequence (capacity precheck, compaction,
 * every rotation step) in a hipGraph and replay it on later calls with the
 * same shape (off by default; no effect while timing is on). */
int pp_batch_set_graphs(pp_batch* p, int on);
/* Graphs captured and graph launches so far. */
int pp_batch_graph_stats(pp_batch* p, int64_t* captures, int64_t* launches);
/* HIP-event timing of every merge-kernel launch (off by default). */
int pp_batch_set_timing(pp_batch* p, int on);
int pp_batch_merge_time(pp_batch* p, double* seconds, int64_t* launches, int reset);

/* Resident layer launches: an execution mode of pp_batch_evolve and
 * pp_evolve (window = 1) with no counterpart in the reference, whose
 * evolve! runs rotation by rotation (spv_evolve.jl:352-385).  One kernel
 * launch runs every rotation of the call with each vector's terms held in the
 * LDS of a cluster of workgroups on one XCD; results are identical to the
 * per-step kernels.  mode 0: off; 1 (default): calls of at least 8 steps
 * whose vectors fit the clusters with room to double; 2: every call the
 * kernel can take (vectors that outgrow their cluster finish the call on the
 * per-step kernels).  cap_limit > 0 lowers the chunk capacity per workgroup
 * (exercises those fallbacks); 0 = the full LDS. */
int pp_batch_set_resident(pp_batch* p, int mode, int cap_limit);
int pp_set_resident(pp_handle* h, int mode, int cap_limit);
/* Resident launches so far, vectors finished by the per-step kernels after a
 * resident launch stopped them, and aborted launches (after an abort the
 * batch uses the per-step kernels only). */
int pp_batch_resident_stats(pp_batch* p, int64_t* launches, int64_t* fallbacks, int64_t* aborts);
/* HIP-event time of resident launches while timing is on (pp_batch_set_timing). */
int pp_batch_resident_time(pp_batch* p, double* seconds, int reset);
/* The resident launches' share of the counters since creation (or the last
 * reset): rotations, terms entering them, terms leaving their merges, and
 * chunk rebalances (a cluster whose chunk would outgrow its LDS redistributes
 * its terms and redoes the rotation). */
int pp_batch_resident_work(pp_batch* p, int64_t* rotations, int64_t* swept, int64_t* merge_out, int64_t* rebalances,
                           int reset);

/* Test hook, no counterpart in the reference: move the merge look-back tag
 * base forward to `epoch` (1 .. 2^22 - 2, not below the tags already used),
 * so that tests reach the 22-bit tag wrap, where every look-back array is
 * cleared and the base restarts at 1, without 4 million launches.  pp_set_epoch
 * also moves the windowed merge's tag. */
int pp_batch_set_epoch(pp_batch* p, uint32_t epoch);
int pp_set_epoch(pp_handle* h, uint32_t epoch);

/* ---- Learned surrogate of ⟨O⟩(θ) -------------------------------------------
 * No counterpart in the reference: an ensemble of E MLPs
 * din -> hidden -> ... (layers hidden layers, GELU) -> 1, trained with Adam on
 * ⟨O⟩ labels produced by the propagation path above, used with an ensemble
 * disagreement gate that falls back to exact propagation.  BF16 matrix-core
 * GEMMs with FP32 accumulation; FP32 master weights and optimizer state.
 * din <= 64; hidden a multiple of 64; batch sizes multiples of 64.
 * Host parameter layout per member (floats, pp_mlp_param_count of them):
 * W_1[hidden][din], b_1[hidden], then (W_l[hidden][hidden], b_l[hidden]) for
 * each further hidden layer, then w_out[hidden], b_out. */
typedef struct pp_mlp pp_mlp;
int pp_mlp_create(int E, int din, int hidden, int layers, pp_mlp** out);
int pp_mlp_destroy(pp_mlp* h);
int pp_mlp_param_count(pp_mlp* h, int64_t* count);
/* params[e*count ..] for every member; resets the Adam state. */
int pp_mlp_set_params(pp_mlp* h, const float* params);
int pp_mlp_get_params(pp_mlp* h, float* params);
/* Training set: x[n][din], y[n], copied to the device. */
int pp_mlp_load_data(pp_mlp* h, const float* x, const float* y, int64_t n);
/* One epoch of floor(n/batch) Adam steps (lr; betas 0.9, 0.999; eps 1e-8) on
 * the mean squared error; member e takes rows in the order perm[e*n ..].
 * loss[e]: mean minibatch loss of member e over the epoch. */
int pp_mlp_train_epoch(pp_mlp* h, const int32_t* perm, int batch, float lr, float* loss);
/* Gradients of each member's mean squared error on one batch x[batch][din],
 * t[batch] (grads[e*count ..], host layout) and the losses (may be NULL). */
int pp_mlp_gradients(pp_mlp* h, const float* x, const float* t, int batch, float* grads, float* loss);
/* Predictions for x[n][din]: members[e*n + i] (may be NULL), ensemble mean
 * and population standard deviation; seconds (may be NULL) = device time. */
int pp_mlp_predict(pp_mlp* h, const float* x, int64_t n, float* members, double* mean, double* sd,
                   double* seconds);

/* Julia's Float64^Int (compensated power by squaring, base/math.jl), for
 * callers that are not Julia and want the same channel tables. */
double pp_julia_powi(double x, int64_t n);

/* Message of the last failed call on this thread ("" if none). */
const char* pp_last_error(void);

/* Library version string. */
const char* pp_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PAULIPROP_H */
