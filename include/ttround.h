/*
 * ttround.h -- C-ABI of libttround, the MI355X (gfx950) TT-rounding library.
 *
 * Drop-in boundary for the rounding hot path of the Julia package
 * TensorTrains (rodrigofj99/Tensor-Train-Julia).  Every entry point names the
 * reference function it replaces (file:line in that repository).  A Julia
 * caller reaches these with ccall; the host mirror in
 * tensor-train-julia_amd/ttround/ reaches them with ctypes.
 *
 * Conventions
 *  - All functions return a status code (TTR_OK == 0); on failure
 *    ttr_last_error(h) describes it.  No C++ exception crosses this ABI.
 *  - Host cores are passed in the reference layout: core k of a TT vector is a
 *    column-major (n_k, r_{k-1}, r_k) array of doubles (TTvector.ttv_vec[k],
 *    src/tt_tools.jl:24-34); core k of a TT operator is (n_k, n_k, R_{k-1}, R_k)
 *    (TToperator.tto_vec[k], src/tt_tools.jl:61-71).  Ranks arrays have d+1
 *    entries with r_0 = r_d = 1.  Host buffers are borrowed for the duration of
 *    the call only.
 *  - TT objects live in device memory owned by the handle and are named by
 *    int64 ids.  Results are new objects; the library never frees caller memory
 *    and callers release objects with ttr_tt_free.
 *  - One handle per host thread; a handle owns one HIP stream and its
 *    workspace.  Every call is synchronous on return.
 */
#ifndef TTROUND_H
#define TTROUND_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  TTR_OK = 0,
  TTR_INVALID_ARG = 1,       /* reference: ArgumentError / AssertionError       */
  TTR_DIM_MISMATCH = 2,      /* reference: DimensionMismatch / AssertionError   */
  TTR_OOM = 3,
  TTR_NUMERICAL_FAILURE = 4, /* reference: LAPACKException                      */
  TTR_DEVICE_ERROR = 5,
  TTR_NOT_FOUND = 6          /* unknown object id                               */
};

enum { TTR_PREC_F64 = 0, TTR_PREC_F32 = 1, TTR_PREC_BF16 = 2 };
enum { TTR_DIR_LEFT = 0, TTR_DIR_RIGHT = 1 };

typedef struct ttr_handle_s* ttr_handle_t;

/* library identification */
const char* ttr_version(void);

/* creates a handle on HIP device `device` (one stream, empty workspace).  The
   handle times the GEMM tile configurations on the first call at each site and
   keeps the fastest (every configuration gives identical bits); with the
   environment variable TTR_GEMM_TUNE=0 set at creation it uses fixed tiles. */
int ttr_create(ttr_handle_t* h, int device);
int ttr_destroy(ttr_handle_t h);
/* message of the last failed call on this handle ("" if none) */
const char* ttr_last_error(ttr_handle_t h);

/* ---------------------------------------------------------------- objects */

/* Uploads a TTvector.  ot may be NULL (all 0).
 * Replaces: the TTvector constructor (src/tt_tools.jl:30-33). */
int ttr_tt_upload(ttr_handle_t h, int64_t d, const int64_t* dims, const int64_t* rks, const int64_t* ot,
                  const double* const* cores, int64_t* id);

/* Metadata of a TT object; any output pointer may be NULL.
 * dims: d entries, rks: d+1, ot: d. */
int ttr_tt_info(ttr_handle_t h, int64_t id, int64_t* d, int64_t* dims, int64_t* rks, int64_t* ot);

/* Copies core k (0-based) to dst in the reference layout (n_k, r_{k-1}, r_k). */
int ttr_tt_download_core(ttr_handle_t h, int64_t id, int64_t k, double* dst);

/* Releases a TT or a TT operator (the ids share one space). */
int ttr_tt_free(ttr_handle_t h, int64_t id);

/* Seeded TT with i.i.d. N(0,1) cores drawn on the device, ranks capped as in
 * rand_tt (src/tt_tools.jl:255-303, r_and_d_to_rks with rmax = max(rks)).
 * Draw e of core k (reference order) is the library's Philox4x32-10 normal
 * number (seed, k, e) of the input stream. */
int ttr_tt_randn(ttr_handle_t h, int64_t d, const int64_t* dims, const int64_t* rks, uint64_t seed, int64_t* id);

/* Uploads a TToperator (MPO).  Replaces: the TToperator constructor
 * (src/tt_tools.jl:67-70). */
int ttr_tto_upload(ttr_handle_t h, int64_t d, const int64_t* dims, const int64_t* rks, const double* const* cores,
                   int64_t* id);

/* ------------------------------------------------------- rank-growing ops */

/* out = sum_j alpha_j * x_j with block-diagonal cores; the scalar multiplies
 * the first core with ot == 0 of each term.
 * Replaces: *(::Vector{TTvector}, ::Vector) (src/tt_operations.jl:252-258),
 * +(::TTvector, ::TTvector) (:81-112), *(::Number, ::TTvector) (:428-441),
 * -(::TTvector, ::TTvector) (:453-455). */
int ttr_lincomb(ttr_handle_t h, int64_t nterms, const double* alpha, const int64_t* ids, int64_t* out_id);

/* out = A * x (MPO times TT), bond index with the operator rank fastest.
 * Replaces: *(::TToperator, ::TTvector) (src/tt_operations.jl:170-194). */
int ttr_mpo_apply(ttr_handle_t h, int64_t op_id, int64_t x_id, int64_t* out_id);

/* out = x_1 .* x_2 .* ... .* x_m (entrywise product of m <= 8 TTs with equal
 * dims): core k of out holds, per mode index, the Kronecker product of the
 * factors' core-k slices, ranks prod_f r_f(k), ot all zero.  first_fastest == 0:
 * the LAST factor's rank index runs fastest (kron(x, y) per slice).
 * Replaces: *(::TTvector, ::TTvector) (src/tt_operations.jl:273-290).
 * first_fastest == 1: the FIRST factor's index runs fastest, the layout of the
 * implicit product in the NTuple methods ttrand_rounding
 * (src/tt_randtools.jl:364-466) and tt_recursive_sketch / tt_sketch /
 * tt_combined_sketch (src/sketches.jl:474-604, 819-909, 1003-1040): rounding or
 * sketching the product built this way gives their results (their sketch
 * tensors W_k are the rows of the product's W_k in this order).  Every product
 * core must stay below 2^31 entries. */
int ttr_hadamard(ttr_handle_t h, int64_t m, const int64_t* ids, int32_t first_fastest, int64_t* out_id);

/* ------------------------------------------------- deterministic rounding */

/* Orthogonalizes n TTs around root (1-based, as the reference) with
 * Householder QR (left cores) and LQ (right cores), FP64.
 * Replaces: orthogonalize(x; i) (src/tt_rounding.jl:65-97). */
int ttr_orthogonalize(ttr_handle_t h, int64_t n, const int64_t* ids, int64_t root, int64_t* out_ids);

/* Deterministic SVD rounding of n TTs (batched), FP64: orthogonalization
 * sweep unless the input is already left- (right-) orthogonal, then the
 * truncated-SVD sweep with the reference's floor rule and rmax.
 * Replaces: tt_rounding(x; tol, rmax, direction) (src/tt_rounding.jl:146-154,
 * 174-269) and tt_rounding!(x; ...) (:158-167). */
int ttr_round_svd(ttr_handle_t h, int64_t n, const int64_t* ids, double tol, int64_t rmax, int direction,
                  int64_t* out_ids);

/* Frobenius norms of n TTs.  Replaces: norm(::TTvector) (src/tt_rounding.jl:107-114). */
int ttr_norm(ttr_handle_t h, int64_t n, const int64_t* ids, double* out);

/* ------------------------------------------------------- inner products */

/* Euclidean inner products <x_i, y_i> of n pairs (x_i and y_i of equal dims),
 * FP64, folded left to right with every pair in the same launches.
 * Replaces: dot(A, B) (src/tt_operations.jl:302-321) and dot_par(A, B)
 * (:339-370), the Gram-Schmidt products of the Krylov solvers
 * (src/tt_solvers.jl:69,90,279-280,315). */
int ttr_dot(ttr_handle_t h, int64_t n, const int64_t* x_ids, const int64_t* y_ids, double* out);

/* Same results as ttr_dot, computed without allocating anything: no hipMalloc,
 * hipFree, pinned-memory allocation or allocating library call, on the first
 * call as on every later one.  The environments, temporaries, split-K partial
 * sums and GEMM descriptors go into a static device array of 64 MiB (one per GPU,
 * held until the process ends) and a pinned host staging buffer of 4 MiB that
 * ttr_create allocates; a call whose environments need more than 32 MiB
 * (2 r_x r_y + r_x n r_y doubles per pair at the largest bond) fails with
 * TTR_INVALID_ARG.  Calls on one GPU share that array: calls that overlap on
 * one GPU (handles used concurrently from several host threads) race.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight solver or serving loops such as
 * the Gram-Schmidt passes of src/tt_solvers.jl:69,90,279-280,315). */
int ttr_dot_noalloc(ttr_handle_t h, int64_t n, const int64_t* x_ids, const int64_t* y_ids, double* out);

/* All partial contractions of x and y: W[k] (r^x_k x r^y_k, column-major,
 * k = 0..d) into W_out[k] (NULL entries skipped); reverse != 0 contracts the
 * cores k..d-1 into W[k] (W[0] = <x, y>, W[d] = 1), reverse == 0 the cores
 * 0..k-1 (W[0] = 1, W[d] = <x, y>).
 * Replaces: partial_contraction(A, B; reverse) (src/sketches.jl:1047-1070). */
int ttr_partial_contraction(ttr_handle_t h, int64_t x_id, int64_t y_id, int reverse, double* const* W_out);

/* Bilinear forms <psi_i|H_i|phi_i> of n triples without forming H phi, FP64.
 * Replaces: dot_operator(psi, H, phi) (src/tt_operations.jl:385-421), the
 * energy evaluations of src/tt_solvers.jl:586,607. */
int ttr_dot_operator(ttr_handle_t h, int64_t n, const int64_t* psi_ids, const int64_t* op_ids, const int64_t* phi_ids,
                     double* out);

/* ---------------------------------------------------- randomized rounding */

typedef struct {
  int64_t nterms;           /* number of terms of the linear combination            */
  const double* alpha;      /* nterms coefficients                                  */
  const int64_t* term_ids;  /* nterms TT ids (same dims)                            */
  int64_t op_id;            /* -1, or an operator id: the first term is then         */
                            /* alpha[0] * H * term_ids[0], sketched and projected    */
                            /* without forming the product                           */
  const int64_t* rks;       /* d+1 target sketch ranks (rks[0] = rks[d] = 1)         */
} ttr_rand_item;

typedef struct {
  int32_t orthogonal;   /* generate_sketch_blocks(...; orthogonal) (src/sketches.jl:44-84)   */
  int32_t sketch_prec;  /* TTR_PREC_*: precision of the sketch contractions                   */
  int64_t block_rks;    /* <= 0 means d (the reference default block_rks = N)                */
  uint64_t seed;        /* the reference default is 1234 (src/tt_randtools.jl:112)           */
  const int64_t* p;     /* optional d+1 block counts (tt_recursive_sketch kwarg p), or NULL   */
  double oversampling;  /* block-count multiplier (kwarg oversampling; 1 = none)              */
  int32_t forward;      /* 0: right-to-left sketch (kwarg reverse=true, the default); 1: the  */
                        /* left-to-right sketch (reverse=false; ttr_sketch, ttr_sketch_blocks) */
  /* Optional caller-supplied sketch blocks, or NULL for the library's draws of `seed`.
   * blocks[k] (k = 0..d-1) holds core k's blocks exactly as the reference's
   * generate_sketch_blocks returns them (src/sketches.jl:44-84, already scaled or
   * orthogonalised), column-major: (bv_{k+1} n_k) x bv_k x p_k for the
   * right-to-left sketch, (bv_k n_k) x bv_{k+1} x p_{k+1} for the left-to-right
   * one (src/sketches.jl:232-237, 279-281), FP64 (cast to sketch_prec on the
   * device).  A Julia caller passes the arrays that seeding default_rng with
   * `seed` and calling generate_sketch_blocks core by core in the reference's
   * order yields, so the rounding uses the reference's own draws; `orthogonal`
   * and `seed` are then not used for the blocks.  Every item of a batch must
   * have the same block counts.  Read during the call only. */
  const double* const* blocks;
  /* ttr_stta / ttr_stta_sketch only: the right-to-left sketch's blocks (seed_right);
   * `blocks` then holds the left-to-right sketch's.  NULL elsewhere. */
  const double* const* blocks_right;
  /* 0: the sketch tensors W_k and the contractions that form them run in
   * sketch_prec (FP32 / BF16 MFMA with FP32 accumulation); 1: only the blocks are
   * rounded to sketch_prec, W_k and the contractions are FP64 (the computation
   * the FP64 oracle performs on the same rounded blocks). */
  int32_t sketch_accum;
} ttr_rand_opts;

/* Randomize-then-orthogonalize rounding of n linear combinations in one
 * batch: per item a right-to-left block Gaussian TT sketch of each term (same
 * draws for every term and item), then one left-to-right sweep of
 * Z_k = sum_t Y_t W_t, Householder QR (FP64) and projection (FP64).  Output
 * cores are left-orthogonal (ot = [1,...,1,0]) with ranks
 * min(n_k r_{k-1}, sketch_rks_k), untruncated.
 * Replaces: ttrand_rounding(y, rks; orthogonal, seed, block_rks)
 * (src/tt_randtools.jl:112-154), ttrand_rounding(y, rmax) (:162-167),
 * ttrand_rounding(alpha, y, rks) (:223-276) and its rmax form (:278-283);
 * with op_id >= 0 also ttrand_rounding(alpha, A, y, rks) (:285-354), its rmax
 * form (:356-361) and ttrand_rounding(A, y, b, rks) (:169-221, as alpha = (1, -1),
 * terms (y, b)).  Operator bonds need n_k * R_k <= 64. */
int ttr_round_rand(ttr_handle_t h, int64_t n, const ttr_rand_item* items, const ttr_rand_opts* opts,
                   int64_t* out_ids);

/* Randomized rounding of n entrywise (Hadamard) products x_1 .* ... .* x_M
 * without forming them.  Item i has nfactors[i] factors (1..8, equal dims,
 * rks[0] = rks[d] = 1); factor_ids lists the items' factors one after the other;
 * rks holds d+1 target sketch ranks per item.  The product's bond k carries
 * the factors' rank indices with the first factor fastest.  The right-to-left
 * sketch contracts one factor at a time per mode index (cost n_k times the sum
 * over factors of the rank products, instead of n_k P_k P_{k+1} for the formed
 * product core); the left-to-right sweep keeps the left part as
 * (r_k n_k) x prod_m r_{m,k+1} and applies the next cores factor by factor.
 * Output as for ttr_round_rand (left-orthogonal, untruncated ranks).
 * Replaces: ttrand_rounding(y::NTuple{M,TTvector}, rks; orthogonal, seed,
 * block_rks) (src/tt_randtools.jl:364-458) and its rmax form (:460-466). */
int ttr_round_rand_tuple(ttr_handle_t h, int64_t n, const int64_t* nfactors, const int64_t* factor_ids,
                         const int64_t* rks, const ttr_rand_opts* opts, int64_t* out_ids);

/* The recursive sketch of the entrywise product of m TTs, never formed:
 * W_out[k] (prod_f r_{f,k} x sketch_rks[k], column-major, the first factor's
 * rank index fastest; NULL entries skipped), sketch ranks into sketch_rks (d+1).
 * opts->forward = 0: right to left (any sketch precision); opts->forward = 1:
 * left to right, FP64, blocks from the forward structure (opts->p and
 * opts->oversampling as for ttr_sketch).
 * Replaces: tt_recursive_sketch(T, A::NTuple, rks; reverse=true|false, ...)
 * (src/sketches.jl:474-604) with contract_sketch_core_kronecker_backwards!
 * (src/sketches.jl:1735-1881) and contract_sketch_core_kronecker_forwards!
 * (src/sketches.jl:1332-1478). */
int ttr_sketch_tuple(ttr_handle_t h, int64_t m, const int64_t* factor_ids, const int64_t* rks,
                     const ttr_rand_opts* opts, double* const* W_out, int64_t* sketch_rks);

/* The recursive sketch of one TT: W[k] (r_k x sketch_rks[k], column-major,
 * k = 0..d) into W_out[k] (NULL entries skipped), the sketch ranks into
 * sketch_rks (d+1).  opts->forward selects the left-to-right sketch (FP64 only).
 * Replaces: tt_recursive_sketch(T, A, rks; reverse=true, ...)
 * (src/sketches.jl:175-240,300-303) and tt_recursive_sketch(T, A, rks;
 * reverse=false, ...) (:241-303). */
int ttr_sketch(ttr_handle_t h, int64_t tt_id, const int64_t* rks, const ttr_rand_opts* opts, double* const* W_out,
               int64_t* sketch_rks);

/* The recursive sketch of H x without forming the product: W[k]
 * ((r_k R_k) x sketch_rks[k], column-major, row index x rank + r_k * H rank).
 * opts->forward selects the left-to-right sketch (FP64; n_k R_k <= 64; as in the
 * reference, opts->oversampling is not applied on that path).
 * Replaces: tt_recursive_sketch(T, A::TToperator, x, rks; reverse=true, ...)
 * (src/sketches.jl:317-389,456-459,463-472) and its reverse=false form
 * (:390-459, contract_sketch_core_forwards! for operators :1240-1309). */
int ttr_sketch_op(ttr_handle_t h, int64_t op_id, int64_t tt_id, const int64_t* rks, const ttr_rand_opts* opts,
                  double* const* W_out, int64_t* sketch_rks);

/* Boundary-only reverse sketch of n TTs (op_ids NULL or op_ids[i] < 0), or of
 * H_i x_i without forming the product (op_ids[i] >= 0), all sharing one draw
 * stream: the block count p = cld(s, bv_1) is the same at every core, and
 * only vec(W_1) / sqrt(p) is kept, bv_1 * p = *s_actual values per item.
 * Item i goes to out[i * s_actual ...]; out_len must be >= n * s_actual.
 * out == NULL only computes *s_actual.  opts->p and opts->oversampling are
 * ignored (p follows from s).  The extra sketch of tt_combined_sketch is this
 * call with seed + 1.
 * Replaces: tt_sketch(T, A, s; reverse=true, orthogonal, seed, block_rks)
 * (src/sketches.jl:645-723) and tt_sketch(T, H, A, s; reverse=true, ...)
 * (:729-813). */
int ttr_tt_sketch(ttr_handle_t h, int64_t n, const int64_t* tt_ids, const int64_t* op_ids, int64_t s,
                  const ttr_rand_opts* opts, double* out, int64_t out_len, int64_t* s_actual);

/* The sketch blocks of core k (0-based) exactly as the device uses them,
 * upcast to FP64, in the reference order (bv_{k+1}, n_k, bv_k, p_k)
 * column-major (the array generate_sketch_blocks returns, src/sketches.jl:83);
 * with opts->forward the blocks of the left-to-right sketch, (bv_k, n_k,
 * bv_{k+1}, p_{k+1}) (src/sketches.jl:279-281).  Writes the product of the 4
 * sizes; shape4 (may be NULL) gets them. */
int ttr_sketch_blocks(ttr_handle_t h, int64_t d, const int64_t* dims, const int64_t* rks, const ttr_rand_opts* opts,
                      int64_t k, double* out, int64_t* shape4);

/* Streaming TT approximation of n linear combinations (two-sided sketches):
 * per item the left-to-right sketch of every term (block counts of rks, times
 * 1.5; seed opts->seed) and the right-to-left one (seed_right), Omega_k =
 * W_L[k+1]^T W_R[k+1] and the sketched cores Psi_k summed over the terms with
 * their alpha, then per bond the pseudo-inverse of Omega_k (singular values
 * above 1e-12 max) applied to Psi_k from the right (Omega_k wide) or to
 * Psi_{k+1} from the left.  FP64; op_id must be -1; opts->p must be NULL.
 * Output ot = 0.
 * Replaces: stta(y, rks; seed_left, seed_right, orthogonal, block_rks)
 * (src/tt_randtools.jl:552-580), its rmax form (:582-589) and
 * stta(alpha, y, rks; ...) (:591-638). */
int ttr_stta(ttr_handle_t h, int64_t n, const ttr_rand_item* items, const ttr_rand_opts* opts, uint64_t seed_right,
             int64_t* out_ids);

/* The sketch stage of ttr_stta for one TT: Omega_out[k] (sl_{k+1} x sr_{k+1},
 * column-major, k = 0..d-2) and Psi_out[k] (reference layout (n_k, sl_k,
 * sr_{k+1}), k = 0..d-1), the left and right sketch ranks into sl, sr (d+1).
 * NULL output entries are skipped.
 * Replaces: stta_sketch(x, rks; seed_left, seed_right, orthogonal, block_rks)
 * (src/sketches.jl:1126-1156). */
int ttr_stta_sketch(ttr_handle_t h, int64_t tt_id, const int64_t* rks, const ttr_rand_opts* opts,
                    uint64_t seed_right, double* const* Omega_out, double* const* Psi_out, int64_t* sl,
                    int64_t* sr);

/* stta_sketch(x, L, R) with explicit sketch TTs L (left) and R (right), all of
 * equal dims: Omega_out[k] (rL_{k+1} x rR_{k+1}, column-major, k = 0..d-2) =
 * the left partial contraction of (x, L) through core k, transposed, times the
 * right one of (x, R) from core k+1, and Psi_out[k] (reference layout
 * (n_k, rL_k, rR_{k+1}), k = 0..d-1) = core k of x contracted with both; NULL
 * entries are skipped.  FP64.
 * Replaces: stta_sketch(x, L, R) (src/sketches.jl:1112-1124). */
int ttr_stta_sketch_tt(ttr_handle_t h, int64_t x_id, int64_t l_id, int64_t r_id, double* const* Omega_out,
                       double* const* Psi_out);

/* Raw N(0,1) draws number 0..count-1 of core k of the sketch stream. */
int ttr_sketch_draws(ttr_handle_t h, int64_t k, int64_t count, uint64_t seed, double* out);

/* ------------------------------------------------------ instrumentation */

/* hipEvent timing of every kernel launch and sweep phase on the handle's
 * stream, with the algorithmic flops and bytes of each launch.  Phase names
 * ("phase:reverse_sketch", "phase:orthogonalization", ...) follow the
 * reference's TimerOutputs sections (src/tt_randtools.jl:113-153,
 * src/sketches.jl:176-304).  Statistics accumulate until ttr_profile_reset.
 * on: 0 off, 1 every timed launch, 2 only the GEMM classes (with their call
 * sites) and the phases, which keeps the host cost of the events off the other
 * launches. */
int ttr_profile_enable(ttr_handle_t h, int on);
int ttr_profile_reset(ttr_handle_t h);
int ttr_profile_count(ttr_handle_t h, int64_t* count);
int ttr_profile_get(ttr_handle_t h, int64_t i, char* name, int64_t name_len, double* ms, int64_t* launches,
                    double* flops, double* bytes);

/* HIP graphs of ttr_round_rand (no counterpart in the reference, whose sweeps
 * run eagerly; the control exists for callers that repeat one call shape, such
 * as a time-stepping loop).  With graphs on, per call key (item count, TT
 * dimensions and ranks, target ranks, options and seed) the first call runs
 * eagerly, the second is captured into a graph with its own workspace, and
 * later calls replay it: only the host side runs, writing the call's
 * descriptors (TT addresses, coefficients) into pinned staging that the graph
 * copies to the device.  Results equal the eager ones bit for bit.  Off by
 * default (or TTR_GRAPHS=1 at ttr_create); ttr_graph_enable(h, 0) frees every
 * cached graph.  Calls with profiling on run eagerly.  At most 8 keys are
 * cached per handle (least recently used evicted); each holds its workspace. */
int ttr_graph_enable(ttr_handle_t h, int on);
/* Counters since ttr_create: graphs captured, replays, replays that ran eagerly
 * because their launches differed from the capture; graphs now cached. */
int ttr_graph_stats(ttr_handle_t h, int64_t* captures, int64_t* replays, int64_t* fallbacks, int64_t* cached);

/* Returns the handle's idle device memory to the runtime: the workspace arena
 * (kept at its largest size so repeated calls allocate nothing) and the freed
 * TT storage held for reuse.  Live TTs, operators and cached graphs are kept.
 * No counterpart in the reference (Julia's GC owns its buffers); for callers
 * that move from one workload to another of a different size, such as a large
 * batch followed by materialised products. */
int ttr_trim(ttr_handle_t h);

#ifdef __cplusplus
}
#endif

#endif /* TTROUND_H */
