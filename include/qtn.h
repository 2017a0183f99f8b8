/*
 * qtn.h -- C ABI of libqtn, the gfx950 (MI355X) contraction & truncation
 * library behind the Qaintensor.jl tensor-network API.
 *
 * Conventions shared by every entry point
 *   - Complex numbers are interleaved (re, im) doubles, bit-compatible with
 *     Julia ComplexF64, C `double _Complex` and hipDoubleComplex.  They are
 *     passed as `double*` pointing at 2*count doubles.
 *   - Arrays are column-major (Julia order): dims are given in Julia order,
 *     the first index varies fastest.
 *   - Sizes and dims are int64_t (Julia Int).  Leg numbers are 1-based.
 *   - Every function returns int32_t status: QTN_OK (0) or a negative code;
 *     qtn_last_error() then returns a message (thread-local).  Nothing throws
 *     or aborts across the ABI: a host allocation failure inside the library
 *     returns QTN_ERR_OOM, any other internal exception QTN_ERR_HIP.
 *   - Host pointers are only read during the call; the library keeps none.
 *     Entry points whose name ends in _dev take device pointers instead.
 *   - Calls are synchronous with respect to the host unless stated.
 *
 * Reference interfaces replaced (paths relative to Qaintensor.jl):
 *   qtn_contract_pair     <- TensorOperations.ncon pairwise step reached from
 *                            contract(net), src/contract.jl:242-264
 *   qtn_svd_trunc         <- LinearAlgebra.svd + truncation rule of
 *                            contract_svd, src/svd.jl:26-33
 *   qtn_svd               <- LinearAlgebra.svd of MPS(psi), MPO(m),
 *                            decompose!, switch! (src/mps.jl:147,157, ...)
 *   qtn_contract_svd      <- contract_svd, src/svd.jl:7-38
 *   qtn_mps_*             <- MPS container / switch! two-site split,
 *                            src/mps.jl:7-89, src/switch.jl:18-56
 *   qtn_mpo_* / zip-up    <- apply_MPO wiring + contract(),
 *                            src/mpo.jl:184-252
 */
#ifndef QTN_H
#define QTN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define QTN_OK              0
#define QTN_ERR_ARG        -1  /* invalid argument (dims, legs, cutoff, ...) */
#define QTN_ERR_HIP        -2  /* HIP runtime error                          */
#define QTN_ERR_OOM        -3  /* device or host allocation failed           */
#define QTN_ERR_NOCONV     -4  /* SVD sweep cap reached before convergence   */
#define QTN_ERR_TRUNC      -5  /* every singular value is within the cutoff  */
#define QTN_ERR_NOTFINITE  -6  /* NaN/Inf met in the input of an SVD         */

/* Jacobi block-pair tilings (qtn_tune_jacobi, qtn_set_jacobi_tiling) */
#define QTN_TILING_W32       0  /* 32x32 Gram blocks, three launches a round */
#define QTN_TILING_W32_FUSED 1  /* 32x32 Gram blocks, one fused launch        */
#define QTN_TILING_W64       2  /* 64x64 Gram blocks, three launches a round */

/* precision forms of the two-site update */
#define QTN_FORM_PLAIN      0  /* ComplexF64 end to end                       */
#define QTN_FORM_CHOSEN     1  /* complex FP32 bulk, FP64 core (see docs)     */
#define QTN_FORM_LEARNED    2  /* chosen + rank-hinted randomized SVD        */

typedef struct qtn_ctx qtn_ctx;
typedef struct qtn_mps qtn_mps;
typedef struct qtn_mpo qtn_mpo;

/* ---- context ----------------------------------------------------------- */
/* Opens device `device`, creates one HIP stream and a workspace pool.      */
int32_t qtn_ctx_create(int32_t device, qtn_ctx** ctx);
/* Destroy every qtn_mps, qtn_mpo and qtn_mlp made on ctx first:
 * their destroy calls use the context.                                    */
int32_t qtn_ctx_destroy(qtn_ctx* ctx);
/* Blocks until all work queued on the context's stream is done.            */
int32_t qtn_ctx_sync(qtn_ctx* ctx);
/* Returns the context's hipStream_t (as void*) so callers can time with
 * events on the stream the kernels run on.                                 */
int32_t qtn_ctx_stream(qtn_ctx* ctx, void** stream);
/* Work counters of the context since creation (or the last reset):
 * out[0] SVD problems solved, out[1] Jacobi sweeps summed over problems,
 * out[2] largest sweep count of one problem, out[3] batched SVD launches,
 * out[4] chosen-form single splits whose FP32-seeded Rayleigh-Ritz result
 * passed its FP64 check, out[5] chosen-form single splits redone by the
 * FP64 engine.  n = number of slots to write (<= 6).  reset != 0 zeroes
 * them after.                                                             */
int32_t qtn_ctx_counters(qtn_ctx* ctx, int64_t* out, int32_t n,
                         int32_t reset);
/* Per-kernel timing with HIP events on the context stream (off by default).
 * Kernel ids: 0 jac_inner, 1 jac_gram, 2 jac_update (the three-launch
 * block Jacobi round, FP64 sweeps), 3 jac_round (the opt-in fused round),
 * 4 jac_inner, 5 jac_gram, 6 jac_update, 7 jac_round (FP32 sweeps of the
 * chosen form; the fused round is their default from 320 to 512 rows).
 * Times come from start / stop events stamped by each dispatch.  qtn_ctx_profile(ctx, 1) enables and clears the accumulators.
 * qtn_ctx_profile_read returns, for one kernel id, the launch count, the
 * summed launch durations (ms), and the algorithmic HBM bytes and flops of
 * those launches (counted for the problems active in each launch).        */
int32_t qtn_ctx_profile(qtn_ctx* ctx, int32_t enable);
int32_t qtn_ctx_profile_read(qtn_ctx* ctx, int32_t kernel, int64_t* count,
                             double* ms, double* bytes, double* flops);
/* Message of the last failing call on this thread ("" if none).            */
const char* qtn_last_error(void);
/* Library version string. */
const char* qtn_version(void);

/* ---- pairwise contraction (the ncon step) ------------------------------ */
/* C = contraction of A and B over nc leg pairs (legsA[i] of A with
 * legsB[i] of B, 1-based).  The free legs of A (in order) followed by the
 * free legs of B (in order) form the natural output; `outperm` (1-based,
 * length ndA+ndB-2nc) lists, for each output axis, which natural axis it
 * is, exactly as permutedims(natural, outperm).  outperm may be NULL for
 * the identity.  Complex FP64.  Replaces one pairwise step of
 * TensorOperations.ncon as used by contract(), src/contract.jl:257,263.    */
int32_t qtn_contract_pair(qtn_ctx* ctx,
                          const double* A, int32_t ndA, const int64_t* dimsA,
                          const double* B, int32_t ndB, const int64_t* dimsB,
                          int32_t nc, const int32_t* legsA,
                          const int32_t* legsB, const int32_t* outperm,
                          double* C);
/* Same with device pointers; enqueued on the context stream, no sync.      */
int32_t qtn_contract_pair_dev(qtn_ctx* ctx,
                              const double* A, int32_t ndA,
                              const int64_t* dimsA, const double* B,
                              int32_t ndB, const int64_t* dimsB, int32_t nc,
                              const int32_t* legsA, const int32_t* legsB,
                              const int32_t* outperm, double* C);
/* qtn_contract_pair_dev that allocates nothing (no hipMalloc / hipFree or
 * allocating routine in the call): its scratch -- the permuted operands
 * and the product in natural leg order -- comes from one fixed 32 MiB
 * device array per GPU that the library holds until the process ends.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  A contraction whose scratch (each piece rounded to
 * 256 bytes) exceeds 32 MiB is refused with QTN_ERR_ARG.  Every call on a
 * GPU shares that GPU's array, so calls that overlap on one GPU (two
 * streams, two contexts) race; calls on one context's stream run in order.
 * Queued on the context's stream like qtn_contract_pair_dev.              */
int32_t qtn_contract_pair_dev_noalloc(qtn_ctx* ctx, const double* A,
                                      int32_t ndA, const int64_t* dimsA,
                                      const double* B, int32_t ndB,
                                      const int64_t* dimsB, int32_t nc,
                                      const int32_t* legsA,
                                      const int32_t* legsB,
                                      const int32_t* outperm, double* C);

/* Axis permutation of a column-major complex tensor:
 * out = permutedims(in, perm) (perm 1-based).  Replaces permutedims at
 * src/svd.jl:20-21, src/switch.jl:28-37, src/contract.jl:244.             */
int32_t qtn_permute(qtn_ctx* ctx, const double* in, int32_t nd,
                    const int64_t* dims, const int32_t* perm, double* out);
int32_t qtn_permute_dev(qtn_ctx* ctx, const double* in, int32_t nd,
                        const int64_t* dims, const int32_t* perm,
                        double* out);

/* Complex FP64 GEMM on device pointers, column-major:
 * C = alpha*op(A)*op(B) + beta*C, op = 'N', 'T' or 'C'.  Batched with
 * element strides (0 = shared operand).                                    */
int32_t qtn_zgemm_dev(qtn_ctx* ctx, char opA, char opB, int64_t m,
                      int64_t n, int64_t k, const double* alpha,
                      const double* A, int64_t lda, int64_t strideA,
                      const double* B, int64_t ldb, int64_t strideB,
                      const double* beta, double* C, int64_t ldc,
                      int64_t strideC, int64_t batch);
/* Complex FP32 variant (interleaved float pairs), FP32 accumulation.       */
int32_t qtn_cgemm_dev(qtn_ctx* ctx, char opA, char opB, int64_t m,
                      int64_t n, int64_t k, const float* alpha,
                      const float* A, int64_t lda, int64_t strideA,
                      const float* B, int64_t ldb, int64_t strideB,
                      const float* beta, float* C, int64_t ldc,
                      int64_t strideC, int64_t batch);

/* ---- truncated SVD ----------------------------------------------------- */
/* A (m x n, leading dim lda) = U S V^H.  Truncation follows
 * src/svd.jl:29-33: keep the fewest leading singular values whose
 * discarded tail has 2-norm <= er, where er = cutoff*||s||_2 when
 * `relative` is nonzero and er = cutoff otherwise; then k = min(k, chimax)
 * (chimax <= 0: no cap).  Outputs: *k, U (m x k, ld m), S (k), Vh (k x n,
 * ld k), *disc_weight = sum of squares of the discarded values.  Any of
 * U/S/Vh may be NULL.  Capacity: U, Vh and S must hold min(m,n) columns /
 * rows / values.  Returns QTN_ERR_TRUNC when every value is within er
 * (the reference errors there), QTN_ERR_ARG for cutoff < 0.                */
int32_t qtn_svd_trunc(qtn_ctx* ctx, int64_t m, int64_t n, const double* A,
                      int64_t lda, double cutoff, int32_t relative,
                      int64_t chimax, int64_t* k, double* U, double* S,
                      double* Vh, double* disc_weight);
/* Thin SVD without truncation: LinearAlgebra.svd as called at
 * src/mps.jl:147,157, src/mpo.jl:53,62, src/decompose.jl:26,38 and
 * src/switch.jl:77.  All k = min(m,n) values are returned, descending:
 * U (m x k, ld m), S (k), Vh (k x n, ld k); absorb = 1 returns diag(S)*V^H
 * in Vh (the product every one of those call sites forms next).  Any
 * output may be NULL.  An exactly zero singular value comes with a zero
 * singular vector (no orthonormal completion).                            */
int32_t qtn_svd(qtn_ctx* ctx, int64_t m, int64_t n, const double* A,
                int64_t lda, int32_t absorb, double* U, double* S,
                double* Vh);
/* Full singular value list (min(m,n) values, descending) of A; no vectors. */
int32_t qtn_svd_values(qtn_ctx* ctx, int64_t m, int64_t n, const double* A,
                       int64_t lda, double* S);

/* contract_svd (src/svd.jl:7-38): contract leg legA (1-based) of A with
 * leg legB of B after replacing each operand, matricised as
 * (other legs) x (leg) and (leg) x (other legs), by its SVD truncated at
 * the absolute tail norm er (src/svd.jl:29-33).  C has the free legs of A
 * then of B, in order, column-major.  Errors: "Error must be positive"
 * (er < 0), "Dimensions of contraction legs do not match", QTN_ERR_TRUNC
 * when an operand has every singular value within er.                    */
int32_t qtn_contract_svd(qtn_ctx* ctx, const double* A, int32_t ndA,
                         const int64_t* dimsA, const double* B, int32_t ndB,
                         const int64_t* dimsB, int32_t legA, int32_t legB,
                         double er, double* C);

/* ---- matrix-product states --------------------------------------------- */
/* An MPS of `nsites` qubits lives on the device in right-canonical form
 * (Vidal-free TEBD with the bond spectra kept beside the site tensors).
 * Site j tensor is (chi_{j-1}, 2, chi_j), column-major; chi_0 = chi_n = 1.
 * Created as |0...0>.  `chimax_cap` sizes the per-site storage.            */
int32_t qtn_mps_create(qtn_ctx* ctx, int32_t nsites, int64_t chimax_cap,
                       qtn_mps** mps);
int32_t qtn_mps_destroy(qtn_mps* mps);
/* Copy state a into b (same nsites and capacity).                          */
int32_t qtn_mps_copy(const qtn_mps* a, qtn_mps* b);
/* Replace the state by host site tensors of any gauge (dims[3*j..3*j+2] =
 * (chi_j, 2, chi_{j+1}) for 0-based site j, boundary bonds 1): an SVD sweep
 * left to right, then one right to left bring it to the right-canonical
 * form with bond spectra, normalised (only exact zeros are dropped, and
 * bonds beyond chimax_cap are truncated to it).  Replaces MPS construction
 * from tensors (src/mps.jl:99-113) for the device state.                  */
int32_t qtn_mps_set_sites(qtn_mps* mps, const int64_t* dims,
                          const double* const* sites);
/* Bond dimensions chi_0..chi_n (n+1 values).                               */
int32_t qtn_mps_bond_dims(const qtn_mps* mps, int64_t* dims);
/* Bond spectrum lambda_bond (bond 1..n-1, between site bond and bond+1).  */
int32_t qtn_mps_bond_spectrum(const qtn_mps* mps, int32_t bond, double* S,
                              int64_t* k);
/* Download site tensor (1-based) to host; dims[3] receives its shape.      */
int32_t qtn_mps_get_site(const qtn_mps* mps, int32_t site, double* out,
                         int64_t* dims);

/* Single-qubit gate u (2x2, column-major) on `site` (1-based); enqueued on
 * the context stream without waiting (later calls on the context follow it).  */
int32_t qtn_mps_apply_1q(qtn_mps* mps, int32_t site, const double* u);
/* Two-qubit gate on sites (site, site+1): g is 4x4 column-major, row/col
 * index = s_site + 2*s_{site+1} (wire w <-> bit w-1, as apply_MPO wires
 * gate bit k-1 to iwire[k], src/mpo.jl:199-214).  Truncation as
 * qtn_svd_trunc with relative cutoff and chimax.  Outputs (may be NULL):
 * kept rank, discarded weight (of the normalised spectrum), and the full
 * sorted spectrum before truncation (capacity 2*chimax_cap).              */
int32_t qtn_mps_apply_2q(qtn_mps* mps, int32_t site, const double* g,
                         double cutoff, int64_t chimax, int32_t form,
                         int64_t* kept, double* disc_weight,
                         double* spectrum, int64_t* spectrum_len);
/* switch!(mps, i, j) (src/switch.jl:63-86) on the device state: wires i
 * and j trade places.  As in the reference, wires are numbered from the
 * most significant (wire w is site n - w + 1, src/switch.jl:69-70) and the
 * exchange is a chain of adjacent swaps (src/switch.jl:18-56, each a
 * two-site update with the SWAP gate, untruncated but for exact zeros, up
 * to the state's capacity).  Errors (QTN_ERR_ARG) carry the reference's
 * messages: "Wire indices `i` and `j` must be positive", "Indices to swap
 * `i` and `j` must be less than or equal to the number of open wires in
 * MPS".                                                                   */
int32_t qtn_mps_switch(qtn_mps* mps, int32_t i, int32_t j);
/* Base.permute!(mps, order) (src/switch.jl:94-106): order (len = n
 * values, a permutation of 1..n) is realised by switch! calls as the
 * reference makes them; its checks and messages are the reference's.     */
int32_t qtn_mps_permute(qtn_mps* mps, const int64_t* order, int64_t len);
/* A two-qubit gate on any two sites i != j (1-based, the numbering of
 * qtn_mps_apply_2q): g is 4x4 column-major with index s_i + 2 s_j.  The
 * qubit of the farther site travels next to the nearer one by SWAP
 * updates, the gate is applied there and the qubit travels back; every
 * step is a two-site update truncated with cutoff / chimax in `form`.
 * kept / disc (may be NULL) are those of the gate's own update.  Replaces
 * applying a gate on non-adjacent wires through extend_MPO's identity
 * pipes and apply_MPO (src/mpo.jl:122-157, 184-252) for the device
 * state, with the swap network of switch! (src/switch.jl:63-86).          */
int32_t qtn_mps_apply_2q_long(qtn_mps* mps, int32_t i, int32_t j,
                              const double* g, double cutoff, int64_t chimax,
                              int32_t form, int64_t* kept,
                              double* disc_weight);
/* A layer of gates on disjoint neighbouring pairs, batched in one pass:
 * sites[i], gates[16*i..] as in apply_2q.  kept/disc per gate.
 * Failure: arguments and gate entries (NaN / Inf) are checked before any
 * gate is applied, so those errors leave the state untouched.  The gates
 * run in independent size groups; an error only the sweeps can detect
 * (QTN_ERR_NOCONV, QTN_ERR_TRUNC for a zero two-site tensor) is returned
 * after the other groups may have applied their gates, so after such an
 * error the state must be treated as undefined.  The same holds for
 * qtn_mps_apply_layer_multi and the learned / predicted layer calls.     */
int32_t qtn_mps_apply_layer(qtn_mps* mps, int32_t ngates,
                            const int32_t* sites, const double* gates,
                            double cutoff, int64_t chimax, int32_t form,
                            int64_t* kept, double* disc_weight);

/* A layer as qtn_mps_apply_layer in the learned form: gate i whose
 * predicted rank rank_hint[i] is positive and whose sketch
 * (rank_hint[i] + 16 columns) is narrower than chimax and at most 3/4 of
 * the full rank first takes the
 * randomized range finder (2 power iterations, complex FP64); its exact
 * discarded weight, |theta - Q Q^H theta|_F^2 summed entry by entry plus
 * the dropped sketch values^2, must meet the cutoff, else the gate falls
 * back to the full Jacobi SVD.  Gates
 * that skip the range finder are solved as in QTN_FORM_CHOSEN (FP32
 * sweeps when the cap binds, FP64 otherwise), so hints of 0 everywhere
 * give the chosen form.
 * accepted (may be NULL): 1 for the gates the range finder decided, -1
 * for those whose sketch failed the check (solved by the full SVD), 0 where
 * no sketch was tried.                                                    */
int32_t qtn_mps_apply_layer_learned(qtn_mps* mps, int32_t ngates,
                                    const int32_t* sites, const double* gates,
                                    double cutoff, int64_t chimax,
                                    const int64_t* rank_hint, int64_t* kept,
                                    double* disc_weight, int32_t* accepted);
/* Many states at once (the circuits of a batch): gate i acts on sites
 * (sites[i], sites[i]+1) of states[i] with the bond cap chimax[i] (<= 0:
 * that state's capacity); gates on one state must be disjoint, all states
 * share one context.  The gates are grouped by the padded size of their
 * two-site matrix and every group is one batched pass (contraction, gate,
 * Jacobi SVD, split), as in qtn_mps_apply_layer.  No counterpart in the
 * reference (it applies gates to one network at a time).                  */
int32_t qtn_mps_apply_layer_multi(int32_t ngates, qtn_mps* const* states,
                                  const int32_t* sites, const double* gates,
                                  const int64_t* chimax, double cutoff,
                                  int32_t form, int64_t* kept,
                                  double* disc_weight);
/* Single-qubit gates us[8*i..] (2x2 column-major) on (states[i],
 * sites[i]), one launch, enqueued on the states' context stream without a
 * wait (the next call that reads or rewrites the states orders after it).
 * The (state, site) pairs must be distinct.                                */
int32_t qtn_mps_apply_1q_multi(int32_t ngates, qtn_mps* const* states,
                               const int32_t* sites, const double* us);

/* Amplitudes <b|psi> for nb bitstrings; bits is nb x nsites bytes (0/1),
 * row-major per bitstring (bits[i*nsites + w-1] is wire w).               */
int32_t qtn_mps_amplitudes(const qtn_mps* mps, int64_t nb,
                           const uint8_t* bits, double* out);
/* <Z_w> for every wire (n doubles), normalised by <psi|psi>.              */
int32_t qtn_mps_expect_z(const qtn_mps* mps, double* out);
/* <a|b> (complex, 2 doubles).                                               */
int32_t qtn_mps_overlap(const qtn_mps* a, const qtn_mps* b, double* out);
/* <psi|psi> (real).                                                        */
int32_t qtn_mps_norm2(const qtn_mps* mps, double* out);

/* ---- whole networks (contract, src/contract.jl:242-264) ------------------ */
/* Contraction of a whole network on the device: the work of
 * TensorOperations.ncon(tensors, labels; order) that contract(net) calls
 * (src/contract.jl:257, 263).  ntensors host arrays (column-major complex
 * doubles, Julia's ComplexF64) with ranks ndims[t]; dims and labels hold
 * every tensor's leg sizes and ncon labels back to back (positive: a
 * contracted pair, in two tensors or twice in one for a trace; negative:
 * an open leg).  order lists every positive label once: the first label
 * not yet contracted names the next two tensors, contracted over all the
 * labels they share (pass the labels in increasing order for ncon's
 * default, or the sequence a planner such as contract_order,
 * src/contract.jl:184-235, returns -- applied as planned).  Tensors left
 * unconnected are joined by outer products.  out (out_len = the product of
 * the open legs' sizes) receives the result with legs ordered -1, -2, ...
 * Intermediates stay on the device; one copy in, one copy out.  Errors:
 * "Dimensions of contraction legs do not match" (src/svd.jl:16) and
 * QTN_ERR_ARG for labels or an order that do not describe a network.     */
int32_t qtn_contract_network(qtn_ctx* ctx, int32_t ntensors,
                             const int32_t* ndims, const int64_t* dims,
                             const int32_t* labels,
                             const double* const* tensors, int32_t norder,
                             const int32_t* order, double* out,
                             int64_t out_len);

/* ---- matrix-product operators (apply_MPO, src/mpo.jl:184-252) ----------- */
/* MPO of nsites tensors W_j (D_j, 2 out, 2 in, D_{j+1}), column-major host
 * arrays, D_0 = D_n = 1 (the middle-tensor layout of MPO(m),
 * src/mpo.jl:66); (W psi)[s'] = sum_s W[s', s] psi[s], wire j <-> bit j-1. */
int32_t qtn_mpo_create(qtn_ctx* ctx, int32_t nsites, const int64_t* bond_dims,
                       const double* const* sites, qtn_mpo** mpo);
int32_t qtn_mpo_destroy(qtn_mpo* mpo);
/* <psi|W|psi> (complex, 2 doubles) by a left-to-right environment sweep on
 * MFMA GEMMs; psi is normalised in the device gauge.                      */
int32_t qtn_mps_expect_mpo(const qtn_mps* mps, const qtn_mpo* mpo,
                           double* out);
/* The same sweep in a precision form: QTN_FORM_PLAIN (= qtn_mps_expect_mpo,
 * complex FP64 throughout) or QTN_FORM_CHOSEN, where each site's two
 * bond-sized contractions run as BF16x3 split-precision GEMMs (three bf16
 * terms per operand, six products, FP32 accumulation on the bf16 matrix
 * cores) once all of their dimensions reach 512; smaller ones stay FP64.  */
int32_t qtn_mps_expect_mpo_form(const qtn_mps* mps, const qtn_mpo* mpo,
                                int32_t form, double* out);
/* C (m x n) = op(A) B in BF16x3 split precision (as above) on host
 * complex FP64 arrays, column-major; opA 'N' (A is m x k) or 'C' (A is
 * k x m, op = conjugate transpose).  No counterpart in the reference.     */
int32_t qtn_zgemm_bf16x3(qtn_ctx* ctx, char opA, int64_t m, int64_t n,
                         int64_t k, const double* A, int64_t lda,
                         const double* B, int64_t ldb, double* C,
                         int64_t ldc);
/* out <- W|in> compressed: zip-up left to right (truncation per bond with
 * the relative cutoff and chimax), then a right-to-left SVD sweep that
 * recompresses with the same rule into out's right-canonical gauge.
 * *norm receives |W in| of the compressed state (out is normalised).
 * Replaces contract(apply_MPO(psi, mpo, iwire)) for chains too long to
 * contract densely.                                                        */
int32_t qtn_mps_apply_mpo(const qtn_mps* in, const qtn_mpo* mpo,
                          double cutoff, int64_t chimax, qtn_mps* out,
                          double* norm);
/* qtn_mps_apply_mpo in a chosen form.  QTN_FORM_CHOSEN: each zip-up split
 * whose rank the cap can decide (min(rows, cols) > its cap) runs its Jacobi
 * sweeps in complex FP32; when FP32 keeps exactly the cap, the leading
 * cap + 64 left vectors are re-orthonormalised in FP64 (three Newton-Schulz
 * steps) and A's projection on them is split in FP64 (Rayleigh-Ritz),
 * else the split is redone in FP64.  The right-to-left recompression stays
 * FP64.  No counterpart in the reference (it has the plain form only).    */
int32_t qtn_mps_apply_mpo_form(const qtn_mps* in, const qtn_mpo* mpo,
                               double cutoff, int64_t chimax, int32_t form,
                               qtn_mps* out, double* norm);

/* ---- learned rank predictor (no counterpart in the reference) ---------- */
/* A 3-layer FP32 MLP, y = W3 gelu(W2 gelu(W1 a + b1) + b2) + b3, whose
 * inputs are columns of nin features (the first min(64, nin) enter as
 * log10(max(x, 1e-16)) / 8).  params: W1 (width x nin), b1 (width), W2 (width x
 * width), b2 (width), W3 (1 x width), b3 (1), column-major, concatenated.
 * Forward, backward (mean squared error) and Adam run in the library's own
 * kernels on the GPU.                                                      */
typedef struct qtn_mlp qtn_mlp;
int32_t qtn_mlp_create(qtn_ctx* ctx, int32_t nin, int32_t width,
                       const float* params, qtn_mlp** mlp);
int32_t qtn_mlp_destroy(qtn_mlp* mlp);
int32_t qtn_mlp_get_params(const qtn_mlp* mlp, float* params);
/* Y (nb) = mlp(X), X nin x nb column-major (host).                         */
int32_t qtn_mlp_forward(qtn_mlp* mlp, int64_t nb, const float* X, float* Y);
/* loss = mean (y - T)^2 and its gradient (same layout as params).          */
int32_t qtn_mlp_loss_grad(qtn_mlp* mlp, int64_t nb, const float* X,
                          const float* T, float* loss, float* grad);
/* `steps` Adam steps (beta 0.9 / 0.999) over contiguous minibatches of the
 * (pre-shuffled) training set; loss_hist (steps floats, may be NULL).       */
int32_t qtn_mlp_train(qtn_mlp* mlp, int64_t nb, const float* X,
                      const float* T, int32_t steps, int64_t minibatch,
                      float lr, float* loss_hist);

/* Operator-Schmidt coefficients of ngates 4x4 gates (gates[32*i..] as in
 * qtn_mps_apply_2q): the singular values, descending, of each gate split
 * into (s_i, s_i') x (s_{i+1}, s_{i+1}') factors, out[4*i..].  The split is
 * the one of src/mpo.jl:40-56 (MPO(m) of a two-site operator).            */
int32_t qtn_gate_schmidt(qtn_ctx* ctx, int32_t ngates, const double* gates,
                         double* out);
/* Autotuner of the Jacobi SVD engine.  Times each block-pair tiling on
 * `batch` synthetic complex Gaussian rows x cols problems (FP64 sweeps, or
 * FP32 sweeps when single != 0) with HIP events, best of two runs after a
 * warm-up, and keeps the fastest for this context, column bucket (cols
 * rounded up to a multiple of 64) and batch class (floor(log2(batch)),
 * capped at 6): every later batched SVD of that bucket whose number of
 * solved problems falls in the class uses it.  best receives the QTN_TILING_* code; us (3
 * doubles, may be NULL) the microseconds per SVD of W32, W32_FUSED and
 * W64 (-1 where a tiling does not apply).  No counterpart in the
 * reference.                                                              */
int32_t qtn_tune_jacobi(qtn_ctx* ctx, int64_t rows, int64_t cols,
                        int32_t batch, int32_t single, int32_t* best,
                        double* us);
/* Sets the tiling of a bucket and batch class (e.g. from a saved table of
 * qtn_tune_jacobi).                                                       */
int32_t qtn_set_jacobi_tiling(qtn_ctx* ctx, int64_t cols, int32_t single,
                              int32_t batch, int32_t tiling);

/* The predictor's input for the gates on (sites[i], sites[i]+1) of a state:
 * out (72 x ngates floats, column-major, host) holds per gate the bond
 * spectrum before the gate (top 64 values, zero padded), the gate's four
 * operator-Schmidt coefficients schmidt[4*i..] (as given), the bond
 * position sites[i]/n and chi_{i-1}, chi_i, chi_{i+1} / chimax (chimax
 * <= 0: the state's capacity).  No counterpart in the reference.          */
int32_t qtn_mps_gate_features(qtn_mps* mps, int32_t ngates,
                              const int32_t* sites, const double* schmidt,
                              int64_t chimax, float* out);
/* A layer in the learned form with the predictor in the library: the
 * features above are gathered on the device, the MLP (nin 72, same
 * context) predicts y per gate, and rank_hint = ceil((y + margin) chimax)
 * clipped to [0, chimax] drives qtn_mps_apply_layer_learned.  hints (may
 * be NULL) receives the hints.  No counterpart in the reference.           */
int32_t qtn_mps_apply_layer_predicted(qtn_mps* mps, int32_t ngates,
                                      const int32_t* sites,
                                      const double* gates,
                                      const double* schmidt, double cutoff,
                                      int64_t chimax, qtn_mlp* mlp,
                                      double margin, int64_t* kept,
                                      double* disc_weight, int32_t* accepted,
                                      int64_t* hints);

#ifdef __cplusplus
}
#endif
#endif /* QTN_H */
