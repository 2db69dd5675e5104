/*
 * mm_hermite.h -- C ABI of the MI355X Bargmann->Fock library (libmmhip.so).
 *
 * Plain pointers and sizes only.  Complex numbers are interleaved
 * (re, im) float64 pairs, i.e. complex128.  Every array pointer is a DEVICE
 * pointer unless the comment says "host".  `stream` is a hipStream_t passed
 * as void* (NULL = the default stream).  All entry points are stream-ordered
 * and never synchronise the host, except the mm_plan_* set-up calls.
 *
 * Return value: MM_OK (0) or a negative status; mm_last_error() then gives a
 * thread-local message.  No entry point frees caller memory.  Workspace
 * buffers are caller-provided: query the size first, pass at least that many
 * bytes (16-byte aligned).
 */
#ifndef MM_HERMITE_H
#define MM_HERMITE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MM_OK 0
#define MM_ERR_INVALID (-1)     /* bad shape, D, sizes, null pointer            */
#define MM_ERR_HIP (-2)         /* a HIP runtime call failed                     */
#define MM_ERR_WORKSPACE (-3)   /* workspace missing or smaller than required    */
#define MM_ERR_UNSUPPORTED (-4) /* valid input outside what the kernels support  */

#define MM_MAX_DIMS 8        /* D <= 8 indices per triple                       */
#define MM_MAX_CUTOFF 100000 /* each shape entry < 100000: the reference's SQRT
                                table length, core.py:22                        */

/* Message for the last failing call on this thread (never NULL). */
const char *mm_last_error(void);

/* Library version string. */
const char *mm_version(void);

/* Upload the per-device constant tables (sqrt(n) and 1/sqrt(n), n < 100000)
 * for `device`.  Called implicitly by every entry point; idempotent. */
int mm_init(int device);

/* ------------------------------------------------------------------ forward */

/* Workspace bytes needed by mm_hermite_fwd for this shape (0 for most). */
size_t mm_hermite_fwd_workspace(const int64_t *shape /* host, D entries */, int D, int stable);

/* G = renormalized multidimensional Hermite tensor of (A, b, c) on `shape`,
 * C order (last index fastest).  stable = 0: first-pivot recurrence;
 * stable = 1: average over all pivots.
 * Replaces vanilla_numba / stable_numba
 *   (mrmustard/math/lattice/strategies/vanilla/core.py:25-124, :127-213)
 * as bound by BackendNumpy.hermite_renormalized (math/backend_numpy.py:381-392).
 * A: D*D complex (row-major), b: D complex, c: 1 complex, G: prod(shape) complex. */
int mm_hermite_fwd(const double *A, const double *b, const double *c, const int64_t *shape, int D,
                   int stable, double *G, void *workspace, size_t ws_bytes, void *stream);

/* Uniform batch: B triples sharing one shape.  A: B*D*D, b: B*D, c: B,
 * G: B*prod(shape) complex.  Replaces vanilla_batch_numba
 *   (vanilla/batch.py:27-61) as bound by BackendNumpy.hermite_renormalized_batched
 *   (math/backend_numpy.py:394-403). */
size_t mm_hermite_batch_fwd_workspace(int64_t B, const int64_t *shape, int D, int stable);
int mm_hermite_batch_fwd(const double *A, const double *b, const double *c, int64_t B,
                         const int64_t *shape, int D, int stable, double *G, void *workspace,
                         size_t ws_bytes, void *stream);

/* ---------------------------------------------------------------------- VJP */

/* Holomorphic vector-Jacobian product of the vanilla forward (no conjugation
 * of dLdG): dA (D*D, symmetrised (M + M^T)/2), db (D), dc (1).
 * Replaces vanilla_vjp_numba (vanilla/gradients.py:25-82), the backward of
 * hermite_renormalized_jax (math/jax_vjps/hermite.py:86-102). */
size_t mm_hermite_vjp_workspace(const int64_t *shape, int D);
int mm_hermite_vjp(const double *G, const double *c, const double *dLdG, const int64_t *shape, int D,
                   double *dA, double *db, double *dc, void *workspace, size_t ws_bytes, void *stream);

/* Uniform-batch VJP.  Replaces vanilla_batch_vjp_numba (vanilla/gradients.py:85-116). */
size_t mm_hermite_batch_vjp_workspace(int64_t B, const int64_t *shape, int D);
int mm_hermite_batch_vjp(const double *G, const double *c, const double *dLdG, int64_t B,
                         const int64_t *shape, int D, double *dA, double *db, double *dc,
                         void *workspace, size_t ws_bytes, void *stream);

/* mm_hermite_batch_vjp without a workspace argument and without any device
 * allocation: the partial sums live in a static device buffer of
 * MM_STATIC_SCRATCH_BYTES per GPU, held until the process ends.  No
 * counterpart in the reference: for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops).  mm_init(device) must have run on
 * the current device first (it looks up the buffer's address); returns
 * MM_ERR_UNSUPPORTED when B triples of this shape need more scratch than the
 * buffer holds (mm_hermite_batch_vjp_workspace(B, shape, D) > the maximum).
 * Every call on one GPU shares that GPU's buffer: calls that overlap on one
 * GPU (different streams or threads) race and must be ordered by the caller.
 * Results equal mm_hermite_batch_vjp's bit for bit. */
#define MM_STATIC_SCRATCH_BYTES (8 << 20)
int mm_hermite_batch_vjp_static(const double *G, const double *c, const double *dLdG, int64_t B,
                                const int64_t *shape, int D, double *dA, double *db, double *dc,
                                void *stream);

/* Fused forward + VJP in one sweep when dL/dG is known up front: the
 * composition vanilla_numba -> vanilla_vjp_numba (core.py:25, gradients.py:25). */
size_t mm_hermite_fused_workspace(const int64_t *shape, int D);
int mm_hermite_fused(const double *A, const double *b, const double *c, const double *dLdG,
                     const int64_t *shape, int D, double *G, double *dA, double *db, double *dc,
                     void *workspace, size_t ws_bytes, void *stream);

/* Uniform-batch fused forward + VJP: the composition vanilla_batch_numba ->
 * vanilla_batch_vjp_numba (batch.py:27, gradients.py:85) in one sweep per triple. */
size_t mm_hermite_batch_fused_workspace(int64_t B, const int64_t *shape, int D);
int mm_hermite_batch_fused(const double *A, const double *b, const double *c, const double *dLdG,
                           int64_t B, const int64_t *shape, int D, double *G, double *dA,
                           double *db, double *dc, void *workspace, size_t ws_bytes, void *stream);

/* VJP with recomputation: (dA, db, dc) from (A, b, c, dLdG) without G.  G is
 * recomputed slab by slab (axis-0 slabs) in a window of three slabs held in
 * the workspace, so resident memory is 3 * prod(shape[1:]) complex values
 * instead of prod(shape).  Lattices of at most 64 MB (4 Mi amplitudes), and
 * those whose slabs fit in LDS, are computed whole in the workspace instead
 * (mm_hermite_vjp_recompute_workspace says how much).  Same results as
 * mm_hermite_vjp applied to mm_hermite_fwd's G (vanilla_vjp_numba,
 * vanilla/gradients.py:25-82, of vanilla_numba's output, core.py:25-124)
 * while every value stays finite.  An
 * inf or NaN in G or dLdG makes dc non-finite (dc is the sum of G * dLdG
 * divided by c); the other entries may then differ from the reference's in
 * which parts are NaN or +-inf.  A caller that needs the reference's values
 * there calls mm_hermite_fwd and mm_hermite_vjp when dc is not finite. */
size_t mm_hermite_vjp_recompute_workspace(const int64_t *shape, int D);
int mm_hermite_vjp_recompute(const double *A, const double *b, const double *c, const double *dLdG,
                             const int64_t *shape, int D, double *dA, double *db, double *dc,
                             void *workspace, size_t ws_bytes, void *stream);

/* ----------------------------------------------------------------- binomial
 *
 * The lattice filled by total photon number w = sum_j k_j, stopping after the
 * first w at which |c|^2 + the squared norms of weights 1..w exceed max_l2, or
 * at w = global_cutoff - 1; entries of larger weight are zero.  Values equal
 * vanilla's.  Replaces binomial (mrmustard/math/lattice/strategies/binomial.py
 * :30-74) as bound by BackendNumpy.hermite_renormalized_binomial
 * (math/backend_numpy.py:405-421).  norm (device, 1 double, may be NULL)
 * receives the accumulated squared norm, binomial's second result.  The sum of
 * (shape - 1) + 1 must be at most 8192.  The squared norms of a weight are
 * summed in parallel (as numba's prange does), so a norm within rounding of
 * max_l2 can stop one weight apart from a sequential sum. */
size_t mm_hermite_binomial_workspace(const int64_t *shape, int D);
int mm_hermite_binomial(const double *A, const double *b, const double *c, const int64_t *shape, int D, double max_l2,
                        int64_t global_cutoff, double *G, double *norm, void *workspace, size_t ws_bytes, void *stream);

/* ---------------------------------------------------------------- autoshape
 *
 * Fock cutoffs of B Gaussian density matrices of M <= 8 modes: for every mode
 * the least k whose diagonal probability sum reaches max_prob (early-stopped
 * diagonal recursion of the single-mode reduced state), clipped to
 * [min_shape, max_shape].  Replaces autoshape_numba
 * (mrmustard/math/lattice/autoshape.py:23-154), called by State.auto_shape
 * (mrmustard/lab/states/base.py:427-434).  A: B x (2M x 2M), indices
 * (bra modes, ket modes); b: B x 2M; c: B; shape: B x M int64 (device).  A
 * mode whose complement matrix is singular gets -1 (the reference raises).
 * The workspace holds each (state, mode)'s single-mode marginal between the
 * two launches (mm_autoshape_workspace bytes). */
size_t mm_autoshape_workspace(int64_t B, int M);
int mm_autoshape(const double *A, const double *b, const double *c, int64_t B, int M, double max_prob, int max_shape,
                 int min_shape, int64_t *shape, void *workspace, size_t ws_bytes, void *stream);

/* ----------------------------------------------------------------- diagonal
 *
 * The photon-number diagonal G[k_0, k_0, k_1, k_1, ...] of the 2M-index
 * lattice of (A, B, c) for M <= 10 modes, without the full lattice.  Replaces
 * hermite_multidimensional_diagonal (mrmustard/math/lattice/strategies/
 * compactFock/inputValidation.py:61-79 -> diagonal_amps.py:200-248) as bound
 * by BackendNumpy.hermite_renormalized_diagonal (math/backend_numpy.py:423-432).
 * A: 2M x 2M; B: 2M x nb (nb = 1 unbatched, else the reference's batch of
 * B's); c: one value, or nb values if c_batched; G: cutoffs x nb (device,
 * complex128, written in full).  reordered != 0 takes A and B in the
 * [modes..., modes...] order and applies reorder_AB_bargmann
 * (backend_numpy.py:367-375) on the fly; 0 takes them already in the
 * [mode0, mode0, mode1, mode1, ...] order.  The workspace holds the four
 * auxiliary sub-lattices, about (4M - 1) x prod(cutoffs) x nb complex128.
 * The inner products are summed in index order, so results differ from the
 * reference's BLAS dot at rounding level. */
size_t mm_hermite_diagonal_workspace(const int64_t *cutoffs, int M, int64_t nb);
int mm_hermite_diagonal(const double *A, const double *B, const double *c, int c_batched, const int64_t *cutoffs, int M,
                        int64_t nb, int reordered, double *G, void *workspace, size_t ws_bytes, void *stream);

/* Forward and vector-Jacobian product of the diagonal strategy in one call
 * (unbatched B): G as mm_hermite_diagonal, then dA (2M x 2M), dB (2M), dc (1)
 * = sum_k dLdG[k] dG[k]/d(A, B, c), holomorphic (dLdG not conjugated), every
 * entry of A independent (not symmetrised), in the caller's index order.
 * Replaces fock_representation_diagonal_grad (compactFock/diagonal_grad.py
 * :19-38, :41-258, :357-384) contracted with the cotangent as in
 * hermite_renormalized_diagonal_jax_bwd (jax_vjps/hermite.py:292-330); that
 * forward-mode Jacobian of prod(cutoffs) x (2M)^2 entries is never formed.
 * Sums are accumulated in a fixed order (deterministic run to run). */
size_t mm_hermite_diagonal_vjp_workspace(const int64_t *cutoffs, int M);
int mm_hermite_diagonal_vjp(const double *A, const double *B, const double *c, const int64_t *cutoffs, int M, int reordered,
                            const double *dLdG, double *G, double *dA, double *dB, double *dc, void *workspace, size_t ws_bytes,
                            void *stream);

/* -------------------------------------------------------- one leftover mode
 *
 * The conditional density matrices of mode 0 (output_cutoff <= 1023) when the
 * Mt <= 7 other modes are PNR-detected: out[m, n, k_1..k_Mt] = G[(m, k), (n, k)] for m, n <=
 * output_cutoff and k_j <= pnr_cutoffs[j], shape (output_cutoff + 1,
 * output_cutoff + 1, pnr_cutoffs + 1) (device, complex128, written in full).
 * Replaces hermite_multidimensional_1leftoverMode (mrmustard/math/lattice/
 * strategies/compactFock/inputValidation.py:103-118 -> singleLeftoverMode_amps
 * .py:369-421) as bound by the JAX backend (math/backend_jax.py:524-534,
 * jax_vjps/hermite.py:343-387); the numpy backend's fast_diagonal
 * (math/backend_numpy.py:434-446, strategies/fast_diagonal.py:32-80) returns
 * the same amplitudes.  A (2L x 2L, L = Mt + 1), b (2L), c (1) in the
 * [modes..., modes...] order when reordered != 0, else [mode0, mode0, mode1,
 * mode1, ...].  Inner products are summed in index order (rounding-level
 * differences from the reference's BLAS dot). */
size_t mm_hermite_1leftover_workspace(int64_t output_cutoff, const int64_t *pnr_cutoffs, int Mt);
int mm_hermite_1leftover(const double *A, const double *b, const double *c, int64_t output_cutoff, const int64_t *pnr_cutoffs,
                         int Mt, int reordered, double *out, void *workspace, size_t ws_bytes, void *stream);

/* Vector-Jacobian product of mm_hermite_1leftover: runs the forward (out is
 * written as by mm_hermite_1leftover), then dA = sum dLdout * d out / dA, db,
 * dc likewise, holomorphic (no conjugation), A's entries independent, dA and
 * db in the caller's order.  Replaces grad_hermite_multidimensional_1leftoverMode
 * (compactFock/inputValidation.py:121-141 -> singleLeftoverMode_grad.py:19-46,
 * :50-225, :254-556, :560-692) with its contraction in
 * hermite_renormalized_1leftoverMode_jax_bwd (jax_vjps/hermite.py:397-440),
 * computed in reverse mode instead of the reference's Jacobian arrays. */
size_t mm_hermite_1leftover_vjp_workspace(int64_t output_cutoff, const int64_t *pnr_cutoffs, int Mt);
int mm_hermite_1leftover_vjp(const double *A, const double *b, const double *c, int64_t output_cutoff, const int64_t *pnr_cutoffs,
                             int Mt, int reordered, const double *dLdout, double *out, double *dA, double *db, double *dc,
                             void *workspace, size_t ws_bytes, void *stream);

/* The Fock array of a triple with derived variables, the lattice never
 * stored whole: out[h] = sum_k G[h, k] w[k], G the lattice of (A, b, c) on
 * `shape` (D axes), h its first n_head axes, k the other D - n_head axes
 * flattened (K = prod(shape[n_head:]) entries of w; out has prod(shape[:n_head])
 * entries).  Replaces CircuitComponent.fock_array's num_derived_vars > 0
 * branch (mrmustard/lab/circuit_components.py:516-528: the lattice of
 * (A, b, 1) on shape + shape_derived_vars, then math.einsum with c over the
 * derived indices) for one triple: the lattice is computed one level-0 slab
 * at a time in a three-slab window (3 * prod(shape[1:]) * 16 bytes of
 * workspace) and each slab is contracted as soon as it exists.  Lattices that
 * fit in LDS are computed whole.  The values of G are those of
 * mm_hermite_fwd; the sums over k are in a fixed order of their own. */
size_t mm_hermite_fwd_contract_workspace(const int64_t *shape, int D);
int mm_hermite_fwd_contract(const double *A, const double *b, const double *c, const int64_t *shape, int D, int n_head,
                            const double *w, double *out, void *workspace, size_t ws_bytes, void *stream);

/* The same for B triples sharing `shape`, contracted with a K x M matrix:
 * out[t] (prod(shape[:n_head]) x M, row-major) = G_t (rows x K) . W_t with
 * W_t = W + t * sW complex elements (K x M row-major; sW = 0 shares one
 * matrix).  Covers a Bargmann state converted to Fock and at once contracted
 * with a Fock array on its trailing wires -- CircuitComponent.contract's
 * to_fock of the Bargmann side followed by ArrayAnsatz.contract
 * (mrmustard/lab/circuit_components.py:445-458,
 * mrmustard/physics/ansatz/array_ansatz.py:159-225, mm_einsum.py:272-292) --
 * and the batched derived variables of fock_array (circuit_components.py:
 * 516-528, M = 1, one w per triple).  Each level-0 slab is contracted as it
 * is made (M = 1: one dot product per row; M > 1: the matrix cores, as
 * mm_zgemm), so the workspace is that of one mm_hermite_fwd_contract call for
 * any B.  A, b, c are packed per triple (D*D, D, 1 complex entries). */
size_t mm_hermite_batch_fwd_contract_workspace(const int64_t *shape, int D);
int mm_hermite_batch_fwd_contract(const double *A, const double *b, const double *c, int64_t B, const int64_t *shape, int D,
                                  int n_head, const double *W, int64_t M, int64_t sW, double *out, void *workspace,
                                  size_t ws_bytes, void *stream);

/* ------------------------------------------------------ contraction (MFMA)
 *
 * C = A * B for complex128 row-major matrices (A: M x K with row stride lda,
 * B: K x N, ldb, C: M x N, ldc; strides in complex elements), on the FP64
 * matrix cores.  The consumer of the amplitudes: a Fock-basis tensor reshaped
 * to a matrix and contracted with another (ArrayAnsatz.contract ->
 * BackendManager.einsum, mrmustard/physics/ansatz/array_ansatz.py:159+,
 * mrmustard/math/backend_manager.py:526-548).  Sums over K in chunks of 16
 * with the instruction's own accumulation order, so results differ from a
 * sequential sum at rounding level. */
int mm_zgemm(const double *A, int64_t lda, const double *B, int64_t ldb, double *C, int64_t ldc, int64_t M, int64_t N,
             int64_t K, void *stream);

/* A batch of products C_b = A_b * B_b (b < batch) with element strides sA,
 * sB, sC between consecutive products (0 broadcasts an operand): the batched
 * contraction that ArrayAnsatz.contract's einsum becomes once its labels are
 * grouped into (batch, rows, contracted, columns) (array_ansatz.py:159-225,
 * backend_manager.py:526-548).  Same summation as mm_zgemm. */
int mm_zgemm_batched(const double *A, int64_t lda, int64_t sA, const double *B, int64_t ldb, int64_t sB, double *C, int64_t ldc,
                     int64_t sC, int64_t M, int64_t N, int64_t K, int64_t batch, void *stream);

/* ------------------------------------------------- work-queue batch (plans)
 *
 * A plan describes n triples of uneven mode count D_k <= 8 and shape_k.
 * Packed layouts (complex elements, triples in the order given):
 *   A, dA: concatenation of D_k*D_k blocks;  b, db: of D_k blocks;
 *   c, dc: n entries;                         G, dLdG: of prod(shape_k) blocks.
 * mm_plan_create sorts the triples largest-first into a device work queue
 * consumed by one persistent kernel (blocks take large triples, single waves
 * take small ones, single lanes take one-index chains).  It allocates the
 * plan's own device tables and synchronises the device once.
 * The reference has no uneven batch: its counterpart is a loop of
 * vanilla_numba calls, or vanilla_batch_numba (batch.py:27) when all shapes
 * agree.  mm_plan_run launches on `stream`; it may be captured in a graph.
 */
typedef struct mm_plan mm_plan;

typedef struct {
    int block_threads;   /* threads per workgroup: 256, 512 or 1024 (0 = auto)  */
    int blocks_per_cu;   /* persistent workgroups per compute unit (0 = auto)    */
    int wave_max_slab;   /* largest axis-0 slab (prod(shape[1:])) given to a single
                            wave instead of a workgroup (0 = auto, -1 = never)    */
    int reserved;
} mm_tune;

int mm_plan_create(int64_t n, const int32_t *dims /* host, n */, const int64_t *shapes /* host, n*8 */,
                   int fused, const mm_tune *tune /* may be NULL */, mm_plan **plan);
int mm_plan_destroy(mm_plan *plan);
/* Totals of the plan (host): number of amplitudes and of complex elements in A and b. */
int mm_plan_info(const mm_plan *plan, int64_t *n_amplitudes, int64_t *a_elems, int64_t *b_elems);
/* How the plan splits its triples (host, 8 entries): workgroup triples, wave
 * triples, lane triples, triples run outside the persistent kernel, grid,
 * threads per workgroup, dynamic LDS bytes per workgroup, LDS bytes per wave. */
int mm_plan_describe(const mm_plan *plan, int64_t *out);
/* dLdG, dA, db, dc must be NULL for a plan created with fused = 0. */
int mm_plan_run(const mm_plan *plan, const double *A, const double *b, const double *c,
                const double *dLdG, double *G, double *dA, double *db, double *dc, void *stream);
/* As mm_plan_run, but through a HIP graph that the plan captures on first use
 * (and re-captures when the buffer pointers change). */
int mm_plan_run_graph(mm_plan *plan, const double *A, const double *b, const double *c,
                      const double *dLdG, double *G, double *dA, double *db, double *dc, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MM_HERMITE_H */
