/*
 * draugr_hip.h — C-ABI of the MI355X (gfx950) AMG solve / resetup backend.
 *
 * Part 1 mirrors the reference C API one to one: same symbol names,
 * argument meaning, return codes and error convention as
 * c_api/draugr_amg.h (implemented there by src/cfunction_interface.jl).
 * A host that links libdraugr.so can link libdraugr_hip.so instead; no Julia
 * runtime is involved, so init_julia/shutdown_julia keep only their place in
 * the lifecycle (see their declarations).
 *
 * Part 2 adds entry points the reference does not have: device-pointer
 * (HBM-resident) vectors with an explicit HIP stream, the FP64 PCG driver,
 * resetup from device values, and inspection / per-kernel hooks used by the
 * parity tests.  Plain pointers and sizes only; a stream is passed as void*
 * (a hipStream_t; NULL is the default stream).
 *
 * Errors: every call returns -1 on failure and stores a message retrievable
 * with draugr_amg_last_error().  The message is per calling thread and valid
 * until that thread's next draugr_* call.
 *
 * Threading: the handle tables are protected by a lock; calls on one
 * hierarchy handle are serialised by a per-hierarchy lock.
 */
#ifndef DRAUGR_HIP_H
#define DRAUGR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ── Part 1: reference C API ─────────────────────────────────────────────── */

/* Runtime lifecycle.  Replace c_api/draugr_amg.h:26-36.  There is no runtime
 * to start: init_julia ignores its arguments and does nothing, and the other
 * entry points do not need it (a host may still call it first, as the
 * reference requires).  shutdown_julia releases every hierarchy, batch and
 * config handle still open (their device memory included), as the end of the
 * reference's runtime would; retcode is ignored and later calls with those
 * handles fail with "Invalid handle". */
void init_julia(int argc, char **argv);
void shutdown_julia(int retcode);

/* Last error message ("" if none).  Replaces c_api/draugr_amg.h:46. */
const char *draugr_amg_last_error(void);

/* Config from JSON; all keys optional, unknown keys ignored, defaults of
 * AMGConfig() (src/types.jl:514-536).  Keys of c_api/draugr_amg.h:58-89 plus
 * "precision" ("fp64" | "fp32" | "bf16": arithmetic of the V/W cycle),
 * "seed" (setup RNG seed), "graphs" (true: the PCG / mixed-precision preconditioner cycle, and
 * each batched PCG iteration, is captured once into a hipGraph and replayed; false: launched
 * eagerly; "auto", the default: graphs for batches, eager launches for one system),
 * "kernel" ("auto" | "stream" | "rowseq" | "pipe" | "stream16" | "sorted" | "stream16pipe" |
 * "sortedpipe" | "sortedquad" | "vector" (lanes per row by row length; FP32 / BF16
 * cycles only, FP64 passes take "stream") | "pattern" (operators whose rows repeat at most
 * 64 column patterns relative to the row, as a grid stencil does: a 16-bit pattern index per
 * row instead of a column per entry; other operators take "stream")), "fuse" ("auto"
 * | "off" | "on": a level's first smoothing step from zero fused with its
 * residual, that step, residual and restriction in one pass, and its
 * prolongation and first post-smoothing step in one pass; Jacobi-type
 * smoothers; "auto" times each on the level's operators), "gpu_setup" (default
 * true: when the hierarchy is uploaded, each level's extended+i interpolation,
 * coarse operator, triple map, transpose map and the PMIS sweeps of its C/F
 * splitting run on the GPU, identical to the host's, and the triple map stays
 * in HBM; levels of fewer than "gpu_setup_min_rows" rows (default 32768)
 * interpolate and sweep on the host),
 * "gs_symmetric" (coloured GS post-smoothing sweeps the colours in reverse
 * order: a symmetric cycle, which PCG requires of GS), "batch_rhs_per_thread"
 * (0 = autotuned, 1, 4, 8 or 16: mapping of the batched passes; -1 or -4: the
 * LDS-staged kernels of mappings 1 and 4),
 * "filtering", "filter_tol".  Returns a handle > 0, or -1.
 * Replaces c_api/draugr_amg.h:94. */
int32_t draugr_amg_config_from_json(const char *json_str);

/* Build the hierarchy on the host (setup phase) and upload it once to HBM.
 * Inputs are copied.  index_base 0 or 1.  Returns a handle > 0, or -1.
 * Replaces c_api/draugr_amg.h:115-120. */
int32_t draugr_amg_setup(int32_t n, int32_t nnz, const int32_t *rowptr, const int32_t *colval,
                         const double *nzval, int32_t config_handle, int32_t index_base,
                         int32_t allow_partial_resetup);

/* partial=1: new values with the same pattern -> Galerkin products, smoothers
 * and coarse factorisation recomputed on the GPU.  partial=0: full rebuild.
 * Returns 0 or -1.  Replaces c_api/draugr_amg.h:143-149. */
int32_t draugr_amg_resetup(int32_t handle, int32_t n, int32_t nnz, const int32_t *rowptr,
                           const int32_t *colval, const double *nzval, int32_t config_handle,
                           int32_t index_base, int32_t partial, int32_t allow_partial_resetup);

/* Stationary AMG iteration on host vectors (x in/out).  Returns the number
 * of iterations (maxiter when not converged), or -1.
 * Replaces c_api/draugr_amg.h:167-170. */
int32_t draugr_amg_solve(int32_t handle, int32_t n, double *x, const double *b, int32_t config_handle,
                         double tol, int32_t maxiter);

/* One V/W cycle on host vectors (x in/out).  Returns 0 or -1.
 * Replaces c_api/draugr_amg.h:178-180. */
int32_t draugr_amg_cycle(int32_t handle, int32_t n, double *x, const double *b, int32_t config_handle);

/* Replace c_api/draugr_amg.h:188 and :194.  Return 0, or -1 for an unknown handle. */
int32_t draugr_amg_free(int32_t handle);
int32_t draugr_amg_config_free(int32_t handle);

/* ── Part 2: backend extensions (no counterpart in the reference) ────────── */

/* Number of visible HIP devices (0 when none; never fails). */
int32_t draugr_hip_device_count(void);

/* Setup on the host only (no GPU needed): same arguments as draugr_amg_setup.
 * The handle supports the inspection calls below; draugr_hip_upload() makes
 * it usable for solves. */
int32_t draugr_hip_setup_host(int32_t n, int32_t nnz, const int32_t *rowptr, const int32_t *colval,
                              const double *nzval, int32_t config_handle, int32_t index_base,
                              int32_t allow_partial_resetup);
int32_t draugr_hip_upload(int32_t handle);

/* out[0..]: n_levels, n_coarse, n_fine, setup microseconds, uploaded (0/1),
 * total operator nnz, total rows.  Returns the number of values written. */
int32_t draugr_hip_hierarchy_info(int32_t handle, int64_t *out, int32_t nout);

/* out[0..]: n, nnz(A), n_coarse, nnz(P), ntriples, num_colors, has_rmap,
 * row blocks of A on the device (0 before upload), A stored with 16-bit
 * block-relative columns on the device (0/1), nonzeros of A in the
 * column-sorted form on the device. */
int32_t draugr_hip_level_info(int32_t handle, int32_t level, int64_t *out, int32_t nout);

/* Copy a host array of level `level` (0-based) into dst, 0-based indices:
 *  0 A.rowptr  1 A.colval  2 A.nzval(f64)  3 P.rowptr  4 P.colval  5 P.nzval(f64)
 *  6 Pt.offsets 7 Pt.fine_rows 8 Pt.p_nz_idx 9 R.nz_offsets(i64) 10 R.pi 11 R.ai
 * 12 R.pj 13 color_offsets 14 color_order
 * level == -1: 0 coarsest rowptr 1 colval 2 nzval 15 dense coarse (col-major f64). */
int32_t draugr_hip_level_export(int32_t handle, int32_t level, int32_t which, void *dst);

/* Hierarchy table as text (src/setup.jl:405-473 layout); returns length. */
int32_t draugr_hip_summary(int32_t handle, char *buf, int32_t buflen);

/* Device-pointer variants: x, b are FP64 arrays in HBM of length n. */
int32_t draugr_hip_cycle_device(int32_t handle, int32_t n, double *x, const double *b, int32_t config_handle,
                                int32_t zero_guess, void *stream);
int32_t draugr_hip_solve_device(int32_t handle, int32_t n, double *x, const double *b, int32_t config_handle,
                                double tol, int32_t maxiter, void *stream, double *relres);
/* FP64 PCG preconditioned by one cycle (cycle precision from the config). */
int32_t draugr_hip_pcg_device(int32_t handle, int32_t n, double *x, const double *b, int32_t config_handle,
                              double tol, int32_t maxiter, void *stream, double *relres);
int32_t draugr_hip_pcg(int32_t handle, int32_t n, double *x, const double *b, int32_t config_handle,
                       double tol, int32_t maxiter, double *relres);

/* draugr_hip_pcg_device without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory during a solve
 * (tight serving loops, code that must not call hipMalloc).  Same arguments,
 * kernels and results (bit for bit) as draugr_hip_pcg_device.  The PCG's
 * device scalars and the partials of its device-side sums live in a static
 * device array of the library (513 KiB per GPU, kept until the process ends;
 * the partials of r.r, which the host sums, go to the hierarchy's own pinned
 * host buffer, made with the hierarchy), so the finest
 * operator may have at most 65536 row blocks (about 16 M rows); larger
 * hierarchies get -1 with a message.  The hierarchy must already be prepared
 * for the config's precision by one call of draugr_hip_pcg_device (which
 * makes the precision mirrors, the kernel choices and the preconditioner
 * graph); otherwise -1.  The current HIP device must be the hierarchy's.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (different hierarchies or streams) race and must be serialised by the
 * caller. */
int32_t draugr_hip_pcg_device_noalloc(int32_t handle, int32_t n, double *x, const double *b, int32_t config_handle,
                                      double tol, int32_t maxiter, void *stream, double *relres);

/* Partial resetup from new fine-level values already in HBM (same pattern). */
int32_t draugr_hip_resetup_device(int32_t handle, int64_t nnz, const double *nzval, void *stream);

/* One hot-path kernel on level `level`, device pointers (prec 0 f64, 1 f32,
 * 2 bf16 values with f32 vectors):
 *  op 0 out = A in1           op 1 out = in2 - A in1
 *  op 2 out = in1 + w d (in2 - A in1)          (Jacobi / l1-Jacobi step)
 *  op 3 out = w d in2, out2 = in2 - A out      (zero-guess step + residual)
 *  op 4 out = P^T in1 (restriction)  op 5 out += P in1 (prolongation)
 *  op 6 out: one coloured Gauss-Seidel sweep in place, rhs in2
 *  op 7 out = A_coarse^-1 in1        op 8 Galerkin values of level+1
 *  op 9 out = inverse diagonal of the smoother (f64; SPAI(0): a_ii / sum a_ij^2)
 *  op 10 out: one Chebyshev smoothing step in place (out holds x), rhs in2
 *  op 11 out: one ILU(0) smoothing step in place (out holds x), rhs in2
 *  op 12 as op 7 with the vector in a global workspace (the form of a coarsest
 *        level whose vector does not fit in LDS)
 *  op 13 refactor the coarsest operator column by column over the chip (the
 *        form used above 256 rows; same factors) */
int32_t draugr_hip_level_op(int32_t handle, int32_t level, int32_t op, int32_t prec, const void *in1,
                            const void *in2, void *out, void *out2, void *stream);

/* draugr_hip_level_op `reps` times back to back on one stream (no host round
 * trip between the launches): the per-launch time of a kernel as a cycle sees
 * it, for measurement. */
int32_t draugr_hip_level_op_repeat(int32_t handle, int32_t level, int32_t op, int32_t prec, const void *in1,
                                   const void *in2, void *out, void *out2, int32_t reps, void *stream);

/* The PCG's preconditioner application (the zero-guess cycle in the config's
 * precision, replayed from its hipGraph when "graphs" is true, else launched
 * eagerly) `reps` times
 * back to back on the hierarchy's own residual and correction vectors, the
 * residual set from the device vector b first: the cycle time a PCG iteration
 * sees, for measurement (draugr_hip_cycle_device adds the caller-vector
 * copies and conversions around it). */
int32_t draugr_hip_precondition_repeat(int32_t handle, int32_t config_handle, const double *b, int32_t reps,
                                       void *stream);

/* Device-side FP64 values to the host: which 0 A.nzval, 1 smoother inverse
 * diagonal, 2 coarse LU factors (col-major), 3 coarsest operator values,
 * 4 Chebyshev bounds (lambda_min, lambda_max) of the level, 5 the level's
 * power-iteration start vector (chebyshev smoother), 6 / 7 the ILU(0)
 * factors L / U on the level's pattern (ilu0 smoother). */
int32_t draugr_hip_export_device_values(int32_t handle, int32_t level, int32_t which, double *dst);

/* Algorithmic bytes of one cycle from a zero guess under the given config,
 * and the HBM bytes the hierarchy holds. */
int32_t draugr_hip_cycle_bytes(int32_t handle, int32_t config_handle, double *bytes);
/* Bytes of one pass over level `level`'s operator `which` (0 A, 1 P, 2 P^T)
 * at `precision` (0 fp64, 1 fp32, 2 bf16): values, column indices, row
 * pointers; vectors excluded.  as_stored = 1: indices in the form the
 * autotuned kernel reads (16-bit block-relative where it chose that form);
 * 0: int32 indices (the algorithmic count, also that of
 * draugr_hip_cycle_bytes).  No counterpart in the reference (measurement). */
int32_t draugr_hip_op_bytes(int32_t handle, int32_t level, int32_t which, int32_t precision, int32_t as_stored,
                            double *bytes);
int32_t draugr_hip_resident_bytes(int32_t handle, int64_t *bytes);

/* ── Differentiation of x = A(k)^-1 b (device pointers) ─────────────────── */

/* Values of the 7-point cell-centred diffusion operator on an m^3 grid
 * (cell c = i + m (j + m z)) from cell permeabilities k: harmonic-mean face
 * transmissibilities, Dirichlet faces through a half-cell term 2 k.  Written
 * in the row layout of the hierarchy's finest operator (columns sorted), ready
 * for draugr_hip_resetup_device.  Requires n == m^3 and that pattern. */
int32_t draugr_hip_assemble_7pt(int32_t handle, int32_t m, const double *k, double *nzval, void *stream);

/* dL/dA_ij = -lambda_i x_j on the finest operator's pattern (grad has nnz). */
int32_t draugr_hip_matrix_grad(int32_t handle, const double *x, const double *lambda, double *grad, void *stream);

/* Chain rule of dL/dA to d L / d log k for the operator of
 * draugr_hip_assemble_7pt; grad has m^3 entries. */
int32_t draugr_hip_grad_logk_7pt(int32_t m, const double *k, const double *x, const double *lambda, double *grad,
                                 void *stream);

/* ── Batches of independent systems (no counterpart in the reference, which
 *    solves one system and one right-hand side per call: src/cycle.jl:12, :110,
 *    c_api/draugr_amg.h:167-180) ──────────────────────────────────────────── */

/* Set up nsys systems (system s: n[s] x n[s] CSR with nnz[s] entries, the
 * arrays of c_api/draugr_amg.h:115-120 per system) with one config, each with
 * its own hierarchy, and pack them for nrhs right-hand sides (a power of two,
 * 1..256).  The smoother must be jacobi or l1_jacobi and every system must have
 * at least one AMG level.  Returns a batch handle > 0, or -1. */
int32_t draugr_hip_batch_setup(int32_t nsys, const int32_t *n, const int32_t *nnz, const int32_t *const *rowptr,
                               const int32_t *const *colval, const double *const *nzval, int32_t config_handle,
                               int32_t index_base, int32_t nrhs);

/* out[0..]: nsys, nrhs, total rows, deepest hierarchy (AMG levels), setup
 * microseconds, resident bytes.  Returns the number of values written. */
int32_t draugr_hip_batch_info(int32_t batch, int64_t *out, int32_t nout);

/* out[0..]: n, row offset in the packed vectors, AMG levels, coarsest size. */
int32_t draugr_hip_batch_system_info(int32_t batch, int32_t sys, int64_t *out, int32_t nout);

/* Vectors of a batch are packed and in HBM: system s occupies rows
 * [offset_s, offset_s + n_s), row i holds its nrhs values contiguously
 * (entry (i, j) at i * nrhs + j).  x = V(0; b), one V(1,1) cycle per system and
 * right-hand side, in the config's precision ("fp64" or "fp32"). */
int32_t draugr_hip_batch_cycle_device(int32_t batch, double *x, const double *b, int32_t config_handle,
                                      void *stream);

/* FP64 PCG with the batched cycle as preconditioner, each (system, rhs)
 * stopping on its own ||r||/||b|| < tol; x is the initial guess and the result.
 * iters[s * nrhs + j], relres[s * nrhs + j] (may be NULL) receive each pair's
 * iteration count and final recurrence residual.  Returns the largest
 * iteration count, or -1. */
int32_t draugr_hip_batch_pcg_device(int32_t batch, double *x, const double *b, int32_t config_handle, double tol,
                                    int32_t maxiter, void *stream, int32_t *iters, double *relres);

int32_t draugr_hip_batch_free(int32_t batch);

/* The batch autotuner's choices (right-hand sides per thread of each level's
 * four passes, with the times it measured) as text; returns the length.  The
 * tuner runs on the first cycle of each precision. */
int32_t draugr_hip_batch_tune_report(int32_t batch, char *buf, int32_t buflen);

/* out[0..]: rows that smooth on packed level `level`, rows of its vectors,
 * nnz(A), nnz(P), rows of P^T.  Returns the number of values written. */
int32_t draugr_hip_batch_level_info(int32_t batch, int32_t level, int64_t *out, int32_t nout);

/* One K-wide Jacobi / l1-Jacobi step on packed level `level` with the tuned
 * mapping: out = x + w d (b - A x) for its smoothing rows (vectors rows x K,
 * row-major; prec 0 FP64, 1 FP32 vectors and values).  Measurement and test
 * hook. */
int32_t draugr_hip_batch_jacobi_pass(int32_t batch, int32_t level, int32_t prec, const void *x, const void *b,
                                     void *out, void *stream);

/* ILU(0) of a CSR operator on the host (0-based, columns sorted), as the
 * ilu0 smoother factorizes each level (src/smoothers.jl:1016-1098): diag[n]
 * receives the diagonal positions, L[nnz] / U[nnz] the factors on A's
 * pattern.  Needs no GPU. */
int32_t draugr_hip_ilu0_factor(int32_t n, const int32_t *rowptr, const int32_t *colval, const double *nzval,
                               int32_t *diag, double *L, double *U);

/* Kernel choices of the autotuner (per level and cycle precision) as text;
 * returns the length.  The tuner runs on the first cycle of each precision. */
int32_t draugr_hip_tune_report(int32_t handle, char *buf, int32_t buflen);

#ifdef __cplusplus
}
#endif

#endif /* DRAUGR_HIP_H */
