/* eigenfns_hip.h — C ABI of the MI355X (gfx950) Maxwell eigensolver kernels.
 *
 * The reference (fhernandez93/eigenfns) has no FFI: its boundary is the Python
 * operator protocol of eigenfns/operator.py and the module functions of
 * solver.py, chebyshev.py and interior.py, whose device work XLA generates.
 * Each entry point below replaces one of those device computations; the
 * citation names the reference code it stands in for.  The Python package
 * eigenfns_amd binds these with ctypes and keeps the reference's Python API.
 *
 * Conventions
 *  - Blocks of vectors are (m, 2, G, G, G) complex64 in device memory, C order,
 *    z fastest, frequencies in np.fft.fftfreq order; "c64" pointers point to
 *    interleaved float pairs, "c128" to interleaved double pairs.
 *  - The caller owns every buffer.  A context owns only its tables and a
 *    workspace allocated at creation; hot calls allocate nothing and do not
 *    synchronise the host.  Calls sharing one context must be ordered on one
 *    stream (they share its workspace).
 *  - stream is a hipStream_t (NULL = the default stream).
 *  - Every call returns 0 on success or a negative EFN_E* code; efn_last_error()
 *    returns the calling thread's last message.  No C++ exception crosses the ABI.
 */
#ifndef EIGENFNS_HIP_H
#define EIGENFNS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EFN_OK 0
#define EFN_EINVAL (-1)       /* invalid argument        -> ValueError   */
#define EFN_EUNSUPPORTED (-2) /* grid size not compiled  -> ValueError   */
#define EFN_EHIP (-3)         /* HIP runtime failure     -> RuntimeError */
#define EFN_ENOMEM (-4)       /* device allocation failed -> MemoryError */

typedef struct efn_ctx efn_ctx;

/* Library version (major*10000 + minor*100 + patch). */
int efn_version(void);
/* Last error message of the calling thread ("" if none). */
const char* efn_last_error(void);
/* Non-zero if the grid edge G has compiled kernels. */
int efn_grid_supported(int grid_size);

/* Operator context for one ε grid and Bloch k-point.
 * Replaces MaxwellOperator.__init__ + make_basis (eigenfns/operator.py:136-151,
 * :44-79): builds the FP32 |k+G| axis tables, the FP64 reference-axis switch of
 * operator.py:57 as a bit mask, ε⁻¹/G³ and ε/G³, and a workspace for
 * `max_chunk` vectors (3·G³ complex64 each; 0 = automatic, ~192 MB; < 0 = no
 * workspace: a table-only context for the slot lists of the *_multi calls).
 * eps_host: G³ floats on the host, C order. k in units of 2π/L. */
int efn_ctx_create(int grid_size, double box_size, double kx, double ky, double kz,
                   const float* eps_host, int max_chunk, efn_ctx** out);
int efn_ctx_destroy(efn_ctx* ctx);
/* kn_floor = 0.5·min nonzero FP32 |k+G| (operator.py:149-151), eps_mean_inv
 * = mean(1/ε) (operator.py:148), chunk = vectors per workspace pass. */
int efn_ctx_query(const efn_ctx* ctx, double* kn_floor, double* eps_mean_inv, int* chunk,
                  size_t* workspace_bytes);
/* |k+G| (G³ floats) and the frame t (2,3,G³ floats) on the device, as the
 * kernels rebuild them (TransverseBasis.kn / .t, operator.py:29-41). */
int efn_ctx_basis(const efn_ctx* ctx, float* kn_dev, float* t_dev, void* stream);

/* out = alpha·Θ·in + beta·in for m vectors.  Θ: operator.py:82-95 / :153-155
 * (alpha = 1, beta = 0); the Chebyshev map B = (2/λmax)Θ − I of
 * interior.py:341-342 and exp_kpm_dos.py:95-96 is alpha = 2/λmax, beta = −1.
 * in and out must not overlap. */
int efn_theta(efn_ctx* ctx, const void* in_c64, void* out_c64, int m, float alpha, float beta,
              void* stream);

/* Multi-operator Θ: one launch set over `nslot` (≤ 64) operators of the same
 * grid size — Bloch k-points, structures or independent requests.  Slot i
 * applies ctxs[i] to rows in_off[i] .. in_off[i]+nrow−1 of in_base and writes
 * rows out_off[i] .. of out_base (offsets in vectors of 2·G³ complex64).  The
 * passes run in ws's workspace.  Replaces a loop of operator.py:153-155 calls
 * over operators; no reference counterpart as a single call. */
int efn_theta_multi(efn_ctx* ws, int nslot, efn_ctx* const* ctxs, const void* in_base,
                    const int64_t* in_off, void* out_base, const int64_t* out_off, int nrow,
                    float alpha, float beta, void* stream);
/* Multi-operator preconditioner (operator.py:157-161 over several operators). */
int efn_precondition_multi(efn_ctx* ws, int nslot, efn_ctx* const* ctxs, const void* in_base,
                           const int64_t* in_off, void* out_base, const int64_t* out_off, int nrow,
                           void* stream);

/* Ragged multi-operator Θ / preconditioner: as efn_theta_multi and
 * efn_precondition_multi, but slot i holds rows[i] vectors (block widths may
 * differ, 0 allowed), so one launch set covers problems of different block
 * widths — e.g. every Bloch k-point of a band-structure path at its own
 * width.  Replaces a loop of operator.py:153-155 / :157-161 calls over
 * operators and widths; no reference counterpart as a single call. */
int efn_theta_ragged(efn_ctx* ws, int nslot, efn_ctx* const* ctxs, const void* in_base,
                     const int64_t* in_off, void* out_base, const int64_t* out_off, const int* rows,
                     float alpha, float beta, void* stream);
int efn_precondition_ragged(efn_ctx* ws, int nslot, efn_ctx* const* ctxs, const void* in_base,
                            const int64_t* in_off, void* out_base, const int64_t* out_off,
                            const int* rows, void* stream);

/* Real-space fields of m vectors (post-solve; operator.py:169-184 and the
 * energy density of run_modes.py:116-124 / run_interior.py:252-262):
 *   mode 0: H(r) = ifftn(a·t̂₁ + b·t̂₂)                      out (m,3,G,G,G) c64
 *   mode 1: E(r) = ε⁻¹·ifftn(−i|k+G|b·t̂₁ + i|k+G|a·t̂₂)/√λ   out (m,3,G,G,G) c64
 *   mode 2: ε·Σ_c |E_c(r)|²                                  out (m,G,G,G) float32
 * lam_dev: m eigenvalues (float32, device) for modes 1–2; NULL means λ = 1.
 * Uses the context's workspace (same passes as Θ up to the x pass). */
int efn_realspace(efn_ctx* ctx, int mode, const void* in_c64, const float* lam_dev, void* out, int m,
                  void* stream);

/* MPB transverse-projection preconditioner, operator.py:104-127 / :157-161. */
int efn_precondition(efn_ctx* ctx, const void* in_c64, void* out_c64, int m, void* stream);

/* Diagonal kinetic preconditioner out = in/(|k+G|² + scale), operator.py:98-101. */
int efn_precondition_simple(efn_ctx* ctx, const void* in_c64, void* out_c64, int m, float scale,
                            void* stream);

/* Fused Chebyshev step (interior.py:347-349):
 *   T0 ← 2·(alpha·Θ·T1 − T1) − T0,   Y ← Y + ck·T0(new).
 * One operator application, T0 and Y updated in place. */
int efn_cheb_step(efn_ctx* ctx, const void* T1_c64, void* T0_c64, void* Y_c64, float alpha,
                  float ck, int m, void* stream);

/* Fused KPM doubling step (scripts/exp/exp_kpm_dos.py:145-152), moments on the
 * device in FP64:  T0 ← 2·(alpha·Θ·T1 − T1) − T0 and, for every row v,
 *   mom[v·ld + 2k]   = 2⟨T1,T1⟩ − mom[v·ld + 0]
 *   mom[v·ld + 2k+1] = 2·Re⟨T0(new),T1⟩ − mom[v·ld + 1]. */
int efn_kpm_step(efn_ctx* ctx, const void* T1_c64, void* T0_c64, double* mom_dev, int ld, int k,
                 float alpha, int m, void* stream);

/* Fused KPM-count step (chebyshev.py:260-263):
 *   T0 ← 2·(alpha·Θ·T1 − T1) − T0,   est[v] += ck·Re⟨Z_v, T0(new)_v⟩ (FP64). */
int efn_kpm_count_step(efn_ctx* ctx, const void* T1_c64, void* T0_c64, const void* Z_c64,
                       double* est_dev, double ck, float alpha, int m, void* stream);

/* Process-wide pass structure of Θ and the preconditioner (and their fused
 * Chebyshev / KPM epilogues); returns the previous engine (engine < 0: query).
 *   EFN_THETA_AUTO  the per-grid choice compiled into the library (default)
 *   3               xin | yzmid | xout: 132 B of compulsory traffic per voxel
 *                   and vector, on grids that have the plane pass (32, 48, 64, 96, 128)
 *   5               zin | y | x·ε·x | y | zout: 228 B per voxel and vector
 * Either form computes Θ of operator.py:82-95 (the two differ only by FP32
 * rounding).  No reference counterpart (XLA schedules the reference's FFTs). */
#define EFN_THETA_AUTO 0
int efn_theta_engine(int engine);
/* The pass structure (3 or 5) Θ uses now at this grid size. */
int efn_theta_passes(int grid_size);
/* 1 if the three-pass form is compiled for this grid size, else 0 (host
 * only, no device needed). */
int efn_theta_three_pass(int grid_size);
/* This context's pass structure: 0 follows the engine, 3 or 5 applies while
 * the engine is EFN_THETA_AUTO (a forced engine still wins); passes < 0 only
 * queries.  Returns the form the context's launches use now (3 or 5), or
 * EFN_EUNSUPPORTED for 3 on a grid without the plane pass.  Set it before the
 * context is used on a stream; the tuning table (eigenfns_amd.autotune) sets
 * it at creation.  No reference counterpart. */
int efn_ctx_set_passes(efn_ctx* ctx, int passes);

/* Gram block out[i][j] = Σ_k conj(A[i][k])·B[j][k] (solver.py:35-43), products
 * of the FP32 inputs summed in FP64; out is ma×mb complex128 on the device.
 * n = complex entries per row (2·G³ for a block).  work is a caller-provided
 * device buffer of at least efn_gram_work_bytes(ma, mb, n) bytes (split
 * partials, summed in a fixed order, so results are bitwise reproducible). */
size_t efn_gram_work_bytes(int ma, int mb, int64_t n);

/* Process-wide Gram engine for efn_gram / efn_gram_multi; returns the previous
 * engine (engine < 0: query only).  All engines use the same split partials
 * and the same fixed-order FP64 sum over them.
 *   EFN_GRAM_VALU_F64  vector ALU: exact FP64 products, FP64 sums.
 *   EFN_GRAM_MFMA_F64  matrix cores (v_mfma_f64_16x16x4f64) on the exact FP64
 *                      images of the inputs: exact products, FP64 sums (default).
 *   EFN_GRAM_MFMA_F32  matrix cores in FP32 (v_mfma_f32_16x16x4f32); FP32 partial
 *                      sums of at most 64 products, compensated by adding them
 *                      into FP64 accumulators.
 * MFMA engines need n even and 16-byte aligned rows; other calls use the
 * vector-ALU kernel.  No reference counterpart (XLA picks the reference's
 * einsum kernels). */
#define EFN_GRAM_VALU_F64 0
#define EFN_GRAM_MFMA_F64 1
#define EFN_GRAM_MFMA_F32 2
int efn_gram_engine(int engine);
int efn_gram(const void* A_c64, int ma, const void* B_c64, int mb, int64_t n, void* out_c128,
             void* work, size_t work_bytes, void* stream);

/* efn_gram without a work buffer and without any allocation, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The split partials go to a static device array of EFN_GRAM_NOALLOC_MAX_BYTES
 * that each device holds for the life of the process; calls needing more
 * (efn_gram_work_bytes(ma, mb, n) > the maximum, e.g. beyond 96 × 96 at 192³)
 * return EFN_EINVAL.  Results are bitwise those of efn_gram (same kernels, same
 * splits).  Every call on one device shares that device's array, so calls that
 * overlap on one device (different streams, host threads) race: order them on
 * one stream. */
#define EFN_GRAM_NOALLOC_MAX_BYTES (40u << 20)
int efn_gram_noalloc(const void* A_c64, int ma, const void* B_c64, int mb, int64_t n,
                     void* out_c128, void* stream);

/* Gram blocks of `nprob` (≤ 64) independent problems in one launch:
 * out[p] = A_pᴴ·B_p (ma×mb complex128 each, contiguous), A_p / B_p device
 * pointers given in host arrays.  solver.py:35-43 per problem. */
size_t efn_gram_multi_work_bytes(int nprob, int ma, int mb, int64_t n);
int efn_gram_multi(int nprob, const void* const* A_c64, const void* const* B_c64, int ma, int mb,
                   int64_t n, void* out_c128, void* work, size_t work_bytes, void* stream);

/* Ragged Gram: problem p is ma[p] × mb[p] (the blocks may differ in row
 * count).  out holds nprob blocks of max(ma) × max(mb) complex128, zero
 * outside each problem's own ma[p] × mb[p]; work as efn_gram_multi_work_bytes
 * at the largest sizes.  hermitian != 0: square problems known to be
 * Hermitian (upper tiles computed, mirrored).  solver.py:35-43 per problem;
 * runs on the matrix-core engines and needs 16-byte aligned blocks. */
int efn_gram_ragged(int nprob, const void* const* A_c64, const void* const* B_c64, const int* ma,
                    const int* mb, int64_t n, int hermitian, void* out_c128, void* work,
                    size_t work_bytes, void* stream);

/* Hermitian form of efn_gram_multi for products known to be Hermitian
 * (A_pᴴA_p, or A_pᴴ·Θ·A_p that the caller symmetrises anyway): only the upper
 * block triangle of output tiles is computed; out[i][j] for i > j is the
 * conjugate of out[j][i] and the diagonal is real, so the result is exactly
 * Hermitian.  m×m outputs; same work buffer size as
 * efn_gram_multi_work_bytes(nprob, m, m, n).  The Rayleigh–Ritz matrix
 * A = Sᴴ·Θ·S of solver.py:275-294 (symmetrised at :294) per problem. */
int efn_gram_multi_herm(int nprob, const void* const* A_c64, const void* const* B_c64, int m,
                        int64_t n, void* out_c128, void* work, size_t work_bytes, void* stream);

/* Linear combination out_i = alpha·X_i + Σ_j C[j][i]·S_j (solver.py:49-57,
 * :106, :125): C is m_in×m_out complex64 on the device, S m_in rows, X
 * (optional, NULL = none) and out m_out rows of n complex64.  out may be S
 * itself (in place) when m_out <= 64 and m_out <= m_in (one row tile: every
 * column is read in full before it is written); otherwise it must not
 * overlap S. */
int efn_combine(const void* C_c64, int m_in, int m_out, const void* S_c64, int64_t n,
                const void* X_c64, float alpha, void* out_c64, void* stream);

/* Batched Hermitian eigendecomposition of nprob small matrices (n ≤ 96) on
 * the device: A (nprob × n × n complex128, row-major; only its upper triangle
 * and the real part of its diagonal are read) → w (nprob × n, ascending) and
 * V (nprob × n × n complex128, column j the eigenvector of w[j]).  Cyclic
 * two-sided Jacobi in FP64, one workgroup per matrix, relative convergence
 * test (small eigenvalues keep their relative accuracy); max_sweeps <= 0
 * means 30.  Replaces the host eigensolves of SVQB and Rayleigh–Ritz
 * (solver.py:119-121, :294-297) so a LOBPCG iteration needs no host round
 * trip for them. */
int efn_heev_batched(int nprob, int n, const void* A_c128, double* w, void* V_c128, int max_sweeps,
                     void* stream);

/* Cholesky-QR coefficients of nprob Hermitian positive semidefinite Gram
 * matrices G (n × n complex128, row-major, n ≤ 96): pivoted Cholesky of G in
 * FP64, stopped where the largest remaining pivot is ≤ drop·max(diag G)
 * (rank[p] = pivots taken); C (n × n complex128) = Π·L⁻ᴴ on those pivots and
 * zero elsewhere, so Cᴴ·G·C = diag(I_rank, 0) and the rows Σ_j C[j][i]·y_j
 * are orthonormal.  The orthonormalisation the reference does by SVQB
 * (solver.py:113-126, :258-274) with a Cholesky factor in place of the
 * eigendecomposition.  G and C must not alias. */
int efn_cholqr_batched(int nprob, int n, const void* G, void* C, double drop, int* rank, void* stream);

/* The same decomposition with rocSOLVER's batched divide and conquer
 * (rocsolver_zheevd_strided_batched, FP64): any n, same layouts (only A's
 * upper triangle is read), results to LAPACK accuracy; measured 1.55 ms for
 * 16 matrices of 72 × 72 (the Jacobi kernel above takes 7.7 ms per matrix).
 * work: efn_heevd_work_bytes(nprob, n) bytes of device scratch (the
 * tridiagonal off-diagonals and per-matrix status); the rocBLAS handle behind
 * it is created once per host thread and device.  Replaces the host
 * eigensolves of solver.py:119-121, :263-265, :301. */
size_t efn_heevd_work_bytes(int nprob, int n);
int efn_heevd_batched(int nprob, int n, const void* A_c128, double* w, void* V_c128, void* work,
                      size_t work_bytes, void* stream);

/* The lowest `lowest` eigenpairs of nprob Hermitian matrices (n ≤ 96, FP64),
 * one workgroup per matrix with everything in LDS: Householder reduction to
 * a real tridiagonal matrix, implicit QL with Wilkinson shifts, and the
 * back-transformation of the selected eigenvectors only.  Allocates nothing
 * (can be captured in a HIP graph).  A: nprob × n × n complex128 row-major,
 * only the lower triangle read; with n_each (host array, nprob ≤ 64) problem
 * p is the leading n_each[p] × n_each[p] block of its n × n slot.  w: nprob ×
 * lowest, ascending, +inf past min(lowest, n_p); V: nprob × n × lowest
 * (V[p][i][j] = component i of eigenvector j, zero past n_p rows and
 * min(lowest, n_p) columns); info (device, nprob ints): 0, or 1 + l when the
 * QL iteration for eigenvalue l did not converge within 30·n_p iterations.
 * A and V must not overlap.  The host and rocSOLVER eigensolves of
 * solver.py:119-121, :263-265, :301 without a library workspace. */
int efn_heev_ql_batched(int nprob, int n, const int* n_each, const void* A_c128, int lowest, double* w,
                        void* V_c128, int* info_dev, void* stream);

/* The same lowest eigenpairs by bisection and inverse iteration (LAPACK
 * zheevx's scheme: dstebz + zstein) after the same Householder reduction:
 * the eigenvalues of the tridiagonal matrix by multisection on Sturm counts,
 * many at once; the selected eigenvectors by inverse iteration, one cluster
 * of close eigenvalues per thread, orthogonalised within the cluster.  Same
 * arguments, layouts and guarantees as efn_heev_ql_batched (a lowest-k solve
 * equals the leading columns of the full one bit for bit), except info: the
 * number of eigenvectors of the problem whose inverse iteration did not pass
 * its stopping test (their vectors are still returned).  Allocates nothing.
 * The host and rocSOLVER eigensolves of solver.py:119-121, :263-265, :301. */
int efn_heevx_batched(int nprob, int n, const int* n_each, const void* A_c128, int lowest, double* w,
                      void* V_c128, int* info_dev, void* stream);

/* Process-wide combine engine: 1 = matrix cores (v_mfma_f32_16x16x4f32,
 * default, for m_out > 8), 0 = vector ALU; FP32 accumulation either way.
 * Returns the previous engine (engine < 0: query only).  No reference
 * counterpart. */
int efn_combine_engine(int engine);

/* combine for `nprob` (≤ 32) problems in one launch: C is nprob×m_in×m_out
 * complex64 on the device; S, X (optional, NULL = none) and out are host
 * arrays of per-problem device pointers.  solver.py:49-57 per problem. */
int efn_combine_multi(int nprob, const void* C_c64, const void* const* S_c64, int m_in, int m_out,
                      int64_t n, const void* const* X_c64, float alpha, void* const* out_c64,
                      void* stream);

/* efn_combine_multi without X, its m_out output rows split over two blocks:
 * row i < split to out[p] + i·n, row i >= split to out2[p] + (i − split)·n.
 * With m_out <= 64 both blocks may lie inside S's rows (one row tile), so
 * the two Rayleigh–Ritz updates X ← S·Cx and P ← S·Cp (solver.py:298-307)
 * land in S's X and P rows in one pass over S.  No reference counterpart (the
 * reference forms X and P in separate products). */
int efn_combine_multi_split(int nprob, const void* C_c64, const void* const* S_c64, int m_in,
                            int m_out, int64_t n, void* const* out_c64, void* const* out2_c64,
                            int split, void* stream);

/* Ragged combine: problem p has m_in[p] rows of S, m_out[p] outputs and its
 * own split (rows < split[p] to out[p], the rest to out2[p], which may be
 * NULL when every split equals m_out); C holds nprob blocks of
 * max(m_in) × max(m_out) complex64 of which each problem reads its leading
 * m_in[p] × m_out[p].  X / alpha as efn_combine_multi.  solver.py:298-307
 * per problem; matrix-core engine. */
int efn_combine_ragged(int nprob, const void* C_c64, const void* const* S_c64, const int* m_in,
                       const int* m_out, int64_t n, const void* const* X_c64, float alpha,
                       void* const* out_c64, void* const* out2_c64, const int* split, void* stream);

/* Row reductions in FP64 (solver.py:228-231, interior.py:75-82):
 *   op 0: out[i] = (Σ_k |A[i][k]|², 0)      op 1: out[i] = Σ_k conj(A[i][k])·B[i][k]
 * out is m complex128 on the device. */
size_t efn_rowreduce_work_bytes(int m, int64_t n);
int efn_rowreduce(int op, const void* A_c64, const void* B_c64, int m, int64_t n, void* out_c128,
                  void* work, size_t work_bytes, void* stream);

/* Strided-row forms for blocks that hold several problems' rows: logical row
 * i (0 ≤ i < m) is physical row (i / nrow)·stride + i % nrow of its array
 * (e.g. the X rows of every slot in slot-major [X; W; P] storage:
 * nrow = block width, stride = 3·block width). */
int efn_rowreduce_strided(int op, const void* A_c64, int64_t stride_a, const void* B_c64,
                          int64_t stride_b, int nrow, int m, int64_t n, void* out_c128, void* work,
                          size_t work_bytes, void* stream);
int efn_row_axpby_strided(const void* X_c64, int64_t stride_x, const double* a_dev,
                          const void* Y_c64, int64_t stride_y, const double* b_dev, int nrow, int m,
                          int64_t n, void* out_c64, int64_t stride_out, void* stream);

/* Fused residual statistics (solver.py:228-238 per row, one read of each
 * block): out4[i] = (Re⟨A_i,B_i⟩, Im⟨A_i,B_i⟩, |A_i|², |B_i|²) in FP64, strided
 * rows as above.  work ≥ efn_rowstats_work_bytes(m, n). */
size_t efn_rowstats_work_bytes(int m, int64_t n);
int efn_rowstats_strided(const void* A_c64, int64_t stride_a, const void* B_c64, int64_t stride_b,
                         int nrow, int m, int64_t n, double* out4, void* work, size_t work_bytes,
                         void* stream);
/* out_i = a[i]·X_i + b[i]·Y_i (strided rows) and norm2[i] = (|out_i|², 0) of the
 * stored FP32 values in FP64, in one pass (the residual R = HX − λX and its
 * norm, solver.py:229-231).  out must not alias X or Y;
 * work ≥ efn_rowreduce_work_bytes(m, n). */
int efn_row_axpby_norm_strided(const void* X_c64, int64_t stride_x, const double* a_dev,
                               const void* Y_c64, int64_t stride_y, const double* b_dev, int nrow,
                               int m, int64_t n, void* out_c64, int64_t stride_out,
                               double* norm2_c128, void* work, size_t work_bytes, void* stream);
/* The two above over selected rows (the same residual statistics,
 * solver.py:228-238, for the occupied rows of slots of different widths):
 * logical row i is row r of slot `slot` with sel_dev[i] = slot << 16 | r
 * (device int32, m entries), physical row slot·stride + r of each operand;
 * an entry with r ≥ nrow or slot ≥ nslot reads and writes nothing and gives
 * zero sums.  0 < nrow < 65536, 0 < nslot < 32768, every stride ≥ nrow.
 * Each selected row's sums are the strided entry point's for that row, bit
 * for bit.  Same work buffers. */
int efn_rowstats_sel(const void* A_c64, int64_t stride_a, const void* B_c64, int64_t stride_b,
                     const int* sel_dev, int nrow, int nslot, int m, int64_t n, double* out4, void* work,
                     size_t work_bytes, void* stream);
int efn_row_axpby_norm_sel(const void* X_c64, int64_t stride_x, const double* a_dev, const void* Y_c64,
                           int64_t stride_y, const double* b_dev, const int* sel_dev, int nrow, int nslot,
                           int m, int64_t n, void* out_c64, int64_t stride_out, double* norm2_c128,
                           void* work, size_t work_bytes, void* stream);

/* out_i = a[i]·X_i + b[i]·Y_i with per-row FP64 coefficients on the device
 * (Y, b may be NULL).  R = HX − λX (solver.py:229), row scaling (:306-307). */
int efn_row_axpby(const void* X_c64, const double* a_dev, const void* Y_c64, const double* b_dev,
                  int m, int64_t n, void* out_c64, void* stream);

/* out = a·X + b·Y elementwise over count complex entries (Y may be NULL). */
int efn_axpby(const void* X_c64, float a, const void* Y_c64, float b, int64_t count, void* out_c64,
              void* stream);

/* Batched copy of `count` complex64 entries for `nprob` (≤ 32) problems:
 * dst[p][e] = src[p][e] (src NULL: zero fill).  Moves rows between the
 * fixed per-slot buffers of the batched solver (solver.py:65-71's locked-block
 * write, per slot). */
int efn_copy_multi(int nprob, const void* const* src_c64, void* const* dst_c64, int64_t count,
                   void* stream);

/* Random block of m rows (rows row0 .. row0+m−1 of the stream `seed`), Γ-masked.
 * kind 0: complex normal (Box–Muller), kind 1: Rademacher ±1.  Replaces the JAX
 * threefry draws of solver.py:196-205, chebyshev.py:243-248, exp_kpm_dos.py:138-140. */
int efn_random_block(const efn_ctx* ctx, int kind, uint64_t seed, int64_t row0, int m,
                     void* out_c64, void* stream);

/* Binary "pen-like" rasterisation of a rod network into ε (G³ floats on the
 * device, C order), rasterize_penlike of structure.py:51-151.  The caller
 * computes the per-rod FP64 parameters with the reference's expressions
 * (eigenfns_amd.structure does): rodf_dev = nrods×8 doubles (p1u[3] — the
 * first endpoint with z divided by the aspect ratio —, the unit axis n[3] in
 * the unwarped space, its length L_u, unused), rodi_dev = nrods×6 ints (first
 * and last voxel index along x, y, z; unwrapped indices when periodic,
 * clipped to [0, G) otherwise).  bb = radius².  Every voxel centre in a rod's
 * box is tested in FP64 with the reference's operation order and no fused
 * multiply-add, so the mask is bit-identical to the host rasteriser's.
 * Voxels inside a rod get eps_rod, all others eps_bg. */
int efn_rasterize(int grid_size, double box_size, double aspect, double bb, const double* rodf_dev,
                  const int* rodi_dev, int nrods, int periodic, float eps_rod, float eps_bg,
                  float* eps_dev, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* EIGENFNS_HIP_H */
