/*
 * dalia_hip.h -- C ABI of the MI355X (gfx950) inner-solve library.
 *
 * One handle owns the dense storage of one block-tridiagonal-with-arrowhead
 * (BTA) SPD matrix of n diagonal blocks of size b and an arrowhead of size a,
 * and runs on one HIP stream.  Every entry point is FP64, takes plain device
 * pointers, returns an int status (0 = ok, < 0 = bad argument or HIP error)
 * and is asynchronous on the handle's stream unless its comment says it
 * synchronises.  A handle is not thread safe.
 *
 * Storage: blocks are row-major with a leading dimension padded to a multiple
 * of 64 (bp >= b); the arrowhead is stored with ap = 64 rows (a <= 64).
 * Padding rows/columns hold the identity on the diagonal blocks and the tip
 * and zeros elsewhere, so the factorisation, solve and selected inversion of
 * the padded matrix equal those of the logical one.  The logical blocks are
 * the [b][b], [a][b] and [a][a] sub-views at offset 0 (see dh_layout).
 *
 * Reference interfaces replaced (DALIA, affifboudaoud/DALIA):
 *   dh_create / dh_destroy   SerinvSolver.__init__   src/dalia/solvers/structured_solver.py:22-70
 *   dh_assemble              Model.construct_Q_prior / construct_Q_conditional +
 *                            SerinvSolver._spmatrix_to_structured
 *                            src/dalia/core/model.py:306-423,
 *                            src/dalia/submodels/spatio_temporal.py:114-149,
 *                            src/dalia/solvers/structured_solver.py:192-468
 *   dh_scatter_coo           SerinvSolver._spmatrix_to_bta / _spmatrix_to_bt + the block
 *                            mapping kernel   structured_solver.py:252-468,
 *                            src/dalia/kernels/blockmapping.py:12-169
 *   dh_cholesky              SerinvSolver.cholesky (pobtaf / pobtf)   structured_solver.py:72-100
 *   dh_solve                 SerinvSolver.solve (pobtas N then C)     structured_solver.py:102-149
 *   dh_logdet                SerinvSolver.logdet                      structured_solver.py:151-166
 *   dh_selected_inversion    SerinvSolver.selected_inversion (pobtasi) structured_solver.py:168-190
 *   dh_qc_matvec             Q_conditional @ x inside DALIA._evaluate_conditional_latent_parameters
 *                            src/dalia/core/dalia.py:1161
 *   dh_scale, dh_dot         the vector algebra of core/model.py:446 and core/dalia.py:1161
 *   dh_contract_fem, dh_contract_fem_vec, dh_contract_ata, dh_prior_eig,
 *   dh_contract_fem_v, dh_contract_fem_vec_v, dh_contract_ata_v, dh_ata_quad_v
 *                            no counterpart: the adjoint gradient that replaces the
 *                            finite-difference gradient of DALIA._objective_function
 *                            src/dalia/core/dalia.py:513-601
 */
#ifndef DALIA_HIP_H
#define DALIA_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define DH_SPARSITY_BT 0
#define DH_SPARSITY_BTA 1

typedef struct dh_handle dh_handle;

typedef struct dh_layout {
  int n, b, a;        /* logical sizes */
  int bp, ap;         /* padded block size (multiple of 64) and arrow rows (64 or 0) */
  double* diag;       /* [n][bp][bp]   diagonal blocks (L_ii, later X_ii) */
  double* lower;      /* [n-1][bp][bp] sub-diagonal blocks (A_{i+1,i}); NULL when n == 1 */
  double* arrow;      /* [n][ap][bp]   arrowhead bottom blocks; NULL when a == 0 */
  double* tip;        /* [ap][ap]      arrowhead tip; NULL when a == 0 */
  double* dinv;       /* [n][bp]       1 / diag(L_ii) after dh_cholesky */
} dh_layout;

/* Device-resident, theta-independent model data of the spatio-temporal
 * Gaussian model; all pointers are device pointers owned by the caller and
 * must outlive every call that receives the struct. */
typedef struct dh_model {
  int ns, nt, nb;
  int fem_nnz;
  const int* fem_rowptr;     /* [ns+1] union pattern of c0, g1, g2, g3 */
  const int* fem_col;        /* [fem_nnz] */
  const double* fem_val;     /* [4][fem_nnz]  c0, g1, g2, g3 on that pattern */
  const double* m_diag;      /* [3][nt]    diagonals of M0, M1, M2 */
  const double* m_off;       /* [3][nt-1]  sub-diagonals of M0, M1, M2 */
  long long ata_nnz;
  const long long* ata_rowptr; /* [ns*nt+1] A_st^T A_st, CSR, global columns */
  const int* ata_col;
  const double* ata_val;
  const double* ata_arrow;   /* [nb][ns*nt] X^T A_st */
  const double* ata_tip;     /* [nb][nb]    X^T X */
  double fe_prec;            /* prior precision of the fixed effects */
} dh_model;

int dh_version(void);

/* Allocates the BTA storage and scratch on the current device. */
int dh_create(int n, int b, int a, void* hip_stream, dh_handle** out);
int dh_destroy(dh_handle* h);
int dh_get_layout(const dh_handle* h, dh_layout* out);
int dh_set_stream(dh_handle* h, void* hip_stream);
/* Bytes of device memory the handle owns. */
long long dh_memory_bytes(const dh_handle* h);

/* Fill the blocks with Q = sum_k sum_m c[k][m] (M_k kron G_m)
 * (+ fe_prec I on the tip for BTA) (+ tau_scale A^T A when with_likelihood).
 * params: DEVICE array of 13 doubles, c[3][4] row-major then tau_scale =
 * exp(theta_obs); read by the kernels at run time, so a captured graph can be
 * replayed after updating them.  sparsity BT drops the arrowhead. */
int dh_assemble(dh_handle* h, const dh_model* m, const double* params,
                int sparsity, int with_likelihood);

/* Coregional assembly (nv = 2 or 3 variates on one mesh): the handle holds
 * b = nv*ns, a = nv*nb.  m[0..nv-1] carry each variate's likelihood parts
 * (ata_*, fe_prec); the FEM and temporal parts are shared and read from m[0].
 * Diagonal/sub-diagonal block rows are variate-major ([variate 0 nodes,
 * variate 1 nodes, ...]), the arrowhead holds the fixed effects of all
 * variates in order.  params: DEVICE array, C_ij[3][4] for every variate pair
 * (i, j) row-major (nv*nv*12 doubles: q_ij = sum_k sum_m C_ij[k][m] M_k kron G_m),
 * then exp(prec_o_i) for i < nv.  Replaces CoregionalModel.construct_Q_prior /
 * construct_Q_conditional with its permutations (models/coregional_model.py:
 * 279-551, 595-639, 849-915). */
int dh_assemble_coreg(dh_handle* h, const dh_model* m, int nv, const double* params, int sparsity,
                      int with_likelihood);
/* y[p] = s_dev[idx[p]] * x[p] (device arrays): the coregional information
 * vector A^T D y from the per-variate A_i^T y_i.  Part of
 * CoregionalModel.construct_information_vector (coregional_model.py:641-680). */
int dh_scale_indexed(dh_handle* h, const double* x, const int* idx, const double* s_dev, double* y, long long len);

/* Fill the blocks from a symmetric sparse matrix of order n*b (+ a) given as
 * COO with global 0-based indices (device arrays, no duplicates): diagonal
 * blocks take both triangles, then the sub-diagonal blocks, the arrow rows
 * and the tip; entries outside the BTA pattern or outside the matrix are
 * ignored, as are the arrowhead rows for BT. */
int dh_scatter_coo(dh_handle* h, long long nnz, const long long* rows, const long long* cols,
                   const double* vals, int sparsity);

/* In-place BTA (pobtaf) or BT (pobtf) Cholesky.  A failing pivot shows up as
 * NaN in dh_logdet and as a positive value from dh_get_info. */
int dh_cholesky(dh_handle* h, int sparsity);

/* In-place solve A x = rhs with the factor; rhs is a device vector of
 * length n*b (+ a for BTA), logical (unpadded) layout. */
int dh_solve(dh_handle* h, double* rhs, int sparsity);

/* logdet = 2 sum log diag(L).  out_dev: device double (async).  NaN when the factorization
 * failed: a non-positive pivot, or a readiness wait of the persistent launch that gave up, in
 * which case the NaN's bits are DH_NAN_GIVEUP (and dh_get_info returns -4 until the next
 * factorization). */
#define DH_NAN_GIVEUP 0x7ff80000dead0004ULL

/* dh_logdet without a handle, for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc); no counterpart in the
 * reference.  logdet = 2 (sum_i sum_k log D_i[k][k] + sum_k log tip[k][k]) of a factor held in
 * the caller's device buffers: block i of n at diag + i * block_stride, b x b with leading
 * dimension ld; the tip (a x a, leading dimension ld_tip) is skipped when a == 0 (BT).  The
 * handle's own layout (dh_get_layout) is diag, n, b, bp, bp * bp, tip, a, ap.  *out_dev is a
 * device double written on hip_stream (asynchronous).  Calls no allocating routine: the n
 * partial sums live in one fixed device array of DH_NOALLOC_MAX_BLOCKS doubles per GPU (512 KiB,
 * held until the process ends) that every call on that GPU shares, so calls that overlap on one
 * GPU race.  n > DH_NOALLOC_MAX_BLOCKS returns DH_EARG. */
#define DH_NOALLOC_MAX_BLOCKS 65536
int dh_logdet_noalloc(const double* diag, int n, int b, int ld, long long block_stride, const double* tip, int a,
                      int ld_tip, double* out_dev, void* hip_stream);

int dh_logdet(dh_handle* h, int sparsity, double* out_dev);

/* Overwrite the factor with the blocks of A^-1 on the BTA pattern; the
 * diagonal blocks and the tip are stored full symmetric. */
int dh_selected_inversion(dh_handle* h, int sparsity);

/* Synchronises the stream; *info = 0 or 1 + global row of the first failing pivot.
 * Returns -4 if a workgroup gave up waiting for a factored diagonal tile
 * inside a fused launch (results are then invalid). */
int dh_get_info(dh_handle* h, int* info);

/* The zero-tile profiles the last persistent factorization used (tuning key 9): out[s * nbt + R]
 * = the first 64-column tile of tile row R of sub-diagonal slot s that held a nonzero entry (nbt:
 * none), s = 0 .. n-2, as stored at that time (two-ended: slots m .. n-2 hold the transposed
 * coupling).  cap >= (n - 1) * nbt ints.  Synchronises the handle's stream.  DH_ESTATE when the
 * last factorization read no profiles (key 9 off, or the per-panel schedule).  For the algorithmic
 * flop count of a profiled factorization.  No counterpart in the reference (Serinv's pobtaf
 * treats every block as dense). */
int dh_get_tile_profile(dh_handle* h, int* out, int cap);

/* y_prior = Q_prior x (spatio-temporal part + fe_prec on the fixed effects),
 * y_like = tau_scale A^T A x; vectors of length ns*nt + nb (device);
 * params as for dh_assemble (device). */
int dh_qc_matvec(dh_handle* h, const dh_model* m, const double* params,
                 const double* x, double* y_prior, double* y_like);

/* y = s * x (device vectors of length len). */
int dh_scale(dh_handle* h, const double* x, double s, double* y, long long len);
/* y = s_dev[0] * x, the scale read from device memory (graph-replay friendly). */
int dh_scale_dev(dh_handle* h, const double* x, const double* s_dev, double* y, long long len);
/* *out_dev = x . y */
int dh_dot(dh_handle* h, const double* x, const double* y, long long len, double* out_dev);

/* After dh_selected_inversion (BTA or BT): out_dev[(2n-1)*4] =
 * <X_block, G_m>_F for the n diagonal blocks then the n-1 sub-diagonal blocks. */
int dh_contract_fem(dh_handle* h, const dh_model* m, double* out_dev);
/* out_dev[(2n-1)*4] = x_t^T G_m x_t for t < n, then x_{t+1}^T G_m x_t. */
int dh_contract_fem_vec(dh_handle* h, const dh_model* m, const double* x, double* out_dev);
/* After dh_selected_inversion (BTA): *out_dev = <X, A^T A> over all blocks. */
int dh_contract_ata(dh_handle* h, const dh_model* m, double* out_dev);
/* The same contractions for a model whose nodes sit at an offset inside larger
 * blocks (variate k of a coregional model: offset k ns, fixed effects at arrow
 * row k nb), the inputs of the coregional adjoint gradient; no counterpart in
 * the reference, which differentiates the coregional f(theta) by central
 * differences (core/dalia.py:513-601).
 *   dh_contract_fem_v: <X sub-block (rows roff.., cols coff..), G_m> per block pair;
 *   dh_contract_fem_vec_v: x_t[roff..]^T G_m x_t'[coff..] with x[t stride + ...];
 *   dh_contract_ata_v: <X, [A X]^T [A X]> for the model's rows and fixed effects;
 *   dh_ata_quad_v: x^T [A X]^T [A X] x, fixed effects read at x[n stride + aoff].
 * Output layouts as for the unshifted calls. */
int dh_contract_fem_v(dh_handle* h, const dh_model* m, int roff, int coff, double* out_dev);
int dh_contract_fem_vec_v(dh_handle* h, const dh_model* m, const double* x, int stride, int roff, int coff,
                          double* out_dev);
int dh_contract_ata_v(dh_handle* h, const dh_model* m, int voff, int aoff, double* out_dev);
int dh_ata_quad_v(dh_handle* h, const dh_model* m, const double* x, int stride, int voff, int aoff,
                  double* out_dev);
/* Kronecker-eigen closed form of the prior log-determinant:
 * out_dev[4] = (sum_j logdet T_j, d/dgamma_s, d/dgamma_t, d/dgamma_st). */
int dh_prior_eig(dh_handle* h, const dh_model* m, const double* lam, const double* gam,
                 double* out_dev);
/* The same with (gamma_s, gamma_t, gamma_st) in device memory, read when the kernel runs (so that a
 * captured graph can be replayed for a new theta).  No counterpart in the reference. */
int dh_prior_eig_dev(dh_handle* h, const dh_model* m, const double* lam, const double* gam_dev,
                     double* out_dev);

/* Marginal variances of the latent field, after dh_selected_inversion (BTA):
 * out[i b + s] = X_ii[s][s], out[n b + j] = X_aa[j][j]  (length n b + a, device).
 * Replaces extract_diagonal(_structured_to_spmatrix(eye, "bta")) in
 * DALIA.get_marginal_variances_latent_parameters (core/dalia.py:939-962). */
int dh_marginal_variances(dh_handle* h, double* out);
/* Marginal variances of the observations, after dh_selected_inversion (BTA):
 * out[o] = a_o^T X a_o with a_o = [A_st row o, xcov row o]; A_st as CSR over
 * the n b spatio-temporal latents (int64 row pointers, int32 columns), xcov
 * dense row-major n_obs x a (may be null when a = 0).  All device pointers.
 * Replaces diag(A X A^T) in DALIA.get_marginal_variances_observations
 * (core/dalia.py:964-1016). */
int dh_obs_variances(dh_handle* h, long long n_obs, const long long* rowptr, const int* col, const double* val,
                     const double* xcov, double* out);

/* Checkpointed evaluation (for matrices whose full factor does not fit the memory budget).  No
 * counterpart in the reference, which keeps the whole factor in the solver's buffers
 * (src/dalia/solvers/structured_solver.py:38-60) and bounds memory only by splitting the time
 * axis over processes (distributed_structured_solver.py:62-64).  A context holds a ring of K + 2
 * block slots, the boundary factor of every segment of K time blocks (L_{j,j-1}, L_{a,j-1}: the
 * checkpoints) and an n-block handle for vector operations (dh_ckpt_vectors: use it for
 * dh_qc_matvec, dh_dot, dh_contract_fem_vec, dh_prior_eig on the whole model).  Memory: about
 * (3 K + 22 + n / K) bp^2 doubles instead of 3 n bp^2.
 *   dh_ckpt_forward: factors Qc (BTA, assembled from the model with the likelihood) or the prior
 *     (BT) segment by segment; *out_dev = logdet (NaN on a failed factorization, as dh_logdet);
 *     BTA with rhs (device, n b + a, logical layout): also the forward sweep of the solve and the
 *     tip solve (rhs is read only).
 *   dh_ckpt_backward: directly after dh_ckpt_forward(BTA, rhs) with the same params (no other
 *     dh_ckpt_forward in between: the last segment's factor is taken from it): recomputes the
 *     segments in reverse from their checkpoints, finishes the solve (rhs = Qc^-1 rhs) and, when
 *     selinv != 0, forms the selected inverse segment by segment and contracts it:
 *     cx_dev[(2n-1)*4] as dh_contract_fem, *ata_dev = <X, A^T A> as dh_contract_ata, and
 *     var_dev (n b + a, may be null) = diag(Qc^-1) as dh_marginal_variances. */
typedef struct dh_ckpt dh_ckpt;
int dh_ckpt_create(int n, int b, int a, int K, void* hip_stream, dh_ckpt** out);
int dh_ckpt_destroy(dh_ckpt* c);
long long dh_ckpt_memory_bytes(const dh_ckpt* c);
dh_handle* dh_ckpt_vectors(dh_ckpt* c);
int dh_ckpt_set_stream(dh_ckpt* c, void* hip_stream);
int dh_ckpt_get_info(dh_ckpt* c, int* info);
/* dh_set_tuning on the context's segment factorizations and recursions (keys 0, 2, 3, 5, 6, 8, 9,
 * 10; keys 4, 7, 12 and 13 are refused; key 11 has no effect: the segments are assembled by the context). */
int dh_ckpt_set_tuning(dh_ckpt* c, int key, int value);
/* dh_timing_enable / dh_timing_read on the context's segment launches. */
int dh_ckpt_timing_enable(dh_ckpt* c, int kind);
int dh_ckpt_timing_read(dh_ckpt* c, double* total_ms, long long* count);
int dh_ckpt_forward(dh_ckpt* c, const dh_model* m, const double* params, int sparsity, const double* rhs,
                    double* out_dev);
int dh_ckpt_backward(dh_ckpt* c, const dh_model* m, const double* params, double* rhs, int selinv, double* cx_dev,
                     double* ata_dev, double* var_dev);

/* Several factorizations in one persistent launch: the handles of a batch (each with its own
 * sizes, assembled beforehand, all with the persistent schedule).  The tasks of all instances are
 * merged into one order in which every instance keeps its own order and progresses in proportion
 * to its length, so the instances' chains of diagonal tiles run side by side.
 * dh_multi_cholesky(mc, sparsity, stream) then does what dh_cholesky(h, sparsity) does for every
 * handle, on that stream (which every handle's preceding and following work must be ordered
 * with).  No counterpart in the reference, which runs such evaluations on separate processes
 * (core/dalia.py:541-573). */
typedef struct dh_multi dh_multi;
int dh_multi_create(dh_handle* const* hs, int count, dh_multi** out);
int dh_multi_destroy(dh_multi* mc);
int dh_multi_cholesky(dh_multi* mc, int sparsity, void* hip_stream);

/* Task table of one block of the persistent pobtaf/pobtf launch for nbt 64-column tiles per block
 * (last != 0: the last block, without sub-diagonal tiles; arrow != 0: with the arrowhead), as the
 * launch decodes it: entry = kind << 24 | r << 12 | c, kinds as in the source's DK_* list.  Writes
 * min(cap, count) entries and the helper and early-task lookaheads; returns the count, or DH_EARG.
 * Host only (no GPU call): the schedule's ordering invariant is tested from it.  No counterpart in
 * the reference. */
int dh_dag_table(int nbt, int last, int arrow, int* out, int cap, int* look, int* slook);

/* Tuning knobs (chosen per shape by the autotuner in dalia_amd/autotune.py):
 * key 0: Schur-update kernel of the per-panel pobtaf schedule, 0 = 64x64
 *        tiles (one launch), 1 = 128x128 tiles with the k range split in two
 *        (+ a tail launch; needs an even number of 64-tiles per block);
 * key 2: pobtaf/pobtf schedule, 0 = one launch per 64-column panel (key 0
 *        applies), 1 = one persistent launch of left-looking tile tasks
 *        synchronised by per-tile readiness flags (default);
 * key 3: selected inversion, 1 = products of the factor alone on a second
 *        stream ahead of the recursion, 0 = one stream (default);
 * key 4: work run beside the persistent factorization (schedule 1, outside
 *        stream capture) on a second stream, chunk by chunk as its blocks
 *        finish: 0 = none (default), 1 = the inverse of the diagonal factors
 *        that dh_solve needs, 2 = also the products of the factor alone that
 *        dh_selected_inversion needs (setting 2 reserves per-block slots,
 *        2 n bp^2 doubles, until dh_destroy; DH_EHIP if they do not fit).
 *        dh_cholesky then returns with that work queued on the handle's
 *        stream as well;
 * key 5: 1 = the next factorizations start with the flag that a timed-out readiness wait
 *        raises already set, so every wait of the persistent launch returns at once, the
 *        launch drains with an invalid factor, dh_logdet writes NaN and dh_get_info returns -4
 *        (a test of that error path in which every kernel exits normally); 0 = off (default);
 * key 6: at most this many workgroups in the persistent launch (0 = one per CU that fits, the
 *        default); a factorization bound by its chain of diagonal tiles needs fewer, and leaves the
 *        other CUs to concurrent work of other handles (other streams).
 * key 7: 1 = two-ended factorization (n >= 3, schedule 1, key 4 = 0; DH_EARG for n < 3): blocks
 *        0 .. m-1 are eliminated downward and blocks n-1 .. m+1 upward in one persistent launch,
 *        m = (n - 1) / 2, then block m and the tip, which halves the chain of diagonal tiles.
 *        The handle's buffers then hold the Cholesky factor of the matrix with its blocks ordered
 *        0 .. m-1, n-1 .. m+1, m: diag[i] and arrow[i] the factor of block i, lower[i] (i >= m)
 *        the coupling of blocks i and i + 1 with the rows of block i.  dh_logdet, dh_solve and
 *        dh_selected_inversion (whose result is in the usual layout) follow that factor;
 *        dh_assemble writes blocks m .. n-2 of the sub-diagonal transposed for it, and a
 *        factorization that is not two-ended (or dh_multi_cholesky) transposes them back
 *        first.  0 = one-ended (default).
 * key 8: operands of the persistent launch's products: 0 = read along the factor's rows, 1 = read
 *        by columns from transposed copies of the final tiles that the launch also writes (they
 *        take 2 n bp^2 + 64 n bp doubles, reserved on first use and kept until dh_destroy),
 *        2 = by columns in two-ended launches only (default).  Columns are faster for products in
 *        bulk and slower on the chain of diagonal tiles, so a one-ended launch, which that chain
 *        bounds, reads by rows.  dh_multi_cholesky reads by columns unless a handle has 0.
 * key 9: 1 = zero-tile profiles (default): before each persistent launch, every sub-diagonal
 *        block as stored is scanned for the first tile column of each 64-row tile row that
 *        holds a nonzero entry; the tiles left of it stay exactly zero in the factor, and the
 *        launch skips every product with them (a banded spatial ordering leaves about half of
 *        the factorization's tile products out).  The factor is the same.  0 = every tile is
 *        treated as nonzero.  dh_multi_cholesky follows the setting its handles had when
 *        dh_multi_create ran.
 * key 10: 1 = dh_selected_inversion in congruence form, for callers that read the selected
 *        inverse on the matrix's pattern only: X_ii = inv(L_ii)^T (I + M) inv(L_ii) with
 *        M = L^T Y + L_a^T Ya, Y = X_qq L + X_aq^T L_a, Ya = X_aq L + X_aa L_a (L = L_{q,i}, q the
 *        successor), X_{a,i} = -Ya inv(L_ii), and X_{q,i} = -Y inv(L_ii) only on the 64x64 tiles
 *        of each coupling block between the first and last tile of a tile row where the matrix
 *        has entries (the others keep stale values).  Products with L skip its zero tiles.  Needs
 *        the profiles of the current factor (key 9); otherwise, and with key 4 = 2, the default
 *        form runs.  0 = the whole blocks (default, as the reference's selected inversion).
 * key 11: 1 = dh_assemble records the zero-tile profiles of the coupling blocks (key 9) while it
 *        writes them (the scan's test, v != 0.0, on the rows it builds), and the next
 *        dh_cholesky uses them instead of scanning the blocks.  Valid for that one factorization;
 *        dh_scatter_coo, dh_assemble_coreg, a change between the one- and two-ended layouts and
 *        dh_multi_cholesky drop them.  Writes to the blocks by other means in between are not
 *        seen.  0 = scan at each factorization (default).
 * key 12: 1 = the first dh_solve after a factorization forms the inverse of the diagonal factors
 *        on the handle's second stream, in chunks of blocks in the order its forward sweep reaches
 *        them (from both ends inward for a two-ended factor), and the sweep waits for each chunk's
 *        event before its first step that reads it; the calling stream has waited for every chunk
 *        when dh_solve returns.  Not under stream capture (the inverse then runs first on the
 *        calling stream).  Bitwise the same results.  0 = the inverse first (default).
 * key 13: 1 = dh_solve records events during its backward sweep (after its first two steps, every
 *        fourth and the last), and a
 *        dh_selected_inversion that follows it (congruence form, key 10) runs on the handle's second
 *        stream: after everything the calling stream held before that sweep, each recursion step
 *        after the first recorded sweep step at or after the one that reads the blocks it
 *        overwrites; the calling stream waits for
 *        the recursion before the call returns.  Work that the caller enqueues on its stream
 *        between the two calls runs beside the recursion, so it must not touch the factor's
 *        storage: the entry points that do (dh_logdet, dh_marginal_variances, dh_obs_variances,
 *        the dh_contract_fem / dh_contract_ata forms, the assemblies, dh_scatter_coo, dh_cholesky,
 *        dh_multi_cholesky, dh_set_tuning) cancel the arrangement, and the recursion then runs on
 *        the calling stream after them; reads through the block views by other means are not
 *        seen.  Not under stream capture.  Bitwise the same results.  0 = off (default).
 * Other keys or values return DH_EARG.  No counterpart in the reference. */
int dh_set_tuning(dh_handle* h, int key, int value);

/* Per-kernel timing with HIP events on the handle's stream, for the roofline
 * of one kernel.  kind: 0 k_syrk (Schur update of the next diagonal block in
 * pobtaf; k_syrk128 for tuning value 1), 1 k_update (panel trailing update),
 * 2 pobtasi tile launches (k_sel_w, k_sel_x, k_sel_s), 3 k_panel, 4 k_leaf,
 * 5 k_assemble_rows, 6 k_syrk_tail, 7 k_pobtaf_dag (the persistent pobtaf
 * launch); kind < 0 disables.  While enabled,
 * each launch of that kind is bracketed by an event pair (at most 16384
 * launches are recorded).  dh_timing_read synchronises the stream and returns
 * the summed kernel time and the number of recorded launches.  No
 * counterpart in the reference (its timers are host wall clock,
 * structured_solver.py:80-100). */
int dh_timing_enable(dh_handle* h, int kind);
int dh_timing_read(dh_handle* h, double* total_ms, long long* count);

#ifdef __cplusplus
}
#endif
#endif
