/*
 * rydhip — C ABI of the MI355X (gfx950) matrix-free Rydberg Hamiltonian engine.
 *
 *   H = Σ_i Ω_i/2 σx_i + Σ_i delta_i n_i + Σ_{i<j} V_ij n_i n_j
 *
 * (reference sign convention: the diagonal carries +delta_i n_i, so the
 * Hamiltonian term −Δ_i n_i is delta = −Δ; complex Ω: g→r takes Ω/2, r→g conj(Ω)/2).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns a
 * status code (RYD_OK = 0) and never throws.  The caller allocates and owns
 * every state vector, workspace and result buffer (device memory); a plan
 * owns only its immutable device tables, so one plan may be used from
 * several threads.  Calls that take a workspace must not share that
 * workspace between concurrently running calls.  `stream` is a hipStream_t
 * (NULL = default stream).  After a failure, ryd_last_error() returns the
 * message of the calling thread's last error.
 *
 * Element types (dtype): RYD_F32 float, RYD_F64 double, RYD_C64 complex64,
 * RYD_C128 complex128 (interleaved re, im).
 *
 * Bit order: site i is bit (n-1-i) of a basis code (site 0 = MSB), bit 1 = |r>.
 */
#ifndef RYDHIP_H
#define RYDHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RYD_ABI_VERSION 12

enum ryd_status {
  RYD_OK = 0,
  RYD_ERR_INVALID = 1,     /* bad argument (maps to ValueError)          */
  RYD_ERR_MEMORY = 2,      /* device allocation failed (MemoryError)     */
  RYD_ERR_DEVICE = 3,      /* HIP runtime / no device (RuntimeError)     */
  RYD_ERR_UNSUPPORTED = 4  /* valid request this build does not support  */
};

enum ryd_dtype { RYD_F32 = 0, RYD_F64 = 1, RYD_C64 = 2, RYD_C128 = 3 };

enum ryd_basis_kind {
  RYD_BASIS_FULL = 0,       /* all 2^n codes                                */
  RYD_BASIS_CODES = 1,      /* explicit ascending int64 codes (number
                               sectors, any subset of codes)               */
  RYD_BASIS_CHAIN_FIB = 2,  /* chain nearest-neighbour blockade sector      */
  RYD_BASIS_GRAPH = 3,      /* independent sets of a blockade graph (edges
                               below), enumerated on the device in
                               ascending order as independent_states,
                               basis.py:24-47; n <= 60, no 2^20 cap        */
  RYD_BASIS_CSR = 4         /* a caller's sparse matrix (ryd_plan_create_csr) */
};

/* Model description (host memory, copied at plan creation). */
typedef struct ryd_model {
  int32_t n;               /* number of sites, 1..62 (full basis 1..40)     */
  const double* om_re;     /* [n] Re Ω_i (MHz)                              */
  const double* om_im;     /* [n] Im Ω_i (MHz); NULL = real drive           */
  const double* delta;     /* [n] diagonal coefficient of n_i (MHz)         */
  int32_t n_pairs;
  const int32_t* pair_i;   /* [n_pairs] site i                               */
  const int32_t* pair_j;   /* [n_pairs] site j                               */
  const double* pair_v;    /* [n_pairs] V_ij (MHz), duplicates are summed    */
  int32_t basis_kind;      /* ryd_basis_kind                                 */
  const int64_t* codes;    /* RYD_BASIS_CODES: ascending unique codes (host) */
  int64_t n_codes;
  /* Pair flip channels (channels.py:119-131), NULL = none: J exchange
   * |r_i g_j> -> |g_i r_j> takes conj(J), the reverse J; A pair creation
   * |g_i g_j> -> |r_i r_j> takes A, the reverse conj(A).  A model with pair
   * flips runs the generic channel kernel (full basis or codes). */
  const double* pair_jr;   /* [n_pairs] Re J_ij                              */
  const double* pair_ji;   /* [n_pairs] Im J_ij                              */
  const double* pair_ar;   /* [n_pairs] Re A_ij                              */
  const double* pair_ai;   /* [n_pairs] Im A_ij                              */
  /* RYD_BASIS_GRAPH: blockade edges (site pairs that may not both be
   * excited), as blockade_edges of build_plan, hamiltonian_plan.py:36-72 */
  int32_t n_edges;
  const int32_t* edge_i;   /* [n_edges] */
  const int32_t* edge_j;   /* [n_edges] */
} ryd_model;

typedef struct ryd_plan ryd_plan;

/* Library ABI version (RYD_ABI_VERSION).  No reference counterpart. */
int ryd_abi_version(void);

/* Copies the calling thread's last error message into buf (NUL-terminated). */
int ryd_last_error(char* buf, size_t size);

/* Builds the device tables of a model on the current device.
 * Replaces the per-call channel enumeration and transition tables of the
 * reference: src/core/manybody/channels.py:91-132 (channel_spec),
 * :135-175 (ChannelKernel.__init__) and matrix_free.py:65-115
 * (_channel_arrays / _group_channels, uploaded at matrix_free.py:185-198). */
int ryd_plan_create(const ryd_model* model, ryd_plan** out);

/* A plan over a caller's sparse Hermitian matrix H (CSR, host arrays: row
 * pointers indptr[dim+1] from 0, column indices[nnz] in [0, dim), values
 * val_re/val_im[nnz], val_im NULL for a real H), copied to the device.
 * Replaces the COO upload of kpm_moments_gpu (src/core/manybody/kpm.py:26-75,
 * `torch.sparse_coo_tensor` at :40-47), which takes an assembled H: the plan
 * serves ryd_hv, ryd_apply, ryd_lanczos, ryd_lanczos_stored and
 * ryd_cheb_moments; the entry points that need site codes (occupations,
 * Pauli strings, correlations, codes/ranks, evolution) refuse it.
 * dim must be below 2^31. */
int ryd_plan_create_csr(int64_t dim, const int64_t* indptr, const int64_t* indices,
                        const double* val_re, const double* val_im, ryd_plan** out);
int ryd_plan_destroy(ryd_plan* plan);

/* Basis dimension of a plan (0 on a NULL plan). */
int64_t ryd_plan_dim(const ryd_plan* plan);
/* How the plan ranks codes: 0 identity (full basis), 1 combinatorial number
 * system, 2 binary search over the code list, 3 blockade-graph tables,
 * 4 chain-sector Zeckendorf rank.  No reference counterpart (diagnostic). */
int ryd_plan_rank_mode(const ryd_plan* plan);

/* Bytes of workspace any single reducing call on this plan needs (per-block
 * partial sums; no initialisation required). */
size_t ryd_workspace_bytes(const ryd_plan* plan);

/* y = H x.  Replaces ChannelKernel.matvec (channels.py:210-221) and the GPU
 * action() of torch_lanczos_ground_state (matrix_free.py:200-204). */
int ryd_hv(const ryd_plan* plan, int dtype, const void* x, void* y, void* stream);

/* Fused update y = ch·H(s·x) + cx·(s·x) + cz·z (z may be NULL), with FP64
 * reductions {Re<s·x,y>, |y|^2, |s·x|^2, Σ_e p_e Re(y_e)} (p_e = ±1 from the
 * optional bit-packed probe, bit set = +1) stored to sums[4] (accumulate = 0)
 * or added to them (accumulate = 1).  sums may be NULL (no reductions).
 * Covers the Chebyshev step of kpm.py:146-148 / :69-72. */
int ryd_apply(const ryd_plan* plan, int dtype, const void* x, const void* z, void* y,
              double s, double ch, double cx, double cz, const uint32_t* probe,
              double* sums, int accumulate, void* ws, void* stream);

/* Three-term Lanczos recurrence, steps j0 .. j0+steps-1, entirely on the
 * device (no host synchronisation per step).  Replaces the Krylov build loop
 * of torch_lanczos_ground_state (matrix_free.py:231-251) and the recurrence of
 * spectrum.lanczos (spectrum.py:76-96).
 *   b0, b1, b2 : three vectors of dim elements; at j0 = 0, b0 holds the start
 *                vector (any norm).  Vector u_j lives in b[j % 3]; on return
 *                the last two (u_{j0+steps-1}, u_{j0+steps}, or the residual
 *                at a breakdown) are there, the third buffer is scratch
 *                (full-basis plans up to n = 13, complex128 12, run every
 *                step of a call in one launch).
 *   alpha, nrm : device FP64 arrays (>= j0+steps and >= j0+steps+1 entries);
 *                alpha_j = alpha[j], beta_j = nrm[j+1], nrm[0] = |start|.
 *   state      : device buffer of ryd_lanczos_state_bytes(); stop flag and
 *                step count after a breakdown (|u_{j+1}| <= tiny).
 *   replay = 1 : recompute the same vectors from stored alpha/nrm and add
 *                Σ_j yk[j]/nrm[j]·u_j to ritz0/ritz1 (either may be NULL;
 *                y0/y1 device FP64);
 *                b0 must again hold the start vector at j0 = 0. */
int ryd_lanczos(const ryd_plan* plan, int dtype, void* b0, void* b1, void* b2, int64_t j0,
                int64_t steps, double* alpha, double* nrm, void* state, double tiny,
                void* ritz0, void* ritz1, const double* y0, const double* y1, int replay,
                void* ws, void* stream);
size_t ryd_lanczos_state_bytes(void);
/* Lanczos from a source vector with full reorthogonalisation, the
 * recurrence of spectrum.lanczos with reortho=True
 * (src/core/manybody/spectrum.py:34-103): on entry row 0 of `basis` (k rows,
 * row stride ld >= dim, the plan's element type) holds phi; it is scaled by
 * phi_scale (= 1/|phi|) in place and rows 1..k-1 receive the unit Lanczos
 * vectors.  Step j: w = H q_j, alpha[j] = Re<q_j, w>, w -= alpha_j q_j +
 * beta_{j-1} q_{j-1}, then `passes` classical Gram-Schmidt sweeps of w
 * against rows 0..j (the reference sweeps twice), nrm[j+1] = beta_j = |w|,
 * q_{j+1} = w / beta_j.  The run stops when beta_j <= tiny (absolute; the
 * rows are unit vectors) and `state` (ryd_lanczos_state_bytes) records the
 * steps, read with ryd_lanczos_status.  w: one vector (the last step's
 * residual on return); gs_ws: ryd_reortho_workspace_bytes(dim, k) bytes of
 * device memory; ws: the plan's workspace.  No host synchronisation. */
size_t ryd_reortho_workspace_bytes(int64_t dim, int64_t k);
int ryd_lanczos_reortho(const ryd_plan* plan, int dtype, void* basis, int64_t ld, void* w,
                        int64_t k, double phi_scale, double* alpha, double* nrm, void* state,
                        double tiny, int passes, void* gs_ws, void* ws, void* stream);

/* ryd_lanczos_reortho with nlock locked rows in front of the start vector:
 * rows [0, nlock) of `basis` hold orthonormal vectors that every
 * Gram-Schmidt sweep includes but the recurrence never uses; row nlock holds
 * phi (scaled by phi_scale, made orthogonal to the locked rows by two sweeps
 * and normalised again), rows nlock+1 .. nlock+k-1 receive the Lanczos
 * vectors.  The tridiagonal is then that of H restricted to the orthogonal
 * complement of the locked rows: with the converged ground vector locked its
 * lowest Ritz value is E1, or E0 again when the ground state is degenerate,
 * the multiplicity the reference's two-vector block Krylov
 * (matrix_free.py:219-266) and eigh (equilibrium.py:65-75, :272) report.
 * gs_ws: ryd_reortho_workspace_bytes(dim, nlock + k). */
int ryd_lanczos_reortho_locked(const ryd_plan* plan, int dtype, void* basis, int64_t ld,
                               int64_t nlock, void* w, int64_t k, double phi_scale,
                               double* alpha, double* nrm, void* state, double tiny, int passes,
                               void* gs_ws, void* ws, void* stream);

/* Reads the state written by ryd_lanczos (synchronises the stream). */
/* The recurrence of ryd_lanczos in the orthogonal complement of a unit
 * vector `lock` (device, plan dtype; e.g. a converged ground vector): step j
 * leaves u_pre = w − (alpha_j/nrm_j) u_j, stores γ_j = <lock, u_pre> to
 * gam[2j], gam[2j+1] (device FP64, >= 2·(j0+steps) entries; imaginary part
 * 0 for real types) and continues with u_{j+1} = u_pre − γ_j lock, nrm[j+1]
 * = |u_{j+1}|.  Removing the lock component of the finished residual keeps
 * <lock, u_j> at rounding level (removing only <lock, w> lets alpha_j·
 * <lock, q_j> grow every step).  The start vector in b0 must be orthogonal
 * to lock.  replay = 1 reads the stored alpha, nrm and gam and accumulates
 * Ritz vectors as ryd_lanczos does (y0, y1: device FP64 coefficients).  The lowest genuine Ritz value is then
 * the next level counted with multiplicity, as eigh's w[1]
 * (src/core/manybody/equilibrium.py:65-75, :272) and the reference's
 * two-vector block Krylov (matrix_free.py:219-266) report it; O(steps)
 * instead of the O(steps²·dim) of ryd_lanczos_reortho_locked. */
int ryd_lanczos_locked(const ryd_plan* plan, int dtype, void* b0, void* b1, void* b2,
                       const void* lock, int64_t j0, int64_t steps, double* alpha, double* nrm,
                       double* gam, void* state, double tiny, void* ritz0, void* ritz1,
                       const double* y0, const double* y1, int replay, void* ws, void* stream);
int ryd_lanczos_status(const void* state, int* stopped, int64_t* steps, void* stream);

/* The same recurrence keeping every Lanczos vector (no replay for Ritz
 * vectors; the reference keeps its block Krylov basis too,
 * matrix_free.py:219-251).  basis: device chunk of rows of ld >= dim
 * elements, u_j at basis + (j - jbase)·ld for the rows this call touches,
 * j in [max(j0-1, jbase), j0 + steps]; at j0 = 0 row 0 holds the start
 * vector.  A chunk that starts a later cycle repeats u_{j0-1}, u_{j0}. */
int ryd_lanczos_stored(const ryd_plan* plan, int dtype, void* basis, int64_t ld, int64_t jbase,
                       int64_t j0, int64_t steps, double* alpha, double* nrm, void* state,
                       double tiny, void* ws, void* stream);
/* out0 (+)= Σ_{j<count} (y0[j] / nrm[j0+j]) u_{j0+j}, likewise out1 with y1
 * (may be NULL), over rows basis + j·ld (count <= 2048; y0/y1 host arrays;
 * accumulate = 0 overwrites).  ws: >= 256 + 32768 bytes.  Returns after the
 * launch has been ordered on `stream` and the coefficients copied. */
int ryd_ritz_stored(int dtype, const void* basis, int64_t ld, int64_t dim, int64_t count,
                    const double* y0, const double* y1, const double* nrm, int64_t j0,
                    int accumulate, void* out0, void* out1, void* ws, void* stream);

/* Chebyshev recurrence of one Rademacher probe on H~ = (H - c)/r:
 * t_0 = p, t_1 = H~ t_0, t_k = 2 H~ t_{k-1} - t_{k-2}, k <= nsteps.
 * For each k >= 1 adds {<t_{k-1},t_k>, |t_k|^2, |t_{k-1}|^2, <p,t_k>} to
 * sums[4k .. 4k+3] (device FP64, (nsteps+1)*4 entries).  Replaces the probe
 * loop of kpm.kpm_moments (kpm.py:139-149) and kpm_moments_gpu (kpm.py:62-73). */
int ryd_cheb_moments(const ryd_plan* plan, int dtype, const uint32_t* probe, void* t0,
                     void* t1, void* t2, int64_t nsteps, double c, double r, double* sums,
                     void* ws, void* stream);

/* Allocation-free form of ryd_cheb_moments (same arguments without the
 * workspace, same results bit for bit): the reduction partials live in a
 * device array the library holds for the life of the process, 8 MiB per GPU
 * (plans whose ryd_workspace_bytes exceeds that, e.g. a full basis beyond
 * n = 31, are refused with RYD_ERR_INVALID).  No counterpart in the
 * reference; it is for callers that cannot allocate device memory: code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  It calls no allocating routine.  The plan's device must be
 * current.  All calls on one GPU share that GPU's array: calls that overlap
 * on one GPU race.  sums accumulate as in ryd_cheb_moments (the caller
 * clears them). */
int ryd_cheb_moments_noalloc(const ryd_plan* plan, int dtype, const uint32_t* probe, void* t0,
                             void* t1, void* t2, int64_t nsteps, double c, double r,
                             double* sums, void* stream);

/* out[i] = Σ_s |x_s|^2 b_i(s) for every site i, out[n] = Σ_s |x_s|^2 (device FP64,
 * n+1 entries; divide by out[n] for <n_i>).  Replaces _gs_observables' n_atom
 * (eq_route.py:108-122) and the bits @ |psi|^2 reduction of
 * engine_ed._reduce (engine_ed.py:261-263).  Full basis and chain sector. */
int ryd_site_occupation(const ryd_plan* plan, int dtype, const void* x, double* out, void* ws,
                        void* stream);

/* Pauli-string expectations <x|P_q|x> for q < count on the plan's basis
 * (any kind).  String q: flip[q] = X|Y sites, sign[q] = Y|Z sites, as bit
 * masks by bit position (site i = bit n-1-i), ny[q] = number of Y sites;
 * out (device FP64, [count][2] = re, im).  The basis path of
 * PauliPlan.expect (observables.py:564-622): targets outside the basis
 * contribute nothing.  ws: the plan's workspace (ryd_workspace_bytes). */
int ryd_pauli_expect(const ryd_plan* plan, int dtype, const void* x, int count,
                     const uint64_t* flip, const uint64_t* sign, const int32_t* ny,
                     double* out, void* ws, void* stream);

/* Site correlation sums on the plan's basis: out[i*n + j] = Σ_s |x_s|^2
 * b_i(s) b_j(s) (both triangles), out[n*n] = Σ_s |x_s|^2 (device FP64,
 * n*n + 1 entries); C_ij = out[i,j]/out[n*n] - <n_i><n_j> as corr_matrix
 * (observables.py:40-47).  scratch: device buffer of scratch_bytes >=
 * (n(n+1)/2 + 1)·8 bytes (more bytes: more blocks, up to 512). */
int ryd_site_correlations(const ryd_plan* plan, int dtype, const void* x, double* out,
                          void* scratch, size_t scratch_bytes, void* stream);

/* Restricted-basis enumeration (chain sector, codes and blockade-graph
 * plans): codes[k] = code of rank ranks[k] (or of rank first + k when ranks
 * is NULL); ranks outside [0, dim) give -1.  Codes come out in the
 * reference's basis order, ascending (independent_states,
 * src/core/manybody/basis.py:24-47; chain: edges (i, i+1), rank(code) =
 * Σ_{site i excited} F(n-i+1)).  Device int64 arrays. */
int ryd_sector_codes(const ryd_plan* plan, const int64_t* ranks, int64_t first, int64_t count,
                     int64_t* codes, void* stream);

/* ranks[k] = rank of codes[k] in the plan's restricted basis, -1 if the code
 * is not in it (replaces the position table / searchsorted lookup of
 * ChannelKernel._indices, channels.py:177-192).  Device int64 arrays. */
int ryd_sector_rank(const ryd_plan* plan, const int64_t* codes, int64_t count, int64_t* ranks,
                    void* stream);

/* Piecewise-constant evolution psi <- Π_k exp(-i·CONV·H_k·dt_k) psi
 * (complex dtypes), each slice by a Chebyshev series:
 *   H_k = om[k]·F + D + dadd[k]·N   (F: the plan's Ω flips, D: its diagonal,
 *                                    N = Σ_i n_i),
 *   psi_out = Σ_{j=0..K} a_kj T_j((H_k - c[k])/r[k]) psi_in,
 * coef = host FP64 [nslices][K+1][re, im] (a_kj, phase and Bessel factors
 * included).  work: 4 vectors of dim elements.  Full basis with Ω/Δ/V and
 * n <= 22: fused slice kernels, any om/dadd.  Every other plan (restricted
 * bases: chain sector, code lists, blockade graphs; J/A models; full basis
 * n > 22) evolves under its own H: om[k] = 1 and dadd[k] = 0, else
 * RYD_ERR_UNSUPPORTED.  Replaces the per-interval expm_multiply of
 * engine_ed._propagate_segments / _propagate_sparse (engine_ed.py:117-147),
 * also on a projected plan's basis_indices (engine_ed.py:191-192, :288). */
int ryd_evolve_forward(const ryd_plan* plan, int dtype, void* psi, int64_t nslices,
                       const double* om, const double* dadd, const double* c, const double* r,
                       int K, const double* coef, void* work, void* stream);

/* psi <- exp(-tau·(i·H + G/2)) psi = exp(-i·CONV·dt·(H - (i/2)·G)) psi with
 * tau = CONV·dt, for the plan's H on its basis and the loss diagonal
 * G = Σ_i gamma[i]·n_i (host FP64 [n], by site; MHz like delta, i.e. the
 * reference's rate γ_i·1e-3/CONV): the non-Hermitian
 * (dissipative) drive of engine_ed._propagate_segments / _propagate_sparse
 * (engine_ed.py:117-147, expm_multiply of −i·CONV·H with the −(i/2)Γ_i n_i
 * diagonal of channels.py:58-77).  nsub Taylor steps of degree m, each term
 * one H·v and one fused update; the caller picks nsub so that tau/nsub times
 * a bound of ‖H − (i/2)G‖ stays below 1 (the truncation is then below
 * 1/(m+1)! per step).  Complex dtypes; work: 2 vectors of dim elements. */
int ryd_evolve_taylor(const ryd_plan* plan, int dtype, void* psi, double tau, int nsub, int m,
                      const double* gamma, void* work, void* stream);

/* y += -(i/2)·G·x, G = Σ_i gamma[i]·n_i over the plan's basis (host FP64
 * [n], by site, MHz like delta): with ryd_hv, the action of the
 * non-Hermitian H of channel_spec with gamma != 0 (channels.py:58-77, the
 * −(i/2)Γ_i n_i diagonal) for hamiltonian_linear_operator
 * (matrix_free.py:51-62).  Complex dtypes. */
int ryd_loss_axpy(const ryd_plan* plan, int dtype, const void* x, void* y, const double* gamma,
                  void* stream);

/* Vector-Jacobian product of one slice of ryd_evolve_forward (the adjoint;
 * no reference counterpart — the reference has no gradient).  Given the
 * slice input psi and g = dL/d(conj psi_out), overwrites g with
 * dL/d(conj psi_in) and stores for j = 0..K-1 the FP64 sums
 * {Re<F b_{j+1}, t_j>, Re<N b_{j+1}, t_j>, |b_j|^2, Re<t_j, b_j>} at
 * sums[4j..4j+3] (t_j: Chebyshev vectors, b_j: Clenshaw adjoints), from which
 * dL/d om = (2/r)(S_0 + 2 Σ_{j>=1} S_j)[0] and likewise for dadd with [1].
 * tstore: K vectors; work: 3 vectors; coef: host [K+1][re, im] of this slice. */
int ryd_evolve_adjoint(const ryd_plan* plan, int dtype, const void* psi, void* g, double om,
                       double dadd, double c, double r, int K, const double* coef, void* tstore,
                       void* work, double* sums, void* ws, void* stream);

/* ryd_evolve_forward (fused slice kernels only: full basis, Ω/Δ/V,
 * n <= 22) that also keeps each slice's Chebyshev vectors for the adjoint:
 * t_0..t_{K-1} of slice k (t_0 = the slice's input state) at
 * tstore + k·K·dim elements.  tstore: nslices·K vectors; work: 1 vector.
 * No reference counterpart (the reference has no gradient); it is the
 * recompute step of a checkpointed adjoint. */
int ryd_evolve_forward_store(const ryd_plan* plan, int dtype, void* psi, int64_t nslices,
                             const double* om, const double* dadd, const double* c,
                             const double* r, int K, const double* coef, void* tstore,
                             void* work, void* stream);

/* Bytes of the partials buffer ryd_evolve_adjoint_stored needs for K terms
 * (-1 for a plan or dtype the fused kernels do not take).  No reference
 * counterpart. */
int64_t ryd_evolve_partials_bytes(const ryd_plan* plan, int dtype, int K);

/* Adjoint of nslices slices stored by ryd_evolve_forward_store, last slice
 * first (no reference counterpart — the reference has no gradient).  g =
 * dL/d(conj psi_out) of the last slice on entry, dL/d(conj psi_in) of the
 * first on exit.  For slice k and term j = 0..K-1 the FP64 sums of
 * ryd_evolve_adjoint go to sums[(k·K + j)·4 ..] in DEVICE memory (no host
 * round trip per slice).  work: 3 vectors; partials: device,
 * ryd_evolve_partials_bytes; coef, om, dadd, c, r: host arrays as in
 * ryd_evolve_forward. */
int ryd_evolve_adjoint_stored(const ryd_plan* plan, int dtype, void* g, int64_t nslices,
                              const double* om, const double* dadd, const double* c,
                              const double* r, int K, const double* coef, const void* tstore,
                              void* work, double* sums, double* partials, void* stream);

/* Batched Lanczos over many independent full-basis instances of uneven n
 * (1 <= n <= 24, real drives), FP32 or FP64 vectors with FP64 alpha/beta.
 * One segmented launch per Lanczos half-step covers every instance; the
 * whole recurrence is captured once into a hipGraph and replayed
 * (use_graph = 1).  Replaces the dense batched eigh of batched_ground_state
 * (src/core/manybody/equilibrium.py:251-311), which needs uniform dim, and
 * the serial loop of run_ensemble (ensemble.py:51-175).
 * The batch object owns its Krylov workspace and a default-priority stream
 * (so that the captured graph has fixed addresses); it is not thread-safe.
 * Start vectors are concatenated in instance order; ryd_batch_offset(b, i)
 * gives instance i's offset (i = count gives the total length). */
typedef struct ryd_batch ryd_batch;
int ryd_batch_create(const ryd_model* models, int32_t count, int dtype, int32_t steps,
                     ryd_batch** out);
/* A batch over caller-supplied sparse Hermitian instances of any size
 * (e.g. the dense stack of batched_ground_state, equilibrium.py:251-311,
 * passed without its zeros): CSR host arrays of the concatenated rows,
 * indptr[total+1] from 0 (total = Σ dims), indices[nnz] the column inside
 * the row's own instance (0 .. dims[b]-1), val_re/val_im[nnz] (val_im NULL
 * for real values; complex values need dtype C64 or C128).  Vectors in
 * dtype (F32, F64, C64, C128) with FP64 alpha/beta; tiles of rows (1024
 * rows, one per thread, or 64 rows, one per wavefront, when rows average 16
 * or more entries).  Every other ryd_batch_* entry point applies except the
 * tile-size ones (autotune, set_tile_bits: RYD_ERR_UNSUPPORTED); occupations
 * are reported for instances whose dim is 2^n (n <= 24), bit n-1-i of the
 * row index being site i.  Replaces the reference's dense batched eigh for
 * instances of any Hamiltonian.  dims[b] < 2^31. */
int ryd_batch_create_csr(int32_t count, const int64_t* dims, const int64_t* indptr,
                         const int64_t* indices, const double* val_re, const double* val_im,
                         int dtype, int32_t steps, ryd_batch** out);
/* CSR batches: per-instance lock vectors g_b (device, the batch element type,
 * concatenated like the start vectors, unit norm), or NULL.  Each step then
 * removes the component <g_b, w> g_b, so the recurrence runs in the
 * complement of g_b: with g_b a converged ground vector the lowest Ritz value
 * is the second level counted with multiplicity, as eigh's w[1]
 * (equilibrium.py:272).  Start vectors must be orthogonal to g_b.  The
 * memory stays the caller's; drops the captured graph.  No reference
 * counterpart (the reference diagonalises densely). */
int ryd_batch_set_lock(ryd_batch* batch, const void* lock);
int ryd_batch_destroy(ryd_batch* batch);
int64_t ryd_batch_offset(const ryd_batch* batch, int32_t index);
/* Runs the recurrence from the device start vectors v0, ordered after and
 * before `stream`. */
int ryd_batch_run(ryd_batch* batch, const void* v0, int use_graph, void* stream);
/* Steps [j0, j1) of the same recurrence without a graph (full-basis batches):
 * j0 = 0 starts it from v0, a later call continues where the last one ended,
 * so the caller can stop converged instances between parts
 * (ryd_batch_stop).  Replaces the fixed-length run of the reference's dense
 * batch (equilibrium.py:251-311) for certified solves: an instance leaves
 * its cycle at the step its Ritz pairs pass, not at the cycle's end. */
int ryd_batch_run_steps(ryd_batch* batch, const void* v0, int32_t j0, int32_t j1, void* stream);
/* Sets the stop flag of every instance b with flags[b] != 0 (host array of
 * `count`): its recurrence stays at the step reached. */
int ryd_batch_stop(ryd_batch* batch, const int32_t* flags);
/* Times the recurrence for tiles of 2^10, 2^11, 2^12 codes and keeps the
 * fastest; ms_out (host, 3 entries) may be NULL. */
int ryd_batch_autotune(ryd_batch* batch, const void* v0, int32_t* chosen_tile_bits,
                       double* ms_out);
/* Host only (no device work): the two lowest genuine Ritz pairs of `count`
 * Lanczos tridiagonals, replacing the dense tridiagonal eigensolve and the
 * ghost filter the reference's Rayleigh-Ritz step does not need
 * (matrix_free.py:251-266 diagonalises its projected block instead).
 * Tridiagonal b: diagonal alpha[b*lda + i] (i < steps[b]), off-diagonal
 * offd[b*ldo + i] (i < steps[b]-1), beta_last[b] = the next Lanczos norm
 * (0 after a breakdown).  Eigenvalues by Sturm bisection; a value within
 * merge_rel·|T| of the previous kept one merges (a ghost copy: 1e-10 for a
 * recurrence in FP64 vectors, 1e-6 in FP32 ones), and the kept value takes
 * the bound and eigenvector of the copy with the smallest bound; a simple value
 * that T without its first row and column also has is spurious
 * (Cullum-Willoughby) unless its bound |beta_last · y[k-1]| is at most
 * conv_tol·max(1,|E|); the lowest value is always kept.  Outputs (host):
 * E[b*2 + q], bound[b*2 + q] (q = 0: E0, 1: E1; NaN / inf when absent) and,
 * if y is non-null, the unit eigenvectors y[(b*2 + q)*ldy + i], i < steps[b]. */
int ryd_ritz_lowest(int32_t count, const int32_t* steps, const double* alpha, int64_t lda,
                    const double* offd, int64_t ldo, const double* beta_last, double conv_tol,
                    double merge_rel, double* E, double* bound, double* y, int64_t ldy);
/* Bytes of a whole Krylov basis of the batch: (steps + 1) rows of all
 * instances in the batch's element type.  No reference counterpart. */
int64_t ryd_batch_basis_bytes(const ryd_batch* batch);
/* Lends the batch device memory (16-byte aligned, at least
 * ryd_batch_basis_bytes) for its whole Krylov basis: the recurrence then
 * keeps every row, and ryd_batch_ritz combines the rows of the last run
 * (steps <= 2048) instead of replaying it.  basis = NULL returns to the
 * batch's own three rotating rows.  The memory stays the caller's: it must
 * outlive its use (detach or destroy the batch first); a caching allocator's
 * block avoids a driver allocation and release of tens of GB per batch.
 * Drops the captured graph.  No reference counterpart. */
int ryd_batch_attach_basis(ryd_batch* batch, void* basis, int64_t bytes);
/* Uses tiles of 2^tile_bits codes (10, 11 or 12) without timing, e.g. the
 * autotuner's choice for another batch of the same instances (an FP64
 * refinement of an FP32 bulk run).  No reference counterpart. */
int ryd_batch_set_tile_bits(ryd_batch* batch, int32_t tile_bits);
/* Host copies: alpha [count][steps], nrm [count][steps+1] (beta_j = nrm[j+1]),
 * stop [count] (1 where the recurrence broke down). */
int ryd_batch_results(const ryd_batch* batch, double* alpha, double* nrm, int32_t* stop);
/* Ritz vectors and site occupations of every instance (the n_atom of
 * batched_ground_state, equilibrium.py:251-311, there from dense
 * eigenvectors).  With a lent basis, combines the rows of the last run
 * (v0 must be that run's start vectors; it is not read), else replays the
 * recurrence from the device start vectors v0 with the chosen tile size;
 * either way x_b = Σ_j ycoef[b][j] q_j
 * (ycoef: host [count][steps], the tridiagonal eigenvector, zero past the
 * instance's steps) into ritz (device, batch element type, concatenated like
 * v0), then occ (device, [count][25]): occ[b][i] = Σ_s |x_b(s)|^2 n_i(s) for
 * site i < n_b, occ[b][24] = |x_b|^2.  Ordered after and before `stream`. */
int ryd_batch_ritz(ryd_batch* batch, const void* v0, const double* ycoef, void* ritz,
                   double* occ, void* stream);

/* sums[0..2] = {Re<a,b>, |b|^2, |a|^2} over n elements (a may be NULL). */
int ryd_dots(int dtype, const void* a, const void* b, int64_t n, double* sums, void* ws,
             void* stream);

/* Element-wise cast between F32<->F64 or C64<->C128 (n elements). */
int ryd_cast(int dtype_in, const void* x, int dtype_out, void* y, int64_t n, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* RYDHIP_H */
