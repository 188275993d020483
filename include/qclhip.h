/*
 * qclhip — MI355X (gfx950) operator backend for the SCBA inner loop of
 * QCLNEGF.jl: a third operator set beside the reference's educational (E0)
 * and production (E1) operators, with the same argument meaning, array
 * layout and error behaviour, callable from Julia (ccall), Python (ctypes)
 * or C/C++.
 *
 * Conventions
 *  - Every field argument is a DEVICE pointer to a complex (NE,Nk,Nb,Nb)
 *    array in the reference layout: column-major, energy fastest, interleaved
 *    (re,im) doubles (Julia ComplexF64): element (e,m,a,b) at
 *    e + NE*(m + Nk*(a + Nb*b)). h is a device (Nk,Nb,Nb) complex array.
 *  - Host arrays are named *_host. Grids (eps, wE, wk) are given once to
 *    qclhip_grid_create and kept on the device.
 *  - Work is enqueued on the context's stream. Entry points that return a
 *    host scalar synchronize that stream; the others do not. Failures that
 *    the device detects (a singular Dyson block, a number functional that is
 *    not real, a Pauli-capacity violation) are latched in the context and
 *    returned by qclhip_ctx_synchronize and by the next synchronizing call.
 *  - Return codes: 0 ok; >0 one-based linear index e + NE*m + 1 of a singular
 *    Dyson block (Julia: SingularException); QCLHIP_ERR_ARGUMENT
 *    (ArgumentError), QCLHIP_ERR_DIMENSION (DimensionMismatch),
 *    QCLHIP_ERR_DOMAIN (DomainError), QCLHIP_ERR_RUNTIME (HIP failure), QCLHIP_ERR_UNSUPPORTED (a size this build does not handle).
 *  - Output fields must not alias inputs unless stated (the reference rejects
 *    aliasing with ArgumentError, production_cache.jl:308-309).
 *
 * Reference citations are paths relative to the QCLNEGF.jl v0.2.0 root.
 */
#ifndef QCLHIP_H
#define QCLHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QCLHIP_OK 0
#define QCLHIP_ERR_ARGUMENT (-1)
#define QCLHIP_ERR_DIMENSION (-2)
#define QCLHIP_ERR_DOMAIN (-3)
#define QCLHIP_ERR_RUNTIME (-4)
#define QCLHIP_ERR_UNSUPPORTED (-5)

typedef struct qclhip_ctx qclhip_ctx;
typedef struct qclhip_grid qclhip_grid;
typedef struct qclhip_shift qclhip_shift;
typedef struct qclhip_kernel qclhip_kernel;
typedef struct qclhip_hilbert qclhip_hilbert;
typedef struct qclhip_batch qclhip_batch;

/* ---- context ------------------------------------------------------------
 * One context per host thread and device; enqueues on the caller's stream
 * (a hipStream_t; NULL is the default stream) and owns the device status
 * and scratch. Replaces the per-worker workspaces of the
 * reference (production_cache.jl:63-117). */
int qclhip_ctx_create(int device, void* stream, qclhip_ctx** out);
int qclhip_ctx_destroy(qclhip_ctx* ctx);
/* Waits for the stream, returns (and clears) the latched device status. */
int qclhip_ctx_synchronize(qclhip_ctx* ctx);
const char* qclhip_ctx_last_error(const qclhip_ctx* ctx);
/* Moves the context to another stream. */
int qclhip_ctx_set_stream(qclhip_ctx* ctx, void* stream);
/* Deferred results: while dev_buffer != NULL, the entries that return host
 * scalars (current_green's out_host, number_functional, boundary_flux,
 * state_populations, the FFT round-off witness) neither synchronize nor
 * write their host outputs; they copy their values, on the stream, to
 * dev_buffer[slot...] (current_green: lambda, number, holes, a, c; the others
 * in their host order) and advance slot. qclhip_ctx_result_slot gives the
 * next slot. NULL ends the mode; every call resets slot to 0. This is what
 * makes a loop iteration capturable as a graph. */
int qclhip_ctx_defer_results(qclhip_ctx* ctx, double* dev_buffer, int capacity);
int qclhip_ctx_result_slot(const qclhip_ctx* ctx);

/* ---- graphs (the plan's graph API) ---------------------------------------
 * Everything enqueued on the context's stream (which must not be the
 * default stream) between begin and end becomes one executable graph;
 * launch replays it on the stream. No allocation may happen during capture:
 * run the captured sequence once eagerly first (it sizes the scratch), and
 * call synchronizing entries only in deferred-results mode. A graph holds
 * the context's scratch as it was at capture: launch returns
 * QCLHIP_ERR_ARGUMENT when an eager call has since reallocated that scratch
 * (capture again), or when the context is not the capturing one. */
typedef struct qclhip_graph qclhip_graph;
int qclhip_graph_begin(qclhip_ctx* ctx);
int qclhip_graph_end(qclhip_ctx* ctx, qclhip_graph** out);
int qclhip_graph_launch(qclhip_ctx* ctx, const qclhip_graph* graph);
int qclhip_graph_destroy(qclhip_graph* graph);
/* Build-time maximum basis size of the Dyson/Keldysh kernels. */
int qclhip_max_basis(void);

/* ---- grids (ModelGrids: reference/grids.jl:10-47) -----------------------
 * eps, wE: energy nodes and trapezoid weights [NE]; wk: annular momentum
 * weights [Nk]; gs: spin degeneracy g_s. */
int qclhip_grid_create(qclhip_ctx* ctx, int NE, int Nk, int Nb, int gs,
                       const double* eps_host, const double* wE_host,
                       const double* wk_host, qclhip_grid** out);
int qclhip_grid_destroy(qclhip_grid* grid);

/* ---- GF-solve precision (the plan's precision option) --------------------
 * The reference solves in FP64 only (production_operators.jl:16-66, LAPACK
 * zgetrf/zgetri). QCLHIP_PRECISION_FP32_REFINE factors each Dyson block in
 * FP32 and applies refine_steps Newton steps R = I - (D/s) X,
 * X <- X + X R with the residual in FP64; a block whose predicted residual
 * ||R_1||^(2^refine_steps) exceeds refine_tol (or whose FP32 elimination
 * meets a zero pivot) is re-solved in FP64. The Keldysh products stay FP64
 * (on the engine of qclhip_grid_set_engine). QCLHIP_PRECISION_FP32_PRODUCTS
 * adds FP32 matrix-core Keldysh products: G</G> then carry FP32 rounding
 * (~1e-7 relative), and the number functionals of the grid gate realness at
 * 1e-5 relative instead of the reference's 1e-10 (numerical_invariants.jl:5).
 * Outputs stay ComplexF64 in the same layout. Applies to retarded_green,
 * dyson_keldysh and current_green on this grid. refine_steps in [0, 4];
 * refine_tol <= 0 selects the default 1e-10. */
#define QCLHIP_PRECISION_FP64 0
#define QCLHIP_PRECISION_FP32_REFINE 1
#define QCLHIP_PRECISION_FP32_PRODUCTS 2
int qclhip_grid_set_precision(qclhip_ctx* ctx, qclhip_grid* grid, int precision,
                              int refine_steps, double refine_tol);
/* Engine of the FP64 Keldysh products: vector ALUs (default) or the FP64
 * matrix cores (results agree to rounding), or FUSED: the FP64 Dyson step
 * and the vector Keldysh products in one kernel per tile, G^R kept in LDS
 * between them (bitwise the vector engine's results; at FP32 precisions the
 * vector engine runs instead). */
#define QCLHIP_ENGINE_VECTOR 0
#define QCLHIP_ENGINE_MATRIX 1
#define QCLHIP_ENGINE_FUSED 2
int qclhip_grid_set_engine(qclhip_ctx* ctx, qclhip_grid* grid, int keldysh_engine);
/* Autotuner: times both FP64 Keldysh engines, the fused kernel and the tile shapes of the FP64
 * vector kernels (energies per workgroup: 4 or 8 for the Dyson kernel, 2 or
 * 4 for the Keldysh kernel, where the LDS of a CU holds two workgroups) on
 * synthetic blocks of the grid's shape (a sub-grid of at most 2048 energies)
 * and selects the fastest of each for the grid; the choice is cached in the
 * context per (Nb, NE, Nk). ms_host (may be NULL) receives the two engine
 * times (vector, matrix) in ms. Results do not depend on the choice.
 * Allocates temporary device memory; not during graph capture. */
int qclhip_grid_autotune(qclhip_ctx* ctx, qclhip_grid* grid, double* ms_host);
/* The autotuner's record for the grid's shape: ms_host[8] = (vector engine,
 * matrix engine, Dyson W=4, Dyson W=8, Keldysh W=2, Keldysh W=4, fused
 * Dyson + Keldysh, vector Keldysh at the chosen waves per block; 0 where the
 * shape is not possible), choice_host[4] = (engine, Dyson W, Keldysh W,
 * Keldysh waves per block); the engine is the fastest launch set.
 * ArgumentError when the shape has not been tuned. */
int qclhip_grid_tuning(qclhip_ctx* ctx, const qclhip_grid* grid, double* ms_host, int* choice_host);
/* Waves per block of the FP64 vector Keldysh kernel (0: the built-in rule,
 * one): S > 1 splits a block's rows over S waves that share its LDS tiles,
 * for N > 24 where two tiles per wave would leave one wave per SIMD.
 * Compiled: 2 or 4 for N in 25..32, 5 for 33..40, 2, 3 or 6 for 41..48;
 * ArgumentError otherwise. Results are bitwise those of one wave per block
 * (same per-element arithmetic and trace order). No counterpart in the
 * reference. */
int qclhip_grid_set_keldysh_split(qclhip_ctx* ctx, qclhip_grid* grid, int waves_per_block);
/* Sets the tile shapes by hand (0: the built-in rule); ArgumentError for a
 * shape that is not compiled or does not fit the LDS for the basis size. */
int qclhip_grid_set_tiles(qclhip_ctx* ctx, qclhip_grid* grid, int dyson_waves, int keldysh_waves);
/* Blocks of the grid's last mixed-precision Dyson call that the guard sent
 * to the FP64 re-solve (0 for an FP64 grid). Synchronizes. */
int qclhip_grid_refine_fallbacks(qclhip_ctx* ctx, const qclhip_grid* grid, int64_t* count_host);

/* ---- A1: _retarded_green_production (production_operators.jl:1337-1385)
 * G^R = inv(D/s)/s per block, D = (eps_e + i eta) I - h_m - SigmaR_em,
 * s = max row sum |D| (written to scale[NE*Nk] when not NULL). sigR may be
 * NULL (Sigma = 0, the seed of seed.jl:151-197). cond(D) (:1380) is a
 * diagnostic that the reference never reads; it is not computed. */
int qclhip_retarded_green(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h,
                          const double* sigR, double eta, double* GR, double* scale);

/* ---- A1+A2: Dyson plus the raw Keldysh products
 * G<raw = G^R S< G^R^dagger, G>raw = G^R S> G^R^dagger (production_state.jl:
 * 14-48), and the weighted block traces for the number functionals.
 * traces may be NULL (context scratch); else 2*NE*Nk complex. */
int qclhip_dyson_keldysh(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h,
                         const double* sigR, const double* sigL, const double* sigG,
                         double eta, double* GR, double* GL, double* GG, double* scale,
                         double* traces);

/* The same launch pair on fields whose (a,b) planes lie `plane` elements
 * apart (element (e,m,a,b) at e + NE m + plane (a + Nb b), plane >= NE*Nk):
 * the layout of a Julia view into a larger array, as the reference's
 * loops over G[e, m, a, b] accept (production_operators.jl:1337-1385,
 * production_state.jl:14-48). A plane stride off a power of two avoids the
 * HBM bank conflicts of NE*Nk = 2^k. FP64 vector engine only. */
int qclhip_dyson_keldysh_strided(qclhip_ctx* ctx, const qclhip_grid* grid, int64_t plane,
                                 const double* h, const double* sigR, const double* sigL,
                                 const double* sigG, double eta, double* GR, double* GL,
                                 double* GG, double* scale, double* traces);

/* ---- A2: _current_green_production (production_state.jl:1-74) ----------
 * total = fam[0] + fam[1] + ... (embedding first, then mechanisms in order;
 * production_operators.jl:1447-1473); Dyson; Keldysh; number/holes
 * (A3, production_operators.jl:1417-1445); paired coefficients (A4,
 * reference/greens.jl:155-185); G< <- aG< - cG>, G> <- (1-c)G> - (1-a)G<;
 * A = i(G^R - G^R^dagger) when A != NULL (A5). famR/famL/famG are HOST
 * arrays of nfam device pointers. out_host (lambda, number, holes, a, c) may
 * be NULL; when given, the call synchronizes. */
int qclhip_current_green(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h,
                         int nfam, const double* const* famR, const double* const* famL,
                         const double* const* famG, double target, double* GR, double* GL,
                         double* GG, double* A, double* scale, double* out_host);
/* qclhip_current_green that also writes the summed families (the total
 * self-energy S^R, S<, S> of the iteration, _total_family_production,
 * production_operators.jl:1447-1473) to totR, totL, totG: in the FP64
 * vector path the kernels store the sums they load (one extra write per
 * component); otherwise a summation pass fills them. For the later passes
 * of an iteration (the residual suite) that would sum the families again. */
int qclhip_current_green_totals(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, int nfam,
                                const double* const* famR, const double* const* famL, const double* const* famG,
                                double target, double* GR, double* GL, double* GG, double* A, double* scale,
                                double* out_host, double* totR, double* totL, double* totG);

/* No counterpart in the reference (a fusion of two of its passes, for the
 * SCBA iteration that measures every iteration): qclhip_current_green (with
 * the totals when totR, totL, totG are given, else NULL) whose FP64 vector
 * Keldysh kernel also runs the residual suite's positivity screens of the raw
 * products -i G<raw and i G>raw (residuals.jl:699-728) on its product tiles
 * and writes one screen word per block to raw_screen (device, NE*Nk bytes),
 * for qclhip_residual_suite_prescreened / _state_prescreened on the same
 * Green state and totals. Returns QCLHIP_ERR_UNSUPPORTED for the matrix-core
 * and fused engines, for FP32 products and above Nb = 32. */
int qclhip_current_green_screened(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, int nfam,
                                  const double* const* famR, const double* const* famL,
                                  const double* const* famG, double target, double* GR, double* GL, double* GG,
                                  double* A, double* scale, double* out_host, double* totR, double* totL,
                                  double* totG, unsigned char* raw_screen);

/* qclhip_current_green that allocates nothing. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results as qclhip_current_green (bitwise),
 * except that the block traces live in a fixed 16 MiB device array of the
 * library (one per device, held until the process ends) instead of the
 * context's growable scratch: NE*Nk <= 524,288 blocks and nfam <= 8, else
 * QCLHIP_ERR_ARGUMENT. Every call on one device shares that array, so calls
 * that overlap on one device (two streams, two contexts, two threads) race;
 * serialize them. Nothing in it needs to start at zero. */
int qclhip_current_green_noalloc(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, int nfam,
                                 const double* const* famR, const double* const* famL, const double* const* famG,
                                 double target, double* GR, double* GL, double* GG, double* A, double* scale,
                                 double* out_host);

/* qclhip_current_green with occupation_normalization = :scalar_lesser, the
 * reference's legacy comparator branch (production_state.jl:54-61; option
 * accepted at algorithm_options.jl:29,63): lambda = N[G<raw]/target must be
 * finite and positive (DomainError otherwise), G< = G<raw / lambda, and
 * G> = G^R - G^R^dagger + G< as _green_state_production computes it without
 * a given G> (production_operators.jl:1387-1415); no paired map and no
 * capacity gates. out_host = (lambda, number, 0, 1, 0). */
int qclhip_current_green_scalar_lesser(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, int nfam,
                                       const double* const* famR, const double* const* famL,
                                       const double* const* famG, double target, double* GR, double* GL,
                                       double* GG, double* A, double* scale, double* out_host);

/* A3: number functional N[G] = Re[-i g_s/(2 pi) sum wE wk tr G] with the
 * realness gate (synchronizes). */
int qclhip_number_functional(qclhip_ctx* ctx, const qclhip_grid* grid, const double* G,
                             double* value_host);
/* A5: A = i(G^R - G^R^dagger) (production_operators.jl:1387-1415). */
int qclhip_spectral(qclhip_ctx* ctx, const qclhip_grid* grid, const double* GR, double* A);
/* _green_state_production without a given G> (production_operators.jl:
 * 1387-1415): G> = G^R - G^R^dagger + G<, A = i(G^R - G^R^dagger); either
 * output may be NULL. */
int qclhip_green_state(qclhip_ctx* ctx, const qclhip_grid* grid, const double* GR,
                       const double* GL, double* GG, double* A);
/* A4 applied with host coefficients (production_state.jl:62-70). */
int qclhip_normalize_pair(qclhip_ctx* ctx, int64_t n, double* GL, double* GG, double a,
                          double c);

/* ---- A6: build_shift_plan / apply_energy_shift! (production_cache.jl:258-325)
 * Non-cyclic two-point interpolation X(eps+delta); zero outside the window. */
int qclhip_shift_create(qclhip_ctx* ctx, const qclhip_grid* grid, double delta,
                        qclhip_shift** out);
int qclhip_shift_destroy(qclhip_shift* shift);
/* Y = S X over ncols columns of length NE; Y must not alias X. */
int qclhip_apply_shift(qclhip_ctx* ctx, const qclhip_shift* shift, int64_t ncols,
                       const double* X, double* Y);

/* ---- A7: _production_static_contraction! (production_operators.jl:402-658)
 * Kernel representations (production_contracts.jl:262-323), HOST matrices,
 * column-major complex:
 *   QCLHIP_KERNEL_DENSE: a = Kflat (Nc x Nc), rows (m,a,b), cols (m',c,d),
 *     Nc = Nk*Nb^2 (production_cache.jl:387-409);
 *   QCLHIP_KERNEL_MOMENTUM_INDEPENDENT: a = block (Nb^2 x Nb^2);
 *   QCLHIP_KERNEL_SEPARABLE_IFR: a = left (Nb^2 x rank), b = right
 *     (rank x Nb^2), c = radial (Nk x Nk);
 *   QCLHIP_KERNEL_LOW_RANK: a = left (Nc x rank), b = right_adjoint (rank x Nc).
 * The radial weights wk of the grid are folded into the device copy. */
#define QCLHIP_KERNEL_DENSE 0
#define QCLHIP_KERNEL_MOMENTUM_INDEPENDENT 1
#define QCLHIP_KERNEL_SEPARABLE_IFR 2
#define QCLHIP_KERNEL_LOW_RANK 3
int qclhip_kernel_create(qclhip_ctx* ctx, const qclhip_grid* grid, int kind, int rank,
                         const double* a_host, const double* b_host, const double* c_host,
                         qclhip_kernel** out);
/* The dense representation from a DEVICE matrix Kflat (Nc x Nc,
 * column-major complex): the weights are folded into the library's own
 * device copy on the stream (at config sizes Kflat is tens of GB, which a
 * host round trip would not carry). */
int qclhip_kernel_create_dense_device(qclhip_ctx* ctx, const qclhip_grid* grid, const double* a_dev,
                                      qclhip_kernel** out);
int qclhip_kernel_destroy(qclhip_kernel* kernel);
/* S <- alpha*qK*K[wk G] + beta*S; beta == 0 does not read S. */
int qclhip_contract(qclhip_ctx* ctx, const qclhip_kernel* kernel, const double* G,
                    double qK, double alpha, double beta, double* S);
/* The momentum-independent branch of _production_static_contraction!
 * (production_operators.jl:535-595) without its broadcast over momenta: Y
 * (NE x Nb^2, one momentum's columns) <- alpha*qK*K[wk G], which every
 * momentum of the full result holds bit for bit. Forward-form kernels of
 * kind QCLHIP_KERNEL_MOMENTUM_INDEPENDENT only (ARGUMENT otherwise). */
int qclhip_contract_compact(qclhip_ctx* ctx, const qclhip_kernel* kernel, const double* G,
                            double qK, double alpha, double* Y);
/* The broadcast of that branch: S(e, m, a, b) = Y(e, a, b) for every
 * momentum m of the grid (Y: NE x Nb^2). */
int qclhip_broadcast_momenta(qclhip_ctx* ctx, const qclhip_grid* grid, const double* Y, double* S);

/* ---- A8: _production_lo_contraction, thermal N_LO (production_operators.jl:
 * 1089-1162): S< = (N+1) K[S+ G<] + N K[S- G<], S> = (N+1) K[S- G>] + N K[S+ G>]
 * with S+- the +-hbar*omega_LO shifts. */
int qclhip_lo_contraction(qclhip_ctx* ctx, const qclhip_kernel* kernel, const double* GL,
                          const double* GG, double qK, double NLO,
                          const qclhip_shift* plus, const qclhip_shift* minus, double* SL,
                          double* SG);

/* ---- A9/A10: FFT Hilbert transform (fft_parallel.jl:27-291;
 * production_operators.jl:735-884). Plan for ncols columns of length NE on a
 * uniform grid; L = nextpow2(2NE-1) <= 262144 (NE <= 131072); verify adds
 * the 2L round-off witness. Up to L = 16384 a column's transform runs in one
 * workgroup's LDS; above, as a four-step transform through HBM scratch
 * (columns in chunks of at most 1 GiB of scratch). */
int qclhip_hilbert_create(qclhip_ctx* ctx, const qclhip_grid* grid, int64_t ncols,
                          int verify, qclhip_hilbert** out);
int qclhip_hilbert_destroy(qclhip_hilbert* plan);
/* Which implementation the plan runs: AUTO (by L, the default), LDS (L <=
 * 16384) or HBM (the four-step transform, L >= 4096). Results agree to
 * rounding; no counterpart in the reference (its FFTW plan has no such
 * choice). get_path reports the resolved choice (LDS or HBM). */
#define QCLHIP_HILBERT_PATH_AUTO 0
#define QCLHIP_HILBERT_PATH_LDS 1
#define QCLHIP_HILBERT_PATH_HBM 2
int qclhip_hilbert_set_path(qclhip_ctx* ctx, qclhip_hilbert* plan, int path);
/* Each column of the plan stands for mult identical columns in the round-off
 * witness's norms (a momentum-independent self-energy transformed once for
 * all Nk momenta, mult = Nk): the witness equals that of the full field's
 * transform (fft_parallel.jl:247-291) up to rounding. Default 1. */
int qclhip_hilbert_set_multiplicity(qclhip_ctx* ctx, qclhip_hilbert* plan, int64_t mult);
int qclhip_hilbert_get_path(qclhip_ctx* ctx, const qclhip_hilbert* plan, int* path_host);
/* Lambda_e = sum_{e' != e} wE_e' Gamma_e' / (2 pi (eps_e - eps_e')).
 * roundoff_host (may be NULL; requires verify) = ||L2 - L|| / (||L|| + 1e-14). */
int qclhip_fft_hilbert(qclhip_ctx* ctx, const qclhip_hilbert* plan, const double* Gamma,
                       double* Lambda, double* roundoff_host);
/* Gamma = i(S> - S<), S^R = H[Gamma] - i Gamma/2; Gamma may be NULL. */
int qclhip_retarded_self_energy(qclhip_ctx* ctx, const qclhip_hilbert* plan,
                                const double* SL, const double* SG, double* SR,
                                double* Gamma, double* roundoff_host);

/* ---- A12: _embedding_component_production (production_operators.jl:886-922)
 * S_em = T (shift X)_em T^dagger, zero outside the shift window. T is a
 * device Nb x Nb complex matrix. */
int qclhip_embedding_component(qclhip_ctx* ctx, const qclhip_grid* grid, const double* T,
                               const double* X, const qclhip_shift* shift, double* S);

/* ---- A13: elementwise family algebra (production_operators.jl:1447-1518) */
/* out = x[0] + x[1] + ... (x is a HOST array of nx device pointers). */
int qclhip_sum_fields(qclhip_ctx* ctx, int64_t n, int nx, const double* const* x, double* out);
/* old <- (1-alpha) old + alpha cand. */
int qclhip_mix(qclhip_ctx* ctx, int64_t n, double* old_, const double* cand, double alpha);
/* The same for npairs (old[p], cand[p]) field pairs in one pass
 * (_mix_mechanisms_production! and the plus/minus families,
 * production_solver.jl:819-832); old and cand are HOST arrays of device
 * pointers. */
int qclhip_mix_fields(qclhip_ctx* ctx, int64_t n, int npairs, double* const* old_,
                      const double* const* cand, double alpha);

/* ---- A14/A15: observables (production_iteration.jl:174-215) */
/* Jbar = Re[g_s/(2 pi) sum_{e sorted by (|e|,e), m} wE wk tr(S>+ G< - S<+ G>)]. */
int qclhip_boundary_flux(qclhip_ctx* ctx, const qclhip_grid* grid, const double* plusL,
                         const double* plusG, const double* GL, const double* GG,
                         double* Jbar_host);
/* p_a = Re[-i g_s/(2 pi) sum wE wk G<_aa] / target, out_host[Nb]. */
int qclhip_state_populations(qclhip_ctx* ctx, const qclhip_grid* grid, const double* GL,
                             double target, double* out_host);

/* ---- A16: seed (seed.jl:28-128) */
/* q_e = g_s/(2 pi) wE_e sum_m wk_m tr A_em; q_host[NE]. */
int qclhip_seed_weights(qclhip_ctx* ctx, const qclhip_grid* grid, const double* GR,
                        double* q_host);
/* G< = -f(eps; mu, kBT) (G^R - G^R^dagger). */
int qclhip_seed_lesser(qclhip_ctx* ctx, const qclhip_grid* grid, const double* GR, double mu,
                       double kBT, double* GL);

/* ---- batched GF solve over independent instances -------------------------
 * No counterpart in the reference, which solves one problem per
 * solve_production call (src/QCLNEGF.jl:982-983); a bias sweep calls it once
 * per point. Instances have their own grids (N_E, N_k, N_b may differ);
 * every kernel of current_green runs one launch per tile class over all
 * instances of that class. FP64 (the grids' precision option is ignored). */
int qclhip_batch_create(qclhip_ctx* ctx, int ninst, const qclhip_grid* const* grids,
                        qclhip_batch** out);
int qclhip_batch_destroy(qclhip_batch* batch);
/* qclhip_current_green for every instance i: h[i], families
 * famR/famL/famG[i*nfam + f] (HOST arrays of device pointers, nfam <= 8, the
 * same count for all instances), target_host[i], outputs GR/GL/GG[i];
 * out_host[5i..5i+4] = (lambda, number, holes, a, c) when not NULL (the call
 * then synchronizes; in deferred-results mode the 5*ninst values go to the
 * result buffer instead). The instance tables are staged to the device when
 * the fields or targets differ from the previous call (a host wait); a call
 * that repeats them launches only, so it may be captured in a graph. */
int qclhip_batch_current_green(qclhip_ctx* ctx, qclhip_batch* batch, const double* const* h,
                               int nfam, const double* const* famR, const double* const* famL,
                               const double* const* famG, const double* target_host,
                               double* const* GR, double* const* GL, double* const* GG,
                               double* out_host);
/* Launch granularity of the batched solve: PERSISTENT runs one launch per
 * kernel and tile class over all instances (the tile work of small
 * instances shares the device); PER_INSTANCE runs each instance's own
 * Dyson and Keldysh launches (the single-instance kernels, with register
 * prefetch) followed by the batched number functionals and paired maps.
 * Results are bitwise the same. */
#define QCLHIP_BATCH_PERSISTENT 0
#define QCLHIP_BATCH_PER_INSTANCE 1
int qclhip_batch_set_mode(qclhip_ctx* ctx, qclhip_batch* batch, int mode);
int qclhip_batch_get_mode(const qclhip_batch* batch);
/* Autotuner of the batch: runs qclhip_batch_current_green on these
 * arguments in both modes (a warm call and two timed calls each, HIP
 * events on the context's stream), keeps the faster mode, and writes the
 * two times (persistent, per instance) in ms to ms_host (may be NULL).
 * Synchronizes; not during capture or in deferred-results mode. */
int qclhip_batch_autotune(qclhip_ctx* ctx, qclhip_batch* batch, const double* const* h, int nfam,
                          const double* const* famR, const double* const* famL, const double* const* famG,
                          const double* target_host, double* const* GR, double* const* GL, double* const* GG,
                          double* ms_host);

/* ---- A19: adjoint (vector-Jacobian products) ----------------------------
 * No counterpart in the reference, which lists the adjoint as missing
 * (algorithm_registry.jl:226-249). For a real functional J of a complex
 * field Z the gradient is Zbar = dJ/dRe Z + i dJ/dIm Z (dJ = Re sum
 * conj(Zbar) dZ). Entries marked "accumulates" add into their outputs. */
/* Reverse of qclhip_current_green: given the forward's self-energy families
 * (famL, famG: HOST arrays of nfam device pointers; Sigma^R is not needed),
 * target, its out_host coefficients coef_host = (lambda, number, holes, a, c),
 * its outputs GR, GL, GG and the gradients GRbar (may be NULL), GLbar,
 * GGbar, writes the gradients SRbar, SLbar, SGbar of the TOTAL self-energy
 * (the gradient of every family) and, when hbar != NULL, the gradient of h
 * (Nk,Nb,Nb). Raw Keldysh fields are recovered from the paired map
 * (DomainError when a = c). */
int qclhip_current_green_adjoint(qclhip_ctx* ctx, const qclhip_grid* grid, int nfam,
                                 const double* const* famL, const double* const* famG,
                                 double target, const double* coef_host, const double* GR,
                                 const double* GL, const double* GG, const double* GRbar,
                                 const double* GLbar, const double* GGbar, double* SRbar,
                                 double* SLbar, double* SGbar, double* hbar);
/* Reverse of qclhip_boundary_flux for the scalar gradient Jbar; accumulates
 * into any non-NULL of plusLbar, plusGbar, GLbar, GGbar. */
int qclhip_boundary_flux_adjoint(qclhip_ctx* ctx, const qclhip_grid* grid, double Jbar,
                                 const double* plusL, const double* plusG, const double* GL,
                                 const double* GG, double* plusLbar, double* plusGbar,
                                 double* GLbar, double* GGbar);
/* X (+)= alpha S^T Y: transpose of apply_energy_shift! (production_cache.jl:
 * 300-325); accumulates when accumulate != 0. */
int qclhip_shift_transpose(qclhip_ctx* ctx, const qclhip_grid* grid, const qclhip_shift* shift,
                           const double* Y, double alpha, int accumulate, double* X);
/* Conjugate transpose (K diag(wk))^dagger of a kernel representation, in the
 * same representation: qclhip_contract with it maps Sbar to Gbar. */
int qclhip_kernel_adjoint(qclhip_ctx* ctx, const qclhip_kernel* kernel, qclhip_kernel** out);
/* Reverse of qclhip_retarded_self_energy: Gamma_bar = -wE * H_raw[SRbar]
 * + (i/2) SRbar (H_raw the transform without the wE weights; its kernel is
 * antisymmetric); accumulates SLbar += i Gamma_bar, SGbar -= i Gamma_bar. */
int qclhip_retarded_self_energy_adjoint(qclhip_ctx* ctx, const qclhip_hilbert* plan,
                                        const double* SRbar, double* SLbar, double* SGbar);
/* X_em <- (X_em - X_em^dagger)/2 in place: the projection onto the
 * anti-Hermitian blocks on which S< and S> live (the tangent space of the
 * SCBA state used by the fixed-point adjoint). */
int qclhip_antihermitian_part(qclhip_ctx* ctx, const qclhip_grid* grid, double* X);
/* Reverse of qclhip_embedding_component; accumulates
 * Xbar += shift^T (T^dagger Sbar T). */
int qclhip_embedding_adjoint(qclhip_ctx* ctx, const qclhip_grid* grid, const double* T,
                             const qclhip_shift* shift, const double* Sbar, double* Xbar);

/* production_residual_suite (src/numerics/optimized/residuals.jl:537-772):
 * out_host = r_D, r_A, r_K, r_PSD, r_caus of one SCBA iteration, from the
 * Green functions (GR, GL, GG, A), h (Nk,Nb,Nb), the families of the total
 * self-energy (summed in list order) and the unmixed candidate S< (families
 * summed likewise). A may be NULL: A = i(G^R - G^R^dagger) is then formed
 * from the G^R tile (bitwise the spectral function), for callers that do not
 * store it. One wavefront per (e,m) block; the positivity and causality
 * checks run Cholesky screens of H + delta I, and the blocks they do not
 * settle get their extreme Hermitian eigenvalues (Householder reduction and
 * Sturm multisection) in a second pass; per-block values reduced in a fixed
 * order (sums for r_A and r_K, maxima otherwise; a non-finite value gives
 * Inf). Deferrable (5 doubles). */
int qclhip_residual_suite(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, const double* GR,
                          const double* GL, const double* GG, const double* A, int nfam,
                          const double* const* famR, const double* const* famL, const double* const* famG,
                          int ncand, const double* const* candL, double* out_host);
/* production_selfenergy_residual (src/numerics/optimized/residuals.jl:
 * 802-865): max over (R, <, >) of ||cand - old||_F / (||cand||_F + 1e-14)
 * over n complex entries each; Inf when any entry is non-finite.
 * Deferrable (1 double). */
int qclhip_selfenergy_residual(qclhip_ctx* ctx, int64_t n, const double* const* old_, const double* const* cand,
                               double* out_host);

/* Linear mixing of the production loop (production_solver.jl:819-832,
 * old_p <- (1 - alpha) old_p + alpha cand_p for every pair p) fused with the
 * self-energy residuals of the unmixed pairs (production_iteration.jl:
 * 134-166 over production_selfenergy_residual): ratios_host receives
 * ||cand_t - old_t||_F / (||cand_t||_F + 1e-14) for the terms t = each pair
 * alone (pair order), then, for each p with partner[p] > p (ascending p),
 * the sum of pairs p and partner[p] (the embedding = plus + minus);
 * partner[p] = -1 for none, symmetric otherwise. Inf for a non-finite term.
 * Same traffic as the mixing alone. alpha = 0 measures only (fields are
 * not written; the residuals before an Anderson step). Deferrable
 * (npairs + couples doubles). */
int qclhip_mix_fields_residual(qclhip_ctx* ctx, int64_t n, int npairs, double* const* old_,
                               const double* const* cand, double alpha, const int* partner, double* ratios_host);

/* The same mixing pass that also writes the totals the next GF solve sums
 * (production_solver.jl:819-832 followed by _total_family_production,
 * production_operators.jl:1447-1518): pairs are grouped into ntot totals
 * (tot_count[t] consecutive entries of tot_members, each a pair index, in
 * summation order; every pair in exactly one total, a couple inside one
 * total); totals[t] <- the left-fold sum of the mixed values of its
 * members, bitwise equal to summing the mixed families afterwards. Ratios
 * as qclhip_mix_fields_residual (same term numbering). Deferrable. */
int qclhip_mix_fields_residual_totals(qclhip_ctx* ctx, int64_t n, int npairs, double* const* old_,
                                      const double* const* cand, double alpha, const int* partner, int ntot,
                                      const int* tot_count, const int* tot_members, double* const* totals,
                                      double* ratios_host);
/* The same pass one family at a time, for an iteration that never holds
 * every candidate at once (the family totals of _total_family_production,
 * production_operators.jl:1447-1518, built across calls): tot_flags[t] is a
 * combination of QCLHIP_TOTAL_ACCUMULATE (the left fold starts from the
 * current value of totals[t], so consecutive calls continue one fold,
 * bitwise equal to a single pass over all families) and
 * QCLHIP_TOTAL_CANDIDATE (totals[t] sums the members' CANDIDATE values; such
 * a total neither mixes nor measures its members, which must each belong to
 * a mixing total as well: the candidate S< sum of the residual suite).
 * Totals must not alias a pair's fields. Ratios as
 * qclhip_mix_fields_residual_totals. Deferrable. */
#define QCLHIP_TOTAL_ACCUMULATE 1
#define QCLHIP_TOTAL_CANDIDATE 2
int qclhip_mix_fields_residual_accumulate(qclhip_ctx* ctx, int64_t n, int npairs, double* const* old_,
                                          const double* const* cand, double alpha, const int* partner, int ntot,
                                          const int* tot_count, const int* tot_members, const int* tot_flags,
                                          double* const* totals, double* ratios_host);
/* The same pass where the candidates of the pairs with compact[p] != 0 are
 * one momentum's columns (NE x Nb^2, the momentum-independent contraction
 * without its broadcast, qclhip_contract_compact) read for every momentum
 * of the grid's fields: the values, totals and ratios of mixing their
 * broadcast fields, bit for bit, without writing or reading those fields
 * (n = NE Nk Nb^2 < 2^31). */
int qclhip_mix_fields_residual_compact(qclhip_ctx* ctx, const qclhip_grid* grid, int64_t n, int npairs,
                                       double* const* old_, const double* const* cand, double alpha,
                                       const int* partner, int ntot, const int* tot_count, const int* tot_members,
                                       const int* tot_flags, double* const* totals, const int* compact,
                                       double* ratios_host);

/* The residual suite in two passes, for an iteration that measures the state
 * before its candidates exist (production_residual_suite, residuals.jl:
 * 537-772, split by the fields each term reads):
 * qclhip_residual_state: every term but r_K (r_D, r_A, r_PSD over the six
 * positivity checks, r_caus) from the Green functions and the total
 * self-energy; A may be NULL, in which case A = i(G^R - G^R^dagger) is
 * formed from the G^R tile (bitwise the spectral function). out_host = the
 * five results with r_K = 0.
 * qclhip_residual_keldysh: r_K = ||G< - G^R S<_cand G^R^dagger|| / (||G<|| +
 * 1e-14) alone (:643-664), the candidate families summed in list order;
 * out_host = the five results with only r_K set. Both deferrable (5 doubles);
 * each term equals the one of qclhip_residual_suite. */
int qclhip_residual_state(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, const double* GR,
                          const double* GL, const double* GG, const double* A, int nfam,
                          const double* const* famR, const double* const* famL, const double* const* famG,
                          double* out_host);
int qclhip_residual_keldysh(qclhip_ctx* ctx, const qclhip_grid* grid, const double* GR, const double* GL,
                            int ncand, const double* const* candL, double* out_host);

/* qclhip_residual_suite and qclhip_residual_state on a Green state that
 * qclhip_current_green_screened produced from the total self-energy given
 * here (famR/famL/famG summing to the totals it loaded): the raw-product
 * screens are taken from raw_screen instead of being formed again (same
 * products, same arithmetic, so every result equals the plain entries'
 * bit for bit); blocks whose raw screens did not settle get their
 * eigenvalues from the products formed again in the second pass. No
 * counterpart in the reference. */
int qclhip_residual_suite_prescreened(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, const double* GR,
                                      const double* GL, const double* GG, const double* A, int nfam,
                                      const double* const* famR, const double* const* famL,
                                      const double* const* famG, int ncand, const double* const* candL,
                                      const unsigned char* raw_screen, double* out_host);
int qclhip_residual_state_prescreened(qclhip_ctx* ctx, const qclhip_grid* grid, const double* h, const double* GR,
                                      const double* GL, const double* GG, const double* A, int nfam,
                                      const double* const* famR, const double* const* famL,
                                      const double* const* famG, const unsigned char* raw_screen,
                                      double* out_host);

/* ---- The Poisson step closing the SCBA loop (production_solver.jl:
 * 1044-1047). All arrays are device memory. */
/* _electron_density_bar (src/physics/observables.jl:1-26): nbar (Nz) from
 * G< on the grid: I_ab = sum_{m,e} wE wk G<_{e,m,a,b}, nbar_g = gs/(2 pi)
 * Re(-i sum_ab chi_ga I_ab conj(chi_gb)); chi Nz x Nb complex column-major. */
int qclhip_electron_density(qclhip_ctx* ctx, const qclhip_grid* grid, const double* GL, int Nz, const double* chi,
                            double gs, double* nbar);
/* solve_periodic_poisson (src/numerics/reference/poisson.jl:12-65): the
 * bordered periodic finite-volume system solved by LU with partial
 * pivoting on one workgroup; U (N) receives the Hartree energy; out_host =
 * zeta, r_P, neutrality (deferrable, 3 doubles). DomainError for a
 * singular system. N <= 4096. */
int qclhip_periodic_poisson(qclhip_ctx* ctx, int N, const double* eps_r, const double* wx, const double* ND,
                            double lambdaP, const double* nbar, double* U, double* out_host);
/* project_hamiltonians (src/numerics/reference/basis.jl:288-308): h (device
 * layout (Nb,Nb,Nk), i.e. reference (Nk,Nb,Nb)) = (S + Phi^dagger diag(U)
 * Phi) + kin_m Minv with S = H0 - F (Z - z0) and kin_m = C kappa_m^2. */
int qclhip_project_hamiltonians(qclhip_ctx* ctx, int Nz, int Nb, int Nk, const double* Phi, const double* U,
                                const double* S, const double* kin, const double* Minv, double* h);
/* Outer mixing (production_solver.jl:1046-1047): U <- (1-alpha) U + alpha
 * cand; rU_host = max |cand - U_old| (deferrable, 1 double). */
int qclhip_mix_potential(qclhip_ctx* ctx, int N, double* U, const double* cand, double alpha, double* rU_host);

/* The same fused pass for a fixed-point iteration's change: mixes every
 * pair and returns sums_host = (sum_p ||cand_p - old_p||^2,
 * sum_p ||cand_p||^2) of the unmixed pairs (fixed-order reductions). No
 * counterpart in the reference: the adjoint's stopping criterion. */
int qclhip_mix_fields_change(qclhip_ctx* ctx, int64_t n, int npairs, double* const* old_, const double* const* cand,
                             double alpha, double* sums_host);
/* X[0..n) <- 0 (complex) on the context's stream, by a kernel (usable under
 * graph capture, where hipMemsetAsync nodes are avoided). */
int qclhip_fill_zero(qclhip_ctx* ctx, int64_t n, double* X);

/* ---- Anderson mixing (production_mixing.jl:135-317). A history entry is
 * one field per (family, component) of the ordered families (mechanisms,
 * plus, minus; R, <, >): nfields pointers, entry-major for histories. */
/* _anderson_append! (:195-229) device part: saved_f <- state_f,
 * resid_f <- cand_f - state_f, and for every field the real inner products
 * with the nold older residuals and with itself: dots_host[f*(nold+1) + h]
 * = Re <old_h,f, resid_f>, [f*(nold+1) + nold] = ||resid_f||^2 (the host
 * sums them over fields in family and component order). nold <= 15. */
int qclhip_anderson_append(qclhip_ctx* ctx, int64_t n, int nfields, const double* const* state,
                           const double* const* cand, double* const* saved, double* const* resid, int nold,
                           const double* const* old_resid, double* dots_host);
/* The Anderson linear combination (:256-268): out_f = sum_h c_h (X_h,f +
 * damping R_h,f) in history order; count <= 16. */
int qclhip_anderson_combine(qclhip_ctx* ctx, int64_t n, int nfields, int count, const double* coef_host,
                            double damping, const double* const* X, const double* const* R, double* const* out);
/* _anderson_guard (:36-111): checks every (family, e, m) block of the nfam
 * families (R, <, >); first_failure = one-based index of the first failing
 * block in the reference's order (family, e, m fastest), 0 when all pass;
 * code = its failure code 1..5 (non-finite, causality of Im S^R, Keldysh
 * anti-Hermitian part, Keldysh positivity, S^R - S^A = S> - S<). */
int qclhip_anderson_guard(qclhip_ctx* ctx, const qclhip_grid* grid, int nfam, const double* const* famR,
                          const double* const* famL, const double* const* famG, int64_t* first_failure,
                          int* code);

#ifdef __cplusplus
}
#endif
#endif
