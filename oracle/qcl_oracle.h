/*
 * CPU oracle for the SCBA inner loop of QCLNEGF.jl (FP64, host memory).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP backend
 * is compared against; only the tests, the smoke check and the CPU-baseline
 * leg of the benchmark load it. The product (libqclhip) never links, calls or
 * falls back to it.
 *
 * Every field uses the reference layout: a complex (NE,Nk,Nb,Nb) array in
 * column-major order, energy fastest, element (e,m,a,b) at
 *   e + NE*(m + Nk*(a + Nb*b)),
 * stored as interleaved (re,im) doubles (Julia ComplexF64). h is (Nk,Nb,Nb).
 * All indices in this interface are 0-based.
 *
 * Return codes: 0 ok; >0 one-based linear block (e + NE*m + 1) of a singular
 * Dyson block; -1 argument error; -2 dimension mismatch; -3 domain error.
 */
#ifndef QCL_ORACLE_H
#define QCL_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* Kernel representation of one scattering mechanism
 * (production_contracts.jl:262-323). kind: 0 dense (a = Kflat, Nc x Nc),
 * 1 momentum independent (a = block, Nb^2 x Nb^2), 2 separable IFR
 * (a = left Nb^2 x rank, b = right rank x Nb^2, c = radial Nk x Nk),
 * 3 low rank (a = left Nc x rank, b = right_adjoint rank x Nc),
 * 4 literal six-index tensor (a = K, (Nk,Nk,Nb,Nb,Nb,Nb)). All column-major. */
typedef struct {
    int kind;
    int rank;
    const double* a;
    const double* b;
    const double* c;
} orc_kernel;

int orc_retarded_green(int NE, int Nk, int Nb, const double* eps, const double* h,
                       const double* sigR, double eta, double* GR, double* scale);
int orc_invert_block(int n, double* A);
void orc_spectral(int NE, int Nk, int Nb, const double* GR, double* A);
void orc_keldysh(int NE, int Nk, int Nb, const double* GR, const double* S, double* out);
double orc_number_functional(int NE, int Nk, int Nb, const double* G, const double* wE,
                             const double* wk, int gs, int* status);
int orc_paired_coefficients(double number, double holes, double target, double* a,
                            double* c, double* lambda);
void orc_normalize_pair(long long n, double* GL, double* GG, double a, double c);
int orc_current_green(int NE, int Nk, int Nb, const double* eps, const double* h,
                      int nfam, const double* const* famR, const double* const* famL,
                      const double* const* famG, const double* wE, const double* wk,
                      int gs, double target, double* GR, double* GL, double* GG,
                      double* A, double* scale, double* out3 /* lambda, number, holes */);

int orc_shift_plan(int NE, const double* eps, double delta, int* left, int* right,
                   double* wr, unsigned char* valid);
int orc_apply_shift(int NE, long long ncols, const int* left, const int* right,
                    const double* wr, const unsigned char* valid, const double* X,
                    double* Y);
int orc_shift_matrix(int NE, const double* eps, double delta, double* W);
void orc_apply_shift_dense(int NE, long long ncols, const double* W, const double* X,
                           double* Y);

int orc_is_momentum_independent(int Nk, int Nb, const double* K6);
void orc_kernel_dense_from_K6(int Nk, int Nb, const double* K6, double* Kflat);
void orc_kernel_mi_from_K6(int Nk, int Nb, const double* K6, double* block);
int orc_contract(int NE, int Nk, int Nb, const orc_kernel* op, const double* G,
                 const double* wk, double qK, int energy_chunk, double alpha,
                 double beta, double* S);

int orc_uniform_spacing(int NE, const double* eps, double* spacing);
int orc_fft_hilbert(int NE, long long ncols, const double* eps, const double* wE,
                    const double* Gamma, double* Lambda, int transform_length);
int orc_fft_hilbert_checked(int NE, long long ncols, const double* eps, const double* wE,
                            const double* Gamma, double* Lambda, double* roundoff);
int orc_direct_hilbert(int NE, long long ncols, const double* eps, const double* wE,
                       const double* Gamma, double* Lambda, int reverse_order);
int orc_retarded_self_energy_fft(int NE, long long ncols, const double* eps,
                                 const double* wE, const double* SL, const double* SG,
                                 double* SR, double* Gamma, double* roundoff);
void orc_fft(long long L, double* data, int inverse);

int orc_lo_contraction(int NE, int Nk, int Nb, const orc_kernel* op, const double* GL,
                       const double* GG, const double* wk, double qK, double NLO,
                       const int* pl, const int* pr, const double* pw,
                       const unsigned char* pv, const int* ml, const int* mr,
                       const double* mw, const unsigned char* mv, int energy_chunk,
                       double* SL, double* SG);
int orc_embedding_component(int NE, int Nk, int Nb, const double* T, const double* G,
                            const int* left, const int* right, const double* wr,
                            const unsigned char* valid, double* S);
void orc_mix(long long n, double* old_, const double* cand, double alpha);
void orc_add(long long n, const double* x, const double* y, double* out);

double orc_boundary_flux_bar(int NE, int Nk, int Nb, const double* eps, const double* wE,
                             const double* wk, int gs, const double* plusL,
                             const double* plusG, const double* GL, const double* GG);
void orc_state_populations(int NE, int Nk, int Nb, const double* wE, const double* wk,
                           int gs, double target, const double* GL, double* out);

double orc_fermi(double e, double mu, double kBT);
int orc_seed_weights(int NE, int Nk, int Nb, const double* GR, const double* wE,
                     const double* wk, int gs, double* q);
int orc_seed_mu(int NE, const double* eps, const double* q, double kBT, double target,
                double* mu);
void orc_seed_lesser(int NE, int Nk, int Nb, const double* GR, const double* eps,
                     double mu, double kBT, double* GL);

/* One production SCBA iteration (production_solver.jl:435-844 with linear
 * mixing and the sparse shift plans). */
typedef struct {
    int NE, Nk, Nb, nmech, gs, energy_chunk;
    const double *eps, *wE, *wk, *h, *Tplus, *Tminus;
    double Ep, hwLO, NLO, target, alpha_mix;
    const orc_kernel* kernels;
    const double* qK;
    const int* is_lo;
} orc_problem;

typedef struct {
    double **scatR, **scatL, **scatG; /* nmech fields each */
    double *plusR, *plusL, *plusG, *minusR, *minusL, *minusG, *embR, *embL, *embG;
    double *GR, *GL, *GG, *A, *scale;
} orc_state;

int orc_scba_iteration(const orc_problem* p, orc_state* s, double* out /* lambda, J,
                       roundoff, number, holes */, double* populations);
/* The same iteration with the fixed-point measurement of the reference
 * (production_iteration.jl:102-166) taken before mixing: meas = r_D, r_A,
 * r_K, r_PSD, r_caus, r_Sigma, r_lambda. */
int orc_scba_iteration_measured(const orc_problem* p, orc_state* s, double* out, double* populations,
                                double* meas);

/* Eigenvalues (ascending) of a Hermitian n x n block, cyclic Jacobi. */
void orc_hermitian_eigvals(int n, const double* H, double* w);
/* production_residual_suite (residuals.jl:537-772): out = r_D, r_A, r_K,
 * r_PSD, r_caus. */
int orc_residual_suite(int NE, int Nk, int Nb, const double* eps, const double* h, const double* GR,
                       const double* GL, const double* GG, const double* A, const double* SR,
                       const double* SL, const double* SG, const double* candSL, double* out);
/* production_selfenergy_residual (residuals.jl:802-865); old and cand are
 * arrays of the three component pointers (R, <, >) of n complex entries. */
double orc_selfenergy_residual(long long n, const double* const* old_, const double* const* cand);

int orc_num_threads(void);
void orc_set_num_threads(int n);

/* Outer loop (the Poisson step closing the SCBA loop). */
int orc_electron_density(int NE, int Nk, int Nb, int Nz, const double* GL, const double* wE, const double* wk,
                         const double* chi, double gs, double* nbar);
int orc_periodic_poisson(int N, const double* eps_r, const double* wx, const double* ND, double lambdaP,
                         const double* nbar, double* U, double* out /* zeta, r_P, neutrality */);
int orc_project_hamiltonians(int Nz, int Nb, int Nk, const double* Phi, const double* U, const double* S,
                             const double* kin, const double* Minv, double* h);
double orc_mix_potential(int N, double* U, const double* cand, double alpha);

/* The iteration up to the mixing: candidates (mechanisms R,<,> in order,
 * then plus R,<,>, minus R,<,>) copied into cand_out, state not mixed. */
int orc_scba_iteration_candidates(const orc_problem* p, orc_state* s, double* out, double* populations,
                                  double* meas, double* const* cand_out);
/* Anderson mixing primitives (production_mixing.jl:36-317). */
double orc_anderson_inner(long long n, int nfields, const double* const* a, const double* const* b);
void orc_anderson_combine(long long n, int nfields, int count, const double* coef, double damping,
                          const double* const* X, const double* const* R, double* const* out);
int orc_anderson_guard(int NE, int Nk, int Nb, int nfam, const double* const* R, const double* const* L,
                       const double* const* G, long long* first, int* code);

#ifdef __cplusplus
}
#endif
#endif
