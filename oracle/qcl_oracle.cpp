// CPU oracle for the SCBA inner loop of QCLNEGF.jl, restated in C++17 from the
// reference's production (E1) operators, with the literal (E0) twins used to
// cross-check them.
//
// TEST INFRASTRUCTURE ONLY: the checker the HIP backend is compared against.
// Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
// load this library. The shipped backend never calls it.
//
// References are to files of QCLNEGF.jl v0.2.0 (paths relative to its root).
#include "qcl_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cd = std::complex<double>;

constexpr double kTwoPi = 6.283185307179586476925286766559;
// src/domain/numerical_invariants.jl:5,10
constexpr double kRealnessLimit = 1.0e-10;
constexpr double kFermiSaturation = 40.0;

inline const cd* C(const double* p) { return reinterpret_cast<const cd*>(p); }
inline cd* C(double* p) { return reinterpret_cast<cd*>(p); }

inline size_t at4(int e, int m, int a, int b, int NE, int Nk, int Nb) {
    return (size_t)e + (size_t)NE * ((size_t)m + (size_t)Nk * ((size_t)a + (size_t)Nb * b));
}

inline double abs1(cd z) { return std::fabs(z.real()) + std::fabs(z.imag()); }

// Partial-pivot LU and column-wise triangular solves of A X = I, the owned
// small inverse of src/numerics/optimized/production_operators.jl:84-139.
// Pivot = first row of maximal |Re|+|Im| (LAPACK izamax, used by zgetrf at
// production_operators.jl:20-41). Returns 0 or the one-based zero pivot.
int lu_inverse(int n, cd* A, int* piv, cd* X) {
    for (int k = 0; k < n; ++k) {
        int p = k;
        double largest = abs1(A[k + n * k]);
        for (int i = k + 1; i < n; ++i) {
            double mag = abs1(A[i + n * k]);
            if (mag > largest) { largest = mag; p = i; }
        }
        if (A[p + n * k] == cd(0.0, 0.0)) return k + 1;
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(A[k + n * j], A[p + n * j]);
        for (int i = k + 1; i < n; ++i) A[i + n * k] /= A[k + n * k];
        for (int j = k + 1; j < n; ++j)
            for (int i = k + 1; i < n; ++i) A[i + n * j] -= A[i + n * k] * A[k + n * j];
    }
    for (int col = 0; col < n; ++col) {
        cd* x = X + (size_t)n * col;
        for (int i = 0; i < n; ++i) x[i] = (i == col) ? cd(1.0, 0.0) : cd(0.0, 0.0);
        for (int k = 0; k < n; ++k) std::swap(x[k], x[piv[k]]);
        for (int i = 0; i < n; ++i) {
            cd v = x[i];
            for (int k = 0; k < i; ++k) v -= A[i + n * k] * x[k];
            x[i] = v;
        }
        for (int i = n - 1; i >= 0; --i) {
            cd v = x[i];
            for (int k = i + 1; k < n; ++k) v -= A[i + n * k] * x[k];
            x[i] = v / A[i + n * i];
        }
    }
    std::memcpy(A, X, sizeof(cd) * (size_t)n * n);
    return 0;
}

// C = A * B (n x n, column-major)
void matmul(int n, const cd* A, const cd* B, cd* Cm) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            cd v(0.0, 0.0);
            for (int c = 0; c < n; ++c) v += A[i + n * c] * B[c + n * j];
            Cm[i + n * j] = v;
        }
}

// C = A * B^dagger
void matmul_adj(int n, const cd* A, const cd* B, cd* Cm) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            cd v(0.0, 0.0);
            for (int c = 0; c < n; ++c) v += A[i + n * c] * std::conj(B[j + n * c]);
            Cm[i + n * j] = v;
        }
}

void load_block(const cd* X, int e, int m, int NE, int Nk, int Nb, cd* blk) {
    for (int b = 0; b < Nb; ++b)
        for (int a = 0; a < Nb; ++a) blk[a + Nb * b] = X[at4(e, m, a, b, NE, Nk, Nb)];
}

void store_block(cd* X, int e, int m, int NE, int Nk, int Nb, const cd* blk) {
    for (int b = 0; b < Nb; ++b)
        for (int a = 0; a < Nb; ++a) X[at4(e, m, a, b, NE, Nk, Nb)] = blk[a + Nb * b];
}

double field_norm(size_t n, const cd* x) {
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += std::norm(x[i]);
    return std::sqrt(s);
}

}  // namespace

extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_invert_block(int n, double* A) {
    std::vector<int> piv(n);
    std::vector<cd> X((size_t)n * n);
    return lu_inverse(n, C(A), piv.data(), X.data());
}

// _retarded_green_production (production_operators.jl:1337-1385): D, row-sum
// scale s, G = inv(D/s)/s. cond(D) (:1380) is diagnostic only and is omitted.
int orc_retarded_green(int NE, int Nk, int Nb, const double* eps, const double* h_,
                       const double* sigR_, double eta, double* GR_, double* scale) {
    if (NE <= 0 || Nk <= 0 || Nb <= 0) return -2;
    const cd* h = C(h_);
    const cd* sigR = C(sigR_);
    cd* GR = C(GR_);
    long long blocks = (long long)NE * Nk;
    long long first_bad = 0;
#pragma omp parallel
    {
        std::vector<cd> D((size_t)Nb * Nb), X((size_t)Nb * Nb);
        std::vector<int> piv(Nb);
#pragma omp for schedule(static)
        for (long long linear = 0; linear < blocks; ++linear) {
            int e = (int)(linear % NE), m = (int)(linear / NE);
            for (int b = 0; b < Nb; ++b)
                for (int a = 0; a < Nb; ++a) {
                    cd diag = (a == b) ? cd(eps[e], eta) : cd(0.0, 0.0);
                    D[a + Nb * b] = diag - h[m + (size_t)Nk * (a + (size_t)Nb * b)] -
                                    (sigR ? sigR[at4(e, m, a, b, NE, Nk, Nb)] : cd(0.0, 0.0));
                }
            double s = 0.0;
            for (int a = 0; a < Nb; ++a) {
                double row = 0.0;
                for (int b = 0; b < Nb; ++b) row += std::abs(D[a + Nb * b]);
                s = std::max(s, row);
            }
            int info = 1;
            if (s > 0) {
                for (auto& v : D) v /= s;
                info = lu_inverse(Nb, D.data(), piv.data(), X.data());
            }
            if (info != 0) {
#pragma omp critical
                { if (first_bad == 0 || linear + 1 < first_bad) first_bad = linear + 1; }
                continue;
            }
            for (int b = 0; b < Nb; ++b)
                for (int a = 0; a < Nb; ++a)
                    GR[at4(e, m, a, b, NE, Nk, Nb)] = D[a + Nb * b] / s;
            if (scale) scale[linear] = s;
        }
    }
    return (int)std::min<long long>(first_bad, 0x7fffffff);
}

// A = i(G^R - G^R^dagger) per block (greens.jl:50-57; production_operators.jl:1402-1408).
void orc_spectral(int NE, int Nk, int Nb, const double* GR_, double* A_) {
    const cd* GR = C(GR_);
    cd* A = C(A_);
    long long blocks = (long long)NE * Nk;
#pragma omp parallel for schedule(static)
    for (long long linear = 0; linear < blocks; ++linear) {
        int e = (int)(linear % NE), m = (int)(linear / NE);
        for (int b = 0; b < Nb; ++b)
            for (int a = 0; a < Nb; ++a) {
                cd d = GR[at4(e, m, a, b, NE, Nk, Nb)] - std::conj(GR[at4(e, m, b, a, NE, Nk, Nb)]);
                A[at4(e, m, a, b, NE, Nk, Nb)] = cd(0.0, 1.0) * d;
            }
    }
}

// G^R S G^R^dagger per block (production_state.jl:29-47; greens.jl:67-75).
void orc_keldysh(int NE, int Nk, int Nb, const double* GR_, const double* S_, double* out_) {
    const cd* GR = C(GR_);
    const cd* S = C(S_);
    cd* out = C(out_);
    long long blocks = (long long)NE * Nk;
#pragma omp parallel
    {
        std::vector<cd> G((size_t)Nb * Nb), Sb((size_t)Nb * Nb), T((size_t)Nb * Nb),
            R((size_t)Nb * Nb);
#pragma omp for schedule(static)
        for (long long linear = 0; linear < blocks; ++linear) {
            int e = (int)(linear % NE), m = (int)(linear / NE);
            load_block(GR, e, m, NE, Nk, Nb, G.data());
            load_block(S, e, m, NE, Nk, Nb, Sb.data());
            matmul(Nb, G.data(), Sb.data(), T.data());
            matmul_adj(Nb, T.data(), G.data(), R.data());
            store_block(out, e, m, NE, Nk, Nb, R.data());
        }
    }
}

// _number_functional_production (production_operators.jl:1417-1445): terms in
// the order linear = e*Nk + m, summed sequentially, then -i g_s/(2 pi).
double orc_number_functional(int NE, int Nk, int Nb, const double* G_, const double* wE,
                             const double* wk, int gs, int* status) {
    const cd* G = C(G_);
    cd value(0.0, 0.0);
    for (int e = 0; e < NE; ++e)
        for (int m = 0; m < Nk; ++m) {
            cd tr(0.0, 0.0);
            for (int a = 0; a < Nb; ++a) tr += G[at4(e, m, a, a, NE, Nk, Nb)];
            value += (wE[e] * wk[m]) * tr;
        }
    value *= cd(0.0, -(double)gs / kTwoPi);
    if (status)
        *status = (std::fabs(value.imag()) <= kRealnessLimit * std::max(std::fabs(value.real()), 1.0)) ? 0 : -3;
    return value.real();
}

// _paired_number_coefficients (reference/greens.jl:155-185).
int orc_paired_coefficients(double number, double holes, double target, double* a,
                            double* c, double* lambda) {
    if (!(std::isfinite(target) && target > 0)) return -1;
    if (!(std::isfinite(number) && number > 0)) return -3;
    if (!(std::isfinite(holes) && holes >= 0)) return -3;
    if (!(target <= number + holes)) return -3;
    if (number >= target) {
        *a = target / number;
        *c = 0.0;
    } else {
        *a = 1.0;
        *c = (target - number) / holes;
    }
    *lambda = number / target;
    return 0;
}

// Paired map G< <- a G< - c G>, G> <- (1-c) G> - (1-a) G< (production_state.jl:62-70).
void orc_normalize_pair(long long n, double* GL_, double* GG_, double a, double c) {
    cd* GL = C(GL_);
    cd* GG = C(GG_);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) {
        cd lesser = GL[i], greater = GG[i];
        GL[i] = a * lesser - c * greater;
        GG[i] = (1.0 - c) * greater - (1.0 - a) * lesser;
    }
}

// _current_green_production (production_state.jl:1-74) with the default
// paired normalization. Families: index 0 is the embedding total, 1.. the
// scattering mechanisms in mechanism order (_total_family_production,
// production_operators.jl:1447-1473).
int orc_current_green(int NE, int Nk, int Nb, const double* eps, const double* h, int nfam,
                      const double* const* famR, const double* const* famL,
                      const double* const* famG, const double* wE, const double* wk, int gs,
                      double target, double* GR, double* GL, double* GG, double* A,
                      double* scale, double* out3) {
    size_t n = (size_t)NE * Nk * Nb * Nb;
    std::vector<cd> TR(n), TL(n), TG(n);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
        cd r = C(famR[0])[i], l = C(famL[0])[i], g = C(famG[0])[i];
        for (int f = 1; f < nfam; ++f) {
            r += C(famR[f])[i];
            l += C(famL[f])[i];
            g += C(famG[f])[i];
        }
        TR[i] = r;
        TL[i] = l;
        TG[i] = g;
    }
    int info = orc_retarded_green(NE, Nk, Nb, eps, h, (const double*)TR.data(), 0.0, GR, scale);
    if (info != 0) return info;
    orc_keldysh(NE, Nk, Nb, GR, (const double*)TL.data(), GL);
    orc_keldysh(NE, Nk, Nb, GR, (const double*)TG.data(), GG);
    int st = 0;
    double number = orc_number_functional(NE, Nk, Nb, GL, wE, wk, gs, &st);
    if (st) return st;
    double lambda = number / target;
    if (!(std::isfinite(lambda) && lambda > 0)) return -3;
    double holes = -orc_number_functional(NE, Nk, Nb, GG, wE, wk, gs, &st);
    if (st) return st;
    double a = 0, c = 0, lam = 0;
    int rc = orc_paired_coefficients(number, holes, target, &a, &c, &lam);
    if (rc) return rc;
    orc_normalize_pair((long long)n, GL, GG, a, c);
    if (A) orc_spectral(NE, Nk, Nb, GR, A);
    if (out3) {
        out3[0] = lambda;
        out3[1] = number;
        out3[2] = holes;
    }
    return 0;
}

// build_shift_plan (production_cache.jl:258-288), 0-based indices.
int orc_shift_plan(int NE, const double* eps, double delta, int* left, int* right,
                   double* wr, unsigned char* valid) {
    if (NE < 2) return -1;
    for (int e = 0; e < NE; ++e)
        if (!std::isfinite(eps[e])) return -1;
    for (int e = 1; e < NE; ++e)
        if (!(eps[e] - eps[e - 1] > 0)) return -1;
    if (!std::isfinite(delta)) return -1;
    int N = NE;
    int j = 0;
    for (int e = 0; e < N; ++e) {
        left[e] = 0;
        right[e] = 0;
        wr[e] = 0.0;
        valid[e] = 0;
        double target = eps[e] + delta;
        if (target < eps[0] || target > eps[N - 1]) continue;
        if (target == eps[N - 1]) {
            left[e] = N - 1;
            right[e] = N - 1;
            valid[e] = 1;
            continue;
        }
        while (j < N - 2 && eps[j + 1] <= target) ++j;
        left[e] = j;
        right[e] = j + 1;
        wr[e] = (target - eps[j]) / (eps[j + 1] - eps[j]);
        valid[e] = 1;
    }
    return 0;
}

// apply_energy_shift! (production_cache.jl:300-325).
int orc_apply_shift(int NE, long long ncols, const int* left, const int* right,
                    const double* wr, const unsigned char* valid, const double* X_,
                    double* Y_) {
    if (X_ == Y_) return -1;
    const cd* X = C(X_);
    cd* Y = C(Y_);
#pragma omp parallel for schedule(static)
    for (long long col = 0; col < ncols; ++col) {
        const cd* x = X + (size_t)col * NE;
        cd* y = Y + (size_t)col * NE;
        for (int e = 0; e < NE; ++e) {
            if (!valid[e]) {
                y[e] = cd(0.0, 0.0);
                continue;
            }
            int l = left[e], r = right[e];
            double t = wr[e];
            y[e] = (l == r || t == 0.0) ? x[l] : (1.0 - t) * x[l] + t * x[r];
        }
    }
    return 0;
}

// build_shift_matrix (reference/periodicity.jl:11-33), dense column-major NE x NE.
int orc_shift_matrix(int NE, const double* eps, double delta, double* W) {
    if (NE < 2) return -1;
    std::fill(W, W + (size_t)NE * NE, 0.0);
    for (int e = 0; e < NE; ++e) {
        double target = eps[e] + delta;
        if (target < eps[0] || target > eps[NE - 1]) continue;
        if (target == eps[NE - 1]) {
            W[e + (size_t)NE * (NE - 1)] = 1.0;
            continue;
        }
        int j = (int)(std::upper_bound(eps, eps + NE, target) - eps) - 1;  // searchsortedlast
        j = std::max(0, std::min(j, NE - 2));
        double th = (target - eps[j]) / (eps[j + 1] - eps[j]);
        W[e + (size_t)NE * j] = 1.0 - th;
        W[e + (size_t)NE * (j + 1)] = th;
    }
    return 0;
}

// apply_energy_shift with a dense matrix (reference/periodicity.jl:93-102).
void orc_apply_shift_dense(int NE, long long ncols, const double* W, const double* X_,
                           double* Y_) {
    const cd* X = C(X_);
    cd* Y = C(Y_);
#pragma omp parallel for schedule(static)
    for (long long col = 0; col < ncols; ++col)
        for (int e = 0; e < NE; ++e) {
            cd v(0.0, 0.0);
            for (int j = 0; j < NE; ++j) v += W[e + (size_t)NE * j] * X[(size_t)col * NE + j];
            Y[(size_t)col * NE + e] = v;
        }
}

static inline size_t at6(int m, int mp, int a, int c, int d, int b, int Nk, int Nb) {
    return (size_t)m + (size_t)Nk * ((size_t)mp + (size_t)Nk * ((size_t)a + (size_t)Nb * ((size_t)c + (size_t)Nb * ((size_t)d + (size_t)Nb * b))));
}

// _is_momentum_independent (production_cache.jl:327-334).
int orc_is_momentum_independent(int Nk, int Nb, const double* K6_) {
    const cd* K = C(K6_);
    for (int b = 0; b < Nb; ++b)
        for (int d = 0; d < Nb; ++d)
            for (int c = 0; c < Nb; ++c)
                for (int a = 0; a < Nb; ++a) {
                    cd ref = K[at6(0, 0, a, c, d, b, Nk, Nb)];
                    for (int mp = 0; mp < Nk; ++mp)
                        for (int m = 0; m < Nk; ++m)
                            if (K[at6(m, mp, a, c, d, b, Nk, Nb)] != ref) return 0;
                }
    return 1;
}

// _production_kernel (production_cache.jl:387-409): rows (m,a,b), columns (m',c,d).
void orc_kernel_dense_from_K6(int Nk, int Nb, const double* K6_, double* Kflat_) {
    const cd* K = C(K6_);
    cd* F = C(Kflat_);
    size_t Nc = (size_t)Nk * Nb * Nb;
    for (int d = 0; d < Nb; ++d)
        for (int c = 0; c < Nb; ++c)
            for (int mp = 0; mp < Nk; ++mp)
                for (int b = 0; b < Nb; ++b)
                    for (int a = 0; a < Nb; ++a)
                        for (int m = 0; m < Nk; ++m) {
                            size_t q = m + (size_t)Nk * (a + (size_t)Nb * b);
                            size_t qp = mp + (size_t)Nk * (c + (size_t)Nb * d);
                            F[q + Nc * qp] = K[at6(m, mp, a, c, d, b, Nk, Nb)];
                        }
}

// Momentum-independent block: rows (a,b), columns (c,d) (production_cache.jl:395-399).
void orc_kernel_mi_from_K6(int Nk, int Nb, const double* K6_, double* block_) {
    const cd* K = C(K6_);
    cd* B = C(block_);
    size_t N2 = (size_t)Nb * Nb;
    for (int d = 0; d < Nb; ++d)
        for (int c = 0; c < Nb; ++c)
            for (int b = 0; b < Nb; ++b)
                for (int a = 0; a < Nb; ++a)
                    B[(a + (size_t)Nb * b) + N2 * (c + (size_t)Nb * d)] = K[at6(0, 0, a, c, d, b, Nk, Nb)];
}

// _production_static_contraction! for each representation
// (production_operators.jl:402-658) and the literal six-index sum
// _static_contraction (reference/kernels.jl:415-435) with factor alpha*qK.
int orc_contract(int NE, int Nk, int Nb, const orc_kernel* op, const double* G_,
                 const double* wk, double qK, int energy_chunk, double alpha, double beta,
                 double* S_) {
    if (energy_chunk <= 0) return -1;
    const cd* G = C(G_);
    cd* S = C(S_);
    const int N2 = Nb * Nb;
    const long long Nc = (long long)Nk * N2;
    const double factor = alpha * qK;
    auto put = [&](size_t idx, cd value) {
        S[idx] = (beta == 0.0) ? factor * value : beta * S[idx] + factor * value;
    };
    const int nchunks = (NE + energy_chunk - 1) / energy_chunk;
    if (op->kind == 4) {
        const cd* K = C(op->a);
#pragma omp parallel for schedule(static)
        for (long long linear = 0; linear < (long long)NE * Nk; ++linear) {
            int e = (int)(linear % NE), m = (int)(linear / NE);
            for (int a = 0; a < Nb; ++a)
                for (int b = 0; b < Nb; ++b) {
                    cd v(0.0, 0.0);
                    for (int mp = 0; mp < Nk; ++mp)
                        for (int c = 0; c < Nb; ++c)
                            for (int d = 0; d < Nb; ++d)
                                v += wk[mp] * K[at6(m, mp, a, c, d, b, Nk, Nb)] * G[at4(e, mp, c, d, NE, Nk, Nb)];
                    put(at4(e, m, a, b, NE, Nk, Nb), v);
                }
        }
        return 0;
    }
    if (op->kind == 0) {  // dense (:402-465)
        const cd* K = C(op->a);
#pragma omp parallel
        {
            std::vector<cd> X((size_t)Nc * energy_chunk);
#pragma omp for schedule(static)
            for (int ch = 0; ch < nchunks; ++ch) {
                int e0 = ch * energy_chunk, nc = std::min(energy_chunk, NE - e0);
                for (int j = 0; j < nc; ++j)
                    for (long long q = 0; q < Nc; ++q)
                        X[q + Nc * j] = wk[q % Nk] * G[(size_t)(e0 + j) + (size_t)NE * q];
                for (int j = 0; j < nc; ++j)
                    for (long long q = 0; q < Nc; ++q) {
                        cd v(0.0, 0.0);
                        for (long long qp = 0; qp < Nc; ++qp) v += K[q + Nc * qp] * X[qp + Nc * j];
                        put((size_t)(e0 + j) + (size_t)NE * q, v);
                    }
            }
        }
        return 0;
    }
    if (op->kind == 1) {  // momentum independent (:535-595)
        const cd* B = C(op->a);
#pragma omp parallel
        {
            std::vector<cd> X((size_t)N2);
#pragma omp for schedule(static)
            for (int e = 0; e < NE; ++e) {
                for (int q = 0; q < N2; ++q) {
                    int c = q % Nb, d = q / Nb;
                    cd v(0.0, 0.0);
                    for (int mp = 0; mp < Nk; ++mp) v += wk[mp] * G[at4(e, mp, c, d, NE, Nk, Nb)];
                    X[q] = v;
                }
                for (int q = 0; q < N2; ++q) {
                    cd y(0.0, 0.0);
                    for (int qp = 0; qp < N2; ++qp) y += B[q + (size_t)N2 * qp] * X[qp];
                    int a = q % Nb, b = q / Nb;
                    for (int m = 0; m < Nk; ++m) put(at4(e, m, a, b, NE, Nk, Nb), y);
                }
            }
        }
        return 0;
    }
    if (op->kind == 2) {  // separable IFR (:467-533)
        const cd* left = C(op->a);
        const cd* right = C(op->b);
        const cd* radial = C(op->c);
        const int J = op->rank;
#pragma omp parallel
        {
            std::vector<cd> T((size_t)J * Nk), Y((size_t)J * Nk);
#pragma omp for schedule(static)
            for (int e = 0; e < NE; ++e) {
                for (int m = 0; m < Nk; ++m)
                    for (int j = 0; j < J; ++j) {
                        cd v(0.0, 0.0);
                        for (int q = 0; q < N2; ++q)
                            v += right[j + (size_t)J * q] * G[at4(e, m, q % Nb, q / Nb, NE, Nk, Nb)];
                        T[j + (size_t)J * m] = v;
                    }
                for (int j = 0; j < J; ++j)
                    for (int m = 0; m < Nk; ++m) {
                        cd v(0.0, 0.0);
                        for (int mp = 0; mp < Nk; ++mp)
                            v += radial[m + (size_t)Nk * mp] * (wk[mp] * T[j + (size_t)J * mp]);
                        Y[j + (size_t)J * m] = v;
                    }
                for (int m = 0; m < Nk; ++m)
                    for (int q = 0; q < N2; ++q) {
                        cd v(0.0, 0.0);
                        for (int j = 0; j < J; ++j) v += left[q + (size_t)N2 * j] * Y[j + (size_t)J * m];
                        put(at4(e, m, q % Nb, q / Nb, NE, Nk, Nb), v);
                    }
            }
        }
        return 0;
    }
    if (op->kind == 3) {  // low rank (:597-658)
        const cd* left = C(op->a);
        const cd* radj = C(op->b);
        const int r = op->rank;
#pragma omp parallel
        {
            std::vector<cd> red((size_t)r);
#pragma omp for schedule(static)
            for (int e = 0; e < NE; ++e) {
                for (int i = 0; i < r; ++i) {
                    cd v(0.0, 0.0);
                    for (long long q = 0; q < Nc; ++q)
                        v += radj[i + (size_t)r * q] * (wk[q % Nk] * G[(size_t)e + (size_t)NE * q]);
                    red[i] = v;
                }
                for (long long q = 0; q < Nc; ++q) {
                    cd v(0.0, 0.0);
                    for (int i = 0; i < r; ++i) v += left[q + (size_t)Nc * i] * red[i];
                    put((size_t)e + (size_t)NE * q, v);
                }
            }
        }
        return 0;
    }
    return -1;
}

// _uniform_energy_spacing (production_operators.jl:724-733).
int orc_uniform_spacing(int NE, const double* eps, double* spacing) {
    if (NE < 2) return -1;
    double d = eps[1] - eps[0];
    if (!(d > 0)) return -1;
    double dev = 0.0, emax = 0.0;
    for (int e = 1; e < NE; ++e) dev = std::max(dev, std::fabs((eps[e] - eps[e - 1]) - d));
    for (int e = 0; e < NE; ++e) emax = std::max(emax, std::fabs(eps[e]));
    double tol = 256.0 * 2.220446049250313e-16 * std::max(std::max(emax, std::fabs(d)), 1.0);
    if (!(dev <= tol)) return -1;
    *spacing = d;
    return 0;
}

// In-place radix-2 complex FFT, forward exp(-2 pi i jk/L); the inverse is
// normalized by 1/L like FFTW.ifft.
void orc_fft(long long L, double* data_, int inverse) {
    cd* x = C(data_);
    for (long long i = 1, j = 0; i < L; ++i) {
        long long bit = L >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(x[i], x[j]);
    }
    for (long long len = 2; len <= L; len <<= 1) {
        long long half = len >> 1;
        for (long long k = 0; k < half; ++k) {
            double ang = -kTwoPi * (double)k / (double)len;
            cd w(std::cos(ang), inverse ? -std::sin(ang) : std::sin(ang));
            for (long long s = 0; s < L; s += len) {
                cd u = x[s + k], v = x[s + k + half] * w;
                x[s + k] = u + v;
                x[s + k + half] = u - v;
            }
        }
    }
    if (inverse) {
        double inv = 1.0 / (double)L;
        for (long long i = 0; i < L; ++i) x[i] *= inv;
    }
}

// _fft_hilbert_with_length (production_operators.jl:735-771) and
// _production_fft_hilbert_apply (fft_parallel.jl:168-225): zero-padded linear
// convolution with the kernel 1/(2 pi lag de), lag = -(NE-1)..NE-1, lag != 0.
int orc_fft_hilbert(int NE, long long ncols, const double* eps, const double* wE,
                    const double* Gamma_, double* Lambda_, int L) {
    double de;
    if (orc_uniform_spacing(NE, eps, &de)) return -1;
    if (L < 2 * NE - 1 || (L & (L - 1))) return -1;
    std::vector<cd> kernel((size_t)L, cd(0.0, 0.0));
    for (int lag = -(NE - 1); lag <= NE - 1; ++lag) {
        if (lag == 0) continue;
        kernel[(size_t)(lag + NE - 1)] = cd(1.0 / (kTwoPi * lag * de), 0.0);
    }
    orc_fft(L, (double*)kernel.data(), 0);
    const cd* Gamma = C(Gamma_);
    cd* Lambda = C(Lambda_);
#pragma omp parallel
    {
        std::vector<cd> buf((size_t)L);
#pragma omp for schedule(static)
        for (long long col = 0; col < ncols; ++col) {
            std::fill(buf.begin(), buf.end(), cd(0.0, 0.0));
            for (int e = 0; e < NE; ++e) buf[e] = wE[e] * Gamma[(size_t)col * NE + e];
            orc_fft(L, (double*)buf.data(), 0);
            for (int i = 0; i < L; ++i) buf[i] *= kernel[i];
            orc_fft(L, (double*)buf.data(), 1);
            for (int e = 0; e < NE; ++e) Lambda[(size_t)col * NE + e] = buf[(size_t)e + NE - 1];
        }
    }
    return 0;
}

static int next_pow2(int n) {
    int L = 1;
    while (L < n) L <<= 1;
    return L;
}

// fft_hilbert_transform with return_roundoff (production_operators.jl:792-809;
// fft_parallel.jl:269-291): L = nextpow2(2NE-1), witness with 2L.
int orc_fft_hilbert_checked(int NE, long long ncols, const double* eps, const double* wE,
                            const double* Gamma, double* Lambda, double* roundoff) {
    int L = next_pow2(2 * NE - 1);
    int rc = orc_fft_hilbert(NE, ncols, eps, wE, Gamma, Lambda, L);
    if (rc || !roundoff) return rc;
    size_t n = (size_t)NE * ncols;
    std::vector<cd> second(n);
    rc = orc_fft_hilbert(NE, ncols, eps, wE, Gamma, (double*)second.data(), 2 * L);
    if (rc) return rc;
    const cd* lam = C(Lambda);
    double dn = 0.0;
    for (size_t i = 0; i < n; ++i) dn += std::norm(second[i] - lam[i]);
    *roundoff = std::sqrt(dn) / (field_norm(n, lam) + 1e-14);
    return 0;
}

// direct_hilbert_transform (reference/kernels.jl:457-510): skip the diagonal
// cell, accumulate symmetrically by distance (or in reverse).
int orc_direct_hilbert(int NE, long long ncols, const double* eps, const double* wE,
                       const double* Gamma_, double* Lambda_, int reverse_order) {
    const cd* Gamma = C(Gamma_);
    cd* Lambda = C(Lambda_);
#pragma omp parallel
    {
        std::vector<int> order;
        order.reserve(NE);
#pragma omp for schedule(static)
        for (int e = 0; e < NE; ++e) {
            order.clear();
            for (int d = 1; d <= std::max(e, NE - 1 - e); ++d) {
                if (e - d >= 0) order.push_back(e - d);
                if (e + d <= NE - 1) order.push_back(e + d);
            }
            if (reverse_order) std::reverse(order.begin(), order.end());
            for (long long col = 0; col < ncols; ++col) Lambda[(size_t)col * NE + e] = cd(0.0, 0.0);
            for (int ep : order) {
                double factor = wE[ep] / (kTwoPi * (eps[e] - eps[ep]));
                for (long long col = 0; col < ncols; ++col)
                    Lambda[(size_t)col * NE + e] += factor * Gamma[(size_t)col * NE + ep];
            }
        }
    }
    return 0;
}

// _retarded_self_energy_fft_production (production_operators.jl:842-884):
// Gamma = i(S> - S<), S^R = H[Gamma] - i Gamma/2.
int orc_retarded_self_energy_fft(int NE, long long ncols, const double* eps,
                                 const double* wE, const double* SL_, const double* SG_,
                                 double* SR_, double* Gamma_, double* roundoff) {
    size_t n = (size_t)NE * ncols;
    const cd* SL = C(SL_);
    const cd* SG = C(SG_);
    std::vector<cd> own;
    cd* Gamma = Gamma_ ? C(Gamma_) : nullptr;
    if (!Gamma) {
        own.resize(n);
        Gamma = own.data();
    }
    for (size_t i = 0; i < n; ++i) Gamma[i] = cd(0.0, 1.0) * (SG[i] - SL[i]);
    int rc = orc_fft_hilbert_checked(NE, ncols, eps, wE, (const double*)Gamma, SR_, roundoff);
    if (rc) return rc;
    cd* SR = C(SR_);
    for (size_t i = 0; i < n; ++i) SR[i] = SR[i] - cd(0.0, 0.5) * Gamma[i];
    return 0;
}

// _production_lo_contraction, thermal population (production_operators.jl:1089-1162).
int orc_lo_contraction(int NE, int Nk, int Nb, const orc_kernel* op, const double* GL,
                       const double* GG, const double* wk, double qK, double NLO,
                       const int* pl, const int* pr, const double* pw,
                       const unsigned char* pv, const int* ml, const int* mr,
                       const double* mw, const unsigned char* mv, int energy_chunk,
                       double* SL, double* SG) {
    long long ncols = (long long)Nk * Nb * Nb;
    std::vector<double> shifted((size_t)2 * NE * ncols);
    double* sh = shifted.data();
    int rc = 0;
    orc_apply_shift(NE, ncols, pl, pr, pw, pv, GL, sh);
    rc |= orc_contract(NE, Nk, Nb, op, sh, wk, qK, energy_chunk, NLO + 1.0, 0.0, SL);
    orc_apply_shift(NE, ncols, ml, mr, mw, mv, GL, sh);
    rc |= orc_contract(NE, Nk, Nb, op, sh, wk, qK, energy_chunk, NLO, 1.0, SL);
    orc_apply_shift(NE, ncols, ml, mr, mw, mv, GG, sh);
    rc |= orc_contract(NE, Nk, Nb, op, sh, wk, qK, energy_chunk, NLO + 1.0, 0.0, SG);
    orc_apply_shift(NE, ncols, pl, pr, pw, pv, GG, sh);
    rc |= orc_contract(NE, Nk, Nb, op, sh, wk, qK, energy_chunk, NLO, 1.0, SG);
    return rc;
}

// _embedding_component_production (production_operators.jl:886-922): zero
// outside the shift window, else T * S[G] * T^dagger.
int orc_embedding_component(int NE, int Nk, int Nb, const double* T_, const double* G_,
                            const int* left, const int* right, const double* wr,
                            const unsigned char* valid, double* S_) {
    const cd* T = C(T_);
    const cd* G = C(G_);
    cd* S = C(S_);
    long long blocks = (long long)NE * Nk;
#pragma omp parallel
    {
        std::vector<cd> X((size_t)Nb * Nb), tmp((size_t)Nb * Nb), res((size_t)Nb * Nb);
#pragma omp for schedule(static)
        for (long long linear = 0; linear < blocks; ++linear) {
            int e = (int)(linear % NE), m = (int)(linear / NE);
            if (!valid[e]) {
                for (int b = 0; b < Nb; ++b)
                    for (int a = 0; a < Nb; ++a) S[at4(e, m, a, b, NE, Nk, Nb)] = cd(0.0, 0.0);
                continue;
            }
            int l = left[e], r = right[e];
            double t = wr[e];
            for (int b = 0; b < Nb; ++b)
                for (int a = 0; a < Nb; ++a)
                    X[a + Nb * b] = (l == r || t == 0.0)
                                        ? G[at4(l, m, a, b, NE, Nk, Nb)]
                                        : (1.0 - t) * G[at4(l, m, a, b, NE, Nk, Nb)] + t * G[at4(r, m, a, b, NE, Nk, Nb)];
            matmul(Nb, T, X.data(), tmp.data());
            matmul_adj(Nb, tmp.data(), T, res.data());
            store_block(S, e, m, NE, Nk, Nb, res.data());
        }
    }
    return 0;
}

// _mix_family_production! (production_operators.jl:1475-1491): old <- (1-a) old + a cand.
void orc_mix(long long n, double* old__, const double* cand_, double alpha) {
    cd* old_ = C(old__);
    const cd* cand = C(cand_);
    double b = 1.0 - alpha;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) old_[i] = b * old_[i] + alpha * cand[i];
}

void orc_add(long long n, const double* x_, const double* y_, double* out_) {
    const cd* x = C(x_);
    const cd* y = C(y_);
    cd* out = C(out_);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) out[i] = x[i] + y[i];
}

// _production_boundary_flux_bar (production_iteration.jl:174-201): energies
// in the order sortperm by (|e|, e), momenta inner.
double orc_boundary_flux_bar(int NE, int Nk, int Nb, const double* eps, const double* wE,
                             const double* wk, int gs, const double* plusL_,
                             const double* plusG_, const double* GL_, const double* GG_) {
    const cd* SL = C(plusL_);
    const cd* SG = C(plusG_);
    const cd* GL = C(GL_);
    const cd* GG = C(GG_);
    std::vector<int> order(NE);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        double ax = std::fabs(eps[x]), ay = std::fabs(eps[y]);
        if (ax != ay) return ax < ay;
        return eps[x] < eps[y];
    });
    cd value(0.0, 0.0);
    for (int e : order)
        for (int m = 0; m < Nk; ++m) {
            cd tr(0.0, 0.0);
            for (int a = 0; a < Nb; ++a) {
                cd out(0.0, 0.0), in(0.0, 0.0);
                for (int b = 0; b < Nb; ++b) {
                    out += SG[at4(e, m, a, b, NE, Nk, Nb)] * GL[at4(e, m, b, a, NE, Nk, Nb)];
                    in += SL[at4(e, m, a, b, NE, Nk, Nb)] * GG[at4(e, m, b, a, NE, Nk, Nb)];
                }
                tr += out - in;
            }
            value += (wE[e] * wk[m]) * tr;
        }
    return ((double)gs * value / kTwoPi).real();
}

// _production_state_populations (production_iteration.jl:204-215).
void orc_state_populations(int NE, int Nk, int Nb, const double* wE, const double* wk,
                           int gs, double target, const double* GL_, double* out) {
    const cd* GL = C(GL_);
    std::vector<cd> pops(Nb, cd(0.0, 0.0));
    cd prefactor(0.0, -(double)gs / kTwoPi);
    for (int e = 0; e < NE; ++e)
        for (int m = 0; m < Nk; ++m) {
            cd w = prefactor * wE[e] * wk[m];
            for (int a = 0; a < Nb; ++a) pops[a] += w * GL[at4(e, m, a, a, NE, Nk, Nb)];
        }
    for (int a = 0; a < Nb; ++a) out[a] = pops[a].real() / target;
}

// _fermi (reference/scba.jl:1-6).
double orc_fermi(double e, double mu, double kBT) {
    double x = (e - mu) / kBT;
    if (x > kFermiSaturation) return 0.0;
    if (x < -kFermiSaturation) return 1.0;
    return 1.0 / (std::exp(x) + 1.0);
}

// _seed_number_weights_production (seed.jl:28-66).
int orc_seed_weights(int NE, int Nk, int Nb, const double* GR_, const double* wE,
                     const double* wk, int gs, double* q) {
    const cd* GR = C(GR_);
    double prefactor = (double)gs / kTwoPi;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int e = 0; e < NE; ++e) {
        double msum = 0.0;
        for (int m = 0; m < Nk; ++m) {
            double trA = 0.0;
            for (int a = 0; a < Nb; ++a) {
                cd z = GR[at4(e, m, a, a, NE, Nk, Nb)];
                trA += (cd(0.0, 1.0) * (z - std::conj(z))).real();
            }
            msum += wk[m] * trA;
        }
        q[e] = prefactor * wE[e] * msum;
        if (!std::isfinite(q[e])) bad = 1;
    }
    return bad ? -3 : 0;
}

static double seed_number(double mu, int NE, const double* eps, const double* q, double kBT) {
    double v = 0.0;
    for (int e = 0; e < NE; ++e) v += q[e] * orc_fermi(eps[e], mu, kBT);
    return v;
}

// Chemical potential of the seed (seed.jl:175-182): bracket
// [e_1 - 50kT, e_N + 50kT], bisection to adjacent floating-point numbers.
int orc_seed_mu(int NE, const double* eps, const double* q, double kBT, double target,
                double* mu) {
    if (!(kBT > 0)) return -1;
    double lo = eps[0] - 50.0 * kBT, hi = eps[NE - 1] + 50.0 * kBT;
    double flo = seed_number(lo, NE, eps, q, kBT) - target;
    double fhi = seed_number(hi, NE, eps, q, kBT) - target;
    if (!(flo * fhi <= 0)) return -1;
    if (flo == 0) { *mu = lo; return 0; }
    if (fhi == 0) { *mu = hi; return 0; }
    for (int it = 0; it < 4000; ++it) {
        double mid = lo + 0.5 * (hi - lo);
        if (mid <= lo || mid >= hi) break;
        double fm = seed_number(mid, NE, eps, q, kBT) - target;
        if (fm == 0) { *mu = mid; return 0; }
        if ((fm < 0) == (flo < 0)) { lo = mid; flo = fm; } else { hi = mid; fhi = fm; }
    }
    *mu = (std::fabs(flo) <= std::fabs(fhi)) ? lo : hi;
    return 0;
}

// _seed_lesser_production! (seed.jl:99-128): G< = -f (G^R - G^R^dagger).
void orc_seed_lesser(int NE, int Nk, int Nb, const double* GR_, const double* eps,
                     double mu, double kBT, double* GL_) {
    const cd* GR = C(GR_);
    cd* GL = C(GL_);
    long long blocks = (long long)NE * Nk;
#pragma omp parallel for schedule(static)
    for (long long linear = 0; linear < blocks; ++linear) {
        int e = (int)(linear % NE), m = (int)(linear / NE);
        double f = orc_fermi(eps[e], mu, kBT);
        for (int b = 0; b < Nb; ++b)
            for (int a = 0; a < Nb; ++a)
                GL[at4(e, m, a, b, NE, Nk, Nb)] =
                    -f * (GR[at4(e, m, a, b, NE, Nk, Nb)] - std::conj(GR[at4(e, m, b, a, NE, Nk, Nb)]));
    }
}

// One production SCBA pass (production_solver.jl:435-844): Dyson/Keldysh on
// the current total, fresh candidates (scattering with FFT Hilbert and its
// round-off witness, embedding), observables from the current state, linear
// mixing of every family, embedding total = plus + minus.
static int scba_iteration_core(const orc_problem* p, orc_state* s, double* out, double* populations,
                               double* meas, double* const* cand_out) {
    const int NE = p->NE, Nk = p->Nk, Nb = p->Nb, M = p->nmech;
    const size_t n = (size_t)NE * Nk * Nb * Nb;
    const long long ncols = (long long)Nk * Nb * Nb;
    std::vector<const double*> fR(M + 1), fL(M + 1), fG(M + 1);
    fR[0] = s->embR;
    fL[0] = s->embL;
    fG[0] = s->embG;
    for (int i = 0; i < M; ++i) {
        fR[i + 1] = s->scatR[i];
        fL[i + 1] = s->scatL[i];
        fG[i + 1] = s->scatG[i];
    }
    double o3[3];
    int rc = orc_current_green(NE, Nk, Nb, p->eps, p->h, M + 1, fR.data(), fL.data(), fG.data(),
                               p->wE, p->wk, p->gs, p->target, s->GR, s->GL, s->GG, s->A,
                               s->scale, o3);
    if (rc) return rc;
    std::vector<int> pl(NE), pr(NE), ml(NE), mr(NE);
    std::vector<double> pw(NE), mw(NE);
    std::vector<unsigned char> pv(NE), mv(NE);
    // Candidate scattering.
    std::vector<std::vector<double>> cR(M), cL(M), cG(M);
    double roundoff = 0.0;
    if (orc_shift_plan(NE, p->eps, +p->hwLO, pl.data(), pr.data(), pw.data(), pv.data())) return -1;
    if (orc_shift_plan(NE, p->eps, -p->hwLO, ml.data(), mr.data(), mw.data(), mv.data())) return -1;
    for (int i = 0; i < M; ++i) {
        cR[i].resize(2 * n);
        cL[i].resize(2 * n);
        cG[i].resize(2 * n);
        if (p->is_lo[i]) {
            rc = orc_lo_contraction(NE, Nk, Nb, &p->kernels[i], s->GL, s->GG, p->wk, p->qK[i], p->NLO,
                                    pl.data(), pr.data(), pw.data(), pv.data(), ml.data(), mr.data(),
                                    mw.data(), mv.data(), p->energy_chunk, cL[i].data(), cG[i].data());
        } else {
            rc = orc_contract(NE, Nk, Nb, &p->kernels[i], s->GL, p->wk, p->qK[i], p->energy_chunk, 1.0,
                              0.0, cL[i].data());
            rc |= orc_contract(NE, Nk, Nb, &p->kernels[i], s->GG, p->wk, p->qK[i], p->energy_chunk, 1.0,
                               0.0, cG[i].data());
        }
        if (rc) return rc;
        double rr = 0.0;
        rc = orc_retarded_self_energy_fft(NE, ncols, p->eps, p->wE, cL[i].data(), cG[i].data(),
                                          cR[i].data(), nullptr, &rr);
        if (rc) return rc;
        roundoff = std::max(roundoff, rr);
    }
    // Candidate embedding.
    if (orc_shift_plan(NE, p->eps, +p->Ep, pl.data(), pr.data(), pw.data(), pv.data())) return -1;
    if (orc_shift_plan(NE, p->eps, -p->Ep, ml.data(), mr.data(), mw.data(), mv.data())) return -1;
    std::vector<double> cp[3], cm[3];
    const double* X[3] = {s->GR, s->GL, s->GG};
    for (int k = 0; k < 3; ++k) {
        cp[k].resize(2 * n);
        cm[k].resize(2 * n);
        orc_embedding_component(NE, Nk, Nb, p->Tplus, X[k], pl.data(), pr.data(), pw.data(), pv.data(),
                                cp[k].data());
        orc_embedding_component(NE, Nk, Nb, p->Tminus, X[k], ml.data(), mr.data(), mw.data(), mv.data(),
                                cm[k].data());
    }
    // Fixed-point measurement before mixing (production_iteration.jl:102-166):
    // the residual suite against the totals of this iteration and the unmixed
    // candidate total, the self-energy residuals of the embedding, plus,
    // minus and every mechanism, and |lambda - 1|.
    if (meas) {
        std::vector<double> tot[3], candL(2 * n), ce[3];
        const double* embp[3] = {s->embR, s->embL, s->embG};
        for (int k = 0; k < 3; ++k) {
            tot[k].assign(embp[k], embp[k] + 2 * n);
            for (int i = 0; i < M; ++i) {
                const double* x = k == 0 ? s->scatR[i] : (k == 1 ? s->scatL[i] : s->scatG[i]);
                orc_add((long long)n, tot[k].data(), x, tot[k].data());
            }
            ce[k].resize(2 * n);
            orc_add((long long)n, cp[k].data(), cm[k].data(), ce[k].data());
        }
        candL = ce[1];
        for (int i = 0; i < M; ++i) orc_add((long long)n, candL.data(), cL[i].data(), candL.data());
        rc = orc_residual_suite(NE, Nk, Nb, p->eps, p->h, s->GR, s->GL, s->GG, s->A, tot[0].data(), tot[1].data(),
                                tot[2].data(), candL.data(), meas);
        if (rc) return rc;
        const double* oe[3] = {s->embR, s->embL, s->embG};
        const double* cep[3] = {ce[0].data(), ce[1].data(), ce[2].data()};
        const double* op[3] = {s->plusR, s->plusL, s->plusG};
        const double* cpp[3] = {cp[0].data(), cp[1].data(), cp[2].data()};
        const double* om[3] = {s->minusR, s->minusL, s->minusG};
        const double* cmp[3] = {cm[0].data(), cm[1].data(), cm[2].data()};
        double rs = orc_selfenergy_residual((long long)n, oe, cep);
        rs = std::max(rs, orc_selfenergy_residual((long long)n, op, cpp));
        rs = std::max(rs, orc_selfenergy_residual((long long)n, om, cmp));
        for (int i = 0; i < M; ++i) {
            const double* os[3] = {s->scatR[i], s->scatL[i], s->scatG[i]};
            const double* cs[3] = {cR[i].data(), cL[i].data(), cG[i].data()};
            rs = std::max(rs, orc_selfenergy_residual((long long)n, os, cs));
        }
        meas[5] = rs;
        meas[6] = std::fabs(o3[0] - 1.0);
    }
    // Observables of the current state (production_iteration.jl:218-236).
    double Jbar = orc_boundary_flux_bar(NE, Nk, Nb, p->eps, p->wE, p->wk, p->gs, s->plusL, s->plusG,
                                        s->GL, s->GG);
    if (populations) orc_state_populations(NE, Nk, Nb, p->wE, p->wk, p->gs, p->target, s->GL, populations);
    if (cand_out) {
        // Hand the unmixed candidates over (mechanisms R,<,>, then plus, minus)
        // for an outside mixing step; the state keeps its self-energies.
        for (int i = 0; i < M; ++i) {
            std::copy(cR[i].begin(), cR[i].end(), cand_out[3 * i + 0]);
            std::copy(cL[i].begin(), cL[i].end(), cand_out[3 * i + 1]);
            std::copy(cG[i].begin(), cG[i].end(), cand_out[3 * i + 2]);
        }
        for (int k = 0; k < 3; ++k) {
            std::copy(cp[k].begin(), cp[k].end(), cand_out[3 * M + k]);
            std::copy(cm[k].begin(), cm[k].end(), cand_out[3 * M + 3 + k]);
        }
        if (out) {
            out[0] = o3[0];
            out[1] = Jbar;
            out[2] = roundoff;
            out[3] = o3[1];
            out[4] = o3[2];
        }
        return 0;
    }
    // Linear mixing (production_solver.jl:819-832).
    for (int i = 0; i < M; ++i) {
        orc_mix((long long)n, s->scatR[i], cR[i].data(), p->alpha_mix);
        orc_mix((long long)n, s->scatL[i], cL[i].data(), p->alpha_mix);
        orc_mix((long long)n, s->scatG[i], cG[i].data(), p->alpha_mix);
    }
    double* plus[3] = {s->plusR, s->plusL, s->plusG};
    double* minus[3] = {s->minusR, s->minusL, s->minusG};
    double* emb[3] = {s->embR, s->embL, s->embG};
    for (int k = 0; k < 3; ++k) {
        orc_mix((long long)n, plus[k], cp[k].data(), p->alpha_mix);
        orc_mix((long long)n, minus[k], cm[k].data(), p->alpha_mix);
        orc_add((long long)n, plus[k], minus[k], emb[k]);
    }
    if (out) {
        out[0] = o3[0];
        out[1] = Jbar;
        out[2] = roundoff;
        out[3] = o3[1];
        out[4] = o3[2];
    }
    return 0;
}

int orc_scba_iteration(const orc_problem* p, orc_state* s, double* out, double* populations) {
    return scba_iteration_core(p, s, out, populations, nullptr, nullptr);
}

int orc_scba_iteration_measured(const orc_problem* p, orc_state* s, double* out, double* populations,
                                double* meas) {
    return scba_iteration_core(p, s, out, populations, meas, nullptr);
}

int orc_scba_iteration_candidates(const orc_problem* p, orc_state* s, double* out, double* populations,
                                  double* meas, double* const* cand_out) {
    return scba_iteration_core(p, s, out, populations, meas, cand_out);
}

}  // extern "C"

namespace {
// Eigenvalues of a Hermitian n x n block (column-major) by cyclic two-sided
// complex Jacobi rotations, ascending. Each rotation J = diag(1, u*) R with
// u = H_pq/|H_pq| and R the real symmetric Schur rotation of
// [[H_pp, |H_pq|], [|H_pq|, H_qq]] (Golub-Van Loan 8.5.2); H <- J^H H J.
void herm_eigvals(int n, const cd* Hin, double* w) {
    std::vector<cd> H((size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) H[i + n * j] = 0.5 * (Hin[i + n * j] + std::conj(Hin[j + n * i]));
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const double v = std::norm(H[i + n * j]);
                tot += v;
                if (i != j) off += v;
            }
        if (off == 0.0 || off <= 1e-34 * tot) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const cd b = H[p + n * q];
                const double ab = std::abs(b);
                if (ab == 0.0) continue;
                const double a = H[p + n * p].real(), d = H[q + n * q].real();
                const double tau = (d - a) / (2.0 * ab);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                const cd u = b / ab;
                for (int i = 0; i < n; ++i) {  // columns: H J
                    const cd hp = H[i + n * p], hq = H[i + n * q];
                    H[i + n * p] = c * hp - s * std::conj(u) * hq;
                    H[i + n * q] = s * hp + c * std::conj(u) * hq;
                }
                for (int j = 0; j < n; ++j) {  // rows: J^H H
                    const cd hp = H[p + n * j], hq = H[q + n * j];
                    H[p + n * j] = c * hp - s * u * hq;
                    H[q + n * j] = s * hp + c * u * hq;
                }
                H[p + n * q] = H[q + n * p] = cd(0.0, 0.0);
                H[p + n * p] = cd(H[p + n * p].real(), 0.0);
                H[q + n * q] = cd(H[q + n * q].real(), 0.0);
            }
    }
    for (int i = 0; i < n; ++i) w[i] = H[i + n * i].real();
    std::sort(w, w + n);
}

bool all_finite(int n, const cd* x) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(x[i].real()) || !std::isfinite(x[i].imag())) return false;
    return true;
}

double frob2(int n2, const cd* x) {
    double s = 0.0;
    for (int i = 0; i < n2; ++i) s += std::norm(x[i]);
    return s;
}

// Positivity ratio of the Hermitian part of f X (residuals.jl:413-452 with
// the witness of positivity_diagnostics.jl:61-96): max(0, -lambda_min -
// 64 n eps max(|lambda|_max, 1)) / max(|lambda|_max, floatmin).
double negative_part(int n, const cd* X, cd f, cd* H, double* w) {
    if (!all_finite(n * n, X)) return INFINITY;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) H[i + n * j] = 0.5 * (f * X[i + n * j] + std::conj(f * X[j + n * i]));
    herm_eigvals(n, H, w);
    double big = 0.0;
    for (int i = 0; i < n; ++i) big = std::max(big, std::fabs(w[i]));
    const double budget = 64.0 * n * 2.220446049250313e-16 * std::max(big, 1.0);
    const double absolute = std::max(0.0, -w[0]);
    return std::max(0.0, absolute - budget) / std::max(big, 2.2250738585072014e-308);
}

// Causality of one retarded block (residuals.jl:455-492): lambda_max of
// Im S^R = (S - S^dagger)/(2i) over the spectral norm of S^R (+1e-14).
double causality_block(int n, const cd* S, cd* H, cd* T, double* w) {
    if (!all_finite(n * n, S)) return INFINITY;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) H[i + n * j] = (S[i + n * j] - std::conj(S[j + n * i])) / cd(0.0, 2.0);
    herm_eigvals(n, H, w);
    const double numerator = std::max(0.0, w[n - 1]);
    if (numerator == 0.0) return 0.0;
    // ||S||_2 = sqrt(lambda_max(S^dagger S))
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            cd v(0.0, 0.0);
            for (int c = 0; c < n; ++c) v += std::conj(S[c + n * i]) * S[c + n * j];
            T[i + n * j] = v;
        }
    herm_eigvals(n, T, w);
    const double norm2 = std::sqrt(std::max(0.0, w[n - 1]));
    return numerator / (norm2 + 1e-14);
}

}  // namespace

extern "C" {

void orc_hermitian_eigvals(int n, const double* H, double* w) { herm_eigvals(n, C(H), w); }

/* production_residual_suite (residuals.jl:537-772): out = r_D, r_A, r_K,
 * r_PSD, r_caus. Totals S* and the unmixed candidate S< of the iteration. */
int orc_residual_suite(int NE, int Nk, int Nb, const double* eps, const double* h_, const double* GR_,
                       const double* GL_, const double* GG_, const double* A_, const double* SR_,
                       const double* SL_, const double* SG_, const double* cand_, double* out) {
    const cd *h = C(h_), *GRf = C(GR_), *GLf = C(GL_), *GGf = C(GG_), *Af = C(A_);
    const cd *SRf = C(SR_), *SLf = C(SL_), *SGf = C(SG_), *Cf = C(cand_);
    const long long nblk = (long long)NE * Nk;
    const int n = Nb, n2 = Nb * Nb;
    std::vector<double> specN(nblk), specD(nblk), kelN(nblk), kelD(nblk), rDv(nblk), psdv(nblk), causv(nblk);
#pragma omp parallel
    {
        std::vector<cd> G(n2), GL(n2), GG(n2), A(n2), SR(n2), SL(n2), SG(n2), CL(n2), D(n2), Gm(n2), T1(n2), T2(n2),
            H(n2), T3(n2);
        std::vector<double> w(n);
#pragma omp for schedule(static)
        for (long long linear = 0; linear < nblk; ++linear) {
            const int e = (int)(linear / Nk), m = (int)(linear % Nk);  // the reference's order: m fastest
            load_block(GRf, e, m, NE, Nk, Nb, G.data());
            load_block(GLf, e, m, NE, Nk, Nb, GL.data());
            load_block(GGf, e, m, NE, Nk, Nb, GG.data());
            load_block(Af, e, m, NE, Nk, Nb, A.data());
            load_block(SRf, e, m, NE, Nk, Nb, SR.data());
            load_block(SLf, e, m, NE, Nk, Nb, SL.data());
            load_block(SGf, e, m, NE, Nk, Nb, SG.data());
            load_block(Cf, e, m, NE, Nk, Nb, CL.data());
            // r_D = ||D G - I||_F / (||D||_F ||G||_F + sqrt(Nb)),  D = eps - h - S^R
            for (int b = 0; b < n; ++b)
                for (int a = 0; a < n; ++a)
                    D[a + n * b] = (a == b ? cd(eps[e], 0.0) : cd(0.0, 0.0)) - h[(size_t)m + (size_t)Nk * (a + (size_t)n * b)] -
                                   SR[a + n * b];
            matmul(n, D.data(), G.data(), T1.data());
            for (int a = 0; a < n; ++a) T1[a + n * a] -= 1.0;
            {
                const double num = std::sqrt(frob2(n2, T1.data()));
                const double den = std::sqrt(frob2(n2, D.data())) * std::sqrt(frob2(n2, G.data())) + std::sqrt((double)n);
                rDv[linear] = std::isfinite(num) && std::isfinite(den) ? num / den : INFINITY;
            }
            // Gamma = i (S> - S<)
            for (int i = 0; i < n2; ++i) Gm[i] = cd(0.0, 1.0) * (SG[i] - SL[i]);
            // r_A terms: ||A - G Gamma G^dagger||^2, ||A||^2
            matmul(n, G.data(), Gm.data(), T1.data());
            matmul_adj(n, T1.data(), G.data(), T2.data());
            for (int i = 0; i < n2; ++i) T2[i] = A[i] - T2[i];
            specN[linear] = frob2(n2, T2.data());
            specD[linear] = frob2(n2, A.data());
            // r_K terms against the unmixed candidate S<
            matmul(n, G.data(), CL.data(), T1.data());
            matmul_adj(n, T1.data(), G.data(), T2.data());
            for (int i = 0; i < n2; ++i) T2[i] = GL[i] - T2[i];
            kelN[linear] = frob2(n2, T2.data());
            kelD[linear] = frob2(n2, GL.data());
            // r_PSD over A, -iG<, iG>, Gamma, -i G S< G^dagger, i G S> G^dagger
            double psd = 0.0;
            psd = std::max(psd, negative_part(n, A.data(), cd(1.0, 0.0), H.data(), w.data()));
            psd = std::max(psd, negative_part(n, GL.data(), cd(0.0, -1.0), H.data(), w.data()));
            psd = std::max(psd, negative_part(n, GG.data(), cd(0.0, 1.0), H.data(), w.data()));
            psd = std::max(psd, negative_part(n, Gm.data(), cd(1.0, 0.0), H.data(), w.data()));
            matmul(n, G.data(), SL.data(), T1.data());
            matmul_adj(n, T1.data(), G.data(), T2.data());
            psd = std::max(psd, negative_part(n, T2.data(), cd(0.0, -1.0), H.data(), w.data()));
            matmul(n, G.data(), SG.data(), T1.data());
            matmul_adj(n, T1.data(), G.data(), T2.data());
            psd = std::max(psd, negative_part(n, T2.data(), cd(0.0, 1.0), H.data(), w.data()));
            psdv[linear] = psd;
            causv[linear] = causality_block(n, SR.data(), H.data(), T3.data(), w.data());
        }
    }
    // The reference's canonical reductions (residuals.jl:495-512): a
    // non-finite entry makes the sum or maximum Inf.
    auto csum = [](const std::vector<double>& v) {
        double r = 0.0;
        for (double x : v) {
            if (!std::isfinite(x)) return (double)INFINITY;
            r += x;
            if (!std::isfinite(r)) return (double)INFINITY;
        }
        return r;
    };
    auto cmax = [](const std::vector<double>& v) {
        double r = 0.0;
        for (double x : v) {
            if (!std::isfinite(x)) return (double)INFINITY;
            r = std::max(r, x);
        }
        return r;
    };
    auto ratio = [](double num, double den) {
        const double v = std::sqrt(num) / (std::sqrt(den) + 1e-14);
        return std::isfinite(v) ? v : (double)INFINITY;
    };
    out[0] = cmax(rDv);
    out[1] = ratio(csum(specN), csum(specD));
    out[2] = ratio(csum(kelN), csum(kelD));
    out[3] = cmax(psdv);
    out[4] = cmax(causv);
    return 0;
}

/* production_selfenergy_residual (residuals.jl:802-865): max over R, <, >
 * of ||cand - old||_F / (||cand||_F + 1e-14). */
double orc_selfenergy_residual(long long n, const double* const* old_, const double* const* cand_) {
    double r = 0.0;
    for (int k = 0; k < 3; ++k) {
        const cd *o = C(old_[k]), *c = C(cand_[k]);
        double dn = 0.0, cn = 0.0;
        bool finite = true;
        for (long long i = 0; i < n; ++i) {
            if (!std::isfinite(c[i].real()) || !std::isfinite(c[i].imag()) || !std::isfinite(o[i].real()) ||
                !std::isfinite(o[i].imag()))
                finite = false;
            dn += std::norm(c[i] - o[i]);
            cn += std::norm(c[i]);
        }
        const double v = std::sqrt(dn) / (std::sqrt(cn) + 1e-14);
        r = std::max(r, finite && std::isfinite(v) ? v : (double)INFINITY);
    }
    return r;
}

}  // extern "C"

extern "C" {

/* _electron_density_bar (src/physics/observables.jl:1-26): integrate every
 * coherence I_ab = sum_{m,e} wE_e wk_m G<_{e,m,a,b} (m outer, e inner), then
 * nbar_g = gs/(2 pi) Re(-i sum_{b,a} chi_{g,a} I_ab conj(chi_{g,b})).
 * chi: Nz x Nb complex, column-major. */
int orc_electron_density(int NE, int Nk, int Nb, int Nz, const double* GL_, const double* wE, const double* wk,
                         const double* chi_, double gs, double* nbar) {
    const cd* GL = C(GL_);
    const cd* chi = C(chi_);
    std::vector<cd> I((size_t)Nb * Nb);
    for (int b = 0; b < Nb; ++b)
        for (int a = 0; a < Nb; ++a) {
            cd v(0.0, 0.0);
            for (int m = 0; m < Nk; ++m)
                for (int e = 0; e < NE; ++e) v += wE[e] * wk[m] * GL[at4(e, m, a, b, NE, Nk, Nb)];
            I[a + (size_t)Nb * b] = v;
        }
    const double pref = gs / (2.0 * M_PI);
    for (int g = 0; g < Nz; ++g) {
        cd v(0.0, 0.0);
        for (int b = 0; b < Nb; ++b)
            for (int a = 0; a < Nb; ++a)
                v += chi[g + (size_t)Nz * a] * I[a + (size_t)Nb * b] * std::conj(chi[g + (size_t)Nz * b]);
        nbar[g] = pref * (cd(0.0, -1.0) * v).real();
    }
    return 0;
}

/* solve_periodic_poisson (src/numerics/reference/poisson.jl:12-65): the
 * cyclic finite-volume matrix L of -d/dx(eps_r du/dx) (harmonic-mean faces,
 * dx = wx_1), b = -lambda_P (N_D - nbar), the bordered system
 * [[L/sP, v], [v^T, 0]] [U; zeta] = [b/sP; 0] (sP the largest absolute row
 * sum, v = 1/sqrt(N)) solved by LU with partial pivoting; out = zeta,
 * r_P = ||L U - b|| / (||b|| + lambda_P 1e-14), neutrality
 * |sum wx (N_D - nbar)| / sum wx N_D. Returns -2 for a singular system. */
int orc_periodic_poisson(int N, const double* eps_r, const double* wx, const double* ND, double lambdaP,
                         const double* nbar, double* U, double* out) {
    const double dx = wx[0];
    std::vector<double> face(N), L((size_t)N * N, 0.0);
    for (int g = 0; g < N; ++g) {
        const int gp = g == N - 1 ? 0 : g + 1;
        face[g] = 2.0 * eps_r[g] * eps_r[gp] / (eps_r[g] + eps_r[gp]);
    }
    for (int g = 0; g < N; ++g) {
        const int gm = g == 0 ? N - 1 : g - 1, gp = g == N - 1 ? 0 : g + 1;
        L[g + (size_t)N * g] = (face[gm] + face[g]) / (dx * dx);
        L[g + (size_t)N * gm] = -face[gm] / (dx * dx);
        L[g + (size_t)N * gp] = -face[g] / (dx * dx);
    }
    double nd2d = 0.0, neut = 0.0;
    for (int g = 0; g < N; ++g) {
        nd2d += wx[g] * ND[g];
        neut += wx[g] * (ND[g] - nbar[g]);
    }
    std::vector<double> b(N);
    for (int g = 0; g < N; ++g) b[g] = -lambdaP * (ND[g] - nbar[g]);
    double sP = 0.0;
    for (int g = 0; g < N; ++g) {
        double r = 0.0;
        for (int j = 0; j < N; ++j) r += std::fabs(L[g + (size_t)N * j]);
        sP = std::max(sP, r);
    }
    const int n = N + 1;
    std::vector<double> M((size_t)n * n, 0.0), x(n);
    const double v = 1.0 / std::sqrt((double)N);
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) M[i + (size_t)n * j] = L[i + (size_t)N * j] / sP;
    for (int i = 0; i < N; ++i) {
        M[i + (size_t)n * N] = v;
        M[N + (size_t)n * i] = v;
        x[i] = b[i] / sP;
    }
    x[N] = 0.0;
    // LU with partial pivoting (largest magnitude, first on ties), then the
    // two triangular solves.
    std::vector<int> piv(n);
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(M[i + (size_t)n * k]) > std::fabs(M[p + (size_t)n * k])) p = i;
        piv[k] = p;
        if (M[p + (size_t)n * k] == 0.0) return -2;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(M[k + (size_t)n * j], M[p + (size_t)n * j]);
        const double d = M[k + (size_t)n * k];
        for (int i = k + 1; i < n; ++i) M[i + (size_t)n * k] /= d;
        for (int j = k + 1; j < n; ++j) {
            const double f = M[k + (size_t)n * j];
            if (f == 0.0) continue;
            for (int i = k + 1; i < n; ++i) M[i + (size_t)n * j] -= M[i + (size_t)n * k] * f;
        }
    }
    for (int k = 0; k < n; ++k)
        if (piv[k] != k) std::swap(x[k], x[piv[k]]);
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) x[i] -= M[i + (size_t)n * j] * x[j];
    for (int j = n - 1; j >= 0; --j) {
        x[j] /= M[j + (size_t)n * j];
        for (int i = 0; i < j; ++i) x[i] -= M[i + (size_t)n * j] * x[j];
    }
    for (int g = 0; g < N; ++g) U[g] = x[g];
    double rn = 0.0, bn = 0.0;
    for (int g = 0; g < N; ++g) {
        double r = 0.0;
        for (int j = 0; j < N; ++j) r += L[g + (size_t)N * j] * U[j];
        rn += (r - b[g]) * (r - b[g]);
        bn += b[g] * b[g];
    }
    out[0] = x[N];
    out[1] = std::sqrt(rn) / (std::sqrt(bn) + lambdaP * 1e-14);
    out[2] = std::fabs(neut) / nd2d;
    return 0;
}

/* project_hamiltonians (src/numerics/reference/basis.jl:288-308):
 * h_m = (S + Phi^dagger diag(U) Phi) + kin_m Minv, with S = H0 - F (Z - z0)
 * precomputed by the caller and kin_m = C kappa_m^2. Phi: Nz x Nb complex
 * column-major; S, Minv: Nb x Nb column-major; h: (Nk, Nb, Nb) column-major. */
int orc_project_hamiltonians(int Nz, int Nb, int Nk, const double* Phi_, const double* U, const double* S_,
                             const double* kin, const double* Minv_, double* h_) {
    const cd *Phi = C(Phi_), *S = C(S_), *Minv = C(Minv_);
    cd* h = C(h_);
    for (int b = 0; b < Nb; ++b)
        for (int a = 0; a < Nb; ++a) {
            cd up(0.0, 0.0);
            for (int g = 0; g < Nz; ++g)
                up += (std::conj(Phi[g + (size_t)Nz * a]) * U[g]) * Phi[g + (size_t)Nz * b];
            const cd st = S[a + (size_t)Nb * b] + up;
            for (int m = 0; m < Nk; ++m)
                h[m + (size_t)Nk * (a + (size_t)Nb * b)] = st + kin[m] * Minv[a + (size_t)Nb * b];
        }
    return 0;
}

/* Outer mixing of the Hartree energy (production_solver.jl:1046-1047):
 * U <- (1 - alpha) U + alpha cand; returns max |cand - U_old|. */
double orc_mix_potential(int N, double* U, const double* cand, double alpha) {
    double r = 0.0;
    for (int g = 0; g < N; ++g) {
        r = std::max(r, std::fabs(cand[g] - U[g]));
        U[g] = (1.0 - alpha) * U[g] + alpha * cand[g];
    }
    return r;
}

}  // extern "C"

namespace {
// ||X||_2 of an n x n block: sqrt(lambda_max(X^dagger X)).
double opnorm2(int n, const cd* X) {
    std::vector<cd> T((size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            cd v(0.0, 0.0);
            for (int c = 0; c < n; ++c) v += std::conj(X[c + (size_t)n * i]) * X[c + (size_t)n * j];
            T[i + (size_t)n * j] = v;
        }
    std::vector<double> w(n);
    herm_eigvals(n, T.data(), w.data());
    return std::sqrt(std::max(0.0, w[n - 1]));
}

// _anderson_block_failure! (production_mixing.jl:36-80), exact norms and
// eigenvalues: 0 admissible, else the first failing code.
int anderson_block_failure(int n, const cd* R, const cd* L, const cd* G) {
    const double eps = 2.220446049250313e-16, tiny = 2.2250738585072014e-308;
    auto budget = [&](double norm) { return 64.0 * n * eps * std::max(norm, 1.0); };
    if (!all_finite(n * n, R)) return 1;
    std::vector<cd> H((size_t)n * n), D((size_t)n * n);
    std::vector<double> w(n);
    for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a)
            H[a + (size_t)n * b] = (R[a + (size_t)n * b] - std::conj(R[b + (size_t)n * a])) / cd(0.0, 2.0);
    const double rnorm = opnorm2(n, R);
    herm_eigvals(n, H.data(), w.data());
    if (!(w[n - 1] <= 64.0 * eps * std::max(rnorm, tiny))) return 2;
    if (!all_finite(n * n, L) || !all_finite(n * n, G)) return 1;
    const cd* blocks[2] = {L, G};
    const cd factors[2] = {cd(0.0, -1.0), cd(0.0, 1.0)};
    double norms[2];
    for (int k = 0; k < 2; ++k) {
        const cd* X = blocks[k];
        const cd f = factors[k];
        for (int i = 0; i < n * n; ++i) D[i] = f * X[i];
        norms[k] = opnorm2(n, D.data());
        const double bud = budget(norms[k]);
        double an = 0.0;
        for (int b = 0; b < n; ++b)
            for (int a = 0; a < n; ++a) {
                const cd value = f * X[a + (size_t)n * b], conjugate = std::conj(f * X[b + (size_t)n * a]);
                an += std::norm(value - conjugate);
                H[a + (size_t)n * b] = (value + conjugate) / 2.0;
            }
        if (!(std::sqrt(an) / 2.0 <= bud)) return 3;
        herm_eigvals(n, H.data(), w.data());
        if (!(w[0] >= -bud)) return 4;
    }
    double d5 = 0.0;
    for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a)
            d5 += std::norm(R[a + (size_t)n * b] - std::conj(R[b + (size_t)n * a]) -
                            (G[a + (size_t)n * b] - L[a + (size_t)n * b]));
    const double lnorm = opnorm2(n, L), gnorm = opnorm2(n, G);
    if (!(std::sqrt(d5) <= budget(std::max(rnorm, std::max(lnorm, gnorm))))) return 5;
    return 0;
}
}  // namespace

extern "C" {

/* _anderson_inner (production_mixing.jl:135-144): sum over the fields, in
 * order, of Re <a_f, b_f>. */
double orc_anderson_inner(long long n, int nfields, const double* const* a_, const double* const* b_) {
    double v = 0.0;
    for (int f = 0; f < nfields; ++f) {
        const cd *a = C(a_[f]), *b = C(b_[f]);
        cd d(0.0, 0.0);
        for (long long i = 0; i < n; ++i) d += std::conj(a[i]) * b[i];
        v += d.real();
    }
    return v;
}

/* The Anderson combination (production_mixing.jl:256-268): out_f =
 * sum_h c_h (X_h,f + damping R_h,f); X, R entry-major (count x nfields). */
void orc_anderson_combine(long long n, int nfields, int count, const double* coef, double damping,
                          const double* const* X_, const double* const* R_, double* const* out_) {
    for (int f = 0; f < nfields; ++f) {
        cd* out = C(out_[f]);
        for (long long i = 0; i < n; ++i) {
            cd v(0.0, 0.0);
            for (int h = 0; h < count; ++h)
                v += coef[h] * (C(X_[(size_t)h * nfields + f])[i] + damping * C(R_[(size_t)h * nfields + f])[i]);
            out[i] = v;
        }
    }
}

/* _anderson_guard (production_mixing.jl:82-111): first failing block in the
 * order family, e, m (one-based; 0 when all pass) and its code. */
int orc_anderson_guard(int NE, int Nk, int Nb, int nfam, const double* const* R_, const double* const* L_,
                       const double* const* G_, long long* first, int* code) {
    const int n = Nb;
    std::vector<cd> R((size_t)n * n), L((size_t)n * n), G((size_t)n * n);
    *first = 0;
    *code = 0;
    long long index = 0;
    for (int f = 0; f < nfam; ++f)
        for (int e = 0; e < NE; ++e)
            for (int m = 0; m < Nk; ++m) {
                ++index;
                load_block(C(R_[f]), e, m, NE, Nk, Nb, R.data());
                load_block(C(L_[f]), e, m, NE, Nk, Nb, L.data());
                load_block(C(G_[f]), e, m, NE, Nk, Nb, G.data());
                const int c = anderson_block_failure(n, R.data(), L.data(), G.data());
                if (c) {
                    *first = index;
                    *code = c;
                    return 0;
                }
            }
    return 0;
}

}  // extern "C"
