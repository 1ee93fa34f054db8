// LNV density-matrix minimisation on device matrices: the LNV functional and
// its gradients (LNV_matrix_multiply, src/mult_module.f90:1478-1823), the
// electron-number correction (correct_electron_number_fixspin,
// src/DMMinModule.f90:1676-1817, SolveCubic :2072-2143) and FindMinDM
// (src/DMMinModule.f90:172-312) with earlyDM (:366-723), lineMinL
// (:1395-1614), lateDM (:809-1339) and DoPulay (src/Pulay.f90:188-262).
//
// Host orchestration over the matrix API only: every product, sum, trace and
// transpose runs on the device through cqmm_matrix_*; scalars (energies,
// residuals, step lengths, Pulay coefficients) are FP64 on the host. One spin
// (spin-unpolarised: totals carry the factor 2).
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/cqmm.h"
#include "../../include/cqmm_matrix.h"
#include "cq_common.h"
#include "cq_temps.h"

namespace {

constexpr double kRdErr = 1.0e-8;     // RD_ERR (src/numbers_module.f90:96)
constexpr double kSpinFactor = 2.0;   // spin-unpolarised
constexpr double kTwoPi = 6.28318530717958647692;

#define DTRY(x)                \
  do {                         \
    const int st_ = (x);       \
    if (st_) return st_;       \
  } while (0)

using cq::Temps;

int trace2(cqmm_ctx *ctx, int32_t a, int32_t b, double *out) { return cqmm_matrix_product_trace(ctx, a, b, out); }

// FP64 products for a scope (the electron-number constraint and the final K,
// energy and electron count), whatever the context's product precision; the
// precision is restored when the scope ends.
class Fp64Products {
 public:
  explicit Fp64Products(cqmm_ctx *ctx) : ctx_(ctx) {
    if (cqmm_ctx_get_product_precision(ctx_, &saved_) != CQMM_OK) saved_ = 0;
    if (saved_ != 0) (void)cqmm_ctx_set_product_precision(ctx_, 0);
  }
  ~Fp64Products() {
    if (saved_ != 0) (void)cqmm_ctx_set_product_precision(ctx_, saved_);
  }
  Fp64Products(const Fp64Products &) = delete;
  Fp64Products &operator=(const Fp64Products &) = delete;

 private:
  cqmm_ctx *ctx_;
  int32_t saved_ = 0;
};

// X -> T X T^T at L range (the S^-1 metric of the gradients).
int precondition(cqmm_ctx *ctx, const cqmm_dmm_setup *su, int32_t T, int32_t Tt, int32_t X, int32_t tmpTL,
                 int32_t out) {
  DTRY(cqmm_matrix_product(ctx, T, X, tmpTL, su->mult_T_L_TL));
  return cqmm_matrix_product(ctx, tmpTL, Tt, out, su->mult_TL_T_L);
}

// SolveCubic (src/DMMinModule.f90:2072-2143): the real root of
// z^3 + a z^2 + b z + c = 0 of smallest magnitude (three real roots), or the
// single real root, replaced by `guess` when it exceeds ten times the guess.
double solve_cubic(double a, double b, double c, double guess) {
  double Q = (a * a - 3.0 * b) / 9.0;
  double R = (2.0 * a * a * a - 9.0 * a * b + 27.0 * c) / 54.0;
  if (R * R < Q * Q * Q) {
    const double theta = std::acos(R / std::sqrt(Q * Q * Q));
    const double T = -2.0 * std::sqrt(Q);
    const double z1 = T * std::cos(theta / 3.0) - a / 3.0;
    const double z2 = T * std::cos((theta + kTwoPi) / 3.0) - a / 3.0;
    const double z3 = T * std::cos((theta - kTwoPi) / 3.0) - a / 3.0;
    if (std::fabs(z1) <= std::fabs(z2) && std::fabs(z1) <= std::fabs(z3)) return z1;
    if (std::fabs(z2) <= std::fabs(z3)) return z2;
    return z3;
  }
  Q = -Q;
  R = -R;
  double S = R + std::sqrt(Q * Q * Q + R * R);
  double T = R - std::sqrt(Q * Q * Q + R * R);
  S = S >= 0.0 ? std::cbrt(S) : -std::cbrt(std::fabs(S));
  T = T >= 0.0 ? std::cbrt(T) : -std::cbrt(std::fabs(T));
  double z = S + T - a / 3.0;
  if (std::fabs(z) > 10.0 * std::fabs(guess)) z = guess;
  return z;
}

// DoPulay2D_nospin (src/Pulay.f90:188-262): alpha_i = sum_j (A^-1)_ji /
// sum_ij (A^-1)_ji; when A is singular, alpha = e_{iPulay}. The inverse is a
// Gauss-Jordan elimination with partial pivoting (the reference uses LAPACK
// sytrf/sytri on the symmetric A).
void do_pulay(int ipulay, int n, const std::vector<double> &A, std::vector<double> &alpha) {
  alpha.assign(n, 0.0);
  std::vector<double> M(A), inv(n * n, 0.0);
  for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
  bool ok = true;
  for (int c = 0; c < n && ok; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(M[r * n + c]) > std::fabs(M[piv * n + c])) piv = r;
    if (M[piv * n + c] == 0.0 || !std::isfinite(M[piv * n + c])) {
      ok = false;
      break;
    }
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(M[piv * n + k], M[c * n + k]);
        std::swap(inv[piv * n + k], inv[c * n + k]);
      }
    const double d = M[c * n + c];
    for (int k = 0; k < n; ++k) {
      M[c * n + k] /= d;
      inv[c * n + k] /= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = M[r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        M[r * n + k] -= f * M[c * n + k];
        inv[r * n + k] -= f * inv[c * n + k];
      }
    }
  }
  if (!ok) {
    alpha[ipulay - 1] = 1.0;
    return;
  }
  double sum = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      alpha[i] += inv[j * n + i];
      sum += inv[j * n + i];
    }
  for (int i = 0; i < n; ++i) alpha[i] /= sum;
}

// The residual g = spin_factor * Tr[M3 . T M3 T] (per atom unless global).
double residual(const cqmm_dmm_setup *su, double tr) {
  return su->global_tolerance ? kSpinFactor * tr : kSpinFactor * tr / (double)su->n_atoms;
}

struct Dmm {
  cqmm_ctx *ctx;
  const cqmm_dmm_setup *su;
  int32_t H, S, T, Tt, L, K, phi;
  double delta_e = 0.0;

  int lnv(uint32_t flags, int32_t M3, int32_t ph, double *electrons, double *energy) {
    return cqmm_lnv(ctx, &su->lnv, flags, H, S, L, K, M3, ph, electrons, energy);
  }

  // correct_electron_number_fixspin, with FP64 products.
  int correct() {
    Fp64Products fp64(ctx);
    Temps t(ctx);
    int32_t TL, phi2, Sphi, Sphi2;
    DTRY(t.alloc(su->range_TL, &TL));
    DTRY(t.alloc(su->lnv.range_L, &phi2));
    DTRY(t.alloc(su->lnv.range_L, &Sphi));
    DTRY(t.alloc(su->lnv.range_L, &Sphi2));
    const double e0 = su->electrons_target;
    bool done = false;
    int iter = 0;
    double el = 0.0, el2 = 0.0, g0 = 0.0, g1 = 0.0;
    while (!done && iter < 20) {
      DTRY(lnv(CQMM_LNV_PHI, -1, phi, &el, nullptr));
      DTRY(precondition(ctx, su, T, Tt, phi, TL, Sphi));
      DTRY(trace2(ctx, Sphi, phi, &g0));
      const double step1 = (e0 - el) / g0;
      if (std::fabs(e0 - el) < 1.0e-9) {
        DTRY(cqmm_matrix_sum(ctx, 1.0, L, step1, Sphi));
        done = true;
      } else {
        const double step = step1;
        DTRY(cqmm_matrix_sum(ctx, 1.0, L, step1, Sphi));
        DTRY(lnv(CQMM_LNV_PHI, -1, phi2, &el2, nullptr));
        DTRY(precondition(ctx, su, T, Tt, phi2, TL, Sphi2));
        DTRY(trace2(ctx, phi, Sphi2, &g1));
        double D = el - e0, C = g0;
        double B = 3.0 * (el2 - el) / (step * step) - (g1 + 2.0 * g0) / step;
        const double recA = (step * step * step) / (2.0 * (el - el2) + (g0 + g1) * step);
        B *= recA;
        C *= recA;
        D *= recA;
        const double truestep = solve_cubic(B, C, D, step);
        DTRY(cqmm_matrix_sum(ctx, 1.0, L, truestep - step, Sphi));
        DTRY(lnv(CQMM_LNV_PHI, -1, phi, &el2, nullptr));
        if (truestep == step) {
          if (std::fabs(el2 - e0) < 1.0e-4) done = true;
          ++iter;
        } else {
          done = true;
        }
      }
    }
    return t.release();
  }

  // Projection of the search direction onto constant electron number:
  // search += (e.n / n.n) T phi T (earlyDM), or M3 -= (e.n/n.n) phi and
  // SM3 -= (e.n/n.n) T phi T (lateDM).
  int project(int32_t SM3, int32_t M3, int32_t Sphi, int32_t TL, int32_t search) {
    double en = 0.0, nn = 0.0;
    DTRY(precondition(ctx, su, T, T, phi, TL, Sphi));
    DTRY(trace2(ctx, SM3, phi, &en));
    DTRY(trace2(ctx, Sphi, phi, &nn));
    if (std::fabs(nn) > kRdErr) {
      if (search >= 0) {
        DTRY(cqmm_matrix_sum(ctx, 1.0, search, en / nn, Sphi));
      } else {
        DTRY(cqmm_matrix_sum(ctx, 1.0, SM3, -en / nn, Sphi));
        DTRY(cqmm_matrix_sum(ctx, 1.0, M3, -en / nn, phi));
      }
    }
    return CQMM_OK;
  }

  // lineMinL: cubic fit of E(L + s D) from E, dE/ds at s = 0 and s = step.
  int line_min(int32_t M3, int32_t D, int32_t TL, int32_t SM3, double e_in, double *e_out, bool *inflex,
               double *interpG) {
    Temps t(ctx);
    int32_t M3old;
    DTRY(t.alloc(su->lnv.range_L, &M3old));
    DTRY(cqmm_matrix_sum(ctx, 0.0, M3old, 1.0, M3));
    double tr = 0.0;
    DTRY(trace2(ctx, D, M3, &tr));
    const double g0 = kSpinFactor * tr, ig0 = g0;
    const double energy0 = e_in;
    double step = delta_e / g0;
    if (std::fabs(step) < 1.0e-2) step = 0.01;
    if (std::fabs(step) > 0.1) step = 0.1;
    DTRY(cqmm_matrix_sum(ctx, 1.0, L, step, D));
    double el = 0.0, en = 0.0;
    DTRY(lnv(CQMM_LNV_K | CQMM_LNV_E | CQMM_LNV_M3, M3, -1, &el, &en));
    const double energy1 = kSpinFactor * en;
    DTRY(trace2(ctx, D, M3, &tr));
    const double g1 = kSpinFactor * tr;
    DTRY(trace2(ctx, SM3, M3, &tr));
    double igcross = kSpinFactor * tr;
    DTRY(precondition(ctx, su, T, T, M3, TL, SM3));
    DTRY(trace2(ctx, SM3, M3, &tr));
    double ig1 = kSpinFactor * tr;
    const double ig1_step = ig1;
    const double Dc = energy0, Cc = g0;
    const double Bc = 3.0 * (energy1 - Dc) / (step * step) - (g1 - 2.0 * Cc) / step;
    const double Ac = (g1 - 2.0 * Bc * step - Cc) / (3.0 * step * step);
    const double SQ = 4.0 * Bc * Bc - 12.0 * Ac * Cc;
    if (SQ < 0.0) {
      *inflex = true;
      return t.release();
    }
    double truestep = std::fabs(Ac) > kRdErr ? (-2.0 * Bc + std::sqrt(SQ)) / (6.0 * Ac) : -Cc / (2.0 * Bc);
    if (truestep < 1.0e-4) truestep = 1.0e-4;
    DTRY(cqmm_matrix_sum(ctx, 1.0, L, truestep - step, D));
    DTRY(cqmm_matrix_symmetrise(ctx, L));
    DTRY(lnv(CQMM_LNV_K | CQMM_LNV_E | CQMM_LNV_M3, M3, -1, &el, &en));
    *e_out = kSpinFactor * en;
    delta_e = *e_out - e_in;
    double zeta;
    if (truestep < step) {
      const double lamG = truestep / step;
      const double iG = ig0 * (1.0 - lamG) * (1.0 - lamG) + ig1 * lamG * lamG + 2.0 * igcross * lamG * (1.0 - lamG);
      zeta = (iG - ig1) / (ig0 - ig1);
    } else {
      DTRY(precondition(ctx, su, T, T, M3, TL, SM3));
      DTRY(trace2(ctx, SM3, M3, &tr));
      ig1 = kSpinFactor * tr;
      DTRY(trace2(ctx, SM3, M3old, &tr));
      igcross = kSpinFactor * tr;
      const double lamG = step / truestep;
      const double iG = ig0 * (1.0 - lamG) * (1.0 - lamG) + ig1 * lamG * lamG + 2.0 * igcross * lamG * (1.0 - lamG);
      zeta = (iG - ig1_step) / (ig0 - ig1_step);
    }
    *interpG = zeta;
    return t.release();
  }

  // earlyDM; returns with *ndone iterations, *inflex set on an inflexion.
  int early(int *ndone, bool *inflex, double *energy_out, double *g_out) {
    Temps t(ctx);
    int32_t M3, SM3, Sphi, TL, search;
    DTRY(t.alloc(su->lnv.range_L, &M3));
    DTRY(t.alloc(su->lnv.range_L, &SM3));
    DTRY(t.alloc(su->lnv.range_L, &Sphi));
    DTRY(t.alloc(su->range_TL, &TL));
    DTRY(t.alloc(su->lnv.range_L, &search));
    const uint32_t all = CQMM_LNV_K | CQMM_LNV_E | CQMM_LNV_M3 | CQMM_LNV_PHI;
    if (su->vary_mu) DTRY(correct());
    double el = 0.0, en = 0.0, tr = 0.0;
    DTRY(lnv(all, M3, phi, &el, &en));
    double energy0 = kSpinFactor * en, energy1 = energy0;
    DTRY(precondition(ctx, su, T, T, M3, TL, SM3));
    DTRY(cqmm_matrix_sum(ctx, 0.0, search, -1.0, SM3));
    if (su->vary_mu) DTRY(project(SM3, M3, Sphi, TL, search));
    delta_e = 0.0;
    double g0 = 0.0;
    *ndone = su->n_L_iterations;
    for (int n = 1; n <= su->n_L_iterations; ++n) {
      DTRY(trace2(ctx, M3, SM3, &tr));
      g0 = residual(su, tr);
      double interpG = 0.0;
      DTRY(line_min(M3, search, TL, SM3, energy0, &energy1, inflex, &interpG));
      if (*inflex) {
        *ndone = n;
        *energy_out = energy0;
        *g_out = g0;
        return t.release();
      }
      const double zeta = interpG;
      if (su->vary_mu) DTRY(correct());
      DTRY(lnv(all, M3, phi, &el, &en));
      energy1 = kSpinFactor * en;
      DTRY(precondition(ctx, su, T, T, M3, TL, SM3));
      energy0 = energy1;
      DTRY(cqmm_matrix_sum(ctx, 0.0, search, -1.0, SM3));
      if (su->vary_mu) DTRY(project(SM3, M3, Sphi, TL, search));
      DTRY(trace2(ctx, M3, SM3, &tr));
      g0 = residual(su, tr);
      if (std::fabs(zeta) < su->lin_tol) {
        *ndone = n;
        break;
      }
    }
    *energy_out = energy0;
    *g_out = g0;
    return t.release();
  }

  // lateDM (Pulay); *done when the residual falls below the tolerance.
  int late(int *ndone, bool *done, double *energy_out, double *g_out) {
    Temps t(ctx);
    const int mp = su->maxpulay;
    std::vector<int32_t> Ls(mp), Gs(mp), SGs(mp);
    for (int i = 0; i < mp; ++i) {
      DTRY(t.alloc(su->lnv.range_L, &Ls[i]));
      DTRY(t.alloc(su->lnv.range_L, &Gs[i]));
      DTRY(t.alloc(su->lnv.range_L, &SGs[i]));
    }
    int32_t M3, SM3, Sphi, TL;
    DTRY(t.alloc(su->lnv.range_L, &M3));
    DTRY(t.alloc(su->lnv.range_L, &SM3));
    DTRY(t.alloc(su->lnv.range_L, &Sphi));
    DTRY(t.alloc(su->range_TL, &TL));
    const uint32_t all = CQMM_LNV_K | CQMM_LNV_E | CQMM_LNV_M3 | CQMM_LNV_PHI;
    double el = 0.0, en = 0.0, tr = 0.0;
    DTRY(lnv(all, M3, phi, &el, &en));
    double energy0 = kSpinFactor * en, energy1 = energy0;
    DTRY(precondition(ctx, su, T, T, M3, TL, SM3));
    if (su->vary_mu) DTRY(project(SM3, M3, Sphi, TL, -1));
    DTRY(cqmm_matrix_sum(ctx, 0.0, SGs[0], -1.0, SM3));
    DTRY(cqmm_matrix_sum(ctx, 0.0, Gs[0], -1.0, M3));
    DTRY(cqmm_matrix_sum(ctx, 0.0, Ls[0], 1.0, L));
    DTRY(trace2(ctx, M3, SM3, &tr));
    double g0 = residual(su, tr), g1 = g0;
    *done = su->n_L_iterations == 0;
    *ndone = 0;
    int stuck = 0;
    std::vector<double> Aij, alph;
    for (int n = 1; n <= su->n_L_iterations; ++n) {
      *ndone = n;
      const int npmod = n % mp + 1;             // 1-based slot of this iteration
      const int pul_mx = std::min(n + 1, mp);
      double step = std::fabs(su->global_tolerance ? delta_e / (g0 + kRdErr)
                                                   : delta_e / ((double)su->n_atoms * g0 + kRdErr));
      if (step < su->min_pulay_step) step = su->min_pulay_step;
      if (step > su->max_pulay_step) step = su->max_pulay_step;
      const int prev = npmod > 1 ? npmod - 1 : pul_mx;
      DTRY(cqmm_matrix_sum(ctx, 0.0, L, 1.0, Ls[prev - 1]));
      DTRY(cqmm_matrix_sum(ctx, 1.0, L, step, SGs[prev - 1]));
      if (su->vary_mu) DTRY(correct());
      DTRY(lnv(all, M3, phi, &el, &en));
      energy1 = kSpinFactor * en;
      DTRY(precondition(ctx, su, T, T, M3, TL, SM3));
      if (su->vary_mu) DTRY(project(SM3, M3, Sphi, TL, -1));
      DTRY(cqmm_matrix_sum(ctx, 0.0, SGs[npmod - 1], -1.0, SM3));
      DTRY(cqmm_matrix_sum(ctx, 0.0, Gs[npmod - 1], -1.0, M3));
      DTRY(cqmm_matrix_sum(ctx, 0.0, Ls[npmod - 1], 1.0, L));
      Aij.assign(pul_mx * pul_mx, 0.0);
      for (int i = 0; i < pul_mx; ++i)
        for (int j = 0; j < pul_mx; ++j) {
          DTRY(trace2(ctx, Gs[j], SGs[i], &tr));
          Aij[j * pul_mx + i] = kSpinFactor * tr;  // Aij(j,i)
        }
      do_pulay(npmod, pul_mx, Aij, alph);
      DTRY(cqmm_matrix_scale(ctx, 0.0, L));
      for (int i = 0; i < pul_mx; ++i) DTRY(cqmm_matrix_sum(ctx, 1.0, L, alph[i], Ls[i]));
      if (su->vary_mu) DTRY(correct());
      DTRY(lnv(all, M3, phi, &el, &en));
      energy1 = kSpinFactor * en;
      DTRY(precondition(ctx, su, T, T, M3, TL, SM3));
      if (su->vary_mu) DTRY(project(SM3, M3, Sphi, TL, -1));
      delta_e = energy1 - energy0;
      DTRY(trace2(ctx, M3, SM3, &tr));
      g1 = residual(su, tr);  // of the projected gradient
      if (g1 < su->tolerance) {
        *done = true;
        break;
      } else if (g1 > 2.0 * g0) {
        break;
      } else if (g1 > 0.99 * g0) {
        if (++stuck > 5) break;
      }
      DTRY(cqmm_matrix_sum(ctx, 0.0, SGs[npmod - 1], -1.0, SM3));
      DTRY(cqmm_matrix_sum(ctx, 0.0, Gs[npmod - 1], -1.0, M3));
      DTRY(cqmm_matrix_sum(ctx, 0.0, Ls[npmod - 1], 1.0, L));
      g0 = g1;
      energy0 = energy1;
    }
    *energy_out = energy1;
    *g_out = g1;
    return t.release();
  }
};

}  // namespace

extern "C" {

int cqmm_lnv(cqmm_ctx *ctx, const cqmm_lnv_setup *su, uint32_t flags, int32_t H, int32_t S, int32_t L, int32_t K,
             int32_t M3, int32_t phi, double *electrons, double *energy) {
  return cq::guarded(__func__, [&]() -> int {
    if (!ctx || !su) return cq::set_error(CQMM_ERR_ARG, "cqmm_lnv: null argument");
    const bool doE = flags & CQMM_LNV_E, doK = (flags & CQMM_LNV_K) || doE, doM3 = flags & CQMM_LNV_M3,
               dophi = flags & CQMM_LNV_PHI;
    if (doK && K < 0) return cq::set_error(CQMM_ERR_ARG, "cqmm_lnv: K or energy requested without a K handle");
    if (doE && !energy) return cq::set_error(CQMM_ERR_ARG, "cqmm_lnv: energy requested without an output");
    if (doM3 && M3 < 0) return cq::set_error(CQMM_ERR_ARG, "cqmm_lnv: M3 requested without a handle");
    if (dophi && (phi < 0 || !electrons)) return cq::set_error(CQMM_ERR_ARG, "cqmm_lnv: phi requested without outputs");
    Temps t(ctx);
    int32_t LS, SL;
    DTRY(t.alloc(su->range_LS, &LS));
    DTRY(t.alloc(su->range_LS, &SL));
    DTRY(cqmm_matrix_product(ctx, L, S, LS, su->mult_L_S_LS));
    DTRY(cqmm_matrix_transpose(ctx, LS, SL));
    int32_t LH = -1, HL = -1, HLS = -1, SLH = -1, A = -1, B = -1, C = -1;
    if (doM3) {
      DTRY(t.alloc(su->range_LH, &LH));
      DTRY(t.alloc(su->range_LH, &HL));
      DTRY(cqmm_matrix_product(ctx, L, H, LH, su->mult_L_H_LH));
      DTRY(cqmm_matrix_transpose(ctx, LH, HL));
      DTRY(t.alloc(su->range_LSL, &HLS));
      DTRY(t.alloc(su->range_LSL, &SLH));
      DTRY(t.alloc(su->range_LSL, &A));
      DTRY(t.alloc(su->range_L, &B));
      DTRY(t.alloc(su->range_L, &C));
      DTRY(cqmm_matrix_product(ctx, HL, S, HLS, su->mult_HL_S_LSL));
      DTRY(cqmm_matrix_transpose(ctx, HLS, SLH));
      DTRY(cqmm_matrix_sum(ctx, 0.0, A, 2.0, HLS));
      DTRY(cqmm_matrix_sum(ctx, 1.0, A, 1.0, SLH));
      DTRY(cqmm_matrix_product(ctx, A, SL, B, su->mult_HLS_LS_L));
      DTRY(cqmm_matrix_transpose(ctx, B, C));
    }
    int32_t SLS = -1, SLSLS = -1;
    if (dophi) {
      DTRY(t.alloc(su->range_SLS, &SLS));
      DTRY(t.alloc(su->range_L, &SLSLS));
      DTRY(cqmm_matrix_product(ctx, SL, S, SLS, su->mult_SL_S_SLS));
      DTRY(cqmm_matrix_product(ctx, SLS, SL, SLSLS, su->mult_SLS_LS_L));
    }
    if (doK) {
      int32_t LSL, LSLSL;
      DTRY(t.alloc(su->range_LSL, &LSL));
      DTRY(t.alloc(su->range_H, &LSLSL));
      DTRY(cqmm_matrix_product(ctx, LS, L, LSL, su->mult_LS_L_LSL));
      DTRY(cqmm_matrix_product(ctx, LSL, LS, LSLSL, su->mult_LSL_SL_H));
      DTRY(cqmm_matrix_sum(ctx, 0.0, K, 3.0, LSL));
      DTRY(cqmm_matrix_sum(ctx, 1.0, K, -2.0, LSLSL));
    }
    if (doE) DTRY(cqmm_matrix_product_trace(ctx, H, K, energy));
    if (doM3) {
      DTRY(cqmm_matrix_sum(ctx, 0.0, M3, 3.0, HLS));
      DTRY(cqmm_matrix_sum(ctx, 1.0, M3, 3.0, SLH));
      DTRY(cqmm_matrix_sum(ctx, 1.0, M3, -1.0, B));
      DTRY(cqmm_matrix_sum(ctx, 1.0, M3, -1.0, C));
    }
    if (dophi) {
      double e1 = 0.0, e2 = 0.0;
      DTRY(cqmm_matrix_sum(ctx, 0.0, phi, 1.0, SLS));
      DTRY(cqmm_matrix_product_trace(ctx, L, phi, &e1));
      DTRY(cqmm_matrix_product_trace(ctx, L, SLSLS, &e2));
      *electrons = 3.0 * e1 - 2.0 * e2;
      DTRY(cqmm_matrix_sum(ctx, 6.0, phi, -6.0, SLSLS));
    }
    return t.release();
  });
}

int cqmm_correct_electrons(cqmm_ctx *ctx, const cqmm_dmm_setup *su, int32_t H, int32_t S, int32_t T, int32_t L,
                           double *electrons) {
  return cq::guarded(__func__, [&]() -> int {
    if (!ctx || !su) return cq::set_error(CQMM_ERR_ARG, "cqmm_correct_electrons: null argument");
    Temps t(ctx);
    int32_t K, phi, Tt;
    DTRY(t.alloc(su->lnv.range_H, &K));
    DTRY(t.alloc(su->lnv.range_L, &phi));
    DTRY(t.alloc(su->range_T, &Tt));
    DTRY(cqmm_matrix_transpose(ctx, T, Tt));
    Dmm d{ctx, su, H, S, T, Tt, L, K, phi};
    DTRY(d.correct());
    if (electrons) DTRY(cqmm_lnv(ctx, &su->lnv, CQMM_LNV_PHI, H, S, L, -1, -1, phi, electrons, nullptr));
    return t.release();
  });
}

int cqmm_dmm(cqmm_ctx *ctx, const cqmm_dmm_setup *su, int32_t H, int32_t S, int32_t T, int32_t L, int32_t K,
             cqmm_dmm_result *res) {
  return cq::guarded(__func__, [&]() -> int {
    if (!ctx || !su || !res) return cq::set_error(CQMM_ERR_ARG, "cqmm_dmm: null argument");
    if (su->maxpulay < 1 || su->maxpulay > 16) return cq::set_error(CQMM_ERR_ARG, "cqmm_dmm: maxpulay outside 1..16");
    if (su->n_atoms <= 0 || su->n_L_iterations < 0) return cq::set_error(CQMM_ERR_ARG, "cqmm_dmm: bad sizes");
    std::memset(res, 0, sizeof(*res));
    const auto t0 = std::chrono::steady_clock::now();
    Temps t(ctx);
    int32_t phi, Tt;
    DTRY(t.alloc(su->lnv.range_L, &phi));
    DTRY(t.alloc(su->range_T, &Tt));
    DTRY(cqmm_matrix_transpose(ctx, T, Tt));
    Dmm d{ctx, su, H, S, T, Tt, L, K, phi};
    d.delta_e = su->delta_e;
    // FindMinDM: early stage (steepest descent with line minimisation) until the
    // line search interpolation says the quadratic regime is reached, then Pulay.
    // An inflexion restarts the early stage (the reference first resets L with
    // InitMcW/McWMin, which callers of this entry point do before calling it).
    bool done = false;
    int rounds = 0;
    double energy = 0.0, g = 0.0;
    while (!done && rounds < 10) {
      ++rounds;
      bool inflex = false;
      int ndone = 0;
      DTRY(d.early(&ndone, &inflex, &energy, &g));
      res->early_iterations += ndone;
      if (inflex) {
        ++res->inflexions;
        continue;
      }
      DTRY(d.late(&ndone, &done, &energy, &g));
      res->late_iterations += ndone;
      if (!done && ndone == su->n_L_iterations) break;
    }
    double el = 0.0, en = 0.0;
    {
      Fp64Products fp64(ctx);  // K, 2Tr[KH] and the electron count in FP64
      DTRY(cqmm_lnv(ctx, &su->lnv, CQMM_LNV_K | CQMM_LNV_E | CQMM_LNV_PHI, H, S, L, K, -1, phi, &el, &en));
    }
    res->done = done ? 1 : 0;
    res->energy = kSpinFactor * en;
    res->electrons = kSpinFactor * el;
    res->residual = g;
    DTRY(t.release());
    DTRY(cqmm_synchronize());
    res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return CQMM_OK;
  });
}

}  // extern "C"
