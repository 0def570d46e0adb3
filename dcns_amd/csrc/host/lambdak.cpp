// lambdak.cpp -- FP64 least-squares fit of the Karlin-Altschul lambda and K
// to maximum local-alignment scores of random fragment pairs.
//
// Reference: the ranSco scores (permutationLsqLambdaK.cpp:32-102) are fitted
// by scripts/scriptToestimateKandLambda.r:4-15 with minpack.lm::nlsLM:
//   x = sorted unique scores, y(x) = #(S >= x) / n,
//   y ~ 1 - exp(-K * N * exp(-lambda * x)),  N = 1e6, start K = 0.3, lambda = 0.2,
//   Levenberg-Marquardt, maxiter 550, tolerances sqrt(machine eps).
// Here: the same model and data, fitted by a Levenberg-Marquardt iteration of
// this library's own (analytic Jacobian, Marquardt diagonal scaling).  For
// pairs of different lengths the search space N varies per pair; the model
// then averages the survival over the pairs' N = m*n, grouped into fine
// log-spaced bins (each bin contributes with its mean N and its weight).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../../include/dcns.h"
#include "../ctx.h"

namespace {

struct Model {
    std::vector<double> x, y;        // data points
    std::vector<double> binN, binW;  // search-space mixture (weights sum to 1)
    // f(x) = 1 - sum_b w_b exp(-K N_b e^{-l x}); writes f and the two partials
    void eval(double K, double l, size_t i, double &f, double &dK, double &dl) const {
        const double xi = x[i], e = std::exp(-l * xi);
        double s = 0, sK = 0, sl = 0;
        for (size_t b = 0; b < binN.size(); ++b) {
            const double u = K * binN[b] * e;
            const double t = binW[b] * std::exp(-u);
            s += t;
            sK += t * binN[b] * e;  // d/dK of (1 - t) = t * N e^{-lx}
            sl += t * u * (-xi);     // d/dl of (1 - t) = -x u t
        }
        f = 1.0 - s;
        dK = sK;
        dl = sl;
    }
    double rss(double K, double l) const {
        double r = 0, f, a, b;
        for (size_t i = 0; i < x.size(); ++i) {
            eval(K, l, i, f, a, b);
            r += (y[i] - f) * (y[i] - f);
        }
        return r;
    }
};

// Levenberg-Marquardt with Marquardt's diagonal scaling; stops when the
// relative RSS decrease or the relative step falls below sqrt(eps).
void lm(const Model &m, double K, double l, int32_t maxiter, double *out) {
    const double tol = std::sqrt(2.220446049250313e-16);
    double cur = m.rss(K, l), mu = 1e-3;
    int32_t it = 0;
    for (; it < maxiter; ++it) {
        double A00 = 0, A01 = 0, A11 = 0, g0 = 0, g1 = 0;
        for (size_t i = 0; i < m.x.size(); ++i) {
            double f, dK, dl;
            m.eval(K, l, i, f, dK, dl);
            const double r = m.y[i] - f;
            A00 += dK * dK;
            A01 += dK * dl;
            A11 += dl * dl;
            g0 += dK * r;
            g1 += dl * r;
        }
        bool stepped = false, converged = false;
        for (int tries = 0; tries < 60; ++tries) {
            const double a = A00 * (1 + mu), b = A01, d = A11 * (1 + mu);
            const double det = a * d - b * b;
            if (!(det != 0) || !std::isfinite(det)) {
                mu *= 10;
                continue;
            }
            const double dK = (d * g0 - b * g1) / det, dl = (a * g1 - b * g0) / det;
            const double Kn = K + dK, ln = l + dl;
            const double nr = m.rss(Kn, ln);
            if (std::isfinite(nr) && nr <= cur) {
                const double rel_f = cur > 0 ? (cur - nr) / cur : 0;
                const double rel_p = std::sqrt(dK * dK + dl * dl) / (std::sqrt(K * K + l * l) + tol);
                K = Kn;
                l = ln;
                converged = rel_f <= tol || rel_p <= tol;
                cur = nr;
                mu = std::max(mu / 10, 1e-12);
                stepped = true;
                break;
            }
            mu *= 10;
        }
        if (!stepped || converged) {
            ++it;
            break;
        }
    }
    out[0] = K;
    out[1] = l;
    out[2] = cur;
    out[3] = (double)it;
}

}  // namespace

extern "C" int dcns_fit_lambda_k(const int32_t *scores, int64_t n, const double *log_space, double fixed_space,
                                 double k0, double lambda0, int32_t maxiter, double *out4) {
    if (!scores || n <= 0 || !out4 || maxiter < 1) return DCNS_ERR_ARG;
    if (!log_space && !(fixed_space > 0)) return DCNS_ERR_ARG;
    Model m;
    // empirical survival over the sorted unique scores
    std::vector<int32_t> s(scores, scores + n);
    std::sort(s.begin(), s.end());
    for (int64_t i = 0; i < n;) {
        int64_t j = i;
        while (j < n && s[(size_t)j] == s[(size_t)i]) ++j;
        m.x.push_back((double)s[(size_t)i]);
        m.y.push_back((double)(n - i) / (double)n);
        i = j;
    }
    if (!log_space) {
        m.binN.push_back(fixed_space);
        m.binW.push_back(1.0);
    } else {
        double lo = log_space[0], hi = log_space[0];
        for (int64_t i = 1; i < n; ++i) {
            lo = std::min(lo, log_space[i]);
            hi = std::max(hi, log_space[i]);
        }
        const int nb = 1024;
        std::vector<double> sumN(nb, 0.0), cnt(nb, 0.0);
        const double width = (hi - lo) / nb;
        for (int64_t i = 0; i < n; ++i) {
            int b = width > 0 ? (int)((log_space[i] - lo) / width) : 0;
            b = std::min(std::max(b, 0), nb - 1);
            sumN[(size_t)b] += std::exp(log_space[i]);
            cnt[(size_t)b] += 1.0;
        }
        for (int b = 0; b < nb; ++b)
            if (cnt[(size_t)b] > 0) {
                m.binN.push_back(sumN[(size_t)b] / cnt[(size_t)b]);
                m.binW.push_back(cnt[(size_t)b] / (double)n);
            }
    }
    // Levenberg-Marquardt from the reference's start (K = k0, lambda = lambda0)
    // and from the linearised Gumbel estimate ln(-ln(1-y)) = ln(K N) - lambda x;
    // the start the reference uses diverges on some data, so the fit with the
    // lower residual sum of squares is returned.
    double Nbar = 0;
    for (size_t b = 0; b < m.binN.size(); ++b) Nbar += m.binW[b] * std::log(m.binN[b]);
    Nbar = std::exp(Nbar);
    double sx = 0, sy = 0, sxx = 0, sxy = 0, cntl = 0;
    for (size_t i = 0; i < m.x.size(); ++i) {
        if (!(m.y[i] > 0 && m.y[i] < 1)) continue;
        const double t = std::log(-std::log(1 - m.y[i]));
        sx += m.x[i];
        sy += t;
        sxx += m.x[i] * m.x[i];
        sxy += m.x[i] * t;
        cntl += 1;
    }
    double K1 = k0, l1 = lambda0;
    if (cntl >= 2 && cntl * sxx - sx * sx > 0) {
        const double slope = (cntl * sxy - sx * sy) / (cntl * sxx - sx * sx);
        const double icpt = (sy - slope * sx) / cntl;
        if (slope < 0) {
            l1 = -slope;
            K1 = std::exp(icpt) / Nbar;
        }
    }
    double best[4] = {0, 0, INFINITY, 0};
    for (int start = 0; start < 2; ++start) {
        double r[4];
        lm(m, start == 0 ? k0 : K1, start == 0 ? lambda0 : l1, maxiter, r);
        if (std::isfinite(r[2]) && r[1] > 0 && r[0] > 0 && r[2] < best[2]) std::copy(r, r + 4, best);
    }
    if (!std::isfinite(best[2])) lm(m, k0, lambda0, maxiter, best);
    const double K = best[0], l = best[1], cur = best[2];
    const int32_t it = (int32_t)best[3];
    out4[0] = K;
    out4[1] = l;
    out4[2] = cur;
    out4[3] = (double)it;
    return DCNS_OK;
}

// Many independent fits with a fixed search space each: group g holds
// scores[g * group_size .. +group_size) and N = space[g]; out4[4g..4g+3] as
// dcns_fit_lambda_k.  Groups run on the host's cores.
extern "C" int dcns_fit_lambda_k_batch(const int32_t *scores, int64_t ngroups, int32_t group_size, const double *space,
                                       double k0, double lambda0, int32_t maxiter, double *out4) {
    if (!scores || ngroups < 0 || group_size < 1 || !space || !out4 || maxiter < 1) return DCNS_ERR_ARG;
    for (int64_t g = 0; g < ngroups; ++g)
        if (!(space[g] > 0)) return DCNS_ERR_ARG;
    dcns::parallel_for(ngroups, [&](int64_t g) {
        dcns_fit_lambda_k(scores + g * group_size, group_size, nullptr, space[g], k0, lambda0, maxiter, out4 + 4 * g);
    });
    return DCNS_OK;
}

// calculateLambda (impl/calculateLambda.cpp:64-82): the ungapped
// Karlin-Altschul lambda, the positive root of
//   sum_ij p_i p_j exp(lambda s_ij) = 1,  s_ij = match (i == j) or mismatch,
// over the A/C/G/T frequencies of seq1 (the reference uses seq1's
// frequencies for both factors), by Newton's method with the reference's
// forward-difference derivative of step epsilon, starting from 1 and doubling
// until the derivative is positive, stopping when a step is at most
// precision.  Differences from the reference, whose behaviour is undefined:
// the frequency arrays start at zero (uninitialised there, :66-69) and the
// step size is compared as a double (the reference's unqualified abs truncates
// it to an integer and stops after one step); at most 10,000 steps.
namespace {
double lambda_f(const double *f, int32_t match, int32_t mismatch, double lam) {
    double v = 0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) v += f[i] * f[j] * std::exp((double)(i == j ? match : mismatch) * lam);
    return v - 1.0;
}
}  // namespace

extern "C" double dcns_calculate_lambda(const int8_t *seq1, const int8_t *seq2, uint32_t len1, uint32_t len2,
                                        int32_t match, int32_t mismatch, double epsilon, double precision) {
    (void)seq2;
    (void)len2;
    if (!seq1 || !(epsilon > 0) || !(precision > 0)) return NAN;
    double f[4] = {0, 0, 0, 0}, n = 0;
    for (uint32_t i = 0; i < len1; ++i) {
        const int c = seq1[i];
        const int k = c == 0 ? 0 : c == 1 ? 1 : c == 3 ? 2 : c == 4 ? 3 : -1;
        if (k >= 0) {
            f[k] += 1;
            n += 1;
        }
    }
    if (n == 0) return NAN;
    for (double &v : f) v /= n;
    auto fp = [&](double lam) { return (lambda_f(f, match, mismatch, lam + epsilon) - lambda_f(f, match, mismatch, lam)) / epsilon; };
    double lam = 1.0;
    for (int k = 0; k < 1100 && fp(lam) < 0; ++k) lam *= 2.0;
    for (int it = 0; it < 10000; ++it) {
        const double next = lam - lambda_f(f, match, mismatch, lam) / fp(lam);
        const double d = std::fabs(next - lam);
        lam = next;
        if (!(d > precision)) break;
    }
    return lam;
}
