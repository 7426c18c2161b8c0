// optimize.cpp -- see optimize.h.
#include "optimize.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <numeric>

#include "phylo_hip.h"
#include "phylo_host.h"
#include "rate_model.h"
#include "simpleML.h"
#include "sub_model.h"

namespace phylo_amd {

nm_result nelder_mead(const std::function<double(const std::vector<double>&)>& f, std::vector<double> x0,
                      const std::vector<double>& lo, const std::vector<double>& hi, const nm_options& opt) {
    const size_t n = x0.size();
    auto clamp = [&](std::vector<double>& x) {
        for (size_t i = 0; i < n; ++i) x[i] = std::min(hi[i], std::max(lo[i], x[i]));
    };
    nm_result res;
    int evals = 0;
    auto eval = [&](const std::vector<double>& x) {
        ++evals;
        const double v = f(x);
        return std::isnan(v) ? -std::numeric_limits<double>::infinity() : v;
    };
    clamp(x0);
    std::vector<std::vector<double>> s(n + 1, x0);
    std::vector<double> fs(n + 1);
    for (size_t i = 0; i < n; ++i) {
        const double step = opt.steps.size() == n ? opt.steps[i] : opt.initial_step;
        s[i + 1][i] += (x0[i] + step <= hi[i]) ? step : -step;
        clamp(s[i + 1]);
    }
    for (size_t i = 0; i <= n; ++i) fs[i] = eval(s[i]);
    std::vector<size_t> idx(n + 1);
    const double dn = (double)n;
    const double gamma = opt.adaptive ? 1.0 + 2.0 / dn : 2.0;
    const double rho = opt.adaptive ? 0.75 - 0.5 / dn : 0.5;
    const double sigma = opt.adaptive ? 1.0 - 1.0 / dn : 0.5;
    while (evals < opt.max_evals) {
        std::iota(idx.begin(), idx.end(), 0);
        std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return fs[a] > fs[b]; });   // best first
        const size_t best = idx[0], worst = idx[n], second = idx[n - 1];
        double extent = 0.0;
        for (size_t i = 0; i <= n; ++i)
            for (size_t j = 0; j < n; ++j)
                extent = std::max(extent, std::fabs(s[i][j] - s[best][j]) / (1.0 + std::fabs(s[best][j])));
        if (std::isfinite(fs[best]) && fs[best] - fs[worst] < opt.ftol_abs && extent < std::sqrt(opt.xtol_rel)) {
            res.converged = true;
            break;
        }
        std::vector<double> c(n, 0.0);
        for (size_t k = 0; k < n; ++k)
            for (size_t j = 0; j < n; ++j) c[j] += s[idx[k]][j] / n;
        auto along = [&](double t) {
            std::vector<double> x(n);
            for (size_t j = 0; j < n; ++j) x[j] = c[j] + t * (s[worst][j] - c[j]);
            clamp(x);
            return x;
        };
        std::vector<double> xr = along(-1.0);
        const double fr = eval(xr);
        if (fr > fs[best]) {
            std::vector<double> xe = along(-gamma);
            const double fe = eval(xe);
            if (fe > fr) { s[worst] = xe; fs[worst] = fe; }
            else { s[worst] = xr; fs[worst] = fr; }
        } else if (fr > fs[second]) {
            s[worst] = xr;
            fs[worst] = fr;
        } else {
            const bool outside = fr > fs[worst];
            std::vector<double> xc = along(outside ? -rho : rho);
            const double fc = eval(xc);
            if (fc > std::max(fr, fs[worst])) {
                s[worst] = xc;
                fs[worst] = fc;
            } else {            // shrink towards the best point
                for (size_t i = 0; i <= n; ++i) {
                    if (i == best) continue;
                    for (size_t j = 0; j < n; ++j) s[i][j] = s[best][j] + sigma * (s[i][j] - s[best][j]);
                    clamp(s[i]);
                    fs[i] = eval(s[i]);
                }
            }
        }
    }
    size_t b = 0;
    for (size_t i = 1; i <= n; ++i)
        if (fs[i] > fs[b]) b = i;
    res.x = s[b];
    res.f = fs[b];
    res.evals = evals;
    return res;
}

// P(a, x): power series below a + 1, Legendre's continued fraction for
// Q = 1 - P above (modified Lentz evaluation).
double gamma_p(double a, double x) {
    if (!(x > 0.0)) return 0.0;
    const double lpre = -x + a * std::log(x) - std::lgamma(a);
    if (x < a + 1.0) {
        double term = 1.0 / a, sum = term;
        for (int k = 1; k < 100000; ++k) {
            term *= x / (a + k);
            sum += term;
            if (term < sum * 1e-17) break;
        }
        return std::min(1.0, sum * std::exp(lpre));
    }
    const double tiny = 1e-300;
    double b = x + 1.0 - a, cc = 1.0 / tiny, d = 1.0 / b, h = d;
    for (int k = 1; k < 100000; ++k) {
        const double an = -k * (k - a);
        b += 2.0;
        d = an * d + b;
        if (std::fabs(d) < tiny) d = tiny;
        cc = b + an / cc;
        if (std::fabs(cc) < tiny) cc = tiny;
        d = 1.0 / d;
        const double del = d * cc;
        h *= del;
        if (std::fabs(del - 1.0) < 1e-16) break;
    }
    return std::max(0.0, 1.0 - std::exp(lpre) * h);
}

// Halley iterations from a Wilson-Hilferty start, kept inside a bracket.
double gamma_p_inv(double a, double p) {
    if (p <= 0.0) return 0.0;
    if (p >= 1.0) return std::numeric_limits<double>::infinity();
    // normal quantile (Acklam's rational approximation) for the start
    auto norm_q = [](double q) {
        static const double A[] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                                   1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00};
        static const double B[] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                                   6.680131188771972e+01, -1.328068155288572e+01};
        static const double Cc[] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                                    -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00};
        static const double D[] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                                   3.754408661907416e+00};
        if (q < 0.02425) {
            const double u = std::sqrt(-2 * std::log(q));
            return (((((Cc[0] * u + Cc[1]) * u + Cc[2]) * u + Cc[3]) * u + Cc[4]) * u + Cc[5]) /
                   ((((D[0] * u + D[1]) * u + D[2]) * u + D[3]) * u + 1);
        }
        if (q > 1 - 0.02425) {
            const double u = std::sqrt(-2 * std::log(1 - q));
            return -(((((Cc[0] * u + Cc[1]) * u + Cc[2]) * u + Cc[3]) * u + Cc[4]) * u + Cc[5]) /
                   ((((D[0] * u + D[1]) * u + D[2]) * u + D[3]) * u + 1);
        }
        const double u = q - 0.5, r = u * u;
        return (((((A[0] * r + A[1]) * r + A[2]) * r + A[3]) * r + A[4]) * r + A[5]) * u /
               (((((B[0] * r + B[1]) * r + B[2]) * r + B[3]) * r + B[4]) * r + 1);
    };
    const double z = norm_q(p);
    const double t = 1.0 - 1.0 / (9.0 * a) + z / (3.0 * std::sqrt(a));
    double x = a * t * t * t;
    if (!(x > 0.0)) x = std::pow(p * std::tgamma(a + 1.0), 1.0 / a);
    double lo = 0.0, hi = std::numeric_limits<double>::infinity();
    const double lga = std::lgamma(a);
    for (int it = 0; it < 200; ++it) {
        const double fx = gamma_p(a, x) - p;
        if (fx < 0) lo = x; else hi = x;
        const double dens = std::exp(-x + (a - 1.0) * std::log(x) - lga);
        if (!(dens > 0.0)) {
            x = std::isfinite(hi) ? 0.5 * (lo + hi) : 2.0 * x + 1.0;
            continue;
        }
        const double u = fx / dens;
        double xn = x - u / (1.0 - 0.5 * std::min(1.0, u * ((a - 1.0) / x - 1.0)));
        if (!(xn > lo && xn < hi)) xn = std::isfinite(hi) ? 0.5 * (lo + hi) : 2.0 * x;
        if (std::fabs(xn - x) <= 1e-15 * x) return xn;
        x = xn;
    }
    return x;
}

std::vector<double> gamma_category_rates(double alpha, int C) {
    std::vector<double> r(C, 1.0);
    if (C <= 1) return r;
    double prev = 0.0;
    for (int k = 0; k < C; ++k) {
        const double I = (k == C - 1) ? 1.0 : gamma_p(alpha + 1.0, gamma_p_inv(alpha, double(k + 1) / C));
        r[k] = (I - prev) * C;
        prev = I;
    }
    return r;
}

double gtr_gamma_lnl(phylo_ctx* ctx, int n_nodes, const double* rel_len, int n_cat, const std::vector<double>& x) {
    double r[6];
    for (int i = 0; i < 5; ++i) r[i] = std::exp(x[i]);
    r[5] = 1.0;
    double pi[4], z = 0.0;
    const double logit[4] = {x[5], x[6], x[7], 0.0};
    const double mx = std::max(std::max(logit[0], logit[1]), std::max(logit[2], logit[3]));
    for (int i = 0; i < 4; ++i) z += (pi[i] = std::exp(logit[i] - mx));
    for (double& p : pi) p /= z;
    double Q[16] = {0};
    int k = 0;
    for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j, ++k) {
            Q[i * 4 + j] = r[k] * pi[j];
            Q[j * 4 + i] = r[k] * pi[i];
        }
    double mu = 0.0;
    for (int i = 0; i < 4; ++i) {
        double s = 0.0;
        for (int j = 0; j < 4; ++j)
            if (j != i) s += Q[i * 4 + j];
        Q[i * 4 + i] = -s;
        mu += pi[i] * s;
    }
    for (double& q : Q) q /= mu;
    const std::vector<double> rates = gamma_category_rates(std::exp(x[8]), n_cat);
    const std::vector<double> w(n_cat, 1.0 / n_cat);
    const double scale = std::exp(x[9]);
    std::vector<double> len(rel_len, rel_len + n_nodes);
    for (double& l : len) l *= scale;
    if (phylo_set_model(ctx, 4, Q, pi, pi, PHYLO_P_EIGEN) || phylo_set_categories(ctx, n_cat, rates.data(), w.data()) ||
        phylo_set_edge_lengths(ctx, len.data()))
        return -std::numeric_limits<double>::infinity();
    double total = 0.0;
    if (phylo_loglik(ctx, nullptr, &total)) return -std::numeric_limits<double>::infinity();
    return std::isfinite(total) ? total : -std::numeric_limits<double>::infinity();
}

gtr_fit fit_gtr_gamma(phylo_ctx* ctx, int n_nodes, const double* rel_len, int n_cat, const std::vector<double>& x0,
                      const nm_options& opt) {
    const auto t0 = std::chrono::steady_clock::now();
    const std::vector<double> lo = {-7, -7, -7, -7, -7, -7, -7, -7, std::log(0.02), std::log(1e-4)};
    const std::vector<double> hi = {7, 7, 7, 7, 7, 7, 7, 7, std::log(50.0), std::log(10.0)};
    auto f = [&](const std::vector<double>& x) { return gtr_gamma_lnl(ctx, n_nodes, rel_len, n_cat, x); };
    // Restart from the optimum with a fresh simplex until a restart gains
    // less than ftol_abs: a collapsed simplex can stall Nelder-Mead away from
    // the optimum in ten dimensions.
    nm_result r = nelder_mead(f, x0, lo, hi, opt);
    int evals = r.evals;
    // A log-scale coordinate can end far out on its flat tail (an
    // exchangeability driven towards zero, where a simplex step moves lnL by
    // less than ftol_abs, so the simplex shrinks there) while the optimum sits
    // at an ordinary value; small restart simplices stay on the tail.  Probe
    // each coordinate reset to its start value (n evaluations) and, when the
    // best probe gains more than ftol_abs, rerun the first run's full-size
    // simplex from it.  (One of 1,000 fits of 32 taxa x 2,000 sites from the
    // default start ended 188.7 lnL units low with one log exchangeability at
    // -6.8; the probe moving it back to 0 gained 67.7 and the rerun reached
    // the optimum a fit from a better start finds.)
    for (int round = 0; round < 3 && evals + (int)x0.size() < opt.max_evals; ++round) {
        double best_f = r.f + opt.ftol_abs;
        std::vector<double> best_x;
        for (size_t i = 0; i < x0.size(); ++i) {
            std::vector<double> y = r.x;
            y[i] = std::min(hi[i], std::max(lo[i], x0[i]));
            if (y[i] == r.x[i]) continue;
            const double fy = f(y);
            ++evals;
            if (fy > best_f) { best_f = fy; best_x = y; }
        }
        if (best_x.empty() || evals >= opt.max_evals) break;
        nm_options o = opt;
        o.max_evals = opt.max_evals - evals;
        const nm_result r2 = nelder_mead(f, best_x, lo, hi, o);
        evals += r2.evals;
        if (r2.f > r.f) r = r2;
    }
    for (int restart = 0; restart < opt.max_restarts && evals < opt.max_evals; ++restart) {
        nm_options o = opt;
        o.max_evals = opt.max_evals - evals;
        o.initial_step = std::min(opt.initial_step, opt.restart_step);
        for (double& st : o.steps) st = std::min(st, opt.restart_step);
        const nm_result r2 = nelder_mead(f, r.x, lo, hi, o);
        evals += r2.evals;
        const bool gained = r2.f > r.f + opt.ftol_abs;
        if (r2.f > r.f) r = r2;
        if (!gained) break;
    }
    gtr_fit out;
    out.x = r.x;
    out.lnl = r.f;
    out.evals = evals;
    out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return out;
}

namespace {
bool contains(const std::vector<unsigned int>& v, unsigned int x) { return std::find(v.begin(), v.end(), x) != v.end(); }
}  // namespace

model_objective::model_objective(simpleML& t_, sub_model& model_, rate_model& rate_mod_,
                                 const std::vector<unsigned int>& fixed_, bool fixed_freq,
                                 const std::vector<unsigned int>& fixed_extra_, char method_)
    : t(t_), model(model_), rate_mod(rate_mod_), fixed(fixed_), fixed_extra(fixed_extra_),
      free_freq(model_.is_tr() && !fixed_freq), method(method_) {
    const double big = 1e6;     // the reference's DBL_MAX, kept finite for the simplex arithmetic
    for (unsigned int i = 0; i < model.get_n_rates(); ++i)
        if (!contains(fixed, i)) { x0.push_back(model.get_rate(i)); lo.push_back(0.0); hi.push_back(big); }
    start_freq = x0.size();
    if (free_freq)
        for (unsigned int i = 0; i + 1 < model.get_n_states(); ++i) {
            x0.push_back(model.get_freq(i));
            lo.push_back(0.0);
            hi.push_back(1.0);
        }
    start_extra = x0.size();
    if (rate_mod.get_n_rates())
        for (unsigned int i = 0; i < rate_mod.get_n_parameters(); ++i) {
            if (contains(fixed_extra, i)) continue;
            lo.push_back(0.0);
            if ((method == 't' || method == 'S') && i < rate_mod.get_n_cut_offs()) {
                hi.push_back(t.longest_to_tip());
                // the reference starts every cut-off at the first one's value
                // (get_cut_off(0), treeator.cpp:774)
                x0.push_back(rate_mod.get_cut_off(0));
            } else {
                hi.push_back(big);
                if (method == 't' || method == 'S') x0.push_back(rate_mod.get_rate_in_time(i - rate_mod.get_n_cut_offs()));
                else x0.push_back(rate_mod.get_rate_clade(i));
            }
        }
}

bool model_objective::apply(const std::vector<double>& x) {
    size_t j = 0;
    for (unsigned int i = 0; i < model.get_n_rates(); ++i)
        if (!contains(fixed, i) && !model.set_rate(i, x[j++])) return false;
    if (free_freq)
        for (unsigned int i = 0; i + 1 < model.get_n_states(); ++i)
            if (!model.set_freq(i, x[start_freq + i])) return false;
    j = start_extra;
    for (unsigned int i = 0; i < rate_mod.get_n_parameters() && rate_mod.get_n_rates(); ++i)
        if (!contains(fixed_extra, i)) rate_mod.set_parameter(i, (float)x[j++]);
    return true;
}

double model_objective::operator()(const std::vector<double>& x) {
    const double ninf = -std::numeric_limits<double>::infinity();
    if (!apply(x)) return ninf;
    if ((method == 't' || method == 'S') && rate_mod.get_n_time_rates() < 1) return ninf;
    const double l = t.calculate_log_likelihood(model, rate_mod);
    return (l == 0.0 || !std::isfinite(l)) ? ninf : l;
}

double optimize_parameters(simpleML& t, sub_model& model, rate_model& rate_mod, const std::vector<unsigned int>& fixed,
                           bool fixed_freq, const std::vector<unsigned int>& fixed_extra, char method,
                           const nm_options& opt, int* evals) {
    model_objective obj(t, model, rate_mod, fixed, fixed_freq, fixed_extra, method);
    if (obj.x0.empty()) return t.calculate_log_likelihood(model, rate_mod);
    const nm_result r = nelder_mead([&](const std::vector<double>& x) { return obj(x); }, obj.x0, obj.lo, obj.hi, opt);
    obj.apply(r.x);
    if (evals) *evals = r.evals;
    return t.calculate_log_likelihood(model, rate_mod);
}

}  // namespace phylo_amd

// ---- C entry points (declared in include/phylo_host.h) ----
extern "C" {

int phylo_gamma_rates(double alpha, int n_cat, double* rates) {
    if (!rates || n_cat < 1 || !(alpha > 0.0)) return PHYLO_ERR_ARG;
    const std::vector<double> r = phylo_amd::gamma_category_rates(alpha, n_cat);
    std::copy(r.begin(), r.end(), rates);
    return PHYLO_OK;
}

int phylo_nm_defaults(phylo_nm_options* o) {
    if (!o) return PHYLO_ERR_ARG;
    const phylo_amd::nm_options d;
    o->max_evals = d.max_evals;
    o->ftol_abs = d.ftol_abs;
    o->xtol_rel = d.xtol_rel;
    o->initial_step = d.initial_step;
    o->adaptive = d.adaptive;
    o->max_restarts = d.max_restarts;
    o->steps = nullptr;
    o->restart_step = d.restart_step;
    return PHYLO_OK;
}

int phylo_fit_gtr_gamma(phylo_ctx* ctx, int n_nodes, const double* rel_len, int n_cat, const double* x0,
                        const phylo_nm_options* nm, double* x_out, double* lnl, int* evals, double* seconds) {
    if (!ctx || !rel_len || !x0 || n_cat < 1) return PHYLO_ERR_ARG;
    phylo_amd::nm_options opt;
    if (nm) {
        if (nm->max_evals < 1 || !(nm->ftol_abs >= 0) || !(nm->xtol_rel >= 0) || !(nm->initial_step > 0) ||
            nm->max_restarts < 0 || !(nm->restart_step > 0))
            return PHYLO_ERR_ARG;
        opt.max_evals = nm->max_evals;
        opt.ftol_abs = nm->ftol_abs;
        opt.xtol_rel = nm->xtol_rel;
        opt.initial_step = nm->initial_step;
        opt.adaptive = nm->adaptive != 0;
        opt.max_restarts = nm->max_restarts;
        opt.restart_step = nm->restart_step;
        if (nm->steps) {
            for (int i = 0; i < 10; ++i)
                if (!(nm->steps[i] > 0)) return PHYLO_ERR_ARG;
            opt.steps.assign(nm->steps, nm->steps + 10);
        }
    }
    const phylo_amd::gtr_fit f =
        phylo_amd::fit_gtr_gamma(ctx, n_nodes, rel_len, n_cat, std::vector<double>(x0, x0 + 10), opt);
    if (x_out) std::copy(f.x.begin(), f.x.end(), x_out);
    if (lnl) *lnl = f.lnl;
    if (evals) *evals = f.evals;
    if (seconds) *seconds = f.seconds;
    return PHYLO_OK;
}

double phylo_gtr_gamma_lnl(phylo_ctx* ctx, int n_nodes, const double* rel_len, int n_cat, const double* x) {
    return phylo_amd::gtr_gamma_lnl(ctx, n_nodes, rel_len, n_cat, std::vector<double>(x, x + 10));
}

}  // extern "C"
