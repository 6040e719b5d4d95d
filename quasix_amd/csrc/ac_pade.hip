// Analytic continuation of the imaginary-axis self-energy (host code).
//
// Reference: quasix_core/src/selfenergy/correlation_ac.rs
//   thiele               :470-509   reciprocal differences, 1e10 for |denominator| < 1e-14
//   pade_thiele          :530-566   continued fraction evaluated at a real frequency
//   ac_pade_thiele_diag  :587-654   fit points 1, 7, ..., 37 then every 4th, even count
// and the per-orbital QP step of the evGW driver that consumes the fit,
// quasix_core/src/qp/evgw.rs:705-820: linearised E = eps + Z (Sigma_x + Re Sigma_c - V_xc)
// with a forward-difference Z (delta 1e-6, clamped), or the Newton solve of
// quasix_core/src/qp/newton_solver.rs (shared qx::newton_core) started from that linearised
// value at iteration 0 or from the previous QP energy, falling back to the linearised value
// at the starting point when Newton fails.  The fits are tiny (<= ~16 points per orbital),
// so they run on the host next to the driver; the Sigma(i nu) data they consume comes from
// qx_sigma_ac_dev.
#include "../../include/qxgw.h"
#include "qx_internal.h"
#include "qx_newton.h"

#include <cmath>
#include <complex>
#include <vector>

using cplx = std::complex<double>;

namespace {

constexpr double kTiny = 1e-14, kHuge = 1e10;

std::vector<cplx> thiele(const cplx* fn, const cplx* zn, int n) {
    std::vector<cplx> g((size_t)n * n), a(n);
    for (int i = 0; i < n; ++i) g[(size_t)i * n] = fn[i];
    for (int j = 1; j < n; ++j)
        for (int i = j; i < n; ++i) {
            const cplx d = (zn[i] - zn[j - 1]) * g[(size_t)i * n + j - 1];
            g[(size_t)i * n + j] = std::abs(d) < kTiny
                                       ? cplx(kHuge, 0.0)
                                       : (g[(size_t)(j - 1) * n + j - 1] - g[(size_t)i * n + j - 1]) / d;
        }
    for (int i = 0; i < n; ++i) a[i] = g[(size_t)i * n + i];
    return a;
}

cplx pade_thiele(double omega, const cplx* zn, const cplx* coeff, int n, int stride = 1) {
    if (n == 0) return cplx(0.0, 0.0);
    if (n == 1) return coeff[0];
    const cplx z(omega, 0.0);
    cplx x = coeff[(size_t)(n - 1) * stride] * (z - zn[n - 2]);
    for (int i = n - 2; i >= 1; --i) {
        const cplx d = 1.0 + x;
        x = std::abs(d) < kTiny ? cplx(kHuge, 0.0) : coeff[(size_t)i * stride] * (z - zn[i - 1]) / d;
    }
    const cplx d = 1.0 + x;
    return std::abs(d) < kTiny ? coeff[0] : coeff[0] / d;
}

std::vector<int> fit_points(int nw) {
    std::vector<int> idx;
    for (int i = 1; i < std::min(40, nw); i += 6) idx.push_back(i);
    if (!idx.empty())
        for (int i = idx.back() + 4; i < nw; i += 4) idx.push_back(i);
    idx.resize(idx.size() / 2 * 2);
    return idx;
}

const cplx* as_c(const double* p) { return reinterpret_cast<const cplx*>(p); }
cplx* as_c(double* p) { return reinterpret_cast<cplx*>(p); }

// Padé model of one orbital as a Newton evaluator (Sigma at E is the fit at E - ef).
struct PadeEval {
    const cplx* zn;
    const cplx* coeff;
    int n, stride;
    double ef;
    bool operator()(const double* E, int nq, qx::C2* out) const {
        for (int q = 0; q < nq; ++q) {
            const cplx v = pade_thiele(E[q] - ef, zn, coeff, n, stride);
            out[q] = {v.real(), v.imag()};
        }
        return false;
    }
};

}  // namespace

extern "C" {

qx_status qx_thiele(int n, const double* fn, const double* zn, double* coeff) {
    if (n < 0 || (n > 0 && (!fn || !zn || !coeff))) return QX_ERR_INVALID_INPUT;
    const std::vector<cplx> a = thiele(as_c(fn), as_c(zn), n);
    for (int i = 0; i < n; ++i) as_c(coeff)[i] = a[i];
    return QX_OK;
}

qx_status qx_pade_thiele(int n, double omega, const double* zn, const double* coeff,
                         double* out) {
    if (n < 0 || !out || (n > 0 && (!zn || !coeff))) return QX_ERR_INVALID_INPUT;
    *as_c(out) = pade_thiele(omega, as_c(zn), as_c(coeff), n);
    return QX_OK;
}

qx_status qx_ac_pade_thiele_diag(int norbs, int nw, const double* sigma, const double* omega,
                                 double* coeff, double* ofit, int* nfit) {
    if (norbs <= 0 || nw <= 0 || !sigma || !omega || !coeff || !ofit || !nfit)
        return QX_ERR_INVALID_INPUT;
    const std::vector<int> idx = fit_points(nw);
    const int nf = (int)idx.size();
    std::vector<cplx> fn(nf), zn(nf);
    for (int p = 0; p < norbs; ++p) {
        for (int i = 0; i < nf; ++i) {
            fn[i] = as_c(sigma)[(size_t)p * nw + idx[i]];
            zn[i] = as_c(omega)[(size_t)p * nw + idx[i]];
            as_c(ofit)[(size_t)p * nf + i] = zn[i];
        }
        const std::vector<cplx> a = thiele(fn.data(), zn.data(), nf);
        for (int i = 0; i < nf; ++i) as_c(coeff)[(size_t)i * norbs + p] = a[i];
    }
    *nfit = nf;
    return QX_OK;
}

qx_status qx_qp_solve_ac(int norbs, int nfit, const double* coeff, const double* ofit,
                         const int* orbs, int nmo, int nocc, const double* mo_energy,
                         const double* sigma_x, const double* vxc, int solver,
                         const double* prev_qp, const qx_newton_config* cfg, double* qp_energy,
                         double* z_factor, double* sigma_re, double* sigma_im, int* state) {
    if (norbs <= 0 || nfit < 0 || !coeff || !ofit || !orbs || !mo_energy || !sigma_x || !vxc ||
        !cfg || !qp_energy || !z_factor || !sigma_re || !sigma_im || nocc <= 0 || nocc >= nmo ||
        (solver != 0 && solver != 1))
        return QX_ERR_INVALID_INPUT;
    qx::NewtonParams np;
    np.max_iterations = cfg->max_iterations;
    np.energy_tolerance = cfg->energy_tolerance;
    np.initial_damping = cfg->initial_damping;
    np.min_damping = cfg->min_damping;
    np.max_damping = cfg->max_damping;
    np.max_step = cfg->max_step;
    np.derivative_delta = cfg->derivative_delta;
    np.z_min = cfg->z_min;
    np.z_max = cfg->z_max;
    np.use_richardson = cfg->use_richardson;
    np.use_bisection = cfg->use_bisection_fallback;
    np.max_bisection_iterations = cfg->max_bisection_iterations;
    np.linearized = 0;
    const double ef = 0.5 * (mo_energy[nocc - 1] + mo_energy[nocc]);
    const double de = 1e-6;
    for (int i = 0; i < norbs; ++i) {
        const int p = orbs[i];
        if (p < 0 || p >= nmo) return QX_ERR_INVALID_INPUT;
        const PadeEval ev{as_c(ofit) + (size_t)i * nfit, as_c(coeff) + i, nfit, norbs, ef};
        const double stat = sigma_x[p] - vxc[p];
        auto linear_at = [&](double e_eval, double& E, double& Z, cplx& s) {
            s = pade_thiele(e_eval - ef, ev.zn, ev.coeff, nfit, norbs);
            const cplx sp = pade_thiele(e_eval - ef + de, ev.zn, ev.coeff, nfit, norbs);
            Z = std::fmin(std::fmax(1.0 / (1.0 - (sp.real() - s.real()) / de), cfg->z_min),
                          cfg->z_max);
            E = mo_energy[p] + Z * (stat + s.real());
        };
        double E, Z;
        cplx s;
        int st = QX_QP_LINEARIZED;
        const double start = prev_qp ? prev_qp[p] : mo_energy[p];
        if (solver == 0) {
            linear_at(start, E, Z, s);
        } else {
            double guess = start;
            if (!prev_qp) {
                double Zg;
                cplx sg;
                linear_at(mo_energy[p], guess, Zg, sg);
            }
            qx::C2 sc{0.0, 0.0};
            int it = 0;
            PadeEval evm = ev;
            const int r = qx::newton_core(evm, np, mo_energy[p], stat, ef, true, guess, E, Z, sc,
                                          it);
            if (r == qx::kNewtonConverged || r == qx::kNewtonBisection) {
                s = cplx(sc.re, sc.im);
                st = r == qx::kNewtonConverged ? QX_QP_NEWTON : QX_QP_BISECTION;
            } else {
                linear_at(guess, E, Z, s);
            }
        }
        qp_energy[i] = E;
        z_factor[i] = Z;
        sigma_re[i] = s.real();
        sigma_im[i] = s.imag();
        if (state) state[i] = st;
    }
    return QX_OK;
}

}  // extern "C"
