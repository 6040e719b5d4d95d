// Newton-Raphson quasiparticle solve, shared by the device kernels (a workgroup per target:
// every thread runs the same iteration on the same reduced values) and the host-driven path
// used when residue poles appear.
//
// Restates quasix_core/src/qp/newton_solver.rs:174-391:
//   f(E) = E - eps_n - static - Re Sigma_c(E)      (static = Sigma_x - V_xc)
//   derivative by Richardson 5-point stencil with delta (:282-303), evaluated at
//   omega = E - ef and mapped back as omega + ef, as the reference's closure does;
//   adaptive damping x0.5 / x1.2 (:249-255), step clip (:246), convergence on |f| before
//   the step (:201), Z = clamp(1/(1 - dSigma/dE), z_min, z_max) (:306-315);
//   bisection fallback with brackets +-2 Ha x 2^k (:317-391).
// and the iteration-0 linearised initial guess of quasix_core/src/qp/evgw.rs:730-741
// (forward difference, clamped Z).  With np.linearized set it returns the linearised QP
// step itself (the driver's default solver, evgw.rs:786-820).
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include "qx_internal.h"

namespace qx {

struct C2 { double re, im; };

// Eval: bool operator()(const double* E, int nq, C2* out) -- evaluates Sigma_c at nq <= 5
// energies; returns true to abort (the device evaluator cannot evaluate residues).
template <class Eval>
__host__ __device__ inline int newton_core(Eval& ev, const NewtonParams& np, double eps_n,
                                           double stat, double ef, bool have_guess,
                                           double guess, double& E_out, double& Z_out,
                                           C2& s_out, int& iters_out) {
    const double de = np.derivative_delta;
    double E;
    C2 s[5];
    if (np.linearized) {
        // Linearised QP step at the start point (the guess, else eps_n) with a forward-
        // difference Z clamped to [z_min, z_max] (quasix_core/src/qp/evgw.rs:786-820).
        const double om = (have_guess ? guess : eps_n) - ef;
        double Es[2] = {om + ef, (om + de) + ef};
        if (ev(Es, 2, s)) return kNewtonNeedsResidue;
        double z = 1.0 / (1.0 - (s[1].re - s[0].re) / de);
        z = fmin(fmax(z, np.z_min), np.z_max);
        E_out = eps_n + z * (stat + s[0].re);
        Z_out = z;
        s_out = s[0];
        iters_out = 0;
        return kNewtonLinearized;
    }
    if (have_guess) {
        E = guess;
    } else {
        const double om = eps_n - ef;
        double Es[2] = {om + ef, (om + de) + ef};
        if (ev(Es, 2, s)) return kNewtonNeedsResidue;
        double z = 1.0 / (1.0 - (s[1].re - s[0].re) / de);
        z = fmin(fmax(z, np.z_min), np.z_max);
        E = eps_n + z * (stat + s[0].re);
    }
    const double e_guess = E;
    double damping = np.initial_damping, prev = INFINITY;
    for (int it = 0; it < np.max_iterations; ++it) {
        const double om = E - ef;
        double Es[5] = {om + ef, (om + de) + ef, (om - de) + ef, (om + 2.0 * de) + ef,
                        (om - 2.0 * de) + ef};
        if (ev(Es, 5, s)) return kNewtonNeedsResidue;
        const double d = np.use_richardson
                             ? (8.0 * (s[1].re - s[2].re) - (s[3].re - s[4].re)) / (12.0 * de)
                             : (s[1].re - s[2].re) / (2.0 * de);
        const double res = E - eps_n - stat - s[0].re;
        const double ares = fabs(res);
        if (ares < np.energy_tolerance) {
            E_out = E;
            Z_out = fmin(fmax(1.0 / (1.0 - d), np.z_min), np.z_max);
            s_out = s[0];
            iters_out = it + 1;
            return kNewtonConverged;
        }
        const double fp = 1.0 - d;
        double step = fabs(fp) < 1e-10 ? -copysign(1.0, res) * fmin(np.max_step, ares) : -res / fp;
        step = fmin(fmax(step, -np.max_step), np.max_step);
        if (ares > prev) damping = fmax(damping * 0.5, np.min_damping);
        else if (ares < 0.5 * prev && damping < np.max_damping)
            damping = fmin(damping * 1.2, np.max_damping);
        E += damping * step;
        prev = ares;
    }
    E_out = E;
    iters_out = np.max_iterations;
    if (!np.use_bisection) return kNewtonFailed;
    // Bisection fallback.
    bool abort = false;
    auto resid = [&](double e) {
        double Es[1] = {(e - ef) + ef};
        C2 r[1];
        if (ev(Es, 1, r)) abort = true;
        return e - eps_n - stat - r[0].re;
    };
    const double width = 2.0;
    double lo = e_guess - width, hi = e_guess + width;
    if (resid(lo) * resid(hi) > 0.0) {
        for (int k = 1; k < 5; ++k) {
            const double w = width * (double)(1 << k);
            lo = e_guess - w;
            hi = e_guess + w;
            if (resid(lo) * resid(hi) <= 0.0) break;
        }
    }
    if (abort) return kNewtonNeedsResidue;
    for (int it = 0; it < np.max_bisection_iterations; ++it) {
        const double mid = 0.5 * (lo + hi);
        const double fm = resid(mid);
        if (abort) return kNewtonNeedsResidue;
        if (fabs(fm) < np.energy_tolerance || 0.5 * (hi - lo) < np.energy_tolerance) {
            const double om = mid - ef;
            double Es[5] = {om + ef, (om + de) + ef, (om - de) + ef, (om + 2.0 * de) + ef,
                            (om - 2.0 * de) + ef};
            if (ev(Es, 5, s)) return kNewtonNeedsResidue;
            const double d = np.use_richardson
                                 ? (8.0 * (s[1].re - s[2].re) - (s[3].re - s[4].re)) / (12.0 * de)
                                 : (s[1].re - s[2].re) / (2.0 * de);
            E_out = mid;
            Z_out = fmin(fmax(1.0 / (1.0 - d), np.z_min), np.z_max);
            s_out = s[0];
            iters_out = it + 1;
            return kNewtonBisection;
        }
        const double fl = resid(lo);
        if (abort) return kNewtonNeedsResidue;
        if (fm * fl < 0.0) hi = mid;
        else lo = mid;
    }
    return kNewtonFailed;
}

}  // namespace qx
