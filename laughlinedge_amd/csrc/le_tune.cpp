// Step-size search of edgeMC::optimal_step (modules/edge_mc/em_bisection.cu:49-153), over a
// caller-supplied acceptance measurement, so that it runs (and is tested) without a GPU.
//
// The reference measures the acceptance of a fresh launch at the current step; if it is
// within tol of the target it keeps it; otherwise it doubles (acceptance above target) or
// halves the step until the target is bracketed, then refines by linear interpolation
// between the bracket ends (false position), replacing the end on the same side of the target
// as the guess.  It loops without bound and exits the process when a guess falls outside the
// bracket.  Here the bracket search and the refinement are capped (40 and 60 measurements)
// and every failure is an error code.
#include <cmath>
#include <string>
#include <utility>

#include "../../include/laughlinedge.h"

namespace le {
void set_error(const std::string& msg);
}

extern "C" int le_step_search(le_accept_fn measure, void* ctx, double step0, double target, double tol,
                              double* step_out, int* n_measured)
{
    int count = 0;
    if (n_measured) *n_measured = 0;
    if (!measure || !step_out || !(step0 > 0) || !(tol > 0) || !(target > 0) || !(target < 1)) {
        le::set_error("le_step_search: need a measure function, step0 > 0, tol > 0, 0 < target < 1");
        return LE_ERR_ARG;
    }
    int rc = LE_OK;
    auto meas = [&](double s) {
        double a = 0;
        if (rc == LE_OK) {
            rc = measure(ctx, s, &a);
            ++count;
            if (n_measured) *n_measured = count;
            if (rc == LE_OK && !(a >= 0 && a <= 1)) {
                le::set_error("le_step_search: measured acceptance outside [0, 1]");
                rc = LE_ERR_ARG;
            }
        }
        return a;
    };
    auto accept = [&](double s) {
        *step_out = s;
        return LE_OK;
    };
    double s2 = step0, s1 = step0;
    double a2 = meas(s2), a1 = a2;
    if (rc) return rc;
    if (std::fabs(a2 - target) < tol) return accept(s2);
    // bracket: double (acceptance above target) or halve until the target is crossed
    for (int it = 0;; ++it) {
        s1 *= (a2 > target) ? 2.0 : 0.5;
        a1 = meas(s1);
        if (rc) return rc;
        if ((a2 - target) * (a1 - target) <= 0) break;
        if (it >= 40) {
            le::set_error("step search failed to bracket the target acceptance");
            return LE_ERR_BISECTION;
        }
    }
    if (std::fabs(a1 - target) < tol) return accept(s1);
    if (s2 < s1) {
        std::swap(s1, s2);
        std::swap(a1, a2);
    }
    for (int it = 0; it < 60; ++it) {
        const double g = ((target - a1) * s2 + (a2 - target) * s1) / (a2 - a1);
        const double ag = meas(g);
        if (rc) return rc;
        if (std::fabs(ag - target) < tol) return accept(g);
        if ((ag - target) * (a1 - target) < 0) {
            a2 = ag;
            s2 = g;
        } else if ((ag - target) * (a2 - target) < 0) {
            a1 = ag;
            s1 = g;
        } else {
            // the reference prints and exits here (em_bisection.cu:146-149)
            le::set_error("step bisection lost its bracket");
            return LE_ERR_BISECTION;
        }
    }
    le::set_error("step bisection did not converge in 60 steps");
    return LE_ERR_BISECTION;
}
