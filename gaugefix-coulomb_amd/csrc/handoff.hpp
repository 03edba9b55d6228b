// handoff.hpp -- the FP32 -> FP64 hand-off policy of the MIXED precision mode
// (host code only; no HIP dependency, so that it can be tested on its own).
#pragma once

#include <cmath>
#include <cstdint>

#include "gfc.h"

namespace gfcx {

// FP32 -> FP64 hand-off of the MIXED mode, shared by the single-lattice driver
// and the serving scheduler so that both decide alike from the same residues.
// The FP32 phase is observed at its probes and at scheduled checks.  It is
// finished when its residue reaches mixed_switch or stops falling: the rate
// over the last interval (decades per sweep) drops below a quarter of the
// last clean rate (an interval is clean at 0.7 of the previous clean rate or
// more), which is the FP32 rounding floor.  A finished phase hands over at
// once, or, with mixed_defer, at the sweep that leaves before the next
// scheduled check the FP64 sweeps needed to bring the residue to the
// tolerance at the clean rate (x1.1 + 10 sweeps): earlier FP64 sweeps cannot
// change the outcome of that check, and FP32 sweeps hold the state at its
// floor for half the cost.  Once the floor is reached the FP32 phase is not
// probed any more (the hand-off sweep is fixed).
struct Handoff {
    double rho = 0;   // decades per sweep of the last clean interval
    double last = 0;  // previous residue (0: none) and its sweep
    int last_sweep = 0;
    bool finished = false;
    bool at_floor = false;  // the residue stopped falling: later probes are moot
    int switch_at = 0x7fffffff;

    void start(double res0) {
        *this = Handoff();
        last = res0;
    }
    // Residue g of the FP32 state after `sweep` sweeps, `due` the next
    // scheduled check.  Returns true when the hand-off is due now.
    bool observe(const gfc_options &o, double tol, int sweep, double g, int due) {
        if (!at_floor && last > 0 && g > 0 && sweep > last_sweep) {
            const double r = g < last ? std::log10(last / g) / (sweep - last_sweep) : 0.0;
            if (rho > 0 && r < 0.25 * rho)
                finished = at_floor = true;
            else if (!finished && r > 0 && (rho == 0 || r >= 0.7 * rho))
                rho = r;
        }
        last = g;
        last_sweep = sweep;
        if (g <= o.mixed_switch) finished = true;
        if (!finished) return false;
        if (!o.mixed_defer || rho <= 0 || g <= tol) return true;
        const int n64 = (int)std::ceil(1.1 * std::log10(g / tol) / rho) + 10;
        if ((int64_t)sweep + n64 >= due) return true;
        switch_at = due - n64;
        return false;
    }
};

}  // namespace gfcx
