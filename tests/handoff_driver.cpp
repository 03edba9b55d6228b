// Drives gfcx::Handoff (csrc/handoff.hpp) from stdin for tests/test_host_handoff.py:
//   O mixed_switch mixed_defer tolerance   set the options
//   S residue0                             start a fix
//   X sweep residue due                    one observation; prints
//                                          "ret finished at_floor switch_at rho"
#include <cstdio>
#include <cstring>

#include "handoff.hpp"

int main() {
    gfc_options o;
    std::memset(&o, 0, sizeof o);
    o.mixed_switch = 1e-9;
    o.mixed_defer = 1;
    double tol = 1e-16;
    gfcx::Handoff h;
    char cmd[8];
    while (std::scanf("%7s", cmd) == 1) {
        if (cmd[0] == 'O') {
            if (std::scanf("%lf %d %lf", &o.mixed_switch, &o.mixed_defer, &tol) != 3) return 2;
        } else if (cmd[0] == 'S') {
            double r0;
            if (std::scanf("%lf", &r0) != 1) return 2;
            h.start(r0);
        } else if (cmd[0] == 'X') {
            int sweep, due;
            double g;
            if (std::scanf("%d %lf %d", &sweep, &g, &due) != 3) return 2;
            const bool ret = h.observe(o, tol, sweep, g, due);
            std::printf("%d %d %d %d %.17g\n", ret ? 1 : 0, h.finished ? 1 : 0,
                        h.at_floor ? 1 : 0, h.switch_at, h.rho);
        } else {
            return 2;
        }
    }
    return 0;
}
