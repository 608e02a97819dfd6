// ic_internal.h -- pieces of the host initial-condition generators that the
// device-side Monte-Carlo Glauber admission (vah_batch_admit_glauber) shares.
#pragma once

#include <vector>

#include "../../include/vah_host.h"

namespace vah_host {

// The nucleon positions of vah_ic_mc_glauber's two nuclei for g (the same
// generator, seed and draw order): nucleus A centred at x = -b/2 in xa, ya,
// nucleus B at +b/2 in xb, yb, g.A each.  Returns VAH_OK or VAH_ERR_ARG.
int mc_glauber_nucleons(const vah_glauber &g, std::vector<double> &xa, std::vector<double> &ya,
                        std::vector<double> &xb, std::vector<double> &yb);

}  // namespace vah_host
