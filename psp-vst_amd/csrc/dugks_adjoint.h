// Adjoint of N DUGKS steps (dugks_adjoint.hip), driven by dugks_adjoint in dugks_api.hip.
#pragma once

#include <string>

#include "dugks_internal.h"

namespace dugks {

struct AdjointIO {
    int n_steps = 0;
    const double *dts = nullptr;         // [n_steps] host
    int ckpt = 1;                        // checkpoint every ckpt steps
    const double *q_weight = nullptr;    // [nF][3] host or null
    const double *mom_weight = nullptr;  // [nF][3] host or null
    // host outputs (any may be null); dg0 / dh0 in the internal cell numbering [Nv][nC]
    double *J = nullptr, *dmu = nullptr, *dTw = nullptr, *dUw = nullptr, *dg0 = nullptr, *dh0 = nullptr;
};

// Runs the forward N steps on the state in d (staged kernels, storage S), the reverse
// sweep, and leaves d at the state after N steps.  Returns 0 or 1 with err set.
template <class S>
int run_adjoint(const Geo &g, DevState &d, hipStream_t st, const AdjointIO &io, std::string &err);

}  // namespace dugks
