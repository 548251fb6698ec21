"""Adaptive time step of the dugksFoam loop (src/setDeltaTvar.H:34-50).

Given the Courant number of the previous step's face velocities (fvDVM::getCoNum,
src/fvDVM/fvDVM/fvDVM.C:2177-2185) the step is rescaled towards maxCo: reductions are
immediate, increases are damped to at most 1 + 0.1 * maxCo/Co and capped at 1.2x, and the
result is capped at maxDeltaT.
"""
SMALL = 1.0e-15


def next_delta_t(dt: float, co_num: float, controls) -> float:
    if not controls.adjustTimeStep:
        return dt
    max_delta_t_fact = controls.maxCo / (co_num + SMALL)
    delta_t_fact = min(min(max_delta_t_fact, 1.0 + 0.1 * max_delta_t_fact), 1.2)
    return min(delta_t_fact * dt, controls.maxDeltaT)
