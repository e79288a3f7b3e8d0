/*
 * FP32-state variant of the Euler plugin: the state is stored as float in
 * HBM and LDS (half the LDS per tile), every operation of the step -- the
 * pressure recovery p = (gamma-1)(E - |m|^2/(2 rho)) included -- is
 * evaluated in double from the float values, and the result is rounded to
 * float once per stage.  Python half: euler.py.
 */
#include "euler.cu"

__device__ void step(float *shared_state, int idx, int globalTimeStep)
{
    eu_step(shared_state, idx, globalTimeStep);
}
