/*
 * FP32-state variant of the heat plugin: float storage, double arithmetic,
 * one rounding to float per stage (stochastic, reproducible: heat_store in
 * heat.cu).  Python half: heat.py.
 */
#include "heat.cu"

__device__ void step(float *shared_state, int idx, int globalTimeStep)
{
    heat_step(shared_state, idx, globalTimeStep);
}
