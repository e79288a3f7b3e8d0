/*
 * Known-answer plugin: midpoint RK2 of the heat equation with the
 * coefficient 1/100 (reference: pysweep/equations/half.cu:20-34).
 */
__device__ __constant__ double DX;
__device__ __constant__ double DY;
__device__ __constant__ double DT;

__device__ void step(double *shared_state, int idx, int globalTimeStep)
{
    const double *u = shared_state;
    const double c = u[idx];
    double d = (u[idx + SY] - 2 * c + u[idx - SY]) / 100.0;
    d = d + (u[idx + 1] - 2 * c + u[idx - 1]) / 100.0;
    if (globalTimeStep % ITS == 0)
        shared_state[idx + TIMES] = shared_state[idx - TIMES] + 1.0 * d;
    else
        shared_state[idx + TIMES] = shared_state[idx] + 0.5 * d;
}
