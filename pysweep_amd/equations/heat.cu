/*
 * Heat equation u_t = alpha*(u_xx + u_yy) on a periodic grid: 5-point
 * Laplacian with forward Euler or midpoint RK2.
 *
 * Device half of the heat plugin pair (heat.py).  The arithmetic follows the
 * numpy step of the reference plugin operation by operation
 * (pysweep/equations/heat.py:48-64): the Laplacian is
 * COURANT*(u[x+1]-2u+u[x-1]) + COURANT*(u[y+1]-2u+u[y-1]) with
 * COURANT = alpha*dt/dx/dx, so the device and numpy results agree bitwise.
 *
 * Contract: step(shared_state, idx, globalTimeStep) writes
 * shared_state[idx + TIMES]; x neighbours are idx +- SY, y neighbours
 * idx +- 1, the level one back is idx - TIMES.
 */
__device__ __constant__ double DX;
__device__ __constant__ double DY;
__device__ __constant__ double DT;
__device__ __constant__ double ALPHA;
__device__ __constant__ double COURANT;
__device__ __constant__ int SCHEME; /* 1: forward Euler, 0: RK2 */

/* arithmetic in double; R is the storage type of the state */
template <class R>
__device__ __forceinline__ double heat_laplace(const R *u, int idx)
{
    const double c = u[idx];
    const double lx = COURANT * ((double)u[idx + SY] - 2 * c + (double)u[idx - SY]);
    const double ly = COURANT * ((double)u[idx + 1] - 2 * c + (double)u[idx - 1]);
    return lx + ly;
}

template <class R>
__device__ __forceinline__ void heat_step(R *shared_state, int idx, int globalTimeStep)
{
    const double d = heat_laplace(shared_state, idx);
    if (SCHEME) {
        shared_state[idx + TIMES] = (R)(d + (double)shared_state[idx]);
    } else if (globalTimeStep % 2 == 0) {
        shared_state[idx + TIMES] = (R)(1.0 * d + (double)shared_state[idx - TIMES]);
    } else {
        shared_state[idx + TIMES] = (R)(0.5 * d + (double)shared_state[idx]);
    }
}

__device__ void step(double *shared_state, int idx, int globalTimeStep)
{
    heat_step(shared_state, idx, globalTimeStep);
}
