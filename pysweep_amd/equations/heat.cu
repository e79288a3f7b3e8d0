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

__device__ double heat_laplace(const double *u, int idx)
{
    const double c = u[idx];
    const double lx = COURANT * (u[idx + SY] - 2 * c + u[idx - SY]);
    const double ly = COURANT * (u[idx + 1] - 2 * c + u[idx - 1]);
    return lx + ly;
}

__device__ void step(double *shared_state, int idx, int globalTimeStep)
{
    const double d = heat_laplace(shared_state, idx);
    if (SCHEME) {
        shared_state[idx + TIMES] = d + shared_state[idx];
    } else if (globalTimeStep % 2 == 0) {
        shared_state[idx + TIMES] = 1.0 * d + shared_state[idx - TIMES];
    } else {
        shared_state[idx + TIMES] = 0.5 * d + shared_state[idx];
    }
}
