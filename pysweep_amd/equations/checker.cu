/*
 * Known-answer plugin: each cell becomes the mean of its four neighbours,
 * which flips a checkerboard every stage (reference:
 * pysweep/equations/checker.cu:10-32; summed x+1, x-1, y+1, y-1 as the
 * numpy half does).
 */
__device__ __constant__ double DX;
__device__ __constant__ double DY;
__device__ __constant__ double DT;
__device__ __constant__ int SCHEME;

__device__ void step(double *shared_state, int idx, int globalTimeStep)
{
    for (int v = 0; v < NV; ++v) {
        const int i = idx + v * VARS;
        double s = shared_state[i + SY];
        s = s + shared_state[i - SY];
        s = s + shared_state[i + 1];
        s = s + shared_state[i - 1];
        shared_state[i + TIMES] = s / 4;
    }
}
