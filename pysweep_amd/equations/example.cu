/*
 * Known-answer plugin: every stage adds 1 (SCHEME=1), or an RK2-shaped
 * update adding 1 on intermediate and 2 (from the level one back) on final
 * stages (SCHEME=0).  Same behaviour as the reference's example plugin
 * (pysweep/equations/example.cu:10-23).
 */
__device__ __constant__ double DX;
__device__ __constant__ double DY;
__device__ __constant__ double DT;
__device__ __constant__ int SCHEME;

__device__ void step(double *shared_state, int idx, int globalTimeStep)
{
    for (int v = 0; v < NV; ++v) {
        const int i = idx + v * VARS;
        if (SCHEME) {
            shared_state[i + TIMES] = shared_state[i] + 1;
        } else if (globalTimeStep % ITS == 0) {
            shared_state[i + TIMES] = shared_state[i - TIMES] + 2;
        } else {
            shared_state[i + TIMES] = shared_state[i] + 1;
        }
    }
}
