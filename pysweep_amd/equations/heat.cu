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

/*
 * Storage rounding of a stage value.  FP64 state keeps the value.  FP32 state
 * rounds it stochastically: a step changes u by about 8*pi^2*alpha*dt*u,
 * which at dt = 0.1*dx^2 on fine grids (8192^2: 1.2e-7*u) is about one FP32
 * ulp, so rounding to nearest would round every step's increment the same
 * way and the error would grow linearly with the steps.  Rounding up with
 * probability equal to the discarded fraction is unbiased, and the error
 * grows like a random walk instead.  The random bits are a 32-bit
 * multiply-xorshift hash (the "lowbias32" constants of C. Wellons' hash
 * prospector) of the FP64 value's two words and the stage, so
 * a given value rounds the same way in every kernel, batch and run: results
 * stay reproducible bit for bit, swept or standard, alone or batched.  The
 * 29 discarded mantissa bits plus a random 29-bit integer, truncated, give
 * the rounded float exactly (carries move into the exponent as they should);
 * values outside the normal float range round to nearest.
 */
__device__ __forceinline__ unsigned heat_mix(unsigned x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    return x ^ (x >> 16);
}

template <class R>
__device__ __forceinline__ R heat_store(double x, int globalTimeStep);

template <>
__device__ __forceinline__ double heat_store<double>(double x, int)
{
    return x;
}

template <>
__device__ __forceinline__ float heat_store<float>(double x, int globalTimeStep)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const int e = (int)((b >> 52) & 0x7ff);
    if (e < 1023 - 126 || e >= 1023 + 127)
        return (float)x;
    constexpr unsigned long long low = (1ULL << 29) - 1;
    const unsigned lo = (unsigned)b, hi = (unsigned)(b >> 32);
    const unsigned r = heat_mix(lo ^ __builtin_rotateleft32(hi, 13) ^ (0x9e3779b9U * (unsigned)(globalTimeStep + 1)));
    return (float)__longlong_as_double((long long)((b + (r & low)) & ~low));
}

template <class R>
__device__ __forceinline__ void heat_step(R *shared_state, int idx, int globalTimeStep)
{
    const double d = heat_laplace(shared_state, idx);
    if (SCHEME) {
        shared_state[idx + TIMES] = heat_store<R>(d + (double)shared_state[idx], globalTimeStep);
    } else if (globalTimeStep % 2 == 0) {
        shared_state[idx + TIMES] = heat_store<R>(1.0 * d + (double)shared_state[idx - TIMES], globalTimeStep);
    } else {
        shared_state[idx + TIMES] = heat_store<R>(0.5 * d + (double)shared_state[idx], globalTimeStep);
    }
}

__device__ void step(double *shared_state, int idx, int globalTimeStep)
{
    heat_step(shared_state, idx, globalTimeStep);
}

/*
 * Point update used by the band kernel (PS_HAS_POINT): one cell from its
 * level r-1 value c, x neighbours xm/xp, y neighbours ym/yp and its level
 * r-2 value prev, all widened to double; the arithmetic and its order are
 * heat_step's, so results are bitwise the same.  The scheme is a template
 * argument V (ps_point_variant() picks it once per launch: 1 forward Euler,
 * 0 RK2), so the kernel's levels carry no branch on it.
 */
#define PS_HAS_POINT 1
#define PS_POINT_VARIANTS 2

__device__ __forceinline__ int ps_point_variant() { return SCHEME ? 1 : 0; }

template <class R, int V>
__device__ __forceinline__ R ps_point(double c, double xm, double xp, double ym, double yp, double prev,
                                      int globalTimeStep)
{
    const double lx = COURANT * (xp - 2 * c + xm);
    const double ly = COURANT * (yp - 2 * c + ym);
    const double d = lx + ly;
    if (V == 1) return heat_store<R>(d + c, globalTimeStep);
    if (globalTimeStep % 2 == 0) return heat_store<R>(1.0 * d + prev, globalTimeStep);
    return heat_store<R>(0.5 * d + c, globalTimeStep);
}

/*
 * Level update used by the pysweep_amd kernels (PS_HAS_LEVEL): the cells of
 * [x0,x1) x [y0,y1) of one level, lane t taking cells t, t + PS_NT, ...
 * (row-major, y fastest).  cur is level r-1 with the tile strides, prev
 * level r-2 (the base of a final RK2 stage) indexed the same way (an LDS
 * slot, or an accessor of the ring), out level r with the tile strides (or
 * null); the arithmetic is heat_step's, so
 * results are bitwise the same.  Separate levels let the swept kernels keep
 * their LDS window circular (no copies).  Each new value is also handed to
 * emit(x, y, values), which writes it to HBM.
 */
#define PS_HAS_LEVEL 1
#define PS_LEVEL_SCRATCH(CELLS, FACES) 1
/* prev is read only at the cell being written, so out may be prev's slot */
#define PS_LEVEL_INPLACE 1

/* t / w for 0 <= t < 2^22, w >= 1, rw within one ulp of 1/w (the hardware
 * reciprocal): the float quotient is within one of t / w; one correction */
__device__ __forceinline__ int heat_div(int t, int w, float rw)
{
    int q = (int)((float)t * rw);
    const int r = t - q * w;
    if (r < 0)
        q -= 1;
    else if (r >= w)
        q += 1;
    return q;
}

template <class R, int MAXC, class Prev, class Emit>
__device__ __forceinline__ void ps_level(const R *cur, const Prev &prev, R *out, int x0, int x1, int y0, int y1,
                                         int globalTimeStep, double *, const Emit &emit)
{
    const bool fin = globalTimeStep % 2 == 0;
    const int w = y1 - y0, n = (x1 - x0) * w;
    // lane t starts at cell t = (i, j) and steps by PS_NT cells, i.e. by
    // (di, dj) with a carry, so only the first cell needs a division (a
    // compile-time width divides by the constant; otherwise the float
    // reciprocal)
    const int t0 = threadIdx.x;
    int i = __builtin_constant_p(w) ? t0 / w : heat_div(t0, w, __builtin_amdgcn_rcpf((float)w));
    int j = t0 - i * w;
    const int di = PS_NT / w, dj = PS_NT - di * w;
    for (int t = t0; t < n; t += PS_NT) {
        const int idx = (x0 + i) * SY + y0 + j;
        const double d = heat_laplace(cur, idx);
        R v[1];
        if (SCHEME)
            v[0] = heat_store<R>(d + (double)cur[idx], globalTimeStep);
        else if (fin)
            v[0] = heat_store<R>(1.0 * d + (double)prev[idx], globalTimeStep);
        else
            v[0] = heat_store<R>(0.5 * d + (double)cur[idx], globalTimeStep);
        if (out)
            out[idx] = v[0];
        emit(x0 + i, y0 + j, v);
        i += di;
        j += dj;
        if (j >= w) {
            j -= w;
            ++i;
        }
    }
}
