/*
 * 2-D compressible Euler equations, conserved variables
 * (rho, rho*u, rho*v, E), finite volume on a periodic grid: minmod
 * pressure-ratio limited reconstruction, central flux plus spectral-radius
 * (Roe-averaged) dissipation, midpoint RK2, radius-2 cross stencil.
 *
 * Device half of the Euler plugin pair (euler.py).  The arithmetic follows
 * the numpy step of the reference plugin operation by operation
 * (pysweep/equations/euler.py:131-313: getFluxInX/Y, evaluateFluxInX/Y,
 * fluxLimiter, evaluateSpectral, getPressure), squares evaluated as a*a, so
 * device results agree bitwise with a numpy evaluation in that order.  The
 * scheme amplifies round-off strongly (the limiter zeroes ratios of
 * round-off-sized pressure differences), so a different operation order
 * would not stay within 1e-12 of the numpy path beyond a few stages.
 *
 * Contract: step(shared_state, idx, globalTimeStep) writes the four
 * variables at shared_state[idx + v*VARS + TIMES]; x neighbours idx +- SY,
 * y neighbours idx +- 1, level one back idx - TIMES.
 */
__device__ __constant__ double DX;
__device__ __constant__ double DY;
__device__ __constant__ double DT;
__device__ __constant__ double GAMMA;
__device__ __constant__ double GAM_M1;
__device__ __constant__ double DTDX;
__device__ __constant__ double DTDY;

struct eu_q {
    double r, m, n, e; /* rho, rho*u, rho*v, E */
};

/* state values are read into double whatever the storage type R */
template <class R>
__device__ __forceinline__ eu_q eu_at(const R *s, int i)
{
    eu_q q;
    q.r = s[i];
    q.m = s[i + VARS];
    q.n = s[i + 2 * VARS];
    q.e = s[i + 3 * VARS];
    return q;
}

/* gM1*(E - (m*m + n*n)/(2*rho))  (euler.py:284-293) */
__device__ __forceinline__ double eu_pressure(const eu_q &q)
{
    return GAM_M1 * (q.e - (q.m * q.m + q.n * q.n) / (2 * q.r));
}

/* q1 + min(Pr,1)/2*(q2-q1), Pr = num/den zeroed when nan, inf, >1e6 or <0
 * (euler.py:250-256) */
__device__ __forceinline__ eu_q eu_limit(const eu_q &q1, const eu_q &q2, double num, double den)
{
    double pr = num / den;
    if (isnan(pr) || isinf(pr) || pr > 1e6 || pr < 0) pr = 0.0;
    const double h = (1.0 < pr ? 1.0 : pr) / 2;
    eu_q o;
    o.r = q1.r + h * (q2.r - q1.r);
    o.m = q1.m + h * (q2.m - q1.m);
    o.n = q1.n + h * (q2.n - q1.n);
    o.e = q1.e + h * (q2.e - q1.e);
    return o;
}

/* central flux of a face, x direction (euler.py:181-197) or y (225-241) */
template <bool XDIR>
__device__ __forceinline__ eu_q eu_flux(const eu_q &L, const eu_q &R)
{
    const double pl = eu_pressure(L), pr = eu_pressure(R);
    eu_q f;
    if (XDIR) {
        f.r = L.m + R.m;
        f.m = L.m * L.m / L.r + pl + R.m * R.m / R.r + pr;
        f.n = L.m * L.n / L.r + R.m * R.n / R.r;
        f.e = (L.e + pl) * L.m / L.r + (R.e + pr) * R.m / R.r;
    } else {
        f.r = L.n + R.n;
        f.m = L.m * L.n / L.r + R.m * R.n / R.r;
        f.n = L.n * L.n / L.r + pl + R.n * R.n / R.r + pr;
        f.e = (L.e + pl) * L.n / L.r + (R.e + pr) * R.n / R.r;
    }
    return f;
}

/* spectral radius of the Roe average times (L - R); radius zeroed when nan
 * (euler.py:258-282) */
template <bool XDIR>
__device__ __forceinline__ eu_q eu_dissipation(const eu_q &L, const eu_q &R)
{
    const double sl = sqrt(L.r), sr = sqrt(R.r);
    const double den = 1 / (sl + sr);
    const double a0 = sl * sr;
    const double a1 = (sl * (L.m / L.r) + sr * (R.m / R.r)) * den;
    const double a2 = (sl * (L.n / L.r) + sr * (R.n / R.r)) * den;
    const double a3 = (sl * (L.e / L.r) + sr * (R.e / R.r)) * den;
    eu_q roe;
    roe.r = a0;
    roe.m = a0 * a1;
    roe.n = a0 * a2;
    roe.e = a0 * a3;
    double rad = sqrt(GAMMA * eu_pressure(roe) / a0) + fabs(XDIR ? a1 : a2);
    if (isnan(rad)) rad = 0.0;
    eu_q d;
    d.r = rad * (L.r - R.r);
    d.m = rad * (L.m - R.m);
    d.n = rad * (L.n - R.n);
    d.e = rad * (L.e - R.e);
    return d;
}

/* flux difference along one direction (euler.py:158-179, 202-223): the
 * west face's flux and dissipation added, the east face's subtracted, *0.5 */
template <bool XDIR, class S>
__device__ __forceinline__ eu_q eu_divergence(const S *s, int idx, int stride)
{
    const eu_q ww = eu_at(s, idx - 2 * stride), w = eu_at(s, idx - stride), c = eu_at(s, idx);
    const eu_q e = eu_at(s, idx + stride), ee = eu_at(s, idx + 2 * stride);
    const double pww = eu_pressure(ww), pw = eu_pressure(w), pc = eu_pressure(c);
    const double pe = eu_pressure(e), pee = eu_pressure(ee);
    eu_q acc;
    acc.r = acc.m = acc.n = acc.e = 0.0;

    eu_q L = eu_limit(w, c, pc - pw, pw - pww);
    eu_q R = eu_limit(c, w, pc - pw, pe - pc);
    eu_q f = eu_flux<XDIR>(L, R);
    acc.r = acc.r + f.r; acc.m = acc.m + f.m; acc.n = acc.n + f.n; acc.e = acc.e + f.e;
    f = eu_dissipation<XDIR>(L, R);
    acc.r = acc.r + f.r; acc.m = acc.m + f.m; acc.n = acc.n + f.n; acc.e = acc.e + f.e;

    L = eu_limit(c, e, pe - pc, pc - pw);
    R = eu_limit(e, c, pe - pc, pee - pe);
    f = eu_flux<XDIR>(L, R);
    acc.r = acc.r - f.r; acc.m = acc.m - f.m; acc.n = acc.n - f.n; acc.e = acc.e - f.e;
    f = eu_dissipation<XDIR>(L, R);
    acc.r = acc.r - f.r; acc.m = acc.m - f.m; acc.n = acc.n - f.n; acc.e = acc.e - f.e;

    acc.r = acc.r * 0.5; acc.m = acc.m * 0.5; acc.n = acc.n * 0.5; acc.e = acc.e * 0.5;
    return acc;
}

/* midpoint RK2 (euler.py:131-151): final stages (even globalTimeStep) start
 * from the level one back with the full step, intermediate stages from the
 * current level with half a step.  Arithmetic in double; R is the storage
 * type of the state. */
template <class R>
__device__ __forceinline__ void eu_step(R *shared_state, int idx, int globalTimeStep)
{
    const eu_q fx = eu_divergence<true>(shared_state, idx, SY);
    const eu_q fy = eu_divergence<false>(shared_state, idx, 1);
    R *out = shared_state + idx + TIMES;
    const eu_q b = eu_at(shared_state, globalTimeStep % 2 == 0 ? idx - TIMES : idx);
    const double h = globalTimeStep % 2 == 0 ? 1.0 : 0.5;
    if (globalTimeStep % 2 == 0) {
        out[0] = (R)(b.r + (fx.r * DTDX + fy.r * DTDY));
        out[VARS] = (R)(b.m + (fx.m * DTDX + fy.m * DTDY));
        out[2 * VARS] = (R)(b.n + (fx.n * DTDX + fy.n * DTDY));
        out[3 * VARS] = (R)(b.e + (fx.e * DTDX + fy.e * DTDY));
    } else {
        out[0] = (R)(b.r + h * (fx.r * DTDX + fy.r * DTDY));
        out[VARS] = (R)(b.m + h * (fx.m * DTDX + fy.m * DTDY));
        out[2 * VARS] = (R)(b.n + h * (fx.n * DTDX + fy.n * DTDY));
        out[3 * VARS] = (R)(b.e + h * (fx.e * DTDX + fy.e * DTDY));
    }
}

__device__ void step(double *shared_state, int idx, int globalTimeStep)
{
    eu_step(shared_state, idx, globalTimeStep);
}
