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

/*
 * Range of states for which the fast quotients and square roots below are
 * bitwise the IEEE ones: rho, E and the pressure in [2^-120, 2^120], each
 * momentum 0 or of magnitude in [2^-400, 2^120].  Every quotient of the
 * step then has operands at least 2^53 above the subnormal range and fewer
 * than 768 binary orders apart, and every denominator but the limiter's is
 * a positive normal number.  A state outside it sets EU_RANGE_FAULT (a
 * module global the host clears before a solve and reads after it: the
 * solve is then repeated with EU_EXACT_DIVISION, where every quotient and
 * square root is the compiler's IEEE `/` and `sqrt`).
 */
__device__ int EU_RANGE_FAULT;
#ifdef EU_EXACT_DIVISION
#define EU_EXACT 1
#else
#define EU_EXACT 0
#endif

__device__ __forceinline__ bool eu_pos_in_range(double v) { return v >= 0x1p-120 && v <= 0x1p120; }

__device__ __forceinline__ bool eu_mom_in_range(double v)
{
    const double a = fabs(v);
    return a <= 0x1p120 && (a >= 0x1p-400 || a == 0.0);
}

__device__ __forceinline__ bool eu_in_range(double r, double m, double n, double e, double p)
{
    return eu_pos_in_range(r) && eu_pos_in_range(e) && eu_pos_in_range(p) && eu_mom_in_range(m) &&
           eu_mom_in_range(n);
}

/* flag a state outside the range (a vector store; a benign race: every
 * writer writes 1) */
__device__ __forceinline__ void eu_range_check(double r, double m, double n, double e, double p)
{
    if (!eu_in_range(r, m, n, e, p))
        *(volatile int *)&EU_RANGE_FAULT = 1;
}

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

/* the same from an object indexed like a level slot (s[i]) */
template <class S>
__device__ __forceinline__ eu_q eu_at(const S &s, int i)
{
    eu_q q;
    q.r = s[i];
    q.m = s[i + VARS];
    q.n = s[i + 2 * VARS];
    q.e = s[i + 3 * VARS];
    return q;
}

/*
 * Division.  n/d is formed as the gfx950 compiler's IEEE double division
 * forms it -- reciprocal estimate, two Newton-Raphson refinements, product,
 * one residual correction, special-case fix-up -- less its operand-scaling
 * steps, which leave every value unchanged unless an operand lies within
 * 2^53 of the subnormal range, d is subnormal or the exponents differ by 768
 * or more (no Euler state comes near).  The quotient is therefore bitwise
 * the one `n / d` gives.  The refined reciprocal depends on d alone, so the
 * quotients of one face that share a denominator share it (6 reciprocals
 * serve the 19 quotients of a face).
 */
struct eu_rcp {
    double d, y; /* denominator, refined reciprocal */
};

__device__ __forceinline__ eu_rcp eu_recip(double d)
{
    double y = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-d, y, 1.0);
    y = __builtin_fma(y, e, y);
    eu_rcp q;
    q.d = d;
    q.y = y;
    return q;
}

/* the reciprocal of 2d from that of d: doubling d halves every step of
 * eu_recip exactly */
__device__ __forceinline__ eu_rcp eu_twice(const eu_rcp &q)
{
    eu_rcp h;
    h.d = 2 * q.d;
    h.y = q.y * 0.5;
    return h;
}

__device__ __forceinline__ double eu_quo(double n, const eu_rcp &q)
{
    if (EU_EXACT)
        return n / q.d;
    double z = n * q.y;
    const double rem = __builtin_fma(-q.d, z, n);
    z = __builtin_fma(rem, q.y, z);
    return __builtin_amdgcn_div_fixup(z, q.d, n);
}

/* n/d for a positive normal denominator and operands in the range above:
 * the special-case fix-up of eu_quo then changes nothing, and is left out */
__device__ __forceinline__ double eu_quo_pos(double n, const eu_rcp &q)
{
    if (EU_EXACT)
        return n / q.d;
    const double z = n * q.y;
    const double rem = __builtin_fma(-q.d, z, n);
    return __builtin_fma(rem, q.y, z);
}

/* sqrt(x) as the compiler's IEEE double square root forms it -- reciprocal
 * square-root estimate, Goldschmidt refinement, two residual corrections,
 * x itself for +-0 and +inf -- less the scaling of x < 2^-767, which changes
 * nothing for larger x (densities and Roe sound speeds never come near) */
__device__ __forceinline__ double eu_sqrt(double x)
{
    if (EU_EXACT)
        return sqrt(x);
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(g, r, g);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return (x == 0.0 || x == __builtin_inf()) ? x : g;
}

/* gM1*(E - (m*m + n*n)/(2*rho))  (euler.py:284-293); r2 is the reciprocal
 * of 2*rho */
__device__ __forceinline__ double eu_pressure(const eu_q &q, const eu_rcp &r2)
{
    return GAM_M1 * (q.e - eu_quo_pos(q.m * q.m + q.n * q.n, r2));
}

__device__ __forceinline__ double eu_pressure(const eu_q &q)
{
    return eu_pressure(q, eu_recip(2 * q.r));
}

/* q1 + min(Pr,1)/2*(q2-q1), Pr = num/den zeroed when nan, inf, >1e6 or <0
 * (euler.py:250-256) */
__device__ __forceinline__ eu_q eu_limit(const eu_q &q1, const eu_q &q2, double num, double den)
{
    double pr = eu_quo(num, eu_recip(den));
    if (isnan(pr) || isinf(pr) || pr > 1e6 || pr < 0) pr = 0.0;
    const double h = (1.0 < pr ? 1.0 : pr) / 2;
    eu_q o;
    o.r = q1.r + h * (q2.r - q1.r);
    o.m = q1.m + h * (q2.m - q1.m);
    o.n = q1.n + h * (q2.n - q1.n);
    o.e = q1.e + h * (q2.e - q1.e);
    return o;
}

/* central flux of a face, x direction (euler.py:181-197) or y (225-241);
 * rl, rr are the reciprocals of L.r and R.r */
template <bool XDIR>
__device__ __forceinline__ eu_q eu_flux(const eu_q &L, const eu_q &R, const eu_rcp &rl, const eu_rcp &rr)
{
    const double pl = eu_pressure(L, eu_twice(rl)), pr = eu_pressure(R, eu_twice(rr));
    eu_q f;
    if (XDIR) {
        f.r = L.m + R.m;
        f.m = eu_quo_pos(L.m * L.m, rl) + pl + eu_quo_pos(R.m * R.m, rr) + pr;
        f.n = eu_quo_pos(L.m * L.n, rl) + eu_quo_pos(R.m * R.n, rr);
        f.e = eu_quo_pos((L.e + pl) * L.m, rl) + eu_quo_pos((R.e + pr) * R.m, rr);
    } else {
        f.r = L.n + R.n;
        f.m = eu_quo_pos(L.m * L.n, rl) + eu_quo_pos(R.m * R.n, rr);
        f.n = eu_quo_pos(L.n * L.n, rl) + pl + eu_quo_pos(R.n * R.n, rr) + pr;
        f.e = eu_quo_pos((L.e + pl) * L.n, rl) + eu_quo_pos((R.e + pr) * R.n, rr);
    }
    return f;
}

/* spectral radius of the Roe average times (L - R); radius zeroed when nan
 * (euler.py:258-282) */
template <bool XDIR>
__device__ __forceinline__ eu_q eu_dissipation(const eu_q &L, const eu_q &R, const eu_rcp &rl, const eu_rcp &rr)
{
    const double sl = eu_sqrt(L.r), sr = eu_sqrt(R.r);
    const double den = eu_quo_pos(1.0, eu_recip(sl + sr));
    const double a0 = sl * sr;
    const double a1 = (sl * eu_quo_pos(L.m, rl) + sr * eu_quo_pos(R.m, rr)) * den;
    const double a2 = (sl * eu_quo_pos(L.n, rl) + sr * eu_quo_pos(R.n, rr)) * den;
    const double a3 = (sl * eu_quo_pos(L.e, rl) + sr * eu_quo_pos(R.e, rr)) * den;
    eu_q roe;
    roe.r = a0;
    roe.m = a0 * a1;
    roe.n = a0 * a2;
    roe.e = a0 * a3;
    const eu_rcp ra = eu_recip(a0);
    double rad = eu_sqrt(eu_quo_pos(GAMMA * eu_pressure(roe, eu_twice(ra)), ra)) + fabs(XDIR ? a1 : a2);
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
    eu_range_check(ww.r, ww.m, ww.n, ww.e, pww);
    eu_range_check(w.r, w.m, w.n, w.e, pw);
    eu_range_check(c.r, c.m, c.n, c.e, pc);
    eu_range_check(e.r, e.m, e.n, e.e, pe);
    eu_range_check(ee.r, ee.m, ee.n, ee.e, pee);
    eu_q acc;
    acc.r = acc.m = acc.n = acc.e = 0.0;

    eu_q L = eu_limit(w, c, pc - pw, pw - pww);
    eu_q R = eu_limit(c, w, pc - pw, pe - pc);
    eu_rcp rl = eu_recip(L.r), rr = eu_recip(R.r);
    eu_q f = eu_flux<XDIR>(L, R, rl, rr);
    acc.r = acc.r + f.r; acc.m = acc.m + f.m; acc.n = acc.n + f.n; acc.e = acc.e + f.e;
    f = eu_dissipation<XDIR>(L, R, rl, rr);
    acc.r = acc.r + f.r; acc.m = acc.m + f.m; acc.n = acc.n + f.n; acc.e = acc.e + f.e;

    L = eu_limit(c, e, pe - pc, pc - pw);
    R = eu_limit(e, c, pe - pc, pee - pe);
    rl = eu_recip(L.r);
    rr = eu_recip(R.r);
    f = eu_flux<XDIR>(L, R, rl, rr);
    acc.r = acc.r - f.r; acc.m = acc.m - f.m; acc.n = acc.n - f.n; acc.e = acc.e - f.e;
    f = eu_dissipation<XDIR>(L, R, rl, rr);
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

/*
 * Optional cooperative level update used by the pysweep_amd kernels when a
 * plugin defines PS_HAS_LEVEL: all lanes of a workgroup evaluate one level
 * of a rectangle of cells S = [x0,x1) x [y0,y1) together.  Each cell's
 * pressure and each face's limited states, flux and dissipation are
 * computed once (the per-cell step evaluates every face from both sides and
 * ten pressures per cell) and shared through LDS; the sums are formed in the
 * per-cell step's order from the same helper functions, so the results are
 * bitwise those of step().
 *
 *   cur    level r-1 in LDS, variable v of cell (x, y) at
 *          cur[v*VARS + x*SY + y];
 *   prev   level r-2 (the base of a final stage), indexed like cur
 *          (prev[v*VARS + x*SY + y]) and read only at the cells of S, by
 *          the lane computing the cell, before that lane writes it: an LDS
 *          slot in the swept kernels, an accessor of the ring in HBM in the
 *          standard kernel;
 *   out    level r in LDS with cur's strides (a slot of its own, so the
 *          caller may keep its window circular; it may be prev's slot), or
 *          null when only emit needs the values;
 *   sc     scratch of PS_LEVEL_SCRATCH(SX * SY, PS_MAXFACES) doubles,
 *          PS_MAXFACES being the most faces of one direction of any
 *          rectangle the harness passes ((x1-x0+1)*(y1-y0) or
 *          (x1-x0)*(y1-y0+1)); the face buffer is component-major so that
 *          consecutive lanes touch consecutive LDS words;
 *   MAXC   upper bound of cells per lane, ceil(|S| / PS_NT);
 *   emit   called as emit(x, y, v) with the tile coordinates and the NV new
 *          values of every cell of S, by the lane that computed it (the
 *          kernels write the cell to HBM from there, without re-reading
 *          LDS).
 */
#define PS_HAS_LEVEL 1
#define PS_LEVEL_SCRATCH(CELLS, FACES) ((CELLS) + 8 * (FACES))
/* prev is read only at the cell being written, by the lane writing it, so
 * out may be prev's own slot (a swept window of ITS levels suffices) */
#define PS_LEVEL_INPLACE 1

/* 1/w for w >= 1 within one ulp: the hardware reciprocal (one instruction
 * where a correctly rounded float division takes ten) */
__device__ __forceinline__ float eu_rcpf(int w) { return __builtin_amdgcn_rcpf((float)w); }

/* t / w for 0 <= t < 2^22, w >= 1, rw within one ulp of 1/w (the hardware
 * reciprocal, eu_rcpf): the float quotient is within one of the true one;
 * one correction makes it exact.  A width known at compile time (the
 * standard tile, unrolled swept levels) divides by the constant instead
 * (measured: Euler 1024^2 Octahedron 88.2 -> 86.0 us, Standard -3 %). */
__device__ __forceinline__ int eu_div(int t, int w, float rw)
{
    if (__builtin_constant_p(w))
        return t / w;
    int q = (int)((float)t * rw);
    const int r = t - q * w;
    if (r < 0)
        q -= 1;
    else if (r >= w)
        q += 1;
    return q;
}

/* exchange the two momentum components */
__device__ __forceinline__ void eu_swap(eu_q &q)
{
    const double t = q.m;
    q.m = q.n;
    q.n = t;
}

/* One face of a level update between the cells at idx - stride and idx:
 * flux in components 0-3 and dissipation in 4-7 of o (component stride
 * PS_MAXFACES), P holding the cells' pressures.  A y face (stride 1) is
 * evaluated as an x face with the momentum components exchanged in its
 * states and exchanged back in its results; the only operations that then
 * run in another order are commutative (m*n as n*m, m*m + n*n as
 * n*n + m*m), so the values are bitwise those of eu_flux<false> and
 * eu_dissipation<false>, and x and y faces share one instruction stream. */
template <class R>
__device__ __forceinline__ void eu_face(const R *cur, const double *P, double *o, int idx, int stride, bool ydir)
{
    constexpr int FS = PS_MAXFACES;
    eu_q wq = eu_at(cur, idx - stride), c = eu_at(cur, idx);
    const double pww = P[idx - 2 * stride], pw = P[idx - stride], pc = P[idx], pe = P[idx + stride];
    if (ydir) {
        eu_swap(wq);
        eu_swap(c);
    }
    const eu_q L = eu_limit(wq, c, pc - pw, pw - pww);
    const eu_q Rr = eu_limit(c, wq, pc - pw, pe - pc);
    const eu_rcp rl = eu_recip(L.r), rr = eu_recip(Rr.r);
    eu_q f = eu_flux<true>(L, Rr, rl, rr), d = eu_dissipation<true>(L, Rr, rl, rr);
    if (ydir) {
        eu_swap(f);
        eu_swap(d);
    }
    o[0] = f.r; o[FS] = f.m; o[2 * FS] = f.n; o[3 * FS] = f.e;
    o[4 * FS] = d.r; o[5 * FS] = d.m; o[6 * FS] = d.n; o[7 * FS] = d.e;
}

/* 0.5*((((0 + flux_a) + diss_a) - flux_b) - diss_b) per component, face a
 * on the low side of the cell and face b on the high side: the per-cell
 * step's order (eu_divergence) */
__device__ __forceinline__ eu_q eu_face_diff(const double *fa, const double *fb)
{
    constexpr int FS = PS_MAXFACES;
    eu_q a;
    a.r = 0.0; a.m = 0.0; a.n = 0.0; a.e = 0.0;
    a.r = a.r + fa[0]; a.m = a.m + fa[FS]; a.n = a.n + fa[2 * FS]; a.e = a.e + fa[3 * FS];
    a.r = a.r + fa[4 * FS]; a.m = a.m + fa[5 * FS]; a.n = a.n + fa[6 * FS]; a.e = a.e + fa[7 * FS];
    a.r = a.r - fb[0]; a.m = a.m - fb[FS]; a.n = a.n - fb[2 * FS]; a.e = a.e - fb[3 * FS];
    a.r = a.r - fb[4 * FS]; a.m = a.m - fb[5 * FS]; a.n = a.n - fb[6 * FS]; a.e = a.e - fb[7 * FS];
    a.r = a.r * 0.5; a.m = a.m * 0.5; a.n = a.n * 0.5; a.e = a.e * 0.5;
    return a;
}

template <class R, int MAXC, class Prev, class Emit>
__device__ __forceinline__ void ps_level(const R *cur, const Prev &prev, R *out, int x0, int x1, int y0, int y1,
                                         int globalTimeStep, double *sc, const Emit &emit)
{
    const int h = x1 - x0, w = y1 - y0;
    const float rw = eu_rcpf(w);
    constexpr int FS = PS_MAXFACES;  /* component stride of the face buffer */
    double *P = sc;                  /* pressures, tile-indexed x*SY + y */
    double *F = sc + SX * SY;        /* face f, component c at F[c*FS + f]: flux 0-3, dissipation 4-7 */

    /* 1. pressures on the cross-shaped footprint */
    {
        const int bh = h + 4, bw = w + 4, n = bh * bw;
        const float rbw = eu_rcpf(bw);
        for (int t = threadIdx.x; t < n; t += PS_NT) {
            const int i = eu_div(t, bw, rbw), j = t - i * bw;
            const bool inx = i >= 2 && i < h + 2, iny = j >= 2 && j < w + 2;
            if (!(inx || iny)) continue;
            const int x = x0 - 2 + i, y = y0 - 2 + j;
            const eu_q q = eu_at(cur, x * SY + y);
            const double p = eu_pressure(q);
            P[x * SY + y] = p;
            eu_range_check(q.r, q.m, q.n, q.e, p);
        }
    }
    __syncthreads();

    /* 2. faces.  x face i (0..h) lies west of cell x0+i, at F[i*w + j]; y
     * face j (0..w) lies south of cell y0+j, at Fy[i*(w+1) + j].  When both
     * sets fit the buffer they are evaluated in one pass over one lane space
     * (the small levels then fill whole wavefronts, and the level needs two
     * barriers fewer); otherwise the x divergences are taken first and the
     * y faces overwrite the x faces. */
    const int nfx = (h + 1) * w, nfy = h * (w + 1), wf = w + 1;
    const float rwf = eu_rcpf(wf);
    const bool merged = nfx + nfy <= FS;
    double *Fy = merged ? F + nfx : F;
    {
        const int n = merged ? nfx + nfy : nfx;
        for (int t = threadIdx.x; t < n; t += PS_NT) {
            const bool xf = t < nfx;
            const int u = xf ? t : t - nfx, wd = xf ? w : wf;
            const int i = eu_div(u, wd, xf ? rw : rwf), j = u - i * wd;
            eu_face(cur, P, F + t, (x0 + i) * SY + y0 + j, xf ? SY : 1, !xf);
        }
    }
    __syncthreads();

    eu_q dfx[MAXC];
    if (!merged) {
        /* 3. x divergence of my cells (cell (i, j): west face i, east face i+1) */
#pragma unroll
        for (int m = 0; m < MAXC; ++m) {
            const int t = threadIdx.x + m * PS_NT;
            if (t < h * w) {
                const int i = eu_div(t, w, rw), j = t - i * w;
                dfx[m] = eu_face_diff(F + (i * w + j), F + ((i + 1) * w + j));
            }
        }
        __syncthreads();
        /* 4. y faces over the x faces */
        for (int t = threadIdx.x; t < nfy; t += PS_NT) {
            const int i = eu_div(t, wf, rwf), j = t - i * wf;
            eu_face(cur, P, Fy + t, (x0 + i) * SY + y0 + j, 1, true);
        }
        __syncthreads();
    }

    /* 5. (x and) y divergence and the RK2 update of my cells */
    const bool fin = globalTimeStep % 2 == 0;
    const double hh = fin ? 1.0 : 0.5;
#pragma unroll
    for (int m = 0; m < MAXC; ++m) {
        const int t = threadIdx.x + m * PS_NT;
        if (t < h * w) {
            const int i = eu_div(t, w, rw), j = t - i * w;
            const eu_q fx = merged ? eu_face_diff(F + (i * w + j), F + ((i + 1) * w + j)) : dfx[m];
            const double *fs = Fy + (i * wf + j);
            const eu_q a = eu_face_diff(fs, fs + 1);
            const int idx = (x0 + i) * SY + (y0 + j);
            const eu_q b = fin ? eu_at(prev, idx) : eu_at(cur, idx);
            R v[4];
            if (fin) {
                v[0] = (R)(b.r + (fx.r * DTDX + a.r * DTDY));
                v[1] = (R)(b.m + (fx.m * DTDX + a.m * DTDY));
                v[2] = (R)(b.n + (fx.n * DTDX + a.n * DTDY));
                v[3] = (R)(b.e + (fx.e * DTDX + a.e * DTDY));
            } else {
                v[0] = (R)(b.r + hh * (fx.r * DTDX + a.r * DTDY));
                v[1] = (R)(b.m + hh * (fx.m * DTDX + a.m * DTDY));
                v[2] = (R)(b.n + hh * (fx.n * DTDX + a.n * DTDY));
                v[3] = (R)(b.e + hh * (fx.e * DTDX + a.e * DTDY));
            }
            if (out) {
                R *o = out + idx;
                o[0] = v[0];
                o[VARS] = v[1];
                o[2 * VARS] = v[2];
                o[3 * VARS] = v[3];
            }
            emit(x0 + i, y0 + j, v);
        }
    }
}

#ifdef PS_BAND_COMMON
/*
 * Band sweep (PS_HAS_BAND): the swept solve's kernel with every level of a
 * launch in registers.  One wavefront owns a band of 64 columns, one column
 * per lane, and sweeps it along x; a launch of PS_BAND_H stages reads level
 * A over the band and writes level A + PS_BAND_H over its middle
 * 64 - 4 * PS_BAND_H columns (the radius-2 stencil costs two halo columns
 * per level on each side, recomputed by the neighbouring bands instead of
 * exchanged).
 *
 * Step tau brings in row xs + tau of level A (xs = x0 - 2 * hh) and, for
 * l = 1..hh in order, computes row xs + tau - 2 * l of level A + l.  Each
 * level keeps the last five rows it produced (states, in the storage type)
 * and the last four rows' pressures: the x face between rows r and r + 1
 * needs the pressures of rows r - 1 .. r + 2 and the states of r and r + 1,
 * and a final RK2 stage needs its cell two levels back, four rows behind.
 * The x face below a row is the one above the row before, carried from the
 * previous step; the y face east of a column is computed by its lane from
 * the neighbouring lanes' states and pressures (DPP), and the west one is
 * the lower lane's east face.  Every face, pressure and update is evaluated
 * by the functions and in the order of ps_level and the per-cell step, so
 * results are bitwise theirs.  Rows and columns that a level cannot yet
 * compute correctly (warm-up steps, the outermost lanes) hold garbage that
 * no valid cell ever reads; only valid cells are range-checked and stored.
 * Launches start on full steps (the first stage of a launch never reads
 * the level two back).
 */
#define PS_HAS_BAND 1
#define EU_BAND_WOUT (64 - 4 * PS_BAND_H)
static_assert(EU_BAND_WOUT >= 8, "PS_BAND_H too deep for a 64-column band");

/* face between cells a (low) and b (high) along x (XDIR) or y: pww, pw, pc
 * and pe are the pressures of the cell below a, a, b and the cell above b;
 * flux and dissipation as in eu_face */
template <bool XDIR>
__device__ __forceinline__ void eu_face_vals(const eu_q &a, const eu_q &b, double pww, double pw, double pc,
                                             double pe, eu_q &f, eu_q &d)
{
    const eu_q L = eu_limit(a, b, pc - pw, pw - pww);
    const eu_q Rr = eu_limit(b, a, pc - pw, pe - pc);
    const eu_rcp rl = eu_recip(L.r), rr = eu_recip(Rr.r);
    f = eu_flux<XDIR>(L, Rr, rl, rr);
    d = eu_dissipation<XDIR>(L, Rr, rl, rr);
}

/* eu_face_diff on values: 0.5*((((0 + fa) + da) - fb) - db) */
__device__ __forceinline__ eu_q eu_band_diff(const eu_q &fa, const eu_q &da, const eu_q &fb, const eu_q &db)
{
    eu_q a;
    a.r = 0.0; a.m = 0.0; a.n = 0.0; a.e = 0.0;
    a.r = a.r + fa.r; a.m = a.m + fa.m; a.n = a.n + fa.n; a.e = a.e + fa.e;
    a.r = a.r + da.r; a.m = a.m + da.m; a.n = a.n + da.n; a.e = a.e + da.e;
    a.r = a.r - fb.r; a.m = a.m - fb.m; a.n = a.n - fb.n; a.e = a.e - fb.e;
    a.r = a.r - db.r; a.m = a.m - db.m; a.n = a.n - db.n; a.e = a.e - db.e;
    a.r = a.r * 0.5; a.m = a.m * 0.5; a.n = a.n * 0.5; a.e = a.e * 0.5;
    return a;
}

__device__ __forceinline__ eu_q eu_lower(const eu_q &q)
{
    eu_q o;
    o.r = ps_from_lower(q.r); o.m = ps_from_lower(q.m); o.n = ps_from_lower(q.n); o.e = ps_from_lower(q.e);
    return o;
}

__device__ __forceinline__ eu_q eu_upper(const eu_q &q)
{
    eu_q o;
    o.r = ps_from_upper(q.r); o.m = ps_from_upper(q.m); o.n = ps_from_upper(q.n); o.e = ps_from_upper(q.e);
    return o;
}

template <class R>
struct eu_band_row {
    R v[4];
    __device__ __forceinline__ eu_q q() const
    {
        eu_q o;
        o.r = (double)v[0]; o.m = (double)v[1]; o.n = (double)v[2]; o.e = (double)v[3];
        return o;
    }
};

/* The last three rows a level produced (slot tau % 3 holds the newest, so
 * nothing moves between steps), their pressures, and each row's pressure
 * difference to the row before it: the x face (r | r+1) takes
 * p(r+1) - p(r), p(r) - p(r-1) and p(r+2) - p(r+1) as its limiter ratios'
 * numerator and denominators. */
template <class R>
struct eu_band_win {
    eu_band_row<R> q[3];
    double p[3], dp[3];
};

/* (the range check of a valid cell only marks `bad`: the fault flag is
 * written once per wavefront, after the sweep, so that the steps hold no
 * branch) */
template <int PH, class R>
__device__ __forceinline__ void eu_band_push(eu_band_win<R> &w, const eu_band_row<R> &row, bool check, bool &bad)
{
    const eu_q q = row.q();
    const double p = eu_pressure(q);
    w.q[PH] = row;
    w.dp[PH] = p - w.p[(PH + 2) % 3];
    w.p[PH] = p;
    bad |= check & !eu_in_range(q.r, q.m, q.n, q.e, p);
}

/* Row r of level l from level l-1's window at step phase PH (its newest
 * row, slot PH, is r + 2): the x face (r | r+1), the y faces of row r, the
 * divergences with the x face below (xf, xd, replaced by the new one) and
 * the RK2 update from base (the cell itself, or for a final stage the cell
 * two levels back). */
template <int PH, bool FINAL, class R>
__device__ __forceinline__ eu_band_row<R> eu_band_cell(const eu_band_win<R> &w, eu_q &xf, eu_q &xd,
                                                       const eu_q &base2)
{
    constexpr int pa = (PH + 1) % 3, pb = (PH + 2) % 3;  /* rows r, r + 1 */
    const eu_q a = w.q[pa].q(), b = w.q[pb].q();
    eu_q f, d;
    {
        const eu_q L = eu_limit(a, b, w.dp[pb], w.dp[pa]);
        const eu_q Rr = eu_limit(b, a, w.dp[pb], w.dp[PH]);
        const eu_rcp rl = eu_recip(L.r), rr = eu_recip(Rr.r);
        f = eu_flux<true>(L, Rr, rl, rr);
        d = eu_dissipation<true>(L, Rr, rl, rr);
    }
    const double pr = w.p[pa], pu = ps_from_upper(pr);
    eu_q fe, de;
    eu_face_vals<false>(a, eu_upper(a), ps_from_lower(pr), pr, pu, ps_from_upper(pu), fe, de);
    const eu_q fx = eu_band_diff(xf, xd, f, d);
    const eu_q fy = eu_band_diff(eu_lower(fe), eu_lower(de), fe, de);
    xf = f;
    xd = d;
    const eu_q bs = FINAL ? base2 : a;
    eu_band_row<R> o;
    if (FINAL) {
        o.v[0] = (R)(bs.r + (fx.r * DTDX + fy.r * DTDY));
        o.v[1] = (R)(bs.m + (fx.m * DTDX + fy.m * DTDY));
        o.v[2] = (R)(bs.n + (fx.n * DTDX + fy.n * DTDY));
        o.v[3] = (R)(bs.e + (fx.e * DTDX + fy.e * DTDY));
    } else {
        o.v[0] = (R)(bs.r + 0.5 * (fx.r * DTDX + fy.r * DTDY));
        o.v[1] = (R)(bs.m + 0.5 * (fx.m * DTDX + fy.m * DTDY));
        o.v[2] = (R)(bs.n + 0.5 * (fx.n * DTDX + fy.n * DTDY));
        o.v[3] = (R)(bs.e + 0.5 * (fx.e * DTDX + fy.e * DTDY));
    }
    return o;
}

#define EU_GLOBAL __attribute__((address_space(1)))

template <class R>
struct eu_band_ctx {
    /* the ring in the global address space: its accesses compile to global
     * loads and stores also where the pointer was read from memory (an
     * instance table), whose address space the compiler cannot infer */
    EU_GLOBAL R *state;
    const EU_GLOBAL R *in;  /* level A */
    size_t cells;
    int s0, T, nx, ny, wlo, hh, x0, rows, xs, col, oc, lane;
    bool out_col;
};

template <class R>
__device__ __forceinline__ void eu_band_load(const eu_band_ctx<R> &c, int row, eu_band_row<R> &o)
{
    const EU_GLOBAL R *p = c.in + (size_t)row * c.ny + c.col;
#pragma unroll
    for (int v = 0; v < 4; ++v) o.v[v] = p[v * c.cells];
}

template <class R>
__device__ __forceinline__ void eu_band_store(const eu_band_ctx<R> &c, int l, int row, const eu_band_row<R> &o)
{
    if (l >= c.wlo && c.out_col && row >= c.x0 && row < c.x0 + c.rows) {
        const int slot = c.s0 + l >= c.T ? c.s0 + l - c.T : c.s0 + l;
        EU_GLOBAL R *dst = c.state + (size_t)slot * NV * c.cells + (size_t)row * c.ny + c.oc;
#pragma unroll
        for (int v = 0; v < 4; ++v) dst[v * c.cells] = o.v[v];
    }
}

/* wrap a row index into [0, nx) (rows stay within one period of the grid
 * around the segment only when nx exceeds the halo, so reduce fully) */
__device__ __forceinline__ int eu_band_wrap(int r, int nx)
{
    r %= nx;
    return r < 0 ? r + nx : r;
}

/* One step (phase PH = tau % 3) of a one-step (two-stage) launch.  FULL:
 * a step of the steady state of a two-level launch that writes level 2
 * only (both levels valid; no test until the final store), so that the
 * compiler may interleave the two levels' independent work. */
template <int PH, bool FULL, class R>
__device__ __forceinline__ void eu_band_step(const eu_band_ctx<R> &c, int tau, eu_band_win<R> &w0,
                                             eu_band_win<R> &w1, eu_q (&xf)[2], eu_q (&xd)[2],
                                             eu_band_row<R> (&pf)[3], int &lr, bool &bad)
{
    /* level A: this step's row (loaded three steps ago) */
    const eu_band_row<R> in = pf[PH];
    eu_band_load(c, lr, pf[PH]);
    lr = lr + 1 == c.nx ? 0 : lr + 1;
    /* the base of the final stage: level A two rows behind level 1's row,
     * re-read (this wavefront loaded it four steps ago) */
    eu_band_row<R> b2;
    const int r2 = c.xs + tau - 4;
    const bool two = FULL || (c.hh >= 2 && tau >= 7);
    if (two) eu_band_load(c, eu_band_wrap(r2, c.nx), b2);
    eu_band_push<PH>(w0, in, true, bad);
    /* level 1 (a half stage), row xs + tau - 2 */
    if (FULL || tau >= 3) {
        const eu_band_row<R> v1 = eu_band_cell<PH, false>(w0, xf[0], xd[0], eu_q());
        const bool ok = (FULL || tau >= 4) && c.lane >= 2 && c.lane < 62;
        eu_band_push<PH>(w1, v1, ok, bad);
        if (!FULL && ok) eu_band_store(c, 1, c.xs + tau - 2, v1);
    }
    /* level 2 (the final stage), row xs + tau - 4 */
    if (two) {
        const eu_band_row<R> v2 = eu_band_cell<PH, true>(w1, xf[1], xd[1], b2.q());
        eu_band_store(c, 2, r2, v2);
    }
}

template <class R>
__device__ __forceinline__ void ps_band_sweep(R *state, int gts0, int A, const ps_band_grid &g, int band, int seg)
{
    static_assert(PS_BAND_H == 2, "the Euler band sweep takes one RK2 step (two stages) per launch");
    constexpr int WOUT = EU_BAND_WOUT;
    const int lane = threadIdx.x & 63;
    (void)gts0;  /* launches start on full steps: level 1 is a half stage, level 2 a final one */
    eu_band_ctx<R> c;
    c.state = (EU_GLOBAL R *)state;  /* device memory */
    c.cells = (size_t)g.nx * g.ny;
    c.s0 = A % g.T;
    c.T = g.T;
    c.nx = g.nx;
    c.ny = g.ny;
    c.in = c.state + (size_t)c.s0 * NV * c.cells;
    c.hh = g.hh;
    c.wlo = g.wlo;
    c.x0 = seg * g.rows;
    c.rows = min(g.rows, g.nx - c.x0);
    c.xs = c.x0 - 2 * g.hh;
    c.lane = lane;
    const int o = lane - 2 * PS_BAND_H;  /* band-local output column */
    c.oc = band * WOUT + o;
    c.col = eu_band_wrap(c.oc, g.ny);
    c.out_col = o >= 0 && o < WOUT && c.oc < g.ny;
    /* defined values in what the warm-up steps read before any level has
     * produced it (never used by a valid cell) */
    eu_band_win<R> w0, w1;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
#pragma unroll
        for (int v = 0; v < 4; ++v) w0.q[k].v[v] = w1.q[k].v[v] = (R)0;
        w0.p[k] = w1.p[k] = w0.dp[k] = w1.dp[k] = 0.0;
    }
    eu_q xf[2], xd[2];
#pragma unroll
    for (int l = 0; l < 2; ++l) {
        xf[l].r = xf[l].m = xf[l].n = xf[l].e = 0.0;
        xd[l] = xf[l];
    }
    eu_band_row<R> pf[3];
    int lr = eu_band_wrap(c.xs, g.nx);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        eu_band_load(c, lr, pf[k]);
        lr = lr + 1 == g.nx ? 0 : lr + 1;
    }
    const int steps = c.rows + 4 * g.hh;
    bool bad = false;
    int tau = 0;
    if (g.hh == 2 && g.wlo == 2) {
        /* warm-up: level 2 starts at step 7, valid cells of level 1 at 4 */
        for (; tau < 9; tau += 3) {
            eu_band_step<0, false>(c, tau, w0, w1, xf, xd, pf, lr, bad);
            eu_band_step<1, false>(c, tau + 1, w0, w1, xf, xd, pf, lr, bad);
            eu_band_step<2, false>(c, tau + 2, w0, w1, xf, xd, pf, lr, bad);
        }
        for (; tau < steps; tau += 3) {
            eu_band_step<0, true>(c, tau, w0, w1, xf, xd, pf, lr, bad);
            eu_band_step<1, true>(c, tau + 1, w0, w1, xf, xd, pf, lr, bad);
            eu_band_step<2, true>(c, tau + 2, w0, w1, xf, xd, pf, lr, bad);
        }
    } else {
        for (; tau < steps; tau += 3) {
            eu_band_step<0, false>(c, tau, w0, w1, xf, xd, pf, lr, bad);
            eu_band_step<1, false>(c, tau + 1, w0, w1, xf, xd, pf, lr, bad);
            eu_band_step<2, false>(c, tau + 2, w0, w1, xf, xd, pf, lr, bad);
        }
    }
    /* a vector store (a benign race: every writer writes 1) */
    if (bad) *(volatile int *)&EU_RANGE_FAULT = 1;
}
#endif /* PS_BAND_COMMON */
