/*
 * hzl_oracle.c -- FP64 CPU oracle for the batched vorticity ensemble engine.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path never calls it.
 *
 * It restates, in plain C99 with double precision, the reference algorithms of
 * pvasbob/HZL-stochasticFlowInference for the solver path:
 *   - full-complex N x N radix-2 FFT, forward scaled by 1/N^2   (src/fft.cpp:9-69)
 *   - wave-number map and strict 2/3 square mask                 (src/fft.cpp:70-72,
 *                                                                  src/solver.cpp:126-135)
 *   - velocity, dealiased Jacobian RHS, classical RK4 with the
 *     forcing held at its midpoint value through the four stages (src/solver.cpp:177-254)
 *   - diagnostics (energy, enstrophy, dissipation, injection, CFL)(src/solver.cpp:381-397)
 *   - energy spectrum                                             (src/solver.cpp:426-433)
 *   - SplitMix stream_seed, mt19937_64 + Box-Muller NormalStream  (src/solver.cpp:41-66)
 *   - bootstrap-PF weighting, log-sum-exp, ESS, systematic
 *     resampling                                                  (src/filter.cpp:38-126)
 *   - serial perturbed-observation EnKF scalar update             (src/filter.cpp:265-321)
 *   - SSPRK(3,3) step, tangent and adjoint drifts                  (src/solver.cpp:235-246,
 *                                                                  264-323)
 *   - OU shell forcing: pair order, stationary start, exact
 *     transition, forcing advanced by dt/2 around each step       (src/solver.cpp:67-119,
 *                                                                  120-139,174-176,217-254)
 * plus the engine's own input definitions that have no reference counterpart
 * (Philox4x32-10 counter RNG, shell white noise, Kolmogorov forcing, Gaussian
 * random-field initial condition), written from their published definitions.
 *
 * Complex arrays are interleaved doubles (re, im), row-major index y*n + x, the
 * reference's Field layout (include/hzl/fft.hpp:7-9).
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   fft2d      -- brute-force DFT, sin(3x) -> -0.5i known answer, round trip,
 *                 reference build (oracle/_ref) bit-level agreement
 *   rhs/step   -- triad brute force, single-shell closed form RK4 response,
 *                 Taylor-Green decay, inviscid invariants, reference build
 *   diagnostics-- Taylor-Green energy pi^2 e^{-4t}, spectrum sum, reference build
 *   mt19937_64 -- 10000th output of the default seed (C++ standard KAT)
 *   philox     -- Random123 known-answer vectors
 *   PF pieces  -- reference bootstrap filter in the deterministic limit,
 *                 long-double LSE, brute-force systematic resampling
 *   EnKF scalar-- dense closed form on tiny ensembles, reference deterministic limit
 *   grf        -- enstrophy 1, Hermitian symmetry, spectrum slope (statistical)
 *   ssprk3     -- single-shell closed form (stability polynomial, fixed point),
 *                 reference build
 *   tangent    -- central difference of the (quadratic) drift, reference build
 *   adjoint    -- adjoint identity <lambda, L d> = <L* lambda, d>, reference build
 *   OU forcing -- reference build (Solver with sigma > 0: trajectory and forcing),
 *                 stationary variance sigma^2 tau / 4, one-step autocorrelation
 *                 e^{-dt/tau}, pair set = forcing_pairs
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

typedef struct {
    double re, im;
} cpx;

static inline cpx cmul(cpx a, cpx b) {
    cpx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}

int orc_wave_number(long index, long n) {
    return index < n / 2 ? (int)index : (int)(index - n);
}

static int is_pow2(long n) {
    return n >= 8 && n <= 4096 && (n & (n - 1)) == 0;
}

/* ---------------------------------------------------------------- FFT ---- */
typedef struct {
    long n;
    long *rev;
    cpx *roots;
    cpx *line;
} fft_plan;

static int plan_make(fft_plan *p, long n) {
    if (!is_pow2(n))
        return -1;
    p->n = n;
    p->rev = (long *)malloc(sizeof(long) * (size_t)n);
    p->roots = (cpx *)malloc(sizeof(cpx) * (size_t)(n / 2));
    p->line = (cpx *)malloc(sizeof(cpx) * (size_t)n);
    if (!p->rev || !p->roots || !p->line)
        return -1;
    int bits = 0;
    while ((1L << bits) < n)
        ++bits;
    for (long i = 0; i < n; ++i) {
        long r = 0, x = i;
        for (int b = 0; b < bits; ++b) {
            r = (r << 1) | (x & 1);
            x >>= 1;
        }
        p->rev[i] = r;
    }
    /* Twiddles exp(-2 pi i j / n), evaluated once in double. */
    for (long j = 0; j < n / 2; ++j) {
        const double a = -2.0 * ORC_PI * (double)j / (double)n;
        p->roots[j].re = cos(a);
        p->roots[j].im = sin(a);
    }
    return 0;
}

static void plan_free(fft_plan *p) {
    free(p->rev);
    free(p->roots);
    free(p->line);
}

/* In-place iterative radix-2 decimation-in-time transform of p->line. */
static void fft_line(fft_plan *p, int inverse) {
    const long n = p->n;
    cpx *a = p->line;
    for (long i = 0; i < n; ++i) {
        const long r = p->rev[i];
        if (i < r) {
            cpx t = a[i];
            a[i] = a[r];
            a[r] = t;
        }
    }
    for (long len = 2; len <= n; len <<= 1) {
        const long half = len / 2, stride = n / len;
        for (long s = 0; s < n; s += len)
            for (long j = 0; j < half; ++j) {
                cpx w = p->roots[j * stride];
                if (inverse)
                    w.im = -w.im;
                const cpx u = a[s + j];
                const cpx v = cmul(w, a[s + j + half]);
                a[s + j].re = u.re + v.re;
                a[s + j].im = u.im + v.im;
                a[s + j + half].re = u.re - v.re;
                a[s + j + half].im = u.im - v.im;
            }
    }
}

static void fft2d_plan(fft_plan *p, cpx *d, int inverse) {
    const long n = p->n;
    for (long y = 0; y < n; ++y) {
        memcpy(p->line, d + y * n, sizeof(cpx) * (size_t)n);
        fft_line(p, inverse);
        memcpy(d + y * n, p->line, sizeof(cpx) * (size_t)n);
    }
    for (long x = 0; x < n; ++x) {
        for (long y = 0; y < n; ++y)
            p->line[y] = d[y * n + x];
        fft_line(p, inverse);
        for (long y = 0; y < n; ++y)
            d[y * n + x] = p->line[y];
    }
    if (!inverse) {
        const double s = (double)(n * n);
        for (long i = 0; i < n * n; ++i) {
            d[i].re /= s;
            d[i].im /= s;
        }
    }
}

/* Reference Fourier2D::forward / inverse (src/fft.cpp:44-69). Returns -1 on bad n. */
int orc_fft2d(long n, double *data, int inverse) {
    fft_plan p;
    if (plan_make(&p, n)) {
        plan_free(&p);
        return -1;
    }
    fft2d_plan(&p, (cpx *)data, inverse);
    plan_free(&p);
    return 0;
}

/* ------------------------------------------------------------- solver ---- */
/* Retained iff 3|kx| < n and 3|ky| < n and k != 0 (src/solver.cpp:126-135). */
static int active_index(long i, long n) {
    const int a = orc_wave_number(i % n, n), b = orc_wave_number(i / n, n);
    return 3 * abs(a) < n && 3 * abs(b) < n && i != 0;
}

void orc_active_mask(long n, unsigned char *mask) {
    for (long i = 0; i < n * n; ++i)
        mask[i] = (unsigned char)active_index(i, n);
}

typedef struct {
    long n;
    fft_plan plan;
    double *kx, *ky, *ksq;
    unsigned char *active;
    cpx *u, *v, *wx, *wy, *nl, *tmp, *k1, *k2, *k3, *k4;
} orc_solver;

static int solver_make(orc_solver *s, long n) {
    memset(s, 0, sizeof(*s));
    if (plan_make(&s->plan, n))
        return -1;
    s->n = n;
    const size_t c = (size_t)(n * n);
    s->kx = (double *)malloc(sizeof(double) * c);
    s->ky = (double *)malloc(sizeof(double) * c);
    s->ksq = (double *)malloc(sizeof(double) * c);
    s->active = (unsigned char *)malloc(c);
    cpx **f[] = {&s->u, &s->v, &s->wx, &s->wy, &s->nl, &s->tmp, &s->k1, &s->k2, &s->k3, &s->k4};
    for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); ++i) {
        *f[i] = (cpx *)calloc(c, sizeof(cpx));
        if (!*f[i])
            return -1;
    }
    for (long i = 0; i < n * n; ++i) {
        const int a = orc_wave_number(i % n, n), b = orc_wave_number(i / n, n);
        s->kx[i] = a;
        s->ky[i] = b;
        s->ksq[i] = (double)(a * a + b * b);
        s->active[i] = (unsigned char)active_index(i, n);
    }
    return 0;
}

static void solver_free(orc_solver *s) {
    plan_free(&s->plan);
    free(s->kx);
    free(s->ky);
    free(s->ksq);
    free(s->active);
    cpx *f[] = {s->u, s->v, s->wx, s->wy, s->nl, s->tmp, s->k1, s->k2, s->k3, s->k4};
    for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); ++i)
        free(f[i]);
}

/* u = i ky psi, v = -i kx psi, psi = w / |k|^2, then inverse FFT (src/solver.cpp:177-185). */
static void solver_velocity(orc_solver *s, const cpx *w) {
    const long c = s->n * s->n;
    for (long i = 0; i < c; ++i) {
        cpx psi = {0, 0};
        if (s->ksq[i] > 0) {
            psi.re = w[i].re / s->ksq[i];
            psi.im = w[i].im / s->ksq[i];
        }
        s->u[i].re = -s->ky[i] * psi.im;
        s->u[i].im = s->ky[i] * psi.re;
        s->v[i].re = s->kx[i] * psi.im;
        s->v[i].im = -s->kx[i] * psi.re;
    }
    fft2d_plan(&s->plan, s->u, 1);
    fft2d_plan(&s->plan, s->v, 1);
}

/* RHS = -FFT[u wx + v wy] - (nu k^2 + alpha) w + F on the mask (src/solver.cpp:186-203). */
static void solver_rhs(orc_solver *s, double nu, double alpha, const cpx *w, const cpx *F,
                       cpx *out) {
    const long c = s->n * s->n;
    solver_velocity(s, w);
    for (long i = 0; i < c; ++i) {
        s->wx[i].re = -s->kx[i] * w[i].im;
        s->wx[i].im = s->kx[i] * w[i].re;
        s->wy[i].re = -s->ky[i] * w[i].im;
        s->wy[i].im = s->ky[i] * w[i].re;
    }
    fft2d_plan(&s->plan, s->wx, 1);
    fft2d_plan(&s->plan, s->wy, 1);
    for (long i = 0; i < c; ++i) {
        s->nl[i].re = s->u[i].re * s->wx[i].re + s->v[i].re * s->wy[i].re;
        s->nl[i].im = 0;
    }
    fft2d_plan(&s->plan, s->nl, 0);
    for (long i = 0; i < c; ++i) {
        if (s->active[i]) {
            const double rate = nu * s->ksq[i] + alpha;
            out[i].re = -s->nl[i].re - rate * w[i].re + (F ? F[i].re : 0.0);
            out[i].im = -s->nl[i].im - rate * w[i].im + (F ? F[i].im : 0.0);
        } else {
            out[i].re = 0;
            out[i].im = 0;
        }
    }
}

/* One classical RK4 step with F held through all stages (src/solver.cpp:217-254). */
static void solver_step(orc_solver *s, double nu, double alpha, double dt, cpx *w, const cpx *F) {
    const long c = s->n * s->n;
    solver_rhs(s, nu, alpha, w, F, s->k1);
    for (long i = 0; i < c; ++i) {
        s->tmp[i].re = w[i].re + dt / 2 * s->k1[i].re;
        s->tmp[i].im = w[i].im + dt / 2 * s->k1[i].im;
    }
    solver_rhs(s, nu, alpha, s->tmp, F, s->k2);
    for (long i = 0; i < c; ++i) {
        s->tmp[i].re = w[i].re + dt / 2 * s->k2[i].re;
        s->tmp[i].im = w[i].im + dt / 2 * s->k2[i].im;
    }
    solver_rhs(s, nu, alpha, s->tmp, F, s->k3);
    for (long i = 0; i < c; ++i) {
        s->tmp[i].re = w[i].re + dt * s->k3[i].re;
        s->tmp[i].im = w[i].im + dt * s->k3[i].im;
    }
    solver_rhs(s, nu, alpha, s->tmp, F, s->k4);
    const double h = dt / 6;
    for (long i = 0; i < c; ++i) {
        w[i].re += h * (s->k1[i].re + 2.0 * s->k2[i].re + 2.0 * s->k3[i].re + s->k4[i].re);
        w[i].im += h * (s->k1[i].im + 2.0 * s->k2[i].im + 2.0 * s->k3[i].im + s->k4[i].im);
        if (!s->active[i]) {
            w[i].re = 0;
            w[i].im = 0;
        }
    }
}

/* One Shu-Osher SSPRK(3,3) step with F held through the three stages
 * (src/solver.cpp:235-246): u1 = w + dt k1, u2 = 3/4 w + 1/4 (u1 + dt k2),
 * w+ = w / 3 + 2/3 (u2 + dt k3). */
static void solver_step_ssprk3(orc_solver *s, double nu, double alpha, double dt, cpx *w,
                               const cpx *F) {
    const long c = s->n * s->n;
    cpx *u1 = s->tmp, *u2 = s->k4;
    solver_rhs(s, nu, alpha, w, F, s->k1);
    for (long i = 0; i < c; ++i) {
        u1[i].re = w[i].re + dt * s->k1[i].re;
        u1[i].im = w[i].im + dt * s->k1[i].im;
    }
    solver_rhs(s, nu, alpha, u1, F, s->k2);
    for (long i = 0; i < c; ++i) {
        u2[i].re = 0.75 * w[i].re + 0.25 * (u1[i].re + dt * s->k2[i].re);
        u2[i].im = 0.75 * w[i].im + 0.25 * (u1[i].im + dt * s->k2[i].im);
    }
    solver_rhs(s, nu, alpha, u2, F, s->k3);
    for (long i = 0; i < c; ++i) {
        w[i].re = w[i].re / 3.0 + 2.0 / 3.0 * (u2[i].re + dt * s->k3[i].re);
        w[i].im = w[i].im / 3.0 + 2.0 / 3.0 * (u2[i].im + dt * s->k3[i].im);
        if (!s->active[i]) {
            w[i].re = 0;
            w[i].im = 0;
        }
    }
}

/* `steps` SSPRK(3,3) steps in place, forcing as orc_rk4_steps. */
int orc_ssprk3_steps(long n, double nu, double alpha, double dt, long steps, double *w,
                     const double *forcing, long forcing_stride) {
    orc_solver s;
    if (solver_make(&s, n)) {
        solver_free(&s);
        return -1;
    }
    for (long k = 0; k < steps; ++k) {
        const cpx *F = forcing ? (const cpx *)(forcing + 2 * forcing_stride * k) : NULL;
        solver_step_ssprk3(&s, nu, alpha, dt, (cpx *)w, F);
    }
    solver_free(&s);
    return 0;
}

/* i kx w and i ky w of a spectral field, transformed to physical space. */
static void solver_gradient(orc_solver *s, const cpx *w, cpx *gx, cpx *gy) {
    const long c = s->n * s->n;
    for (long i = 0; i < c; ++i) {
        gx[i].re = -s->kx[i] * w[i].im;
        gx[i].im = s->kx[i] * w[i].re;
        gy[i].re = -s->ky[i] * w[i].im;
        gy[i].im = s->ky[i] * w[i].re;
    }
    fft2d_plan(&s->plan, gx, 1);
    fft2d_plan(&s->plan, gy, 1);
}

/* Tangent of the drift along `direction` at `state` (src/solver.cpp:264-290):
 * -FFT[u(d) . grad(state) + u(state) . grad(d)] - (nu k^2 + alpha) d on the
 * mask, with u(x) the velocity of x's streamfunction. */
int orc_linearized_drift(long n, double nu, double alpha, const double *state,
                         const double *direction, double *out) {
    orc_solver s;
    if (solver_make(&s, n)) {
        solver_free(&s);
        return -1;
    }
    const long c = n * n;
    const cpx *w = (const cpx *)state, *d = (const cpx *)direction;
    cpx *o = (cpx *)out;
    cpx *bu = s.k1, *bv = s.k2; /* velocity of the state */
    solver_velocity(&s, w);
    memcpy(bu, s.u, sizeof(cpx) * (size_t)c);
    memcpy(bv, s.v, sizeof(cpx) * (size_t)c);
    solver_gradient(&s, w, s.wx, s.wy);
    solver_velocity(&s, d);
    for (long i = 0; i < c; ++i) {
        s.nl[i].re = s.u[i].re * s.wx[i].re + s.v[i].re * s.wy[i].re;
        s.nl[i].im = 0;
    }
    solver_gradient(&s, d, s.wx, s.wy);
    for (long i = 0; i < c; ++i)
        s.nl[i].re += bu[i].re * s.wx[i].re + bv[i].re * s.wy[i].re;
    fft2d_plan(&s.plan, s.nl, 0);
    for (long i = 0; i < c; ++i) {
        o[i].re = o[i].im = 0;
        if (s.active[i]) {
            const double rate = nu * s.ksq[i] + alpha;
            o[i].re = -s.nl[i].re - rate * d[i].re;
            o[i].im = -s.nl[i].im - rate * d[i].im;
        }
    }
    solver_free(&s);
    return 0;
}

/* Adjoint of the drift's tangent at `state` applied to `cotangent`
 * (src/solver.cpp:291-323): FFT[u(state) . grad(lambda)]
 * + i ky / k^2 FFT[lambda wx] - i kx / k^2 FFT[lambda wy]
 * - (nu k^2 + alpha) lambda on the mask, lambda the physical cotangent and
 * (wx, wy) the physical gradient of the state. */
int orc_adjoint_drift(long n, double nu, double alpha, const double *state,
                      const double *cotangent, double *out) {
    orc_solver s;
    if (solver_make(&s, n)) {
        solver_free(&s);
        return -1;
    }
    const long c = n * n;
    const cpx *w = (const cpx *)state, *lam = (const cpx *)cotangent;
    cpx *o = (cpx *)out;
    cpx *phys = s.tmp, *lx = s.k1, *ly = s.k2;
    solver_velocity(&s, w);
    solver_gradient(&s, w, s.wx, s.wy);
    memcpy(phys, lam, sizeof(cpx) * (size_t)c);
    fft2d_plan(&s.plan, phys, 1);
    solver_gradient(&s, lam, lx, ly);
    for (long i = 0; i < c; ++i) {
        s.nl[i].re = s.u[i].re * lx[i].re + s.v[i].re * ly[i].re;
        s.nl[i].im = 0;
        s.wx[i].re = phys[i].re * s.wx[i].re;
        s.wx[i].im = 0;
        s.wy[i].re = phys[i].re * s.wy[i].re;
        s.wy[i].im = 0;
    }
    fft2d_plan(&s.plan, s.nl, 0);
    fft2d_plan(&s.plan, s.wx, 0);
    fft2d_plan(&s.plan, s.wy, 0);
    for (long i = 0; i < c; ++i) {
        o[i].re = o[i].im = 0;
        if (s.active[i]) {
            const double rate = nu * s.ksq[i] + alpha, q = 1.0 / s.ksq[i];
            /* i ky/k^2 a = (-ky/k^2 a.im, ky/k^2 a.re) */
            o[i].re = s.nl[i].re - s.ky[i] * q * s.wx[i].im + s.kx[i] * q * s.wy[i].im -
                      rate * lam[i].re;
            o[i].im = s.nl[i].im + s.ky[i] * q * s.wx[i].re - s.kx[i] * q * s.wy[i].re -
                      rate * lam[i].im;
        }
    }
    solver_free(&s);
    return 0;
}

/* Drift RHS of one state; F may be NULL. Returns -1 on bad n. */
int orc_rhs(long n, double nu, double alpha, const double *w, const double *F, double *out) {
    orc_solver s;
    if (solver_make(&s, n)) {
        solver_free(&s);
        return -1;
    }
    solver_rhs(&s, nu, alpha, (const cpx *)w, (const cpx *)F, (cpx *)out);
    solver_free(&s);
    return 0;
}

/* `steps` RK4 steps in place; forcing is either NULL, one field held for all
 * steps (forcing_stride = 0) or one field per step (forcing_stride = n*n). */
int orc_rk4_steps(long n, double nu, double alpha, double dt, long steps, double *w,
                  const double *forcing, long forcing_stride) {
    orc_solver s;
    if (solver_make(&s, n)) {
        solver_free(&s);
        return -1;
    }
    for (long k = 0; k < steps; ++k) {
        const cpx *F = forcing ? (const cpx *)(forcing + 2 * forcing_stride * k) : NULL;
        solver_step(&s, nu, alpha, dt, (cpx *)w, F);
    }
    solver_free(&s);
    return 0;
}

/* out[0..4] = energy, enstrophy, dissipation, injection, cfl_rate (src/solver.cpp:381-397). */
int orc_diagnostics(long n, double nu, const double *wd, const double *Fd, double *out) {
    orc_solver s;
    if (solver_make(&s, n)) {
        solver_free(&s);
        return -1;
    }
    const cpx *w = (const cpx *)wd, *F = (const cpx *)Fd;
    const double area = 4 * ORC_PI * ORC_PI;
    double e = 0, z = 0, inj = 0, cfl = 0;
    for (long i = 1; i < n * n; ++i) {
        const double norm = w[i].re * w[i].re + w[i].im * w[i].im;
        e += area / 2 * norm / s.ksq[i];
        z += area / 2 * norm;
        if (F)
            inj += area * (w[i].re * F[i].re + w[i].im * F[i].im) / s.ksq[i];
    }
    solver_velocity(&s, w);
    for (long i = 0; i < n * n; ++i) {
        const double r = (fabs(s.u[i].re) + fabs(s.v[i].re)) * (double)n / (2 * ORC_PI);
        if (r > cfl)
            cfl = r;
    }
    out[0] = e;
    out[1] = z;
    out[2] = 2 * nu * z;
    out[3] = inj;
    out[4] = cfl;
    solver_free(&s);
    return 0;
}

/* Shell spectrum binned by nearest |k| (src/solver.cpp:426-433); result has n bins. */
void orc_energy_spectrum(long n, const double *wd, double *bins) {
    const cpx *w = (const cpx *)wd;
    memset(bins, 0, sizeof(double) * (size_t)n);
    for (long i = 1; i < n * n; ++i) {
        const int a = orc_wave_number(i % n, n), b = orc_wave_number(i / n, n);
        const double k2 = (double)(a * a + b * b);
        const long shell = (long)floor(sqrt(k2) + 0.5);
        bins[shell] += 2 * ORC_PI * ORC_PI * (w[i].re * w[i].re + w[i].im * w[i].im) / k2;
    }
}

/* ------------------------------------------------------------ streams ---- */
/* SplitMix64-style logical stream identity (src/solver.cpp:61-66). */
uint64_t orc_stream_seed(uint64_t base, uint64_t index) {
    uint64_t x = base + 0x9e3779b97f4a7c15ULL * (index + 1);
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

/* mt19937_64 as specified by the C++ standard ([rand.eng.mers], [rand.predef]). */
typedef struct {
    uint64_t mt[312];
    int idx;
    int has_spare;
    double spare;
} orc_mt;

size_t orc_mt_size(void) {
    return sizeof(orc_mt);
}

void orc_mt_seed(orc_mt *s, uint64_t seed) {
    s->mt[0] = seed;
    for (int i = 1; i < 312; ++i)
        s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
    s->idx = 312;
    s->has_spare = 0;
    s->spare = 0;
}

uint64_t orc_mt_next(orc_mt *s) {
    const uint64_t upper = 0xFFFFFFFF80000000ULL, lower = 0x7FFFFFFFULL;
    if (s->idx >= 312) {
        for (int i = 0; i < 312; ++i) {
            const uint64_t x = (s->mt[i] & upper) | (s->mt[(i + 1) % 312] & lower);
            uint64_t xa = x >> 1;
            if (x & 1)
                xa ^= 0xB5026F5AA96619E9ULL;
            s->mt[i] = s->mt[(i + 156) % 312] ^ xa;
        }
        s->idx = 0;
    }
    uint64_t y = s->mt[s->idx++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= y >> 43;
    return y;
}

/* NormalStream: open-interval uniform and Box-Muller with a cached second
 * draw (src/solver.cpp:41-54). */
double orc_normal(orc_mt *s) {
    if (s->has_spare) {
        s->has_spare = 0;
        return s->spare;
    }
    const double u1 = ((double)(orc_mt_next(s) >> 11) + 0.5) * 0x1.0p-53;
    const double r = sqrt(-2 * log(u1));
    const double u2 = ((double)(orc_mt_next(s) >> 11) + 0.5) * 0x1.0p-53;
    const double angle = 2 * ORC_PI * u2;
    s->spare = r * sin(angle);
    s->has_spare = 1;
    return r * cos(angle);
}

/* ------------------------------------------------ OU shell forcing ---- */
/* forcing_pairs (src/solver.cpp:67-83) in the reference's order: grid index
 * i = y n + x of each conjugate pair's first mode (i < partner) inside the
 * mask and kmin^2 <= k^2 <= kmax^2.  Returns the count; first may be NULL. */
int orc_ou_pairs(long n, int kmin, int kmax, long *first) {
    int count = 0;
    for (long y = 0; y < n; ++y)
        for (long x = 0; x < n; ++x) {
            const int a = orc_wave_number(x, n), b = orc_wave_number(y, n);
            const long i = y * n + x, partner = ((n - y) % n) * n + (n - x) % n;
            const int k2 = a * a + b * b;
            if (3 * abs(a) < n && 3 * abs(b) < n && i < partner && k2 >= kmin * kmin &&
                k2 <= kmax * kmax) {
                if (first)
                    first[count] = i;
                ++count;
            }
        }
    return count;
}

static long ou_partner(long n, long i) {
    const long y = i / n, x = i % n;
    return ((n - y) % n) * n + (n - x) % n;
}

/* initialize_ou (src/solver.cpp:84-92): F = sigma sqrt(tau / 4) (xi1 + i xi2)
 * per pair, conjugate on the partner; two normals per pair, in pair order. */
void orc_ou_initialize(orc_mt *s, long n, int npairs, const long *first, double sigma,
                       double tau, double *Fd) {
    cpx *F = (cpx *)Fd;
    const double scale = sigma * sqrt(tau / 4);
    for (int j = 0; j < npairs; ++j) {
        const double re = orc_normal(s), im = orc_normal(s);
        F[first[j]].re = scale * re;
        F[first[j]].im = scale * im;
        F[ou_partner(n, first[j])].re = scale * re;
        F[ou_partner(n, first[j])].im = -(scale * im);
    }
}

/* transition_ou (src/solver.cpp:93-105): exact OU step over dt,
 * F <- e^{-dt/tau} F + sigma sqrt(-tau expm1(-2 dt / tau) / 4) (xi1 + i xi2). */
void orc_ou_transition(orc_mt *s, long n, int npairs, const long *first, double sigma,
                       double tau, double dt, double *Fd) {
    cpx *F = (cpx *)Fd;
    const double decay = exp(-dt / tau), scale = sigma * sqrt(-tau * expm1(-2 * dt / tau) / 4);
    for (int j = 0; j < npairs; ++j) {
        const double re = orc_normal(s), im = orc_normal(s);
        cpx *f = &F[first[j]];
        f->re = decay * f->re + scale * re;
        f->im = decay * f->im + scale * im;
        F[ou_partner(n, first[j])].re = f->re;
        F[ou_partner(n, first[j])].im = -f->im;
    }
}

/* Solver(n, p, seed) + `steps` Solver::step(dt) from the spectral state w (in
 * place): the stream mt19937_64(seed) draws the stationary OU start, then each
 * step advances the forcing by dt/2, takes one RK4 step with that midpoint
 * forcing held, and advances it by dt/2 (src/solver.cpp:120-139,217-254).
 * F_out receives the final forcing (n*n complex). Returns -1 on bad input. */
int orc_run_ou(long n, double nu, double alpha, double sigma, double tau, int kmin, int kmax,
               uint64_t seed, double dt, long steps, double *w, double *F_out) {
    const int np = orc_ou_pairs(n, kmin, kmax, NULL);
    if (np <= 0 || !(tau > 0))
        return -1;
    long *first = (long *)malloc(sizeof(long) * (size_t)np);
    orc_ou_pairs(n, kmin, kmax, first);
    orc_mt st;
    orc_mt_seed(&st, seed);
    memset(F_out, 0, sizeof(double) * 2 * (size_t)(n * n));
    orc_ou_initialize(&st, n, np, first, sigma, tau, F_out);
    orc_solver s;
    if (solver_make(&s, n)) {
        solver_free(&s);
        free(first);
        return -1;
    }
    for (long k = 0; k < steps; ++k) {
        orc_ou_transition(&st, n, np, first, sigma, tau, dt / 2, F_out);
        solver_step(&s, nu, alpha, dt, (cpx *)w, (const cpx *)F_out);
        orc_ou_transition(&st, n, np, first, sigma, tau, dt / 2, F_out);
    }
    solver_free(&s);
    free(first);
    return 0;
}

/* Half-open uniform used by the filter's resampler (src/filter.cpp:11-13). */
double orc_mt_uniform(orc_mt *s) {
    return (double)(orc_mt_next(s) >> 11) * 0x1.0p-53;
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Two standard normals from one Philox block: 53-bit open uniforms and
 * Box-Muller (cos, sin). */
void orc_philox_normal2(uint64_t key, uint32_t c0, uint64_t c12, uint32_t tag, double *g) {
    const uint32_t ctr[4] = {c0, (uint32_t)c12, (uint32_t)(c12 >> 32), tag};
    const uint32_t k[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, k, o);
    const uint64_t a = (((uint64_t)o[0] << 32) | o[1]) >> 11;
    const uint64_t b = (((uint64_t)o[2] << 32) | o[3]) >> 11;
    const double u1 = ((double)a + 0.5) * 0x1.0p-53, u2 = ((double)b + 0.5) * 0x1.0p-53;
    const double r = sqrt(-2.0 * log(u1)), t = 2.0 * ORC_PI * u2;
    g[0] = r * cos(t);
    g[1] = r * sin(t);
}

#define ORC_TAG_NOISE 0x4E4F4953u /* "NOIS" */
#define ORC_TAG_GRF 0x47524631u   /* "GRF1" */

/* Canonical shell modes kmin^2 <= kx^2+ky^2 <= kmax^2 of the half plane
 * (kx > 0, or kx == 0 and ky > 0), ordered by kx then ky.  The pair set is the
 * reference's forcing_pairs (src/solver.cpp:67-83); the order is the engine's. */
int orc_shell_modes(long n, int kmin, int kmax, int *kx, int *ky) {
    int count = 0;
    for (int a = 0; a <= kmax; ++a)
        for (int b = -kmax; b <= kmax; ++b) {
            const int k2 = a * a + b * b;
            if ((a > 0 || b > 0) && k2 >= kmin * kmin && k2 <= kmax * kmax && 3 * a < n &&
                3 * abs(b) < n) {
                if (kx)
                    kx[count] = a;
                if (ky)
                    ky[count] = b;
                ++count;
            }
        }
    return count;
}

static long full_index(long n, int a, int b) {
    const long x = ((a % n) + n) % n, y = ((b % n) + n) % n;
    return y * n + x;
}

/* Held forcing of one step in the full layout:
 *   F = amp * sin(kf y)  +  sigma / sqrt(2 dt) (xi1 + i xi2) on every shell pair,
 * xi from Philox keyed by member_key with counter (pair, step, "NOIS"). */
int orc_step_forcing(long n, double amp, int kf, double sigma, int kmin, int kmax, double dt,
                     uint64_t member_key, uint64_t step, double *Fd) {
    cpx *F = (cpx *)Fd;
    memset(F, 0, sizeof(cpx) * (size_t)(n * n));
    if (amp != 0 && kf > 0) {
        F[full_index(n, 0, kf)].im += -amp / 2;
        F[full_index(n, 0, -kf)].im += amp / 2;
    }
    if (sigma > 0) {
        const double scale = sigma / sqrt(2.0 * dt);
        int count = 0;
        for (int a = 0; a <= kmax; ++a)
            for (int b = -kmax; b <= kmax; ++b) {
                const int k2 = a * a + b * b;
                if (!((a > 0 || b > 0) && k2 >= kmin * kmin && k2 <= kmax * kmax))
                    continue;
                double g[2];
                orc_philox_normal2(member_key, (uint32_t)count, step, ORC_TAG_NOISE, g);
                const long i = full_index(n, a, b), j = full_index(n, -a, -b);
                F[i].re += scale * g[0];
                F[i].im += scale * g[1];
                F[j].re += scale * g[0];
                F[j].im -= scale * g[1];
                ++count;
            }
    }
    return 0;
}

/* Gaussian random field: on canonical retained modes (ordered by kx in 0..K,
 * then ky in -K..K) w = a(k) (xi1 + i xi2)/sqrt(2), a(k)^2 = k^5 exp(-(k/4)^2)
 * (so E(k) ~ k^4 exp(-(k/4)^2)), mirrored to -k, then scaled so that the
 * reference's enstrophy (A/2) sum |w|^2 equals 1. */
int orc_grf(long n, uint64_t member_key, double *wd) {
    if (!is_pow2(n))
        return -1;
    cpx *w = (cpx *)wd;
    memset(w, 0, sizeof(cpx) * (size_t)(n * n));
    const int K = (int)((n - 1) / 3);
    uint32_t q = 0;
    double sum = 0;
    for (int a = 0; a <= K; ++a)
        for (int b = -K; b <= K; ++b) {
            if (!(a > 0 || b > 0))
                continue;
            const double k2 = (double)(a * a + b * b), k = sqrt(k2);
            const double amp = pow(k, 2.5) * exp(-k2 / 32.0) / sqrt(2.0);
            double g[2];
            orc_philox_normal2(member_key, q, 0, ORC_TAG_GRF, g);
            ++q;
            const long i = full_index(n, a, b), j = full_index(n, -a, -b);
            w[i].re = amp * g[0];
            w[i].im = amp * g[1];
            w[j].re = amp * g[0];
            w[j].im = -amp * g[1];
            sum += w[i].re * w[i].re + w[i].im * w[i].im;
        }
    const double scale = 1.0 / sqrt(4 * ORC_PI * ORC_PI * sum);
    for (long i = 0; i < n * n; ++i) {
        w[i].re *= scale;
        w[i].im *= scale;
    }
    return 0;
}

/* One member trajectory of the engine's synthetic workload in FP64: GRF
 * initial field, then `steps` RK4 steps with per-step Kolmogorov + shell
 * white-noise forcing.  The step counter of step k is step0 + k. */
int orc_member_trajectory(long n, double nu, double alpha, double dt, long steps, uint64_t step0,
                          double amp, int kf, double sigma, int kmin, int kmax,
                          uint64_t member_key, int init_grf, double *w) {
    orc_solver s;
    if (solver_make(&s, n)) {
        solver_free(&s);
        return -1;
    }
    if (init_grf)
        orc_grf(n, member_key, w);
    double *F = (double *)malloc(sizeof(double) * 2 * (size_t)(n * n));
    for (long k = 0; k < steps; ++k) {
        orc_step_forcing(n, amp, kf, sigma, kmin, kmax, dt, member_key, step0 + (uint64_t)k, F);
        solver_step(&s, nu, alpha, dt, (cpx *)w, (const cpx *)F);
    }
    free(F);
    solver_free(&s);
    return 0;
}

/* Final states of members first..first+count-1 of the synthetic workload,
 * member-parallel over `threads` OpenMP threads; out holds count n*n fields. */
int orc_trajectories(long n, double nu, double alpha, double dt, long steps, double amp, int kf,
                     double sigma, int kmin, int kmax, uint64_t seed, long first, long count,
                     int threads, double *out) {
    int status = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) reduction(| : status)
    for (long m = 0; m < count; ++m)
        status |= orc_member_trajectory(n, nu, alpha, dt, steps, 0, amp, kf, sigma, kmin, kmax,
                                        orc_stream_seed(seed, (uint64_t)(first + m)), 1,
                                        out + 2 * n * n * m);
    return status;
}

/* CPU baseline: `members` independent trajectories, member-parallel over
 * `threads` OpenMP threads (the scheduling-independent pattern of the
 * reference ensemble(), src/inference.cpp:555-590).  Final energies out. */
int orc_ensemble(long n, double nu, double alpha, double dt, long steps, double amp, int kf,
                 double sigma, int kmin, int kmax, uint64_t seed, long members, int threads,
                 double *energy) {
    int status = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) reduction(| : status)
    for (long m = 0; m < members; ++m) {
        double *w = (double *)malloc(sizeof(double) * 2 * (size_t)(n * n));
        double d[5];
        status |= orc_member_trajectory(n, nu, alpha, dt, steps, 0, amp, kf, sigma, kmin, kmax,
                                        orc_stream_seed(seed, (uint64_t)m), 1, w);
        orc_diagnostics(n, nu, w, NULL, d);
        energy[m] = d[0];
        free(w);
    }
    return status;
}

/* ----------------------------------------------------- particle filter ---- */
/* Gaussian log-weight increment of one particle (src/filter.cpp:66-74):
 * logw = log(w_prev) - sum_o [ r^2/2 + log sd + log(2 pi)/2 ]. */
double orc_pf_logweight(double prev_weight, long nobs, const double *pred, const double *value,
                        const double *sd) {
    double v = prev_weight > 0 ? log(prev_weight) : -INFINITY;
    for (long o = 0; o < nobs; ++o) {
        const double r = (pred[o] - value[o]) / sd[o];
        v -= 0.5 * r * r + log(sd[o]) + 0.5 * log(2 * ORC_PI);
    }
    return v;
}

/* Log-sum-exp normalisation, weights and ESS (src/filter.cpp:76-92).
 * Returns the increment (or -inf when every weight is -inf). */
double orc_pf_normalize(long count, const double *logs, double *weights, double *ess,
                        double *max_weight) {
    double m = -INFINITY;
    for (long i = 0; i < count; ++i)
        if (logs[i] > m)
            m = logs[i];
    if (!isfinite(m))
        return -INFINITY;
    double total = 0;
    for (long i = 0; i < count; ++i)
        total += exp(logs[i] - m);
    const double inc = m + log(total);
    double sq = 0, mw = 0;
    for (long i = 0; i < count; ++i) {
        weights[i] = exp(logs[i] - inc);
        sq += weights[i] * weights[i];
        if (weights[i] > mw)
            mw = weights[i];
    }
    if (ess)
        *ess = 1 / sq;
    if (max_weight)
        *max_weight = mw;
    return inc;
}

/* Systematic resampling with one uniform u in [0,1): position_i = u/N + i/N,
 * ancestor advances while position > cumulative (src/filter.cpp:102-120). */
void orc_systematic_resample(long count, const double *weights, double u, int *ancestors) {
    long a = 0;
    double cum = weights[0];
    const double start = u / (double)count;
    for (long i = 0; i < count; ++i) {
        const double pos = start + (double)i / (double)count;
        while (pos > cum && a + 1 < count)
            cum += weights[++a];
        ancestors[i] = (int)a;
    }
}

/* ---------------------------------------------------------------- EnKF ---- */
/* One serial perturbed-observation scalar update on `dim` complex state
 * entries per member (src/filter.cpp:265-321, localisation excluded):
 *   cov_j = sum_i (x_ij - mean_j) (h_i - hbar) / (M-1)
 *   var   = sum_i (h_i - hbar)^2 / (M-1)
 *   x_ij += cov_j (y + sd*eps_i - h_i) / (var + sd^2).
 * states: members x dim complex (interleaved), h: predictions, eps: normals. */
void orc_enkf_scalar_update(long members, long dim, double *states, const double *h, double y,
                            double sd, const double *eps) {
    cpx *x = (cpx *)states;
    const double M = (double)members;
    double hbar = 0;
    for (long i = 0; i < members; ++i)
        hbar += h[i] / M;
    cpx *mean = (cpx *)calloc((size_t)dim, sizeof(cpx));
    cpx *cov = (cpx *)calloc((size_t)dim, sizeof(cpx));
    for (long i = 0; i < members; ++i)
        for (long j = 0; j < dim; ++j) {
            mean[j].re += x[i * dim + j].re / M;
            mean[j].im += x[i * dim + j].im / M;
        }
    double var = 0;
    for (long i = 0; i < members; ++i) {
        const double d = h[i] - hbar;
        var += d * d / (M - 1);
        for (long j = 0; j < dim; ++j) {
            cov[j].re += (x[i * dim + j].re - mean[j].re) * (d / (M - 1));
            cov[j].im += (x[i * dim + j].im - mean[j].im) * (d / (M - 1));
        }
    }
    const double den = var + sd * sd;
    for (long i = 0; i < members; ++i) {
        const double g = (y + sd * eps[i] - h[i]) / den;
        for (long j = 0; j < dim; ++j) {
            x[i * dim + j].re += cov[j].re * g;
            x[i * dim + j].im += cov[j].im * g;
        }
    }
    free(mean);
    free(cov);
}
