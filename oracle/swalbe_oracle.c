/*
 * swalbe_oracle.c -- CPU FP64 restatement of Swalbe.jl's 2D thin-film LBM
 * step (reference v1.0.6).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product.  Only tests/,
 * the smoke check and the benchmark's CPU-baseline leg load this library.
 *
 * Every op is a whole-array sweep in the same order as the reference's
 * broadcasts and circshift! calls, so this file is also the
 * reference-equivalent CPU timing path (single thread, like Swalbe).
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off).
 *
 * Parity status of each function is recorded where it is defined.  "Pinned"
 * means a reference known-answer test or a mathematical identity checks it
 * in tests/test_oracle_*.py.
 */
#include "swalbe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NQ 9
/* D2Q9 velocities; population k is Julia plane k+1 (src/collide.jl:270-282),
 * streaming shifts from src/collide.jl:92-100. */
static const int CX[NQ] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int CY[NQ] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

static inline int64_t wrap(int64_t a, int64_t L) {
    a %= L;
    return a < 0 ? a + L : a;
}

/* ------------------------------------------------------------------ */
/* Portable math.  The product implements the same published algorithms
 * (Philox4x32-10 of Salmon et al. 2011; Box-Muller; atanh-series log;
 * Taylor-series sin/cos of pi*r on |r| <= 1/4), op for op, so that FP64
 * results agree to the bit.  Pinned: Philox by the Random123 known-answer
 * vectors, log/sincospi by libm within 2 ulp and by exact special values. */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

static const double LOG_C[12] = {
    2.0, 0.6666666666666666, 0.4, 0.2857142857142857, 0.2222222222222222,
    0.18181818181818182, 0.15384615384615385, 0.13333333333333333,
    0.11764705882352941, 0.10526315789473684, 0.09523809523809523,
    0.08695652173913043};
static const double LN2_HI = 0.6931471805599453;
static const double LN2_LO = 2.3190468138462996e-17;
static const double SQRT2 = 1.4142135623730951;

/* ln(x) for finite normal x > 0: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 atanh(s), s = (m-1)/(m+1). */
double orc_log(double x) {
    uint64_t b;
    memcpy(&b, &x, 8);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    uint64_t mb = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m;
    memcpy(&m, &mb, 8);
    if (m > SQRT2) {
        m = m * 0.5;
        e = e + 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double z = s * s;
    double q = LOG_C[11];
    for (int k = 10; k >= 1; --k) q = LOG_C[k] + z * q;
    double lnm = s * (LOG_C[0] + z * q);
    double de = (double)e;
    return de * LN2_HI + (lnm + de * LN2_LO);
}

static const double SIN_C[10] = {
    3.141592653589793, -5.16771278004997, 2.5501640398773455, -0.5992645293207921,
    0.08214588661112823, -0.0073704309457143504, 0.00046630280576761255,
    -2.1915353447830217e-05, 7.952054001475513e-07, -2.2948428997269873e-08};
static const double COS_C[11] = {
    1.0, -4.934802200544679, 4.0587121264167685, -1.3352627688545895,
    0.2353306303588932, -0.02580689139001406, 0.0019295743094039231,
    -0.0001046381049248457, 4.303069587032947e-06, -1.3878952462213771e-07,
    3.604730797462501e-09};

/* sin(pi x), cos(pi x): reduce to r in [-1/4, 1/4] by quarter turns. */
void orc_sincospi(double x, double *s, double *c) {
    double ax = fabs(x);
    if (ax >= 2.0) ax = ax - 2.0 * floor(ax * 0.5);
    double k = floor(2.0 * ax + 0.5);
    double r = ax - 0.5 * k;
    double z = r * r;
    double sp = SIN_C[9];
    for (int i = 8; i >= 0; --i) sp = SIN_C[i] + z * sp;
    sp = r * sp;
    double cp = COS_C[10];
    for (int i = 9; i >= 0; --i) cp = COS_C[i] + z * cp;
    int q = ((int)k) & 3;
    double so, co;
    switch (q) {
    case 0: so = sp; co = cp; break;
    case 1: so = cp; co = -sp; break;
    case 2: so = -sp; co = -cp; break;
    default: so = -cp; co = sp; break;
    }
    if (x < 0) so = -so;
    *s = so;
    *c = co;
}

double orc_cospi(double x) {
    double s, c;
    orc_sincospi(x, &s, &c);
    return c;
}

/* Counter layout: {c0, tag, step_lo, step_hi}, key {seed_lo, seed_hi}.
 * Normals: stream index s; one call per index pair (c0 = s / 2); parity
 * b = s % 2 uses the 32-bit uniforms (w + 1/2) 2^-32 of words 2b and 2b+1,
 * then Box-Muller.  Lattice node (x, y) draws from stream index
 * orc_noise_index(x, y, Lx) = 2 (x + Lx floor(y/2)) + y mod 2, so the two
 * nodes of a vertical pair (x, 2m), (x, 2m+1) share one Philox call.
 * Uniforms: the node's own counter (c0 = x + Lx y), 53 bits of words 0, 1. */
uint32_t orc_noise_index(int x, int y, int Lx) {
    return 2u * ((uint32_t)x + (uint32_t)Lx * (uint32_t)(y >> 1)) + (uint32_t)(y & 1);
}

static void philox_ctr(uint64_t seed, uint64_t step, uint32_t c0, uint32_t tag, uint32_t r[4]) {
    uint32_t ctr[4] = {c0, tag, (uint32_t)step, (uint32_t)(step >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    orc_philox4x32_10(ctr, key, r);
}

static inline double u53(uint32_t hi, uint32_t lo) {
    uint64_t v = (((uint64_t)hi << 32) | lo) >> 11;
    return ((double)v + 0.5) * 1.1102230246251565e-16; /* 2^-53 */
}

static inline double u32(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386963e-10; }

void orc_normal_pair(uint64_t seed, uint64_t step, uint32_t node, uint32_t tag, double *z0,
                     double *z1) {
    uint32_t r[4];
    philox_ctr(seed, step, node >> 1, tag, r);
    int b = (int)(node & 1u);
    double u1 = u32(r[2 * b]);
    double u2 = u32(r[2 * b + 1]);
    double rad = sqrt(-2.0 * orc_log(u1));
    double s, c;
    orc_sincospi(2.0 * u2, &s, &c);
    *z0 = rad * c;
    *z1 = rad * s;
}

double orc_uniform(uint64_t seed, uint64_t step, uint32_t node, uint32_t tag) {
    uint32_t r[4];
    philox_ctr(seed, step, node, tag, r);
    return u53(r[0], r[1]);
}

/* ------------------------------------------------------------------ */
/* Julia circshift!(dst, src, (sx, sy)): dst[i,j] = src[i-sx, j-sy], periodic.
 * Pinned: streaming KATs of test/collide.jl:27-123. */
void orc_circshift(double *dst, const double *src, int Lx, int Ly, int sx, int sy) {
    for (int j = 0; j < Ly; ++j) {
        int64_t js = wrap(j - sy, Ly);
        for (int i = 0; i < Lx; ++i) {
            int64_t is = wrap(i - sx, Lx);
            dst[i + (int64_t)Lx * j] = src[is + (int64_t)Lx * js];
        }
    }
}

/* equilibrium!(feq, height, velx, vely, vsq, g): src/equilibrium.jl:63-116.
 * Pinned: test/equilibrium.jl:1-85 and the doctest src/equilibrium.jl:35-54. */
void orc_equilibrium(double *feq, const double *h, const double *ux, const double *uy,
                     double *vsq, double g, int Lx, int Ly) {
    const int64_t N = (int64_t)Lx * Ly;
    const double g0 = 1.5 * g;
    const double w1 = 1.0 / 9.0;
    const double w5 = 1.0 / 36.0;
    for (int64_t x = 0; x < N; ++x) vsq[x] = ux[x] * ux[x] + uy[x] * uy[x];
    double *f0 = feq, *f1 = feq + N, *f2 = feq + 2 * N, *f3 = feq + 3 * N, *f4 = feq + 4 * N;
    double *f5 = feq + 5 * N, *f6 = feq + 6 * N, *f7 = feq + 7 * N, *f8 = feq + 8 * N;
    for (int64_t x = 0; x < N; ++x)
        f0[x] = h[x] * ((1.0 - (5.0 / 6.0) * g * h[x]) - (2.0 / 3.0) * vsq[x]);
    for (int64_t x = 0; x < N; ++x)
        f1[x] = w1 * h[x] * (((g0 * h[x] + 3.0 * ux[x]) + 4.5 * (ux[x] * ux[x])) - 1.5 * vsq[x]);
    for (int64_t x = 0; x < N; ++x)
        f2[x] = w1 * h[x] * (((g0 * h[x] + 3.0 * uy[x]) + 4.5 * (uy[x] * uy[x])) - 1.5 * vsq[x]);
    for (int64_t x = 0; x < N; ++x)
        f3[x] = w1 * h[x] * (((g0 * h[x] - 3.0 * ux[x]) + 4.5 * (ux[x] * ux[x])) - 1.5 * vsq[x]);
    for (int64_t x = 0; x < N; ++x)
        f4[x] = w1 * h[x] * (((g0 * h[x] - 3.0 * uy[x]) + 4.5 * (uy[x] * uy[x])) - 1.5 * vsq[x]);
    for (int64_t x = 0; x < N; ++x) {
        double a = ux[x] + uy[x];
        f5[x] = w5 * h[x] * (((g0 * h[x] + 3.0 * a) + 4.5 * (a * a)) - 1.5 * vsq[x]);
    }
    for (int64_t x = 0; x < N; ++x) {
        double a = uy[x] - ux[x];
        f6[x] = w5 * h[x] * (((g0 * h[x] + 3.0 * a) + 4.5 * (a * a)) - 1.5 * vsq[x]);
    }
    for (int64_t x = 0; x < N; ++x) {
        double a = ux[x] + uy[x];
        f7[x] = w5 * h[x] * (((g0 * h[x] - 3.0 * a) + 4.5 * (a * a)) - 1.5 * vsq[x]);
    }
    for (int64_t x = 0; x < N; ++x) {
        double a = ux[x] - uy[x];
        f8[x] = w5 * h[x] * (((g0 * h[x] + 3.0 * a) + 4.5 * (a * a)) - 1.5 * vsq[x]);
    }
}

/* BGKandStream!(fout, feq, ftemp, Fx, Fy, τ): src/collide.jl:70-105.
 * Collide into fout, circshift! each plane into ftemp, then fout .= ftemp.
 * Pinned: test/collide.jl:27-123 (bit-exact ==) and doctests. */
void orc_bgk_and_stream(double *fout, const double *feq, double *ftemp, const double *Fx,
                        const double *Fy, double tau, int Lx, int Ly) {
    const int64_t N = (int64_t)Lx * Ly;
    const double omeg = 1.0 - 1.0 / tau;
    const double itau = 1.0 / tau;
    const double c3 = 1.0 / 3.0, c24 = 1.0 / 24.0;
    const double *fe[NQ];
    double *ft[NQ], *fo[NQ];
    for (int k = 0; k < NQ; ++k) {
        fe[k] = feq + k * N;
        ft[k] = ftemp + k * N;
        fo[k] = fout + k * N;
    }
    for (int64_t x = 0; x < N; ++x) fo[0][x] = omeg * ft[0][x] + itau * fe[0][x];
    for (int64_t x = 0; x < N; ++x) fo[1][x] = (omeg * ft[1][x] + itau * fe[1][x]) + c3 * Fx[x];
    for (int64_t x = 0; x < N; ++x) fo[2][x] = (omeg * ft[2][x] + itau * fe[2][x]) + c3 * Fy[x];
    for (int64_t x = 0; x < N; ++x) fo[3][x] = (omeg * ft[3][x] + itau * fe[3][x]) - c3 * Fx[x];
    for (int64_t x = 0; x < N; ++x) fo[4][x] = (omeg * ft[4][x] + itau * fe[4][x]) - c3 * Fy[x];
    for (int64_t x = 0; x < N; ++x)
        fo[5][x] = (omeg * ft[5][x] + itau * fe[5][x]) + c24 * (Fx[x] + Fy[x]);
    for (int64_t x = 0; x < N; ++x)
        fo[6][x] = (omeg * ft[6][x] + itau * fe[6][x]) + c24 * (Fy[x] - Fx[x]);
    for (int64_t x = 0; x < N; ++x)
        fo[7][x] = (omeg * ft[7][x] + itau * fe[7][x]) - c24 * (Fx[x] + Fy[x]);
    for (int64_t x = 0; x < N; ++x)
        fo[8][x] = (omeg * ft[8][x] + itau * fe[8][x]) + c24 * (Fx[x] - Fy[x]);
    for (int k = 0; k < NQ; ++k) orc_circshift(ft[k], fo[k], Lx, Ly, CX[k], CY[k]);
    memcpy(fout, ftemp, sizeof(double) * NQ * N);
}

/* moments!(height, velx, vely, fout): src/moments.jl:43-52: sum! over populations 0..8 in order, then
 * u = Σ c f / h.  Pinned: test/moments.jl:1-135 (bit-exact ==). */
void orc_moments(double *h, double *ux, double *uy, const double *fout, int Lx, int Ly) {
    const int64_t N = (int64_t)Lx * Ly;
    const double *f[NQ];
    for (int k = 0; k < NQ; ++k) f[k] = fout + k * N;
    for (int64_t x = 0; x < N; ++x) h[x] = 0.0;
    for (int k = 0; k < NQ; ++k)
        for (int64_t x = 0; x < N; ++x) h[x] = h[x] + f[k][x];
    for (int64_t x = 0; x < N; ++x)
        ux[x] = (((((f[1][x] - f[3][x]) + f[5][x]) - f[6][x]) - f[7][x]) + f[8][x]) / h[x];
    for (int64_t x = 0; x < N; ++x)
        uy[x] = (((((f[2][x] - f[4][x]) + f[5][x]) + f[6][x]) - f[7][x]) - f[8][x]) / h[x];
}

/* power_broad(arg, n): src/pressure.jl:363-369.  Pinned: test/pressure.jl:139-161. */
double orc_power_broad(double arg, int n) {
    double t = 1.0;
    for (int i = 0; i < n; ++i) t *= arg;
    return t;
}
/* fast_93 / fast_32: src/pressure.jl:392-422. */
double orc_fast_93(double arg) {
    double t = (arg * arg) * arg;
    return ((t * t) * t) - t;
}
double orc_fast_32(double arg) { return ((arg * arg) * arg) - (arg * arg); }

/* The 8 neighbour copies of src/pressure.jl:129-139 (same shifts in
 * src/forcing.jl:169-179 and src/differences.jl:59-67):
 * planes hip, hjp, him, hjm, hipjp, himjp, himjm, hipjm. */
static const int NB_SX[8] = {1, 0, -1, 0, 1, -1, -1, 1};
static const int NB_SY[8] = {0, 1, 0, -1, 1, 1, -1, -1};
static void neighbours(double *dgrad, const double *f, int Lx, int Ly) {
    const int64_t N = (int64_t)Lx * Ly;
    for (int q = 0; q < 8; ++q) orc_circshift(dgrad + q * N, f, Lx, Ly, NB_SX[q], NB_SY[q]);
}

/* Laplacian part shared by filmpressure! and ∇²f!:
 * 2/3 (hjp+hip+him+hjm) + 1/6 (hipjp+himjp+himjm+hipjm) - 10/3 f. */
static inline double lap9(const double *d, int64_t N, int64_t x, double fx) {
    const double hip = d[x], hjp = d[N + x], him = d[2 * N + x], hjm = d[3 * N + x];
    const double hipjp = d[4 * N + x], himjp = d[5 * N + x], himjm = d[6 * N + x],
                 hipjm = d[7 * N + x];
    return ((2.0 / 3.0) * (((hjp + hip) + him) + hjm) +
            (1.0 / 6.0) * (((hipjp + himjp) + himjm) + hipjm)) -
           (10.0 / 3.0) * fx;
}

static inline double kappa(double theta, int n, int m, double hmin) {
    return (((1.0 - orc_cospi(theta)) * (double)(n - 1)) * (double)(m - 1)) /
           ((double)(n - m) * hmin);
}

/* filmpressure!(state, sys; θ, γ, n, m, hmin, hcrit): src/pressure.jl:119-155
 * (power_broad form used by every State method).  θ may be a field.
 * Pinned: test/pressure.jl:3-53 and the doctest src/pressure.jl:39-63. */
void orc_filmpressure(double *p, const double *h, double *dgrad, double gamma, double theta,
                      const double *theta_field, int n, int m, double hmin, double hcrit,
                      int Lx, int Ly) {
    const int64_t N = (int64_t)Lx * Ly;
    neighbours(dgrad, h, Lx, Ly);
    for (int64_t x = 0; x < N; ++x) {
        double th = theta_field ? theta_field[x] : theta;
        double a = hmin / (h[x] + hcrit);
        p[x] = -gamma * (kappa(th, n, m, hmin) * (orc_power_broad(a, n) - orc_power_broad(a, m)));
    }
    for (int64_t x = 0; x < N; ++x) p[x] = p[x] - gamma * lap9(dgrad, N, x, h[x]);
}

/* filmpressure!(output, f, dgrad, γ, θ, n, m, hmin, hcrit): src/pressure.jl:72-115.
 * fast_93 / fast_32 only; any other (n, m) is a DomainError. */
int orc_filmpressure_raw(double *p, const double *h, double *dgrad, double gamma, double theta,
                         const double *theta_field, int n, int m, double hmin, double hcrit,
                         int Lx, int Ly) {
    const int64_t N = (int64_t)Lx * Ly;
    int which;
    if (n == 9 && m == 3) which = 93;
    else if (n == 3 && m == 2) which = 32;
    else return ORC_EDOMAIN;
    neighbours(dgrad, h, Lx, Ly);
    for (int64_t x = 0; x < N; ++x) {
        double th = theta_field ? theta_field[x] : theta;
        double a = hmin / (h[x] + hcrit);
        double pw = which == 93 ? orc_fast_93(a) : orc_fast_32(a);
        p[x] = -gamma * (kappa(th, n, m, hmin) * pw);
    }
    for (int64_t x = 0; x < N; ++x) p[x] = p[x] - gamma * lap9(dgrad, N, x, h[x]);
    return ORC_OK;
}

/* ∇f!(outx, outy, f[, a]): src/differences.jl:153-206 (3-, 4- and 5-argument
 * forms; a == NULL is the 3-argument form).  Pinned: test/differences.jl:3-76. */
void orc_grad(double *outx, double *outy, const double *f, const double *a, int Lx, int Ly) {
    const int64_t N = (int64_t)Lx * Ly;
    double *d = (double *)malloc(sizeof(double) * 8 * N);
    neighbours(d, f, Lx, Ly);
    for (int64_t x = 0; x < N; ++x) {
        const double fip = d[x], fjp = d[N + x], fim = d[2 * N + x], fjm = d[3 * N + x];
        const double fipjp = d[4 * N + x], fimjp = d[5 * N + x], fimjm = d[6 * N + x],
                     fipjm = d[7 * N + x];
        double gx = (-1.0 / 3.0) * (fip - fim) - (1.0 / 12.0) * (((fipjp - fimjp) - fimjm) + fipjm);
        double gy = (-1.0 / 3.0) * (fjp - fjm) - (1.0 / 12.0) * (((fipjp + fimjp) - fimjm) - fipjm);
        outx[x] = a ? a[x] * gx : gx;
        outy[x] = a ? a[x] * gy : gy;
    }
    free(d);
}

/* ∇²f!(output, f, γ): src/differences.jl:57-75.  Pinned: test/differences.jl:104-124. */
void orc_laplacian(double *out, const double *f, double gamma, int Lx, int Ly) {
    const int64_t N = (int64_t)Lx * Ly;
    double *d = (double *)malloc(sizeof(double) * 8 * N);
    neighbours(d, f, Lx, Ly);
    for (int64_t x = 0; x < N; ++x) out[x] = gamma * lap9(d, N, x, f[x]);
    free(d);
}

/* h∇p!(state): src/forcing.jl:168-187.  Pinned: test/forcing.jl:104-139
 * (bit-exact ==) and the doctest src/forcing.jl:144-158. */
void orc_hgradp(double *hgpx, double *hgpy, const double *p, const double *h, double *dgrad,
                int Lx, int Ly) {
    const int64_t N = (int64_t)Lx * Ly;
    neighbours(dgrad, p, Lx, Ly);
    const double *d = dgrad;
    for (int64_t x = 0; x < N; ++x) {
        const double fip = d[x], fim = d[2 * N + x];
        const double fipjp = d[4 * N + x], fimjp = d[5 * N + x], fimjm = d[6 * N + x],
                     fipjm = d[7 * N + x];
        hgpx[x] = h[x] * ((-1.0 / 3.0) * (fip - fim) -
                          (1.0 / 12.0) * (((fipjp - fimjp) - fimjm) + fipjm));
    }
    for (int64_t x = 0; x < N; ++x) {
        const double fjp = d[N + x], fjm = d[3 * N + x];
        const double fipjp = d[4 * N + x], fimjp = d[5 * N + x], fimjm = d[6 * N + x],
                     fipjm = d[7 * N + x];
        hgpy[x] = h[x] * ((-1.0 / 3.0) * (fjp - fjm) -
                          (1.0 / 12.0) * (((fipjp + fimjp) - fimjm) - fipjm));
    }
}

/* slippage!(slipx, slipy, h, ux, uy, δ, μ): src/forcing.jl:42-46.
 * Pinned: test/forcing.jl:13-69. */
void orc_slippage(double *slipx, double *slipy, const double *h, const double *ux,
                  const double *uy, double delta, double mu, int Lx, int Ly) {
    const int64_t N = (int64_t)Lx * Ly;
    const double c6mu = 6.0 * mu, c6d = 6.0 * delta, c3dd = 3.0 * (delta * delta);
    for (int64_t x = 0; x < N; ++x)
        slipx[x] = (c6mu * h[x] * ux[x]) / ((2.0 * (h[x] * h[x]) + c6d * h[x]) + c3dd);
    for (int64_t x = 0; x < N; ++x)
        slipy[x] = (c6mu * h[x] * uy[x]) / ((2.0 * (h[x] * h[x]) + c6d * h[x]) + c3dd);
}

/* thermal!(fx, fy, h, kbt, μ, δ): src/forcing.jl:297-311, with randn!
 * replaced by counter-based normals keyed (seed, step, node).  The noise
 * amplitude is pinned by test/forcing.jl:141-164 (mean, var = 2 kbt/11);
 * the normals themselves have no reference counterpart (randn! stream). */
void orc_thermal(double *kx, double *ky, const double *h, double kbt, double mu, double delta,
                 uint64_t seed, uint64_t step, int Lx, int Ly) {
    const int64_t N = (int64_t)Lx * Ly;
    for (int64_t x = 0; x < N; ++x) {
        double z0, z1;
        orc_normal_pair(seed, step, orc_noise_index((int)(x % Lx), (int)(x / Lx), Lx), 0u, &z0,
                        &z1);
        kx[x] = z0;
        ky[x] = z1;
    }
    for (int64_t x = 0; x < N; ++x)
        kx[x] = kx[x] * sqrt((((2.0 * kbt) * mu) * 6.0) * h[x] /
                             ((((2.0 * h[x]) * h[x]) + ((6.0 * h[x]) * delta)) + ((3.0 * delta) * delta)));
    for (int64_t x = 0; x < N; ++x)
        ky[x] = ky[x] * sqrt((((2.0 * kbt) * mu) * 6.0) * h[x] /
                             ((((2.0 * h[x]) * h[x]) + ((6.0 * h[x]) * delta)) + ((3.0 * delta) * delta)));
}

/* inclination!(α, state; t, tstart, tsmooth): src/forcing.jl:363-368;
 * fac = 0.5 + 0.5 tanh((t - tstart)/tsmooth) is a host scalar.
 * Pinned: test/forcing.jl:166-184. */
void orc_inclination(double *Fx, double *Fy, const double *h, double ax, double ay, double fac,
                     int Lx, int Ly) {
    const int64_t N = (int64_t)Lx * Ly;
    for (int64_t x = 0; x < N; ++x) Fx[x] = Fx[x] + (h[x] * ax) * fac;
    for (int64_t x = 0; x < N; ++x) Fy[x] = Fy[x] + (h[x] * ay) * fac;
}

/* Fx .= -h∇px .- slipx (.- kbtx): src/simulate.jl:18-19,
 * scripts/Rivulet_stability.jl:123-124. */
void orc_force_sum(double *Fx, double *Fy, const double *hgpx, const double *hgpy,
                   const double *slipx, const double *slipy, const double *kbtx,
                   const double *kbty, int Lx, int Ly) {
    const int64_t N = (int64_t)Lx * Ly;
    for (int64_t x = 0; x < N; ++x) Fx[x] = -hgpx[x] - slipx[x];
    for (int64_t x = 0; x < N; ++x) Fy[x] = -hgpy[x] - slipy[x];
    if (kbtx) {
        for (int64_t x = 0; x < N; ++x) Fx[x] = Fx[x] - kbtx[x];
        for (int64_t x = 0; x < N; ++x) Fy[x] = Fy[x] - kbty[x];
    }
}

/* Seeded initial conditions (no reference counterpart: they replace
 * randinterface!'s randn!, src/initialvalues.jl:23-33, by a counter stream). */
void orc_init_uniform(double *h, int Lx, int Ly, uint64_t seed, double base, double amp) {
    const int64_t N = (int64_t)Lx * Ly;
    for (int64_t x = 0; x < N; ++x) h[x] = base + amp * orc_uniform(seed, 0, (uint32_t)x, 2u);
}
void orc_init_normal(double *h, int Lx, int Ly, uint64_t seed, double base, double amp) {
    const int64_t N = (int64_t)Lx * Ly;
    for (int64_t x = 0; x < N; ++x) {
        double z0, z1;
        orc_normal_pair(seed, 0, orc_noise_index((int)(x % Lx), (int)(x / Lx), Lx), 1u, &z0, &z1);
        h[x] = base + amp * z0;
    }
}

/* time_loop (src/simulate.jl:6-96): the 7-call loop, with the mass every
 * tdump, the Δh series, the θ-field variant, the measurement callback
 * (wetted! or inclination!) and the thermal force as options. */
int orc_time_loop(orc_state *s, const orc_params *p, int64_t nsteps, const orc_loop_opts *o) {
    const int Lx = s->Lx, Ly = s->Ly;
    const int64_t N = (int64_t)Lx * Ly;
    int64_t nmass = 0;
    for (int64_t it = 1; it <= nsteps; ++it) {
        const int64_t t = o->t0 + it;
        if (p->tdump > 0 && t % p->tdump == 0 && o->mass_series) {
            double mass = 0.0;
            for (int64_t x = 0; x < N; ++x) mass += s->height[x];
            o->mass_series[nmass++] = mass;
        }
        if (o->dh_series) {
            double mx = s->height[0], mn = s->height[0];
            for (int64_t x = 1; x < N; ++x) {
                mx = s->height[x] > mx ? s->height[x] : mx;
                mn = s->height[x] < mn ? s->height[x] : mn;
            }
            o->dh_series[it - 1] = mx - mn;
        }
        orc_filmpressure(s->pressure, s->height, s->dgrad, p->gamma, p->theta, o->theta_field,
                         p->n, p->m, p->hmin, p->hcrit, Lx, Ly);
        orc_hgradp(s->hgpx, s->hgpy, s->pressure, s->height, s->dgrad, Lx, Ly);
        orc_slippage(s->slipx, s->slipy, s->height, s->velx, s->vely, p->delta, p->mu, Lx, Ly);
        if (o->thermal)
            orc_thermal(s->kbtx, s->kbty, s->height, p->kbt, p->mu, p->delta, o->seed,
                        (uint64_t)t, Lx, Ly);
        orc_force_sum(s->Fx, s->Fy, s->hgpx, s->hgpy, s->slipx, s->slipy,
                      o->thermal ? s->kbtx : NULL, o->thermal ? s->kbty : NULL, Lx, Ly);
        if (o->wet_series) {
            int64_t c = 0;
            for (int64_t x = 0; x < N; ++x) c += s->height[x] > o->wet_thresh;
            o->wet_series[it - 1] = c;
        }
        if (o->body)
            orc_inclination(s->Fx, s->Fy, s->height, o->body_ax, o->body_ay, o->body_fac, Lx, Ly);
        orc_equilibrium(s->feq, s->height, s->velx, s->vely, s->vsq, p->g, Lx, Ly);
        orc_bgk_and_stream(s->fout, s->feq, s->ftemp, s->Fx, s->Fy, p->tau, Lx, Ly);
        orc_moments(s->height, s->velx, s->vely, s->fout, Lx, Ly);
    }
    return ORC_OK;
}

/* ∇f_simple!(outx, outy, f, dgrad): src/measures.jl:175-192, the 9-point
 * gradient written with -1/3 and -1/12 weights in this order.
 * Pinned: test/measures.jl:67-78 (exact Xgrad / Ygrad of reshape(1:25, 5, 5)). */
void orc_grad_simple(double *outx, double *outy, const double *f, int Lx, int Ly) {
    for (int y = 0; y < Ly; ++y)
        for (int x = 0; x < Lx; ++x) {
            const int xm = (x - 1 + Lx) % Lx, xp = (x + 1) % Lx;
            const int ym = (y - 1 + Ly) % Ly, yp = (y + 1) % Ly;
#define F_(a, b) f[(a) + (int64_t)Lx * (b)]
            const double fip = F_(xm, y), fim = F_(xp, y), fjp = F_(x, ym), fjm = F_(x, yp);
            const double fipjp = F_(xm, ym), fimjp = F_(xp, ym), fimjm = F_(xp, yp),
                         fipjm = F_(xm, yp);
#undef F_
            const int64_t i = x + (int64_t)Lx * y;
            outx[i] = ((-1.0 / 3.0) * (fip - fim)) - ((1.0 / 12.0) * (((fipjp - fimjp) - fimjm) + fipjm));
            outy[i] = ((-1.0 / 3.0) * (fjp - fjm)) - ((1.0 / 12.0) * (((fipjp + fimjp) - fimjm) - fipjm));
        }
}

/* wetted! (src/measures.jl:6-17), fluid_dry! (:23-28) and surfacearea!
 * (:126-168) of one height field: out = {count(h > thresh), maximum(h),
 * Σ sqrt(gx² + gy² + 1) over h > thresh, that sum minus count·cos πθ
 * (theta_field NULL) or Σ cos πθ over h > thresh}.  Sums run sequentially in
 * column-major order (the reference's sum over a masked vector).  mask may
 * be NULL.  Pinned: test/measures.jl:2-97. */
void orc_measure(const double *h, const double *theta_field, double theta, int Lx, int Ly,
                 double thresh, double out[4], unsigned char *mask) {
    const int64_t N = (int64_t)Lx * Ly;
    double *gx = (double *)malloc(sizeof(double) * N), *gy = (double *)malloc(sizeof(double) * N);
    orc_grad_simple(gx, gy, h, Lx, Ly);
    double cnt = 0.0, mx = -INFINITY, surf = 0.0, cs = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        const int wet = h[i] > thresh;
        mx = h[i] > mx ? h[i] : mx;
        if (mask) mask[i] = (unsigned char)wet;
        if (!wet) continue;
        cnt += 1.0;
        surf += sqrt((gx[i] * gx[i] + gy[i] * gy[i]) + 1.0);
        if (theta_field) cs += orc_cospi(theta_field[i]);
    }
    out[0] = cnt;
    out[1] = mx;
    out[2] = surf;
    out[3] = surf - (theta_field ? cs : cnt * orc_cospi(theta));
    free(gx);
    free(gy);
}

/* t0(; hᵦ, γ, μ, θ) (src/measures.jl:56-61): the spinodal dewetting time
 * scale 3μ/(γ q²²) with q² = hᵦ(1 - cos πθ)(2 - 3hᵦ).  Pinned:
 * test/measures.jl:117-120 (t0() ≈ 177428.32340802). */
void orc_t0(double hb, double gamma, double mu, double theta, double out[2]) {
    const double qsq = ((hb * (1.0 - orc_cospi(theta))) * (2.0 - 3.0 * hb));
    out[0] = (3.0 * mu) / (gamma * (qsq * qsq));
    out[1] = qsq;
}
