/*
 * swalbe_cpu_fused.c -- TEST AND BASELINE INFRASTRUCTURE ONLY: the 2D step of
 * swalbe_oracle.c reorganised for the CPU as three row-parallel OpenMP sweeps
 * per step (pressure; forces; collide at the destination, pull-stream and
 * moments) instead of the reference's ~56 whole-array sweeps.  It is the
 * multi-core CPU baseline that bench.py reports beside the single-thread
 * oracle.  Every per-node expression keeps the oracle's evaluation order, so
 * its results equal orc_time_loop's bit for bit (tests/test_oracle_fused.py).
 * Nothing in the product links or calls it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "swalbe_oracle.h"

#define NQF 9
static const int FCX[NQF] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int FCY[NQF] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

static inline int wrapf(int a, int L) { return a < 0 ? a + L : (a >= L ? a - L : a); }

int orc_fused_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* equilibrium population k of one node (src/equilibrium.jl:63-116, the
 * oracle's expressions) */
static inline double feq_k(int k, double h, double ux, double uy, double g) {
    const double g0 = 1.5 * g, w1 = 1.0 / 9.0, w5 = 1.0 / 36.0;
    const double vsq = ux * ux + uy * uy;
    double a;
    switch (k) {
        case 0: return h * ((1.0 - (5.0 / 6.0) * g * h) - (2.0 / 3.0) * vsq);
        case 1: return w1 * h * (((g0 * h + 3.0 * ux) + 4.5 * (ux * ux)) - 1.5 * vsq);
        case 2: return w1 * h * (((g0 * h + 3.0 * uy) + 4.5 * (uy * uy)) - 1.5 * vsq);
        case 3: return w1 * h * (((g0 * h - 3.0 * ux) + 4.5 * (ux * ux)) - 1.5 * vsq);
        case 4: return w1 * h * (((g0 * h - 3.0 * uy) + 4.5 * (uy * uy)) - 1.5 * vsq);
        case 5: a = ux + uy; return w5 * h * (((g0 * h + 3.0 * a) + 4.5 * (a * a)) - 1.5 * vsq);
        case 6: a = uy - ux; return w5 * h * (((g0 * h + 3.0 * a) + 4.5 * (a * a)) - 1.5 * vsq);
        case 7: a = ux + uy; return w5 * h * (((g0 * h - 3.0 * a) + 4.5 * (a * a)) - 1.5 * vsq);
        default: a = ux - uy; return w5 * h * (((g0 * h + 3.0 * a) + 4.5 * (a * a)) - 1.5 * vsq);
    }
}

/* the BGK force term of population k (src/collide.jl:79-89) */
static inline double force_k(int k, double base, double Fx, double Fy) {
    const double c3 = 1.0 / 3.0, c24 = 1.0 / 24.0;
    switch (k) {
        case 0: return base;
        case 1: return base + c3 * Fx;
        case 2: return base + c3 * Fy;
        case 3: return base - c3 * Fx;
        case 4: return base - c3 * Fy;
        case 5: return base + c24 * (Fx + Fy);
        case 6: return base + c24 * (Fy - Fx);
        case 7: return base - c24 * (Fx + Fy);
        default: return base + c24 * (Fx - Fy);
    }
}

/* nsteps of the plain time loop (src/simulate.jl:6-24, plus the thermal
 * force when `thermal`), steps t0+1 .. t0+nsteps; scalar θ. */
int orc_fused_time_loop(orc_state *s, const orc_params *p, int64_t nsteps, int thermal,
                        uint64_t seed, int64_t t0) {
    const int Lx = s->Lx, Ly = s->Ly;
    const int64_t N = (int64_t)Lx * Ly;
    double *P = (double *)malloc(sizeof(double) * N);
    double *FX = (double *)malloc(sizeof(double) * N), *FY = (double *)malloc(sizeof(double) * N);
    double *H2 = (double *)malloc(sizeof(double) * N), *U2 = (double *)malloc(sizeof(double) * N);
    double *V2 = (double *)malloc(sizeof(double) * N);
    double *FT2 = (double *)malloc(sizeof(double) * NQF * N);
    if (!P || !FX || !FY || !H2 || !U2 || !V2 || !FT2) {
        free(P); free(FX); free(FY); free(H2); free(U2); free(V2); free(FT2);
        return ORC_EDOMAIN;
    }
    double *h = s->height, *ux = s->velx, *uy = s->vely, *ft = s->ftemp;
    const double kap = (((1.0 - orc_cospi(p->theta)) * (double)(p->n - 1)) * (double)(p->m - 1)) /
                       ((double)(p->n - p->m) * p->hmin);
    const double omeg = 1.0 - 1.0 / p->tau, itau = 1.0 / p->tau;
    const double c6mu = 6.0 * p->mu, c6d = 6.0 * p->delta, c3dd = 3.0 * (p->delta * p->delta);
    const double kb26 = ((2.0 * p->kbt) * p->mu) * 6.0;
    for (int64_t it = 1; it <= nsteps; ++it) {
        const uint64_t t = (uint64_t)(t0 + it);
#pragma omp parallel for schedule(static)
        for (int y = 0; y < Ly; ++y) {
            const int ym = wrapf(y - 1, Ly), yp = wrapf(y + 1, Ly);
            for (int x = 0; x < Lx; ++x) {
                const int xm = wrapf(x - 1, Lx), xp = wrapf(x + 1, Lx);
#define HH(a, b) h[(a) + (int64_t)Lx * (b)]
                const double hc = HH(x, y);
                const double hip = HH(xm, y), hjp = HH(x, ym), him = HH(xp, y), hjm = HH(x, yp);
                const double hipjp = HH(xm, ym), himjp = HH(xp, ym), himjm = HH(xp, yp),
                             hipjm = HH(xm, yp);
#undef HH
                const double a = p->hmin / (hc + p->hcrit);
                double v = -p->gamma * (kap * (orc_power_broad(a, p->n) - orc_power_broad(a, p->m)));
                const double lap = ((2.0 / 3.0) * (((hjp + hip) + him) + hjm) +
                                    (1.0 / 6.0) * (((hipjp + himjp) + himjm) + hipjm)) -
                                   (10.0 / 3.0) * hc;
                P[x + (int64_t)Lx * y] = v - p->gamma * lap;
            }
        }
#pragma omp parallel for schedule(static)
        for (int y = 0; y < Ly; ++y) {
            const int ym = wrapf(y - 1, Ly), yp = wrapf(y + 1, Ly);
            for (int x = 0; x < Lx; ++x) {
                const int xm = wrapf(x - 1, Lx), xp = wrapf(x + 1, Lx);
                const int64_t i = x + (int64_t)Lx * y;
#define PP(a, b) P[(a) + (int64_t)Lx * (b)]
                const double fip = PP(xm, y), fjp = PP(x, ym), fim = PP(xp, y), fjm = PP(x, yp);
                const double fipjp = PP(xm, ym), fimjp = PP(xp, ym), fimjm = PP(xp, yp),
                             fipjm = PP(xm, yp);
#undef PP
                const double hc = h[i];
                const double hgx = hc * ((-1.0 / 3.0) * (fip - fim) -
                                         (1.0 / 12.0) * (((fipjp - fimjp) - fimjm) + fipjm));
                const double hgy = hc * ((-1.0 / 3.0) * (fjp - fjm) -
                                         (1.0 / 12.0) * (((fipjp + fimjp) - fimjm) - fipjm));
                const double den = (2.0 * (hc * hc) + c6d * hc) + c3dd;
                const double sx = (c6mu * hc * ux[i]) / den, sy = (c6mu * hc * uy[i]) / den;
                double Fx = -hgx - sx, Fy = -hgy - sy;
                if (thermal) {
                    double z0, z1;
                    orc_normal_pair(seed, t, orc_noise_index(x, y, Lx), 0u, &z0, &z1);
                    const double amp = sqrt(kb26 * hc / ((((2.0 * hc) * hc) + ((6.0 * hc) * p->delta)) +
                                                         ((3.0 * p->delta) * p->delta)));
                    Fx = Fx - z0 * amp;
                    Fy = Fy - z1 * amp;
                }
                FX[i] = Fx;
                FY[i] = Fy;
            }
        }
#pragma omp parallel for schedule(static)
        for (int y = 0; y < Ly; ++y) {
            for (int x = 0; x < Lx; ++x) {
                const int64_t i = x + (int64_t)Lx * y;
                double f[NQF];
                for (int k = 0; k < NQF; ++k) {
                    const int sxx = wrapf(x - FCX[k], Lx), syy = wrapf(y - FCY[k], Ly);
                    const int64_t j = sxx + (int64_t)Lx * syy;
                    const double base = omeg * ft[k * N + j] + itau * feq_k(k, h[j], ux[j], uy[j], p->g);
                    f[k] = force_k(k, base, FX[j], FY[j]);
                    FT2[k * N + i] = f[k];
                }
                double hn = 0.0;
                for (int k = 0; k < NQF; ++k) hn = hn + f[k];
                H2[i] = hn;
                U2[i] = (((((f[1] - f[3]) + f[5]) - f[6]) - f[7]) + f[8]) / hn;
                V2[i] = (((((f[2] - f[4]) + f[5]) + f[6]) - f[7]) - f[8]) / hn;
            }
        }
        double *tmp;  /* the new state becomes current (pointer swap, no copy) */
        tmp = ft; ft = FT2; FT2 = tmp;
        tmp = h; h = H2; H2 = tmp;
        tmp = ux; ux = U2; U2 = tmp;
        tmp = uy; uy = V2; V2 = tmp;
    }
    if (h != s->height) {  /* an odd number of steps: the results sit in the scratch arrays */
        memcpy(s->height, h, sizeof(double) * N);
        memcpy(s->velx, ux, sizeof(double) * N);
        memcpy(s->vely, uy, sizeof(double) * N);
        memcpy(s->ftemp, ft, sizeof(double) * NQF * N);
        /* the scratch arrays are the ones malloc'ed above: free those */
        double *t1 = h, *t2 = ux, *t3 = uy, *t4 = ft;
        h = H2; ux = U2; uy = V2; ft = FT2;
        H2 = t1; U2 = t2; V2 = t3; FT2 = t4;
    }
    memcpy(s->fout, s->ftemp, sizeof(double) * NQF * N);
    free(P); free(FX); free(FY); free(H2); free(U2); free(V2); free(FT2);
    return ORC_OK;
}
