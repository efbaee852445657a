/*
 * swalbe_oracle_1d.c -- TEST INFRASTRUCTURE ONLY: the FP64 CPU restatement of
 * Swalbe.jl's one-dimensional D1Q3 thin-film step (the 1D methods of src/),
 * used by the tests as the checker of the GPU path.  Nothing in the product
 * links or calls it.
 *
 * Layout as the reference: a field is a Vector of L nodes, a distribution an
 * (L, 3) matrix (population k at offset k*L).  Velocities c = 0, +1, -1.
 * Every expression keeps the reference's evaluation order (compile with
 * -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "swalbe_oracle.h"

/* equilibrium!(feq, height, velocity, gravity): src/equilibrium.jl:169-180.
 * Pinned: test/equilibrium.jl:87-153 (five cases, exact). */
void orc1d_equilibrium(double *feq, const double *h, const double *u, double g, int L) {
    for (int x = 0; x < L; ++x) {
        const double uu = u[x] * u[x];
        feq[x] = h[x] * ((1.0 - (0.5 * g) * h[x]) - uu);
        feq[L + x] = h[x] * (((0.25 * g) * h[x] + 0.5 * u[x]) + 0.5 * uu);
        feq[2 * L + x] = h[x] * (((0.25 * g) * h[x] - 0.5 * u[x]) + 0.5 * uu);
    }
}

/* BGKandStream!(fout, feq, ftemp, F::Vector, τ): src/collide.jl:179-205; with
 * border masks it is BGKandStream!(state::StateWithBound_1D, sys) (:214-249):
 * the part of population 1 (2) flagged by border1 (border2) bounces back into
 * population 2 (1) of the same node instead of streaming.
 * Pinned: test/collide.jl:125-226. */
void orc1d_bgk_and_stream(double *fout, const double *feq, double *ftemp, const double *F,
                          double tau, const double *border1, const double *border2, int L) {
    const double omeg = 1.0 - 1.0 / tau, itau = 1.0 / tau;
    double *fo = (double *)malloc(sizeof(double) * 3 * L);
    double *fb = (double *)malloc(sizeof(double) * 2 * L);
    for (int x = 0; x < L; ++x) {
        fo[x] = omeg * ftemp[x] + itau * feq[x];
        fo[L + x] = (omeg * ftemp[L + x] + itau * feq[L + x]) + 0.5 * F[x];
        fo[2 * L + x] = (omeg * ftemp[2 * L + x] + itau * feq[2 * L + x]) - 0.5 * F[x];
    }
    if (border1) {
        for (int x = 0; x < L; ++x) {
            fb[x] = fo[L + x] * border1[x];
            fb[L + x] = fo[2 * L + x] * border2[x];
        }
        for (int x = 0; x < L; ++x) {
            fo[L + x] = fo[L + x] - fb[x];
            fo[2 * L + x] = fo[2 * L + x] - fb[L + x];
        }
    }
    for (int x = 0; x < L; ++x) {
        ftemp[x] = fo[x];
        ftemp[L + x] = fo[L + (x - 1 + L) % L];       /* circshift(fo1, 1) */
        ftemp[2 * L + x] = fo[2 * L + (x + 1) % L];   /* circshift(fo2, -1) */
    }
    if (border1) {
        for (int x = 0; x < L; ++x) {
            ftemp[L + x] = ftemp[L + x] + fb[L + x];
            ftemp[2 * L + x] = ftemp[2 * L + x] + fb[x];
        }
    }
    memcpy(fout, ftemp, sizeof(double) * 3 * L);
    free(fo);
    free(fb);
}

/* moments!(height::Vector, vel, fout): src/moments.jl:54-61 (sum! adds the
 * populations in order).  Pinned: test/moments.jl:1-135. */
void orc1d_moments(double *h, double *u, const double *f, int L) {
    for (int x = 0; x < L; ++x) {
        h[x] = (f[x] + f[L + x]) + f[2 * L + x];
        u[x] = (f[L + x] - f[2 * L + x]) / h[x];
    }
}

static inline double kappa1d(double theta, int n, int m, double hmin) {
    return (((1.0 - orc_cospi(theta)) * (double)(n - 1)) * (double)(m - 1)) /
           ((double)(n - m) * hmin);
}

/* filmpressure!(state::LBM_state_1D, sys; θ, n, m, hmin, hcrit, γ):
 * src/pressure.jl:231-255 (power_broad form); θ may be a field
 * (run_dropletpatterned 1D, src/simulate.jl:442-460) and so may γ (the
 * State_gamma_1D method, src/pressure.jl:284-315, called with γ = gamma by
 * run_gamma, src/simulate.jl:545).  gamma_field == NULL uses the scalar.
 * Pinned: test/pressure.jl:56-103; a constant γ field reproduces the scalar. */
void orc1d_filmpressure(double *p, const double *h, double gamma, const double *gamma_field,
                        double theta, const double *theta_field, int n, int m, double hmin,
                        double hcrit, int L) {
    for (int x = 0; x < L; ++x) {
        const double th = theta_field ? theta_field[x] : theta;
        const double gm = gamma_field ? gamma_field[x] : gamma;
        const double a = hmin / (h[x] + hcrit);
        p[x] = -gm * (kappa1d(th, n, m, hmin) * (orc_power_broad(a, n) - orc_power_broad(a, m)));
    }
    for (int x = 0; x < L; ++x) {
        const double gm = gamma_field ? gamma_field[x] : gamma;
        const double hip = h[(x - 1 + L) % L], him = h[(x + 1) % L];
        p[x] = p[x] - gm * ((hip - 2.0 * h[x]) + him);
    }
}

/* ∇γ!(state) (src/forcing.jl:423-432): -3/2 (γ[x-1] - γ[x+1]) / 2 (form 0), and
 * ∇γ!(state, sys) (:434-447): (2h² + 6δh + 3δ²)/(6h) · h/2 · (γ[x-1] - γ[x+1])/2
 * (form 1).  Pinned by a linear γ ramp (exact slope) and a constant γ (zero). */
void orc1d_grad_gamma(double *out, const double *gam, const double *h, double delta, int form,
                      int L) {
    for (int x = 0; x < L; ++x) {
        const double fip = gam[(x - 1 + L) % L], fim = gam[(x + 1) % L];
        if (form == 0) {
            out[x] = -1.5 * ((fip - fim) / 2.0);
        } else {
            const double a = ((2.0 * (h[x] * h[x])) + ((6.0 * delta) * h[x])) + 3.0 * (delta * delta);
            out[x] = (((a / (6.0 * h[x])) * h[x]) / 2.0) * ((fip - fim) / 2.0);
        }
    }
}

/* filmpressure!(output::Vector, f, dgrad, γ, θ, n, m, hmin, hcrit):
 * src/pressure.jl:196-229, fast_93 / fast_32 only, DomainError otherwise. */
int orc1d_filmpressure_raw(double *p, const double *h, double gamma, double theta, int n, int m,
                           double hmin, double hcrit, int L) {
    int which;
    if (n == 9 && m == 3) which = 93;
    else if (n == 3 && m == 2) which = 32;
    else return ORC_EDOMAIN;
    for (int x = 0; x < L; ++x) {
        const double a = hmin / (h[x] + hcrit);
        const double pw = which == 93 ? orc_fast_93(a) : orc_fast_32(a);
        p[x] = -gamma * (kappa1d(theta, n, m, hmin) * pw);
    }
    for (int x = 0; x < L; ++x) {
        const double hip = h[(x - 1 + L) % L], him = h[(x + 1) % L];
        p[x] = p[x] - gamma * ((hip - 2.0 * h[x]) + him);
    }
    return ORC_OK;
}

/* ∇f!(output::Vector, f, dgrad[, a]): src/differences.jl:208-228, the
 * central difference -1/2 (f[i-1] - f[i+1]) (times a).  a == NULL is the
 * three-argument form.  Pinned: test/differences.jl:79-100. */
void orc1d_grad(double *out, const double *f, const double *a, int L) {
    for (int x = 0; x < L; ++x) {
        const double fip = f[(x - 1 + L) % L], fim = f[(x + 1) % L];
        out[x] = a ? (a[x] * -0.5) * (fip - fim) : -0.5 * (fip - fim);
    }
}

/* ∇²f!(output, f::Vector, dgrad): src/differences.jl:77-85.
 * Pinned: test/differences.jl:126-139. */
void orc1d_laplacian(double *out, const double *f, int L) {
    for (int x = 0; x < L; ++x) {
        const double hip = f[(x - 1 + L) % L], him = f[(x + 1) % L];
        out[x] = (hip - 2.0 * f[x]) + him;
    }
}

/* h∇p!(state::LBM_state_1D): src/forcing.jl:189-198 (height .* -0.5 .* (fip .- fim)).
 * Pinned: test/forcing.jl:104-139. */
void orc1d_hgradp(double *hgp, const double *p, const double *h, int L) {
    for (int x = 0; x < L; ++x) {
        const double fip = p[(x - 1 + L) % L], fim = p[(x + 1) % L];
        hgp[x] = (h[x] * -0.5) * (fip - fim);
    }
}

/* slippage!(slip, height, vel, δ, μ): src/forcing.jl:68-71.
 * Pinned: test/forcing.jl:71-102. */
void orc1d_slippage(double *slip, const double *h, const double *u, double delta, double mu,
                    int L) {
    const double c6mu = 6.0 * mu, c6d = 6.0 * delta, c3dd = 3.0 * (delta * delta);
    for (int x = 0; x < L; ++x)
        slip[x] = (c6mu * h[x] * u[x]) / ((2.0 * (h[x] * h[x]) + c6d * h[x]) + c3dd);
}

/* thermal!(fluc, height, kᵦT, μ, δ): src/forcing.jl:322-332, randn! replaced
 * by the counter-based normal of stream index x (its pair's first normal).
 * The amplitude is pinned by test/forcing.jl:141-164 (mean, var = 2 kbt/11). */
void orc1d_thermal(double *fluc, const double *h, double kbt, double mu, double delta,
                   uint64_t seed, uint64_t step, int L) {
    for (int x = 0; x < L; ++x) {
        double z0, z1;
        orc_normal_pair(seed, step, (uint32_t)x, 0u, &z0, &z1);
        fluc[x] = z0;
    }
    for (int x = 0; x < L; ++x)
        fluc[x] = fluc[x] * sqrt((((2.0 * kbt) * mu) * 6.0) * h[x] /
                                 ((((2.0 * h[x]) * h[x]) + ((6.0 * h[x]) * delta)) +
                                  ((3.0 * delta) * delta)));
}

/* inclination!(α::Float64, state::State_1D; t, tstart, tsmooth): src/forcing.jl:379-383,
 * fac = 0.5 + 0.5 tanh((t - tstart)/tsmooth).  Pinned: test/forcing.jl:166-184. */
void orc1d_inclination(double *F, const double *h, double alpha, double fac, int L) {
    for (int x = 0; x < L; ++x) F[x] = F[x] + (h[x] * alpha) * fac;
}

/* state.F .= -state.h∇p .- state.slip (src/simulate.jl:111), minus the noise
 * when it is on. */
void orc1d_force_sum(double *F, const double *hgp, const double *slip, const double *kbt, int L) {
    for (int x = 0; x < L; ++x) F[x] = -hgp[x] - slip[x];
    if (kbt)
        for (int x = 0; x < L; ++x) F[x] = F[x] - kbt[x];
}

/* time_loop(sys::SysConst_1D, state::State_1D[, θ | Δh | f, measure]):
 * src/simulate.jl:98-178, and the bounce-back loop (:181-205) when borders
 * are given. */
int orc1d_time_loop(orc1d_state *st, const orc_params *p, int64_t nsteps,
                    const orc1d_loop_opts *o) {
    const int L = st->L;
    orc1d_loop_opts zero;
    memset(&zero, 0, sizeof(zero));
    if (!o) o = &zero;
    int64_t km = 0;
    for (int64_t it = 1; it <= nsteps; ++it) {
        const int64_t t = o->t0 + it;
        if (o->mass_series && p->tdump > 0 && t % p->tdump == 0) {
            /* Σ(h - interior h) with walls (src/simulate.jl:189), else Σh (:101) */
            double s = 0.0;
            for (int x = 0; x < L; ++x)
                s += o->interior ? st->height[x] - o->interior[x] * st->height[x] : st->height[x];
            o->mass_series[km++] = s;
        }
        if (o->dh_series) {
            double mx = st->height[0], mn = st->height[0];
            for (int x = 1; x < L; ++x) {
                mx = st->height[x] > mx ? st->height[x] : mx;
                mn = st->height[x] < mn ? st->height[x] : mn;
            }
            o->dh_series[it - 1] = mx - mn;
        }
        orc1d_filmpressure(st->pressure, st->height, p->gamma, o->gamma_field, p->theta,
                           o->theta_field, p->n, p->m, p->hmin, p->hcrit, L);
        orc1d_hgradp(st->hgp, st->pressure, st->height, L);
        orc1d_slippage(st->slip, st->height, st->vel, p->delta, p->mu, L);
        if (o->thermal)
            orc1d_thermal(st->kbt, st->height, p->kbt, p->mu, p->delta, o->seed, (uint64_t)t, L);
        orc1d_force_sum(st->F, st->hgp, st->slip, o->thermal ? st->kbt : NULL, L);
        /* run_gamma: F .= -h∇p .- slip .- ∇γ (src/simulate.jl:548); the
         * coalescence scans add it (analyze/droplet_coalescence.jl:148) */
        if (o->grad_gamma && o->grad_gamma_sign < 0)
            for (int x = 0; x < L; ++x) st->F[x] = st->F[x] - o->grad_gamma[x];
        if (o->grad_gamma && o->grad_gamma_sign > 0)
            for (int x = 0; x < L; ++x) st->F[x] = st->F[x] + o->grad_gamma[x];
        if (o->wet_series) {
            int64_t c = 0;
            for (int x = 0; x < L; ++x) c += st->height[x] > o->wet_thresh ? 1 : 0;
            o->wet_series[it - 1] = c;
        }
        if (o->body) orc1d_inclination(st->F, st->height, o->alpha, o->body_fac, L);
        orc1d_equilibrium(st->feq, st->height, st->vel, p->g, L);
        orc1d_bgk_and_stream(st->fout, st->feq, st->ftemp, st->F, p->tau, o->border1, o->border2,
                             L);
        orc1d_moments(st->height, st->vel, st->fout, L);
    }
    return ORC_OK;
}
