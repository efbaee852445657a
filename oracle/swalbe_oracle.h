/*
 * swalbe_oracle.h -- CPU FP64 restatement of Swalbe.jl's 2D thin-film
 * lattice Boltzmann time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product (libswalbe_hip.so).  Only the test suite, the smoke check and the
 * CPU-baseline leg of the benchmark may load it.  Nothing in the product
 * links, calls or falls back to it.
 *
 * Arrays are Julia column-major: a field (Lx, Ly) stores node (i, j)
 * (0-based) at i + Lx*j; a distribution (Lx, Ly, 9) stores population k at
 * plane offset k*Lx*Ly.  Everything is Float64, compiled with
 * -ffp-contract=off, and every expression is evaluated in the order Julia
 * evaluates the reference broadcast (left to right, no fused multiply-add).
 */
#ifndef SWALBE_ORACLE_H
#define SWALBE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OK 0
#define ORC_EDOMAIN 1

/* Taumucs{T} (src/initialize.jl:43-60), Float64. */
typedef struct orc_params {
    double tau, mu, delta, kbt, gamma, hmin, hcrit, theta, g;
    int32_t n, m;
    int64_t tdump;
} orc_params;

/* State{T,3} (src/initialize.jl:149-168) plus State_thermal's kbtx/kbty
 * (src/initialize.jl:182-186). */
typedef struct orc_state {
    int32_t Lx, Ly;
    double *fout, *ftemp, *feq;
    double *height, *velx, *vely, *vsq, *pressure;
    double *Fx, *Fy, *slipx, *slipy, *hgpx, *hgpy;
    double *dgrad; /* (Lx, Ly, 8) */
    double *kbtx, *kbty;
} orc_state;

/* Options of the four 2D time_loop methods (src/simulate.jl:6-96) plus the
 * thermal force of scripts/Rivulet_stability.jl:123-124. */
typedef struct orc_loop_opts {
    const double *theta_field; /* time_loop(sys, state, θ) (simulate.jl:26-45) */
    int32_t thermal;           /* F -= kbt from thermal! each step */
    int32_t body;              /* inclination! callback (simulate.jl:89, forcing.jl:363-368) */
    uint64_t seed;
    int64_t t0;                /* steps run are t0+1 .. t0+nsteps */
    double body_ax, body_ay, body_fac;
    double *dh_series;         /* Δh variant (simulate.jl:56); NULL = off */
    int64_t *wet_series;       /* wetted! callback (measures.jl:13-17); NULL = off */
    double wet_thresh;
    double *mass_series;       /* sum(height) when t % tdump == 0 (simulate.jl:8-14) */
} orc_loop_opts;

/* ---- portable math shared with the product's spec ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_log(double x);
void orc_sincospi(double x, double *s, double *c);
double orc_cospi(double x);
/* stream index of lattice node (x, y) for the normals: 2 (x + Lx floor(y/2)) + y mod 2 */
uint32_t orc_noise_index(int x, int y, int Lx);
void orc_normal_pair(uint64_t seed, uint64_t step, uint32_t node, uint32_t tag,
                     double *z0, double *z1);
double orc_uniform(uint64_t seed, uint64_t step, uint32_t node, uint32_t tag);

/* ---- ops (one function per reference function) ---- */
void orc_circshift(double *dst, const double *src, int Lx, int Ly, int sx, int sy);
void orc_equilibrium(double *feq, const double *h, const double *ux, const double *uy,
                     double *vsq, double g, int Lx, int Ly);
void orc_bgk_and_stream(double *fout, const double *feq, double *ftemp, const double *Fx,
                        const double *Fy, double tau, int Lx, int Ly);
void orc_moments(double *h, double *ux, double *uy, const double *fout, int Lx, int Ly);
double orc_power_broad(double arg, int n);
double orc_fast_93(double arg);
double orc_fast_32(double arg);
void orc_filmpressure(double *p, const double *h, double *dgrad, double gamma, double theta,
                      const double *theta_field, int n, int m, double hmin, double hcrit,
                      int Lx, int Ly);
int orc_filmpressure_raw(double *p, const double *h, double *dgrad, double gamma, double theta,
                         const double *theta_field, int n, int m, double hmin, double hcrit,
                         int Lx, int Ly);
void orc_grad(double *outx, double *outy, const double *f, const double *a, int Lx, int Ly);
void orc_laplacian(double *out, const double *f, double gamma, int Lx, int Ly);
void orc_hgradp(double *hgpx, double *hgpy, const double *p, const double *h, double *dgrad,
                int Lx, int Ly);
void orc_slippage(double *slipx, double *slipy, const double *h, const double *ux,
                  const double *uy, double delta, double mu, int Lx, int Ly);
void orc_thermal(double *kx, double *ky, const double *h, double kbt, double mu, double delta,
                 uint64_t seed, uint64_t step, int Lx, int Ly);
void orc_inclination(double *Fx, double *Fy, const double *h, double ax, double ay,
                     double fac, int Lx, int Ly);
void orc_force_sum(double *Fx, double *Fy, const double *hgpx, const double *hgpy,
                   const double *slipx, const double *slipy, const double *kbtx,
                   const double *kbty, int Lx, int Ly);
void orc_init_uniform(double *h, int Lx, int Ly, uint64_t seed, double base, double amp);
void orc_init_normal(double *h, int Lx, int Ly, uint64_t seed, double base, double amp);

/* ---- driver ---- */
int orc_time_loop(orc_state *st, const orc_params *p, int64_t nsteps, const orc_loop_opts *o);

/* ---- one dimension: D1Q3 (State_1D, src/initialize.jl:276-290); fields of L
 * nodes, distributions (L, 3), population k at offset k*L ---- */
int orc_fused_time_loop(orc_state *s, const orc_params *p, int64_t nsteps, int thermal,
                        uint64_t seed, int64_t t0);
int orc_fused_threads(void);
void orc_grad_simple(double *outx, double *outy, const double *f, int Lx, int Ly);
void orc_measure(const double *h, const double *theta_field, double theta, int Lx, int Ly,
                 double thresh, double out[4], unsigned char *mask);
void orc_t0(double hb, double gamma, double mu, double theta, double out[2]);

typedef struct orc1d_state {
    int32_t L;
    double *fout, *ftemp, *feq;
    double *height, *vel, *pressure, *F, *slip, *hgp, *kbt;
} orc1d_state;

typedef struct orc1d_loop_opts {
    const double *theta_field; /* time_loop(sys, state, θ) with a θ vector */
    int32_t thermal;           /* F -= thermal! noise each step */
    int32_t body;              /* inclination!(α, state) callback */
    double alpha, body_fac;
    const double *border1, *border2; /* bounce-back masks (SysConstWithBound_1D), or NULL */
    uint64_t seed;
    int64_t t0;
    double *dh_series;
    int64_t *wet_series;
    double wet_thresh;
    double *mass_series;
    const double *gamma_field; /* γ field in filmpressure! (run_gamma) */
    const double *grad_gamma;  /* ∇γ field, applied to F with grad_gamma_sign */
    int32_t grad_gamma_sign;   /* -1: F .-= ∇γ, +1: F .+= ∇γ */
    const double *interior;    /* wall interior mask for the mass, or NULL */
} orc1d_loop_opts;

void orc1d_equilibrium(double *feq, const double *h, const double *u, double g, int L);
void orc1d_bgk_and_stream(double *fout, const double *feq, double *ftemp, const double *F,
                          double tau, const double *border1, const double *border2, int L);
void orc1d_moments(double *h, double *u, const double *f, int L);
void orc1d_filmpressure(double *p, const double *h, double gamma, const double *gamma_field,
                        double theta, const double *theta_field, int n, int m, double hmin,
                        double hcrit, int L);
void orc1d_grad_gamma(double *out, const double *gam, const double *h, double delta, int form,
                      int L);
int orc1d_filmpressure_raw(double *p, const double *h, double gamma, double theta, int n, int m,
                           double hmin, double hcrit, int L);
void orc1d_grad(double *out, const double *f, const double *a, int L);
void orc1d_laplacian(double *out, const double *f, int L);
void orc1d_hgradp(double *hgp, const double *p, const double *h, int L);
void orc1d_slippage(double *slip, const double *h, const double *u, double delta, double mu,
                    int L);
void orc1d_thermal(double *fluc, const double *h, double kbt, double mu, double delta,
                   uint64_t seed, uint64_t step, int L);
void orc1d_inclination(double *F, const double *h, double alpha, double fac, int L);
void orc1d_force_sum(double *F, const double *hgp, const double *slip, const double *kbt, int L);
int orc1d_time_loop(orc1d_state *st, const orc_params *p, int64_t nsteps,
                    const orc1d_loop_opts *o);

#ifdef __cplusplus
}
#endif
#endif
