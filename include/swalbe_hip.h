/*
 * swalbe_hip.h -- C ABI of libswalbe_hip, the MI355X (gfx950) implementation
 * of Swalbe.jl's thin-film lattice Boltzmann time step (2D D2Q9 and 1D D1Q3).
 *
 * Plain C types only: handles, device/host pointers and sizes.  Every entry
 * point returns an int status (SWALBE_OK = 0) and never aborts; the text of
 * the last error of the calling thread is swalbe_last_error().
 *
 * Fields use the reference's layout: Julia column-major arrays, a field
 * (Lx, Ly) stores node (i, j) (0-based) at i + Lx*j, a distribution
 * (Lx, Ly, 9) stores population k at plane offset k*Lx*Ly.
 *
 * Each entry point names the reference function it replaces
 * (paths relative to the Swalbe.jl v1.0.6 source tree).
 */
#ifndef SWALBE_HIP_H
#define SWALBE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define SWALBE_OK 0
#define SWALBE_EDOMAIN 1 /* DomainError of src/pressure.jl:101-107 */
#define SWALBE_EINVAL 2  /* bad argument (size, field id, null pointer) */
#define SWALBE_ENOMEM 3  /* device allocation failed */
#define SWALBE_EHIP 4    /* HIP runtime error */

/* ---- element types and state kinds (Sys(sysc, device; T, kind),
 *      src/initialize.jl:491-572) -------------------------------------- */
#define SWALBE_F64 0
#define SWALBE_F32 1
#define SWALBE_SIMPLE 0
#define SWALBE_THERMAL 1

/* ---- field ids: the arrays of State{T,3} (src/initialize.jl:149-168),
 *      State_thermal's kbtx/kbty (:182-186) and a contact-angle field θ
 *      (time_loop(sys, state, θ), src/simulate.jl:26-45) ---------------- */
enum swalbe_field {
    SWALBE_FOUT = 0,
    SWALBE_FTEMP = 1,
    SWALBE_FEQ = 2,
    SWALBE_HEIGHT = 3,
    SWALBE_VELX = 4,
    SWALBE_VELY = 5,
    SWALBE_VSQ = 6,
    SWALBE_PRESSURE = 7,
    SWALBE_FX = 8,
    SWALBE_FY = 9,
    SWALBE_SLIPX = 10,
    SWALBE_SLIPY = 11,
    SWALBE_HGRADPX = 12,
    SWALBE_HGRADPY = 13,
    SWALBE_KBTX = 14,
    SWALBE_KBTY = 15,
    SWALBE_THETA = 16,
    SWALBE_NFIELDS = 17
};

/* Taumucs{T} (src/initialize.jl:43-60).  μ is taken as given (the reference
 * derives it as cs^2 (τ - 1/2), :49). */
typedef struct swalbe_params {
    int64_t Tmax, tdump;
    double tau, cs, mu, delta, kbt, gamma;
    int32_t n, m;
    double hmin, hcrit, theta, g;
} swalbe_params;

typedef struct swalbe_state swalbe_state;

const char *swalbe_last_error(void);
int swalbe_version(void);
/* Number of visible GPUs (0 on a host without one). */
int swalbe_device_count(int32_t *n);

/* ---- state lifetime: Sys(sysc, "GPU"; T, kind) (src/initialize.jl:491-572).
 * Allocates every field on the current device, zeroed, with height = 1. */
int swalbe_create(int32_t Lx, int32_t Ly, int32_t dtype, int32_t kind, swalbe_state **out);
int swalbe_destroy(swalbe_state *st);
/* FP32 states store height and distributions as deviations from a film at
 * rest of height href (default 1); this re-encodes the resident fields.
 * No effect on FP64 states.  g is the gravity of the rest equilibrium. */
int swalbe_set_reference_height(swalbe_state *st, double href, double g);

/* ---- field access: reads/writes of state.<field> (host <-> device copy).
 * bytes selects the host element type: N*8 for Float64, N*4 for Float32,
 * with N = Lx*Ly (x9 for distributions).  Synchronous. */
int swalbe_upload(swalbe_state *st, int32_t field, const void *host, size_t bytes);
int swalbe_download(swalbe_state *st, int32_t field, void *host, size_t bytes);
/* the stored representation (elements x 4 bytes for FP32 states, x 8 for
 * FP64; FP32 heights and distributions as deviations from the reference
 * film): an exact save / restore of a state, e.g. for a bit-exact restart */
int swalbe_download_raw(swalbe_state *st, int32_t field, void *host, size_t bytes);
int swalbe_upload_raw(swalbe_state *st, int32_t field, const void *host, size_t bytes);
int swalbe_get_reference_height(swalbe_state *st, double *href, double *g);
/* device-side copy of one field into another of the same kind (both
 * distributions, both heights or both plain planes): the broadcast
 * assignments of the drivers, e.g. `state.ftemp .= state.feq` */
int swalbe_copy(swalbe_state *st, int32_t dst_field, int32_t src_field);
/* state.<field> .= value (src/simulate.jl:241, test/collide.jl:19-22). */
int swalbe_fill(swalbe_state *st, int32_t field, double value);
int swalbe_synchronize(swalbe_state *st);
/* Enqueue the state's work on a caller's HIP stream (a hipStream_t; NULL
 * restores the state's own stream).  Work already enqueued completes first.
 * The caller keeps the stream alive while the state uses it. */
int swalbe_set_stream(swalbe_state *st, void *stream);

/* ---- one op per reference function, on a state ------------------------ */
/* equilibrium!(state, sys) -- src/equilibrium.jl:63-128 */
int swalbe_equilibrium(swalbe_state *st, double g);
/* BGKandStream!(state, sys) -- src/collide.jl:70-105, :151-161 */
int swalbe_bgk_and_stream(swalbe_state *st, double tau);
/* moments!(state) -- src/moments.jl:43-52, :64-73 */
int swalbe_moments(swalbe_state *st);
/* filmpressure!(state, sys; θ, γ, n, m, hmin, hcrit) -- src/pressure.jl:119-193
 * (power_broad form, any n > m).  use_theta_field = 1 takes θ from the
 * resident SWALBE_THETA field.  raw = 1 selects the raw-array method
 * filmpressure!(p, h, dgrad, γ, θ, n, m, hmin, hcrit) (src/pressure.jl:72-117)
 * with fast_93 / fast_32, returning SWALBE_EDOMAIN for other (n, m). */
int swalbe_filmpressure(swalbe_state *st, double gamma, double theta, int32_t n, int32_t m,
                        double hmin, double hcrit, int32_t use_theta_field, int32_t raw);
/* h∇p!(state) -- src/forcing.jl:168-187, :200-221 */
int swalbe_hgradp(swalbe_state *st);
/* slippage!(state, sys) -- src/forcing.jl:42-66 */
int swalbe_slippage(swalbe_state *st, double delta, double mu);
/* thermal!(state, sys) -- src/forcing.jl:297-320.  randn! is replaced by
 * counter-based normals keyed (seed, step, node). */
int swalbe_thermal(swalbe_state *st, double kbt, double mu, double delta, uint64_t seed,
                   uint64_t step);
/* inclination!(α, state; t, tstart, tsmooth) -- src/forcing.jl:363-377, with
 * fac = 0.5 + 0.5 tanh((t - tstart)/tsmooth) computed by the caller. */
int swalbe_inclination(swalbe_state *st, double ax, double ay, double fac);
/* state.Fx .= -state.h∇px .- state.slipx [.- state.kbtx]
 * -- src/simulate.jl:18-19, scripts/Rivulet_stability.jl:123-124 */
int swalbe_force_sum(swalbe_state *st, int32_t with_thermal);

/* ---- fused time loop: time_loop(sys, state[, θ | Δh | f, measure])
 * -- src/simulate.jl:6-96.  One kernel per step (collide + pull-stream +
 * moments, with pressure, h∇p, slip, thermal and body forces computed from
 * an LDS halo tile).  Steps run are t0+1 .. t0+nsteps. */
typedef struct swalbe_step_opts {
    int32_t use_theta_field; /* filmpressure!(state, sys, θ = field) (src/simulate.jl:35) */
    int32_t thermal;         /* F -= thermal!(…) noise each step */
    int32_t body;            /* inclination! callback (src/simulate.jl:89, 480) */
    int32_t materialize;     /* last step also writes feq, vsq, Fx, Fy, slip, h∇p, kbt */
    uint64_t seed;
    int64_t t0;
    double body_ax, body_ay, body_fac;
    double wet_thresh;
    double *dh_series;     /* host [nsteps]: max(h) - min(h) before each step (:56), or NULL */
    int64_t *wet_series;   /* host [nsteps]: count(h > wet_thresh) (measures.jl:13-17), or NULL */
    double *mass_series;   /* host [nsteps/tdump + 1]: sum(h) when t % tdump == 0 (:8-14), or NULL */
    int32_t tile_y;        /* rows per workgroup tile: 0 = default (16 for FP32; FP64 16 where
                              h, ux, uy exceed 32 MB, else 8);
                              8 or 16 (the same results either way); variant 3: rows per
                              strip segment, 0 = default (64), 4..4096 */
    int32_t variant;       /* FP32 kernel: 0 = default (2), 1 = LDS tiles, collide at the
                              source node, 2 = LDS tiles, collide at the destination node,
                              3 = LDS-free strips (one wavefront per 58-column strip) */
    int32_t nontemporal;   /* fused-step stores, FP32 and FP64 (the same bits either way):
                              0 = cached, 1 = every output streams past the caches
                              (non-temporal), 2 = the distributions and the pressure
                              stream while h, ux, uy (the next step's input) stay cached */
    /* FP32 states: 0 (default) evaluates the disjoining pressure
     * κ [(hmin/(h+hcrit))^n - (hmin/(h+hcrit))^m] in FP64 (the stored height
     * exactly, an IEEE division) at nodes with h + hcrit < 4 hmin, where the
     * film nears hmin and the term dominates the pressure; 1 evaluates it in
     * FP32 everywhere.  Batch films always use the FP64 evaluation. */
    int32_t no_fp64_guard;
    /* snapshot!(snap, height, t; dumping) (src/measures.jl:99-105): after every
     * step t with t % snap_every == 0 the height is copied to snap_series
     * (host, one Lx*Ly column-major block per snapshot, in step order), through
     * a device ring and a second stream so the copies overlap later steps.
     * 0 / NULL = off. */
    int64_t snap_every;
    double *snap_series;
    /* measurement callbacks of time_loop(sys, state, f, measure) (src/simulate.jl:69-96)
     * taken on the device on each step's input height, copied back after the loop:
     * wetted!(area, maxheight, h, t) (src/measures.jl:6-11), fluid_dry!(fluid, dummy,
     * h, t) (:23-28), surfacearea!(area_lv, red_energy, h, θ, …, t) (:126-168).
     * NULL = off. */
    double *maxh_series;    /* host [nsteps]: maximum(h) */
    double *area_series;    /* host [nsteps]: Σ sqrt(|∇h|² + 1) over h > surf_thresh */
    double *energy_series;  /* host [nsteps]: area - Σ cos πθ over h > surf_thresh */
    uint8_t *fluid_series;  /* host [nsteps * Lx * Ly]: h > surf_thresh per step */
    double surf_thresh;     /* htresh / hthresh of those (0 = 0.055) */
    int32_t surf_theta_field; /* reduced energy from the θ field (else cos πθ of params) */
    /* 1: with τ = 1, write the distributions and the pressure on the last
     * step only (no step reads them, so every result is the same); the
     * other fields are written every step.  Ignored when τ ≠ 1. */
    int32_t final_outputs_only;
} swalbe_step_opts;
int swalbe_time_loop(swalbe_state *st, const swalbe_params *p, int64_t nsteps,
                     const swalbe_step_opts *opts);
/* Same, but enqueues only (no host series) and reports the device time of
 * the nsteps launches measured with HIP events on the state's stream. */
int swalbe_time_loop_timed(swalbe_state *st, const swalbe_params *p, int64_t nsteps,
                           const swalbe_step_opts *opts, double *ms_out);

/* ---- reductions: sum(state.height) (src/simulate.jl:10), maximum/minimum
 * (:56) and count(h .> thresh) (src/measures.jl:13-17), in FP64.
 * out = {sum, max, min, count}.  Deterministic order. */
int swalbe_reduce(swalbe_state *st, int32_t field, double thresh, double out[4]);

/* ---- measurements of the height on the device (src/measures.jl):
 * out = {count(h > thresh), maximum(h), Σ sqrt(|∇h|² + 1) over h > thresh
 * (∇f_simple!, :175-192), that area minus Σ cos πθ over h > thresh} --
 * wetted! (:6-17) and surfacearea! (:126-168); surface = 0 skips the last
 * two.  θ is the scalar `theta` or, with use_theta_field, the θ field.
 * host_mask (Lx*Ly bytes, or NULL) receives fluid_dry!'s mask (:23-28). */
int swalbe_measure(swalbe_state *st, double thresh, int32_t use_theta_field, double theta,
                   int32_t surface, double out[4], uint8_t *host_mask, size_t mask_bytes);

/* ---- seeded initial height on the device: h = base + amp * U(0,1)
 * (dist 0) or base + amp * N(0,1) (dist 1), Philox keyed (seed, node);
 * the counterpart of randinterface! (src/initialvalues.jl:23-33). */
int swalbe_init_random(swalbe_state *st, int32_t dist, uint64_t seed, double base, double amp);

/* ---- batches of independent FP32 films (no reference counterpart: the
 * reference steps one film per process; these run many films of different
 * sizes and parameters in one launch per time step).  A film is admitted
 * with its parameters, an initial height and a step budget; it advances with
 * every swalbe_batch_step until the budget is spent and keeps its state
 * until removed.  Each film's arithmetic is the single-film FP32 kernel's,
 * so a film's result is the same alone, in any mix, and via swalbe_time_loop
 * with opts.variant = 2 and opts.tile_y 8 or 16 (both heights give the same
 * bits).
 * Initial height kinds follow src/initialvalues.jl (flat, randinterface!,
 * singledroplet, two_droplets); τ = 1 films keep one distribution buffer.
 * add / remove / step only enqueue work on the batch's stream (admissions
 * never wait for the GPU); download waits for the film's last step only and
 * copies on a second stream, so it overlaps later steps.  One batch is used
 * by one host thread at a time. */
#define SWALBE_INIT_FLAT 0     /* h = base */
#define SWALBE_INIT_NORMAL 1   /* h = base + amp N(0,1) (Philox, seed) */
#define SWALBE_INIT_CAP 2      /* spherical cap r1, θ1 at (c1x, c1y) on a precursor */
#define SWALBE_INIT_TWO_CAPS 3 /* cap 1 + cap 2 - precursor */
typedef struct swalbe_film_init {
    int32_t kind;
    int32_t thermal; /* subtract thermal noise from F (kbt from the params) */
    int32_t body;    /* body force α h fac (inclination!) */
    int32_t clamp;   /* caps: 1 = raise every h below the precursor to it; 0 = the
                        reference's rule (only negative h becomes the precursor),
                        which leaves h ≪ precursor at cap rims and necks */
    uint64_t seed;   /* noise and initial-height stream */
    double base, amp, precursor, href;
    double r1, theta1, c1x, c1y; /* 1-based centres, θ in units of π */
    double r2, theta2, c2x, c2y;
    double body_ax, body_ay, body_fac;
} swalbe_film_init;
typedef struct swalbe_batch swalbe_batch;
int swalbe_batch_create(int32_t dtype, size_t capacity_bytes, int32_t max_films, swalbe_batch **out);
int swalbe_batch_destroy(swalbe_batch *b);
/* device bytes one film takes in a batch arena */
int swalbe_batch_film_bytes(int32_t Lx, int32_t Ly, double tau, size_t *bytes);
int swalbe_batch_add(swalbe_batch *b, const swalbe_params *p, int32_t Lx, int32_t Ly,
                     const swalbe_film_init *init, int64_t nsteps, int32_t *slot);
int swalbe_batch_remove(swalbe_batch *b, int32_t slot);
/* advance every resident film by nsteps (films stop at their budget) */
int swalbe_batch_step(swalbe_batch *b, int64_t nsteps);
int swalbe_batch_step_timed(swalbe_batch *b, int64_t nsteps, double *ms);
/* replay runs of steps_per_graph steps as one captured HIP graph (0 = off) */
int swalbe_batch_set_graph(swalbe_batch *b, int32_t steps_per_graph);
/* stores of the batch's films: 0, 1 or 2 as swalbe_step_opts.nontemporal (default 2) */
int swalbe_batch_set_nontemporal(swalbe_batch *b, int32_t on);
/* Autotuned tile heights: every film runs in tiles of 64 columns by 8 or 16
 * rows (16 by default).  With autotune on, a film admitted afterwards takes
 * the height measured fastest for its shape: a temporary batch of copies of
 * the film (about 4 M nodes) is timed with each height, once per device and
 * (Lx, Ly, τ ≠ 1, noise, body force), and the result cached.  Timing a new
 * shape synchronises the batch's stream first.  The heights give the same
 * bits, so only the speed changes. */
int swalbe_batch_set_autotune(swalbe_batch *b, int32_t on);
/* films admitted afterwards use tiles of `rows` rows (8 or 16); 0 restores
 * the default (16, or the tuned height with autotune on) */
int swalbe_batch_set_tile_y(swalbe_batch *b, int32_t rows);
/* tile height the tuner picks for this film shape on the current device
 * (measured now, or cached), with the milliseconds per batched step of the
 * copies at 8 and 16 rows (either pointer may be null) */
int swalbe_batch_tune_shape(const swalbe_params *p, int32_t Lx, int32_t Ly,
                            const swalbe_film_init *init, int32_t *tile_y, double *ms_rows8,
                            double *ms_rows16);
/* tile height (8 or 16) of a resident film */
int swalbe_batch_film_tile_y(swalbe_batch *b, int32_t slot, int32_t *tile_y);
/* Films admitted after this call with τ = 1 write their distributions and
 * pressure on their last step only (their heights and velocities every
 * step): the same final state for less traffic.  Downloading FOUT or
 * PRESSURE of such a film before it finishes is an error.  The native
 * server sets it (its requests return heights). */
int swalbe_batch_set_final_outputs_only(swalbe_batch *b, int32_t on);
/* block until at most `steps_behind` of the enqueued batch steps are still
 * pending on the device (bounds how far the host runs ahead) */
int swalbe_batch_wait(swalbe_batch *b, int64_t steps_behind);
int swalbe_batch_sync(swalbe_batch *b);
int swalbe_batch_film_status(swalbe_batch *b, int32_t slot, int64_t *steps_done,
                             int32_t *finished);
/* field: SWALBE_HEIGHT, VELX, VELY, PRESSURE or FOUT; bytes as swalbe_download */
int swalbe_batch_download(swalbe_batch *b, int32_t slot, int32_t field, void *host, size_t bytes);
int swalbe_batch_stats(swalbe_batch *b, int32_t *nfilms, int64_t *nodes, size_t *bytes_used,
                       size_t *bytes_free);
/* ---- serving: continuous batching of solve requests over a batch (no
 * reference counterpart; the reference runs sweeps as a serial host loop,
 * scripts/Rivulet_stability.jl:141-190).  Requests are admitted in
 * submission order while their film fits the arena and a slot is free.
 * Each poll admits, enqueues one tick of steps (the tick cut to the smallest
 * remaining budget) and hands back films that finished in an earlier tick;
 * finished heights wait on the host until taken (peek, then take). */
typedef struct swalbe_server swalbe_server;
int swalbe_serve_create(size_t capacity_bytes, int32_t max_films, int32_t tick_steps,
                        int32_t graph_steps, swalbe_server **out);
int swalbe_serve_destroy(swalbe_server *s);
int swalbe_serve_submit(swalbe_server *s, const swalbe_params *p, int32_t Lx, int32_t Ly,
                        const swalbe_film_init *init, int64_t nsteps, int64_t *request_id);
int swalbe_serve_poll(swalbe_server *s, int32_t *n_done, int32_t *n_resident, int32_t *n_queued);
/* id and size of the oldest finished request */
int swalbe_serve_peek(swalbe_server *s, int64_t *request_id, int32_t *Lx, int32_t *Ly);
/* its final height, Lx*Ly doubles column-major; removes it */
int swalbe_serve_take(swalbe_server *s, double *height, size_t bytes);

/* initial height of a single-film state from the same spec as a batch film
 * (singledroplet / two_droplets / randinterface!, src/initialvalues.jl) */
int swalbe_init_height(swalbe_state *st, const swalbe_film_init *init);
/* free and total device memory (admission by HBM footprint) */
int swalbe_device_memory(size_t *free_bytes, size_t *total_bytes);


/* ======================================================================
 * One-dimensional films (D1Q3, velocities c = 0, +1, -1).
 * ====================================================================== */

/* Fields of a 1D film: the arrays of State_1D (src/initialize.jl:276-290),
 * kbt of State_thermal_1D (:338-341), γ and ∇γ of State_gamma_1D (:304-308),
 * θ as time_loop(sys, state, θ) takes it (src/simulate.jl:118), the wall
 * masks of SysConstWithBound_1D (:112-120; border1 = obsright, border2 =
 * obsleft of obslist1D, src/obstacle.jl:6-35) and its interior mask.
 * Distributions are (L, 3) column-major: population k at offset k*L. */
enum swalbe1d_field {
    SWALBE1D_FOUT = 0,
    SWALBE1D_FTEMP = 1,
    SWALBE1D_FEQ = 2,
    SWALBE1D_HEIGHT = 3,
    SWALBE1D_VEL = 4,
    SWALBE1D_PRESSURE = 5,
    SWALBE1D_F = 6,
    SWALBE1D_SLIP = 7,
    SWALBE1D_HGRADP = 8,
    SWALBE1D_KBT = 9,
    SWALBE1D_THETA = 10,     /* optional fields: allocated on first upload / fill */
    SWALBE1D_GAMMA = 11,
    SWALBE1D_GRADGAMMA = 12,
    SWALBE1D_BORDER1 = 13,
    SWALBE1D_BORDER2 = 14,
    SWALBE1D_INTERIOR = 15,
    SWALBE1D_NFIELDS = 16
};

typedef struct swalbe_film1d swalbe_film1d;

/* Sys(sysc::Consts_1D; T, kind) (src/initialize.jl:587-616): all fields zero
 * except height = 1.  kind SWALBE_SIMPLE or SWALBE_THERMAL ("gamma" and
 * "gamma_bound" are a simple film plus the optional γ / wall fields).
 * L >= 3.  A film takes one device allocation for its fixed fields and
 * enqueues on one of a few HIP streams the library keeps per device and
 * shares between films (so thousands of live films cost no per-film runtime
 * queues); synchronising a film waits for the other films on its stream. */
int swalbe1d_create(int32_t L, int32_t dtype, int32_t kind, swalbe_film1d **out);
int swalbe1d_destroy(swalbe_film1d *st);
int swalbe1d_length(swalbe_film1d *st, int32_t *L);
/* bytes = n * 8 (host doubles) or n * 4 (host floats), n = L or 3L */
int swalbe1d_upload(swalbe_film1d *st, int32_t field, const void *host, size_t bytes);
int swalbe1d_download(swalbe_film1d *st, int32_t field, void *host, size_t bytes);
int swalbe1d_fill(swalbe_film1d *st, int32_t field, double value);
int swalbe1d_synchronize(swalbe_film1d *st);

/* equilibrium!(state::State_1D, sys; g) -- src/equilibrium.jl:169-191 */
int swalbe1d_equilibrium(swalbe_film1d *st, double g);
/* BGKandStream!(state, sys) -- src/collide.jl:179-212; bound = 1 is the
 * bounce-back method of StateWithBound_1D (:214-249) with BORDER1/BORDER2 */
int swalbe1d_bgk_and_stream(swalbe_film1d *st, double tau, int32_t bound);
/* moments!(state) -- src/moments.jl:54-76 */
int swalbe1d_moments(swalbe_film1d *st);
/* filmpressure!(state, sys; θ, n, m, hmin, hcrit, γ) -- src/pressure.jl:231-315,
 * raw = 1 the Vector method (:196-229, fast_93 / fast_32, SWALBE_EDOMAIN
 * otherwise).  use_theta_field / use_gamma_field take θ / γ from the
 * resident THETA / GAMMA fields (θ and γ broadcast as vectors). */
int swalbe1d_filmpressure(swalbe_film1d *st, double gamma, double theta, int32_t n, int32_t m,
                          double hmin, double hcrit, int32_t use_theta_field,
                          int32_t use_gamma_field, int32_t raw);
/* ∇γ!(state) (form 0) / ∇γ!(state, sys) (form 1) -- src/forcing.jl:423-447:
 * GRADGAMMA from GAMMA (and the height, form 1) */
int swalbe1d_grad_gamma(swalbe_film1d *st, int32_t form, double delta);
/* h∇p!(state) -- src/forcing.jl:189-198, 223-232 */
int swalbe1d_hgradp(swalbe_film1d *st);
/* slippage!(state, sys) -- src/forcing.jl:68-82 */
int swalbe1d_slippage(swalbe_film1d *st, double delta, double mu);
/* thermal!(state, sys) -- src/forcing.jl:322-335, randn! replaced by the
 * counter-based normal of (seed, step, node x) */
int swalbe1d_thermal(swalbe_film1d *st, double kbt, double mu, double delta, uint64_t seed,
                     uint64_t step);
/* inclination!(α, state; t, tstart, tsmooth) -- src/forcing.jl:379-389,
 * fac = 0.5 + 0.5 tanh((t - tstart)/tsmooth) */
int swalbe1d_inclination(swalbe_film1d *st, double alpha, double fac);
/* F .= -h∇p .- slip [.- kbt] [.- ∇γ | .+ ∇γ] (src/simulate.jl:111, 548;
 * analyze/droplet_coalescence.jl:148 adds ∇γ): grad_gamma -1, 0 or +1 */
int swalbe1d_force_sum(swalbe_film1d *st, int32_t with_thermal, int32_t grad_gamma);
/* {Σ, max, min, count(> thresh)} of a field in FP64; use_interior: Σ(h - interior h)
 * as the walled loop's mass (src/simulate.jl:189) */
int swalbe1d_reduce(swalbe_film1d *st, int32_t field, double thresh, int32_t use_interior,
                    double out[4]);

/* ---- fused 1D time loop: time_loop(sys::SysConst_1D, state[, θ | Δh | f,
 * measure]) and the walled loop (src/simulate.jl:98-205), run_gamma's loop
 * (:518-560).  Steps run are t0+1 .. t0+nsteps. */
typedef struct swalbe1d_step_opts {
    int32_t use_theta_field; /* θ from THETA (src/simulate.jl:118-136) */
    int32_t use_gamma_field; /* γ from GAMMA (filmpressure!(state, sys, γ = gamma), :545) */
    int32_t grad_gamma;      /* F ∓ ∇γ from GRADGAMMA: -1 run_gamma (:548), +1 the scans, 0 off */
    int32_t thermal;         /* F .-= thermal! noise each step */
    int32_t body;            /* inclination!(α, state) after the force sum */
    int32_t bound;           /* bounce-back walls (BORDER1/2); mass Σ(h - interior h) */
    int32_t materialize;     /* afterwards pressure, h∇p, slip, kbt, F, feq hold the last step's */
    int32_t variant;         /* 0 = resident when the film fits in LDS, else one launch per
                                step; 1 = one launch per step; 2 = resident (error if too big) */
    uint64_t seed;
    int64_t t0;
    double alpha, body_fac;
    double wet_thresh;
    double *dh_series;     /* host [nsteps]: max(h) - min(h) before each step (:147), or NULL */
    int64_t *wet_series;   /* host [nsteps]: count(h > wet_thresh) before each step, or NULL */
    double *mass_series;   /* host: Σh before each step t with t % tdump == 0 (:100-105) */
    int64_t snap_every;    /* snapshot!(fluid, h, t; dumping) (src/measures.jl:99-105) */
    double *snap_series;   /* host [nsnap * L], in step order, or NULL */
    int32_t threads;       /* resident kernel: threads per film, 0 (by L), 256, 512 or 1024;
                              an ensemble uses opts[0].threads */
    int32_t pad_;
} swalbe1d_step_opts;
int swalbe1d_time_loop(swalbe_film1d *st, const swalbe_params *p, int64_t nsteps,
                       const swalbe1d_step_opts *opts);
/* Same, reporting the device time of the loop (HIP events on the film's stream). */
int swalbe1d_time_loop_timed(swalbe_film1d *st, const swalbe_params *p, int64_t nsteps,
                             const swalbe1d_step_opts *opts, double *ms_out);
/* Many 1D films stepped together in one launch, one workgroup per film,
 * each with its own parameters and options (params[i], opts[i]; opts may be
 * NULL).  No reference counterpart: the coalescence scans
 * (analyze/droplet_coalescence.jl) loop one film at a time.  Each film's
 * result equals its swalbe1d_time_loop with variant 2.  All films share a
 * dtype and a device and each must fit in LDS. */
int swalbe1d_ensemble_time_loop(swalbe_film1d *const *films, int32_t nfilms,
                                const swalbe_params *params, int64_t nsteps,
                                const swalbe1d_step_opts *opts);
/* Same, reporting the device time of the launch (HIP events). */
int swalbe1d_ensemble_time_loop_timed(swalbe_film1d *const *films, int32_t nfilms,
                                      const swalbe_params *params, int64_t nsteps,
                                      const swalbe1d_step_opts *opts, double *ms_out);
/* Largest L that runs resident (whole film in LDS) on the current device. */
int swalbe1d_resident_max_length(int32_t dtype, int32_t tau_is_one, int32_t *L);
/* ∇f!(output::Vector, f, dgrad[, a]) (src/differences.jl:208-228) and
 * ∇²f!(output, f::Vector, dgrad) (:77-85) on caller-owned device vectors */
int swalbe1d_grad_ptr(int32_t dtype, void *out, const void *f, const void *a, int32_t L,
                      void *stream);
int swalbe1d_laplacian_ptr(int32_t dtype, void *out, const void *f, int32_t L, void *stream);

/* ---- raw-array twins on caller-owned device arrays (plain values of the
 * given dtype, Julia layout), enqueued on `stream` (NULL = default stream).
 * These replace the positional-array methods the reference exposes. */
int swalbe_equilibrium_ptr(int32_t dtype, void *feq, const void *h, const void *ux,
                           const void *uy, void *vsq, double g, int32_t Lx, int32_t Ly,
                           void *stream); /* src/equilibrium.jl:63-116 */
int swalbe_bgk_and_stream_ptr(int32_t dtype, void *fout, const void *feq, void *ftemp,
                              const void *Fx, const void *Fy, double tau, int32_t Lx,
                              int32_t Ly, void *stream); /* src/collide.jl:70-105 */
int swalbe_moments_ptr(int32_t dtype, void *h, void *ux, void *uy, const void *fout, int32_t Lx,
                       int32_t Ly, void *stream); /* src/moments.jl:43-52 */
int swalbe_filmpressure_ptr(int32_t dtype, void *p, const void *h, double gamma, double theta,
                            const void *theta_field, int32_t n, int32_t m, double hmin,
                            double hcrit, int32_t raw, int32_t Lx, int32_t Ly,
                            void *stream); /* src/pressure.jl:72-155 */
int swalbe_grad_ptr(int32_t dtype, void *outx, void *outy, const void *f, const void *a,
                    int32_t Lx, int32_t Ly, void *stream); /* ∇f! src/differences.jl:153-206 */
int swalbe_laplacian_ptr(int32_t dtype, void *out, const void *f, double gamma, int32_t Lx,
                         int32_t Ly, void *stream); /* ∇²f! src/differences.jl:57-75 */
int swalbe_hgradp_ptr(int32_t dtype, void *hgpx, void *hgpy, const void *p, const void *h,
                      int32_t Lx, int32_t Ly, void *stream); /* src/forcing.jl:168-187 */
int swalbe_slippage_ptr(int32_t dtype, void *slipx, void *slipy, const void *h, const void *ux,
                        const void *uy, double delta, double mu, int32_t Lx, int32_t Ly,
                        void *stream); /* src/forcing.jl:42-46 */
int swalbe_thermal_ptr(int32_t dtype, void *kbtx, void *kbty, const void *h, double kbt,
                       double mu, double delta, uint64_t seed, uint64_t step, int32_t Lx,
                       int32_t Ly, void *stream); /* src/forcing.jl:297-311 */
int swalbe_inclination_ptr(int32_t dtype, void *Fx, void *Fy, const void *h, double ax,
                           double ay, double fac, int32_t Lx, int32_t Ly,
                           void *stream); /* src/forcing.jl:363-368 */
/* surfacearea! / wetted! / fluid_dry! on a caller's height array (device
 * pointers, plain values): out_dev = 4 doubles as swalbe_measure; mask
 * (device, Lx*Ly bytes) may be NULL.  Synchronises `stream`. */
int swalbe_measure_ptr(int32_t dtype, const void *h, const void *theta_field, double theta,
                       int32_t Lx, int32_t Ly, double thresh, void *mask, double *out_dev,
                       void *stream);
/* The same measurement with no device allocation: no counterpart in the
 * reference.  It is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  The 512 block partials
 * live in a 16 KiB device array the library holds per GPU until the process
 * ends, so any film size fits; the call only enqueues work on `stream` (no
 * synchronisation) and out_dev holds the result once the stream reaches it.
 * Calls on one GPU share that array: calls that overlap in time (on
 * different streams or host threads) race, while calls on one stream run in
 * order. */
int swalbe_measure_ptr_noalloc(int32_t dtype, const void *h, const void *theta_field,
                               double theta, int32_t Lx, int32_t Ly, double thresh, void *mask,
                               double *out_dev, void *stream);
/* ∇f_simple!(outx, outy, f, dgrad) -- src/measures.jl:175-192 */
int swalbe_grad_simple_ptr(int32_t dtype, void *outx, void *outy, const void *f, int32_t Lx,
                           int32_t Ly, void *stream);
/* ---- host-side draws of the same counter-based streams (initial
 * conditions built on the host): node i of n gets the pair / uniform of
 * counter (seed, step, i, tag).  cospi as used for κ(θ). */
int swalbe_host_normals(uint64_t seed, uint64_t step, uint32_t tag, int64_t n, double *z0,
                        double *z1);
int swalbe_host_uniforms(uint64_t seed, uint64_t step, uint32_t tag, int64_t n, double *u);
double swalbe_host_cospi(double x);

/* device buffers for the raw-array forms */
int swalbe_malloc(void **dptr, size_t bytes);
int swalbe_free(void *dptr);
int swalbe_memcpy_h2d(void *dst, const void *src, size_t bytes);
int swalbe_memcpy_d2h(void *dst, const void *src, size_t bytes);

#ifdef __cplusplus
}
#endif
#endif /* SWALBE_HIP_H */
