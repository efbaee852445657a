"""Batched FP32 films on the GPU: each film of a batch gives the same bits as
the same film stepped alone (single-film fused loop), whatever the mix, the
admission time, the graph capture and the arena placement; films agree
with the FP64 oracle to the FP32 tolerance; the serving loop returns the
alone results."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import swalbe
    if swalbe.device_count() < 1:
        pytest.skip("no GPU")
    return swalbe


def relL2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def caps_host(Lx, Ly, fi):
    """Host restatement of the cap initial height (src/initialvalues.jl:203-224,
    two caps as cap1 + cap2 - precursor, :277-323), 1-based centres."""
    i = np.arange(1, Lx + 1)[:, None]
    j = np.arange(1, Ly + 1)[None, :]

    def cap(R, th, cx, cy):
        r = np.sqrt((i - cx) ** 2 + (j - cy) ** 2)
        s = np.minimum(r / R, 1.0)
        h = np.where(r <= R, (np.sqrt(1 - s * s) - math.cos(math.pi * th)) * R, fi.precursor)
        return np.where(h < 0, fi.precursor, h)

    h = cap(fi.r1, fi.theta1, fi.c1x, fi.c1y)
    if fi.kind == 3:
        h = h + cap(fi.r2, fi.theta2, fi.c2x, fi.c2y) - fi.precursor
        h = np.where(h < 0, fi.precursor, h)
    if fi.clamp:
        h = np.maximum(h, fi.precursor)
    return h


def film_specs(S):
    """(Lx, Ly, Taumucs, FilmInit, nsteps): a mix of sizes and features."""
    T = S.Taumucs
    fi = S.film_init
    specs = [
        (256, 128, T(Tmax=1, γ=0.01, θ=1 / 9), fi(S.INIT_TWO_CAPS, precursor=0.1, href=0.1,
                                                 cap1=(40.0, 1 / 9, 114.8, 64.5),
                                                 cap2=(40.0, 1 / 9, 142.2, 64.5),
                                                 clamp=True), 300),
        (200, 100, T(Tmax=1, kbt=1e-6, γ=0.01), fi(S.INIT_NORMAL, thermal=True, seed=11, base=1.0,
                                                   amp=1e-3, href=1.0), 300),
        (199, 61, T(Tmax=1, kbt=1e-5, γ=0.02), fi(S.INIT_NORMAL, thermal=True, seed=12, base=1.0,
                                                  amp=1e-3, href=1.0), 200),
        (40, 20, T(Tmax=1, kbt=1e-6), fi(S.INIT_NORMAL, thermal=True, seed=13, base=1.0,
                                         amp=1e-3, href=1.0), 300),
        (192, 96, T(Tmax=1, τ=0.8, γ=0.01, θ=1 / 6), fi(S.INIT_CAP, precursor=0.1, href=0.1,
                                                       cap1=(30.0, 1 / 6, 96.5, 48.5)), 250),
        (256, 96, T(Tmax=1, γ=0.01), fi(S.INIT_NORMAL, seed=14, base=1.0, amp=1e-3, href=1.0,
                                        body=(2e-5, -1e-5, 1.0)), 300),
        (64, 64, T(Tmax=1, kbt=1e-6, γ=0.005), fi(S.INIT_CAP, thermal=True, seed=15,
                                                  precursor=0.1, href=0.1, clamp=True,
                                                  cap1=(20.0, 1 / 12, 32.5, 32.5)), 100),
        (130, 22, T(Tmax=1, γ=0.01), fi(S.INIT_FLAT, base=0.5, href=0.5), 50),
        # three by two nodes: every neighbour of every node wraps
        (3, 2, T(Tmax=1, kbt=1e-6, γ=0.01), fi(S.INIT_NORMAL, thermal=True, seed=16, base=1.0,
                                                amp=1e-3, href=1.0), 40),
    ]
    return specs


FIELDS = ("height", "velx", "vely", "pressure", "fout")


def alone(S, Lx, Ly, prm, init, nsteps):
    from swalbe import serve
    r = serve.Request(0, Lx, Ly, prm, init, nsteps, fields=FIELDS)
    return serve.solve_alone(r, "fused")


def assert_same(got, ref, what):
    for name in FIELDS:
        a, b = got[name], ref[name]
        assert a.shape == b.shape, (what, name)
        bad = np.flatnonzero(a.ravel(order="F") != b.ravel(order="F"))
        assert bad.size == 0, f"{what}: {name} differs at {bad.size} entries, first {bad[:5]}"


def test_initial_height_matches_host_formula(S):
    for Lx, Ly, prm, init, _ in film_specs(S):
        if init.kind not in (S.INIT_CAP, S.INIT_TWO_CAPS):
            continue
        b = S.Batch(64 << 20, 8)
        slot = b.add(prm, Lx, Ly, init, 0)
        h = b.get(slot, "height")
        b.close()
        ref = caps_host(Lx, Ly, init)
        assert np.max(np.abs(h - ref)) <= 4e-6 * max(1.0, ref.max()), (Lx, Ly)


def test_batch_matches_each_film_alone_bitwise(S):
    specs = film_specs(S)
    b = S.Batch(1 << 30, 64)
    slots = [b.add(prm, Lx, Ly, init, n) for Lx, Ly, prm, init, n in specs]
    b.step(max(s[4] for s in specs) + 7)  # films stop at their own budgets
    for slot, (Lx, Ly, prm, init, n) in zip(slots, specs):
        print(f"[batch] film {Lx}x{Ly} tau {prm.τ} steps {n}", flush=True)
        assert b.status(slot) == (n, True)
        got = {name: b.get(slot, name) for name in FIELDS}
        assert_same(got, alone(S, Lx, Ly, prm, init, n), f"{Lx}x{Ly}")
    b.close()


def test_large_film_alone_matches_batch(S):
    """A film with more 64 x 16 tiles than five per CU hold at once (2048 x
    1024: 2048 tiles) steps alone with p and the (Fx, Fy) pairs sharing one
    LDS buffer; the batch kernel keeps the two apart.  Both give the same
    bits: τ = 1 with noise and τ = 0.8, alone with every output written each
    step and with the final outputs only.  1500 x 1031 (1560 tiles) adds
    partial tiles on the right and top edges and an odd Ly (one noise call
    per node)."""
    from swalbe import serve
    T, fi = S.Taumucs, S.film_init
    specs = [
        (2048, 1024, T(Tmax=1, kbt=1e-6, γ=0.01), fi(S.INIT_NORMAL, thermal=True, seed=21,
                                                     base=1.0, amp=1e-3, href=1.0), 40),
        (2048, 1024, T(Tmax=1, τ=0.8, γ=0.01, θ=1 / 9), fi(S.INIT_CAP, precursor=0.1, href=0.1,
                                                          cap1=(150.0, 1 / 9, 1024.5, 512.5)),
         30),
        (1500, 1031, T(Tmax=1, kbt=1e-6, γ=0.01), fi(S.INIT_NORMAL, thermal=True, seed=23,
                                                     base=1.0, amp=1e-3, href=1.0), 25),
    ]
    b = S.Batch(1 << 30, 8)
    slots = [b.add(prm, Lx, Ly, init, n) for Lx, Ly, prm, init, n in specs]
    b.step(max(s[4] for s in specs))
    for slot, (Lx, Ly, prm, init, n) in zip(slots, specs):
        got = {name: b.get(slot, name) for name in FIELDS}
        assert np.isfinite(got["height"]).all()
        for last_only in (False, True):
            print(f"[large] film {Lx}x{Ly} tau {prm.τ} final_outputs_only {last_only}", flush=True)
            r = serve.Request(0, Lx, Ly, prm, init, n, fields=FIELDS)
            ref = serve.solve_alone(r, "fused", final_outputs_only=last_only)
            assert_same(got, ref, f"large {Lx}x{Ly} tau {prm.τ} last_only {last_only}")
    b.close()


def test_final_outputs_only_batch_matches_alone(S):
    """Films admitted with final_outputs_only write distributions and pressure
    on their last step only: every field at the end is the same bits as the
    film alone; reading FOUT / PRESSURE of a running τ = 1 film is an error."""
    specs = film_specs(S)
    b = S.Batch(1 << 30, 64, final_outputs_only=True)
    slots = [b.add(prm, Lx, Ly, init, n) for Lx, Ly, prm, init, n in specs]
    b.step(20)
    flat = next(i for i, sp in enumerate(specs) if sp[2].τ == 1.0 and sp[4] > 20)
    with pytest.raises(S.SwalbeError):
        b.get(slots[flat], "fout")
    assert b.get(slots[flat], "height").shape == (specs[flat][0], specs[flat][1])
    b.step(max(s[4] for s in specs))
    for slot, (Lx, Ly, prm, init, n) in zip(slots, specs):
        got = {name: b.get(slot, name) for name in FIELDS}
        assert_same(got, alone(S, Lx, Ly, prm, init, n), f"final-only {Lx}x{Ly}")
    b.close()


def test_staggered_admission_and_slot_reuse(S):
    """Films admitted after others have run, into memory freed by a retired
    film, still give their alone bits; graph replays and single launches mix."""
    specs = film_specs(S)
    b = S.Batch(96 << 20, 16, graph_steps=32)
    A = b.add(*_args(specs[0]))
    b.step(70)
    B = b.add(*_args(specs[1]))
    C_ = b.add(*_args(specs[3]))
    b.step(100)
    D = b.add(*_args(specs[4]))
    b.step(40)
    assert b.status(C_) == (140, False)
    b.step(200)
    # A: 300 steps, done at launch 300; B, C: 300 steps from launch 70
    for slot, sp in ((A, specs[0]), (B, specs[1]), (C_, specs[3]), (D, specs[4])):
        Lx, Ly, prm, init, n = sp
        assert b.status(slot)[0] == min(n, {A: 410, B: 340, C_: 340, D: 240}[slot])
    got_a = {name: b.get(A, name) for name in FIELDS}
    assert_same(got_a, alone(S, *specs[0]), "A")
    b.remove(A)
    E = b.add(*_args(specs[2]))  # reuses A's memory while B, C, D keep running
    b.step(300)
    for slot, sp, what in ((B, specs[1], "B"), (C_, specs[3], "C"), (D, specs[4], "D"),
                           (E, specs[2], "E")):
        got = {name: b.get(slot, name) for name in FIELDS}
        assert_same(got, alone(S, *sp), what)
    st = b.stats()
    assert st["films"] == 4
    b.close()


def test_kernel_class_switches_with_tau(S):
    """Batches whose films all have τ = 1 run a kernel without the f_t path
    (one more workgroup per CU); admitting a τ ≠ 1 film switches to the full
    kernel and re-captures the graph, removing it switches back.  Every film
    keeps its alone bits across both switches."""
    specs = film_specs(S)
    tau1, taux = specs[0], specs[4]
    assert tau1[2].τ == 1.0 and taux[2].τ != 1.0
    b = S.Batch(96 << 20, 8, graph_steps=16)
    A = b.add(*_args(tau1))             # 300 steps from launch 0
    b.step(40)
    D = b.add(*_args(taux))             # 250 steps from launch 40: the full kernel
    b.step(260)
    assert b.status(D) == (250, True)
    got_d = {name: b.get(D, name) for name in FIELDS}
    b.remove(D)                         # τ = 1 only again
    E = b.add(*_args(specs[5]))         # 300 steps from launch 300
    b.step(310)
    assert b.status(A) == (300, True) and b.status(E) == (300, True)
    assert_same(got_d, alone(S, *taux), "tau 0.8 film")
    for slot, sp, what in ((A, tau1, "A"), (E, specs[5], "E")):
        got = {name: b.get(slot, name) for name in FIELDS}
        assert_same(got, alone(S, *sp), what)
    b.close()


def _args(sp):
    Lx, Ly, prm, init, n = sp
    return prm, Lx, Ly, init, n


def test_graph_and_plain_launches_agree(S):
    specs = film_specs(S)[:4]
    outs = []
    for g in (0, 16, 64):
        b = S.Batch(256 << 20, 8, graph_steps=g)
        slots = [b.add(*_args(sp)) for sp in specs]
        for n in (5, 64, 100, 131):
            b.step(n)
        outs.append([b.get(s, "height") for s in slots])
        b.close()
    for o in outs[1:]:
        for a, ref in zip(o, outs[0]):
            assert np.array_equal(a, ref)


def test_tile_heights_and_autotune_give_the_alone_bits(S):
    """Films in 8-row tiles, 16-row tiles and tuned tiles (mixed in one
    launch) are the same bits as each film alone; the tuner picks 8 or 16
    for a shape, once, and reports both timings."""
    specs = film_specs(S) + [(130, 20, S.Taumucs(Tmax=1, kbt=1e-6), S.film_init(
        S.INIT_NORMAL, thermal=True, seed=16, base=1.0, amp=1e-3, href=1.0), 80)]
    b = S.Batch(1 << 30, 64, graph_steps=16)
    slots = []
    for i, sp in enumerate(specs):
        if i % 3 == 0:
            b.set_tile_y(8)
            b.set_autotune(False)
        elif i % 3 == 1:
            b.set_tile_y(16)
        else:
            b.set_tile_y(0)
            b.set_autotune(True)
        slots.append(b.add(*_args(sp)))
    for i, (slot, sp) in enumerate(zip(slots, specs)):
        ty = b.tile_y(slot)
        assert ty in (8, 16)
        if i % 3 == 0:
            assert ty == 8
        elif i % 3 == 1:
            assert ty == 16
        else:
            t = S.tune_shape(sp[2], sp[0], sp[1], sp[3])
            assert t["tile_y"] == ty and t["ms_rows8"] > 0 and t["ms_rows16"] > 0
            assert S.tune_shape(sp[2], sp[0], sp[1], sp[3]) == t  # cached
    b.step(max(s[4] for s in specs) + 3)
    for slot, (Lx, Ly, prm, init, n) in zip(slots, specs):
        print(f"[batch] tile rows {b.tile_y(slot)} film {Lx}x{Ly} steps {n}", flush=True)
        got = {name: b.get(slot, name) for name in FIELDS}
        assert_same(got, alone(S, Lx, Ly, prm, init, n), f"rows {b.tile_y(slot)} {Lx}x{Ly}")
    with pytest.raises(S.SwalbeError):
        b.set_tile_y(12)
    b.close()


def test_nontemporal_modes_give_the_same_bits(S):
    """The store mode is a cache hint only: every mode gives the same bits,
    also when it changes between runs of a captured graph (the graph is
    captured again with the new mode)."""
    specs = film_specs(S)[:4]
    outs = []
    for mode in (0, 1, 2):
        b = S.Batch(256 << 20, 8, graph_steps=16)
        slots = [b.add(*_args(sp)) for sp in specs]
        b.step(40)
        b.set_nontemporal(mode)
        b.step(57)
        outs.append([{k: b.get(s, k) for k in ("height", "velx", "fout")} for s in slots])
        with pytest.raises(S.SwalbeError):
            b.set_nontemporal(3)
        b.close()
    for o in outs[1:]:
        for a, ref in zip(o, outs[0]):
            for k in ref:
                assert np.array_equal(a[k], ref[k]), k


def test_batch_film_vs_fp64_oracle(S):
    """FP32 batch film against the FP64 oracle from the same initial height:
    rel-L2 of h <= 1e-5 after 400 steps; mass to 1e-6."""
    Lx, Ly, prm, init, _ = film_specs(S)[0]
    b = S.Batch(64 << 20, 4)
    slot = b.add(prm, Lx, Ly, init, 400)
    h0 = b.get(slot, "height")
    b.step(400)
    h = b.get(slot, "height")
    b.close()
    ost = O.State(Lx, Ly)
    ost.height[...] = h0
    ost.feq[...], _ = O.equilibrium(ost.height, ost.velx, ost.vely, 0.0)
    p = O.params(gamma=prm.γ, theta=prm.θ, tdump=10 ** 9)
    O.time_loop(ost, p, 400)
    e = relL2(h, ost.height)
    print(f"[batch] fp32 vs fp64 oracle rel-L2 h = {e:.3e}", flush=True)
    assert e <= 1e-5
    assert abs(h.sum() - ost.height.sum()) <= 1e-6 * ost.height.sum()


def test_config4_full_batch(S):
    """Config 4 at full size: the 256 coalescence films of the batch benchmark
    (Lx in {128, ..., 1024} step 64, Ly = Lx/2, two tangent caps, per-film
    θ, γ and radius; 57.2 M nodes), graph-replayed for 300 steps with every
    output written.  Every film stays finite and keeps its mass to 1e-5;
    four films (the first, two from the middle, the last) give the bits of
    the same film solved alone."""
    from dataclasses import replace
    from swalbe import serve
    nsteps = 300
    reqs = serve.coalescence_sweep(256, seed=1, nsteps=nsteps)
    cap = sum(S.film_bytes(r.Lx, r.Ly) for r in reqs)
    print(f"[batch] config 4: 256 films, {sum(r.Lx * r.Ly for r in reqs)} nodes", flush=True)
    b = S.Batch(cap, 256, graph_steps=100)
    slots = [b.add(r.param, r.Lx, r.Ly, r.init, r.nsteps) for r in reqs]
    m0 = [float(b.get(s, "height").sum()) for s in slots]
    b.step(nsteps)
    for k, s in enumerate(slots):
        assert b.status(s) == (nsteps, True), k
        h = b.get(s, "height")
        assert np.isfinite(h).all(), k
        assert abs(float(h.sum()) - m0[k]) <= 1e-5 * m0[k], k
    for k in (0, 85, 170, 255):
        r = reqs[k]
        print(f"[batch] config 4 film {k} ({r.Lx}x{r.Ly}) against the film alone", flush=True)
        got = {name: b.get(slots[k], name) for name in FIELDS}
        ref = serve.solve_alone(replace(r, fields=FIELDS), "fused")
        assert_same(got, ref, f"config 4 film {k} {r.Lx}x{r.Ly}")
    b.close()


def test_batch_errors(S):
    from swalbe._lib import SwalbeError
    b = S.Batch(1 << 20, 2)
    prm = S.Taumucs(Tmax=1)
    fi = S.film_init(S.INIT_FLAT, base=1.0, href=1.0)
    with pytest.raises(SwalbeError, match="no room"):
        b.add(prm, 1024, 1024, fi, 10)
    s0 = b.add(prm, 32, 16, fi, 10)
    b.add(prm, 32, 16, fi, 10)
    with pytest.raises(SwalbeError, match="slot"):
        b.add(prm, 32, 16, fi, 10)
    b.remove(s0)
    with pytest.raises(SwalbeError):
        b.remove(s0)
    b.step(3)
    assert b.stats()["films"] == 1
    b.close()


def test_serving_results_equal_alone(S):
    """Continuous batching with staggered arrivals: every served result is
    bit-identical to the same request solved alone."""
    from swalbe import serve
    reqs = serve.poisson_requests(12, rate=2000.0, seed=9, lx_range=(128, 320),
                                  steps=(100, 300, 600))
    be = serve.BatchBackend(1 << 30, 64, graph_steps=50)
    srv = serve.Server(be, tick=120)
    done = srv.serve(reqs)
    be.close()
    assert len(done) == len(reqs)
    assert srv.max_resident > 1
    for r in done:
        ref = serve.solve_alone(r, "fused")
        assert np.array_equal(r.result["height"], ref["height"]), r.rid


def test_native_server_results_equal_alone(S):
    """The library's own scheduler (swalbe_serve_*) gives the alone results."""
    from swalbe import serve
    reqs = serve.poisson_requests(10, rate=2000.0, seed=21, lx_range=(128, 256),
                                  steps=(50, 200, 400))
    srv = serve.NativeServer(1 << 30, 64, tick=90, graph_steps=32)
    done = srv.serve(reqs)
    srv.close()
    assert sorted(r.rid for r in done) == list(range(len(reqs)))
    assert srv.max_resident > 1
    for r in done:
        ref = serve.solve_alone(r, "fused")
        assert np.array_equal(r.result["height"], ref["height"]), r.rid


def test_serving_config5_shapes(S):
    """Serving at the full sizes of the Poisson serving workload: 2048 x 1024
    requests at T = 10^4 and 5 10^4, a 2048 x 2048 one at 10^3 and small ones,
    seeded θ/γ/radius, arriving while others run.  Each served height is the
    bits of the same request solved alone; a 5 % sample (here the 5 10^4-step
    2048 x 1024 request) against the FP64 solve: mass, maximum height and
    the pointwise rel-L2 within the bounds stated below."""
    from swalbe import serve
    rng = np.random.default_rng(44)
    spec = [(2048, 1024, 10_000, 0.0), (128, 64, 50_000, 0.01), (2048, 1024, 50_000, 0.02),
            (640, 640, 10_000, 0.03), (2048, 2048, 1_000, 0.5)]
    reqs = [serve.coalescence_request(k, Lx, Ly, T, rng, serve.domain_seed(44, k), arrival=t,
                                      max_gamma_h=serve.MAX_GAMMA_H)
            for k, (Lx, Ly, T, t) in enumerate(spec)]
    be = serve.BatchBackend(8 << 30, 64, graph_steps=100)
    srv = serve.Server(be, tick=500)
    print("[serve] config-5 shapes: 5 requests up to 2048x2048, T up to 5e4", flush=True)
    done = srv.serve(reqs)
    be.close()
    assert sorted(r.rid for r in done) == list(range(len(reqs)))
    assert srv.max_resident >= 3
    for r in done:
        print(f"[serve] request {r.rid} {r.Lx}x{r.Ly} T={r.nsteps}: alone", flush=True)
        assert np.isfinite(r.result["height"]).all(), r.rid
        ref = serve.solve_alone(r, "fused")
        assert np.array_equal(r.result["height"], ref["height"]), r.rid
    r = [q for q in done if q.nsteps == 50_000 and q.Lx == 2048][0]
    print("[serve] FP64 solve of the sample request", flush=True)
    ref = serve.solve_alone(r, "fused", T=np.float64)["height"]
    h = r.result["height"]
    e = relL2(h, ref)
    dm = abs(h.sum() - ref.sum()) / ref.sum()
    dx = abs(h.max() - ref.max()) / ref.max()
    print(f"[serve] FP64 sample {r.Lx}x{r.Ly} T={r.nsteps}: rel-L2 h {e:.3e}, mass {dm:.2e}, "
          f"max h {dx:.2e}", flush=True)
    assert np.isfinite(h).all()
    assert dm <= 1e-3
    assert dx <= 1e-2
    assert e <= 0.5
