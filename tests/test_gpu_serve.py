"""Batched server on the GPU: every box must compute exactly what the
single-box solver computes for the same request, whether it runs alone or in
a mix, under graph or eager launches, after queueing or next to a cancelled
box.  Bit-exact comparisons: the per-node work is shared code and each node's
update depends only on its own box."""
import numpy as np
import pytest

from m3lb_amd import abi, cases
from m3lb_amd.serve import Server
from m3lb_amd.solver import FluidSolver

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def tg(shape, steps, omega=1.7, u0=0.05, **kw):
    return abi.make_request(shape, steps, omega=omega, init=abi.INIT_TAYLOR_GREEN,
                            u=(u0, 0.0, 0.0), **kw)


def channel(shape, steps, omega=1.4, g=1e-5, **kw):
    return abi.make_request(shape, steps, omega=omega, periodic=(True, False, True),
                            walls=abi.WALLS_Y, force=(g, 0.0, 0.0), **kw)


def solo(scfg, req, flags=None):
    """The same request on the single-box solver: populations, mass, momentum."""
    cfg = abi.request_config(scfg, req)
    shape = (req.nx, req.ny, req.nz)
    with FluidSolver(cfg) as s:
        if flags is not None:
            s.set_geometry(flags)
        elif req.walls == abi.WALLS_Y:
            s.set_geometry(cases.channel_flags(shape))
        if req.init == abi.INIT_TAYLOR_GREEN:
            s.initialize_taylor_green(req.u[0])
        else:
            s.initialize(req.rho0, *req.u)
        s.step(req.steps)
        return (s.gather_populations(), s.total_mass(),
                tuple(s.total_momentum(a) for a in range(3)))


def run_server(scfg, reqs, keep=True, between=0):
    """Submits reqs in order (with `between` progress calls between submits),
    drains, returns {index: (populations, result)}."""
    out = {}
    with Server(scfg) as srv:
        ids = []
        for r in reqs:
            r.keep = 1 if keep else 0
            ids.append(srv.submit(r))
            for _ in range(between):
                srv.progress(wait_us=-1)
        srv.drain()
        for k, rid in enumerate(ids):
            res = srv.result(rid)
            assert res["state"] == abi.REQ_DONE
            assert res["steps"] == reqs[k].steps
            out[k] = (srv.fetch_populations(rid), res)
            srv.release(rid)
    return out


MIX_D3 = [
    ("tg-cube", lambda: tg((32, 32, 32), 21)),
    ("tg-odd", lambda: tg((33, 40, 29), 17, omega=1.9)),
    ("tg-wide", lambda: tg((300, 9, 7), 6)),         # two x-chunks per row
    ("tg-64", lambda: tg((64, 20, 12), 11)),
    ("channel", lambda: channel((40, 34, 24), 19)),
    ("channel-odd", lambda: channel((37, 23, 31), 9, g=3e-6)),
]


@pytest.mark.parametrize("shape_policy,run", [(abi.SHAPE_ROWS, 64), (abi.SHAPE_LINEAR, 8),
                                              (abi.SHAPE_AUTO, 0),
                                              (abi.SHAPE_ROWS, abi.TILE_BLOCK_RANGES),
                                              (abi.SHAPE_LINEAR, abi.TILE_BLOCK_RANGES)])
def test_boxes_match_single_solver_cm_f32(shape_policy, run):
    scfg = abi.make_server_config(budget_bytes=512 * MiB, tile_shape=shape_policy,
                                  tile_run=run, steps_per_chunk=4)
    reqs = [mk() for _, mk in MIX_D3]
    got = run_server(scfg, reqs)
    for k, (name, mk) in enumerate(MIX_D3):
        pops, mass, mom = solo(scfg, mk())
        print(f"shape policy {shape_policy}: {name}", flush=True)
        assert np.array_equal(got[k][0], pops), name
        res = got[k][1]
        assert abs(res["mass"] - mass) <= 1e-12 * abs(mass), name
        for a in range(3):
            assert abs(res["momentum"][a] - mom[a]) <= 1e-12 * abs(mass), name


@pytest.mark.parametrize("lattice,op,storage,dtype", [
    (abi.D3Q27, abi.OP_BGK, abi.RAW, abi.F64),
    (abi.D3Q27, abi.OP_TRT, abi.SHIFTED, abi.F32),
    (abi.D3Q27, abi.OP_MRT, abi.SHIFTED, abi.F64),
    (abi.D3Q27, abi.OP_BGK_PRODUCT, abi.SHIFTED, abi.F32),
    (abi.D2Q9, abi.OP_CM, abi.SHIFTED, abi.F64),
    (abi.D2Q9, abi.OP_BGK, abi.RAW, abi.F32),
])
def test_boxes_match_single_solver_other_configs(lattice, op, storage, dtype):
    scfg = abi.make_server_config(lattice=lattice, op=op, storage=storage, dtype=dtype,
                                  budget_bytes=256 * MiB, steps_per_chunk=3)
    if lattice == abi.D2Q9:
        mks = [lambda: tg((48, 40, 1), 13), lambda: channel((70, 33, 1), 10),
               lambda: tg((256, 64, 1), 5)]
    else:
        mks = [lambda: tg((24, 20, 18), 9, omega_minus=1.2),
               lambda: channel((30, 26, 20), 7, omega_minus=1.1)]
    reqs = [mk() for mk in mks]
    got = run_server(scfg, reqs)
    for k, mk in enumerate(mks):
        pops, _, _ = solo(scfg, mk())
        assert np.array_equal(got[k][0], pops), k


def test_no_force_server_matches_force_free_solver():
    scfg = abi.make_server_config(forcing=abi.FORCE_NONE, budget_bytes=128 * MiB)
    mk = lambda: tg((28, 24, 20), 10)
    got = run_server(scfg, [mk()])
    pops, _, _ = solo(scfg, mk())
    assert np.array_equal(got[0][0], pops)


def test_user_flags_match_single_solver():
    scfg = abi.make_server_config(budget_bytes=128 * MiB)
    shape = (30, 28, 26)
    x, y, z = cases.interior_xyz(abi.make_config(abi.D3Q27, shape=shape))
    flags = (((x - 15) ** 2 + (y - 14) ** 2 + (z - 13) ** 2) <= 36).astype(np.uint8)
    req = channel(shape, 12)
    with Server(scfg) as srv:
        req.keep = 1
        rid = srv.submit(req, flags=flags)
        srv.drain()
        got = srv.fetch_populations(rid)
    ref_flags = flags.copy().reshape(shape[2], shape[1], shape[0])
    pops, _, _ = solo(scfg, channel(shape, 12), flags=ref_flags.reshape(-1))
    assert np.array_equal(got, pops)


def test_same_alone_and_in_the_mix():
    """The request of interest joins while other boxes are mid-run, shares
    chunks with boxes that finish before and after it, and must match the run
    where it is alone, bit for bit (populations and completion sums)."""
    scfg = abi.make_server_config(budget_bytes=512 * MiB, steps_per_chunk=4)
    probe = lambda: channel((36, 30, 22), 23)
    alone = run_server(scfg, [probe()])[0]
    others = [tg((40, 32, 24), 40), tg((33, 33, 33), 5), channel((50, 20, 20), 30),
              tg((64, 16, 16), 12)]
    mix = others[:2] + [probe()] + others[2:]
    got = run_server(scfg, mix, between=1)
    assert np.array_equal(got[2][0], alone[0])
    for key in ("mass", "momentum"):
        assert got[2][1][key] == alone[1][key]


def test_graph_and_eager_agree():
    """Each chunk is replayed from one captured graph many times with the
    server's own eager work (seeding, summaries, copies) in between; every
    box must match the eager launches exactly."""
    reqs = lambda: [tg((32, 24, 20), 37), channel((28, 30, 18), 29), tg((45, 17, 19), 44)]
    g = run_server(abi.make_server_config(budget_bytes=256 * MiB, steps_per_chunk=4,
                                          use_graph=True), reqs(), between=2)
    e = run_server(abi.make_server_config(budget_bytes=256 * MiB, steps_per_chunk=4,
                                          use_graph=False), reqs(), between=2)
    for k in g:
        assert np.array_equal(g[k][0], e[k][0]), k


def test_long_chunks_match_single_solver():
    """Chunks of 32 steps (the length the serving bench uses), each one
    captured graph of 32 launches: boxes whose step counts are not multiples
    of 32 stop inside a chunk, a box joins mid-run, and every box matches the
    single-box solver bit for bit."""
    scfg = abi.make_server_config(budget_bytes=512 * MiB, steps_per_chunk=32)
    mks = [lambda: tg((32, 24, 20), 37), lambda: channel((28, 30, 18), 70),
           lambda: tg((300, 9, 7), 33), lambda: tg((45, 17, 19), 100)]
    got = run_server(scfg, [mk() for mk in mks], between=1)
    for k, mk in enumerate(mks):
        pops, mass, _ = solo(scfg, mk())
        assert np.array_equal(got[k][0], pops), k
        assert abs(got[k][1]["mass"] - mass) <= 1e-12 * abs(mass), k


def test_budget_queues_and_rejects():
    scfg = abi.make_server_config(budget_bytes=64 * MiB, steps_per_chunk=2)
    mk = lambda: tg((40, 40, 40), 6)  # 7 MB of FP32 populations
    from m3lb_amd import serve
    per = serve.footprint(scfg, mk())
    n_fit = (64 * MiB) // per
    reqs = [mk() for _ in range(n_fit + 3)]
    with Server(scfg) as srv:
        ids = [srv.submit(r) for r in reqs]
        srv.progress(0)
        st = srv.stats()
        assert st["resident"] == n_fit
        assert st["queued"] == 3
        assert st["bytes_used"] <= st["bytes_budget"]
        with pytest.raises(RuntimeError, match="budget"):
            srv.submit(tg((200, 200, 200), 1))
        assert srv.stats()["rejected"] == 1
        srv.drain()
        res = srv.poll()
        assert sorted(r["id"] for r in res) == sorted(ids)
        for r in res:
            assert r["state"] == abi.REQ_DONE
            assert r["arrival"] <= r["admitted"] <= r["completed"]
        masses = {r["mass"] for r in res}
        assert len(masses) == 1  # identical requests, identical sums
        st = srv.stats()
        assert st["done"] == len(reqs) and st["resident"] == 0 and st["bytes_used"] == 0
        assert st["lattice_updates"] == len(reqs) * 40 ** 3 * 6


def test_cancel_leaves_others_unchanged():
    from m3lb_amd import serve
    a = lambda: tg((30, 30, 30), 20)
    b = lambda: channel((34, 22, 26), 18)
    v = lambda: tg((32, 32, 32), 1000)
    big = lambda: tg((56, 56, 56), 1000)
    probe = abi.make_server_config()
    fa, fb, fv, fbig = (serve.footprint(probe, r()) for r in (a, b, v, big))
    # `big` fits the budget, but neither next to a, victim and b nor after the
    # victim is cancelled, so it stays queued while b is admitted past it
    assert fv + fa // 2 > fb
    budget = fa + fb + fv + (fbig - fv - fa // 2)
    scfg = abi.make_server_config(budget_bytes=budget, steps_per_chunk=2)
    ref = run_server(abi.make_server_config(budget_bytes=256 * MiB, steps_per_chunk=2),
                     [a(), b()])
    with Server(scfg) as srv:
        ra, rb = a(), b()
        ra.keep = rb.keep = 1
        ia = srv.submit(ra)
        victim = srv.submit(v())
        queued = srv.submit(big())
        ib = srv.submit(rb)
        srv.progress(wait_us=-1)
        assert srv.result(victim)["state"] == abi.REQ_RESIDENT
        assert srv.result(queued)["state"] == abi.REQ_QUEUED
        assert srv.result(ib)["state"] == abi.REQ_RESIDENT
        srv.cancel(victim)
        srv.progress(0)
        assert srv.result(queued)["state"] == abi.REQ_QUEUED
        srv.cancel(queued)
        srv.drain()
        assert np.array_equal(srv.fetch_populations(ia), ref[0][0])
        assert np.array_equal(srv.fetch_populations(ib), ref[1][0])
        states = {r["id"]: r["state"] for r in srv.poll()}
        assert states[victim] == abi.REQ_CANCELLED and states[queued] == abi.REQ_CANCELLED
        assert srv.stats()["cancelled"] == 2


def test_autotune_then_serve():
    scfg = abi.make_server_config(budget_bytes=1024 * MiB, steps_per_chunk=4)
    mk = lambda: channel((36, 30, 22), 11)
    ref = run_server(scfg, [mk()])[0]
    with Server(scfg) as srv:
        t = srv.autotune(steps=2)
        print("autotune", t, flush=True)
        assert t["candidates"] == 36 and t["best_glups"] > 0 and t["initial_glups"] > 0
        st = srv.stats()
        assert st["tile_shape"] == t["tile_shape"] and st["tile_run"] == t["tile_run"]
        assert st["bytes_used"] == 0
        r = mk()
        r.keep = 1
        rid = srv.submit(r)
        srv.drain()
        assert np.array_equal(srv.fetch_populations(rid), ref[0])


def test_fetch_macroscopic_matches_solver():
    scfg = abi.make_server_config(budget_bytes=128 * MiB)
    mk = lambda: channel((26, 24, 20), 15)
    req = mk()
    req.keep = 1
    with Server(scfg) as srv:
        rid = srv.submit(req)
        srv.drain()
        got = srv.fetch_macroscopic(rid)
    cfg = abi.request_config(scfg, mk())
    with FluidSolver(cfg) as s:
        s.set_geometry(cases.channel_flags((26, 24, 20)))
        s.initialize(1.0)
        s.step(15)
        want = s.compute_macroscopic()
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


def test_taylor_green_initializer_matches_host_field():
    """m3lb_initialize_taylor_green against the host formula through
    initialize_field (device and host sin/cos may differ in the last bit)."""
    cfg = abi.make_config(abi.D3Q27, abi.OP_CM, abi.SHIFTED, abi.F64, (24, 20, 16), omega=1.5)
    with FluidSolver(cfg) as s, FluidSolver(cfg) as h:
        s.initialize_taylor_green(0.05)
        h.initialize_field(*cases.taylor_green_box(cfg, 0.05))
        a, b = s.gather_populations(), h.gather_populations()
    assert np.max(np.abs(a - b)) <= 1e-15


def test_resident_node_cap():
    """With a cap of two boxes' nodes, two boxes run at a time; the rest wait
    in arrival order, and each result is the one it has alone."""
    mk = lambda: tg((30, 26, 22), 9)
    nodes = 30 * 26 * 22
    scfg = abi.make_server_config(budget_bytes=256 * MiB, steps_per_chunk=4,
                                  max_resident_nodes=2 * nodes)
    ref = run_server(abi.make_server_config(budget_bytes=256 * MiB), [mk()])[0]
    with Server(scfg) as srv:
        ids = []
        for _ in range(5):
            r = mk()
            r.keep = 1
            ids.append(srv.submit(r))
        srv.progress(0)
        st = srv.stats()
        assert st["resident"] == 2 and st["queued"] == 3
        assert st["resident_nodes"] == 2 * nodes
        srv.drain()
        done = srv.poll()
        assert [r["id"] for r in sorted(done, key=lambda r: r["admitted"])] == ids
        for rid in ids:
            assert np.array_equal(srv.fetch_populations(rid), ref[0])
        # a box larger than the cap is still admitted into an empty server
        big = tg((40, 40, 40), 2)
        big.keep = 1
        rb = srv.submit(big)
        srv.drain()
        assert srv.result(rb)["state"] == abi.REQ_DONE


def test_passing_never_delays_the_head():
    """Under a node cap, a head H that does not fit beside resident box A
    waits for A's 40 steps.  Behind it, S (8 steps) fits now and will have
    finished by then, so it is admitted at once; L (200 steps) would hold its
    nodes past that point, so it waits and is admitted after H.  S computes
    what it computes alone."""
    A, H = tg((30, 26, 22), 40), tg((40, 40, 40), 4)
    S, L = tg((16, 16, 16), 8), tg((16, 16, 16), 200)
    cap = 30 * 26 * 22 + 16 ** 3 + 100
    scfg = abi.make_server_config(budget_bytes=256 * MiB, steps_per_chunk=4,
                                  max_resident_nodes=cap)
    ref = run_server(abi.make_server_config(budget_bytes=256 * MiB), [tg((16, 16, 16), 8)])[0]
    with Server(scfg) as srv:
        ids = []
        for r in (A, H, S, L):
            r.keep = 1
            ids.append(srv.submit(r))
        srv.progress(0)
        st = srv.stats()
        assert st["resident"] == 2 and st["queued"] == 2
        srv.drain()
        res = {rid: srv.result(rid) for rid in ids}
        a, h, s_, l_ = (res[i] for i in ids)
        assert all(r["state"] == abi.REQ_DONE for r in res.values())
        assert s_["admitted"] <= h["admitted"] and s_["completed"] <= h["admitted"]
        assert a["completed"] <= h["admitted"] <= l_["admitted"]
        assert np.array_equal(srv.fetch_populations(ids[2]), ref[0])
        assert s_["mass"] == ref[1]["mass"]



@pytest.mark.parametrize("dtype", [abi.E2M13, abi.F16])
def test_16bit_boxes_match_single_solver(dtype):
    """16-bit boxes (shifted storage): each box gets the scales the
    single-box solver chooses for the same request (seed, measure, seed
    again) and then the same bits, alone or in a mix of Taylor-Green boxes
    and forced channels; completion sums likewise."""
    scfg = abi.make_server_config(dtype=dtype, budget_bytes=256 * MiB, steps_per_chunk=4)
    reqs = [tg((32, 24, 20), 13), channel((30, 26, 18), 9), tg((300, 9, 7), 6, omega=1.9)]
    got = run_server(scfg, [r for r in reqs], between=1)
    for k, r in enumerate([tg((32, 24, 20), 13), channel((30, 26, 18), 9),
                           tg((300, 9, 7), 6, omega=1.9)]):
        pops, mass, mom = solo(scfg, r)
        assert np.array_equal(got[k][0], pops), k
        # the sums reduce in a different order from the solver's
        assert abs(got[k][1]["mass"] - mass) <= 1e-12 * abs(mass), k
        for a in range(3):
            assert abs(got[k][1]["momentum"][a] - mom[a]) <= 1e-12 * abs(mass), k
