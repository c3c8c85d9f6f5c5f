"""GPU checks of the step graphs and of the serving loop.

* graph replay: steps replayed from the captured two-step hipGraph equal
  eagerly launched steps bit for bit, replay after replay, with downloads
  and synchronisations in between;
* serving: every request of a mixed stream gives bit-identical results to
  running it alone (its grouped kernels touch only its own buffers, every
  kernel is deterministic), and the grouped step kernels give the ordinary
  kernels' results; FIFO admission respects the slot limit and the HBM
  budget, admitted by exact footprints from one arena.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_graph_replay_matches_eager(W, precision):
    p = W.dam_break_params(3)
    st = W.dam_break_state(W.dam_break_desc(3, 0.04, tank_l=2.0, col_h=1.0, tank_h=1.5), p)
    outs = {}
    for graphs in (True, False):
        with W.Simulation.from_state(p, st, precision=precision) as sim:
            sim.set_graphs(graphs)
            sim.init()
            snaps = []
            for _ in range(4):
                sim.step(10)
                snaps.append(sim.download())
                sim.step(3)  # odd counts end with an eager step
                sim.sync()
            outs[graphs] = (snaps, sim.stats())
    for a, b in zip(outs[True][0], outs[False][0]):
        for k in a:
            assert np.array_equal(a[k], b[k]), k
    assert outs[True][1]["pairs_total"] == outs[False][1]["pairs_total"]


def small_stream(W, precision="mixed"):
    spec = [(2, 30_000, 2.0, 0.05, 160, 0.00), (3, 25_000, 1.5, 0.12, 120, 0.02),
            (2, 60_000, 3.5, 0.18, 200, 0.05), (3, 40_000, 2.5, 0.08, 90, 0.06),
            (2, 20_000, 1.0, 0.02, 150, 0.30)]
    return [W.make_request(d, n, a, al, s, t, precision) for d, n, a, al, s, t in spec]


KEYS = ("steps", "pairs", "sim_time", "front", "column", "kinetic", "n")


def test_serving_results_equal_alone(W):
    reqs = small_stream(W)
    mix = W.serve(reqs, max_slots=3, chunk_steps=25)
    assert all(r["status"] == 0 for r in mix)
    for k, r in enumerate(reqs):
        r.arrival_s = 0.0
        alone = W.serve([r], max_slots=1, chunk_steps=25)[0]
        for key in KEYS:
            assert alone[key] == mix[k][key], (k, key)
        assert mix[k]["steps"] == r.steps
        assert mix[k]["latency_s"] > 0 and mix[k]["start_s"] >= mix[k]["arrival_s"] - 1e-9


def test_serving_fifo_one_slot(W):
    reqs = small_stream(W, precision="fp32")
    res = W.serve(reqs, max_slots=1, chunk_steps=50)
    order = np.argsort([r["arrival_s"] for r in res], kind="stable")
    for a, b in zip(order[:-1], order[1:]):
        assert res[b]["start_s"] >= res[a]["end_s"] - 1e-9


def test_serving_budget_serialises(W):
    reqs = small_stream(W, precision="fp32")
    # a budget that holds any one request but never the two largest at once,
    # by the exact footprints the loop admits by
    sizes = [W.request_bytes(r) for r in reqs]
    budget = max(sizes) + min(sizes) - 1
    res = W.serve(reqs, max_slots=4, hbm_budget_bytes=budget, chunk_steps=50)
    ev = sorted([(r["start_s"], 1, sizes[k]) for k, r in enumerate(res)] + [(r["end_s"], -1, sizes[k]) for k, r in
                enumerate(res)], key=lambda e: (e[0], e[1]))
    used = 0
    for _, sign, sz in ev:
        used += sign * sz
        assert used <= budget
    assert all(r["status"] == 0 for r in res)


@pytest.mark.parametrize("precision", ["mixed", "fp32", "fp64"])
def test_serving_footprint_is_exact(W, precision):
    # an arena of exactly one request's footprint holds all its buffers (the
    # figure is not an underestimate), and one byte less is refused up front
    r = W.make_request(2, 20_000, 1.0, 0.05, 20, 0.0, precision)
    b = W.request_bytes(r)
    res = W.serve([r], max_slots=1, hbm_budget_bytes=b, chunk_steps=10)[0]
    assert res["status"] == 0 and res["steps"] == 20
    with pytest.raises(W.WcsphError, match="budget"):
        W.serve([r], max_slots=1, hbm_budget_bytes=b - 1)


@pytest.mark.parametrize("wall_mode", [0, 1])
@pytest.mark.parametrize("dim", [2, 3])
def test_grouped_steps_equal_simulation(W, dim, wall_mode):
    # the grouped step kernels (serving) give the results of the ordinary
    # step kernels (a Simulation stepped directly), bit for bit, also when
    # the step count is odd (the last step runs as a one-member launch),
    # with frozen walls (wall_mode 1) as without
    r = W.make_request(dim, 30_000, 1.5, 0.1, 121, 0.0, "mixed")
    r.params.wall_mode = wall_mode
    res = W.serve([r], max_slots=1, chunk_steps=40)[0]
    assert res["status"] == 0 and res["precision_used"] == W.PREC_MIXED
    st = W.dam_break_state(r.geom, r.params)
    band = 2 * r.geom.dx
    with W.Simulation.from_state(r.params, st, precision="mixed") as sim:
        sim.init()
        sim.step(121)
        s = sim.stats()
        d = sim.diagnostics(band, band, r.geom.tank_l - band)
    assert res["steps"] == s["step"] == 121
    assert res["pairs"] == s["pairs_total"]
    assert res["sim_time"] == s["time"]
    assert res["front"] == d["front"] and res["column"] == d["column"] and res["kinetic"] == d["kinetic"]


def test_serving_single_step_chunks(W):
    # chunk_steps 1: every grouped member advances by one-member single
    # steps, and the results equal those of two-step graph chunks bit for bit
    spec = [(2, 20_000, 1.0, 0.02, 11), (2, 30_000, 2.0, 0.05, 6), (3, 25_000, 1.5, 0.12, 7)]
    reqs = [W.make_request(d, n, a, al, s, 0.0, "mixed") for d, n, a, al, s in spec]
    one = W.serve(reqs, max_slots=3, chunk_steps=1)
    big = W.serve(reqs, max_slots=3, chunk_steps=25)
    for k, r in enumerate(reqs):
        assert one[k]["status"] == 0 and one[k]["steps"] == r.steps
        for key in KEYS:
            assert one[k][key] == big[k][key], (k, key)


def test_serving_rejects_oversized_request(W):
    reqs = small_stream(W)[:1]
    with pytest.raises(W.WcsphError, match="budget"):
        W.serve(reqs, hbm_budget_bytes=1 << 20)


def test_serving_c4_stream_results_equal_alone(W):
    # a stream of the serving configuration's shape: seeded Poisson arrivals
    # (4/s), 2D and 3D dam breaks of 20k-2M particles (log-uniform), column
    # aspect 1-4, alpha 0.02-0.2, 500-5000 steps, plus one 3D request of 2M
    # particles; every request's result in the mix equals its result served
    # alone, bit for bit, and the completion latencies are reported
    import time
    reqs = W.poisson_workload(2026, 6, 4.0, precision="mixed")
    big = W.make_request(3, 2_000_000, 2.0, 0.1, 600, reqs[2].arrival_s, "mixed")
    reqs.insert(3, big)
    t0 = time.perf_counter()
    mix = W.serve(reqs, max_slots=4, chunk_steps=50)
    wall = time.perf_counter() - t0
    assert all(r["status"] == 0 for r in mix), [r["status"] for r in mix]
    assert max(r["n"] for r in mix) >= 1_900_000 and min(r["steps"] for r in mix) >= 500
    lat = np.array([r["latency_s"] for r in mix])
    pairs = sum(r["pairs"] for r in mix)
    makespan = max(r["end_s"] for r in mix)
    print(f"C4 stream: {len(reqs)} requests, n {[r['n'] for r in mix]}, steps {[r['steps'] for r in mix]}; "
          f"p50 {np.percentile(lat, 50):.3f} s, p99 {np.percentile(lat, 99):.3f} s, makespan {makespan:.2f} s, "
          f"{pairs / makespan:.3e} pairs/s (wall {wall:.1f} s)", flush=True)
    for k, r in enumerate(reqs):
        r.arrival_s = 0.0
        alone = W.serve([r], max_slots=1, chunk_steps=50)[0]
        for key in KEYS:
            assert alone[key] == mix[k][key], (k, key)
