"""GPU tests of the serving loop (request queues, continuous batching over three networks).

Served results match the reference Fortran on every network of the mix (a seeded stream with
t_end up to 1e7 yr).  Each result is the same alone and in the mix: a request answered by a queue while other
requests of three networks arrive (Poisson stream) and run beside it is bit-identical to
the same point solved alone by ucg_solve_batch (the per-point computation never depends
on which workgroup runs it or on what else is resident).  Error paths are forced from the
host (a closed queue, a full queue, an idle limit the host lets expire)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nets():
    from uclchem_amd import gpu
    from uclchem_amd import makerates_hip as mh
    from uclchem_amd import synthetic as sy
    out = {"default": gpu.Network(mh.load_network(mh.network_path("default")))}
    for name in ("synthetic_150_1500", "synthetic_500_6000"):
        out[name] = gpu.Network(sy.load_or_build(name))
    return out


def test_results_same_alone_and_in_mix(nets):
    from uclchem_amd import serving
    reqs = serving.poisson_requests(18, rate_per_s=40.0, networks=list(nets), seed=3, t_end_range=(1e3, 3e4))
    slots = serving.plan_slots(nets, n_cu=64, hbm_budget_bytes=8e9)
    srv = serving.Server(nets, slots, capacity=64)
    try:
        print(f"[serving] slots {slots}", flush=True)
        done = serving.run_stream(srv, reqs, timeout_s=300.0,
                                  progress=lambda d, s, t: print(f"[serving] {d}/{s} done at {t:.0f}s", flush=True))
    finally:
        srv.close()
    assert len(done) == len(reqs)
    by_arrival = {round(t.t_arrival - min(x.t_arrival for x in done), 9): t for t in done}
    assert len(by_arrival) == len(reqs)
    for (arr, name, p), t in zip(reqs, sorted(done, key=lambda x: x.t_arrival)):
        assert t.network == name
        alone = nets[name].solve(p[None])
        assert t.status == 0 and alone["status"][0] == 0
        assert np.array_equal(t.y, alone["y"][0]), name
        assert np.array_equal(t.stats, alone["stats"][0]), name
        assert t.t_done >= t.t_arrival


def test_served_results_match_the_reference(nets):
    """The seeded 24-request stream of tests/golden/ref_serving.npz (t_end log-uniform
    1e3..1e7 yr over the three networks) served with continuous admission: every result within
    1% of the reference Fortran compiled on the request's network (the default network's
    shipped sources; the synthetic ones written by uclchem_amd/reference_format.py) on every
    abundance above 1e-14."""
    import os
    from uclchem_amd import serving
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_serving.npz")
    if not os.path.exists(path):
        pytest.skip("ref_serving.npz not generated")
    f = np.load(path)
    from uclchem_amd.params import widen
    reqs = [(float(a), str(n), p) for a, n, p in zip(f["arrival"], f["network"], widen(f["params"]))]
    slots = serving.plan_slots(nets, n_cu=64, hbm_budget_bytes=8e9)
    srv = serving.Server(nets, slots, capacity=64)
    try:
        done = serving.run_stream(srv, reqs, timeout_s=500.0,
                                  progress=lambda d, s, t: print(f"[serving] {d}/{s} done at {t:.0f}s", flush=True))
    finally:
        srv.close()
    done = sorted(done, key=lambda x: x.t_arrival)
    bad = []
    for q, ((arr, name, p), t) in enumerate(zip(reqs, done)):
        assert t.network == name and t.status == 0 and f["rc"][q] == 0
        ns = nets[name].nspec
        ref = f["final"][q, 4:4 + ns]
        m = ref > 1e-14
        r = np.abs(t.y[:ns][m] - ref[m]) / ref[m]
        if r.max() > 0.01:
            bad.append((q, name, float(p[7]), float(r.max())))
    assert not bad, bad


def test_queue_errors_are_reported(nets):
    from uclchem_amd import serving
    from uclchem_amd.gpu import GpuError
    from uclchem_amd.params import make_params
    net = nets["synthetic_150_1500"]
    p = make_params(initialdens=1e4, initialtemp=10.0, finaltime=1e3)
    q = serving.RequestQueue(net, capacity=2, n_slots=2)
    try:
        r0 = q.submit(p)
        r1 = q.submit(p)
        with pytest.raises(GpuError):
            q.submit(p)                       # capacity reached
        while not (q.done(r0) and q.done(r1)):
            time.sleep(0.01)
        y0, s0, _ = q.result(r0)
        y1, s1, _ = q.result(r1)
        assert s0 == 0 and np.array_equal(y0, y1)
        with pytest.raises(GpuError):
            q.submit(make_params(initialdens=-1.0))   # invalid parameters never reach the device
    finally:
        q.close()
    # a queue whose workgroups hit their idle limit reports requests it can no longer run
    q = serving.RequestQueue(net, capacity=4, n_slots=1, idle_timeout_s=0.2)
    try:
        time.sleep(1.5)
        r = q.submit(p)
        with pytest.raises(GpuError):
            for _ in range(200):
                if q.done(r):
                    break
                time.sleep(0.01)
    finally:
        q.close()


def _max_overlap(intervals):
    """Largest number of [start, end) intervals that are open at one time."""
    ev = sorted([(a, 1) for a, b in intervals] + [(b, -1) for a, b in intervals], key=lambda e: (e[0], e[1]))
    n = best = 0
    for _, d in ev:
        n += d
        best = max(best, n)
    return best


def test_shared_queue_gives_a_burst_every_slot(nets):
    """One queue shared by the three networks (ucg_queue_open_shared): a burst of requests on
    the default network alone runs on all of the queue's slots at once (device timestamps:
    every slot holds one of them at some moment; a per-network partition would give it a
    third), requests of the other networks interleave with it, and every result is
    bit-identical to the same point solved alone."""
    from uclchem_amd import serving
    from uclchem_amd.params import make_params
    names = list(nets)
    n_slots = 12
    q = serving.SharedQueue([nets[k] for k in names], capacity=64, n_slots=n_slots)
    burst = [make_params(initialdens=10 ** (3.5 + 0.05 * i), initialtemp=10.0, finaltime=1e5) for i in range(2 * n_slots)]
    others = [(names.index(k), make_params(initialdens=1e4, initialtemp=12.0, finaltime=3e3))
              for k in names if k != "default"]
    try:
        reqs = [(names.index("default"), p, q.submit(p, names.index("default"))) for p in burst]
        reqs += [(k, p, q.submit(p, k)) for k, p in others]
        t0 = time.time()
        while not all(q.done(r) for _, _, r in reqs):
            assert time.time() - t0 < 240, "shared queue did not finish"
            time.sleep(0.01)
        res = [(k, p, q.result(r, k), q.times(r)) for k, p, r in reqs]
    finally:
        q.close()
    spans = [t for k, _, _, t in res if k == names.index("default")]
    assert all(b > a for a, b in spans)
    busy = _max_overlap(spans)
    print(f"[shared] {busy} of {n_slots} slots ran default-network requests at once", flush=True)
    assert busy == n_slots
    for k, p, (y, st, stats), _ in res:
        alone = nets[names[k]].solve(p[None])
        assert st == 0 and alone["status"][0] == 0
        assert np.array_equal(y, alone["y"][0]) and np.array_equal(stats, alone["stats"][0]), names[k]
