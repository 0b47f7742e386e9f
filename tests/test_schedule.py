"""Host side of the point scheduling order of Network.solve_device.  The plumbing tests run
on CPU tensors with a stand-in library, so they let solve_device accept CPU tensors
(gpu.DEVICE_TYPE); with the real setting it refuses them."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def test_densest_first_is_a_stable_descending_permutation():
    from uclchem_amd.gpu import schedule_order
    from uclchem_amd.params import INDEX
    from uclchem_amd import workloads
    p = workloads.static_cloud_grid(64)
    p[10, INDEX["initialdens"]] = p[3, INDEX["initialdens"]]      # a tie: index order kept
    perm = schedule_order(torch.tensor(p), "densest").numpy()
    assert sorted(perm.tolist()) == list(range(64))
    d = p[perm, INDEX["initialdens"]]
    assert np.all(d[:-1] >= d[1:])
    assert list(perm).index(3) < list(perm).index(10)


def test_explicit_order_checked():
    from uclchem_amd.gpu import schedule_order
    p = torch.zeros((5, 64), dtype=torch.float64)
    assert schedule_order(p, torch.tensor([4, 3, 2, 1, 0])).tolist() == [4, 3, 2, 1, 0]
    for bad in ([0, 1, 2, 3], [0, 1, 2, 3, 3], [0, 1, 2, 3, 5]):
        with pytest.raises(ValueError):
            schedule_order(p, torch.tensor(bad))
    with pytest.raises(ValueError):
        schedule_order(p, "random")


def test_ordered_solve_scatters_outputs_back(monkeypatch):
    """The gather/launch/scatter plumbing of solve_device(order=...), on host tensors with a
    stand-in for the library entry point: the stand-in writes, for each position k of the
    (permuted) batch, values that depend only on that point's parameters, plus the position
    itself as its start stamp; after the scatter every row must be the one of its own point
    and the start stamps must follow the requested order."""
    import ctypes
    from uclchem_amd import gpu
    from uclchem_amd.params import INDEX
    n, neq, nprm = 6, 4, 64
    seen = {}

    def arr(ptr, count, ctype):
        return np.ctypeslib.as_array((ctype * count).from_address(ptr.value))

    class Stub:
        def ucg_solve_batch_device(self, h, nn, p, yi, yo, mr, rec, nrec, st, stats, t0, t1, bits, slots, stream):
            k = nn.value
            prm = arr(p, k * nprm, ctypes.c_double).reshape(k, nprm)
            y = arr(yo, k * neq, ctypes.c_double).reshape(k, neq)
            y[:] = prm[:, INDEX["initialdens"], None] * np.arange(1, neq + 1)
            arr(st, k, ctypes.c_int)[:] = (prm[:, INDEX["initialdens"]] % 7).astype(np.int32)
            arr(t0, k, ctypes.c_double)[:] = np.arange(k)
            seen["dens"] = prm[:, INDEX["initialdens"]].copy()
            return 0

    monkeypatch.setattr(gpu, "_lib", Stub())
    monkeypatch.setattr(gpu, "DEVICE_TYPE", "cpu")
    net = gpu.Network.__new__(gpu.Network)
    net.handle, net.neq = 0, neq
    p = torch.zeros((n, nprm), dtype=torch.float64)
    p[:, INDEX["initialdens"]] = torch.tensor([3.0, 50.0, 7.0, 900.0, 1.0, 20.0])
    y = torch.zeros((n, neq), dtype=torch.float64)
    st = torch.zeros(n, dtype=torch.int32)
    stats = torch.zeros((n, 8), dtype=torch.int32)
    t0 = torch.zeros(n, dtype=torch.float64)
    net.solve_device(p, y, st, stats, t_start=t0, order="densest")
    dens = p[:, INDEX["initialdens"]].numpy()
    assert seen["dens"].tolist() == sorted(dens.tolist(), reverse=True)
    assert np.array_equal(y.numpy(), dens[:, None] * np.arange(1, neq + 1))
    assert np.array_equal(st.numpy(), (dens % 7).astype(np.int32))
    assert t0.numpy().tolist() == [4, 1, 3, 0, 5, 2]          # rank of each point in densest-first order
    t0.zero_()
    net.solve_device(p, y, st, stats, t_start=t0, order=torch.tensor([5, 4, 3, 2, 1, 0]))
    assert t0.numpy().tolist() == [5, 4, 3, 2, 1, 0]


def test_host_solve_order_round_trip(monkeypatch):
    """Network.solve(order="densest") hands the library the points densest first and returns
    every output in index order (stand-in library as above)."""
    import ctypes
    from uclchem_amd import gpu
    from uclchem_amd.params import INDEX
    n, neq, nprm, max_rec, nspec = 5, 4, 64, 2, 3
    seen = {}

    def arr(ptr, count, ctype):
        return np.ctypeslib.as_array((ctype * count).from_address(ptr.value))

    class Stub:
        def ucg_solve_batch(self, h, nn, p, yi, yo, mr, rec, nrec, st, stats, bits):
            k = nn.value
            prm = arr(p, k * nprm, ctypes.c_double).reshape(k, nprm)
            d = prm[:, INDEX["initialdens"]]
            arr(yo, k * neq, ctypes.c_double).reshape(k, neq)[:] = d[:, None] + np.arange(neq)
            arr(rec, k * max_rec * (4 + nspec), ctypes.c_double).reshape(k, max_rec, 4 + nspec)[:] = d[:, None, None]
            arr(nrec, k, ctypes.c_int)[:] = (d % 3).astype(np.int32)
            arr(st, k, ctypes.c_int)[:] = np.arange(k)          # position in the launch
            arr(stats, k * 8, ctypes.c_int).reshape(k, 8)[:] = (d[:, None] % 11).astype(np.int32)
            seen["dens"] = d.copy()
            return 0

    monkeypatch.setattr(gpu, "_lib", Stub())
    net = gpu.Network.__new__(gpu.Network)
    net.handle, net.neq, net.nspec = 0, neq, nspec
    p = np.zeros((n, nprm))
    dens = np.array([5.0, 400.0, 30.0, 2.0, 30.0])
    p[:, INDEX["initialdens"]] = dens
    out = net.solve(p, max_rec=max_rec, order="densest")
    assert seen["dens"].tolist() == [400.0, 30.0, 30.0, 5.0, 2.0]
    assert np.array_equal(out["y"], dens[:, None] + np.arange(neq))
    assert np.all(out["rec"] == dens[:, None, None])
    assert np.array_equal(out["n_rec"], (dens % 3).astype(np.int32))
    assert out["status"].tolist() == [3, 0, 1, 4, 2]       # rank in densest-first order, ties by index
    assert np.array_equal(out["stats"], np.repeat((dens[:, None] % 11).astype(np.int32), 8, axis=1))
    with pytest.raises(ValueError):
        net.solve(p, order="random")


def test_device_solve_refuses_mismatched_buffers(monkeypatch):
    """solve_device checks shapes, dtypes, layout and device before it hands pointers to the
    kernel (which indexes them by point): nothing reaches the library on a mismatch."""
    from uclchem_amd import gpu
    calls = []

    class Stub:
        def ucg_solve_batch_device(self, *a):
            calls.append(a)
            return 0

    monkeypatch.setattr(gpu, "_lib", Stub())
    net = gpu.Network.__new__(gpu.Network)
    net.handle, net.neq, net.nspec = 0, 5, 4
    n = 3
    host = dict(params=torch.zeros((n, 64), dtype=torch.float64), y_out=torch.zeros((n, 5), dtype=torch.float64),
                status=torch.zeros(n, dtype=torch.int32), stats=torch.zeros((n, 8), dtype=torch.int32))
    with pytest.raises(ValueError):          # host tensors never reach the kernel
        net.solve_device(**host)
    assert not calls
    monkeypatch.setattr(gpu, "DEVICE_TYPE", "cpu")
    good = dict(params=torch.zeros((n, 64), dtype=torch.float64), y_out=torch.zeros((n, 5), dtype=torch.float64),
                status=torch.zeros(n, dtype=torch.int32), stats=torch.zeros((n, 8), dtype=torch.int32))
    net.solve_device(**good)
    assert len(calls) == 1
    bad = [("params", torch.zeros((n, 64), dtype=torch.float32)), ("params", torch.zeros((n, 24), dtype=torch.float64)),
           ("y_out", torch.zeros((n, 4), dtype=torch.float64)), ("status", torch.zeros(n, dtype=torch.int64)),
           ("stats", torch.zeros((n, 8), dtype=torch.int32).t().contiguous().t()),
           ("stats", torch.zeros((n + 1, 8), dtype=torch.int32))]
    for key, val in bad:
        args = dict(good)
        args[key] = val
        with pytest.raises(ValueError):
            net.solve_device(**args)
    with pytest.raises(ValueError):
        net.solve_device(**good, t_done=torch.zeros(n, dtype=torch.float32))
    with pytest.raises(ValueError):
        net.solve_device(**good, rec=torch.zeros((n, 2, 8), dtype=torch.float64), max_rec=2)   # n_rec missing
    net.solve_device(**good, rec=torch.zeros((n, 2, 8), dtype=torch.float64), max_rec=2,
                     n_rec=torch.zeros(n, dtype=torch.int32))
    assert len(calls) == 2


def test_piece_wrappers_check_widths(monkeypatch):
    """The rates / RHS / Jacobian / Newton-solve wrappers refuse arrays whose width differs
    from what the library reads per point (NPARAM, 8 physics values, nreac, neq)."""
    from uclchem_amd import gpu
    monkeypatch.setattr(gpu, "_lib", object())        # any call into the library would fail loudly
    net = gpu.Network.__new__(gpu.Network)
    net.handle, net.neq, net.nspec, net.nreac = 0, 5, 4, 7
    p, ph, y, r = np.zeros((2, 64)), np.zeros((2, 8)), np.zeros((2, 5)), np.zeros((2, 7))
    for call in (lambda: net.rates(np.zeros((2, 24)), ph, y), lambda: net.rates(p, np.zeros((2, 7)), y),
                 lambda: net.rates(p, ph, np.zeros((2, 4))), lambda: net.rhs(p, ph, np.zeros((2, 6)), y),
                 lambda: net.jacobian(p, ph, r, np.zeros((3, 5))),
                 lambda: net.newton_solve(p, ph, y, 1.0, np.zeros((2, 6)))):
        with pytest.raises(ValueError):
            call()
