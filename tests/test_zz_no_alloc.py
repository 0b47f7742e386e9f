"""GPU check of ucg_solve_batch_device_noalloc (the entry point that allocates nothing): through
the C ABI, two calls in a row (the second with fewer points) give results bit-identical to
ucg_solve_batch_device on the same inputs, and slot counts beyond the fixed workspace are
refused before anything is launched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _buffers(net, params):
    import torch
    dev = torch.device("cuda:0")
    n = len(params)
    return dict(params=torch.tensor(params, device=dev),
                y_out=torch.full((n, net.neq), -1.0, dtype=torch.float64, device=dev),
                status=torch.full((n,), 99, dtype=torch.int32, device=dev),
                stats=torch.full((n, 8), -1, dtype=torch.int32, device=dev),
                n_rec=torch.zeros(n, dtype=torch.int32, device=dev),
                rec=torch.zeros((n, 32, 4 + net.nspec), dtype=torch.float64, device=dev))


def _noalloc(net, b, n_slots=0, lu_bits=64):
    import torch
    from uclchem_amd import gpu
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    n = b["params"].shape[0]
    stream = torch.cuda.current_stream(b["params"].device)
    return gpu.lib().ucg_solve_batch_device_noalloc(
        ctypes.c_int(net.handle), ctypes.c_int(n), p(b["params"]), None, p(b["y_out"]), ctypes.c_int(32),
        p(b["rec"]), p(b["n_rec"]), p(b["status"]), p(b["stats"]), None, None, ctypes.c_int(lu_bits),
        ctypes.c_int(n_slots), ctypes.c_void_p(stream.cuda_stream))


def _reference_entry(net, b, lu_bits=64):
    import torch
    stream = torch.cuda.current_stream(b["params"].device)
    net.solve_device(b["params"], b["y_out"], b["status"], b["stats"], rec=b["rec"], n_rec=b["n_rec"], max_rec=32,
                     lu_bits=lu_bits, stream=stream.cuda_stream)


@pytest.mark.parametrize("lu_bits", [64, 32])
def test_noalloc_matches_device_entry_on_two_calls(gpu_net, lu_bits):
    import torch
    from uclchem_amd import workloads
    from uclchem_amd.params import INDEX
    params = workloads.static_cloud_grid(24)
    params[:, INDEX["finaltime"]] = 1e3           # 28 output rows each (fits rec): quick
    for n in (24, 10):                             # second call: fewer points, same fixed workspace
        print(f"noalloc lu_bits={lu_bits} n={n}", flush=True)
        a, b = _buffers(gpu_net, params[:n]), _buffers(gpu_net, params[:n])
        _reference_entry(gpu_net, a, lu_bits)
        rc = _noalloc(gpu_net, b, lu_bits=lu_bits)
        assert rc == 0, gpu_net_error()
        torch.cuda.synchronize()
        for k in ("y_out", "status", "stats", "n_rec", "rec"):
            assert torch.equal(a[k], b[k]), k
        assert bool((b["status"] == 0).all())


def test_noalloc_refuses_slots_beyond_the_workspace(gpu_net):
    import torch
    from uclchem_amd import workloads
    b = _buffers(gpu_net, workloads.static_cloud_grid(400))     # 400 slots of ~296 KB > 80 MiB
    rc = _noalloc(gpu_net, b, n_slots=400)
    assert rc == -102                              # UCG_E_ARG
    assert b"fixed workspace" in gpu_net_error()
    torch.cuda.synchronize()
    assert bool((b["status"] == 99).all())         # nothing was launched


def gpu_net_error():
    from uclchem_amd import gpu
    return gpu.lib().ucg_last_error()
