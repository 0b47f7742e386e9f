"""The allocation-free device entry (ucg_solve_batch_device_noalloc) captured in a HIP graph:
its launch pair (work-counter reset kernel, solve kernel) records no memset, no allocation and
no synchronisation, so a caller can capture it once and replay it.  Each replay — with eager
work of the program and synchronisations between replays, and the outputs cleared eagerly
before each — must give results bit-identical to the same launch run eagerly."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_captured_noalloc_launch_replays_like_eager(gpu_net):
    import torch
    from uclchem_amd import gpu, workloads
    from uclchem_amd.params import INDEX
    dev = torch.device("cuda:0")
    params = workloads.static_cloud_grid(16)
    params[:, INDEX["finaltime"]] = 1e3
    n = len(params)
    P = torch.tensor(params, device=dev)
    outs = dict(y_out=torch.zeros((n, gpu_net.neq), dtype=torch.float64, device=dev),
                status=torch.zeros(n, dtype=torch.int32, device=dev),
                stats=torch.zeros((n, 8), dtype=torch.int32, device=dev))
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731

    def launch(stream):
        return gpu.lib().ucg_solve_batch_device_noalloc(
            ctypes.c_int(gpu_net.handle), ctypes.c_int(n), ptr(P), None, ptr(outs["y_out"]), ctypes.c_int(0), None,
            None, ptr(outs["status"]), ptr(outs["stats"]), None, None, ctypes.c_int(64), ctypes.c_int(0),
            ctypes.c_void_p(stream.cuda_stream))

    def clear():
        outs["y_out"].fill_(-1.0)
        outs["status"].fill_(99)
        outs["stats"].fill_(-1)

    clear()
    assert launch(torch.cuda.current_stream(dev)) == 0
    torch.cuda.synchronize()
    eager = {k: v.clone() for k, v in outs.items()}
    assert bool((eager["status"] == 0).all())
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream(device=dev)
    side.wait_stream(torch.cuda.current_stream(dev))
    rc = []
    with torch.cuda.graph(g, stream=side, capture_error_mode="relaxed"):
        rc.append(launch(side))
    assert rc == [0]
    torch.cuda.synchronize()
    for replay in range(3):
        print(f"[graph] replay {replay}", flush=True)
        clear()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for k in outs:
            assert torch.equal(outs[k], eager[k]), (replay, k)
        other = gpu_net.solve(params[:2])          # eager work of the program between replays
        assert np.all(other["status"] == 0)
