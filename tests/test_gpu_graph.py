"""HIP graph capture of the batched solve (the u, v, w + p sequence of one substep).

After a warmup call has sized the library's buffers (and created its second stream), the call records
into a torch CUDA graph (a hipGraph) with its Dirichlet and Neumann groups. The Neumann group's stream
joins the capture through the library's fork/join events. The solve path records no memset or copy
node. Each replay, with eager work and synchronisations between the replays, must equal the eager
solve to the bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    import duct_amd
    torch.cuda.set_device(0)
    duct_amd.init(0)
    yield duct_amd
    duct_amd.shutdown()


@pytest.mark.parametrize("fold,overlap", [(0, True), (2, True), (2, False)])
def test_batch_solve_replays_in_a_graph(dev, fold, overlap):
    import torch
    m, M = 129, 8
    dev.setup(m)
    rng = np.random.default_rng(11 + fold)
    f = [torch.from_numpy(rng.standard_normal((M, m, m)) + 1j * rng.standard_normal((M, m, m))).cuda()
         for _ in range(4)]
    bounds, alams = "dddn", [1000.0, 1000.0, 500.0, 0.0]
    print("graph capture m=%d M=%d fold=%d overlap=%d" % (m, M, fold, overlap), flush=True)
    try:
        dev.set_fold(fold)
        dev.set_overlap(overlap)
        eager = [torch.zeros_like(x) for x in f]
        dev.helm3d_batch(f, eager, bounds, alams)      # warmup: sizes the buffers (rhs boundary zeroed)
        torch.cuda.synchronize()
        sol = [torch.zeros_like(x) for x in f]
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=side):
            dev.helm3d_batch(f, sol, bounds, alams)
        torch.cuda.synchronize()
        for rep in range(3):
            for x in sol:
                x.zero_()
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            for a, b in zip(sol, eager):
                assert torch.equal(a, b), (fold, overlap, rep)
            scratch = [torch.zeros_like(x) for x in f]  # eager work between the replays
            dev.helm3d_batch(f, scratch, bounds, alams)
            torch.cuda.synchronize()
            assert all(torch.equal(a, b) for a, b in zip(scratch, eager))
        del g
    finally:
        dev.set_fold(0)
        dev.set_overlap(True)
