"""The single-pass z-sweeps replayed from a captured graph.

The sweep's ring entries carry tags made from a launch count that a
one-thread kernel advances before every launch; inside a captured graph
that kernel is replayed too, so every replay must see fresh tags.  Two steps
(an even number: the host's A/B buffer index returns to where capture
started) are captured once and replayed several times, with eager work and
synchronizations between replays; after each replay the state equals, bit
for bit, a level of the same initial state stepped eagerly.
"""
import numpy as np
import pytest

from marbles_amd import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage,kernel", [("f64", "k_sweep"), ("f32s", "k_sweep_f32"), ("f32s", "k_sweep")])
def test_sweep_graph_replays_equal_eager(engine, storage, kernel):
    import torch

    cfg = configs.channel_sphere((48, 24, 24), radius=5.0, center=(16.0, 12.0, 12.0))
    side = torch.cuda.Stream()
    levels = []
    for _ in range(2):
        E = configs.build_level(cfg, layout="compact", storage=storage)
        E.set_hot_kernel(kernel)
        assert E.hot_kernel_name() == kernel
        levels.append(E)
    G, R = levels  # G: replayed from the graph, R: stepped eagerly
    # one eager step pair on each first (the launch's one-time set-up, such
    # as the occupancy query, happens outside the capture)
    for E in levels:
        E.advance(2, stream=side.cuda_stream)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        G.advance(2, stream=side.cuda_stream)
    torch.cuda.synchronize()
    scratch = torch.zeros(1 << 20, dtype=torch.float64, device="cuda")
    for rep in range(4):
        graph.replay()
        scratch.add_(1.0)  # eager work between replays
        torch.cuda.synchronize()
        R.advance(2, stream=side.cuda_stream)
        torch.cuda.synchronize()
        for name in ("f", "g"):
            a, b = G.valid(name), R.valid(name)
            assert np.isfinite(b).all(), (storage, kernel, rep)
            assert np.array_equal(a, b), (storage, kernel, rep, name)
    del graph
    for E in levels:
        E.close()
