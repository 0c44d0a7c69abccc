"""Graph replay of single transform calls (the reference's per-transform graph cache,
SpeedyTransforms/ext/gpu_graphs_common.jl:293-331): replays equal direct launches bitwise,
follow their inputs across replays with eager work between them, stop caching at 64 graphs per
direction, and a failed graph falls back to direct launches."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def make(sw, T=31, K=8, graphs=True):
    n, lmax, mmax = {31: (24, 33, 32), 127: (96, 129, 128)}[T]
    return sw.SpectralTransform(sw.Spectrum(lmax, mmax), sw.OctahedralGaussianGrid(n), nlayers=K,
                                gpu_graphs=graphs)


def spec(sw, S, K, seed):
    a = sw.seeded_spectrum(S.spectrum, K, seed)
    return sw.LowerTriangularArray(torch.from_numpy(a).to(torch.complex64).cuda(), S.spectrum)


@pytest.mark.parametrize("T, K", [(31, 8), (127, 24)])
def test_replays_equal_direct_launches(sw, gpu, T, K):
    G, D = make(sw, T, K, graphs=True), make(sw, T, K, graphs=False)
    c = spec(sw, G, K, seed=1)
    fg = sw.Field(torch.zeros((K, G.grid.npoints), device="cuda"), G.grid)
    bg = sw.LowerTriangularArray.zeros(G.spectrum, K, device="cuda")
    for rep in range(4):
        print(f"T{T} K={K} replay {rep}", flush=True)
        # new input values in the same buffers, eager work between the replays
        c.data.copy_(spec(sw, G, K, seed=1 + rep).data)
        sw.transform_(fg, c, G)
        sw.transform_(bg, fg, G)
        x = torch.randn(1 << 16, device="cuda")
        (x * 2).sum().item()
        fd = sw.transform(c, D)
        bd = sw.transform(fd, D)
        torch.cuda.synchronize()
        assert torch.equal(fg.data, fd.data) and torch.equal(bg.data, bd.data), rep
        assert G.graph_count() == (2, 0)
    G.close()
    D.close()


def test_cache_capped_at_64_per_direction(sw, gpu):
    K = 2
    G, D = make(sw, 31, K, graphs=True), make(sw, 31, K, graphs=False)
    c = spec(sw, G, K, seed=3)
    ref = sw.transform(c, D)
    outs = [sw.Field(torch.zeros((K, G.grid.npoints), device="cuda"), G.grid) for _ in range(70)]
    for i, f in enumerate(outs):
        sw.transform_(f, c, G)
        if i in (0, 63, 64, 69):
            print(f"synthesis into buffer {i}: cached {G.graph_count()}", flush=True)
    torch.cuda.synchronize()
    assert G.graph_count() == (64, 0)
    for f in outs:  # calls beyond the cap launched directly, same results
        assert torch.equal(f.data, ref.data)
    # the analysis direction has its own 64
    b = sw.LowerTriangularArray.zeros(G.spectrum, K, device="cuda")
    sw.transform_(b, outs[0], G)
    torch.cuda.synchronize()
    assert G.graph_count() == (65, 0)
    assert torch.equal(b.data, sw.transform(ref, D).data)
    G.close()
    D.close()


def test_failed_graph_falls_back_to_direct(sw, gpu):
    K = 4
    G, D = make(sw, 31, K, graphs=False), make(sw, 31, K, graphs=False)
    G.set_graphs(True, force_failure=True)
    c = spec(sw, G, K, seed=4)
    f = sw.Field(torch.zeros((K, G.grid.npoints), device="cuda"), G.grid)
    for _ in range(3):
        sw.transform_(f, c, G)
    torch.cuda.synchronize()
    assert G.graph_count() == (0, 1)  # one failed key, launched directly afterwards
    assert torch.equal(f.data, sw.transform(c, D).data)
    G.close()
    D.close()
