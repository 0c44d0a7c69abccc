"""Uneven batches (one launch per stage, explicit hipGraph) against single transforms and the
FP64 oracle, including graph replays with eager work between them and the forced fallback."""
import numpy as np
import pytest
import torch

from sht_cases import oracle, rel

pytestmark = pytest.mark.gpu

SIZES = {31: (24, 33, 32), 63: (48, 65, 64), 127: (96, 129, 128), 255: (192, 257, 256)}
_PLANS = {}


def plan(sw, T, nlayers=64, precision="fp32"):
    key = (T, nlayers, precision)
    if key not in _PLANS:
        n, lmax, mmax = SIZES[T]
        _PLANS[key] = sw.SpectralTransform(sw.Spectrum(lmax, mmax), sw.OctahedralGaussianGrid(n), nlayers=nlayers,
                                           precision=precision)
    return _PLANS[key]


def spec(sw, S, K, seed):
    a = sw.seeded_spectrum(S.spectrum, K, seed)
    return sw.LowerTriangularArray(torch.from_numpy(a).to(torch.complex64).cuda(), S.spectrum)


def test_mixed_batch_equals_single_calls(sw, gpu):
    S31, S63, S127 = plan(sw, 31), plan(sw, 63), plan(sw, 127)
    items, expect = [], []
    for i, (S, K) in enumerate([(S31, 5), (S63, 12), (S127, 3), (S31, 33), (S127, 16)]):
        c = spec(sw, S, K, seed=10 + i)
        f = sw.Field(torch.full((K, S.grid.npoints), float("nan"), device="cuda"), S.grid)
        items.append((S, c, f))
        expect.append(sw.transform(c, S).data)
    # analysis items and a round trip, two of them on the same plan as above
    g = sw.transform(spec(sw, S63, 7, seed=20), S63)
    out_a = sw.LowerTriangularArray.zeros(S63.spectrum, 7, device="cuda")
    items.append((S63, g, out_a))
    expect.append(sw.transform(g, S63).data)
    c_rt = spec(sw, S127, 9, seed=21)
    mid = sw.Field(torch.zeros((9, S127.grid.npoints), device="cuda"), S127.grid)
    out_rt = sw.LowerTriangularArray.zeros(S127.spectrum, 9, device="cuda")
    items.append((S127, c_rt, out_rt, mid))
    expect.append(sw.transform(sw.transform(c_rt, S127), S127).data)
    torch.cuda.synchronize()
    for graph in (True, False):
        B = sw.SpectralBatch(items, graph=graph)
        assert B.uses_graph == graph
        B.launch()
        torch.cuda.synchronize()
        for it, ref in zip(items, expect):
            got = it[2].data
            assert torch.equal(got, ref), (it[0], got.shape)
        B.close()


def test_graph_replays_follow_inputs(sw, gpu):
    S = plan(sw, 63)
    c = spec(sw, S, 8, seed=1)
    mid = sw.Field(torch.zeros((8, S.grid.npoints), device="cuda"), S.grid)
    out = sw.LowerTriangularArray.zeros(S.spectrum, 8, device="cuda")
    B = sw.SpectralBatch([(S, c, out, mid)])
    assert B.uses_graph
    for rep in range(4):
        # eager work between replays: new inputs, then the reference result eagerly
        c.data.copy_(torch.from_numpy(sw.seeded_spectrum(S.spectrum, 8, seed=100 + rep)).to(torch.complex64))
        ref = sw.transform(sw.transform(c, S), S).data.clone()
        torch.cuda.synchronize()
        B.launch()
        torch.cuda.synchronize()
        assert torch.equal(out.data, ref), rep
        assert torch.equal(mid.data, sw.transform(c, S).data)
    B.close()


def test_batch_rejects_mismatched_layer_counts(sw, gpu):
    """Every output (and a round trip's grid) must hold as many layers as the input: the
    kernels write K columns of each, so a smaller one raises before anything launches."""
    S = plan(sw, 31)
    c = spec(sw, S, 4, seed=5)
    small_f = sw.Field(torch.zeros((1, S.grid.npoints), device="cuda"), S.grid)
    with pytest.raises(sw.DimensionMismatch):
        sw.SpectralBatch([(S, c, small_f)])
    mid = sw.Field(torch.zeros((4, S.grid.npoints), device="cuda"), S.grid)
    small_c = sw.LowerTriangularArray.zeros(S.spectrum, 2, device="cuda")
    with pytest.raises(sw.DimensionMismatch):
        sw.SpectralBatch([(S, c, small_c, mid)])
    with pytest.raises(sw.DimensionMismatch):
        sw.SpectralBatch([(S, c, sw.LowerTriangularArray.zeros(S.spectrum, 4, device="cuda"), small_f)])


def test_forced_graph_failure_falls_back(sw, gpu):
    S = plan(sw, 31)
    c = spec(sw, S, 4, seed=3)
    f = sw.Field(torch.zeros((4, S.grid.npoints), device="cuda"), S.grid)
    B = sw.SpectralBatch([(S, c, f)], force_graph_failure=True)
    assert not B.uses_graph
    B.launch()
    torch.cuda.synchronize()
    assert torch.equal(f.data, sw.transform(c, S).data)
    B.close()


def test_uneven_batch_of_64_round_trips(sw, gpu):
    # 64 transforms: truncation from {31, 63, 127, 255}, layers U{1..16}, fields U{1..4}, seed 1234
    rng = np.random.default_rng(1234)
    items, refs = [], []
    for i in range(64):
        T = int(rng.choice([31, 63, 127, 255]))
        K = int(rng.integers(1, 17)) * int(rng.integers(1, 5))
        S = plan(sw, T)
        c = spec(sw, S, K, seed=1000 + i)
        mid = sw.Field(torch.zeros((K, S.grid.npoints), device="cuda"), S.grid)
        out = sw.LowerTriangularArray.zeros(S.spectrum, K, device="cuda")
        items.append((S, c, out, mid))
        refs.append((T, K, i))
    B = sw.SpectralBatch(items)
    B.launch()
    torch.cuda.synchronize()
    checked = 0
    for (S, c, out, mid), (T, K, i) in zip(items, refs):
        if T in (31, 127) and checked < 4:  # FP64 oracle on a few members
            n, lmax, mmax = SIZES[T]
            o = oracle("octahedral", n, lmax, mmax)
            a = c.data.cpu().numpy().astype(np.complex128)
            assert rel(out.data.cpu().numpy(), o.analysis(o.synthesis(a))) < 1e-5
            checked += 1
        else:  # single-call path, bitwise
            assert torch.equal(out.data, sw.transform(sw.transform(c, S), S).data)
    B.close()


@pytest.mark.parametrize("precision", ["fp16x3", "bf16x3"])
def test_shared_column_groups_mixed_flags(sw, gpu, precision):
    """Items on one plan share Legendre column groups (column counts that straddle the groups'
    32 columns), split by the unscale_coslat flag: syntheses with and without it, analyses and
    a round trip on the same plan, in a split-precision mode, bitwise equal to single calls."""
    S = plan(sw, 127, nlayers=48, precision=precision)
    items, expect = [], []
    for i, (K, unscale) in enumerate([(5, False), (40, True), (7, False), (33, True), (1, False)]):
        c = spec(sw, S, K, seed=60 + i)
        f = sw.Field(torch.full((K, S.grid.npoints), float("nan"), device="cuda"), S.grid)
        items.append((S, c, f, unscale))
        expect.append(sw.transform(c, S, unscale_coslat=unscale).data)
    for i, K in enumerate((9, 30)):
        g = sw.transform(spec(sw, S, K, seed=70 + i), S)
        out = sw.LowerTriangularArray(torch.full((K, S.spectrum.nharm), float("nan"), dtype=torch.complex64,
                                                 device="cuda"), S.spectrum)
        items.append((S, g, out))
        expect.append(sw.transform(g, S).data)
    c = spec(sw, S, 11, seed=80)
    mid = sw.Field(torch.zeros((11, S.grid.npoints), device="cuda"), S.grid)
    out = sw.LowerTriangularArray.zeros(S.spectrum, 11, device="cuda")
    items.append((S, c, out, mid))
    expect.append(sw.transform(sw.transform(c, S), S).data)
    torch.cuda.synchronize()
    B = sw.SpectralBatch(items)
    B.launch()
    torch.cuda.synchronize()
    for it, ref in zip(items, expect):
        assert torch.equal(it[2].data, ref), (it[1].data.shape, ref.shape)
    B.close()
