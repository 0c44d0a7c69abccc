"""GPU tests of the minimal dry primitive-equation core and the learned column correction.

No reference run is possible (the reference is Julia, absent here), so the core is pinned
by properties the equations fix: the spectral curl / divergence invert UV_from_vordiv, an
atmosphere at rest with horizontally uniform temperature has zero tendency, integrations
stay bounded, the global mean of ln ps is conserved, and the learned correction reduces the
truncation-tendency error on runs it was not trained on."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods(sw, gpu):
    return sw.dycore, sw.learned


def test_curl_divergence_invert_uv(sw, mods):
    dy, _ = mods
    m = dy.DryPrimitiveModel(31, 8, nmembers=2)
    a = sw.seeded_spectrum(m.spectrum, 16, 3)
    vor = torch.from_numpy(a[:8]).to(torch.complex64).cuda() * 1e-5 * m.ops.keep
    div = torch.from_numpy(a[8:]).to(torch.complex64).cuda() * 1e-6 * m.ops.keep
    vor[:, 0] = 0
    div[:, 0] = 0
    U, V = m.uv(vor, div)
    uv = m.synth(torch.cat([U, V]), unscale=True)
    s = m.analyse(torch.cat([uv[:8] * m.coslat_inv, uv[8:] * m.coslat_inv]))
    vor2 = m.ops.curl(s[:8], s[8:]) / m.a
    div2 = m.ops.divergence(s[:8], s[8:]) / m.a
    assert float((vor2 - vor).abs().max() / vor.abs().max()) < 1e-4
    assert float((div2 - div).abs().max() / div.abs().max()) < 1e-4
    m.close()


def test_rest_state_has_no_tendency(mods):
    dy, _ = mods
    m = dy.DryPrimitiveModel(31, 8, nmembers=1)
    st = m.jet_state([0], u0=0.0, noise_K=0.0)  # at rest, T(sigma) only
    t = m.tendencies(st)
    # round-off of the engine only: far below one day's change of any field
    assert float(t["vor"].abs().max()) < 1e-12
    assert float(t["div"].abs().max()) < 1e-10
    assert float(t["temp"].abs().max()) < 1e-6
    assert float(t["lnps"].abs().max()) < 1e-10
    m.close()


def test_jet_run_bounded_and_mass_conserved(mods):
    dy, _ = mods
    m = dy.DryPrimitiveModel(31, 8, nmembers=2)
    m.set_state(m.jet_state([1, 2]))
    q0 = m.state["lnps"][:, 0].clone()
    m.run(int(86400 / m.dt))
    _, g = m.tendencies(m.state, with_grid=True)
    assert torch.isfinite(g["u"]).all() and torch.isfinite(g["T"]).all()
    assert float(g["u"].abs().max()) < 80.0
    assert 150.0 < float(g["T"].min()) and float(g["T"].max()) < 320.0
    assert torch.allclose(m.state["lnps"][:, 0], q0, rtol=1e-6, atol=0)
    m.close()


def test_truncation_tendency_learned(mods):
    dy, ln = mods
    torch.manual_seed(0)
    x, y = ln.generate(list(range(16)), nsamples=2)
    xv, yv = ln.generate(list(range(500, 508)), nsamples=2)
    net, nx, ny, losses = ln.fit(x, y, steps=1000, batch=4096)
    assert losses[-1] < losses[0]
    held = ln.ab_score(net, nx, ny, xv, yv)
    # the corrected coarse tendency is closer to the truncated fine-run tendency than the
    # plain one on runs not used for training (mean squared error ratio, normalised targets)
    assert held["all_normalised"] < 0.95, held
    # and the corrected coarse model's one-hour forecast is closer to the truncated fine run
    fc = ln.forecast_score(net, nx, ny, list(range(700, 708)))
    assert fc["vor"] < 1.0 and fc["div"] < 1.0 and fc["temp"] < 1.0, fc
    # the coupling variants: chosen on validation runs; the baroclinic divergence and
    # temperature corrections have no column mean (the surface-pressure tendency and the
    # column's mean temperature are left to the plain model)
    name, kw, val = ln.choose_corrector(net, nx, ny, list(range(800, 808)))
    assert name in ln.CORRECTOR_VARIANTS and set(val) == set(ln.CORRECTOR_VARIANTS)
    m = dy.DryPrimitiveModel(31, 8, nmembers=2)
    m.set_state(m.jet_state([3, 4]))
    tend, g = m.tendencies(m.state, with_grid=True)
    t0 = {k: v.clone() for k, v in tend.items()}
    ln.corrector(net, nx, ny, baroclinic_div=True, baroclinic_temp=True)(m, tend, g)
    for f in ("div", "temp"):
        dd = (tend[f] - t0[f]).view(2, 8, -1)
        # FP32: the difference carries the rounding of the sums t0 + correction (half an ulp
        # of the tendency) besides the round-off of the correction's own mean removal
        tol = 1e-6 * (float(t0[f].abs().max()) + float(dd.abs().max()))
        assert float(dd.mean(dim=1).abs().max()) <= tol, f
        assert float(dd.abs().max()) > 0, f
    assert torch.equal(tend["lnps"], t0["lnps"])
    m.close()
    # a corrected coarse run stays bounded
    m = dy.DryPrimitiveModel(31, 8, nmembers=2)
    m.set_state(m.jet_state([7, 8]))
    m.run(int(6 * 3600 / m.dt), correction=ln.corrector(net, nx, ny))
    _, g = m.tendencies(m.state, with_grid=True)
    assert torch.isfinite(g["u"]).all() and float(g["u"].abs().max()) < 80.0
    m.close()
    net.close()
