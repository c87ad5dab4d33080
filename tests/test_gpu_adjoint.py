"""GPU: the hand-written adjoint (include/hc.h hc_adj_*).

Cases: fully wet 40x40 fractal grids at 5 m draining to their south edge,
per-cell log-normal Manning n, 8 rain blocks (2 x 4), 200 steps at a fixed
dt = 0.5 s; the loss is the outlet-hydrograph misfit against a target made
with n perturbed by 10 % smooth noise.  "steep" (1 % slope, 2 m relief,
0.3 m deep) runs near the Froude cap; "mild" (0.2 % slope, 0.3 m relief,
0.5 m deep, Froude number about 0.4) keeps the limiters off.

  * the forward is the FP64 plain form stepped at fixed dt: its discharge
    series matches the oracle driven call by call (same rain, same dt, mask
    every 5th step) to 1e-10;
  * gradients agree with central finite differences of the forward loss:
    eps = 1e-4 on the mild case to 2e-4 relative (n, along a smooth
    direction) and 1e-6 (each rain block); eps = 1e-6 on the steep case to
    1e-6, since there eps = 1e-4 already moves some faces across a Froude or
    depth-cap branch (the loss is only piecewise smooth in n);
  * the checkpoint interval does not change the gradients (bit-identical);
  * the torch.autograd.Function wrapper returns the same gradients.
"""
import numpy as np
import pytest

import hail_caesar_amd.inputs as inputs
from hail_caesar_amd import hc
from hcutil import hc_params_from
from pyoracle import OracleModel, pad, params_from_dict

pytestmark = pytest.mark.gpu

N, DX, DT, STEPS = 40, 5.0, 0.5, 200


def _smooth(rng, shape, sigma=3.0):
    """Gaussian-smoothed white noise scaled to unit max (FFT convolution)."""
    xi = rng.standard_normal(shape)
    ky = np.fft.fftfreq(shape[0])[:, None]
    kx = np.fft.fftfreq(shape[1])[None, :]
    g = np.exp(-2 * (np.pi * sigma) ** 2 * (kx ** 2 + ky ** 2))
    s = np.real(np.fft.ifft2(np.fft.fft2(xi) * g))
    return s / np.abs(s).max()


def _case(seed=5000, slope=0.01, relief=2.0, depth=0.3, channel=0.5):
    rng = np.random.Generator(np.random.PCG64(seed))
    dem = inputs.planar_dem(N, N, DX, slope, relief, 0.6, seed, channel_depth=channel)
    n = inputs.lognormal_mannings(N, N, 0.04, 0.2, rng)
    params = inputs.params_text_dict({"spatially_variable_mannings_on": "yes"})
    p = params_from_dict(params, N, N, DX)
    wet = pad(np.where(dem > -9999, depth, 0.0), 0.0)
    block = -np.ones((N + 2, N + 2), np.int32)
    rows = (np.arange(N) * 2) // N
    cols = (np.arange(N) * 4) // N
    block[1:-1, 1:-1] = rows[:, None] * 4 + cols[None, :]
    t = np.arange(STEPS)
    rain = (20.0 + 30.0 * np.sin(t / 40.0) ** 2) / 3.6e6  # m/s
    npert = n * (1 + 0.1 * _smooth(rng, n.shape))
    return p, pad(dem, -9999.0), pad(n, 0.0), pad(npert, 0.0), wet, block, rain


def _context(p, elev, n, wet):
    ctx = hc.Context(p.imax, p.jmax, p.DX, hc_params_from(p), hc.HC_FP64)
    ctx.upload("elev", elev)
    ctx.upload("water_depth", wet)
    ctx.upload("mannings", n)
    return ctx


def _oracle_forward(p, elev, n, wet, block, rain, mult):
    """The same fixed-dt forward, driven call by call through the oracle."""
    o = OracleModel(p)
    o.field("elev")[...] = elev
    o.field("mannings")[...] = n
    h = o.field("water_depth")
    h[...] = wet
    o.call("scan_area")
    live = (block >= 0) & (elev > -9999)
    Q = np.zeros(STEPS)
    for s in range(STEPS):
        h[live] = h[live] + rain[s] * DT * np.asarray(mult)[block[live]]
        o.scalar("time_step", DT)
        o.scalar("maxdepth", 0.1)  # keeps the Courant cap above dt
        o.call("flow_route")
        o.call("depth_update")
        Q[s] = o.call("water_flux_out")
        if (s + 1) % 5 == 0:
            o.call("scan_area")
    return Q, o.field("water_depth").copy()


@pytest.fixture(scope="module")
def case():
    return _case()


@pytest.fixture(scope="module")
def mild():
    return _case(slope=0.002, relief=0.3, depth=0.5, channel=0.1)


def _adjoint(case, C=20):
    p, elev, n, npert, wet, block, rain = case
    ctx = _context(p, elev, npert, wet)
    adj = hc.Adjoint(ctx, STEPS, DT, block, rain, checkpoint_every=C)
    _, qt = adj.forward(np.ones(8))
    adj.set_target(qt)
    ctx.upload("mannings", n)
    return ctx, adj


MULT = 1.0 + 0.05 * np.arange(8)


def test_forward_matches_oracle(case):
    p, elev, n, npert, wet, block, rain = case
    ctx, adj = _adjoint(case)
    loss, Q = adj.forward(MULT)
    Qo, ho = _oracle_forward(p, elev, n, wet, block, rain, MULT)
    assert Qo.max() > 0
    np.testing.assert_allclose(Q, Qo, rtol=1e-10, atol=1e-12 * Qo.max())
    assert loss > 0


def _dn(adj, ctx, n, v, eps):
    ctx.upload("mannings", n * (1 + eps * v))
    lp, _ = adj.forward(MULT)
    ctx.upload("mannings", n * (1 - eps * v))
    lm, _ = adj.forward(MULT)
    ctx.upload("mannings", n)
    return (lp - lm) / (2 * eps)


def test_gradient_vs_central_differences(mild):
    p, elev, n, npert, wet, block, rain = mild
    ctx, adj = _adjoint(mild)
    loss0, _ = adj.forward(MULT)
    gn, gm = adj.backward()
    eps = 1e-4
    # Manning n along a smooth relative direction v: d/de L(n (1 + e v))
    v = pad(_smooth(np.random.Generator(np.random.PCG64(9)), (N, N)), 0.0)
    fd = _dn(adj, ctx, n, v, eps)
    an = float(np.sum(gn * n * v))
    print(f"n: adjoint {an:.10e} fd {fd:.10e} loss {loss0:.6e}")
    assert abs(an - fd) <= 2e-4 * abs(fd)
    for b in range(8):
        m = MULT.copy()
        m[b] += eps
        lp, _ = adj.forward(m)
        m[b] -= 2 * eps
        lm, _ = adj.forward(m)
        fd = (lp - lm) / (2 * eps)
        print(f"block {b}: adjoint {gm[b]:.10e} fd {fd:.10e}")
        assert abs(gm[b] - fd) <= 1e-6 * abs(fd)


def test_gradient_near_limiters(case):
    """Near the Froude cap the loss is piecewise smooth: the finite
    difference converges to the adjoint once eps is inside every branch."""
    p, elev, n, npert, wet, block, rain = case
    ctx, adj = _adjoint(case)
    adj.forward(MULT)
    gn, _ = adj.backward()
    v = pad(_smooth(np.random.Generator(np.random.PCG64(9)), (N, N)), 0.0)
    an = float(np.sum(gn * n * v))
    fd = _dn(adj, ctx, n, v, 1e-6)
    print(f"steep n: adjoint {an:.10e} fd(1e-6) {fd:.10e}")
    assert abs(an - fd) <= 1e-6 * abs(fd)


def test_checkpoint_interval_invariance(case):
    grads = []
    for C in (1, 13, STEPS):
        ctx, adj = _adjoint(case, C=C)
        adj.forward(MULT)
        grads.append(adj.backward())
    for gn, gm in grads[1:]:
        assert np.array_equal(gn, grads[0][0])
        assert np.array_equal(gm, grads[0][1])


def test_autograd_function(case):
    torch = pytest.importorskip("torch")
    from hc_autograd import HydrographMisfit
    p, elev, n, npert, wet, block, rain = case
    ctx, adj = _adjoint(case)
    nt = torch.tensor(n, dtype=torch.float64, requires_grad=True)
    mt = torch.tensor(MULT, dtype=torch.float64, requires_grad=True)
    loss = HydrographMisfit.apply(nt, mt, adj)
    loss.backward()
    adj.forward(MULT)
    gn, gm = adj.backward()
    assert np.array_equal(nt.grad.numpy(), gn)
    assert np.array_equal(mt.grad.numpy(), gm)


# ---- the adjoint workload at its stated shape: 512^2 at 5 m, 2000 steps at
# dt = 0.5 s, checkpoints every 100 steps (inputs.adjoint_case)
BIG, BIG_STEPS = 512, 2000


def _big(channel):
    c = inputs.adjoint_case(size=BIG, steps=BIG_STEPS, channel=channel)
    p = hc.default_params(hflow_threshold=0.001, edgeslope=0.01, spatially_var_mannings=1)
    ctx = hc.Context(BIG, BIG, c["DX"], p, hc.HC_FP64)
    ctx.upload("elev", pad(c["dem"], -9999.0))
    ctx.upload("water_depth", pad(c["water"], 0.0))
    ctx.upload("mannings", pad(c["n_target"], 0.0))
    return c, ctx


def _big_adjoint(c, ctx, nsteps):
    adj = hc.Adjoint(ctx, nsteps, 0.5, c["block"], c["rain"][:nsteps], checkpoint_every=100)
    _, qt = adj.forward(np.ones(8))
    adj.set_target(qt)
    n = pad(c["n"], 0.0)
    ctx.upload("mannings", n)
    return adj, n


def _fd_blocks(adj, eps):
    out = []
    for b in range(8):
        m = np.ones(8)
        m[b] += eps
        lp, _ = adj.forward(m)
        m[b] -= 2 * eps
        lm, _ = adj.forward(m)
        out.append((lp - lm) / (2 * eps))
    return np.array(out)


def test_adjoint_workload_vs_central_differences():
    """The 512^2 / 2000-step workload on its fractal DEM: the gradients with
    respect to all 8 rain-block multipliers and to n along a smooth relative
    direction agree with central differences at eps = 1e-4, all taken at the
    multipliers 1."""
    c, ctx = _big(channel=0.0)
    adj, n = _big_adjoint(c, ctx, BIG_STEPS)
    loss, q = adj.forward(np.ones(8))
    gn, gm = adj.backward()
    fd = _fd_blocks(adj, 1e-4)
    v = pad(inputs.smooth_noise(np.random.Generator(np.random.PCG64(9)), (BIG, BIG), 8.0), 0.0)
    eps = 1e-4
    ctx.upload("mannings", n * (1 + eps * v))
    lp, _ = adj.forward(np.ones(8))
    ctx.upload("mannings", n * (1 - eps * v))
    lm, _ = adj.forward(np.ones(8))
    ctx.upload("mannings", n)
    fdn = (lp - lm) / (2 * eps)
    an = float(np.sum(gn * n * v))
    print(f"loss {loss:.10e}")
    for b in range(8):
        print(f"block {b}: adjoint {gm[b]:.10e} fd {fd[b]:.10e} rel {abs(gm[b]-fd[b])/abs(fd[b]):.2e}")
    print(f"n: adjoint {an:.10e} fd {fdn:.10e} rel {abs(an-fdn)/abs(fdn):.2e}")
    # measured: blocks within 2.9e-4 (the O(eps) curvature of the central
    # difference and a few faces crossing a limiter), n within 7e-5
    assert loss > 0
    assert np.all(np.abs(gm - fd) <= 1e-3 * np.abs(fd))
    assert abs(an - fdn) <= 1e-3 * abs(fdn)
    adj.close()
    ctx.close()


def test_adjoint_workload_forward_matches_oracle():
    """The same 2000-step forward against the oracle driven call by call (the
    rain added to each block's cells, flow_route, depth_update, water_flux_out
    at dt = 0.5 s, scan_area every 5th step): discharge series to 1e-10."""
    import os
    c, ctx = _big(channel=0.0)
    adj, n = _big_adjoint(c, ctx, BIG_STEPS)
    _, q = adj.forward(np.ones(8))
    params = inputs.params_text_dict({"spatially_variable_mannings_on": "yes",
                                      "hflow_threshold": "0.001", "slope_on_edge_cell": "0.01"})
    p = params_from_dict(params, BIG, BIG, c["DX"], nthreads=min(16, os.cpu_count() or 1))
    o = OracleModel(p)
    elev = pad(c["dem"], -9999.0)
    o.field("elev")[...] = elev
    o.field("mannings")[...] = n
    h = o.field("water_depth")
    h[...] = pad(c["water"], 0.0)
    o.call("scan_area")
    block, rain = c["block"], c["rain"]
    live = (block >= 0) & (elev > -9999)
    qo = np.zeros(BIG_STEPS)
    for s in range(BIG_STEPS):
        h[live] = h[live] + rain[s] * 0.5 * 1.0
        o.scalar("time_step", 0.5)
        o.scalar("maxdepth", 0.1)
        o.call("flow_route")
        o.call("depth_update")
        qo[s] = o.call("water_flux_out")
        if (s + 1) % 5 == 0:
            o.call("scan_area")
    np.testing.assert_allclose(q, qo, rtol=1e-10, atol=1e-12 * qo.max())
    adj.close()
    ctx.close()


def test_adjoint_exact_in_a_chaotic_channel():
    """The same workload with a 2 m channel carved down the middle: there the
    fixed-dt scheme is chaotic (a 1e-9 change of one rain multiplier grows by
    about 2 % per step to metres of depth), so the loss is not smooth at the
    scale of eps = 1e-4 and the tangent is large.  After 1000 steps the
    adjoint still equals the central difference taken inside the smooth
    piece (eps = 1e-8) for every block: the reverse sweep is the exact
    derivative of the forward it differentiates."""
    c, ctx = _big(channel=2.0)
    adj, n = _big_adjoint(c, ctx, 1000)
    adj.forward(np.ones(8))
    gn, gm = adj.backward()
    fd8 = _fd_blocks(adj, 1e-8)
    fd4 = _fd_blocks(adj, 1e-4)
    for b in range(8):
        print(f"block {b}: adjoint {gm[b]:.8e} fd(1e-8) {fd8[b]:.8e} fd(1e-4) {fd4[b]:.8e}")
    assert np.all(np.abs(gm - fd8) <= 1e-3 * np.abs(gm) + 0.02)
    adj.close()
    ctx.close()
