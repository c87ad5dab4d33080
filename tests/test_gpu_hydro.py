"""GPU parity of the hydro hot path (libhc.so through the C-ABI) against the
oracle, one method call at a time and over many steps.

Tolerances: the FP64 form follows the reference operation order; the only
rounding difference is hflow^3 (libm pow vs a correctly rounded cube), so
per-call results agree to the last bit in all but a handful of cells and
every field is held to 1e-12 relative (absolute floor 1e-15).  Multi-step
runs are held to the 1e-10 relative bar.
"""
import numpy as np
import pytest

from hail_caesar_amd import hc
from hcutil import rel_err, to_context, wet_flow_state

pytestmark = pytest.mark.gpu


def _close(a, b, rtol, atol):
    np.testing.assert_allclose(a, b, rtol=rtol, atol=atol)


@pytest.mark.parametrize("spatial_n", [False, True])
def test_flow_route_depth_update_per_call(tmp_path, spatial_n):
    w, o = wet_flow_state(tmp_path, size=64, steps=400, spatial_n=spatial_n)
    ctx = to_context(o)
    dt = o.call("get_flow_timestep")
    o.call("flow_route")
    ctx.flow_route(dt)
    for f in ("qx", "qy"):
        _close(ctx.download(f), o.field(f), 1e-12, 1e-15)
    md_o = o.call("depth_update")
    md_g = ctx.depth_update(dt)
    _close(ctx.download("water_depth"), o.field("water_depth"), 1e-12, 1e-15)
    assert abs(md_g - md_o) <= 1e-12 * md_o


def test_fused_matches_separate(tmp_path):
    w, o = wet_flow_state(tmp_path, size=96, steps=300)
    a = to_context(o)
    b = to_context(o)
    dt = o.call("get_flow_timestep")
    a.flow_route(dt)
    md_a = a.depth_update(dt)
    md_b = b.flow_depth(dt)
    for f in ("water_depth", "qx", "qy"):
        x, y = a.download(f), b.download(f)
        assert np.array_equal(x, y), f
    assert md_a == md_b


def test_scan_area_exact(tmp_path):
    w, o = wet_flow_state(tmp_path, size=80, steps=200)
    ctx = to_context(o)
    o.call("scan_area")
    ctx.scan_area()
    assert np.array_equal(ctx.download("mask"), o.field_int("mask"))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_scan_area_exact_ragged(tmp_path, prec):
    """scan_area over several strips of rows (62 rows per work item for 4-byte
    depths, 30 for 8-byte ones) and several 64-column mask words, on a ragged
    wet/dry pattern: 60% of the cells dried at random, and single wet cells
    in dry surroundings on the word seams (columns 63/64, 127/128), on the
    strips' first and last rows and on the grid's edges."""
    from hail_caesar_amd import hc
    w, o = wet_flow_state(tmp_path, size=203, steps=50)
    h = o.field("water_depth")  # a view of the oracle's state
    rng = np.random.Generator(np.random.PCG64(77))
    h[rng.random(h.shape) < 0.6] = 0.0
    for x0, y0 in [(40, 63), (40, 64), (100, 127), (100, 128), (62, 20), (63, 150),
                   (30, 150), (31, 170), (1, 5), (203, 7), (9, 1), (11, 203)]:
        h[max(x0 - 2, 0):x0 + 3, max(y0 - 2, 0):y0 + 3] = 0.0
        h[x0, y0] = 0.25
    ctx = to_context(o, precision=hc.HC_FP32 if prec == "fp32" else hc.HC_FP64)
    o.call("scan_area")
    ctx.scan_area()
    assert np.array_equal(ctx.download("mask"), o.field_int("mask"))


def test_water_flux_out(tmp_path):
    w, o = wet_flow_state(tmp_path, size=64, steps=500)
    # wet every edge so that all four edges and the corners drain
    h = o.field("water_depth")
    h[1, 1:-1] += 0.05
    h[-2, 1:-1] += 0.03
    h[1:-1, 1] += 0.02
    h[1:-1, -2] += 0.04
    ctx = to_context(o)
    ts = 3.7
    o.scalar("time_step", ts)
    t_o = o.call("water_flux_out")
    t_g = ctx.water_flux_out(ts)
    assert np.array_equal(ctx.download("water_depth"), o.field("water_depth"))
    assert abs(t_g - t_o) <= 1e-13 * t_o


def test_point_inputs(tmp_path):
    w, o = wet_flow_state(tmp_path, size=64, steps=0)
    ctx = to_context(o)
    h = o.field("water_depth").copy()
    xs, ys = [1, 5, 64, 10], [1, 7, 64, 33]
    amts = [0.001, 0.25, 1e-9, 0.05]
    for x, y, a in zip(xs, ys, amts):
        h[x, y] += a
    ctx.add_point_inputs(xs, ys, amts)
    assert np.array_equal(ctx.download("water_depth"), h)


@pytest.mark.parametrize("precision", [0, 1])
def test_point_inputs_repeated_cells(tmp_path, precision):
    """A cell listed more than once takes its amounts in turn, in list order
    (the reference's sequential loop, LCM.cpp:3463-3477): 3000 points over
    200 cells, against the same adds made one by one on the host (FP32
    contexts round after each add, as the device does)."""
    w, o = wet_flow_state(tmp_path, size=64, steps=0)
    ctx = to_context(o, precision=precision)
    rng = np.random.Generator(np.random.PCG64(77))
    cells = rng.integers(0, 66, size=(200, 2))
    pick = rng.integers(0, 200, size=3000)
    xs, ys = cells[pick, 0], cells[pick, 1]
    amts = rng.uniform(1e-6, 1e-2, size=3000)
    dt = np.float64 if precision == 0 else np.float32
    h = ctx.download("water_depth").astype(dt)
    for x, y, a in zip(xs, ys, amts):
        h[x, y] = dt(np.float64(h[x, y]) + a)
    ctx.add_point_inputs(xs, ys, amts)
    assert np.array_equal(ctx.download("water_depth"), h.astype(np.float64))


@pytest.mark.parametrize("precision", [0, 1])
def test_persistent_point_list(tmp_path, precision):
    """hc_points_set + hc_points_add (the list kept on the device, amounts by
    rain cell) give the bits of hc_add_point_inputs with each point's amount
    written out, over three steps with a list change between them; a repeated
    cell, an out-of-grid point and an add before any list are refused."""
    w, o = wet_flow_state(tmp_path, size=64, steps=0)
    a = to_context(o, precision=precision)
    b = to_context(o, precision=precision)
    rng = np.random.Generator(np.random.PCG64(5))
    with pytest.raises(hc.HCError):
        a.points_add([0.0, 1.0])
    for step in range(3):
        if step != 1:  # a new list on steps 0 and 2
            flat = rng.choice(66 * 66, size=1500, replace=False)
            xs, ys = flat // 66, flat % 66
            rc = rng.integers(0, 4, size=len(xs))
            a.points_set(xs, ys, rc)
        amt = rng.uniform(0, 1e-3, size=4)
        a.points_add(amt)
        b.add_point_inputs(xs, ys, amt[rc])
        assert np.array_equal(a.download("water_depth"), b.download("water_depth")), step
    with pytest.raises(hc.HCError):
        a.points_set([3, 3], [4, 4], [1, 1])
    with pytest.raises(hc.HCError):
        a.points_set([70], [1], [1])


def test_multistep_fused_vs_oracle(tmp_path):
    """200 oracle iterations replayed on the device kernels with the oracle's
    own dt sequence and input-point injections reproduced from its state."""
    w, o = wet_flow_state(tmp_path, size=64, steps=600)
    ctx = to_context(o)
    for k in range(200):
        dt = o.call("get_flow_timestep")
        o.call("flow_route")
        o.call("depth_update")
        ctx.flow_depth(dt)
        if k % 5 == 4:
            o.call("scan_area")
            ctx.scan_area()
    for f in ("water_depth", "qx", "qy"):
        assert rel_err(ctx.download(f), o.field(f), 1e-12) < 1e-10, f


def test_fixed_point_elevations_high_datum(tmp_path):
    """Fixed-point elevations keep the FP32 form's face differences exact where
    single-precision elevations cannot: the oracle state raised by 2000 m
    (float32 spacing there 1.2e-4 m).  30 steps on the oracle's own dt
    sequence: the fixed-point and FP64-elevation forms stay within 1e-5 of the
    oracle (FP64) in relative depth, and within 5e-6 of each other (both carry
    the FP32 arithmetic's own ~1e-6); the single-precision-elevation form is at
    least ten times further off."""
    w, o = wet_flow_state(tmp_path, size=64, steps=600)
    e = o.field("elev")
    e[e > -9999] += 2000.0
    forms = {"fixed": dict(elev_fixed=True), "z64": dict(flat_dz=-1.0), "z32": dict(flat_dz=0.0)}
    ctxs = {k: to_context(o, 1, **kw) for k, kw in forms.items()}
    for k in range(30):
        dt = o.call("get_flow_timestep")
        o.call("flow_route")
        o.call("depth_update")
        for c in ctxs.values():
            c.flow_depth(dt)
        if k % 5 == 4:
            o.call("scan_area")
            for c in ctxs.values():
                c.scan_area()
    h = o.field("water_depth")
    wet = h > 1e-3
    err = {k: float(np.max(np.abs(c.download("water_depth") - h)[wet] / h[wet]))
           for k, c in ctxs.items()}
    hf, hz = ctxs["fixed"].download("water_depth"), ctxs["z64"].download("water_depth")
    between = float(np.max(np.abs(hf - hz)[wet] / h[wet]))
    print(f"relative depth error vs the oracle: {err}, fixed vs z64 {between}")
    assert err["fixed"] < 1e-5 and err["z64"] < 1e-5
    assert between < 5e-6
    assert err["z32"] > 10 * err["fixed"]


@pytest.mark.parametrize("relief", ["flat", "alpine"])
def test_fixed_point_elevations_extreme_relief(relief):
    """Fixed-point quantisation at the ends of its range: a DEM with no relief
    (the scale is capped, so both forms see dz = 0 and agree exactly), and a
    gently sloping plateau at 8000 m with one pit at -500 m, so the quantum is
    coarse (8500 m / 2^31 = 4e-6 m, against face differences of ~5 cm).  The
    uploaded DEM downloads unchanged; 20 steps track the FP64-elevation form
    to 1e-4 relative depth (flat: exactly)."""
    from hail_caesar_amd import hc
    n = 96
    rng = np.random.Generator(np.random.PCG64(7))
    if relief == "flat":
        dem = np.full((n, n), 12.5)
    else:
        yy, xx = np.mgrid[0:n, 0:n]
        dem = 8000.0 - 0.05 * yy + rng.normal(0, 0.01, (n, n))
        dem[n // 2, 5] = -500.0
    e = np.full((n + 2, n + 2), -9999.0)
    e[1:-1, 1:-1] = dem
    e[1:-1, 1] = -9999.0  # a no-data column inside the grid
    hgt = np.zeros_like(e)
    hgt[1:-1, 1:-1] = 0.3 + 0.1 * rng.random((n, n))
    mask = np.zeros(e.shape, np.int32)
    mask[1:-1, 1:-1] = 1
    ctxs = {}
    for name, kw in (("fixed", dict(elev_fixed=1)), ("z64", dict(flat_dz=-1.0))):
        p = hc.default_params(hflow_threshold=0.001, edgeslope=0.005, mannings=0.04, **kw)
        c = hc.Context(n, n, 5.0, p, hc.HC_FP32)
        c.upload("elev", e)
        c.upload("water_depth", hgt)
        c.upload("mask", mask)
        ctxs[name] = c
    assert np.array_equal(ctxs["fixed"].download("elev"), e)
    for _ in range(20):
        for c in ctxs.values():
            c.flow_depth(0.02)
    hf, hz = ctxs["fixed"].download("water_depth"), ctxs["z64"].download("water_depth")
    wet = hz > 1e-3
    err = float(np.max(np.abs(hf - hz)[wet] / hz[wet]))
    print(f"{relief}: fixed vs FP64 elevations, max relative depth difference {err:.3g}")
    assert err == 0.0 if relief == "flat" else err < 1e-4


def test_health_check(tmp_path):
    w, o = wet_flow_state(tmp_path, size=32, steps=0)
    ctx = to_context(o)
    assert ctx.check_health() == 0
    h = o.field("water_depth").copy()
    h[3, 4] = np.nan
    h[5, 6] = -1.0
    ctx.upload("water_depth", h)
    assert ctx.check_health() == 2


@pytest.mark.parametrize("name", ["flow_step_64.npz", "flow_step_spatial_n_48.npz"])
def test_per_call_vs_reference_golden(name):
    """Vectors written by the compiled reference (tests/golden/make_golden.py)."""
    import json
    import os
    from hail_caesar_amd import hc
    from hcutil import hc_params_from
    from pyoracle import params_from_dict
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name))
    params = json.loads(str(g["params_json"]))
    p = params_from_dict(params, int(g["imax"]), int(g["jmax"]), float(g["DX"]))
    ctx = hc.Context(p.imax, p.jmax, p.DX, hc_params_from(p))
    ctx.upload("elev", g["elev"])
    ctx.upload("water_depth", g["h0"])
    ctx.upload("qx", g["qx0"])
    ctx.upload("qy", g["qy0"])
    ctx.upload("mask", g["mask0"])
    if "n" in g.files:
        ctx.upload("mannings", g["n"])
    dt = float(g["dt"])
    ctx.flow_route(dt)
    _close(ctx.download("qx"), g["qx1"], 1e-12, 1e-15)
    _close(ctx.download("qy"), g["qy1"], 1e-12, 1e-15)
    md = ctx.depth_update(dt)
    _close(ctx.download("water_depth"), g["h1"], 1e-12, 1e-15)
    assert abs(md - float(g["maxdepth1"])) <= 1e-12 * md
    ctx.scan_area()
    assert np.array_equal(ctx.download("mask"), g["mask1"])
    t = ctx.water_flux_out(float(g["time_step0"]))
    assert abs(t - float(g["temptot1"])) <= 1e-12 * max(abs(t), 1e-300)
    _close(ctx.download("water_depth"), g["h2"], 1e-12, 1e-15)
