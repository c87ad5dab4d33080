"""End-to-end GPU runs through the C++ run-loop driver (include/hc_model.h),
compared with the oracle and the reference's known-good answers.

Bars: FP64 plain form within 1e-10 relative of the oracle (absolute floor
1e-13 for fields, 1e-12 m^3/s for discharges); device-resident stepping
bit-identical to host-stepped; FP32 storage form within 1e-4 of the FP64 form
in max-norm relative depth error.
"""
import os
import shutil

import numpy as np
import pytest

import hail_caesar_amd.inputs as inputs
from hail_caesar_amd import hc
from hail_caesar_amd.model import Model
from pyoracle import OracleModel

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _flow(tmp, size, hours=2, seed=1001):
    return inputs.flow_only_256(os.path.join(str(tmp), f"flow{size}"), size=size, hours=hours, seed=seed)


def _assert_rel(a, b, rtol, atol, what):
    np.testing.assert_allclose(a, b, rtol=rtol, atol=atol, err_msg=what)


def test_driver_fp64_vs_oracle(tmp_path):
    w = _flow(tmp_path, 64)
    o = OracleModel.from_workload(w)
    ot, oc, od = o.iterate(10 ** 7)
    m = Model(w.path, w.params_file)
    gt, gc, gd = m.iterate(10 ** 7)
    assert m.finished and len(gt) == len(ot)
    _assert_rel(gc, oc, 1e-12, 0, "cycle")
    _assert_rel(gd, od, 1e-10, 0, "dt")
    _assert_rel(gt, ot, 1e-10, 1e-12, "temptot")
    for f in ("water_depth", "qx", "qy"):
        _assert_rel(m.download(f), o.field(f), 1e-10, 1e-13, f)
    _assert_rel(m.series(), o.series(), 1e-10, 1e-12, "series")


def test_driver_fused_equals_separate(tmp_path):
    w = _flow(tmp_path, 48, hours=1)
    a = Model(w.path, w.params_file, fused=True)
    b = Model(w.path, w.params_file, fused=False)
    ta, _, _ = a.iterate(10 ** 7)
    tb, _, _ = b.iterate(10 ** 7)
    assert np.array_equal(ta, tb)
    for f in ("water_depth", "qx", "qy"):
        assert np.array_equal(a.download(f), b.download(f)), f


@pytest.mark.parametrize("use_graph,chunks", [(False, (333,)), (True, (333,)),
                                              (True, (10, 40, 1500, 7, 2600, 20))])
def test_device_loop_equals_host_loop(tmp_path, use_graph, chunks):
    """Device-resident stepping replays exactly the host loop; the chunk
    pattern with growing sizes makes the record buffer grow under captured
    graphs, which must then be captured again."""
    w = _flow(tmp_path, 64)
    host = Model(w.path, w.params_file)
    th, ch, dh = host.iterate(10 ** 7)
    dev = Model(w.path, w.params_file, device_loop=True, use_graph=use_graph)
    parts = []
    k = 0
    while True:
        t, c, d = dev.iterate(chunks[k % len(chunks)])
        k += 1
        parts.append((t, c, d))
        if dev.finished:
            break
    td = np.concatenate([p[0] for p in parts])
    cd = np.concatenate([p[1] for p in parts])
    dd = np.concatenate([p[2] for p in parts])
    assert len(td) == len(th)
    assert np.array_equal(cd, ch) and np.array_equal(dd, dh)
    np.testing.assert_allclose(td, th, rtol=1e-13, atol=1e-15)
    for f in ("water_depth", "qx", "qy"):
        assert np.array_equal(dev.download(f), host.download(f)), f
    np.testing.assert_allclose(dev.series(), host.series(), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("fixed", [True, False])
def test_fp32_device_loop_equals_host_loop(tmp_path, fixed):
    """FP32 forms (fixed-point and FP64 elevations): the device loop, whose
    fused kernel adds the rain and drains the edges and whose scan_area
    dilates the wet-cell bits that kernel stores on scan bodies, replays the
    host loop (point adds, edge drain and scan from the depths as separate
    calls) bit for bit."""
    w = _flow(tmp_path, 64)
    kw = dict(precision=hc.HC_FP32, flat_dz=0.0 if fixed else -1.0, elev_fixed=fixed)
    host = Model(w.path, w.params_file, **kw)
    th, ch, dh = host.iterate(10 ** 7)
    dev = Model(w.path, w.params_file, device_loop=True, use_graph=True, **kw)
    parts = [dev.iterate(1500)]
    while not dev.finished:
        parts.append(dev.iterate(1500))
    td = np.concatenate([p[0] for p in parts])
    cd = np.concatenate([p[1] for p in parts])
    dd = np.concatenate([p[2] for p in parts])
    assert len(td) == len(th)
    assert np.array_equal(cd, ch) and np.array_equal(dd, dh)
    np.testing.assert_allclose(td, th, rtol=1e-13, atol=1e-15)
    for f in ("water_depth", "qx", "qy"):
        assert np.array_equal(dev.download(f), host.download(f)), f


def test_wet_bit_scan_equals_depth_scan(tmp_path):
    """After a device-loop body that rebuilds the active mask (every 5th),
    the mask dilated from the fused kernel's wet bits equals scan_area run
    on the depths that body stored; checked at several points of a
    partially wet FP32 run."""
    w = _flow(tmp_path, 96)
    dev = Model(w.path, w.params_file, precision=hc.HC_FP32, elev_fixed=True, device_loop=True,
                use_graph=True)
    ctx = dev.context()
    dev.iterate(1)
    for n in (25, 100, 1000, 1400):  # the run ends after 2701 bodies
        dev.iterate(n - int(dev.scalar("counter")) % 5)  # end on a scan body
        assert int(dev.scalar("counter")) % 5 == 0 and not dev.finished
        m_loop = ctx.download("mask")
        ctx.scan_area()
        m_scan = ctx.download("mask")
        inner = m_scan[1:-1, 1:-1]
        assert 0 < inner.sum() < inner.size, "the test needs a partially wet grid"
        assert np.array_equal(m_loop, m_scan), n


def _bos(tmp):
    src = os.path.join(GOLDEN, "boscastle")
    dst = os.path.join(str(tmp), "bos")
    os.makedirs(os.path.join(dst, "out"), exist_ok=True)
    for f in ("boscastle_square_50m.asc", "boscastle_72hr_rain_u.txt"):
        shutil.copy(os.path.join(src, f), dst)
    out = []
    for ln in open(os.path.join(src, "boscastle_test_72hr_50m_u.params")).read().splitlines():
        if ln.startswith("read_path:"):
            ln = "read_path: ./"
        elif ln.startswith("write_path:"):
            ln = "write_path: ./out/"
        elif ln.startswith("raster_output_interval:"):
            ln = "raster_output_interval: 3360"
        out.append(ln)
    open(os.path.join(dst, "bos.params"), "w").write("\n".join(out) + "\n")
    return dst, "bos.params"


def test_driver_boscastle_known_good_answers(tmp_path):
    """The reference's regression case end to end on the device: hydrograph
    and the WaterDepths3360 raster against the known-good answers
    (test/basic_kga_tests.py compares them at rtol 1e-3)."""
    path, pf = _bos(tmp_path)
    m = Model(path, pf, write_files=True)
    t, c, d = m.iterate(10 ** 7)
    assert m.finished and len(t) == 88450
    kga = np.loadtxt(os.path.join(GOLDEN, "boscastle", "boscastle_50m_72hr_u.dat"))
    got = np.loadtxt(os.path.join(path, "out", "boscastle_50m_72hr_u.dat"))
    assert got.shape == kga.shape
    np.testing.assert_allclose(got[:, 1], kga[:, 1], rtol=1e-3, atol=1e-6)
    # full-precision series agree with the printed KGA to its last digit
    s = m.series()
    assert np.max(np.abs(s[:, 1] - kga[:, 1])) <= 5.01e-7
    ras = inputs.read_asc(os.path.join(path, "out", "WaterDepths3360.asc"))
    ref = inputs.read_asc(os.path.join(GOLDEN, "boscastle", "WaterDepths3360.asc"))
    np.testing.assert_allclose(ras, ref, rtol=1e-3, atol=1e-6)
    # device-resident stepping gives the same hydrograph
    dev = Model(path, pf, device_loop=True, use_graph=True)
    while True:
        dev.iterate(20000)
        if dev.finished:
            break
    np.testing.assert_allclose(dev.series(), s, rtol=1e-12, atol=1e-14)


def _errs(a, b):
    # interior cells: the padding corner can hold the reference's stale-point
    # NaN (an empty input-point list injects 0/0 at cell (0,0))
    a, b = a[1:-1, 1:-1], b[1:-1, 1:-1]
    d = a - b
    return {"max": float(np.abs(d).max() / np.abs(b).max()),
            "l2": float(np.sqrt((d ** 2).sum() / (b ** 2).sum())),
            "vol": float(abs(a.sum() - b.sum()) / b.sum())}


def test_fp32_storage_form_error(tmp_path):
    """FP32 storage form against the FP64 form on the shallow rain-fed
    flow-only workload (128^2, 1 h).  That flow is ill-conditioned: in the
    FP64 oracle a 1e-8 m perturbation of the DEM already moves the depth
    field by ~1.5e-4 (L2) and ~7e-4 (max), through isolated threshold and
    limiter flips.  The bar is therefore the problem's own sensitivity floor
    (FP64, 1e-8 m DEM noise), with volume conserved to 1e-6."""
    w = _flow(tmp_path, 128, hours=1)
    a = OracleModel.from_workload(w, nthreads=8)
    a.iterate(10 ** 7)
    b = OracleModel.from_workload(w, nthreads=8, initialise=False)
    e = b.field("elev")
    m = e > -9999
    e[m] += 1e-8 * np.random.Generator(np.random.PCG64(3)).standard_normal(m.sum())
    b.L.orc_initialise(b.h)
    b.iterate(10 ** 7)
    floor = _errs(b.field("water_depth"), a.field("water_depth"))
    ref = Model(w.path, w.params_file, precision=hc.HC_FP64)
    ref.iterate(10 ** 7)
    h64 = ref.download("water_depth")
    res = {}
    for name, fdz, fixed in (("z32", 0.0, False), ("flat0.01", 0.01, False), ("exact_z", -1.0, False),
                             ("fixed_z", 0.0, True)):
        f32 = Model(w.path, w.params_file, precision=hc.HC_FP32, flat_dz=fdz, elev_fixed=fixed)
        f32.iterate(10 ** 7)
        res[name] = _errs(f32.download("water_depth"), h64)
    print(f"fp32 depth error vs fp64 (floor {floor}): {res}")
    for name in ("exact_z", "fixed_z"):
        assert res[name]["l2"] <= max(1e-4, 4 * floor["l2"]), name
        assert res[name]["max"] <= max(1e-4, 4 * floor["max"]), name
        assert res[name]["vol"] < 1e-6, name


def test_fp32_fully_wet_error(tmp_path):
    """Well-conditioned case: a uniformly wet (0.2 m) 128^2 fractal grid,
    200 loop bodies; the FP32 form (FP64 elevation differences) stays within
    1e-4 of the FP64 form in max-norm relative depth error."""
    dem = inputs.planar_dem(128, 128, 2.0, 0.002, 30.0, 0.6, 1003, channel_depth=3.0)
    wi = np.where(dem > -9999, 0.2, 0.0)
    n = inputs.lognormal_mannings(128, 128, 0.035, 0.3, np.random.Generator(np.random.PCG64(1)))
    w = inputs.make_workload(os.path.join(str(tmp_path), "wet"), dem, 2.0, np.array([8.0, 0.0]), 15, 1,
                             mannings=n, water_init=wi, overrides={"min_q_for_depth_calc": "0.01"},
                             name="wet")
    ref = Model(w.path, w.params_file, precision=hc.HC_FP64)
    ref.iterate(200)
    h64 = ref.download("water_depth")
    res = {}
    for name, fdz, fixed in (("z32", 0.0, False), ("exact_z", -1.0, False), ("fixed_z", 0.0, True)):
        f32 = Model(w.path, w.params_file, precision=hc.HC_FP32, flat_dz=fdz, elev_fixed=fixed)
        f32.iterate(200)
        res[name] = _errs(f32.download("water_depth"), h64)
        res[name]["qx_max"] = float(np.abs(f32.download("qx") - ref.download("qx")).max()
                                    / np.abs(ref.download("qx")).max())
    print(f"fp32 fully wet depth error vs fp64: {res}")
    assert res["exact_z"]["max"] < 1e-4
    assert res["fixed_z"]["max"] < 1e-4


@pytest.mark.parametrize("rfnum", [1, 3])
def test_driver_complex_rainfall_vs_oracle(tmp_path, rfnum):
    """spatially_complex_rainfall_on through the driver: the per-cell
    TOPMODEL on the host (host libm, as the reference), every cell's water
    added on the device (hc_runoff_add); FP64 within 1e-10 of the oracle
    (pinned bitwise to the compiled reference) over the whole run -- step
    records, fields, the water input and the series -- and the FP32 form
    (fixed-point elevations) against the FP64 form at the bar of the
    rain-fed shallow flow above: 1e-4 or four times the FP64 run's own
    sensitivity to 1e-8 m of DEM noise, volume to 1e-6."""
    w = inputs.complex_rain(os.path.join(str(tmp_path), f"c{rfnum}"), rfnum)
    o = OracleModel.from_workload(w)
    ot, oc, od = o.iterate(10 ** 7)
    m = Model(w.path, w.params_file)
    gt, gc, gd = m.iterate(10 ** 7)
    assert m.finished and len(gt) == len(ot) > 1000
    _assert_rel(gc, oc, 1e-12, 0, "cycle")
    _assert_rel(gd, od, 1e-10, 0, "dt")
    _assert_rel(gt, ot, 1e-10, 1e-12, "temptot")
    for f in ("water_depth", "qx", "qy"):
        _assert_rel(m.download(f), o.field(f), 1e-10, 1e-13, f)
    _assert_rel(m.series(), o.series(), 1e-10, 1e-12, "series")
    _assert_rel(m.scalar("waterinput"), o.scalar("waterinput"), 1e-10, 0, "waterinput")
    assert o.scalar("waterinput") > 0
    b = OracleModel.from_workload(w, initialise=False)
    e = b.field("elev")
    land = e > -9999
    e[land] += 1e-8 * np.random.Generator(np.random.PCG64(3)).standard_normal(land.sum())
    b.L.orc_initialise(b.h)
    b.iterate(10 ** 7)
    floor = _errs(b.field("water_depth"), o.field("water_depth"))
    f32 = Model(w.path, w.params_file, precision=hc.HC_FP32, elev_fixed=True)
    f32.iterate(10 ** 7)
    assert f32.finished
    err = _errs(f32.download("water_depth"), o.field("water_depth"))
    print(f"complex rain fp32 depth error {err}, floor {floor}")
    assert err["max"] <= max(1e-4, 4 * floor["max"]) and err["l2"] <= max(1e-4, 4 * floor["l2"])
    assert err["vol"] < 1e-6


def _reach_ledger(w, **kw):
    """Per-iteration volume ledger of a host-stepped reach-mode run: the
    change of the stored volume sum(h) DX^2 against what entered (the reach
    discharge of LCM.cpp:3420-3440 over the flow timestep, and the TOPMODEL
    rain at the input points, waterinput over the flow timestep) and what
    was drained (temptot over time_step).  Returns (sum of the residuals,
    final volume)."""
    m = Model(w.path, w.params_file, **kw)
    dx, div, rdt = w.dx, float(w.params["divide_inputs_by"]), float(w.params["reach_input_data_timestep"])
    rows = [np.loadtxt(os.path.join(w.path, w.params[f"reach{k}_input_file"])).astype(np.float32)
            for k in (1, 2)]
    v0 = m.download("water_depth")[1:-1, 1:-1].sum() * dx * dx
    total = 0.0
    while True:
        tt, cy, dt = m.iterate(1)
        if len(tt) == 0:
            break
        i = int(cy[0] / rdt)
        prop = (((i + 1) * rdt) - cy[0]) / rdt
        reach = sum(float(r[i, 1]) + (float(r[i + 1, 1]) - float(r[i, 1])) * (1 - prop) for r in rows)
        vin = (reach / div) * dt[0] + m.scalar("waterinput") * dt[0]
        v1 = m.download("water_depth")[1:-1, 1:-1].sum() * dx * dx
        total += (v1 - v0) - vin + tt[0] * m.scalar("time_step")
        v0 = v1
        if m.finished:
            break
    m.close()
    return total, v0


def test_driver_reach_mode_vs_oracle(tmp_path):
    """reach_mode (reach_water_and_sediment_input, LSDCatchmentModel.cpp:
    3301-3452) through the driver, hydro-only: two reach points' discharge
    added on the device (hc_reach_inputs); FP64 within 1e-10 of the oracle
    (pinned bitwise to the compiled reference) over the whole run, and the
    FP32 form (fixed-point elevations) within 1e-4 of it or four times the
    FP64 run's sensitivity to 1e-8 m of DEM noise.

    Volume, by a per-iteration ledger (inflow, drainage and the stored
    volume): the FP64 run closes it to rounding (measured 3e-16 of the final
    volume).  The FP32 run drifts by -9.5e-6 over the run, nearly all of it
    the rounding of the small TOPMODEL baseflow amounts added at the input
    points to FP32 depths: the amount is nearly the same each step and the
    depth at a point changes slowly, so the same fraction of an ulp is lost
    step after step (an FP64 carry on those adds, tried outside the product,
    closes the ledger to 2.6e-7; the reach point adds themselves contribute
    2e-7 m^3).  The bars: FP64 ledger 1e-12, FP32 ledger and final volume
    against FP64 1.2e-5."""
    w = inputs.reach_inputs(os.path.join(str(tmp_path), "rh"))
    o = OracleModel.from_workload(w)
    ot, oc, od = o.iterate(10 ** 7)
    m = Model(w.path, w.params_file)
    gt, gc, gd = m.iterate(10 ** 7)
    assert m.finished and len(gt) == len(ot) > 100
    _assert_rel(gc, oc, 1e-12, 0, "cycle")
    _assert_rel(gd, od, 1e-10, 0, "dt")
    _assert_rel(gt, ot, 1e-10, 1e-12, "temptot")
    for f in ("water_depth", "qx", "qy"):
        _assert_rel(m.download(f), o.field(f), 1e-10, 1e-13, f)
    _assert_rel(m.series(), o.series(), 1e-10, 1e-12, "series")
    b = OracleModel.from_workload(w, initialise=False)
    e = b.field("elev")
    land = e > -9999
    e[land] += 1e-8 * np.random.Generator(np.random.PCG64(3)).standard_normal(land.sum())
    b.L.orc_initialise(b.h)
    b.iterate(10 ** 7)
    floor = _errs(b.field("water_depth"), o.field("water_depth"))
    f32 = Model(w.path, w.params_file, precision=hc.HC_FP32, elev_fixed=True)
    f32.iterate(10 ** 7)
    err = _errs(f32.download("water_depth"), o.field("water_depth"))
    print(f"reach fp32 depth error {err}, floor {floor}")
    assert err["max"] <= max(1e-4, 4 * floor["max"]) and err["vol"] < 1.2e-5
    led64, vol64 = _reach_ledger(w)
    led32, vol32 = _reach_ledger(w, precision=hc.HC_FP32, elev_fixed=True)
    print(f"reach volume ledger: fp64 {led64 / vol64:.2e}, fp32 {led32 / vol32:.2e}")
    assert abs(led64) <= 1e-12 * vol64
    assert abs(led32) <= 1.2e-5 * vol32


def test_driver_reach_mode_erosion_runs(tmp_path):
    """reach_mode in an erosion run through the driver: the sediment of the
    reach files enters at the two points on the device (suspended fraction
    into Vsusptot, the others into the grain rows and the bed) and the run
    stays finite.  The amounts are the oracle's restatement, pinned per call
    against the compiled reference on the CPU."""
    w = inputs.reach_inputs(os.path.join(str(tmp_path), "re"), erosion=True, sed_scale=50.0)
    m = Model(w.path, w.params_file)
    e0 = m.download("elev")
    m.iterate(300)
    V, e, h = m.download("Vsusptot"), m.download("elev"), m.download("water_depth")
    assert np.isfinite(V).all() and np.isfinite(h).all() and np.isfinite(e).all()
    pts = [(2, 40 // 3), (2, (2 * 40) // 3)]
    assert all(V[p] > 0 or e[p] > e0[p] for p in pts)
