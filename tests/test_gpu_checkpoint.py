"""Exact binary checkpoint/restart on the device: hcm_checkpoint/hcm_restart
(include/hc_model.h) and the context state image hc_state_* (include/hc.h).

The reference has only a lossy hot start (water depth re-read from a raster
written at 6 significant figures, grain state from a Grainz file; qx, qy,
vel_dir and the clocks not restored: LSDCatchmentModel.cpp:331-373, 590-683),
so the bar here is identity with the device's own uninterrupted run: a run
stopped part-way, checkpointed, and continued by a new model created from the
same inputs matches it bit for bit -- per-iteration records, every field,
the interval series and the clocks.  The stop falls on an odd iteration, so
the restored run starts on the other generation of the ping-pong state.
"""
import os

import numpy as np
import pytest

import hail_caesar_amd.inputs as inputs
from hail_caesar_amd import hc
from hail_caesar_amd.model import Model

pytestmark = pytest.mark.gpu

HYDRO_FIELDS = ("water_depth", "qx", "qy", "elev")
SED_FIELDS = ("Vsusptot", "qxs", "qys", "bedrock", "Tau", "Vel", "ss", "grain", "strata", "hasgrain")
SCALARS = ("time_step", "cycle", "counter", "maxdepth", "temptot", "waterinput", "waterOut",
           "baseflow", "totalinputpoints", "erode_mult", "globalsediq")


def _erosion_workload(path, size=48, seed=1013):
    rng = np.random.Generator(np.random.PCG64(seed))
    dem = inputs.planar_dem(size, size, 10.0, 0.01, 20.0, 0.6, seed)
    props = inputs.dirichlet_props(size, size, rng)
    water = np.full((size, size), 0.3)
    return inputs.make_workload(path, dem, 10.0, inputs.triangular_storm(6, 40.0, 15), 15, 2,
                                overrides={"hydro_model_only": "no", "active_layer_thickness": "1"},
                                water_init=water, graindata=props, name="ckero")


def _case(tmp_path, case):
    if case == "erosion-partial":  # grain rows made during the run on both sides
        w = inputs.erosion_run(os.path.join(str(tmp_path), "erp"), grain_cover=0.5)
        return w, {}, HYDRO_FIELDS + SED_FIELDS, 160, 53
    if case in ("erosion", "erosion-fp32"):
        w = _erosion_workload(os.path.join(str(tmp_path), "ero"))
        kw = {"precision": hc.HC_FP32} if case == "erosion-fp32" else {}
        return w, kw, HYDRO_FIELDS + SED_FIELDS, 160, 53
    if case == "complex-rain":
        w = inputs.complex_rain(os.path.join(str(tmp_path), "cr"), 3)
        return w, {}, HYDRO_FIELDS, 700, 333
    w = inputs.flow_only_256(os.path.join(str(tmp_path), "flow"), size=48, hours=1)
    kw = {} if case == "fp64" else {"precision": hc.HC_FP32, "elev_fixed": case == "fp32-fixed"}
    return w, kw, HYDRO_FIELDS, 400, 137


@pytest.mark.parametrize("case", ["fp64", "fp32", "fp32-fixed", "erosion", "erosion-fp32",
                                  "erosion-partial", "complex-rain"])
def test_restart_continues_bit_identical(tmp_path, case):
    w, kw, fields, n, k = _case(tmp_path, case)
    whole = Model(w.path, w.params_file, **kw)
    ta, ca, da = whole.iterate(n)
    assert len(ta) == n and not whole.finished

    first = Model(w.path, w.params_file, **kw)
    t1, _, _ = first.iterate(k)
    assert np.array_equal(t1, ta[:k])
    ek = first.download("elev")
    ck = os.path.join(str(tmp_path), "run.ckpt")
    first.checkpoint(ck)
    first.close()

    cont = Model(w.path, w.params_file, **kw)
    cont.restart(ck)
    tc, cc, dc = cont.iterate(n - k)
    assert np.array_equal(tc, ta[k:]) and np.array_equal(cc, ca[k:]) and np.array_equal(dc, da[k:])
    for f in fields:
        assert np.array_equal(cont.download(f), whole.download(f)), f
    assert np.array_equal(cont.series(), whole.series())
    for s in SCALARS:
        assert cont.scalar(s) == whole.scalar(s), s
    if case.startswith("erosion"):  # grain-row numbers continue as they would have
        ia, na = whole.context().grain_index_get()
        ic, nc = cont.context().grain_index_get()
        assert np.array_equal(ic, ia) and nc == na
    if case.startswith("erosion"):  # the bed moved on both sides of the checkpoint
        e0 = Model(w.path, w.params_file, **kw).download("elev")
        assert not np.array_equal(e0, ek) and not np.array_equal(ek, whole.download("elev"))


def test_restart_refusals(tmp_path):
    """A checkpoint of another grid, a truncated or foreign file, and a
    model whose device-resident loop has started are refused with a message;
    a refused restart leaves the model as it was (it then runs exactly as a
    fresh model)."""
    w = inputs.flow_only_256(os.path.join(str(tmp_path), "a"), size=48, hours=1)
    w2 = inputs.flow_only_256(os.path.join(str(tmp_path), "b"), size=40, hours=1)
    a = Model(w.path, w.params_file)
    a.iterate(25)
    ck = os.path.join(str(tmp_path), "a.ckpt")
    a.checkpoint(ck)

    other = Model(w2.path, w2.params_file)
    with pytest.raises(hc.HCError, match="row count|grid"):
        other.restart(ck)
    data = open(ck, "rb").read()
    short = os.path.join(str(tmp_path), "short.ckpt")
    with open(short, "wb") as f:
        f.write(data[: len(data) // 2])
    fresh = Model(w.path, w.params_file)
    with pytest.raises(hc.HCError, match="truncated"):
        fresh.restart(short)
    junk = os.path.join(str(tmp_path), "junk.ckpt")
    with open(junk, "wb") as f:
        f.write(b"not a checkpoint at all" * 4)
    with pytest.raises(hc.HCError, match="not a checkpoint"):
        fresh.restart(junk)
    ref = Model(w.path, w.params_file)
    tr, _, _ = ref.iterate(30)
    tf, _, _ = fresh.iterate(30)
    assert np.array_equal(tf, tr)
    assert np.array_equal(fresh.download("water_depth"), ref.download("water_depth"))

    dev = Model(w.path, w.params_file, device_loop=True)
    dev.iterate(5)
    with pytest.raises(hc.HCError, match="device-resident"):
        dev.checkpoint(os.path.join(str(tmp_path), "dev.ckpt"))


def test_context_state_image_round_trip():
    """hc_state_save / hc_state_load on a bare context: an image taken
    before two flow steps restores the state exactly, the steps then repeat
    bit for bit, and an image is refused by a context of another precision
    or grid."""
    imax, jmax = 37, 70
    rng = np.random.Generator(np.random.PCG64(17))
    prm = hc.default_params()
    ctx = hc.Context(imax, jmax, 5.0, prm, hc.HC_FP64)
    z = 10.0 + rng.random((imax + 2, jmax + 2))
    h = 0.05 + 0.1 * rng.random((imax + 2, jmax + 2))
    ctx.upload("elev", z)
    ctx.upload("water_depth", h)
    ctx.scan_area()
    ctx.flow_depth(0.5)
    img = ctx.state_save()
    ctx.flow_depth(0.5)
    ctx.flow_depth(0.5)
    after = {f: ctx.download(f) for f in ("water_depth", "qx", "qy")}
    ctx.state_load(img)
    ctx.flow_depth(0.5)
    ctx.flow_depth(0.5)
    for f, v in after.items():
        assert np.array_equal(ctx.download(f), v), f
    c32 = hc.Context(imax, jmax, 5.0, prm, hc.HC_FP32)
    with pytest.raises(hc.HCError):
        c32.state_load(img)
    other = hc.Context(imax + 1, jmax, 5.0, prm, hc.HC_FP64)
    with pytest.raises(hc.HCError):
        other.state_load(img)


BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hail-caesar_amd",
                   "bin", "hail-caesar-mi355x")


def _cli(cwd, *args):
    import subprocess
    r = subprocess.run([BIN, "./", *args], cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_cli_time_slices_continue_bit_identical(tmp_path):
    """The command-line driver run as two time slices (--max-iterations with
    --checkpoint, then --restart) writes the same time series and rasters,
    byte for byte, as one uninterrupted run; rows a killed slice appended
    after its last checkpoint are dropped on restart."""
    import shutil
    src = inputs.flow_only_256(os.path.join(str(tmp_path), "src"), size=48, hours=1)
    whole = os.path.join(str(tmp_path), "whole")
    sliced = os.path.join(str(tmp_path), "sliced")
    shutil.copytree(src.path, whole)
    shutil.copytree(src.path, sliced)
    pf = src.params_file
    assert "FINISHED" in _cli(whole, pf)
    out = _cli(sliced, pf, "--checkpoint", "run.ckpt", "--checkpoint-every", "97",
               "--max-iterations", "250")
    assert "stopped after 250" in out
    series = os.path.join(sliced, "out", "flow256.dat")
    with open(series, "a") as f:  # what a slice killed after its checkpoint leaves
        f.write("9 1.000000 2.000000 3.000000 4.0000000000\n")
    assert "FINISHED" in _cli(sliced, pf, "--restart", "run.ckpt", "--checkpoint", "run.ckpt")
    outs = sorted(os.listdir(os.path.join(whole, "out")))
    assert outs == sorted(os.listdir(os.path.join(sliced, "out"))) and "flow256.dat" in outs
    for name in outs:
        with open(os.path.join(whole, "out", name), "rb") as fa, \
                open(os.path.join(sliced, "out", name), "rb") as fb:
            assert fa.read() == fb.read(), name


def test_checkpoint_write_is_atomic(tmp_path):
    """A checkpoint is written beside its target and renamed over it when
    complete: a write cut short (a truncated `.tmp` left behind) does not
    touch the previous checkpoint, which still restarts bit for bit, and the
    next checkpoint replaces both."""
    w, kw, fields, n, k = _case(tmp_path, "fp64")
    a = Model(w.path, w.params_file, **kw)
    a.iterate(k)
    ck = os.path.join(str(tmp_path), "run.ckpt")
    a.checkpoint(ck)
    good = open(ck, "rb").read()
    with open(ck + ".tmp", "wb") as f:  # what a slice killed mid-write leaves
        f.write(good[: len(good) // 3])
    assert open(ck, "rb").read() == good
    b = Model(w.path, w.params_file, **kw)
    b.restart(ck)
    tb, _, _ = b.iterate(20)
    ta, _, _ = a.iterate(20)
    assert np.array_equal(ta, tb)
    a.checkpoint(ck)  # replaces the stale .tmp and the old checkpoint
    assert not os.path.exists(ck + ".tmp")
    c = Model(w.path, w.params_file, **kw)
    c.restart(ck)
    assert c.scalar("counter") == a.scalar("counter")


BIN_ASAN = BIN + "-asan"


def _run_bin(binary, cwd, *args, env=None):
    import subprocess
    r = subprocess.run([binary, "./", *args], cwd=cwd, capture_output=True, text=True, timeout=300,
                       env=env)
    return r.returncode, r.stdout, r.stderr


@pytest.mark.parametrize("case", ["erosion-grainz-ckpt", "flow-graph"])
def test_driver_host_code_under_asan(tmp_path, case):
    """The command-line driver with its host code (parameter and raster
    parsing, the run loop, TOPMODEL, input points, the series, raster and
    Grainz writers, checkpoints and restarts) built with AddressSanitizer and
    UndefinedBehaviorSanitizer, on the device as usual: no report, and every
    output file byte-identical to the ordinary build's.  Cases: a whole
    erosion run with half the cells in its grain data file (rows made as it
    goes), Grainz output and a checkpoint every 500 iterations, cut at 900
    iterations and restarted; a flow-only run in device-resident graphs."""
    import shutil
    if case == "erosion-grainz-ckpt":
        src = inputs.erosion_run(os.path.join(str(tmp_path), "src"), grain_cover=0.5,
                                 extra_params={"write_grainsize_file": "yes",
                                               "grainsize_file": "grainz",
                                               "raster_output_interval": "30",
                                               "dem_write_extension": "asc"})
        runs = [("--checkpoint", "run.ckpt", "--checkpoint-every", "500", "--max-iterations", "900"),
                ("--restart", "run.ckpt", "--checkpoint", "run.ckpt")]
    else:
        src = inputs.flow_only_256(os.path.join(str(tmp_path), "src"), size=64, hours=1)
        runs = [("--graph",)]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:verify_asan_link_order=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    dirs = {}
    for kind, binary in (("plain", BIN), ("asan", BIN_ASAN)):
        d = os.path.join(str(tmp_path), kind)
        shutil.copytree(src.path, d)
        for args in runs:
            print(f"{case} {kind}: {' '.join(args)}", flush=True)
            rc, out, err = _run_bin(binary, d, src.params_file, *args, env=env)
            assert rc == 0, (kind, args, out[-2000:], err[-4000:])
            assert "AddressSanitizer" not in err and "runtime error" not in err, err[-4000:]
        dirs[kind] = d
    outs = sorted(os.listdir(os.path.join(dirs["plain"], "out")))
    assert outs == sorted(os.listdir(os.path.join(dirs["asan"], "out"))) and outs
    if case == "erosion-grainz-ckpt":
        assert {"grainz0.asc", "grainz30.asc", "grainz60.asc"} <= set(outs)
    for name in outs:
        with open(os.path.join(dirs["plain"], "out", name), "rb") as fa, \
                open(os.path.join(dirs["asan"], "out", name), "rb") as fb:
            assert fa.read() == fb.read(), name
