"""GPU: hc_add_point_inputs_noalloc (include/hc.h), the point-input entry
point that allocates nothing (its staging is a static device array), called
through the C-ABI and compared bit for bit with hc_add_point_inputs on the
same state: two calls in a row on one context, the second with fewer points,
in FP64 and FP32, with distinct and with repeated cells; then the limit."""
import numpy as np
import pytest

from hail_caesar_amd import hc
from hcutil import to_context, wet_flow_state

pytestmark = pytest.mark.gpu


def _points(rng, n, ncell, repeats):
    if repeats:
        cells = rng.integers(1, ncell - 1, size=(max(1, n // 7), 2))
        pick = rng.integers(0, len(cells), size=n)
        return cells[pick, 0], cells[pick, 1], rng.uniform(1e-6, 1e-2, size=n)
    flat = rng.choice(ncell * ncell, size=n, replace=False)
    return flat // ncell, flat % ncell, rng.uniform(1e-6, 1e-2, size=n)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("repeats", [False, True])
def test_noalloc_matches_add_point_inputs(tmp_path, precision, repeats):
    w, o = wet_flow_state(tmp_path, size=128, steps=0)
    a = to_context(o, precision=precision)
    b = to_context(o, precision=precision)
    rng = np.random.Generator(np.random.PCG64(2024 + precision + 2 * repeats))
    for n in (hc.HC_NOALLOC_MAX_POINTS, 1500):  # the second call with fewer points
        xs, ys, amts = _points(rng, n, 130, repeats)
        a.add_point_inputs(xs, ys, amts)
        b.add_point_inputs_noalloc(xs, ys, amts)
        ha, hb = a.download("water_depth"), b.download("water_depth")
        assert np.array_equal(ha, hb), n
    a.close()
    b.close()


def test_noalloc_limit(tmp_path):
    w, o = wet_flow_state(tmp_path, size=64, steps=0)
    c = to_context(o)
    n = hc.HC_NOALLOC_MAX_POINTS + 1
    with pytest.raises(hc.HCError) as e:
        c.add_point_inputs_noalloc(np.ones(n, int), np.ones(n, int), np.zeros(n))
    assert e.value.code == hc.HC_ERR_ARG and "HC_NOALLOC_MAX_POINTS" in str(e.value)
    with pytest.raises(hc.HCError):
        c.add_point_inputs_noalloc([0, 70], [0, 1], [1.0, 1.0])  # outside the padded grid
    c.close()
