"""CPU: the synthetic input generators behave as their docstrings state."""
import numpy as np

import hail_caesar_amd.inputs as inputs


def test_d8_network_drains_to_the_south_edge():
    """Every carved channel cell has a strictly lower channel cell among its
    8 neighbours, except on the south edge: water on the network reaches
    the outlet edge without meeting a pit."""
    z0 = inputs.planar_dem(256, 256, 2.0, 0.002, 30.0, 0.6, 7, channel_depth=0.0)
    z, ch = inputs.carve_d8_network(z0, block=16, depth=3.0)
    assert ch.sum() > 256  # a network, not a single channel
    n = z.shape[0]
    for r, c in zip(*np.nonzero(ch)):
        if r == n - 1:
            continue
        win = (slice(max(r - 1, 0), r + 2), slice(max(c - 1, 0), c + 2))
        assert ((z[win] < z[r, c]) & ch[win] & (z[win] > -9999)).any(), (r, c)
    # carving lowers terrain, never raises it, and leaves NODATA alone
    nod = z0 == -9999
    assert np.array_equal(z[nod], z0[nod])
    assert (z[~nod] <= z0[~nod] + 1e-12).sum() > 0.99 * (~nod).sum()


def test_ensemble_member_sizes_and_seeds():
    specs = inputs.ensemble_members(8, hours=1)
    sides = [a["dem"].shape[0] for a, _ in specs]
    assert all(128 <= s <= 1024 for s in sides)
    again = inputs.ensemble_members(8, hours=1)
    for (a, p), (b, q) in zip(specs, again):
        assert np.array_equal(a["dem"], b["dem"]) and p == q
    ns = [float(p["mannings_n"]) for _, p in specs]
    assert all(0.02 <= v <= 0.08 for v in ns)
