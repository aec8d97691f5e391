"""Geometry bitmaps of the oracle against a vectorised numpy construction
of the same sets and the closed-form bit index
(ScatterGen_RayTracing_CircularHole/scatter.cu:1040-1074, :1440-1540, :1939-2009)."""
import numpy as np


def unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def bars_raw(p):
    n = int(np.floor(p.det[0] + 1e-4))
    rec = p.det[1:1 + 12 * n].reshape(n, 12).astype(np.float64)
    c = rec[:, 0:3]
    r = np.sqrt(rec[:, 3] ** 2 + rec[:, 5] ** 2 + rec[:, 4] ** 2) / 2.0
    return c, r


def crystal_sets(p):
    """{(i, j, k)} flagged by the angular test from i or from j."""
    c, r = bars_raw(p)
    n = len(r)
    out = set()
    for i in range(n):
        for j in range(i + 1, n):
            ks = np.array([k for k in range(n) if k != i and k != j])
            lij = np.linalg.norm(c[i] - c[j])
            u_ki = unit(c[ks] - c[i])
            u_kj = unit(c[ks] - c[j])
            at_i = np.linalg.norm(u_ki - (c[j] - c[i]) / lij, axis=1) <= r[ks] / np.linalg.norm(c[ks] - c[i], axis=1) + r[j] / lij
            at_j = np.linalg.norm(u_kj - (c[i] - c[j]) / lij, axis=1) <= r[ks] / np.linalg.norm(c[ks] - c[j], axis=1) + r[i] / lij
            for k in ks[at_i | at_j]:
                out.add((i, j, int(k)))
    return out


def decode_crystal(words, n):
    bits = np.unpackbits(words.view(np.uint8), bitorder="little")
    out = set()
    for b in np.nonzero(bits)[0]:
        pair, k = divmod(int(b), n)
        i = 0
        while (i + 1) * n - (i + 1) * (i + 2) // 2 + (i + 1) <= pair:
            i += 1
        j = pair - (i * n - i * (i + 1) // 2)
        out.add((i, j, k))
    return out


def test_crystal_bitmap_matches_sets(oracle_mod, params_mod):
    p = params_mod.config(1)
    p.det[0] = 24  # first 24 bars: 3 rows of the 8 x 8 array
    o = oracle_mod.Oracle(p)
    words = o.bitmap_crystal()
    assert words.size == (24 * 25 // 2 * 24 + 31) // 32
    assert decode_crystal(words, 24) == crystal_sets(p)


def test_collimator_bitmap_matches_sets(oracle_mod, params_mod):
    p = params_mod.config(1)
    p.det[0] = 20
    p.coll[10] = 3
    o = oracle_mod.Oracle(p)
    words = o.bitmap_collimator()
    c, r = bars_raw(p)
    n = 20
    expect = np.zeros(3 * n * n, bool)
    for h in range(3):
        hp = p.coll[100 + 9 * h:109 + 9 * h].astype(np.float64)
        ch = np.array([hp[0], (hp[1] + hp[2]) / 2.0, hp[3]])
        for j in range(n):
            lj = np.linalg.norm(c[j] - ch)
            lk = np.linalg.norm(c - ch, axis=1)
            hit = np.linalg.norm(unit(c - ch) - (c[j] - ch) / lj, axis=1) <= r / lk + r[j] / lj
            expect[(h * n + j) * n:(h * n + j + 1) * n] = hit
    bits = np.unpackbits(words.view(np.uint8), bitorder="little")[: 3 * n * n].astype(bool)
    assert np.array_equal(bits, expect)
    # k == j is always flagged (distance 0 on the unit sphere)
    for h in range(3):
        for j in range(n):
            assert bits[(h * n + j) * n + j]


def test_angular_test_semantics(oracle_mod, params_mod):
    """The test is angular, seen from either end: a bar on the line through i
    and j is flagged (between them or beyond one end); a bar well off the
    line is not."""
    p = params_mod.Params.empty()
    p.det[0] = 4
    for k, (x, y) in enumerate(((-10.0, 20.0), (0.0, 20.0), (10.0, 20.0), (0.0, 40.0))):
        params_mod.set_bar(p, k, x, y, 0.0, 2.0, 2.0, 2.0, 0.25, 0.18, 0.07, 0.1)
    params_mod.set_image(p, 2, 2, 1, 1.0, 1.0, 1.0)
    params_mod.set_physics(p)
    o = oracle_mod.Oracle(p)
    assert o.crystal_pred(0, 2, 1) == 1   # between
    assert o.crystal_pred(0, 1, 2) == 1   # beyond j, same direction seen from i
    assert o.crystal_pred(0, 2, 3) == 0   # 20 mm off the line
    words = o.bitmap_crystal()
    assert oracle_mod.lib().smg_o_bit_crystal(words, 4, 2, 0, 1) == 1  # symmetric pair lookup
