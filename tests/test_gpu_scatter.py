"""GPU parity of the geometry bitmaps and the primary-Compton scatter
matrices against the FP64 oracle.  Mirrors the three scatter() launch
wrappers (ScatterGen_RayTracing_CircularHole/scatter.cu:2013-2378,
ScatterGen_Collimator/scatter.cu:1511-1747, ScatterGen_Crystal/scatter.cu:1528-1738)."""
import numpy as np
import pytest

from conftest import TOL_FP32, TOL_FP64, assert_parity

pytestmark = pytest.mark.gpu


def test_bitmaps_cfg1_bit_exact(smg_mod, oracle_mod, params_mod):
    p = params_mod.config(1)
    o = oracle_mod.Oracle(p)
    with smg_mod.open_context(p) as ctx:
        bc = ctx.bitmap_crystal()
        bl = ctx.bitmap_collimator()
    assert np.array_equal(bc, o.bitmap_crystal())
    assert np.array_equal(bl, o.bitmap_collimator())


def test_bitmaps_ragged_bit_exact(smg_mod, oracle_mod, params_mod):
    """N = 37 bars (words straddle pair rows), 5 layer-0 holes."""
    p = params_mod.config(1)
    p.det[0] = 37
    p.coll[10] = 5
    o = oracle_mod.Oracle(p)
    with smg_mod.open_context(p) as ctx:
        assert np.array_equal(ctx.bitmap_crystal(), o.bitmap_crystal())
        assert np.array_equal(ctx.bitmap_collimator(), o.bitmap_collimator())


def test_bitmaps_cfg2_sampled_bits(smg_mod, oracle_mod, params_mod):
    """2048 bars, 400 holes: 537 MB and 210 MB bitmaps on the GPU, a seeded
    sample of bits against the oracle's predicate."""
    p = params_mod.config(2)
    o = oracle_mod.Oracle(p)
    N = o.n_det
    rng = np.random.default_rng(4)
    with smg_mod.open_context(p) as ctx:
        bc = ctx.bitmap_crystal()
        bl = ctx.bitmap_collimator()
    # random triples plus triples near each other (where bits are set)
    i = rng.integers(0, N, 4000)
    j = np.clip(i + rng.integers(1, 200, 4000), 0, N - 1)
    k = np.clip(i + rng.integers(-100, 300, 4000), 0, N - 1)
    keep = (i < j) & (k != i) & (k != j)
    n_set = 0
    for a, b, c in zip(i[keep], j[keep], k[keep]):
        pair = a * N - a * (a + 1) // 2 + b
        bit = pair * N + c
        got = (int(bc[bit // 32]) >> (bit % 32)) & 1
        assert got == o.crystal_pred(int(a), int(b), int(c)), (a, b, c)
        n_set += got
    assert n_set > 10
    h = rng.integers(0, o.n_holes_l0, 3000)
    j = rng.integers(0, N, 3000)
    k = np.clip(j + rng.integers(-80, 80, 3000), 0, N - 1)
    for a, b, c in zip(h, j, k):
        bit = (int(a) * N + int(b)) * N + int(c)
        got = (int(bl[bit // 32]) >> (bit % 32)) & 1
        assert got == o.coll_pred(int(a), int(b), int(c)), (a, b, c)


def _word_bits(words, bits):
    """Bits at the given indices of a device bitmap (torch int32 tensor)."""
    import torch
    idx = torch.as_tensor(np.asarray(bits, np.int64) // 32, device=words.device)
    w = words[idx].cpu().numpy().view(np.uint32).astype(np.int64)
    return (w >> (np.asarray(bits, np.int64) % 32)) & 1


def test_bitmaps_6000_bars(smg_mod, oracle_mod, params_mod):
    """6,000 bars (3 layers of 50x40) and 400 holes: the 13.5 GB crystal
    bitmap (3.4e9 words, word indices past 2^31) and the 1.8 GB collimator
    bitmap, built into device memory.  Whole rows of k for random pairs, the last pair row,
    and seeded triples near each other are checked against the oracle's
    predicate."""
    import torch
    p = params_mod.large_detector()
    o = oracle_mod.Oracle(p)
    N, H = o.n_det, o.n_holes_l0
    assert N == 6000 and H == 400
    rng = np.random.default_rng(8)
    with smg_mod.open_context(p) as ctx:
        n_c = ctx._L.smg_bitmap_crystal_words(N)
        n_l = ctx._L.smg_bitmap_collimator_words(H, N)
        assert n_c > 2 ** 32 // 2
        ms_c = ctx.build_bitmap(False)
        ms_l = ctx.build_bitmap(True)
        bc = torch.empty(n_c, dtype=torch.int32, device="cuda")
        ctx.bitmap_crystal(device_ptr=bc.data_ptr())
        bl = torch.empty(n_l, dtype=torch.int32, device="cuda")
        ctx.bitmap_collimator(device_ptr=bl.data_ptr())
        torch.cuda.synchronize()
    print(f"\n6000-bar bitmaps: crystal {n_c} words {ms_c:.0f} ms, collimator {n_l} words {ms_l:.0f} ms")

    def pair_base(a, b):
        return (a * N - a * (a + 1) // 2 + b) * N

    # whole rows: 12 random pairs, neighbours, and the last pair (N-2, N-1)
    pairs = [tuple(sorted(rng.choice(N, 2, replace=False))) for _ in range(12)]
    pairs += [(10, 11), (1000, 1050), (N - 2, N - 1)]
    ks = np.arange(N)
    n_set = 0
    for a, b in pairs:
        got = _word_bits(bc, pair_base(a, b) + ks)
        want = [0 if c in (a, b) else o.crystal_pred(int(a), int(b), int(c)) for c in ks]
        bad = np.nonzero(got != np.array(want))[0]
        assert bad.size == 0, (a, b, bad[:5])
        n_set += int(got.sum())
    # the bit after the last pair's row does not exist; the last word's tail is 0
    total = N * (N + 1) // 2 * N
    tail = _word_bits(bc, np.arange(total, n_c * 32))
    assert not tail.any()
    # triples near each other
    i = rng.integers(0, N - 1, 3000)
    j = np.clip(i + rng.integers(1, 300, 3000), 0, N - 1)
    k = np.clip(i + rng.integers(-150, 450, 3000), 0, N - 1)
    keep = (i < j) & (k != i) & (k != j)
    got = _word_bits(bc, pair_base(i[keep], j[keep]) + k[keep])
    want = np.array([o.crystal_pred(int(a), int(b), int(c)) for a, b, c in zip(i[keep], j[keep], k[keep])])
    assert np.array_equal(got, want)
    assert n_set + int(got.sum()) > 50
    # collimator: whole rows (h, j) and seeded triples
    for h, b in [(0, 0), (17, 2999), (H - 1, N - 1)] + [tuple(x) for x in zip(rng.integers(0, H, 4), rng.integers(0, N, 4))]:
        got = _word_bits(bl, (h * N + b) * N + ks)
        want = np.array([o.coll_pred(int(h), int(b), int(c)) for c in ks])
        assert np.array_equal(got, want), (h, b)
    h = rng.integers(0, H, 3000)
    j = rng.integers(0, N, 3000)
    k = np.clip(j + rng.integers(-100, 100, 3000), 0, N - 1)
    got = _word_bits(bl, (h * N + j) * N + k)
    want = np.array([o.coll_pred(int(a), int(b), int(c)) for a, b, c in zip(h, j, k)])
    assert np.array_equal(got, want)
    assert want.sum() > 50


def _sample(o, n, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, o.n_det, n), rng.integers(0, o.n_vox, n)


@pytest.mark.parametrize("variant", ["combined", "standalone"])
@pytest.mark.parametrize("prec,tol", [("fp32", TOL_FP32), ("fp64", TOL_FP64)])
def test_collimator_scatter_cfg3(smg_mod, oracle_mod, params_mod, variant, prec, tol):
    """Config 3 geometry (32x32x16 image, 64 bars, 16 layer-0 holes)."""
    p = params_mod.config(3)
    o = oracle_mod.Oracle(p)
    d, v = _sample(o, 3000, 5)
    with smg_mod.open_context(p) as ctx:
        got, st = ctx.scatter_collimator(0, variant, prec)
        bm = ctx.bitmap_collimator()
    ref = o.collimator_scatter_pairs(bm, d, v, variant=0 if variant == "combined" else 1)
    assert (ref > 0).mean() > 0.5
    assert_parity(got[d, v], ref, tol, f"collimator scatter {variant} {prec}")


@pytest.mark.parametrize("prec,tol", [("fp32", TOL_FP32), ("fp64", TOL_FP64)])
def test_crystal_scatter_cfg3(smg_mod, oracle_mod, params_mod, prec, tol):
    """Crystal scatter driven by the GPU PE block of the same geometry."""
    p = params_mod.config(3)
    o = oracle_mod.Oracle(p)
    d, v = _sample(o, 3000, 6)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(4, 4, 4)
        pe, _ = ctx.pe_forward(0, "fp32")
        got, st = ctx.scatter_crystal(pe, 0, prec)
        bm = ctx.bitmap_crystal()
    ref = o.crystal_scatter_pairs(pe, bm, d, v)
    assert (ref > 0).mean() > 0.05
    assert_parity(got[d, v], ref, tol, f"crystal scatter {prec}")


def test_crystal_scatter_subcells_and_loaded_bitmap(smg_mod, oracle_mod, params_mod):
    """2x2x2 detection sub-cells, a bitmap installed from a cache instead of
    computed, and a random PE block (the scatter is linear in it)."""
    p = params_mod.config(1)
    o = oracle_mod.Oracle(p)
    rng = np.random.default_rng(8)
    pe = rng.uniform(0, 1e-9, (o.n_det, o.n_vox)).astype(np.float32)
    bm = o.bitmap_crystal()
    d, v = _sample(o, 600, 9)
    with smg_mod.open_context(p) as ctx:
        ctx.set_scatter_subsamples(2, 2, 2)
        ctx.load_bitmap_crystal(bm)
        got, _ = ctx.scatter_crystal(pe, 0, "fp32")
    ref = o.crystal_scatter_pairs(pe, bm, d, v, sub=(2, 2, 2))
    assert_parity(got[d, v], ref, TOL_FP32, "crystal scatter 2x2x2")


def test_scatter_deterministic(smg_mod, params_mod):
    p = params_mod.config(1)
    with smg_mod.open_context(p) as ctx:
        a, _ = ctx.scatter_collimator(0, "combined", "fp32")
        b, _ = ctx.scatter_collimator(0, "combined", "fp32")
    assert np.array_equal(a, b)


def test_compton_off_gives_zero(smg_mod, params_mod):
    """Phys[0] != 1 disables the scatter kernels (scatter.cu:1226)."""
    p = params_mod.config(1)
    p.phys[0] = 0
    with smg_mod.open_context(p) as ctx:
        a, _ = ctx.scatter_collimator(0, "combined", "fp32")
    assert np.all(a == 0)


def mu_tables(power_law: bool, e0: float = 140.5):
    """64 log-spaced nodes on 20-200 keV (crystal pe, crystal Compton,
    collimator pe, collimator Compton): smg.params.synthetic_mu_tables."""
    from smg import params
    return params.synthetic_mu_tables(power_law, e0)


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-6), ("fp64", 1e-12)])
def test_power_law_tables_equal_power_laws(smg_mod, params_mod, prec, tol):
    """Per-process power-law tables give the reference's scaling: the same
    matrices as without tables, up to the rounding of exp(log)."""
    p = params_mod.config(3)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(2, 2, 2)
        pe, _ = ctx.pe_forward(0, "fp32")
        a_coll, _ = ctx.scatter_collimator(0, "combined", prec)
        a_cry, _ = ctx.scatter_crystal(pe, 0, prec)
        ctx.set_mu_tables(*mu_tables(True))
        b_coll, _ = ctx.scatter_collimator(0, "combined", prec)
        b_cry, _ = ctx.scatter_crystal(pe, 0, prec)
        ctx.set_mu_tables(None)
        c_coll, _ = ctx.scatter_collimator(0, "combined", prec)
    assert np.array_equal(a_coll, c_coll)
    for a, b in ((a_coll, b_coll), (a_cry, b_cry)):
        m = a > 0
        assert m.mean() > 0.05
        assert np.array_equal(a == 0, b == 0)
        assert np.max(np.abs(b[m] / a[m] - 1)) <= tol


@pytest.mark.parametrize("prec,tol", [("fp32", TOL_FP32), ("fp64", TOL_FP64)])
def test_mu_tables_vs_oracle(smg_mod, oracle_mod, params_mod, prec, tol):
    """Smooth non-power-law tables: collimator and crystal scatter against
    the oracle with the same tables."""
    p = params_mod.config(3)
    o = oracle_mod.Oracle(p)
    tabs = mu_tables(False)
    o.set_mu_tables(*tabs)
    d, v = _sample(o, 1500, 12)
    with smg_mod.open_context(p) as ctx:
        ctx.set_mu_tables(*tabs)
        ctx.set_pe_subsamples(2, 2, 2)
        pe, _ = ctx.pe_forward(0, "fp32")
        got_coll, _ = ctx.scatter_collimator(0, "combined", prec)
        got_cry, _ = ctx.scatter_crystal(pe, 0, prec)
        bl = ctx.bitmap_collimator()
        bc = ctx.bitmap_crystal()
    assert_parity(got_coll[d, v], o.collimator_scatter_pairs(bl, d, v, variant=0), tol,
                  f"collimator scatter, mu(E) tables, {prec}")
    assert_parity(got_cry[d, v], o.crystal_scatter_pairs(pe, bc, d, v), tol,
                  f"crystal scatter, mu(E) tables, {prec}")


@pytest.mark.parametrize("prec,tol", [("fp32", TOL_FP32), ("fp64", TOL_FP64)])
def test_collimator_grid_scatter(smg_mod, oracle_mod, params_mod, prec, tol):
    """Grid scatterers over the layer-0 tungsten (2 mm pitch on the config-3
    geometry), first-leg and second-leg collimator attenuation, with and
    without mu(E) tables, against the oracle's restatement of the same model."""
    p = params_mod.config(3)
    o = oracle_mod.Oracle(p)
    pts, dv = o.coll_grid_points(2.0)
    d, v = _sample(o, 200, 13)
    with smg_mod.open_context(p) as ctx:
        got, st, n = ctx.scatter_collimator_grid(2.0, 0, prec)
        ctx.set_mu_tables(*mu_tables(False))
        got_t, _, _ = ctx.scatter_collimator_grid(2.0, 0, prec)
    assert n == len(pts)
    ref = o.coll_grid_scatter_pairs(2.0, d, v)
    assert (ref > 0).mean() > 0.5
    assert_parity(got[d, v], ref, tol, f"collimator grid scatter {prec}")
    o.set_mu_tables(*mu_tables(False))
    ref_t = o.coll_grid_scatter_pairs(2.0, d, v)
    assert_parity(got_t[d, v], ref_t, tol, f"collimator grid scatter, mu(E) tables, {prec}")
    assert not np.allclose(ref, ref_t, rtol=1e-3, atol=0.0)


def test_collimator_grid_chunks_and_determinism(smg_mod, params_mod):
    """Fine pitch: the scatter points run in several chunks whose sums are
    added in a fixed order (bitwise reproducible); the matrix is finite and
    non-negative."""
    p = params_mod.config(1)
    with smg_mod.open_context(p) as ctx:
        a, st, n = ctx.scatter_collimator_grid(0.5, 0, "fp32")
        b, _, _ = ctx.scatter_collimator_grid(0.5, 0, "fp32")
    assert n > 100000
    assert np.array_equal(a, b)
    assert np.isfinite(a).all() and (a >= 0).all() and (a > 0).mean() > 0.5


@pytest.mark.parametrize("prec,tol", [("fp32", TOL_FP32), ("fp64", TOL_FP64)])
def test_crystal_grid_scatter(smg_mod, oracle_mod, params_mod, prec, tol):
    """Grid scatterers at the centres of 2x4x2 cells of every bar (config 1):
    first-leg weights from the PE kernel's point mode, second leg through the
    crystal bitmap and out of the scatter bar, for a subset of detector rows,
    with and without mu(E) tables, against the oracle."""
    p = params_mod.config(1)
    o = oracle_mod.Oracle(p)
    grid = (2, 4, 2)
    rows = np.array([0, 9, 27, 40, 63], np.int32)
    rng = np.random.default_rng(21)
    ri = rng.integers(0, len(rows), 300)
    v = rng.integers(0, o.n_vox, 300)
    with smg_mod.open_context(p) as ctx:
        got, st = ctx.scatter_crystal_grid(grid, rows, 0, prec)
        bm = ctx.bitmap_crystal()
        ctx.set_mu_tables(*mu_tables(False))
        got_t, _ = ctx.scatter_crystal_grid(grid, rows, 0, prec)
    assert got.shape == (len(rows), o.n_vox)
    ref = o.crystal_grid_scatter_pairs(bm, grid, rows[ri], v)
    assert (ref > 0).mean() > 0.5
    assert_parity(got[ri, v], ref, tol, f"crystal grid scatter {prec}")
    o.set_mu_tables(*mu_tables(False))
    ref_t = o.crystal_grid_scatter_pairs(bm, grid, rows[ri], v)
    assert_parity(got_t[ri, v], ref_t, tol, f"crystal grid scatter, mu(E) tables, {prec}")
    assert not np.allclose(ref, ref_t, rtol=1e-3, atol=0.0)


def test_crystal_grid_rows_match_full_matrix(smg_mod, params_mod):
    """A row subset gives the same rows as the full matrix, and a repeat run
    gives the same bits; the matrix is finite and non-negative."""
    p = params_mod.config(1)
    rows = np.array([63, 5, 5, 30], np.int32)
    with smg_mod.open_context(p) as ctx:
        full, _ = ctx.scatter_crystal_grid((2, 2, 2), None, 0, "fp32")
        again, _ = ctx.scatter_crystal_grid((2, 2, 2), None, 0, "fp32")
        part, _ = ctx.scatter_crystal_grid((2, 2, 2), rows, 0, "fp32")
    assert np.array_equal(full, again)
    assert np.isfinite(full).all() and (full >= 0).all() and (full > 0).mean() > 0.5
    np.testing.assert_allclose(part, full[rows], rtol=2e-6, atol=0.0)


@pytest.mark.parametrize("prec,tol", [("fp32", TOL_FP32), ("fp64", TOL_FP64)])
def test_collimator_grid_scatter_half_mm_pitch(smg_mod, oracle_mod, params_mod, prec, tol):
    """The benchmark's collimator grid: 0.5 mm pitch over the 60x5x60 mm
    tungsten (143,440 points), config-3 image (32x32x16) and 64 bars, with
    the benchmark's 64-node mu(E) tables, sampled elements against the
    oracle."""
    p = params_mod.config(3)
    o = oracle_mod.Oracle(p)
    d, v = _sample(o, 48, 17)
    tabs = mu_tables(True)
    with smg_mod.open_context(p) as ctx:
        ctx.set_mu_tables(*tabs)
        got, st, n = ctx.scatter_collimator_grid(0.5, 0, prec)
    assert n > 140000
    o.set_mu_tables(*tabs)
    ref = o.coll_grid_scatter_pairs(0.5, d, v)
    assert (ref > 0).all()
    assert_parity(got[d, v], ref, tol,
                  f"collimator grid 0.5 mm, mu(E) tables, {prec} (fallbacks {st['fp64_fallbacks']})")


@pytest.mark.parametrize("prec,tol", [("fp32", TOL_FP32), ("fp64", TOL_FP64)])
def test_crystal_grid_scatter_bench_size(smg_mod, params_mod, prec, tol):
    """The benchmark's crystal grid workload itself: scatter points at the
    centres of 10x32x10 cells (0.25 mm) of all 2048 bars (6,553,600 points),
    the 8 detector rows it times, the 64-node mu(E) tables, against committed
    FP64-oracle values of 12 elements (tests/golden/make_scatter_golden.py:
    the oracle needs ~1 min per element here)."""
    import os
    from conftest import ROOT
    fx = np.load(os.path.join(ROOT, "tests", "golden", "cfg4_crystal_grid_0p25mm.npz"))
    p = params_mod.config(4)
    rows = fx["rows"]
    with smg_mod.open_context(p) as ctx:
        ctx.set_mu_tables(*mu_tables(True))
        got, st = ctx.scatter_crystal_grid(tuple(int(x) for x in fx["grid"]), rows, 0, prec)
    ri = np.searchsorted(rows, fx["d"])
    assert np.array_equal(rows[ri], fx["d"])
    assert (fx["value"] > 0).all()
    assert_parity(got[ri, fx["v"]], fx["value"], tol,
                  f"crystal grid 0.25 mm, 2048 bars, {prec} (fallbacks {st['fp64_fallbacks']})")


@pytest.mark.parametrize("prec,tol", [("fp32", TOL_FP32), ("fp64", TOL_FP64)])
def test_crystal_scatter_2048_bars(smg_mod, oracle_mod, params_mod, prec, tol):
    """The benchmark's bar-centre crystal scatter (ScatterGen_Crystal path):
    32x32x1 voxels x 2048 bars, 2047 scatter points per element, PE block
    from the GPU, power laws and 64-node mu(E) tables, sampled elements
    against the oracle."""
    p = params_mod.config(4)
    o = oracle_mod.Oracle(p)
    d, v = _sample(o, 300, 61)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(*p.meta["pe_subsamples"])
        pe, _ = ctx.pe_forward(0, "fp32")
        got, st = ctx.scatter_crystal(pe, 0, prec)
        ctx.set_mu_tables(*mu_tables(True))
        got_t, _ = ctx.scatter_crystal(pe, 0, prec)
        bm = ctx.bitmap_crystal()
    ref = o.crystal_scatter_pairs(pe, bm, d, v)
    assert (ref > 0).mean() > 0.5
    assert_parity(got[d, v], ref, tol, f"crystal scatter, 2048 bars, {prec} "
                                        f"(fallbacks {st['fp64_fallbacks']})")
    o.set_mu_tables(*mu_tables(True))
    ref_t = o.crystal_scatter_pairs(pe, bm, d, v)
    assert_parity(got_t[d, v], ref_t, tol, f"crystal scatter, 2048 bars, mu(E) tables, {prec}")


@pytest.mark.parametrize("prec,tol", [("fp32", TOL_FP32), ("fp64", TOL_FP64)])
def test_crystal_grid_scatter_2048_bars(smg_mod, oracle_mod, params_mod, prec, tol):
    """Crystal grid scatter on the benchmark's detector (2 layers of 32x32
    bars, 32x32x1 voxels) with 2x8x2 cells per bar (65,536 points), four
    detector rows, sampled elements against the oracle (the GPU crystal
    bitmap, bit-exact elsewhere, gates both)."""
    p = params_mod.config(4)
    o = oracle_mod.Oracle(p)
    rows = np.array([0, 700, 1400, 2047], np.int32)
    rng = np.random.default_rng(23)
    ri = rng.integers(0, len(rows), 48)
    v = rng.integers(0, o.n_vox, 48)
    with smg_mod.open_context(p) as ctx:
        got, st = ctx.scatter_crystal_grid((2, 8, 2), rows, 0, prec)
        bm = ctx.bitmap_crystal()
    ref = o.crystal_grid_scatter_pairs(bm, (2, 8, 2), rows[ri], v)
    assert (ref > 0).all()
    assert_parity(got[ri, v], ref, tol,
                  f"crystal grid 2x8x2, 2048 bars, {prec} (fallbacks {st['fp64_fallbacks']})")


@pytest.mark.parametrize("prec,tol", [("fp32", TOL_FP32), ("fp64", TOL_FP64)])
def test_collimator_grid_scatter_three_layers(smg_mod, oracle_mod, params_mod, prec, tol):
    """Grid scatterers in the tungsten of three plates (each with its own
    holes and its own Compton mu), rotated bars, ragged tiles: against the
    oracle, and the three-layer matrix equals the sum of the one-layer ones."""
    from conftest import small_geometry
    p = small_geometry(params_mod, beta=0.05, layers=3)
    for m in range(3):
        p.coll[(m + 1) * 10 + 7] = 0.35 + 0.1 * m
    o = oracle_mod.Oracle(p)
    d, v = _sample(o, 120, 27)
    with smg_mod.open_context(p) as ctx:
        got, st, n = ctx.scatter_collimator_grid(1.0, 0, prec, layers=0b111)
        parts = [ctx.scatter_collimator_grid(1.0, 0, prec, layers=1 << m) for m in range(3)]
        with pytest.raises(smg_mod.SmgError):
            ctx.scatter_collimator_grid(1.0, 0, prec, layers=0b1000)
    assert n == sum(q[2] for q in parts) == len(o.coll_grid_points(1.0, layers=0b111)[0])
    ref = o.coll_grid_scatter_pairs(1.0, d, v, layers=0b111)
    assert (ref > 0).mean() > 0.5
    assert_parity(got[d, v], ref, tol, f"three-layer collimator grid scatter {prec}")
    total = sum(q[0].astype(np.float64) for q in parts)
    np.testing.assert_allclose(got, total, rtol=1e-5, atol=1e-7 * np.abs(total).max())


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_scatter_sum_on_gpu(smg_mod, params_mod, prec):
    """smg_scatter_sum forms the combined generator's crystal + collimator
    scatter and the drivers' PE + scatter block on the GPU: the same floats
    as the separate calls added on the host (ScatterGen_RayTracing_CircularHole/
    scatter.cu:2354-2357, ScatterGen_CircularHole.cpp:211-230)."""
    p = params_mod.config(3)
    with smg_mod.open_context(p) as ctx:
        ctx.set_pe_subsamples(2, 2, 2)
        pe, _ = ctx.pe_forward(0, "fp32")
        pe_add = (pe * np.float32(1.25)).astype(np.float32)
        cry, _ = ctx.scatter_crystal(pe, 0, prec)
        comb, _ = ctx.scatter_collimator(0, "combined", prec)
        stand, _ = ctx.scatter_collimator(0, "standalone", prec)
        sc, sm, st = ctx.scatter_sum(["crystal", "combined"], pe, pe_add, 0, prec)
        sc2, sm2, _ = ctx.scatter_sum(["standalone"], None, pe_add, 0, prec)
        _, sm3, _ = ctx.scatter_sum(["crystal"], pe, pe, 0, prec, scatter=False)
        with pytest.raises(smg_mod.SmgError):
            ctx.scatter_sum(["combined", "standalone"], None, pe, 0, prec)
    assert np.array_equal(sc, cry + comb)
    assert np.array_equal(sm, pe_add + (cry + comb))
    assert np.array_equal(sc2, stand) and np.array_equal(sm2, pe_add + stand)
    assert np.array_equal(sm3, pe + cry)
    assert st["rays"] > 0 and (sc > 0).mean() > 0.5
