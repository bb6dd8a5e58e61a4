"""Host-side logic of pg_gpu_amd that needs no GPU: band tiling, fixed-point
decoding, reciprocal tables, generated statistic plans, and that libpgld.so
loads and exports every entry point include/pgld.h declares."""
import ctypes
import importlib.util
import math
import re
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from oracle import ld_oracle as o
from pg_gpu_amd import engine
from pg_gpu_amd._lib import LIB_PATH

ROOT = Path(__file__).resolve().parent.parent


def load_gen():
    spec = importlib.util.spec_from_file_location("gen_stats", ROOT / "pg_gpu_amd/csrc/gen_stats.py")
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


@pytest.mark.parametrize("seed,V,md", [(0, 1, 10.0), (1, 2, 5.0), (2, 300, 40.0), (3, 1000, 3.5),
                                       (4, 777, 1e9), (5, 500, 0.0)])
def test_band_tiles_cover_every_pair_once(seed, V, md):
    rng = np.random.default_rng(seed)
    pos = np.sort(rng.integers(0, 3 * V + 1, V)).astype(np.float64)
    tiles = engine.band_tiles(pos, md)
    assert len({tuple(t) for t in tiles}) == len(tiles)
    ii, jj = o.pairs_within_distance(pos, md)
    have = {tuple(t) for t in tiles}
    for i, j in zip(ii // engine.TILE, jj // engine.TILE):
        assert (i, j) in have
    assert np.all(tiles[:, 1] >= tiles[:, 0]) if len(tiles) else True


def to_limbs(v):
    """Python mirror of pgld_to_limbs (pgld_common.h)."""
    t = abs(v) * 256.0
    h = math.floor(t)
    r = (t - h) * 4294967296.0
    m = math.floor(r)
    s = (r - m) * 4294967296.0
    sg = -1 if v < 0 else 1
    return sg * int(h), sg * int(m), sg * int(np.rint(s))


def test_fixed_point_limbs_sum_exactly():
    rng = np.random.default_rng(3)
    vals = np.concatenate([rng.normal(0, 1e-3, 2000), rng.normal(0, 50, 50), [0.0, -0.0, 1e-30]])
    L = [0, 0, 0]
    for v in vals:
        x2, x1, x0 = to_limbs(float(v))
        assert abs(x1) < 2**32 and abs(x0) <= 2**32
        assert abs(Fraction(x2 * 2**64 + x1 * 2**32 + x0, 2**72) - Fraction(float(v))) <= Fraction(1, 2**73)
        L[0] += x0
        L[1] += x1
        L[2] += x2
    raw = np.array([[[L[0], L[1], L[2]], [len(vals), 0, 0]]], np.int64)
    sums, counts = engine.decode_limbs(raw, 1)
    exact = sum(Fraction(float(v)) for v in vals)
    assert abs(Fraction(sums[0, 0]) - exact) <= len(vals) * Fraction(1, 2**73) + abs(exact) * 2**-52
    assert counts[0] == len(vals)


def test_reciprocal_table():
    tab = engine.reciprocal_table(10).reshape(-1, 4)
    assert np.all(tab[0] == 0) and tab[1, 0] == 1.0 and np.all(tab[1, 1:] == 0)
    assert tab[4, 3] == 1.0 / 24 and tab[3, 3] == 0 and tab[3, 2] == 1.0 / 6
    assert tab[10, 3] == 1.0 / 5040


@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_generated_stat_plan_matches_specs(P):
    """The merged raw-call plan the device code is generated from evaluates to
    the reference's weighted spec sums."""
    gen = load_gen()
    values, stats = gen.stat_set_plan(P, gen.hap_key)
    rng = np.random.default_rng(P)
    counts = [rng.integers(0, 25, (300, 4)) for _ in range(P)]
    pops = [o.HapPop(c) for c in counts]
    fn = {
        "hap_dd_same": lambda a: o.hap_dd([a], 0, 0),
        "hap_dd_pair": lambda a, b: o.hap_dd([a, b], 0, 1),
        "hap_dz_same": lambda a: o.hap_dz([a], 0, 0, 0),
        "hap_dz_iik": lambda a, b: o.hap_dz([a, b], 0, 0, 1),
        "hap_dz_iji": lambda a, b: o.hap_dz([a, b], 0, 1, 0),
        "hap_dz_ijj": lambda a, b: o.hap_dz([a, b], 0, 1, 1),
        "hap_dz_ijk": lambda a, b, c: o.hap_dz([a, b, c], 0, 1, 2),
        "hap_pi2_same": lambda a: o.hap_pi2([a], 0, 0, 0, 0),
        "hap_pi2_triple": lambda t, s: o.hap_pi2([t, s], 0, 0, 0, 1),
        "hap_pi2_iikk": lambda a, b: o.hap_pi2([a, b], 0, 0, 1, 1),
        "hap_pi2_iikl": lambda a, b, c: o.hap_pi2([a, b, c], 0, 0, 1, 2),
        "hap_pi2_ijkk": lambda a, b, c: o.hap_pi2([a, b, c], 0, 1, 2, 2),
        "hap_pi2_ijij": lambda a, b: o.hap_pi2([a, b], 0, 1, 0, 1),
        "hap_pi2_shared": lambda s, a, b: o.hap_pi2([s, a, b], 0, 1, 0, 2),
        "hap_pi2_alldiff": lambda a, b, c, d: o.hap_pi2([a, b, c, d], 0, 1, 2, 3),
    }
    vals = [fn[f](*[pops[x] for x in args]) for f, args in values]
    got = np.stack([sum(w * vals[v] for v, w in terms) for terms in stats], axis=1)
    ref = o.multi_pop_stats_hap(counts, [None] * P)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)) < 1e-13


@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_generated_geno_stat_plan_matches_specs(P):
    """Same for the genotype plan: merging raw calls by geno_key (one device
    function per case of genotype_kernels.py) loses nothing."""
    from oracle import geno_oracle as go

    gen = load_gen()
    values, stats = gen.stat_set_plan(P, gen.geno_key)
    rng = np.random.default_rng(10 + P)
    counts = [rng.integers(0, 12, (200, 9)) for _ in range(P)]
    pops = [go.GenoPop(c) for c in counts]
    rep = {
        "geno_dd_same": ("dd", (0, 0)), "geno_dd_pair": ("dd", (0, 1)),
        "geno_dz_same": ("dz", (0, 0, 0)), "geno_dz_iik": ("dz", (0, 0, 1)),
        "geno_dz_iji": ("dz", (0, 1, 0)), "geno_dz_ijj": ("dz", (0, 1, 1)),
        "geno_dz_ijk": ("dz", (0, 1, 2)), "geno_pi2_same": ("pi2", (0, 0, 0, 0)),
        "geno_pi2_iiij": ("pi2", (0, 0, 0, 1)), "geno_pi2_ijii": ("pi2", (0, 1, 0, 0)),
        "geno_pi2_iikk": ("pi2", (0, 0, 1, 1)), "geno_pi2_iikl": ("pi2", (0, 0, 1, 2)),
        "geno_pi2_ijkk": ("pi2", (0, 1, 2, 2)), "geno_pi2_ijij": ("pi2", (0, 1, 0, 1)),
        "geno_pi2_shared": ("pi2", (0, 1, 0, 2)), "geno_pi2_alldiff": ("pi2", (0, 1, 2, 3)),
    }
    vals = []
    for f, args in values:
        kind, pp = rep[f]
        vals.append(go.kernel_call([pops[x] for x in args], kind, pp))
    got = np.stack([sum(w * vals[v] for v, w in terms) for terms in stats], axis=1)
    ref = go.multi_pop_stats_geno(counts, [None] * P)
    assert got.shape == ref.shape
    scale = np.maximum(np.abs(ref), 1e-12)
    assert np.max(np.abs(got - ref) / scale) < 1e-12


def header_functions():
    text = (ROOT / "include/pgld.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pgld_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    if not LIB_PATH.exists():
        pytest.fail("libpgld.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(LIB_PATH))
    names = header_functions()
    assert "pgld_binned_fused" in names and "pgld_pack" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    from pg_gpu_amd import _lib
    assert set(_lib.exported_symbols()) == set(names)


@pytest.mark.parametrize("square", [1, 3, 16, 64])
def test_band_tile_squares_are_permutations(square):
    """The square ordering changes only the order of the band tiles."""
    from pg_gpu_amd.engine import band_tiles

    rng = np.random.default_rng(5)
    pos = np.sort(rng.choice(np.arange(1, 3_000_000), 5_000, replace=False)).astype(float)
    ref = band_tiles(pos, 150_000)
    got = band_tiles(pos, 150_000, square)
    assert got.shape == ref.shape
    assert set(map(tuple, got)) == set(map(tuple, ref))
    if square == 1:   # squares of one tile: row-major
        np.testing.assert_array_equal(got, ref[np.lexsort((ref[:, 1], ref[:, 0]))])


def _band_tiles_spec(pos, md, sq):
    """The band tile list by definition: every 32-site tile pair (r, c >= r)
    holding a pair within the cutoff, ordered by (square row, square column,
    row, column)."""
    V = len(pos)
    T = engine.TILE
    have = set()
    ii, jj = o.pairs_within_distance(pos, md)
    for i, j in zip(ii // T, jj // T):
        have.add((int(i), int(j)))
    return sorted(have, key=lambda t: (t[0] // sq, t[1] // sq, t[0], t[1])), len(ii)


@pytest.mark.parametrize("seed,V,md,sq", [(0, 0, 5.0, 16), (1, 1, 5.0, 16), (2, 2, 5.0, 1),
                                          (3, 700, 40.0, 3), (4, 999, 1e9, 16),
                                          (5, 1500, 0.0, 8), (6, 2100, 200.0, 64)])
def test_band_tiles_order_and_pair_count(seed, V, md, sq):
    """pgld_band_tiles (host planner) against the definition: the set of
    tiles, their square order and the number of pairs within the cutoff."""
    rng = np.random.default_rng(seed)
    pos = np.sort(rng.integers(0, 3 * V + 1, V)).astype(np.float64)
    tiles, n_pairs = engine.band_tiles(pos, md, sq, with_pairs=True)
    want, n_want = _band_tiles_spec(pos, md, sq)
    assert [tuple(t) for t in tiles.tolist()] == want
    assert n_pairs == n_want


def test_band_tiles_rejects_bad_input():
    with pytest.raises(ValueError):
        engine.band_tiles(np.array([3.0, 1.0, 2.0]), 10.0)
    with pytest.raises(ValueError):
        engine.band_tiles(np.arange(100.0), -1.0)
    with pytest.raises(ValueError):
        engine.band_tiles(np.arange(100.0), 10.0, square=0)


def test_smoke_case_runs_on_cpu():
    """The oracle side of __graft_entry__.smoke() (workload and expected sums)."""
    import importlib.util

    spec = importlib.util.spec_from_file_location(
        "graft_entry", Path(__file__).resolve().parents[1] / "__graft_entry__.py")
    ge = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ge)
    w, ref, cnt, n_enum, absr = ge.smoke_case()
    assert ref.shape == (len(w.bins) - 1, 15) and cnt.sum() <= n_enum and n_enum > 0
    assert np.all(absr >= np.abs(ref) * (1 - 1e-12))


def _header_constants():
    """{raw call: k} for every `hap_X_num = hap_X_raw(...) * k` of the device
    header (the constants the fused kernel moves to the limb conversion)."""
    src = (ROOT / "pg_gpu_amd/csrc/hap_estimators.h").read_text()
    pat = re.compile(r"double (hap_\w+)_num\([^)]*\)\s*\{\s*return \1_raw\([^;]*\) \* ([0-9.]+);")
    return {name: float(k) for name, k in pat.findall(src)}


@pytest.mark.parametrize("P", [1, 2, 3, 4])
def test_generated_accumulation_reproduces_reference_stats(P):
    """Evaluates the generated hap_acc_P*_g* bodies and hap_scale_group tables
    of stats_hap.inc on oracle numerators: raw numerators from the header's
    constants, reciprocal products, the FMA chains and the 2^-e limb scale
    together must give the reference's statistics (a dropped 1/2 or a wrong
    exponent fails by a factor of two)."""
    gen = load_gen()
    inner = _header_constants()
    assert set(gen.HAP_INNER) | {"hap_dz_iik_iji_avg"} == set(inner)
    assert all(inner[k] == v for k, v in gen.HAP_INNER.items())
    assert inner["hap_dz_iik_iji_avg"] == gen.HAP_AVG_INNER
    inc = (ROOT / "pg_gpu_amd/csrc/stats_hap.inc").read_text()
    rng = np.random.default_rng(40 + P)
    counts = [rng.integers(1, 25, (200, 4)) for _ in range(P)]
    pops = [o.HapPop(c) for c in counts]
    ns = [c.sum(axis=1).astype(np.float64) for c in counts]

    def recip(p, k):
        return 1.0 / np.prod([ns[p] - j for j in range(k)], axis=0)

    def ref_call(fn, args):  # the reference value num * R of one raw call
        sub = [pops[x] for x in args]
        idx = list(range(len(args)))
        if fn.startswith("hap_dd"):
            return o.hap_dd(sub, idx[0], idx[-1])
        if fn.startswith("hap_dz"):
            i, j, k = {"same": (0, 0, 0), "iik": (0, 0, 1), "iji": (0, 1, 0), "ijj": (0, 1, 1),
                       "ijk": (0, 1, 2)}[fn[7:]]
            return o.hap_dz(sub, i, j, k)
        q = {"same": (0, 0, 0, 0), "triple": (0, 0, 0, 1), "iikk": (0, 0, 1, 1),
             "iikl": (0, 0, 1, 2), "ijkk": (0, 1, 2, 2), "ijij": (0, 1, 0, 1),
             "shared": (0, 1, 0, 2), "alldiff": (0, 1, 2, 3)}[fn[8:]]
        return o.hap_pi2(sub, *q)

    def num(fn, args):  # reference numerator: value / prod 1/(n)_k
        r = np.prod([recip(args[pos], k) for pos, k in gen.HAP_RECIP[fn]], axis=0)
        return ref_call(fn, args) / r

    ref = o.multi_pop_stats_hap(counts, [None] * P)
    got, mag = np.full(ref.shape, np.nan), np.full(ref.shape, np.nan)
    for k in range(int(re.search(rf"#define HAP_NGROUPS_P{P} (\d+)", inc).group(1))):
        body = re.search(rf"void hap_acc_P{P}_g{k}\(.*?\n(.*?)\n\}}", inc, re.S).group(1)
        env = {}
        for name, expr in re.findall(r"const double (R\d+) = (.*?) \* wt;", body):
            env[name] = np.prod([recip(int(p), int(kk))
                                 for p, kk in re.findall(r"p\[(\d)\]\.r(\d)", expr)], axis=0)
        for name, fn, form, args in re.findall(
                r"const double ([nm]\d+) = (hap_\w+?)_(num|raw)\(([^)]*)\);", body):
            a = [int(x) for x in re.findall(r"p\[(\d)\]", args)]
            if fn == "hap_dz_iik_iji_avg":
                v = 0.5 * (num("hap_dz_iik", a) + num("hap_dz_iji", a))
            else:
                v = num(fn, a)
            env[name] = v / inner[fn] if form == "raw" else v
        accs, absl = {}, {}
        for s, chain in re.findall(r"acc\[(\d+)\] = (fma\(.*\));", body):
            terms = re.findall(r"fma\(((?:(?!fma\().)*?), (R\d+), ", chain)
            vals = [eval(e, {}, dict(env)) * env[r] for e, r in terms]
            accs[int(s)] = sum(vals)
            absl[int(s)] = sum(np.abs(v) for v in vals)
        words = re.search(rf"hap_scale_group<{P}, {k}>\(int q\) \{{\n  const unsigned long long w = "
                          r"(.*?);", inc).group(1)
        w = [int(x, 16) for x in re.findall(r"0x([0-9a-f]+)ull", words)]
        # the emitted output-column table of the group (7-bit fields, 9 per word)
        cwords = re.search(rf"hap_stat_col<{P}, {k}>\(int q\) \{{\n.*?\n"
                           r"  const unsigned long long w = (.*?);", inc, re.S).group(1)
        cw = [int(x, 16) for x in re.findall(r"0x([0-9a-f]+)ull", cwords)]
        for s in sorted(accs):
            e = (w[s // 16] >> (4 * (s % 16))) & 15
            c = (cw[s // 9] >> (7 * (s % 9))) & 127
            assert np.isnan(got[0, c])           # every column written once
            got[:, c] = accs[s] * 2.0 ** -e
            mag[:, c] = absl[s] * 2.0 ** -e
    assert not np.isnan(got).any()
    # relative to the terms' own magnitude where orientations cancel, with a
    # floor at 1e-4 of the statistic's largest value (the numerators here are
    # recovered from rounded reference values, which need not vanish exactly
    # where the reference's own sum does)
    floor = 1e-4 * np.max(np.abs(ref), axis=0)
    scale = np.maximum(np.maximum(np.abs(ref), mag), np.maximum(floor, 1e-300))
    assert np.max(np.abs(got - ref) / scale) < 1e-12


def header_status_codes():
    text = (ROOT / "include/pgld.h").read_text()
    return {k: int(v) for k, v in re.findall(r"\b(PGLD_E[A-Z]+|PGLD_OK)\s*=\s*(\d+)", text)}


def test_status_codes_agree_across_header_binding_and_docs():
    """include/pgld.h's status enum, the ctypes binding's constants and the
    binding documented for pg_gpu maintainers name the same codes, and the
    documented _check stub raises the exception the reference raises."""
    from pg_gpu_amd import _lib

    codes = header_status_codes()
    assert codes == {"PGLD_OK": 0, "PGLD_EINVAL": 1, "PGLD_EDEVICE": 2, "PGLD_EUNSUPPORTED": 3}
    for name, v in codes.items():
        assert getattr(_lib, name) == v
    doc = (ROOT / "INTEGRATION.md").read_text()
    for name, v in codes.items():
        assert f"{v} `{name}`" in doc
    # the documented stub: run its _check against a stand-in library
    block = re.search(r"```python\n(.*?)```", doc, flags=re.S).group(1)
    src = block[block.index("PGLD_OK, PGLD_EINVAL"):]

    class FakeLib:
        @staticmethod
        def pgld_last_error():
            return b"msg"

    ns = {"_lib": FakeLib()}
    exec(src, ns)
    ns["_check"](0)
    for rc, exc in ((1, ValueError), (2, RuntimeError), (3, NotImplementedError)):
        with pytest.raises(exc):
            ns["_check"](rc)
        with pytest.raises(exc):
            _check_via_lib(rc)


def _check_via_lib(rc):
    from pg_gpu_amd import _lib

    class L:
        @staticmethod
        def pgld_last_error():
            return b"msg"

    saved = _lib._lib
    _lib._lib = L()
    try:
        _lib.check(rc)
    finally:
        _lib._lib = saved


def test_estimate_ld_chunk_size_hand_values():
    """Pairs per chunk = clamp(mem // ((4 H P + 150 P) * 3), 1e5, 1e7)
    (ld_pipeline.py:16-49), worked by hand."""
    from pg_gpu_amd.ld_pipeline import estimate_ld_chunk_size

    # H = 500, P = 2: (4000 + 300) * 3 = 12,900 bytes per pair
    assert estimate_ld_chunk_size(500, 12_900 * 2_000_000, 2) == 2_000_000
    assert estimate_ld_chunk_size(500, 12_900 * 2_000_000 + 12_899, 2) == 2_000_000
    assert estimate_ld_chunk_size(500, 10**9, 2) == 100_000          # 77,519 -> floor
    assert estimate_ld_chunk_size(10, 10**12, 1) == 10_000_000       # ceiling
    # H = 1000, P = 4: (16,000 + 600) * 3 = 49,800
    assert estimate_ld_chunk_size(1000, 49_800 * 123_456, 4) == 123_456


def header_struct_fields(name):
    """Field names of `typedef struct { ... } name;` in include/pgld.h."""
    text = (ROOT / "include/pgld.h").read_text()
    end = re.search(r"\}\s*" + name + ";", text).start()
    body = text[text.rindex("typedef struct {", 0, end) + len("typedef struct {"):end]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    return re.findall(r"(\w+)\s*(?:\[[^\]]*\])?\s*;", body)


@pytest.mark.parametrize("cname,pyname", [("pgld_binned_args", "BinnedArgs"),
                                          ("pgld_window_args", "WindowArgs"),
                                          ("pgld_panel_desc", "PanelDesc")])
def test_ctypes_structs_match_the_header(cname, pyname):
    """The ctypes mirrors of the argument structs list the header's fields in
    the header's order (a missing or reordered field shifts every later one)."""
    from pg_gpu_amd import _lib

    want = header_struct_fields(cname)
    got = [f[0] for f in getattr(_lib, pyname)._fields_]
    assert got == want
    if pyname in ("BinnedArgs", "PanelDesc"):
        doc = (ROOT / "INTEGRATION.md").read_text()
        stub = re.search(r"class " + pyname + r"\(ctypes.Structure\):\n\s*_fields_ = \[(.*?)\]\n",
                         doc, flags=re.S).group(1)
        assert re.findall(r'\("(\w+)"', stub) == want


def test_window_launch_pick_prefers_short_turns_within_the_tie():
    from pg_gpu_amd.autotune import WINDOW_TIE, pick_window_launch

    def c(tile, group):
        return dict(tile=tile, group=group)

    t = [(c(32, 16), 1.03), (c(32, 32), 1.015), (c(32, 64), 1.00), (c(16, 64), 2.0)]
    assert pick_window_launch(t) == c(32, 32)           # 16 is outside the 2 % tie
    t = [(c(32, 64), 1.00), (c(32, 16), 1.0 + 0.99 * WINDOW_TIE)]
    assert pick_window_launch(t) == c(32, 16)
    t = [(c(16, 16), 1.01), (c(32, 16), 1.00)]          # equal groups: the faster
    assert pick_window_launch(t) == c(32, 16)
    assert pick_window_launch(t, tie=0.0) == c(32, 16)
    assert pick_window_launch([(c(32, 16), 1.5), (c(32, 64), 1.0)], tie=0.0) == c(32, 64)
