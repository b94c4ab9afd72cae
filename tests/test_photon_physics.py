"""Pins the photon walk's sampling decisions against what the mathematics
fixes, independently of the code that makes them:

- the isotropic direction sampler: unit length; E[u] = 0, E[u_i^2] = 1/3 and
  E[u_x u_y] = 0 within 5 sigma; u_z and the azimuth uniform by chi-square
  (Archimedes: u_z of an isotropic direction is uniform on [-1, 1]);
- roulette: survival frequency r = w / target within 5 sigma, survivors at
  exactly the target weight, so the expected weight after roulette is w;
- splitting: the copy count is the largest power of two <= min(r, max_split)
  (computed here from the ratio alone) and copies x weight = w exactly;
- the whole walk: roulette is weight-neutral in expectation, E[gained - lost]
  = 0 over independent batches, within 5 standard errors.

Each check is also run against the oracle rebuilt with a planted mistake in a
copy of photon.h (a dropped factor, a wrong sign, an inverted comparison),
and must then fail: the checks are sharp enough to catch those slips.
"""

import ctypes
import math
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest
from scipy import stats as sps

import oracle_lib as O

ROOT = Path(__file__).resolve().parents[1]
SPEC = ROOT / "harmonize_amd" / "specs" / "photon.h"

KEEP, SURVIVE, DIE, SPLIT = 0, 1, 2, 3

_F = ctypes.c_float
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64
_VP = ctypes.c_void_p


def _bind(h):
    h.oracle_photon_directions.restype = None
    h.oracle_photon_directions.argtypes = [_U64, _U32, _U32, _U32, _U32, _VP]
    h.oracle_photon_window.restype = None
    h.oracle_photon_window.argtypes = [_F, _U32, _F, _F, _U32, _U64, _U32, _VP, _VP, _VP]
    return h


def directions(h, n, copy=0, seed=0x5EED, event=0, lineage=0):
    out = np.zeros((n, 3), dtype=np.float32)
    h.oracle_photon_directions(seed, event, lineage, copy, n, out.ctypes.data)
    return out


def window(h, w, region, n, split_above=2.0, roulette_below=0.1, max_split=8, seed=0x5EED):
    fate = np.zeros(n, dtype=np.uint32)
    copies = np.zeros(n, dtype=np.uint32)
    w_out = np.zeros(n, dtype=np.float32)
    h.oracle_photon_window(w, region, split_above, roulette_below, max_split, seed, n, fate.ctypes.data,
                           copies.ctypes.data, w_out.ctypes.data)
    return fate, copies, w_out


# ------------------------------------------------------------------ checks
# Each returns a list of failure descriptions (empty = pass).

def direction_failures(h, n=1_000_000, copy=0):
    u = directions(h, n, copy=copy).astype(np.float64)
    fails = []
    norm = np.sqrt((u * u).sum(axis=1))
    if np.abs(norm - 1.0).max() > 1e-6:
        fails.append(f"|u| deviates from 1 by {np.abs(norm - 1.0).max():.3g}")
    ux, uy, uz = u[:, 0], u[:, 1], u[:, 2]

    def moment(name, values, expect, var):
        m = values.mean()
        if abs(m - expect) > 5.0 * math.sqrt(var / n):
            fails.append(f"E[{name}] = {m:.6f}, expected {expect:.6f}")

    moment("ux", ux, 0.0, 1.0 / 3.0)
    moment("uy", uy, 0.0, 1.0 / 3.0)
    moment("uz", uz, 0.0, 1.0 / 3.0)
    # E[u_i^4] = 1/5 for an isotropic direction, so Var(u_i^2) = 1/5 - 1/9.
    moment("uz^2", uz * uz, 1.0 / 3.0, 1.0 / 5.0 - 1.0 / 9.0)
    moment("ux^2", ux * ux, 1.0 / 3.0, 1.0 / 5.0 - 1.0 / 9.0)
    # E[ux^2 uy^2] = 1/15.
    moment("ux*uy", ux * uy, 0.0, 1.0 / 15.0)
    moment("ux*uz", ux * uz, 0.0, 1.0 / 15.0)
    bins = 40
    crit = sps.chi2.isf(1e-6, bins - 1)
    counts, _ = np.histogram(uz, bins=bins, range=(-1.0, 1.0))
    chi = ((counts - n / bins) ** 2 / (n / bins)).sum()
    if chi > crit:
        fails.append(f"u_z not uniform on [-1, 1]: chi2 = {chi:.1f} > {crit:.1f}")
    phi = np.arctan2(uy, ux)
    counts, _ = np.histogram(phi, bins=bins, range=(-math.pi, math.pi))
    chi = ((counts - n / bins) ** 2 / (n / bins)).sum()
    if chi > crit:
        fails.append(f"azimuth not uniform: chi2 = {chi:.1f} > {crit:.1f}")
    return fails


def roulette_failures(h, n=1_000_000):
    fails = []
    region = 3
    target = 2.0 ** -region
    for r in (0.01, 0.05, 0.0999):
        w = np.float32(r * target)
        ratio = float(w) / target  # the ratio the walk sees (w * 2^region is exact)
        fate, copies, w_out = window(h, float(w), region, n)
        if not np.isin(fate, (SURVIVE, DIE)).all():
            fails.append(f"r={r}: outcomes other than survive/die")
            continue
        alive = fate == SURVIVE
        freq = alive.mean()
        if abs(freq - ratio) > 5.0 * math.sqrt(ratio * (1.0 - ratio) / n):
            fails.append(f"r={r}: survival frequency {freq:.5f}, expected {ratio:.5f}")
        if alive.any() and not (w_out[alive] == np.float32(target)).all():
            fails.append(f"r={r}: survivors not at the target weight")
        if not (copies == 1).all():
            fails.append(f"r={r}: roulette made copies")
        mean_w = np.where(alive, w_out.astype(np.float64), 0.0).mean()
        if abs(mean_w - float(w)) > 5.0 * target * math.sqrt(ratio * (1.0 - ratio) / n):
            fails.append(f"r={r}: E[w after roulette] = {mean_w:.6g}, expected {float(w):.6g}")
    # At or above the roulette threshold (and up to the split threshold) nothing happens.
    for r in (0.1, 0.5, 1.0, 2.0):
        w = np.float32(r * target)
        fate, copies, w_out = window(h, float(w), region, 1000)
        if not ((fate == KEEP) & (copies == 1) & (w_out == w)).all():
            fails.append(f"r={r}: expected no roulette and no split")
    return fails


def expected_copies(ratio, max_split):
    """Largest power of two <= min(ratio, max_split), from the ratio alone."""
    limit = min(ratio, float(max_split))
    return 1 << int(math.floor(math.log2(limit)))


def split_failures(h):
    fails = []
    region = 4
    target = 2.0 ** -region
    ratios = [2.0, float(np.nextafter(np.float32(2.0), np.float32(3.0))), 2.5, 3.0,
              float(np.nextafter(np.float32(4.0), np.float32(0.0))), 4.0, 5.5, 7.99, 8.0, 9.0, 15.9, 16.0, 33.0,
              1000.0, 1e6]
    for max_split in (1, 2, 8, 64, 1024):
        for r in ratios:
            w = np.float32(r * target)
            ratio = float(w) / target
            fate, copies, w_out = window(h, float(w), region, 64, max_split=max_split)
            c = expected_copies(ratio, max_split) if ratio > 2.0 else 1
            if not (copies == c).all():
                fails.append(f"r={ratio}, max_split={max_split}: {copies[0]} copies, expected {c}")
                continue
            if not (w_out.astype(np.float64) * copies == float(w)).all():
                fails.append(f"r={ratio}, max_split={max_split}: copies x weight != w")
            want_fate = SPLIT if c > 1 else KEEP
            if not (fate == want_fate).all():
                fails.append(f"r={ratio}, max_split={max_split}: fate {fate[0]}, expected {want_fate}")
    return fails


# ------------------------------------------------------------ real oracle

@pytest.fixture(scope="module")
def oracle():
    return _bind(O.lib())


@pytest.mark.parametrize("copy", [0, 1, 5])
def test_direction_sampler_is_isotropic(oracle, copy):
    assert direction_failures(oracle, copy=copy) == []


def test_directions_differ_between_copies_and_events(oracle):
    a = directions(oracle, 4096, copy=0)
    b = directions(oracle, 4096, copy=1)
    c = directions(oracle, 4096, copy=0, event=1)
    assert (a != b).any(axis=1).mean() > 0.99
    assert (a != c).any(axis=1).mean() > 0.99


def test_roulette_survival_and_expected_weight(oracle):
    assert roulette_failures(oracle) == []


def test_split_copy_count(oracle):
    assert split_failures(oracle) == []


def test_regions_beyond_32_use_the_true_power_of_two(oracle):
    # region 40: target 2^-40; r = w * 2^40 must be computed exactly (a 32-bit
    # shift would wrap and give the wrong ratio).
    region = 40
    target = 2.0 ** -region
    w = np.float32(0.05 * target)
    fate, _, w_out = window(oracle, float(w), region, 200_000)
    ratio = float(w) / target
    freq = (fate == SURVIVE).mean()
    assert np.isin(fate, (SURVIVE, DIE)).all()
    assert abs(freq - ratio) < 5.0 * math.sqrt(ratio * (1 - ratio) / 200_000)
    assert (w_out[fate == SURVIVE] == np.float32(target)).all()
    w = np.float32(12.0 * target)
    fate, copies, w_out = window(oracle, float(w), region, 16, max_split=64)
    assert (copies == 8).all() and (fate == SPLIT).all()
    assert (w_out.astype(np.float64) * 8 == float(w)).all()


def test_roulette_is_weight_neutral_over_batches():
    """E[gained - lost] = 0: over independent batches (different seeds) of
    the default walk, the mean of (roulette gained - roulette lost) is within
    5 standard errors of zero, while roulette is frequent."""
    nets, lost_total = [], 0.0
    for b in range(40):
        p = O.photon_params(photons=5000, seed=0x1000 + b)
        tally, _, _ = O.photon_program(p, count=5000, threads=0)
        lost, gained = tally[-3], tally[-2]
        nets.append(gained - lost)
        lost_total += lost
    nets = np.array(nets)
    se = nets.std(ddof=1) / math.sqrt(len(nets))
    assert lost_total > 100.0  # roulette is exercised
    assert abs(nets.mean()) < 5.0 * se, (nets.mean(), se)


# -------------------------------------------------- planted mistakes (CPU)

MUTATIONS = {
    # Marsaglia's u_z = 1 - 2s becomes 1 - s: directions lean towards +z.
    "uz_one_minus_s": ("uz = (float)(1.0 - 2.0 * s);", "uz = (float)(1.0 - s);", "direction"),
    # Dropping the factor 2 of t = 2 sqrt(1 - s): directions are not unit vectors.
    "t_without_factor_two": ("double t = 2.0 * hz::det::det_sqrt(1.0 - s);",
                             "double t = hz::det::det_sqrt(1.0 - s);", "direction"),
    # A transposed operand: u_y built from v1 instead of v2.
    "uy_from_v1": ("uy = (float)(v2 * t);", "uy = (float)(v1 * t);", "direction"),
    # A dropped square root in t = 2 sqrt(1 - s).
    "t_without_sqrt": ("double t = 2.0 * hz::det::det_sqrt(1.0 - s);", "double t = 2.0 * (1.0 - s);",
                       "direction"),
    # Surviving roulette with probability 1 - r instead of r.
    "roulette_inverted": ("if (u2 >= ratio) {", "if (u2 < ratio) {", "roulette"),
    # Split copy count capped one power of two too low.
    "split_cap": ("while ((2u << k) <= st.max_split && (float)(2u << k) <= ratio) ++k;",
                  "while ((2u << k) < st.max_split && (float)(2u << k) <= ratio) ++k;", "split"),
    # Split copy count rounded past the ratio.
    "split_round_up": ("while ((2u << k) <= st.max_split && (float)(2u << k) <= ratio) ++k;",
                       "while ((2u << k) <= st.max_split && (float)(1u << k) <= ratio) ++k;", "split"),
}


def _build_variant(tmp: Path, old: str, new: str) -> ctypes.CDLL:
    src = SPEC.read_text()
    assert src.count(old) == 1, f"mutation anchor not found exactly once: {old}"
    inc = tmp / "inc"
    inc.mkdir()
    (inc / "photon.h").write_text(src.replace(old, new))
    out = tmp / "libphoton_variant.so"
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    cmd = [gxx, "-std=c++17", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           f"-I{inc}", f"-I{ROOT / 'harmonize_amd' / 'specs'}", f"-I{ROOT / 'harmonize_amd' / 'include'}",
           f"-I{ROOT / 'oracle'}", str(ROOT / "oracle" / "photon_physics.cpp"), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    return _bind(ctypes.CDLL(str(out)))


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_planted_mistake_is_caught(tmp_path, name):
    old, new, check = MUTATIONS[name]
    h = _build_variant(tmp_path, old, new)
    if check == "direction":
        fails = direction_failures(h, n=200_000) + direction_failures(h, n=200_000, copy=3)
    elif check == "roulette":
        fails = roulette_failures(h, n=200_000)
    else:
        fails = split_failures(h)
    assert fails, f"the checks did not catch the planted mistake {name}"


def test_unmutated_variant_build_passes(tmp_path):
    # The variant build itself (same recipe, no mutation) passes every check,
    # so a caught mutation is the mutation's doing.
    h = _build_variant(tmp_path, "uz = (float)(1.0 - 2.0 * s);", "uz = (float)(1.0 - 2.0 * s);")
    assert direction_failures(h, n=200_000) == []
    assert roulette_failures(h, n=200_000) == []
    assert split_failures(h) == []
