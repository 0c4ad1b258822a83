"""Pins the CPU oracle (oracle/sarbp_oracle.c) before anything is compared with it.

Each check is chosen so that a plausible slip fails it: the aperture tables are derived by hand
from the IDL statement quoted at sar_focuser_puntual.py:57-66; the closed forms are evaluated in
50-digit mpmath from the raw inputs (a true-pi phase, a dropped beta term, a flipped exp sign, a
missing R^2 weight or a sum instead of a mean all fail them); the range-bin ties are exact by
construction; linearity and the point-target peak hold for the whole pipeline; and the C oracle
must agree bit-for-bit with the independent NumPy restatement (oracle/bp_numpy.py).
"""
import math

import mpmath
import numpy as np
import pytest

from oracle import bp_numpy as B
from oracle import oracle_lib as O
from sarbp import synth

mpmath.mp.dps = 50
PI32 = mpmath.mpf(3.141592741012573242)
C = mpmath.mpf(2.9979246e8)
ALFA = mpmath.mpf(3.335987619777693e11)
BETA = 4 * PI32 * ALFA / (C * C)


def test_constants_match_reference_values():
    # float32(pi) held in a double, and beta = 4 pi alfa / c^2 from the reference's constants
    # (sar_focuser_puntual.py:35-40) evaluated in FP64
    assert float(np.float32(np.pi)) == 3.141592741012573242
    assert abs(B.BETA_F64 - 4.664368939847232e-05) < 1e-20


# Hand-derived apertures.  IDL: jj = indgen(ceil(s/2.)*2+1); index1 = index + jj - ceil(abs(s)/2.)
# so s >= 0 gives [I-h, I+h] with h = ceil(s/2); s = -1 gives one pulse at I-1 (J = 1, h = 1);
# s <= -2 gives J <= -1, i.e. no pulse.
APERTURES = {
    0: (0, 0), 1: (-1, 1), 2: (-1, 1), 3: (-2, 2), 4: (-2, 2), 5: (-3, 3), 12: (-6, 6),
    -1: (-1, -1), -2: None, -3: None, -5: None,
}


@pytest.mark.parametrize("s", sorted(APERTURES))
@pytest.mark.parametrize("index", [10, 0, 1, 2, 18, 19, 20, -3, 25])
def test_aperture_table(s, index):
    Naz = 21
    tx = np.zeros(Naz)
    ty = np.full(Naz, -3000.0)
    tz = np.full(Naz, 3000.0)
    pulses, bins, _ = O.trace(0.0, 0.0, 0.0, index, s, tx, ty, tz, 100000, Naz, 0.0, 0.5)
    spec = APERTURES[s]
    want = [] if spec is None else [a for a in range(index + spec[0], index + spec[1] + 1) if 0 < a < Naz - 1]
    assert list(pulses) == want
    # the NumPy restatement and the integer pp count agree
    a0, J = B.aperture(index, s)
    assert [a for a in range(a0, a0 + max(J, 0)) if 0 < a < Naz - 1] == want
    m = np.ones(1, np.uint8)
    assert O.count_pp(m, np.array([index]), np.array([s]), Naz) == len(want)
    assert B.count_pp(m, np.array([index]), np.array([s]), Naz) == len(want)


def test_sint_sentinel_and_mask_give_zero():
    sc = synth.tiny_scene(seed=5)
    xf, yf, zf, mf, inf, sf = sc.flat()
    sf = sf.copy()
    mf = mf.copy()
    sf[:5] = -9999
    mf[5:10] = 0
    mf[10:15] = 2  # only mask == 1 focuses (sar_focuser_puntual.py:46)
    out, cnt = O.focus_flat(xf, yf, zf, mf, inf, sf, sc.tx, sc.ty, sc.tz, sc.raw, sc.near, sc.dra, sc.lam)
    assert np.all(out[:15] == 0) and np.all(cnt[:15] == 0)
    assert O.count_pp(mf[:15], inf[:15], sf[:15], sc.Naz) == 0


def _closed_form(x, y, z, tx, ty, tz, d, near, dra, lam, Nra):
    dx, dy, dz = mpmath.mpf(x) - mpmath.mpf(tx), mpmath.mpf(y) - mpmath.mpf(ty), mpmath.mpf(z) - mpmath.mpf(tz)
    r2 = dx * dx + dy * dy + dz * dz
    r = mpmath.sqrt(r2)
    q = (r - near) / dra
    fl = mpmath.floor(q)
    frac = q - fl
    rb = int(fl) + (1 if frac > 0.5 else 0)  # exact ties do not occur for these random inputs
    if not (0 <= rb < Nra - 1):
        return None, rb
    ph = (-4 * PI32 / mpmath.mpf(lam)) * r + BETA * r2
    f = mpmath.mpc(mpmath.cos(ph), -mpmath.sin(ph))
    return r2 * mpmath.mpc(float(d.real), float(d.imag)) * f, rb


@pytest.mark.parametrize("seed", range(6))
def test_closed_form_single_and_three_pulses(seed):
    rng = np.random.default_rng(100 + seed)
    Naz, Nra = 64, 4096
    near, dra, lam = 4000.0, 0.5, 0.031
    tx = rng.uniform(-50, 50, Naz)
    ty = -3000.0 + rng.normal(0, 0.1, Naz)
    tz = 3000.0 + rng.normal(0, 0.1, Naz)
    raw = (rng.normal(size=(Nra, Naz)) + 1j * rng.normal(size=(Nra, Naz))).astype(np.complex64)
    x, y, z = rng.uniform(-20, 20), rng.uniform(0, 900), rng.uniform(0, 50)
    idx = int(rng.integers(5, Naz - 5))
    for s, pulses in ((0, [idx]), (2, [idx - 1, idx, idx + 1])):
        terms = []
        for a in pulses:
            R = math.sqrt((x - tx[a]) ** 2 + (y - ty[a]) ** 2 + (z - tz[a]) ** 2)
            rb = int(round((R - near) / dra))
            t, rbx = _closed_form(x, y, z, tx[a], ty[a], tz[a], raw[rb, a], near, dra, lam, Nra)
            assert rbx == rb
            terms.append(t)
        want = sum(terms) / len(terms)
        out, cnt = O.focus_flat([x], [y], [z], [1], [idx], [s], tx, ty, tz, raw, near, dra, lam)
        assert cnt[0] == len(pulses)
        got = mpmath.mpc(out[0].real, out[0].imag)
        # FP64 rounding of a ~2e6 rad phase bounds the agreement at a few 1e-10
        assert abs(got - want) / abs(want) < 2e-9


def _tie_setup(heights, Nra=8):
    """Sensor straight above the pixel: R = H exactly, so (R-near)/dra hits exact half-integers."""
    Naz = len(heights) + 2
    tx = np.zeros(Naz)
    ty = np.zeros(Naz)
    tz = np.concatenate([[5000.0], np.asarray(heights, np.float64), [5000.0]])
    raw = np.zeros((Nra, Naz), np.complex64)
    raw[:, :] = (np.arange(Nra, dtype=np.float32) + 1.0)[:, None]  # bin r holds r + 1
    return tx, ty, tz, raw


# (H, expected bin or None).  near = 4000, dra = 0.5, Nra = 8 -> valid bins 0..6.
TIES = [
    (4000.25, 0), (4000.75, 2), (4001.25, 2), (4001.75, 4), (3999.75, 0), (3999.25, None),
    (4003.0, 6), (4003.25, 6), (4003.5, None), (4003.75, None), (4002.75, 6),
]


def test_range_bin_ties_half_even_and_last_bin_excluded():
    heights = [h for h, _ in TIES]
    tx, ty, tz, raw = _tie_setup(heights)
    n = len(heights)
    ones = np.ones(n, np.uint8)
    out, cnt = O.focus_flat(np.zeros(n), np.zeros(n), np.zeros(n), ones, np.arange(1, n + 1, dtype=np.int32),
                            np.zeros(n, np.int32), tx, ty, tz, raw, 4000.0, 0.5, 0.031)
    for k, (h, want) in enumerate(TIES):
        if want is None:
            assert cnt[k] == 0 and out[k] == 0, h
        else:
            assert cnt[k] == 1, h
            assert round(abs(out[k]) / (h * h)) - 1 == want, h
    # the NumPy restatement applies the same rule
    for k, (h, want) in enumerate(TIES):
        v, c = B.process_point(0.0, 0.0, 0.0, 1, k + 1, 0, tx, ty, tz, raw, 4000.0, 0.5, 0.031)
        assert c == (0 if want is None else 1)


def test_linearity_in_raw_and_zero_when_no_contributor():
    sc = synth.tiny_scene(seed=7)
    rng = np.random.default_rng(3)
    A = sc.raw
    Bm = (rng.normal(size=A.shape) + 1j * rng.normal(size=A.shape)).astype(np.complex64)
    al, be = 0.75, -1.25  # exact in float32, so al*A + be*B is the same complex64 sum on both sides
    mix = np.asfortranarray((np.complex64(al) * A + np.complex64(be) * Bm).astype(np.complex64))
    f = lambda raw: O.focus(sc.x, sc.y, sc.z, sc.mask, sc.index, sc.tx, sc.ty, sc.tz, raw, sc.near, sc.dra,
                            sc.lam, sc.sint)
    oa, ca = f(A)
    ob, cb = f(Bm)
    om, cm = f(mix)
    assert np.array_equal(ca, cb) and np.array_equal(ca, cm)
    lin = al * oa + be * ob
    scale = np.abs(lin).max()
    # mix is rounded to complex64, so agreement is at float32 relative precision of the data
    assert np.abs(om - lin).max() <= 1e-6 * scale
    assert np.all(om[cm == 0] == 0)
    assert (cm == 0).any() and (cm > 0).any()


def test_point_target_peaks_at_target_pixel():
    sc = synth.tiny_scene(seed=11, nx=40, ny=16, naz=1001, nra=200, ntarget=1, snr_db=None, sint_lo=300,
                          sint_hi=300, edge_cases=False)
    sc.mask[:] = 1
    out, _ = O.focus(*sc.args())
    tgx, tgy = sc.targets[0, 0], sc.targets[0, 1]
    i, j = np.unravel_index(np.argmax(np.abs(out)), out.shape)
    assert abs(sc.x[i, j] - tgx) <= 1.0 and abs(sc.y[i, j] - tgy) <= 4.0


def test_c_oracle_equals_numpy_restatement_bitwise():
    sc = synth.tiny_scene(seed=13)
    oc, cc = O.focus(*sc.args(), nthreads=1)
    on, cn = B.focus(*sc.args())
    assert np.array_equal(cc, cn)
    assert np.array_equal(oc.real, on.real) and np.array_equal(oc.imag, on.imag)


def test_threads_do_not_change_results():
    sc = synth.small_scene(seed=3)
    o1, c1 = O.focus(*sc.args(), nthreads=1)
    o8, c8 = O.focus(*sc.args(), nthreads=4)
    assert np.array_equal(c1, c8) and np.array_equal(o1.real, o8.real) and np.array_equal(o1.imag, o8.imag)


def test_count_pp_matches_brute_force():
    rng = np.random.default_rng(9)
    n, Naz = 500, 300
    idx = rng.integers(-50, Naz + 50, n).astype(np.int32)
    s = rng.integers(-6, 200, n).astype(np.int32)
    s[::17] = -9999
    m = (rng.random(n) < 0.9).astype(np.uint8)
    brute = 0
    for i in range(n):
        if m[i] != 1 or s[i] == -9999:
            continue
        a0, J = B.aperture(int(idx[i]), int(s[i]))
        brute += sum(1 for a in range(a0, a0 + max(J, 0)) if 0 < a < Naz - 1)
    assert O.count_pp(m, idx, s, Naz) == brute == B.count_pp(m, idx, s, Naz)


def test_pair_arithmetic_matches_numpy_and_mpmath():
    # the per-pair oracle (used to check the device arithmetic bit for bit) equals NumPy's
    # elementwise evaluation in the reference's order, and its phase agrees with 50-digit mpmath
    rng = np.random.default_rng(5)
    naz, nra, near, dra, lam = 501, 4762, 4000.0, 0.5, 0.031
    tx = np.arange(naz) * 0.1 + rng.normal(0, 0.05, naz)
    ty = -3000.0 + rng.normal(0, 0.05, naz)
    tz = 3000.0 + rng.normal(0, 0.05, naz)
    n = 20000
    x, y, z = rng.uniform(-1500, 1900, n), rng.uniform(0, 1000, n), rng.uniform(0, 50, n)
    pulse = rng.integers(0, naz, n).astype(np.int32)
    got = O.pairs(x, y, z, pulse, tx, ty, tz, nra, near, dra, lam, nthreads=2)
    dx, dy, dz = x - tx[pulse], y - ty[pulse], z - tz[pulse]
    r2 = dx * dx + dy * dy + dz * dz
    r = np.sqrt(r2)
    q = np.round((r - near) / dra)
    ph = (-4.0 * B.PI_F64 / lam) * r + B.BETA_F64 * r2
    assert np.array_equal(got["r2"], r2) and np.array_equal(got["r"], r) and np.array_equal(got["ph"], ph)
    assert np.array_equal(got["bin"], np.where((q < nra - 1) & (q >= 0), q, -1).astype(np.int32))
    assert np.array_equal(got["sin"], np.sin(ph)) and np.array_equal(got["cos"], np.cos(ph))
    mpmath.mp.dps = 50
    for i in range(0, n, 2000):  # sin/cos of the (exactly represented) phase, to 1 ulp
        p = mpmath.mpf(float(ph[i]))
        assert abs(float(mpmath.sin(p)) - got["sin"][i]) <= 2.3e-16
        assert abs(float(mpmath.cos(p)) - got["cos"][i]) <= 2.3e-16


def _golden():
    import os
    f = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tiny_16x8_naz257_nra64.npz")
    with np.load(f) as z:
        return {k: z[k] for k in z.files}


def test_c_oracle_reproduces_golden_fixture():
    g = _golden()
    out, cnt = O.focus(g["x"], g["y"], g["z"], g["mask"], g["index"], g["tx"], g["ty"], g["tz"], g["raw"],
                       float(g["near"]), float(g["dra"]), float(g["lam"]), g["sint"], nthreads=2)
    assert np.array_equal(cnt, g["ncontrib"])
    assert np.array_equal(out, g["out"])
    assert (cnt == 0).any() and (cnt > 0).sum() > 64  # the fixture has empty and focused pixels


def test_ind_line_ra_min_gate():
    """process_point's ind_line_ra_min (sar_focuser_puntual.py:27, 106, 138) in both restatements:
    the C oracle equals the NumPy restatement bit for bit, and the contributor count with a floor L
    is the number of the pixel's in-swath bins (the oracle's own trace of bins, taken without the
    floor) that are >= max(L, 0) -- negative floors keep every bin, since the accumulation loop drops
    negative bins anyway; 0 reproduces the plain call."""
    sc = synth.edge_grid_scene(end=False)
    x, y, z, m, idx, s = sc.flat()
    checked = 0
    for i in range(0, x.size, 3):
        _, bins, _ = O.trace(float(x[i]), float(y[i]), float(z[i]), int(idx[i]), int(s[i]), sc.tx, sc.ty, sc.tz,
                             sc.Nra, sc.Naz, sc.near, sc.dra)
        if m[i] != 1 or s[i] == -9999:
            bins = bins[:0]
        for lo in (-4, 0, 1, 2, 30, 62, 63):
            v, n = O.point(x[i], y[i], z[i], m[i], idx[i], s[i], sc.tx, sc.ty, sc.tz, sc.raw, sc.near, sc.dra,
                           sc.lam, ind_line_ra_min=lo)
            assert n == int(np.sum(bins >= max(lo, 0))), (i, lo)
            w, nb = B.process_point(x[i], y[i], z[i], m[i], idx[i], s[i], sc.tx, sc.ty, sc.tz, sc.raw, sc.near,
                                    sc.dra, sc.lam, ind_line_ra_min=lo)
            assert nb == n and w == v, (i, lo)
            if lo == 0:
                assert (v, n) == O.point(x[i], y[i], z[i], m[i], idx[i], s[i], sc.tx, sc.ty, sc.tz, sc.raw, sc.near,
                                         sc.dra, sc.lam)
            checked += n > 0
    assert checked > 50
