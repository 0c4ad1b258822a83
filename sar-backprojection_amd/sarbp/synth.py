"""Seeded synthetic scenes in the reference's data model.

A scene is the 13 inputs that ``FocalizzatoreBpPx`` takes (sar_focuser_puntual.py:160): DEM
``x, y, z`` (float64, shape (DimRg, DimAz)), ``mask`` (uint8), ``index``/``sint`` (int32), the
trajectory ``tx, ty, tz`` (float64[Naz]), RawData (complex64, shape (Nra, Naz), column-major as
the reference stores it, sar_focuser_puntual.py:480-491) and the scalars near range, range
spacing and wavelength.

Workloads:

* :func:`small_scene`  -- 500 x 50 pixels over a 5001-pulse track, 600 range bins, 40 targets.
* :func:`full_scene`   -- 4000 x 250 pixels, 40001 pulses at 0.1 m, 4762 range bins,
  2000 targets; ``strip`` keeps only the pixel columns of a sub-range for CPU checks.
* :func:`tiny_scene`   -- a few-hundred-pixel scene for unit tests.

Raw echoes are ``A * sinc((r - R)/dra) * exp(-j*4*pi*R/lambda)`` per target (r = near + bin*dra),
truncated to +-``SINC_HALF`` bins around the peak and to pulses within a +-3 degree squint beam,
plus circular complex Gaussian noise at the stated SNR (noise power = mean signal power / SNR).
"""
from __future__ import annotations

import dataclasses
import math
import os

import numpy as np

SINC_HALF = 8  # range bins either side of a target's peak that receive its echo
BEAM_DEG = 3.0  # squint half-width of the illuminating beam (degrees)


@dataclasses.dataclass
class Scene:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    mask: np.ndarray
    index: np.ndarray
    sint: np.ndarray
    tx: np.ndarray
    ty: np.ndarray
    tz: np.ndarray
    raw: np.ndarray  # complex64 (Nra, Naz), Fortran-ordered
    near: float
    dra: float
    lam: float
    targets: np.ndarray | None = None  # (ntarget, 4): x, y, z, amplitude
    name: str = ""

    @property
    def shape(self):
        return self.x.shape

    @property
    def Nra(self) -> int:
        return int(self.raw.shape[0])

    @property
    def Naz(self) -> int:
        return int(self.raw.shape[1])

    def args(self):
        """Arguments in FocalizzatoreBpPx order (sar_focuser_puntual.py:160)."""
        return (self.x, self.y, self.z, self.mask, self.index, self.tx, self.ty, self.tz, self.raw,
                self.near, self.dra, self.lam, self.sint)

    @property
    def dimx(self) -> int:
        return int(self.x.shape[0])

    def flat(self):
        """F-order flattened pixel arrays (x, y, z, mask, index, sint)."""
        f = lambda a: np.ascontiguousarray(np.asarray(a).flatten(order="F"))
        return f(self.x), f(self.y), f(self.z), f(self.mask), f(self.index), f(self.sint)

    def sub(self, ix0: int, ix1: int, iy0: int = 0, iy1: int | None = None) -> "Scene":
        """The pixel sub-grid [ix0:ix1, iy0:iy1] sharing this scene's raw data and trajectory."""
        sl = (slice(ix0, ix1), slice(iy0, iy1))
        g = lambda a: np.asfortranarray(a[sl])
        return dataclasses.replace(self, x=g(self.x), y=g(self.y), z=g(self.z), mask=g(self.mask),
                                   index=g(self.index), sint=g(self.sint),
                                   name=f"{self.name}[{ix0}:{ix1},{iy0}:{iy1}]")


def _simulate_raw(rng, targets, tx, ty, tz, Nra, near, dra, lam, snr_db):
    Naz = tx.size
    raw = np.zeros((Nra, Naz), dtype=np.complex64, order="F")
    offs = np.arange(-SINC_HALF, SINC_HALF + 1)
    tanb = math.tan(math.radians(BEAM_DEG))
    for (px, py, pz, amp) in targets:
        # pulses whose along-track offset is inside the beam at this target's ground range
        ground = math.hypot(py - float(np.mean(ty)), pz - float(np.mean(tz)))
        half = tanb * ground
        sel = np.nonzero(np.abs(tx - px) <= half)[0]
        if sel.size == 0:
            continue
        R = np.sqrt((px - tx[sel]) ** 2 + (py - ty[sel]) ** 2 + (pz - tz[sel]) ** 2)
        b0 = np.rint((R - near) / dra).astype(np.int64)
        rb = b0[:, None] + offs[None, :]
        ok = (rb >= 0) & (rb < Nra)
        rr = near + rb * dra
        val = amp * np.sinc((rr - R[:, None]) / dra) * np.exp(-1j * 4.0 * math.pi * R / lam)[:, None]
        pa = np.broadcast_to(sel[:, None], rb.shape)
        raw[rb[ok], pa[ok]] += val[ok].astype(np.complex64)
    p_sig = float(np.mean(np.abs(raw.astype(np.complex128, copy=False)) ** 2)) if raw.size < 5e7 else \
        float(np.mean(np.abs(raw[:, :: max(1, Naz // 2000)]) ** 2))
    if snr_db is not None and p_sig > 0:
        sigma = math.sqrt(p_sig / (10.0 ** (snr_db / 10.0)) / 2.0)
        for a0 in range(0, Naz, 4096):  # chunked to bound memory
            a1 = min(Naz, a0 + 4096)
            noise = rng.standard_normal((2, Nra, a1 - a0), dtype=np.float32) * np.float32(sigma)
            raw[:, a0:a1] += noise[0] + 1j * noise[1]
    return raw


def _mask(rng, shape, p=0.95):
    return np.asfortranarray((rng.random(shape) < p).astype(np.uint8))


def small_scene(seed: int = 0, snr_db: float | None = 20.0) -> Scene:
    """500 x 50 pixels, 5001 pulses, 600 range bins, 40 targets, Sint = 400.

    DEM x = 0..2495 m and y = 0..245 m at 5 m, z = 20*exp(-((x-1250)^2+(y-125)^2)/(2*300^2)).
    Trajectory x = linspace(-500, 3000, 5001), y = -3000, z = 3000, each with N(0, 0.05^2) jitter.
    lambda = 0.031 m, near = 4000 m, dra = 0.5 m.  Index = the first pulse whose squint from the
    pixel is within +-3 degrees; mask ~ Bernoulli(0.95)."""
    rng = np.random.default_rng(seed)
    xs = np.arange(500) * 5.0
    ys = np.arange(50) * 5.0
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    Z = 20.0 * np.exp(-((X - 1250.0) ** 2 + (Y - 125.0) ** 2) / (2.0 * 300.0 ** 2))
    Naz = 5001
    tx = np.linspace(-500.0, 3000.0, Naz) + rng.normal(0, 0.05, Naz)
    ty = -3000.0 + rng.normal(0, 0.05, Naz)
    tz = 3000.0 + rng.normal(0, 0.05, Naz)
    near, dra, lam, Nra = 4000.0, 0.5, 0.031, 600
    # Index: first pulse with |squint| <= 3 deg.  squint = asin(along-track offset / slant range)
    sin3 = math.sin(math.radians(3.0))
    index = np.zeros(X.shape, np.int32)
    for j in range(X.shape[1]):
        dxs = tx[None, :] - X[:, j:j + 1]
        R = np.sqrt(dxs ** 2 + (ty[None, :] - Y[:, j:j + 1]) ** 2 + (tz[None, :] - Z[:, j:j + 1]) ** 2)
        inside = np.abs(dxs / R) <= sin3
        index[:, j] = np.argmax(inside, axis=1)
    ntg = 40
    tgx = rng.uniform(0, xs[-1], ntg)
    tgy = rng.uniform(0, ys[-1], ntg)
    tgz = 20.0 * np.exp(-((tgx - 1250.0) ** 2 + (tgy - 125.0) ** 2) / (2.0 * 300.0 ** 2))
    amp = np.abs(rng.normal(0, 1, ntg))
    targets = np.stack([tgx, tgy, tgz, amp], axis=1)
    raw = _simulate_raw(rng, targets, tx, ty, tz, Nra, near, dra, lam, snr_db)
    sint = np.full(X.shape, 400, np.int32, order="F")
    return Scene(np.asfortranarray(X), np.asfortranarray(Y), np.asfortranarray(Z),
                 _mask(rng, X.shape), np.asfortranarray(index), sint, tx, ty, tz, raw,
                 near, dra, lam, targets, "small-500x50")


def hills(X, Y, rng, nhill=8, zmax=50.0):
    """Smooth random DEM: sum of ``nhill`` Gaussian hills scaled to [0, zmax] m."""
    Z = np.zeros_like(X)
    x0, x1 = float(X.min()), float(X.max())
    y0, y1 = float(Y.min()), float(Y.max())
    for _ in range(nhill):
        cx, cy = rng.uniform(x0, x1), rng.uniform(y0, y1)
        s = rng.uniform(150.0, 600.0)
        a = rng.uniform(0.3, 1.0)
        Z += a * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * s * s))
    return Z * (zmax / max(float(Z.max()), 1e-12))


def full_scene(seed: int = 1, snr_db: float | None = 20.0, ntarget: int = 2000,
               nx: int = 4000, ny: int = 250, naz: int = 40001, nra: int = 4762) -> Scene:
    """4000 x 250 pixels at 1 m (x) x 4 m (y) on a DEM of 8 Gaussian hills (0-50 m).

    40001 pulses along x at 0.1 m spacing, altitude 3000 m, cross-track offset -3000 m.
    RawData 4762 x 40001 complex64 (~1.5 GB), 2000 targets + 20 dB noise; lambda = 0.031 m,
    dra = 0.5 m, near = 4000 m.  Index = the broadside pulse (round(x/0.1)); Sint ~ U[2500,3500];
    mask ~ Bernoulli(0.95).  ``nx/ny/naz/nra`` scale the same geometry down for tests."""
    rng = np.random.default_rng(seed)
    xs = np.arange(nx) * 1.0
    ys = np.arange(ny) * 4.0
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    Z = hills(X, Y, rng)
    tx = np.arange(naz) * 0.1
    ty = np.full(naz, -3000.0)
    tz = np.full(naz, 3000.0)
    near, dra, lam = 4000.0, 0.5, 0.031
    index = np.asfortranarray(np.clip(np.rint(X / 0.1), 0, naz - 1).astype(np.int32))
    sint = np.asfortranarray(rng.integers(2500, 3501, X.shape).astype(np.int32))
    tgx = rng.uniform(0, xs[-1], ntarget)
    tgy = rng.uniform(0, ys[-1], ntarget)
    # target heights from the same DEM (nearest pixel)
    tgz = Z[np.clip(np.rint(tgx).astype(int), 0, nx - 1), np.clip(np.rint(tgy / 4).astype(int), 0, ny - 1)]
    amp = np.abs(rng.normal(0, 1, ntarget))
    targets = np.stack([tgx, tgy, tgz, amp], axis=1)
    raw = _simulate_raw(rng, targets, tx, ty, tz, nra, near, dra, lam, snr_db)
    return Scene(np.asfortranarray(X), np.asfortranarray(Y), np.asfortranarray(Z), _mask(rng, X.shape),
                 index, sint, tx, ty, tz, raw, near, dra, lam, targets, f"full-{nx}x{ny}x{naz}")


def tiny_scene(seed: int = 2, nx: int = 24, ny: int = 12, naz: int = 801, nra: int = 80,
               sint_lo: int = 40, sint_hi: int = 120, ntarget: int = 4, snr_db: float | None = 20.0,
               edge_cases: bool = True) -> Scene:
    """A small scene in the full-scene geometry scaled down (pixel 1 m x 4 m, pulses 0.1 m apart,
    R ~ 4.2-4.3 km so the phase magnitude matches the full workload).  With ``edge_cases`` it
    plants Sint = -9999 / -1 / -2 / 0 / 1, masked pixels, apertures clipped at both track ends
    and pixels beyond the last range bin."""
    rng = np.random.default_rng(seed)
    xs = 30.0 + np.arange(nx) * 1.0
    ys = np.arange(ny) * 4.0
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    Z = hills(X, Y, rng, nhill=3, zmax=20.0)
    tx = np.arange(naz) * 0.1 + rng.normal(0, 0.02, naz)
    ty = np.full(naz, -3000.0) + rng.normal(0, 0.02, naz)
    tz = np.full(naz, 3000.0) + rng.normal(0, 0.02, naz)
    near, dra, lam = 4230.0, 0.5, 0.031
    index = np.asfortranarray(np.clip(np.rint(X / 0.1), 0, naz - 1).astype(np.int32))
    sint = np.asfortranarray(rng.integers(sint_lo, sint_hi + 1, X.shape).astype(np.int32))
    mask = _mask(rng, X.shape)
    if edge_cases:
        specials = [-9999, -1, -2, 0, 1, -3, 2, 5]
        for k, s in enumerate(specials):
            sint[k % nx, (k * 3) % ny] = s
        mask[1 % nx, 1 % ny] = 0
        index[0, 2 % ny] = 0              # aperture clipped at the track start
        index[nx - 1, 3 % ny] = naz - 1   # ... and at the track end
        index[2 % nx, 4 % ny] = -50       # centre before the first pulse
        index[3 % nx, 5 % ny] = naz + 30  # centre after the last pulse
    tgx = rng.uniform(xs[0], xs[-1], ntarget)
    tgy = rng.uniform(ys[0], ys[-1], ntarget)
    tgz = Z[np.clip(np.rint(tgx - xs[0]).astype(int), 0, nx - 1), np.clip(np.rint(tgy / 4).astype(int), 0, ny - 1)]
    amp = np.abs(rng.normal(0, 1, ntarget)) + 0.5
    targets = np.stack([tgx, tgy, tgz, amp], axis=1)
    raw = _simulate_raw(rng, targets, tx, ty, tz, nra, near, dra, lam, snr_db)
    return Scene(np.asfortranarray(X), np.asfortranarray(Y), np.asfortranarray(Z), mask, index, sint,
                 tx, ty, tz, raw, near, dra, lam, targets, f"tiny-{nx}x{ny}")


EDGE_SINTS = tuple(range(-5, 13)) + (-9999, 400, 2000)


def edge_grid_scene(seed: int = 3, naz: int = 4001, nra: int = 64, near: float = 4300.0, dra: float = 0.5,
                    lam: float = 0.031, end: bool = False) -> Scene:
    """Every aperture and range-bin edge case on one grid: columns take the aperture centres
    Index = -3..3 (``end=False``) or Naz-4..Naz+2 (``end=True``), rows the Sint values
    ``EDGE_SINTS`` (-5..12, -9999, and two long apertures, 400 and 2000 pulses, the second with a
    range migration that crosses a bin edge inside the aperture), in three bands of rows whose
    broadside range bins step by 0.3 bins across bin 0, across the last bin Nra-1 (excluded by the
    reference) and inside the swath.  Pixels 1 m apart along the track near its start (or end), z = 0; track along x at 0.1 m
    with 2 cm jitter, y = -3000 m, z = 3000 m; RawData is seeded complex Gaussian noise.  Grid shape
    (7, 3 * len(EDGE_SINTS)), F order."""
    rng = np.random.default_rng(seed)
    tx = np.arange(naz) * 0.1 + rng.normal(0, 0.02, naz)
    ty = np.full(naz, -3000.0) + rng.normal(0, 0.02, naz)
    tz = np.full(naz, 3000.0) + rng.normal(0, 0.02, naz)
    idx = np.arange(naz - 4, naz + 3) if end else np.arange(-3, 4)
    x0 = (naz - 1) * 0.1 - 6.0 if end else 0.0
    ns = len(EDGE_SINTS)
    # per band, the bin edge it straddles: bin 0 (the first that counts), Nra-1 (the first that does
    # not), and the swath's middle.  Short apertures step across the edge by 0.3 bins from edge-2.4;
    # the 400-pulse aperture sits 0.4 bins below it, the 2000-pulse one 1.4 bins below, so that its
    # range migration (~2.3 bins over 100 m of track) carries it across the edge.
    edges = (0.0, nra - 1.0, nra / 2.0)
    offset = {400: -0.4, 2000: -1.4}
    nx, ny = idx.size, 3 * ns
    X = np.empty((nx, ny))
    Y = np.empty((nx, ny))
    index = np.empty((nx, ny), np.int32)
    sint = np.empty((nx, ny), np.int32)
    for b, e in enumerate(edges):
        for r, s in enumerate(EDGE_SINTS):
            j = b * ns + r
            R = near + (e + offset.get(s, -2.4 + 0.3 * r)) * dra
            Y[:, j] = np.sqrt(R * R - 3000.0 ** 2) - 3000.0
            X[:, j] = x0 + np.arange(nx)
            index[:, j] = idx
            sint[:, j] = s
    Z = np.zeros((nx, ny))
    mask = np.ones((nx, ny), np.uint8)
    raw = np.asfortranarray((rng.standard_normal((nra, naz)) + 1j * rng.standard_normal((nra, naz))).astype(np.complex64))
    f = np.asfortranarray
    return Scene(f(X), f(Y), f(Z), f(mask), f(index), f(sint), tx, ty, tz, raw, near, dra, lam, None,
                 f"edge-grid-{'end' if end else 'start'}")


def write_long_track(data_dir: str, dim_rg: int = 1000, dim_az: int = 32000, naz: int = 403484, nra: int = 4762,
                     ntraj: int | None = None, sint_range=(2000, 3000), ntarget: int = 20,
                     noise_columns=(10, 11, 16000, 16001, 31990, 31991), seed: int = 5) -> dict:
    """Write the long-track dataset of main_gpu.py (main_gpu.py:304-333: RawData 4762 x 403484,
    trajectory files one sample longer, 403485) as the reference's input files, without holding it in
    memory.  Track along x at 0.1 m (2 cm jitter), y = -3000 m, z = 3000 m; DEM of dim_rg x dim_az
    pixels, 1 m across track (y) and 0.1 * naz / dim_az m along it (x), z = 20 + 10 sin(x/500) cos(y/300)
    m; Index = the broadside pulse, Sint uniform in ``sint_range``, mask Bernoulli(0.95); lambda 0.031
    m, near range 4000 m, 0.5 m bins.

    RawData is a sparse file: zeros, plus the echoes of ``ntarget`` point targets (as in
    :func:`_simulate_raw`, without noise) and complex Gaussian noise over the apertures of the DEM
    columns in ``noise_columns`` (so that those columns focus to non-zero values).  Returns the names,
    the trajectory, the scalars and the noise columns."""
    rng = np.random.default_rng(seed)
    os.makedirs(data_dir, exist_ok=True)
    ntraj = naz + 1 if ntraj is None else ntraj
    near, dra, lam = 4000.0, 0.5, 0.031
    tx = np.arange(ntraj) * 0.1 + rng.normal(0, 0.02, ntraj)
    ty = np.full(ntraj, -3000.0) + rng.normal(0, 0.02, ntraj)
    tz = np.full(ntraj, 3000.0) + rng.normal(0, 0.02, ntraj)
    from . import formats
    names = formats.grid_names(dim_rg, dim_az, underscore=True)
    names.update(formats.traj_names(ntraj))
    names["raw"] = formats.raw_name(nra, naz)
    for k, t in (("tx", tx), ("ty", ty), ("tz", tz)):
        t.astype(np.float64).tofile(os.path.join(data_dir, names[k]))
    for f, v in zip(formats.SCALAR_FILES, (near, dra, lam)):
        np.array([v], np.float64).tofile(os.path.join(data_dir, f))
    dx = 0.1 * naz / dim_az
    ys = np.arange(dim_rg) * 1.0
    files = {k: open(os.path.join(data_dir, names[k]), "wb") for k in ("x", "y", "z", "index", "mask", "sint")}
    try:
        for a0 in range(0, dim_az, 2000):  # F order: whole columns, one block at a time
            xs = np.arange(a0, min(dim_az, a0 + 2000)) * dx
            X, Y = np.meshgrid(xs, ys, indexing="xy")  # (dim_rg, cols): rows across track, columns along it
            Z = 20.0 + 10.0 * np.sin(X / 500.0) * np.cos(Y / 300.0)
            idx = np.clip(np.rint(X / 0.1), 0, naz - 1).astype(np.int32)
            sint = rng.integers(sint_range[0], sint_range[1] + 1, X.shape).astype(np.int32)
            mask = (rng.random(X.shape) < 0.95).astype(np.uint8)
            for k, a in (("x", X), ("y", Y), ("z", Z), ("index", idx), ("mask", mask), ("sint", sint)):
                np.asarray(a).flatten(order="F").tofile(files[k])
    finally:
        for f in files.values():
            f.close()
    raw_path = os.path.join(data_dir, names["raw"])
    with open(raw_path, "wb") as f:
        f.truncate(nra * naz * 8)  # a sparse file of zeros
    raw = np.memmap(raw_path, dtype=np.complex64, mode="r+", shape=(nra, naz), order="F")
    sigma = 1.0
    for j in noise_columns:
        c = int(round(j * dx / 0.1))
        a0, a1 = max(0, c - sint_range[1] // 2 - 100), min(naz, c + sint_range[1] // 2 + 100)
        for b0 in range(a0, a1, 1024):
            b1 = min(a1, b0 + 1024)
            n = rng.standard_normal((2, nra, b1 - b0), dtype=np.float32) * np.float32(sigma)
            raw[:, b0:b1] = n[0] + 1j * n[1]
    tgx = rng.uniform(0, (dim_az - 1) * dx, ntarget)
    tgy = rng.uniform(0, ys[-1], ntarget)
    tgz = 20.0 + 10.0 * np.sin(tgx / 500.0) * np.cos(tgy / 300.0)
    amp = np.abs(rng.normal(0, 1, ntarget)) + 1.0
    offs = np.arange(-SINC_HALF, SINC_HALF + 1)
    tanb = math.tan(math.radians(BEAM_DEG))
    for (px, py, pz, a) in zip(tgx, tgy, tgz, amp):
        half = tanb * math.hypot(py + 3000.0, pz - 3000.0)
        sel = np.nonzero(np.abs(tx[:naz] - px) <= half)[0]
        R = np.sqrt((px - tx[sel]) ** 2 + (py - ty[sel]) ** 2 + (pz - tz[sel]) ** 2)
        rb = np.rint((R - near) / dra).astype(np.int64)[:, None] + offs[None, :]
        val = a * np.sinc((near + rb * dra - R[:, None]) / dra) * np.exp(-1j * 4.0 * math.pi * R / lam)[:, None]
        ok = (rb >= 0) & (rb < nra)
        # only the touched samples are written: the file stays sparse outside the echoes
        np.add.at(raw, (rb[ok], np.broadcast_to(sel[:, None], rb.shape)[ok]), val[ok].astype(np.complex64))
    raw.flush()
    del raw
    return {"names": names, "tx": tx, "ty": ty, "tz": tz, "near": near, "dra": dra, "lam": lam,
            "noise_columns": tuple(noise_columns), "dim_rg": dim_rg, "dim_az": dim_az, "nra": nra, "naz": naz,
            "ntraj": ntraj}
