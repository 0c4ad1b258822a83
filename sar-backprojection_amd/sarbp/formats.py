"""IDL-style column-major ``.dat`` inputs and the reference's output formats.

Inputs (README.md:53-87; readers sar_focuser_puntual.py:420-512, main_gpu.py:303-333, 372-385):

=================================  ===========  =====================================
file                               dtype        shape (Fortran order)
=================================  ===========  =====================================
Dem_{x,y,z}_Dbl<R>x<A>.dat         float64      (DimRg, DimAz)
Traiett_{x,y,z}_Dbl<Naz>.dat       float64      (Naz,)
Index_Long_<R>x<A>.dat             int32        (DimRg, DimAz)
Mask_Byte_<R>x<A>.dat              uint8        (DimRg, DimAz)
Sint_Long_<R>x<A>.dat              int32        (DimRg, DimAz)
RawData_Cmplx_<Nra>x<Naz>.dat      complex64    (Nra, Naz), interleaved (re, im) float32
NearRangeDbl_1_elemento.dat        float64      (1,)   (also DeltaRangeDbl_, LambdaDbl_)
=================================  ===========  =====================================

Two naming schemes exist: ``Dem_x_Dbl4000x250.dat`` (sar_focuser_puntual.py:392) and
``Dem_x_Dbl_4000x32000.dat`` with a ``Traiett_x_Dbl403485.dat`` one sample longer than the raw
azimuth (main_gpu.py:304-333).

Outputs: ``<base>_SplitF32.dat`` (all real parts then all imaginary parts, float32, F order),
``<base>_InterleavedF32.dat`` and ``<base>.npy`` (complex128) -- sar_focuser_puntual.py:548-590,
main_gpu.py:745-797; ``SlcMatr_real.dat`` / ``SlcMatr_imag.dat`` (float64, F order) and
``SlcMatr_complex.dat`` (README.md:99-105; read back by img_visualizer.py:19-21).
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

SCALAR_FILES = ("NearRangeDbl_1_elemento.dat", "DeltaRangeDbl_1_elemento.dat", "LambdaDbl_1_elemento.dat")


def grid_names(dimrg: int, dimaz: int, underscore: bool = False) -> dict:
    sep = "_" if underscore else ""
    g = f"{dimrg}x{dimaz}"
    return {
        "x": f"Dem_x_Dbl{sep}{g}.dat", "y": f"Dem_y_Dbl{sep}{g}.dat", "z": f"Dem_z_Dbl{sep}{g}.dat",
        "index": f"Index_Long_{g}.dat", "mask": f"Mask_Byte_{g}.dat", "sint": f"Sint_Long_{g}.dat",
    }


def traj_names(ntraj: int) -> dict:
    return {k: f"Traiett_{k[1]}_Dbl{ntraj}.dat" for k in ("tx", "ty", "tz")}


def raw_name(nra: int, naz: int) -> str:
    return f"RawData_Cmplx_{nra}x{naz}.dat"


def read_dat(path: str, dtype, shape=None, mmap: bool = False):
    """Read a headerless binary file; ``shape`` is interpreted in Fortran order."""
    dtype = np.dtype(dtype)
    if mmap:
        if shape is None:
            shape = (os.path.getsize(path) // dtype.itemsize,)
        return np.memmap(path, dtype=dtype, mode="r", shape=tuple(shape), order="F")
    a = np.fromfile(path, dtype=dtype)
    if shape is not None:
        need = int(np.prod(shape))
        if a.size != need:
            raise ValueError(f"{os.path.basename(path)}: {a.size} elements, expected {need} for shape {shape}")
        a = a.reshape(shape, order="F")
    return a


def write_dat(path: str, a, dtype=None) -> None:
    a = np.asarray(a)
    if dtype is not None:
        a = a.astype(dtype)
    a.flatten(order="F").tofile(path)


def validate_raw_dims(path: str, nra: int, naz: int) -> tuple[int, int]:
    """Size check of an interleaved complex64 RawData file, reconciling one dimension from the file
    size when it differs (main_gpu.py:208-248).  Raises FileNotFoundError / ValueError."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"Raw data file not found: {path}")
    size = os.path.getsize(path)
    if size % 8:
        raise ValueError(f"RawData file size ({size}) is not a multiple of complex64 size (8 bytes)")
    total = size // 8
    if total == nra * naz:
        return nra, naz
    if nra > 0 and total % nra == 0:
        return nra, total // nra
    if naz > 0 and total % naz == 0:
        return total // naz, naz
    raise ValueError(f"Could not reconcile file size ({total} complex elements) with {nra}x{naz}")


def load_raw(path: str, nra: int, naz: int, mmap: bool = False) -> np.ndarray:
    """Interleaved complex64 RawData, (nra, naz) in Fortran order (sar_focuser_puntual.py:480-491)."""
    return read_dat(path, np.complex64, (nra, naz), mmap=mmap)


def load_raw_tolerant(path: str, expected_nra: int, expected_naz: int):
    """The tolerant loader of gpu_small_dataset.py:368-451: drops a trailing odd float, derives
    Naz from the file size, else searches the largest factor <= sqrt(total) as Nra, else truncates.
    Returns (raw complex64 (Nra, Naz), Nra, Naz)."""
    total_floats = os.path.getsize(path) // 4
    data = np.fromfile(path, dtype=np.float32)
    if total_floats % 2:
        data = data[:-1]
    total = data.size // 2
    nra = expected_nra
    if total % nra == 0:
        naz = total // nra
    else:
        f = int(math.sqrt(total))
        while f > 1 and total % f:
            f -= 1
        if f > 1:
            nra, naz = f, total // f
        else:
            naz = total // nra
            data = data[: 2 * nra * naz]
    raw = (data[0::2].reshape((nra, naz), order="F") + 1j * data[1::2].reshape((nra, naz), order="F"))
    return raw.astype(np.complex64), nra, naz


def read_scalars(data_dir: str) -> tuple[float, float, float]:
    vals = [float(np.fromfile(os.path.join(data_dir, f), dtype=np.float64)[0]) for f in SCALAR_FILES]
    return vals[0], vals[1], vals[2]


def write_scene(scene, data_dir: str, underscore: bool = False, ntraj: int | None = None) -> dict:
    """Write a :class:`sarbp.synth.Scene` as the 13 reference input files; returns the names."""
    os.makedirs(data_dir, exist_ok=True)
    dimrg, dimaz = scene.x.shape
    names = grid_names(dimrg, dimaz, underscore)
    write_dat(os.path.join(data_dir, names["x"]), scene.x, np.float64)
    write_dat(os.path.join(data_dir, names["y"]), scene.y, np.float64)
    write_dat(os.path.join(data_dir, names["z"]), scene.z, np.float64)
    write_dat(os.path.join(data_dir, names["index"]), scene.index, np.int32)
    write_dat(os.path.join(data_dir, names["mask"]), scene.mask, np.uint8)
    write_dat(os.path.join(data_dir, names["sint"]), scene.sint, np.int32)
    nt = scene.tx.size if ntraj is None else ntraj
    tn = traj_names(nt)
    for k in ("tx", "ty", "tz"):
        t = getattr(scene, k)
        if nt > t.size:  # the large dataset ships a trajectory one sample longer (main_gpu.py:331)
            t = np.concatenate([t, np.full(nt - t.size, t[-1])])
        np.asarray(t[:nt], np.float64).tofile(os.path.join(data_dir, tn[k]))
    names.update(tn)
    rn = raw_name(scene.raw.shape[0], scene.raw.shape[1])
    np.asfortranarray(scene.raw, dtype=np.complex64).flatten(order="F").tofile(os.path.join(data_dir, rn))
    names["raw"] = rn
    for f, v in zip(SCALAR_FILES, (scene.near, scene.dra, scene.lam)):
        np.array([v], np.float64).tofile(os.path.join(data_dir, f))
    return names


def load_inputs(data_dir: str, dimrg: int, dimaz: int, nra: int, naz: int, underscore: bool = False,
                ntraj: int | None = None, mmap_raw: bool = False):
    """Read the 13 input files; returns the FocalizzatoreBpPx argument tuple
    (xxg1, yyg1, zzg1, mask, index, wgsx, wgsy, wgsz, Raw_Data, near, dra, lambda, Sint)."""
    g = grid_names(dimrg, dimaz, underscore)
    shp = (dimrg, dimaz)
    p = lambda f: os.path.join(data_dir, f)
    missing = [f for f in list(g.values()) + [raw_name(nra, naz)] + list(SCALAR_FILES) if not os.path.exists(p(f))]
    tn = traj_names(naz if ntraj is None else ntraj)
    missing += [f for f in tn.values() if not os.path.exists(p(f))]
    if missing:
        raise FileNotFoundError("missing input files: " + ", ".join(missing))
    x = read_dat(p(g["x"]), np.float64, shp)
    y = read_dat(p(g["y"]), np.float64, shp)
    z = read_dat(p(g["z"]), np.float64, shp)
    index = read_dat(p(g["index"]), np.int32, shp)
    mask = read_dat(p(g["mask"]), np.uint8, shp)
    sint = read_dat(p(g["sint"]), np.int32, shp)
    tx, ty, tz = (np.fromfile(p(tn[k]), dtype=np.float64) for k in ("tx", "ty", "tz"))
    raw = load_raw(p(raw_name(nra, naz)), nra, naz, mmap=mmap_raw)
    near, dra, lam = read_scalars(data_dir)
    return x, y, z, mask, index, tx, ty, tz, raw, near, dra, lam, sint


# ----------------------------------------------------------------------------- outputs

def write_outputs(slc, base: str, split=True, interleaved=True, numpy=True) -> list[str]:
    """SplitF32 / InterleavedF32 / .npy writers (sar_focuser_puntual.py:548-590)."""
    slc = np.asarray(slc)
    written = []
    re = slc.real.astype(np.float32).flatten(order="F")
    im = slc.imag.astype(np.float32).flatten(order="F")
    if split:
        fn = base + "_SplitF32.dat"
        with open(fn, "wb") as f:
            re.tofile(f)
            im.tofile(f)
        written.append(fn)
    if interleaved:
        fn = base + "_InterleavedF32.dat"
        il = np.empty(re.size * 2, np.float32)
        il[0::2] = re
        il[1::2] = im
        il.tofile(fn)
        written.append(fn)
    if numpy:
        fn = base + ".npy"
        np.save(fn, slc.astype(np.complex128))
        written.append(fn)
    return written


def write_slcmatr(slc, out_dir: str = ".", prefix: str = "SlcMatr") -> list[str]:
    """README.md:99-105: <prefix>_real.dat / _imag.dat (float64, F order) and _complex.dat
    (complex128, interleaved, F order)."""
    slc = np.asarray(slc, np.complex128)
    names = [os.path.join(out_dir, f"{prefix}_{k}.dat") for k in ("real", "imag", "complex")]
    write_dat(names[0], slc.real, np.float64)
    write_dat(names[1], slc.imag, np.float64)
    write_dat(names[2], slc, np.complex128)
    return names


def read_slcmatr(out_dir: str, shape, prefix: str = "SlcMatr"):
    re = read_dat(os.path.join(out_dir, f"{prefix}_real.dat"), np.float64, shape)
    im = read_dat(os.path.join(out_dir, f"{prefix}_imag.dat"), np.float64, shape)
    return re + 1j * im


def read_split_f32(path: str, shape):
    a = np.fromfile(path, dtype=np.float32)
    h = a.size // 2
    return (a[:h].reshape(shape, order="F") + 1j * a[h:].reshape(shape, order="F")).astype(np.complex64)


def read_interleaved_f32(path: str, shape):
    a = np.fromfile(path, dtype=np.float32)
    return (a[0::2].reshape(shape, order="F") + 1j * a[1::2].reshape(shape, order="F")).astype(np.complex64)


# ----------------------------------------------------------------------------- checkpoints

def save_px_checkpoint(path: str, focused_sample, next_idx: int) -> None:
    """``{focused_sample: complex128[dim], next_idx: int64}`` (sar_focuser_puntual.py:290-292)."""
    tmp = path + ".tmp.npz"
    np.savez(tmp, focused_sample=np.asarray(focused_sample, np.complex128), next_idx=np.int64(next_idx))
    os.replace(tmp, path)


def load_px_checkpoint(path: str, dim: int):
    """Returns (focused_sample, next_idx) or None when absent, unreadable or shape-mismatched
    (sar_focuser_puntual.py:218-234: those cases start from zero)."""
    if not os.path.exists(path):
        return None
    try:
        with np.load(path, allow_pickle=False) as ck:
            fs = ck["focused_sample"]
            if fs.shape != (dim,):
                return None
            nxt = int(ck["next_idx"])
            if not (0 <= nxt <= dim):
                return None
            return np.array(fs, np.complex128), nxt
    except Exception:
        return None


def save_json(path: str, obj) -> None:
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(obj, f, indent=4)
    os.replace(tmp, path)
