"""Reference file formats and checkpoints (CPU only)."""
import os

import numpy as np
import pytest

from sarbp import formats, synth


def test_scene_roundtrip_both_namings(tmp_path):
    sc = synth.tiny_scene(seed=1)
    for underscore, ntraj in ((False, None), (True, sc.Naz + 1)):
        d = tmp_path / f"u{int(underscore)}"
        names = formats.write_scene(sc, str(d), underscore=underscore, ntraj=ntraj)
        assert ("Dem_x_Dbl_" in names["x"]) == underscore
        args = formats.load_inputs(str(d), sc.shape[0], sc.shape[1], sc.Nra, sc.Naz, underscore=underscore,
                                   ntraj=ntraj)
        x, y, z, mask, index, tx, ty, tz, raw, near, dra, lam, sint = args
        assert np.array_equal(x, sc.x) and np.array_equal(z, sc.z) and np.array_equal(sint, sc.sint)
        assert np.array_equal(mask, sc.mask) and np.array_equal(index, sc.index)
        assert np.array_equal(raw, sc.raw) and raw.dtype == np.complex64
        assert np.array_equal(tx[: sc.Naz], sc.tx) and tx.size == (ntraj or sc.Naz)
        assert (near, dra, lam) == (sc.near, sc.dra, sc.lam)


def test_raw_is_interleaved_column_major(tmp_path):
    raw = (np.arange(12, dtype=np.float32) + 1j * -np.arange(12, dtype=np.float32)).astype(np.complex64).reshape(
        (3, 4), order="F")
    p = tmp_path / "RawData_Cmplx_3x4.dat"
    formats.write_dat(str(p), raw, np.complex64)
    f = np.fromfile(p, np.float32)
    # (r, a) = (1, 0) is element 1: floats 2, 3 are its re, im (sar_focuser_puntual.py:487-488)
    assert f[2] == raw[1, 0].real and f[3] == raw[1, 0].imag
    assert np.array_equal(formats.load_raw(str(p), 3, 4), raw)
    assert np.array_equal(np.asarray(formats.load_raw(str(p), 3, 4, mmap=True)), raw)


def test_validate_raw_dims_reconciles(tmp_path):
    p = tmp_path / "r.dat"
    np.zeros(6 * 10, np.complex64).tofile(p)
    assert formats.validate_raw_dims(str(p), 6, 10) == (6, 10)
    assert formats.validate_raw_dims(str(p), 6, 11) == (6, 10)    # Naz from the size
    assert formats.validate_raw_dims(str(p), 7, 10) == (6, 10)    # Nra from the size
    with pytest.raises(ValueError):
        formats.validate_raw_dims(str(p), 7, 11)
    with pytest.raises(FileNotFoundError):
        formats.validate_raw_dims(str(tmp_path / "missing.dat"), 6, 10)
    with open(p, "ab") as f:
        f.write(b"\0\0\0\0")
    with pytest.raises(ValueError):
        formats.validate_raw_dims(str(p), 6, 10)


def test_tolerant_loader(tmp_path):
    raw = (np.random.default_rng(0).normal(size=(4, 6)) + 1j).astype(np.complex64)
    p = tmp_path / "r.dat"
    formats.write_dat(str(p), raw, np.complex64)
    with open(p, "ab") as f:  # odd trailing float is dropped (gpu_small_dataset.py:384-391)
        f.write(np.float32(7).tobytes())
    got, nra, naz = formats.load_raw_tolerant(str(p), 4, 999)
    assert (nra, naz) == (4, 6) and np.array_equal(got, raw)
    got, nra, naz = formats.load_raw_tolerant(str(p), 5, 999)  # 24 = largest factor <= sqrt(24) -> 4 x 6
    assert (nra, naz) == (4, 6)


def test_output_writers_roundtrip(tmp_path):
    rng = np.random.default_rng(1)
    slc = np.asfortranarray(rng.normal(size=(7, 5)) + 1j * rng.normal(size=(7, 5)))
    base = str(tmp_path / "out")
    files = formats.write_outputs(slc, base)
    assert len(files) == 3
    assert np.array_equal(formats.read_split_f32(base + "_SplitF32.dat", slc.shape), slc.astype(np.complex64))
    assert np.array_equal(formats.read_interleaved_f32(base + "_InterleavedF32.dat", slc.shape),
                          slc.astype(np.complex64))
    assert np.array_equal(np.load(base + ".npy"), slc)
    formats.write_slcmatr(slc, str(tmp_path))
    assert np.array_equal(formats.read_slcmatr(str(tmp_path), slc.shape), slc)
    # the reference viewer reads SlcMatr_real as float64 (250, 4000) C-order == F-order transposed
    re = np.fromfile(tmp_path / "SlcMatr_real.dat", np.float64).reshape(slc.shape[::-1])
    assert np.array_equal(re.T, slc.real)


def test_px_checkpoint(tmp_path):
    p = str(tmp_path / "ck.npz")
    v = np.arange(10) * (1 + 1j)
    formats.save_px_checkpoint(p, v, 4)
    fs, nxt = formats.load_px_checkpoint(p, 10)
    assert nxt == 4 and np.array_equal(fs, v)
    assert formats.load_px_checkpoint(p, 11) is None        # size mismatch: fresh start
    open(p, "wb").write(b"garbage")
    assert formats.load_px_checkpoint(p, 10) is None        # unreadable: fresh start
    assert formats.load_px_checkpoint(str(tmp_path / "none.npz"), 10) is None
