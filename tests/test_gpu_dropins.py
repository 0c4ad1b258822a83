"""Reference-compatible drivers on the GPU: gpu_small_dataset.py and main_gpu.py entry points,
against the oracle and against each other; the drop-ins read the caller's arrays as they are at
each call; process_point's ind_line_ra_min."""
import os

import numpy as np
import pytest

from oracle import oracle_lib as O
from sarbp import formats, synth


def _against_oracle(got, sc, what, raw=None):
    """A FocalizzatoreBpPx-shaped result against the oracle: exact zeros where no pulse counts,
    <= 1e-10 relative elsewhere."""
    ref, rcnt = O.focus(sc.x, sc.y, sc.z, sc.mask, sc.index, sc.tx, sc.ty, sc.tz, sc.raw if raw is None else raw,
                        sc.near, sc.dra, sc.lam, sc.sint)
    nz = rcnt > 0
    assert np.all(got[~nz] == 0), what
    rel = np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])
    print(what, "pixels", int(nz.sum()), "max rel err vs oracle", float(rel.max()), flush=True)
    assert rel.max() <= 1e-10, what

pytestmark = pytest.mark.gpu


def test_cupy_entry_points_equal_pixel_focuser(gpu, tmp_path):
    from sarbp.gpu_small_dataset import FocalizzatoreBpCuPy, FocalizzatoreBpCuPy_Batched, load_raw_data
    from sarbp.sar_focuser_puntual import FocalizzatoreBpPx
    sc = synth.tiny_scene(seed=4)
    os.chdir(tmp_path)
    ref = FocalizzatoreBpPx(*sc.args())
    a = FocalizzatoreBpCuPy(*sc.args())
    b = FocalizzatoreBpCuPy_Batched(*sc.args(), batch_size=50)
    assert np.array_equal(a, ref) and np.array_equal(b, ref)
    for got, what in ((ref, "FocalizzatoreBpPx"), (a, "FocalizzatoreBpCuPy"), (b, "FocalizzatoreBpCuPy_Batched")):
        _against_oracle(got, sc, what)
    names = formats.write_scene(sc, str(tmp_path / "d"))
    raw, nra, naz = load_raw_data(str(tmp_path / "d"), names["raw"], sc.Nra, sc.Naz)
    assert (nra, naz) == (sc.Nra, sc.Naz) and np.array_equal(raw, sc.raw)
    with pytest.raises(ValueError):
        FocalizzatoreBpCuPy(sc.x, sc.y, sc.z, sc.mask, sc.index, sc.tx, sc.ty, sc.tz, sc.raw.ravel(), sc.near,
                            sc.dra, sc.lam, sc.sint)


def _configure(f, sc):
    f.config.update(dim_rg=sc.shape[0], dim_az=sc.shape[1], nra=sc.Nra, naz=sc.Naz, traj_len=sc.Naz + 1,
                    batch_size_azimuth=5, azimuth_overlap=2)


def test_super_optimized_focuser_resume_and_export(gpu, tmp_path):
    from sarbp.main_gpu import SARSuperOptimizedFocuser
    from sarbp.sar_focuser_puntual import FocalizzatoreBpPx
    sc = synth.tiny_scene(seed=5, nx=24, ny=17)
    d = tmp_path / "data"
    formats.write_scene(sc, str(d), underscore=True, ntraj=sc.Naz + 1)
    os.chdir(tmp_path)
    ref = FocalizzatoreBpPx(*sc.args())

    f = SARSuperOptimizedFocuser()
    _configure(f, sc)
    f.process_dataset(str(d))
    full = f._read_slc()
    assert np.array_equal(full, ref)
    _against_oracle(full, sc, "process_dataset")
    files = f.export_to_formats()
    assert np.array_equal(np.load("SlcCuPy_SuperOptimized.npy"), ref)
    assert np.array_equal(formats.read_split_f32("SlcCuPy_SuperOptimized_SplitF32.dat", sc.shape),
                          ref.astype(np.complex64))
    assert len(files) == 3
    assert f.generate_preview() is not None
    f.close()

    # interrupted after two batches, then resumed from the JSON checkpoint
    os.remove(f.checkpoint_file)
    g = SARSuperOptimizedFocuser(output_file="again.h5", checkpoint_base_name="again")
    _configure(g, sc)
    real_save = g.save_checkpoint

    def stop_after_two(a0, a1):
        real_save(a0, a1)
        if g.processed_batches == 2:
            raise KeyboardInterrupt

    g.save_checkpoint = stop_after_two
    with pytest.raises(KeyboardInterrupt):
        g.process_dataset(str(d))
    assert os.path.exists(g.checkpoint_file)
    g.close()
    h = SARSuperOptimizedFocuser(output_file="again.h5", checkpoint_base_name="again")
    _configure(h, sc)
    h.process_dataset(str(d))
    assert np.array_equal(h._read_slc(), ref)
    _against_oracle(h._read_slc(), sc, "process_dataset resumed")
    h.close()


def test_dropins_read_arrays_fresh_each_call(gpu, tmp_path):
    """A caller may change RawData or the trajectory in place between calls (the reference reads
    them at every call): the next call focuses the new contents."""
    from sarbp.gpu_small_dataset import FocalizzatoreBpCuPy
    from sarbp.sar_focuser_puntual import FocalizzatoreBpPx, process_point
    sc = synth.tiny_scene(seed=8)
    os.chdir(tmp_path)
    raw = sc.raw.copy(order="F")
    args = list(sc.args())
    args[8] = raw
    first = FocalizzatoreBpPx(*args)
    f = [np.asarray(v).flatten(order="F") for v in (sc.x, sc.y, sc.z, sc.mask, sc.index, sc.sint)]
    i = int(np.argmax(np.abs(first).flatten(order="F")))
    pargs = lambda: (f[0][i], f[1][i], f[2][i], f[3][i], f[4][i], f[5][i], sc.tx, sc.ty, sc.tz, raw, sc.near,  # noqa
                     sc.dra, sc.lam, sc.Nra, sc.Naz)
    p1 = process_point(*pargs())
    assert p1 == first.flatten(order="F")[i]
    raw *= np.complex64(1j)  # in place: same buffer, new contents
    for fn, what in ((FocalizzatoreBpPx, "FocalizzatoreBpPx"), (FocalizzatoreBpCuPy, "FocalizzatoreBpCuPy")):
        got = fn(*args)
        assert np.array_equal(got, 1j * first), what
        _against_oracle(got, sc, what + " after an in-place change", raw=raw)
    assert process_point(*pargs()) == 1j * p1
    sc.tx[:] += 0.5  # the trajectory too
    moved = FocalizzatoreBpPx(*args)
    _against_oracle(moved, sc, "FocalizzatoreBpPx after a trajectory change", raw=raw)
    o, _ = O.point(*pargs()[:13])
    assert abs(process_point(*pargs()) - o) <= 1e-10 * abs(o)


def test_process_point_ind_line_ra_min(gpu):
    """process_point's ind_line_ra_min (sar_focuser_puntual.py:27, 106) against the oracle."""
    from sarbp.sar_focuser_puntual import process_point
    sc = synth.tiny_scene(seed=9, nx=6, ny=4)
    f = [np.asarray(v).flatten(order="F") for v in (sc.x, sc.y, sc.z, sc.mask, sc.index, sc.sint)]
    for i in range(f[0].size):
        t = O.trace(float(f[0][i]), float(f[1][i]), float(f[2][i]), int(f[4][i]), int(f[5][i]), sc.tx, sc.ty,
                    sc.tz, sc.Nra, sc.Naz, sc.near, sc.dra)
        bins = t[1]
        for lo in sorted({-5, 0, 3} | ({int(np.median(bins))} if len(bins) else set())):
            args = [v[i] for v in f] + [sc.tx, sc.ty, sc.tz, sc.raw, sc.near, sc.dra, sc.lam]
            o, n = O.point(*args, ind_line_ra_min=lo)
            got = process_point(*args, sc.Nra, sc.Naz, ind_line_ra_min=lo, index_min=7)
            assert (got == 0) if n == 0 else abs(got - o) <= 1e-10 * abs(o), (i, lo, got, o, n)
