"""Reference-compatible drivers on the GPU: gpu_small_dataset.py and main_gpu.py entry points."""
import os

import numpy as np
import pytest

from sarbp import formats, synth

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
    h.close()
