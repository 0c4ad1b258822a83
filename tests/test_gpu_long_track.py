"""The long-track dataset of main_gpu.py at its size through SARSuperOptimizedFocuser (GPU).

RawData 4762 range bins x 403484 pulses (15.4 GB, main_gpu.py:321-333) in the reference's file
names, with trajectory files one sample longer (403485: the focuser reconciles the lengths,
main_gpu.py:387-398), focused over a 1000 x 32000 DEM in 16 batches of 2000 columns
(main_gpu.py:415) after a streamed upload of RawData from the memory-mapped file (pulse blocks,
64-bit column addressing up to 15 GB).  Checks: the contributor counts of every batch add up to the
oracle's pixel-pulse pair count (every slant range lies in the swath); RawData x 2, uploaded again
through the streamed path, gives exactly twice the stored image over the last batch (whose pulses
lie at the end of the 15 GB); and DEM columns near the start, the middle and the end of the track
against the oracle (<= 1e-10 relative, exact zeros without contributors).  The files are written
sparse (zeros outside the targets' echoes and the noise over the checked columns' apertures).
"""
import os
import shutil

import numpy as np
import pytest

from oracle import oracle_lib as O
from sarbp import formats, synth
from sarbp.core import Dataset

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

DIM_RG, DIM_AZ, NRA, NAZ = 1000, 32000, 4762, 403484


def test_long_track_process_dataset(gpu, tmp_path):
    from sarbp.main_gpu import SARSuperOptimizedFocuser
    free = shutil.disk_usage(tmp_path).free
    if free < 24e9:
        pytest.skip(f"needs ~6 GB of disk for the sparse 15 GB RawData file, the DEM and the output "
                    f"(24 GB free required in case the file system does not keep files sparse); {free / 1e9:.1f} GB free")
    d = tmp_path / "data"
    print("case long track: writing 4762 x 403484 RawData (sparse) and a 1000 x 32000 DEM", flush=True)
    meta = synth.write_long_track(str(d), DIM_RG, DIM_AZ, NAZ, NRA)
    os.chdir(tmp_path)
    try:
        f = SARSuperOptimizedFocuser(output_file="long.h5", checkpoint_base_name="long")
        f.config.update(dim_rg=DIM_RG, dim_az=DIM_AZ, nra=NRA, naz=NAZ, traj_len=NAZ + 1, underscore_names=True)
        print("case long track: process_dataset (streamed upload + 16 batches)", flush=True)
        f.process_dataset(str(d))
        slc = f._read_slc()
        assert slc.shape == (DIM_RG, DIM_AZ)
        ds = f._dataset
        assert (ds.nra, ds.naz) == (NRA, NAZ)  # the 403485-sample trajectory cut to the RawData length
        names = meta["names"]
        grid = {k: formats.read_dat(str(d / names[k]), dt, (DIM_RG, DIM_AZ), mmap=True)
                for k, dt in (("x", np.float64), ("y", np.float64), ("z", np.float64), ("index", np.int32),
                              ("mask", np.uint8), ("sint", np.int32))}

        def flat(a0, a1):
            return [np.ascontiguousarray(np.asarray(grid[k][:, a0:a1]).ravel(order="F"))
                    for k in ("x", "y", "z", "mask", "index", "sint")]

        # every pixel-pulse pair contributes: the counts of all batches add up to the oracle's count
        total = 0
        for a0 in range(0, DIM_AZ, 2000):
            fl = flat(a0, a0 + 2000)
            out, cnt = ds.focus_flat(*fl, with_counts=True, dimx=DIM_RG)
            assert np.array_equal(out, slc[:, a0:a0 + 2000].ravel(order="F")), a0  # the stored batch
            total += int(cnt.sum())
        pp = O.count_pp(np.asarray(grid["mask"]).ravel(order="F"), np.asarray(grid["index"]).ravel(order="F"),
                        np.asarray(grid["sint"]).ravel(order="F"), NAZ)
        print("pixel-pulse pairs", pp, "contributors", total, flush=True)
        assert total == pp

        # RawData x 2 through the streamed upload: twice the last batch, bit for bit
        raw = formats.load_raw(str(d / names["raw"]), NRA, NAZ, mmap=True)
        tr = {k: np.fromfile(str(d / names[k]), np.float64)[:NAZ] for k in ("tx", "ty", "tz")}
        print("case long track: RawData x 2 upload", flush=True)
        ds2 = Dataset.streamed(lambda a0, a1: np.asfortranarray(raw[:, a0:a1]) * np.complex64(2), NRA, NAZ,
                               tr["tx"], tr["ty"], tr["tz"], meta["near"], meta["dra"], meta["lam"])
        a0 = DIM_AZ - 2000
        got = ds2.focus_flat(*flat(a0, DIM_AZ), dimx=DIM_RG)
        assert np.array_equal(got, 2 * slc[:, a0:].ravel(order="F"))
        ds2.close()

        # DEM columns near the start, the middle and the end of the track against the oracle
        for j in meta["noise_columns"]:
            fl = flat(j, j + 1)
            ref, rcnt = O.focus_flat(*fl, tr["tx"], tr["ty"], tr["tz"], raw, meta["near"], meta["dra"], meta["lam"])
            got = slc[:, j]
            nz = rcnt > 0
            rel = np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])
            print("column", j, "pixels", int(nz.sum()), "max rel err vs oracle", float(rel.max()), flush=True)
            assert nz.sum() > 900 and rel.max() <= 1e-10 and np.all(got[~nz] == 0)
        f.close()
    finally:
        shutil.rmtree(str(d), ignore_errors=True)
        for p in ("long.npy", "long.npy.attrs.json"):
            if os.path.exists(p):
                os.remove(p)
