"""Drop-in for ``SARSuperOptimizedFocuser`` of main_gpu.py, computed on the MI355X.

Same class name, constructor arguments, ``config`` keys, checkpoint file and methods
(main_gpu.py:160-829).  What changes is where the data lives: the reference re-stages a RawData
window (memmap slice -> float32 planes -> H2D) for every 2000-column DEM batch (main_gpu.py:465-515);
here RawData is uploaded to HBM once, in pulse blocks streamed from the memmap, and every batch
is one focusing launch against it.  Batching over DEM columns, the output store, the JSON
checkpoint and resume-at-``last_end - azimuth_overlap`` (main_gpu.py:404-421) are kept.

Output store: HDF5 dataset ``slc`` with the reference's attributes when h5py is importable
(main_gpu.py:339-357); otherwise a Fortran-ordered complex128 ``.npy`` memmap next to the
requested file name plus a ``.attrs.json`` side file holding the same attributes.

Semantics follow the Numba path (half-to-even bins, last bin excluded), not the CUDA kernel of
main_gpu.py:127-130, so a dataset focused here equals ``FocalizzatoreBpPx`` on the same inputs.
"""
from __future__ import annotations

import json
import os
import time
from datetime import datetime, timedelta

import numpy as np

from . import formats
from .core import Dataset

try:  # optional, as in the reference (main_gpu.py:7)
    import h5py  # type: ignore
except Exception:  # pragma: no cover - absent in this image
    h5py = None


class _NpyStore:
    """complex128 (dimrg, dimaz) F-order .npy memmap + JSON attributes, h5py-like subset."""

    def __init__(self, path, shape, create):
        self.path = path
        self.attr_path = path + ".attrs.json"
        if create or not os.path.exists(path):
            self.slc = np.lib.format.open_memmap(path, mode="w+", dtype=np.complex128, shape=shape,
                                                 fortran_order=True)
            self.attrs = {}
        else:
            self.slc = np.load(path, mmap_mode="r+", allow_pickle=False)
            if self.slc.shape != tuple(shape):
                raise ValueError(f"existing output {path} has shape {self.slc.shape}, expected {shape}")
            self.attrs = json.load(open(self.attr_path)) if os.path.exists(self.attr_path) else {}

    def flush(self):
        self.slc.flush()
        formats.save_json(self.attr_path, self.attrs)

    def close(self):
        self.flush()
        del self.slc


def preview_window(dim_az, preview_size=1000, processed_az=None):
    """Columns [start, end) of the preview (main_gpu.py:641-659): the whole image when it has at
    most ``preview_size`` columns; else ``preview_size`` columns centred on the processed part
    (``processed_azimuth`` columns done) when more than ``preview_size`` are processed, the
    processed columns when fewer, and the image centre when the attribute is absent."""
    if dim_az <= preview_size:
        return 0, dim_az
    if processed_az is not None:
        processed_az = int(processed_az)
        if processed_az > preview_size:
            start = max(0, min(processed_az, dim_az) // 2 - preview_size // 2)
            return start, min(dim_az, start + preview_size)
        return 0, min(dim_az, processed_az)
    start = max(0, dim_az // 2 - preview_size // 2)
    return start, min(dim_az, start + preview_size)


def preview_scale(amplitude):
    """Display range of the linear preview (main_gpu.py:677-690): vmax = min(15 x median, max)
    of the amplitudes, vmin = the 1st percentile of the non-zero amplitudes (0 when none)."""
    a = np.asarray(amplitude, np.float64).ravel()
    a = a[np.isfinite(a)]
    if a.size == 0:
        return 0.0, 0.0
    vmax = min(15.0 * float(np.median(a)), float(a.max()))
    nz = a[a > 0]
    vmin = float(np.percentile(nz, 1)) if nz.size else 0.0
    return vmin, vmax


def _store_path(output_file):
    root, ext = os.path.splitext(output_file)
    return output_file if (h5py is not None or ext == ".npy") else root + ".npy"


class SARSuperOptimizedFocuser:
    def __init__(self, output_file="SlcCuPy_SuperOptimized.h5",
                 checkpoint_base_name="SlcCuPy_SuperOptimized_checkpoint"):
        self.output_file = output_file
        self.checkpoint_dir = "checkpoints"
        self.checkpoint_file = os.path.join(self.checkpoint_dir, f"{checkpoint_base_name}.json")
        os.makedirs(self.checkpoint_dir, exist_ok=True)
        # main_gpu.py:167-174; raw_chunk_size / threads_per_block / max_gpu_memory are kept for
        # compatibility, the resident-RawData design does not need them.  The dataset shape keys
        # default to the reference's hardcoded massive dataset (main_gpu.py:304-333).
        self.config = {
            "batch_size_azimuth": 2000,
            "batch_size_points": 2000000,
            "raw_chunk_size": 500,
            "max_gpu_memory": 230.0,
            "threads_per_block": 64,
            "azimuth_overlap": 100,
            "dim_rg": 4000,
            "dim_az": 32000,
            "nra": 4762,
            "naz": 403484,
            "traj_len": 403485,
            "precision": "f64",
            "underscore_names": True,
        }
        self.processed_batches = 0
        self.start_time = None
        self.checkpoint_data = {}
        self.raw_data_mmap = None
        self._dataset = None
        self._dataset_key = None

    # -- checkpoint (main_gpu.py:250-289) ---------------------------------------------------
    def load_checkpoint(self):
        if os.path.exists(self.checkpoint_file):
            try:
                with open(self.checkpoint_file) as f:
                    self.checkpoint_data = json.load(f)
                self.processed_batches = self.checkpoint_data.get("processed_batches", 0)
                self.start_time = time.time() - self.checkpoint_data.get("total_elapsed_time_seconds", 0)
                return True
            except Exception as e:
                print(f"Error loading checkpoint from {self.checkpoint_file}: {e}")
                self.checkpoint_data = {}
        self.start_time = time.time()
        return False

    def save_checkpoint(self, current_az_start, current_az_end):
        if self.start_time is None:
            self.start_time = time.time()
        formats.save_json(self.checkpoint_file, {
            "last_processed_az_start": int(current_az_start),
            "last_processed_az_end": int(current_az_end),
            "processed_batches": self.processed_batches,
            "total_elapsed_time_seconds": time.time() - self.start_time,
            "timestamp": datetime.now().isoformat(),
        })

    # -- inputs ----------------------------------------------------------------------------
    def _upload_raw(self, raw_path, nra, naz, tx, ty, tz, near, dra, lam):
        key = (os.path.abspath(raw_path), os.path.getmtime(raw_path), nra, naz, near, dra, lam)
        if self._dataset is not None and self._dataset_key == key:
            return self._dataset
        mm = formats.load_raw(raw_path, nra, int(os.path.getsize(raw_path) // 8 // nra), mmap=True)
        self.raw_data_mmap = mm
        fill = lambda a0, a1: np.asfortranarray(mm[:, a0:a1])  # noqa: E731
        if self._dataset is not None:
            self._dataset.close()
        self._dataset = Dataset.streamed(fill, nra, naz, tx, ty, tz, near, dra, lam)
        self._dataset_key = key
        return self._dataset

    def process_dataset(self, data_dir="."):
        """Focus the whole DEM grid in azimuth batches (main_gpu.py:291-621); returns None."""
        cfg = self.config
        resumed = self.load_checkpoint()
        t_run = time.time()
        dimrg, dimaz = int(cfg["dim_rg"]), int(cfg["dim_az"])
        nra_exp, naz_exp = int(cfg["nra"]), int(cfg["naz"])
        near, dra, lam = formats.read_scalars(data_dir)
        tn = formats.traj_names(int(cfg["traj_len"]))
        tx, ty, tz = (np.fromfile(os.path.join(data_dir, tn[k]), dtype=np.float64) for k in ("tx", "ty", "tz"))
        raw_path = os.path.join(data_dir, formats.raw_name(nra_exp, naz_exp))
        nra, naz = formats.validate_raw_dims(raw_path, nra_exp, naz_exp)
        # trajectory vs raw azimuth length reconciliation (main_gpu.py:387-398)
        if tx.size > naz:
            tx, ty, tz = tx[:naz], ty[:naz], tz[:naz]
        elif tx.size < naz:
            naz = tx.size
        ds = self._upload_raw(raw_path, nra, naz, tx, ty, tz, near, dra, lam)

        names = formats.grid_names(dimrg, dimaz, bool(cfg["underscore_names"]))
        grid = {k: formats.read_dat(os.path.join(data_dir, names[k]), dt, (dimrg, dimaz), mmap=True)
                for k, dt in (("x", np.float64), ("y", np.float64), ("z", np.float64), ("index", np.int32),
                              ("mask", np.uint8), ("sint", np.int32))}

        store = self._open_store((dimrg, dimaz), resumed, near, dra, lam)
        bsz = int(cfg["batch_size_azimuth"])
        start_az = 0
        if resumed and "last_processed_az_end" in self.checkpoint_data:
            start_az = max(0, int(self.checkpoint_data["last_processed_az_end"]) - int(cfg["azimuth_overlap"]))
        nb = (dimaz + bsz - 1) // bsz
        try:
            for b in range(nb):
                a0, a1 = b * bsz, min((b + 1) * bsz, dimaz)
                if resumed and a1 <= start_az:
                    continue
                sub = {k: np.asfortranarray(v[:, a0:a1]) for k, v in grid.items()}
                f = {k: np.ascontiguousarray(v.flatten(order="F")) for k, v in sub.items()}
                f["mask"][f["sint"] == -9999] = 0  # main_gpu.py:436-439
                out = ds.focus_flat(f["x"], f["y"], f["z"], f["mask"], f["index"], f["sint"],
                                    cfg["precision"], dimx=dimrg)
                self._write(store, a0, a1, np.reshape(out, (dimrg, a1 - a0), order="F"))
                self.processed_batches += 1
                self.save_checkpoint(a0, a1)
                self._set_attrs(store, processed_azimuth=a1,
                                last_update=datetime.now().strftime("%Y-%m-%d %H:%M:%S"),
                                processed_percentage=float(a1) / dimaz * 100)
            self._set_attrs(store, completion_date=datetime.now().strftime("%Y-%m-%d %H:%M:%S"),
                            processing_time_seconds=time.time() - t_run)
        finally:
            self._close_store(store)
        print(f"Total processing time: {timedelta(seconds=int(time.time() - t_run))}")
        return None

    # -- output store ----------------------------------------------------------------------
    def _open_store(self, shape, resumed, near, dra, lam):
        path = _store_path(self.output_file)
        attrs = dict(creation_date=datetime.now().strftime("%Y-%m-%d %H:%M:%S"), dim_range=shape[0],
                     dim_azimuth=shape[1], near_range=near, delta_range=dra, **{"lambda": lam}, fortran_order=True)
        if h5py is not None:
            f = h5py.File(path, "a" if resumed else "w")
            if "slc" not in f:
                f.create_dataset("slc", shape=shape, dtype=np.complex128, chunks=(shape[0], min(1000, shape[1])),
                                 compression="gzip", compression_opts=4)
                f.attrs.update(attrs)
            return f
        st = _NpyStore(path, shape, create=not resumed)
        if not st.attrs:
            st.attrs.update(attrs)
        return st

    @staticmethod
    def _write(store, a0, a1, block):
        if h5py is not None and not isinstance(store, _NpyStore):
            store["slc"][:, a0:a1] = block
            store.flush()
        else:
            store.slc[:, a0:a1] = block
            store.flush()

    @staticmethod
    def _set_attrs(store, **kw):
        store.attrs.update(kw)
        if isinstance(store, _NpyStore):
            store.flush()

    @staticmethod
    def _close_store(store):
        store.close()

    def _read_slc(self):
        path = _store_path(self.output_file)
        if h5py is not None and not path.endswith(".npy"):
            with h5py.File(path, "r") as f:
                return f["slc"][:]
        return np.asfortranarray(np.load(path, allow_pickle=False))

    # -- exports (main_gpu.py:745-797) -------------------------------------------------------
    def export_to_formats(self, export_split=True, export_interleaved=True, export_numpy=True):
        slc = self._read_slc()
        base = "SlcCuPy_SuperOptimized"
        written = formats.write_outputs(slc, base, split=export_split, interleaved=export_interleaved,
                                        numpy=export_numpy)
        if export_numpy:
            formats.save_json(f"{base}_metadata.json", {"shape": list(slc.shape), "fortran_order": True,
                                                        "dtype": "complex128"})
        return written

    def _read_attrs(self):
        path = _store_path(self.output_file)
        if h5py is not None and not path.endswith(".npy"):
            with h5py.File(path, "r") as f:
                return dict(f.attrs)
        ap = path + ".attrs.json"
        return json.load(open(ap)) if os.path.exists(ap) else {}

    def generate_preview(self, preview_size=1000):
        """Amplitude previews of the focused image (main_gpu.py:635-743): a gray linear-scale
        window (the columns of :func:`preview_window`, the amplitude range of
        :func:`preview_scale`), the same window on a log10 scale, and, for images up to 10000
        columns, the whole image at the window's scale.  Returns the linear preview's path (None
        without matplotlib)."""
        try:
            import matplotlib
            matplotlib.use("Agg")
            import matplotlib.pyplot as plt
        except Exception:
            print("Matplotlib not installed. Cannot generate preview.")
            return None
        slc = self._read_slc()
        dim_rg, dim_az = slc.shape
        a0, a1 = preview_window(dim_az, preview_size, self._read_attrs().get("processed_azimuth"))
        if dim_az > preview_size:
            print(f"Generating preview for columns {a0}-{a1}...")
        amp = np.abs(slc[:, a0:a1])
        bad = int(np.sum(~np.isfinite(amp)))
        if bad:
            print(f"Warning: {bad} invalid (NaN/Inf) values detected in preview region")
            amp = np.nan_to_num(amp, nan=0.0, posinf=0.0, neginf=0.0)
        if amp.size == 0:
            print("Error: No valid data for preview generation")
            return None
        vmin, vmax = preview_scale(amp)
        print(f"Amplitude statistics: mean={amp.mean():.2e}, median={np.median(amp):.2e}, max={amp.max():.2e}")
        out = []
        for kind, data, kw, cbar, title, name in (
                ("linear", amp, dict(cmap="gray", vmin=vmin, vmax=vmax), "Amplitude",
                 "SAR Image Amplitude (Preview)", "SlcCuPy_SuperOptimized_Preview.png"),
                ("log", np.log10(amp + 1e-10), dict(cmap="viridis"), "Log10(Amplitude)",
                 "SAR Image Amplitude (Log Scale)", "SlcCuPy_SuperOptimized_LogScale_Preview.png")):
            fig = plt.figure(figsize=(12, 10))
            plt.imshow(data, aspect="auto", **kw)
            plt.colorbar(label=cbar)
            plt.title(f"{title}\nColumns {a0}-{a1} of {dim_az}")
            plt.xlabel("Azimuth Samples")
            plt.ylabel("Range Samples")
            plt.tight_layout()
            fn = os.path.join(self.checkpoint_dir, name)
            plt.savefig(fn, dpi=150)
            plt.close(fig)
            print(f"{'Preview' if kind == 'linear' else 'Log-scale preview'} saved to {fn}")
            out.append(fn)
        if dim_az <= 10000:
            full = np.nan_to_num(np.abs(slc), nan=0.0, posinf=0.0, neginf=0.0)
            fig = plt.figure(figsize=(16, 12))
            plt.imshow(full, cmap="gray", vmin=vmin, vmax=vmax, aspect="auto")
            plt.colorbar(label="Amplitude")
            plt.title("SAR Image Amplitude (Full Resolution)")
            plt.xlabel("Azimuth Samples")
            plt.ylabel("Range Samples")
            plt.tight_layout()
            fn = os.path.join(self.checkpoint_dir, "SlcCuPy_SuperOptimized_Full.png")
            plt.savefig(fn, dpi=150)
            plt.close(fig)
            print(f"Full resolution image saved to {fn}")
        return out[0]

    def close(self):
        if self._dataset is not None:
            self._dataset.close()
            self._dataset = None
