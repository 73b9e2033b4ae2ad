"""Import IGM's Keras-H5 emulator files (``model.h5`` + ``fieldin.dat``).

IGM ships pretrained emulators as Keras 2 HDF5 files
(igm/processes/iceflow/emulate/emulate.py:64-93, saved by :331-359), named
``pinnbp_{Nz}_{vs}_{arch}_{nb_layers}_{nb_out_filter}_{dim_arrhenius}_1_{a|e}``.
The CNN's weights sit at /model_weights/conv2d[_i]/conv2d[_i]/{kernel:0,bias:0}
(kernel HWIO).  h5py is not available in this image, so the datasets are
exported losslessly with the HDF5 command-line tool ``h5dump -b LE`` (a data
dump: nothing in the file is executed) and read with numpy.

``convert(model_dir, out_path)`` writes the weights as safetensors
(``kernel_{i}``, ``bias_{i}``, i = 0..nb_layers) with the fieldin list in the
metadata; ``igm_amd.emulator`` loads that format.
"""
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np

H5DUMP_CANDIDATES = ("h5dump", "/opt/conda/bin/h5dump")


def _h5dump():
    for c in H5DUMP_CANDIDATES:
        p = shutil.which(c) or (c if os.path.exists(c) else None)
        if p:
            return p
    raise FileNotFoundError("h5dump (HDF5 tools) is needed to read Keras .h5 files")


def _dataset(h5, path, name):
    """Shape + float32 data of one dataset, via a binary little-endian dump."""
    exe = _h5dump()
    head = subprocess.run([exe, "-H", "-d", name, path], capture_output=True, text=True, check=True).stdout
    m = re.search(r"DATASPACE\s+SIMPLE\s*\{\s*\(([^)]*)\)", head)
    shape = tuple(int(s) for s in m.group(1).split(","))
    assert "H5T_IEEE_F32LE" in head, f"{name}: expected float32 data"
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "d.bin")
        subprocess.run([exe, "-d", name, "-b", "LE", "-o", out, path], capture_output=True, check=True)
        data = np.fromfile(out, dtype="<f4")
    assert data.size == int(np.prod(shape)), (name, data.size, shape)
    return data.reshape(shape)


def load_keras_cnn(model_dir):
    """[(kernel HWIO float32, bias float32), ...] in layer order + fieldin list."""
    h5 = os.path.join(model_dir, "model.h5")
    listing = subprocess.run([_h5dump(), "-n", h5], capture_output=True, text=True, check=True).stdout
    layers = sorted({m for m in re.findall(r"/model_weights/(conv2d(?:_\d+)?)/", listing)},
                    key=lambda s: 0 if s == "conv2d" else int(s.split("_")[1]))
    weights = []
    for ln in layers:
        k = _dataset(h5, h5, f"/model_weights/{ln}/{ln}/kernel:0")
        b = _dataset(h5, h5, f"/model_weights/{ln}/{ln}/bias:0")
        weights.append((k, b))
    fieldin = []
    fi = os.path.join(model_dir, "fieldin.dat")
    if os.path.exists(fi):
        with open(fi) as f:
            fieldin = [ln.split()[0] for ln in f if ln.strip()]
    return weights, fieldin


def convert(model_dir, out_path, nz=None, vert_spacing=None):
    from safetensors.numpy import save_file
    weights, fieldin = load_keras_cnn(model_dir)
    tensors = {}
    for i, (k, b) in enumerate(weights):
        tensors[f"kernel_{i}"] = np.ascontiguousarray(k, dtype=np.float32)
        tensors[f"bias_{i}"] = np.ascontiguousarray(b, dtype=np.float32)
    meta = {"fieldin": ",".join(fieldin), "source": os.path.basename(os.path.normpath(model_dir)),
            "architecture": "cnn", "nb_layers": str(len(weights) - 1)}
    if nz is not None:
        meta["Nz"] = str(nz)
    if vert_spacing is not None:
        meta["vert_spacing"] = str(vert_spacing)
    save_file(tensors, out_path, metadata=meta)
    return out_path


if __name__ == "__main__":
    import sys
    print(convert(sys.argv[1], sys.argv[2]))
