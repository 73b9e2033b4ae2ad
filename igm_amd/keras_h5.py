"""Import and export IGM's Keras-H5 emulator files (``model.h5`` + ``fieldin.dat``).

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

``write_keras_cnn(path, weights, ...)`` goes the other way: it writes a Keras 2
HDF5 model file of the reference's CNN (neural_network.py:9-66) with the HDF5
C library (``libhdf5``, through ctypes), with the structure that Keras's
``model.save`` gives (the file layout of the emulators IGM ships): root
attributes ``backend``, ``keras_version`` and the ``model_config`` JSON of the
functional model, and ``/model_weights`` with ``layer_names``, one group per
layer holding ``weight_names`` and, for the convolutions, the datasets
``<layer>/kernel:0`` (HWIO) and ``<layer>/bias:0``.
"""
import ctypes
import json
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


# ------------------------------------------------------------------ export --

HDF5_CANDIDATES = ("libhdf5.so", "/opt/conda/lib/libhdf5.so")
# version of Keras recorded in the file: that of the emulators IGM ships,
# which the reference's TensorFlow loads
KERAS_VERSION = "2.8.0"


class _H5:
    """The few HDF5 C calls the writer needs (HDF5 1.10: hid_t is int64)."""

    def __init__(self):
        lib = None
        for c in HDF5_CANDIDATES:
            try:
                lib = ctypes.CDLL(c)
                break
            except OSError:
                continue
        if lib is None:
            raise FileNotFoundError("the HDF5 C library (libhdf5.so) is needed to write Keras .h5 files")
        hid, herr, sz = ctypes.c_int64, ctypes.c_int, ctypes.c_size_t
        vp, cp = ctypes.c_void_p, ctypes.c_char_p
        sigs = {
            "H5open": (herr, []),
            "H5Fcreate": (hid, [cp, ctypes.c_uint, hid, hid]),
            "H5Fclose": (herr, [hid]),
            "H5Gcreate2": (hid, [hid, cp, hid, hid, hid]),
            "H5Gclose": (herr, [hid]),
            "H5Screate": (hid, [ctypes.c_int]),
            "H5Screate_simple": (hid, [ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), vp]),
            "H5Sclose": (herr, [hid]),
            "H5Tcopy": (hid, [hid]),
            "H5Tset_size": (herr, [hid, sz]),
            "H5Tset_cset": (herr, [hid, ctypes.c_int]),
            "H5Tclose": (herr, [hid]),
            "H5Acreate2": (hid, [hid, cp, hid, hid, hid, hid]),
            "H5Awrite": (herr, [hid, hid, vp]),
            "H5Aclose": (herr, [hid]),
            "H5Dcreate2": (hid, [hid, cp, hid, hid, hid, hid, hid]),
            "H5Dwrite": (herr, [hid, hid, hid, hid, hid, vp]),
            "H5Dclose": (herr, [hid]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        if lib.H5open() < 0:
            raise RuntimeError("H5open failed")
        self.lib = lib
        g = lambda n: ctypes.c_int64.in_dll(lib, n).value  # noqa: E731
        self.C_S1, self.NATIVE_FLOAT = g("H5T_C_S1_g"), g("H5T_NATIVE_FLOAT_g")
        self.NATIVE_DOUBLE, self.IEEE_F32LE = g("H5T_NATIVE_DOUBLE_g"), g("H5T_IEEE_F32LE_g")

    def ok(self, r, what):
        if r < 0:
            raise RuntimeError(f"HDF5 call failed: {what}")
        return r

    def _space(self, dims):
        if dims is None:
            return self.ok(self.lib.H5Screate(0), "H5Screate")  # H5S_SCALAR
        arr = (ctypes.c_uint64 * max(1, len(dims)))(*dims)
        return self.ok(self.lib.H5Screate_simple(len(dims), arr, None), "H5Screate_simple")

    def str_attr(self, loc, name, value, utf8=False):
        """Variable-length string attribute: a scalar (str) or a 1-D array (list)."""
        L = self.lib
        t = self.ok(L.H5Tcopy(self.C_S1), "H5Tcopy")
        self.ok(L.H5Tset_size(t, ctypes.c_size_t(-1).value), "H5Tset_size")  # H5T_VARIABLE
        self.ok(L.H5Tset_cset(t, 1 if utf8 else 0), "H5Tset_cset")
        vals = [value] if isinstance(value, str) else list(value)
        space = self._space(None if isinstance(value, str) else [len(vals)])
        a = self.ok(L.H5Acreate2(loc, name.encode(), t, space, 0, 0), f"H5Acreate2 {name}")
        if vals:
            bufs = [ctypes.create_string_buffer(v.encode()) for v in vals]
            ptrs = (ctypes.c_char_p * len(vals))(*[ctypes.cast(b, ctypes.c_char_p) for b in bufs])
            self.ok(L.H5Awrite(a, t, ptrs), f"H5Awrite {name}")
        L.H5Aclose(a)
        L.H5Sclose(space)
        L.H5Tclose(t)

    def empty_attr(self, loc, name):
        """An empty 1-D float64 attribute (how Keras stores an empty name list)."""
        space = self._space([0])
        a = self.ok(self.lib.H5Acreate2(loc, name.encode(), self.NATIVE_DOUBLE, space, 0, 0), name)
        self.lib.H5Aclose(a)
        self.lib.H5Sclose(space)

    def group(self, loc, name):
        return self.ok(self.lib.H5Gcreate2(loc, name.encode(), 0, 0, 0), f"H5Gcreate2 {name}")

    def dataset_f32(self, loc, name, array):
        a = np.ascontiguousarray(array, dtype=np.float32)
        space = self._space(list(a.shape))
        d = self.ok(self.lib.H5Dcreate2(loc, name.encode(), self.IEEE_F32LE, space, 0, 0, 0),
                    f"H5Dcreate2 {name}")
        self.ok(self.lib.H5Dwrite(d, self.NATIVE_FLOAT, 0, 0, 0, a.ctypes.data), f"H5Dwrite {name}")
        self.lib.H5Dclose(d)
        self.lib.H5Sclose(space)


def cnn_model_config(nb_inputs, nb_outputs, nb_layers, nb_out_filter, conv_ker_size=3,
                     activation="LeakyReLU", dropout_rate=0.0, weight_initialization="glorot_uniform",
                     dropout_layers=None):
    """Keras functional ``model_config`` of the reference's CNN
    (neural_network.py:9-66): Input -> nb_layers x [Conv2D 'same' -> the one
    shared activation layer -> Dropout when dropout_rate > 0] -> Conv2D 1x1.
    Layer names follow Keras's automatic naming.  ``dropout_layers`` forces
    the Dropout layers on or off (the emulators IGM ships were built with
    them at rate 0).  Returns (config dict, layer names in model order)."""
    if dropout_layers is None:
        dropout_layers = dropout_rate > 0
    init_class = {"glorot_uniform": "GlorotUniform", "glorot_normal": "GlorotNormal",
                  "he_uniform": "HeUniform", "he_normal": "HeNormal"}.get(weight_initialization)
    if init_class is None:
        raise NotImplementedError(f"weight_initialization {weight_initialization!r}")

    def conv(i, filters, k, padding, inbound):
        name = "conv2d" if i == 0 else f"conv2d_{i}"
        return {"class_name": "Conv2D",
                "config": {"name": name, "trainable": True, "dtype": "float32", "filters": filters,
                           "kernel_size": [k, k], "strides": [1, 1], "padding": padding,
                           "data_format": "channels_last", "dilation_rate": [1, 1], "groups": 1,
                           "activation": "linear", "use_bias": True,
                           "kernel_initializer": {"class_name": init_class, "config": {"seed": None}},
                           "bias_initializer": {"class_name": "Zeros", "config": {}},
                           "kernel_regularizer": None, "bias_regularizer": None,
                           "activity_regularizer": None, "kernel_constraint": None,
                           "bias_constraint": None},
                "name": name, "inbound_nodes": [[[inbound[0], inbound[1], 0, {}]]]}

    if activation == "LeakyReLU":
        act_name = "leaky_re_lu"
        act = {"class_name": "LeakyReLU",
               "config": {"name": act_name, "trainable": True, "dtype": "float32",
                          "alpha": float(np.float32(0.01))}}
    else:
        act_name = "activation"
        act = {"class_name": "Activation",
               "config": {"name": act_name, "trainable": True, "dtype": "float32", "activation": activation}}
    layers = [{"class_name": "InputLayer",
               "config": {"batch_input_shape": [None, None, None, nb_inputs], "dtype": "float32",
                          "sparse": False, "ragged": False, "name": "input_1"},
               "name": "input_1", "inbound_nodes": []}]
    order = ["input_1"]
    act["name"] = act_name
    act["inbound_nodes"] = []
    prev = ("input_1", 0)
    for i in range(nb_layers):
        c = conv(i, nb_out_filter, conv_ker_size, "same", prev)
        layers.append(c)
        order.append(c["name"])
        act["inbound_nodes"].append([[c["name"], 0, 0, {}]])
        if i == 0:
            layers.append(act)
            order.append(act_name)
        prev = (act_name, i)
        if dropout_layers:
            dn = "dropout" if i == 0 else f"dropout_{i}"
            layers.append({"class_name": "Dropout",
                           "config": {"name": dn, "trainable": True, "dtype": "float32",
                                      "rate": dropout_rate, "noise_shape": None, "seed": None},
                           "name": dn, "inbound_nodes": [[[prev[0], prev[1], 0, {}]]]})
            order.append(dn)
            prev = (dn, 0)
    head = conv(nb_layers, nb_outputs, 1, "valid", prev)
    layers.append(head)
    order.append(head["name"])
    cfg = {"class_name": "Functional",
           "config": {"name": "model", "layers": layers, "input_layers": [["input_1", 0, 0]],
                      "output_layers": [[head["name"], 0, 0]]}}
    return cfg, order


def write_keras_cnn(path, weights, activation="LeakyReLU", dropout_rate=0.0,
                    weight_initialization="glorot_uniform", dropout_layers=None):
    """Write ``weights`` [(kernel HWIO, bias), ...] (nb_layers 3x3 layers then
    the 1x1 head) as a Keras 2 HDF5 model file of the reference's CNN."""
    k0 = np.asarray(weights[0][0])
    nb_layers = len(weights) - 1
    cfg, order = cnn_model_config(k0.shape[2], np.asarray(weights[-1][0]).shape[3], nb_layers,
                                  k0.shape[3], k0.shape[0], activation, dropout_rate,
                                  weight_initialization, dropout_layers)
    conv_names = ["conv2d" if i == 0 else f"conv2d_{i}" for i in range(nb_layers + 1)]
    h5 = _H5()
    L = h5.lib
    f = h5.ok(L.H5Fcreate(os.fsencode(path), 2, 0, 0), f"H5Fcreate {path}")  # H5F_ACC_TRUNC
    try:
        h5.str_attr(f, "backend", "tensorflow", utf8=True)
        h5.str_attr(f, "keras_version", KERAS_VERSION, utf8=True)
        h5.str_attr(f, "model_config", json.dumps(cfg))
        mw = h5.group(f, "model_weights")
        h5.str_attr(mw, "backend", "tensorflow")
        h5.str_attr(mw, "keras_version", KERAS_VERSION)
        h5.str_attr(mw, "layer_names", order)
        for name in order + ["top_level_model_weights"]:
            g = h5.group(mw, name)
            if name in conv_names:
                i = conv_names.index(name)
                h5.str_attr(g, "weight_names", [f"{name}/kernel:0", f"{name}/bias:0"])
                inner = h5.group(g, name)
                h5.dataset_f32(inner, "bias:0", weights[i][1])
                h5.dataset_f32(inner, "kernel:0", weights[i][0])
                L.H5Gclose(inner)
            else:
                h5.empty_attr(g, "weight_names")
            L.H5Gclose(g)
        L.H5Gclose(mw)
    finally:
        L.H5Fclose(f)
    return path


if __name__ == "__main__":
    import sys
    print(convert(sys.argv[1], sys.argv[2]))
