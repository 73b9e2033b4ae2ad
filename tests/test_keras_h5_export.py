"""Keras-H5 export of the emulator (igm_amd.keras_h5.write_keras_cnn) on the CPU.

Pinned against the structure of an emulator file the reference ships
(pinnbp_10_4_cnn_16_32_2_1_a/model.h5, saved by Keras; the fixture holds its
groups, attributes with their values, and dataset types and shapes): an
export of the same architecture must reproduce that structure exactly,
model_config JSON included.  The weights round-trip bit for bit through the
importer (igm_amd.keras_h5.load_keras_cnn, h5dump-based).  Needs h5dump and
libhdf5 (present in this image).
"""
import json
import os

import numpy as np
import pytest

from igm_amd import keras_h5 as K
from tests.helpers import h5_structure

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "keras_h5_pinnbp_10_4_cnn_16_32.json")


def random_weights(nb_layers, filters, nb_in, nb_out, k=3, seed=0):
    rng = np.random.default_rng(seed)
    w = [(rng.standard_normal((k, k, nb_in, filters)), rng.standard_normal(filters))]
    w += [(rng.standard_normal((k, k, filters, filters)), rng.standard_normal(filters))
          for _ in range(nb_layers - 1)]
    w += [(rng.standard_normal((1, 1, filters, nb_out)), rng.standard_normal(nb_out))]
    return [(a.astype(np.float32), b.astype(np.float32)) for a, b in w]


def test_export_matches_the_shipped_keras_file(tmp_path):
    with open(FIXTURE) as f:
        ref = json.load(f)
    path = str(tmp_path / "model.h5")
    K.write_keras_cnn(path, random_weights(16, 32, 5, 20), dropout_layers=True)
    ours = json.loads(json.dumps(h5_structure(path)))  # same normalisation as the fixture
    assert ours["attrs"]["model_config"] == ref["attrs"]["model_config"]
    assert ours == ref


def test_weights_round_trip(tmp_path):
    for nb_layers, filters, nz in ((16, 32, 10), (4, 8, 3)):
        w = random_weights(nb_layers, filters, 5, 2 * nz, seed=nb_layers)
        d = tmp_path / f"m{nb_layers}"
        d.mkdir()
        K.write_keras_cnn(str(d / "model.h5"), w)
        back, _ = K.load_keras_cnn(str(d))
        assert len(back) == len(w)
        for (k0, b0), (k1, b1) in zip(w, back):
            assert k1.dtype == np.float32 and np.array_equal(k0, k1) and np.array_equal(b0, b1)


def test_model_config_without_dropout_layers():
    """The current reference builder (neural_network.py:9-66) adds Dropout only
    for a positive rate: the convolutions then feed each other through the
    one shared activation layer."""
    cfg, order = K.cnn_model_config(5, 20, 3, 16)
    layers = {l["name"]: l for l in cfg["config"]["layers"]}
    assert order == ["input_1", "conv2d", "leaky_re_lu", "conv2d_1", "conv2d_2", "conv2d_3"]
    assert layers["conv2d"]["inbound_nodes"] == [[["input_1", 0, 0, {}]]]
    assert layers["conv2d_1"]["inbound_nodes"] == [[["leaky_re_lu", 0, 0, {}]]]
    assert layers["conv2d_3"]["inbound_nodes"] == [[["leaky_re_lu", 2, 0, {}]]]
    assert [n[0][0] for n in layers["leaky_re_lu"]["inbound_nodes"]] == ["conv2d", "conv2d_1", "conv2d_2"]
    assert layers["conv2d_3"]["config"]["kernel_size"] == [1, 1]
    assert layers["conv2d_3"]["config"]["padding"] == "valid"
    assert cfg["config"]["output_layers"] == [["conv2d_3", 0, 0]]
    cfg2, order2 = K.cnn_model_config(5, 20, 2, 16, dropout_rate=0.1)
    assert order2 == ["input_1", "conv2d", "leaky_re_lu", "dropout", "conv2d_1", "dropout_1", "conv2d_2"]


def test_unknown_initializer_raises():
    with pytest.raises(NotImplementedError):
        K.cnn_model_config(5, 20, 2, 16, weight_initialization="orthogonal")
