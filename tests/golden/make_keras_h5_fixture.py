"""Writes keras_h5_pinnbp_10_4_cnn_16_32.json: the HDF5 structure (groups,
attributes with their values, dataset types and shapes -- no weight values)
of one of the Keras emulator files the reference ships,
igm/processes/iceflow/emulate/emulators/pinnbp_10_4_cnn_16_32_2_1_a/model.h5
(saved by Keras 2.8 from the reference's CNN, 16 layers of 32 filters, with
Dropout layers at rate 0).  Read with h5dump (tests.helpers.h5_structure).

Run from the repository root: python -m tests.golden.make_keras_h5_fixture
"""
import json
import os

from tests.helpers import h5_structure

SRC = "/root/reference/igm/processes/iceflow/emulate/emulators/pinnbp_10_4_cnn_16_32_2_1_a/model.h5"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "keras_h5_pinnbp_10_4_cnn_16_32.json")

if __name__ == "__main__":
    with open(OUT, "w") as f:
        json.dump(h5_structure(SRC), f, indent=0, sort_keys=True)
    print(OUT)
