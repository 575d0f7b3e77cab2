"""Writes reference_kats.json: the known answers the reference's own
tests hold for this path (values copied as data, with their source).

  hermitian mirror   aux/test/test_dfttools.cxx:131-265 (inputs and the
                     asserted relations, expanded to explicit outputs)
  zero-padding KAT   sigproc/test/test_zero_padding.cxx:26 (the
                     10-sample resampling of {1,2,3,2,1}, 6 digits)
  IDFT conformance   aux/test/test_idft.cxx:28-237 (sizes and shapes)
"""
import json
import os

even = [[1, 11], [2, 22], [3, 33], [4, 44], [5, 55], [6, 66]]
odd = even[:5]
kats = {
    "hermitian_mirror": {
        "source": "aux/test/test_dfttools.cxx:131-265",
        "even_in": even,
        "even_out": [[1, 0], [2, 22], [3, 33], [4, 0], [3, -33], [2, -22]],
        "odd_in": odd,
        "odd_out": [[1, 0], [2, 22], [3, 33], [3, -33], [2, -22]],
    },
    "zero_padding": {
        "source": "sigproc/test/test_zero_padding.cxx:26",
        "input": [1, 2, 3, 2, 1],
        "output": [1, 1.35279, 2, 2.69443, 3, 2.69443, 2, 1.35279, 1, 0.905573],
        "digits": 6,
    },
    "idft": {
        "source": "aux/test/test_idft.cxx:205-237",
        "oned_size": 1024,
        "twod_shape": [128, 128],
        "oneb_shapes": [[2, 8], [8, 2]],
        "transpose_shapes": [[2, 8], [8, 2]],
        "tolerance": 1e-8,
    },
}
here = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(here, "reference_kats.json"), "w") as f:
    json.dump(kats, f, indent=1)
