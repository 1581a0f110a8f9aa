"""Runs one compiled-reference call of the oracle in a fresh interpreter
(TEST INFRASTRUCTURE ONLY): python -m oracle._child DIR reads DIR/in.json
(function name, scalar arguments) and DIR/in.npz (array arguments), calls
oracle.<name>(..., reference=True) and writes the outputs to DIR/out.npz."""
import json
import os
import sys

import numpy as np

import oracle

_OUTPUTS = {"covd_resid": ("resid", "covd", "iperm", "ierr")}


def main(d):
    with open(os.path.join(d, "in.json")) as f:
        spec = json.load(f)
    with np.load(os.path.join(d, "in.npz"), allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    name = spec["name"]
    res = getattr(oracle, name)(**spec["scalars"], **arrays, reference=True)
    np.savez(os.path.join(d, "out.npz"), **{k: np.asarray(v) for k, v in zip(_OUTPUTS[name], res)})


if __name__ == "__main__":
    main(sys.argv[1])
