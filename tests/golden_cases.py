"""Loads the golden fixtures (tests/golden/*.npz + manifest.json) as scenario inputs and expected outputs."""
import hashlib
import json
import os

import numpy as np

from pharos_amd import _lib
from pharos_amd.scenario import Scenario, Channel

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)["cases"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load(name):
    """(scenario, meta, rows [epochs, nch], masks, arrays) of one case."""
    meta = manifest()[name]
    z = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    arrays = {k: z[k] for k in z.files}
    chans = [Channel(_lib.SIGNALS[int(s)], int(p)) for s, p in zip(arrays["signals"], arrays["prns"])]
    sc = Scenario(name, meta["samples_per_epoch"], meta["format"], meta["epochs"], chans, noise=meta["noise"],
                  seed=meta["seed"])
    rows = arrays["rows"].view(_lib.ROW_DTYPE).reshape(meta["epochs"], len(chans))
    return sc, meta, rows, arrays["masks"], arrays
