"""Writes the golden fixtures of the synthesis path: inputs (rows, masks, channel list) and the oracle's outputs.

Each case is a few epochs of a c0-, c1- or c2-like channel mix in one output format.  The inputs come from the
library's scenario stage; the outputs from the CPU oracle (oracle/), a restatement of the reference's synthesis
kernel (src/products/run/synthesis/kernel.zig).  Small cases keep their full output bytes; every case keeps the
SHA-256 of its bytes, per-epoch clip counts and pre-quantisation sums, in manifest.json.

    python tests/golden/make_golden.py      (from the repository root, after the library and oracle are built)
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import binding as O  # noqa: E402
from pharos_amd import _lib  # noqa: E402
from pharos_amd import scenario as S  # noqa: E402


def _mini(sc, nch=None, fmt=None, noise=None, seed=None):
    chans = sc.channels[:nch] if nch else sc.channels
    return S.Scenario(sc.name, sc.samples_per_epoch, fmt or sc.format, 4, chans,
                      noise=sc.noise if noise is None else noise, seed=sc.seed if seed is None else seed)


# (name, scenario, first epoch, epochs, gain, keep full outputs)
def cases():
    c0, c1, c2 = S.c0(seed=3), S.c1(seed=3), S.c2(seed=3)
    out = [
        ("c0_iq8_noisefree", _mini(c0), 1234, 3, 1.0, True),
        ("c0_iq2_noise", _mini(c0, fmt="IQ2", noise=True, seed=11), 77, 2, 1.0, True),
        ("c0_iq4_noise", _mini(c0, fmt="IQ4", noise=True, seed=12), 5001, 2, 0.7, True),
        ("c0_iq16_noise", _mini(c0, fmt="IQ16", noise=True, seed=13), 9998, 2, 1.3, True),
        ("c1_iq4", _mini(c1), 30_000, 2, 0.9, False),
        ("c1_iq8_noisefree", _mini(c1, nch=16, fmt="IQ8", noise=False), 41_999, 2, 1.0, False),
        ("c2_iq16", _mini(c2), 12_345, 1, 1.0, False),
        ("c2_iq2", _mini(c2, fmt="IQ2"), 100, 1, 1.0, False),
    ]
    return out


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    manifest = {"generator": "tests/golden/make_golden.py", "oracle": "oracle/ (C restatement of kernel.zig)",
                "cases": {}}
    for name, sc, first, epochs, gain, keep in cases():
        rows, masks = sc.window(first, epochs)
        tabs = O.Tables([(c.signal, c.prn) for c in sc.channels])
        out, clips, sre, sim = O.synth_window(tabs, rows, masks, sc.samples_per_epoch, sc.format_id,
                                              0 if sc.noise else 1, sc.seed, first, gain, threads=8, want_sums=True)
        arrays = {
            "rows": rows.view(np.uint8).reshape(-1),
            "masks": masks,
            "signals": np.array([_lib.SIGNAL[c.signal] for c in sc.channels], dtype=np.uint32),
            "prns": np.array([c.prn for c in sc.channels], dtype=np.uint32),
            "clips": clips,
        }
        if keep:
            arrays["bytes"] = out
            arrays["sums_re"] = sre
            arrays["sums_im"] = sim
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **arrays)
        manifest["cases"][name] = {
            "samples_per_epoch": sc.samples_per_epoch, "format": sc.format, "noise": bool(sc.noise),
            "seed": int(sc.seed), "first_epoch": int(first), "epochs": int(epochs), "gain": float(np.float32(gain)),
            "channels": len(sc.channels), "sha256_bytes": sha(out), "sha256_sums_re": sha(sre),
            "sha256_sums_im": sha(sim), "clips": [int(c) for c in clips], "full_outputs": keep,
        }
        print(f"{name}: {len(sc.channels)} ch x {epochs} epochs, {out.nbytes} bytes, clips {list(map(int, clips))}")
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
