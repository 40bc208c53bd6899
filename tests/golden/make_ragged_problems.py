"""Writes ragged_problems.json: the 2000-problem ragged batch used by the
benchmark (image sides log-uniform in [64, 1024], independent per axis; odd
filter sides uniform in {3, 5, ..., 15}, independent per axis; DEFAULT or
EXPAND with probability 1/2 each). Entries are [W, H, fw, fh, mode]."""
import json
import os

import numpy as np


def main(seed=20261202 + 4, n=2000):
    rng = np.random.default_rng(seed)
    sides = np.floor(np.exp(rng.uniform(np.log(64), np.log(1024), (n, 2)))).astype(int)
    filt = rng.choice(np.arange(3, 16, 2), (n, 2))
    modes = rng.integers(0, 2, n)
    probs = [[int(w), int(h), int(fw), int(fh), int(m)]
             for (w, h), (fw, fh), m in zip(sides, filt, modes)]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ragged_problems.json")
    with open(out, "w") as f:
        json.dump({"seed": seed, "format": "[W, H, fw, fh, mode (0 DEFAULT, 1 EXPAND)]",
                   "problems": probs}, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
