"""Writes tests/golden/les_closure_64.npz: the trained closure of the
512^3 -> 64^3 learned-LES experiment (weights, channel scales, gain, spectral
tilt, backscatter clipping) with the filtered DNS's mean shell spectrum and
resolved energy of its training record, by running the experiment with its
default configuration (hit3d.les_experiment; about 10 minutes on one MI355X).

Run from the repository root: python tests/golden/make_les_closure.py [result.json]"""

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "3dhit_amd"))

from hit3d import les_experiment  # noqa: E402

if __name__ == "__main__":
    argv = ["--save", os.path.join(HERE, "les_closure_64.npz")]
    if len(sys.argv) > 1:
        argv += ["--out", sys.argv[1]]
    les_experiment.main(argv)
