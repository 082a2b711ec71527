"""Writes the golden fixtures under tests/golden/ from the reference's data files.

Run in a container that has the reference checkout at /root/reference:
    python tests/golden/make_fixtures.py
Copies are data only (saved scenes and exported traces):
  * saved_states/*.bin  - small saved scenes with their embedded experiment scripts
                          (reference saved_states/, flatbuffer State tables)
  * traces.npz          - the exported traces of those experiments (reference data/*.csv),
                          one float64 array [rows, 12] per file, keyed by file stem
"""
import csv
import glob
import os
import shutil

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

SCENES = ["TEXPS", "PBEXPS", "Chain_EXP", "Chain_TEXP", "CHTEXPDET", "TEXP_Base", "Chain_Base", "Poly_Base"]
TRACE_GLOBS = ["TEXP_*.csv", "pb_experiment_*_1.csv", "Chain_[0-9]*.csv", "Chain_T_*.csv", "Chain_DET*.csv"]


def main():
    os.makedirs(os.path.join(HERE, "saved_states"), exist_ok=True)
    for s in SCENES:
        shutil.copyfile(os.path.join(REF, "saved_states", s + ".bin"), os.path.join(HERE, "saved_states", s + ".bin"))
    traces = {}
    for g in TRACE_GLOBS:
        for path in sorted(glob.glob(os.path.join(REF, "data", g))):
            with open(path) as f:
                rows = list(csv.reader(f))
            header, body = rows[0], rows[1:]
            assert header[0] == "Timestamp" and len(header) == 12, path
            traces[os.path.splitext(os.path.basename(path))[0]] = np.array([[float(x) for x in r] for r in body])
    np.savez_compressed(os.path.join(HERE, "traces.npz"), **traces)
    print("scenes:", len(SCENES), "traces:", len(traces))


if __name__ == "__main__":
    main()
