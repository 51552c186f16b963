"""Writes reference_runs.json: the block averages of the reference program on the decks of
decks.py, for the statistical parity tests of the host driver.

The program is oracle/_ref/pimc, compiled from the upstream sources by oracle/ref.mk (run
`make -C oracle ref` first; the reference tree is only present in the build container).  It
runs single-threaded in a scratch directory; each deck's .eng lines (one per block after the
equilibration blocks) are kept whole, so the test can form means and standard errors itself.

    python tests/golden/make_reference_runs.py
"""
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "moribs-pimc_amd"))
sys.path.insert(0, HERE)

import decks  # noqa: E402


def run(name):
    exe = os.path.join(ROOT, "oracle", "_ref", "pimc")
    with tempfile.TemporaryDirectory() as d:
        decks.materialise(name, d)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        t0 = time.time()
        with open(os.path.join(d, "log.txt"), "w") as log:
            subprocess.run([exe], cwd=d, env=env, stdout=log, stderr=subprocess.STDOUT, timeout=3600)
        wall = time.time() - t0
        prefix = "CO2_monomer" if name == "co2_100K" else "he4co2"
        eng = decks.read_eng(os.path.join(d, prefix + ".eng"))
    return dict(deck=decks.DECKS[name].format(blocks=decks.BLOCKS[name]), columns=["block", *decks.ENG_COLUMNS],
                eng=eng.tolist(), wall_s=round(wall, 1), program="oracle/_ref/pimc, OMP_NUM_THREADS=1")


def main():
    out = {name: run(name) for name in decks.DECKS}
    with open(os.path.join(HERE, "reference_runs.json"), "w") as f:
        json.dump(out, f, indent=1)
    for name, r in out.items():
        print(name, len(r["eng"]), "blocks", r["wall_s"], "s")


if __name__ == "__main__":
    main()
