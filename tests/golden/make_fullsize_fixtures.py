"""Writes the oracle's canonical tet sets of the full-size benchmark inputs
into inputs_sha256.json (count + SHA-256 of the canonical real-only list).

Run in the build container (CPU only):

    python tests/golden/make_fullsize_fixtures.py [name ...]

Each input comes from wtet_amd/workloads.py, its byte hash is recorded beside
the tet-set hash, and the set is the oracle's serial Bowyer-Watson (exact
predicates + SoS, oracle/delaunay.c) of the normalized points: every quad
sorted ascending, the list sorted lexicographically, as little-endian u32.
These fixtures are what the GPU tests compare the device result with at full
size; the oracle itself does not travel to the GPU run.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from wtet_amd import workloads as W  # noqa: E402

GOLD_PATH = os.path.join(ROOT, "tests", "golden", "inputs_sha256.json")

INPUTS = {
    "uniform_1000000_seed1": lambda: W.uniform(1_000_000, 1),
    "sphere_1000000_seed2": lambda: W.sphere(1_000_000, 2),
    "lattice_128": lambda: W.lattice(128),
    "anisotropic_5000000_seed5": lambda: W.anisotropic(5_000_000, 5),
}


def canonical_sha(pts):
    pn, _, _ = O.normalize_points(pts)
    tets = O.delaunay(pn)
    can = O.canonical_real(tets, len(pts))
    return int(len(can)), hashlib.sha256(can.astype("<u4").tobytes()).hexdigest()


def main(names):
    gold = json.load(open(GOLD_PATH))
    for name in names:
        pts = INPUTS[name]()
        t0 = time.perf_counter()
        count, sha = canonical_sha(pts)
        dt = time.perf_counter() - t0
        h = W.sha256(pts)
        assert gold.get(name, h) == h, f"{name}: input bytes changed"
        gold[name] = h
        gold["oracle_canonical_" + name] = {"count": count, "sha256": sha}
        print(f"{name}: {count} real tets, oracle {dt:.1f} s", flush=True)
        with open(GOLD_PATH, "w") as f:
            json.dump(gold, f, indent=1)
            f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1:] or list(INPUTS))
