"""Writes tests/golden/plane256.json: digests of the oracle's outputs on the
plane-256 workload (256x256 receivers, 16x16 light samples, 10k triangles,
seed 0).  Regenerate only when the oracle's contract changes on purpose."""
import importlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle"), os.path.dirname(HERE)]

import oracle  # noqa: E402
from test_oracle import digests  # noqa: E402

wl = importlib.import_module("4dzb_amd.workloads")
out = digests(oracle.run(wl.plane_256()))
with open(os.path.join(HERE, "plane256.json"), "w") as f:
    json.dump(out, f, indent=1, sort_keys=True)
print(out)
