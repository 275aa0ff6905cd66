"""The committed rocprofv3 evidence the bench cites: every traffic file the
bench reads (one per preconditioner) has its PMC summary beside it, stating
the same per-launch traffic, and a kernel-trace summary of the same workload."""
import glob
import json
import os
import re

from conftest import ROOT

PROF = os.path.join(ROOT, "profiles")


def test_traffic_files_match_their_pmc_summaries():
    files = sorted(glob.glob(os.path.join(PROF, "*_traffic.json")))
    assert files
    for f in files:
        tr = json.load(open(f))
        stem = f[: -len("_traffic.json")]
        assert tr["precond"] in os.path.basename(stem)
        pmc = open(stem + "_pmc.md").read()
        assert os.path.exists(stem + "_kernel_trace.md")
        m = re.search(r"measured HBM traffic per launch: fetch ([\d.]+) MB \+ write ([\d.]+) MB = ([\d.]+) MB", pmc)
        assert m, f
        fetch, write, total = (float(g) for g in m.groups())
        assert abs(fetch - tr["fetch_bytes"] / 1e6) <= 0.1
        assert abs(write - tr["write_bytes"] / 1e6) <= 0.1
        assert abs(total - tr["traffic_bytes"] / 1e6) <= 0.1
        assert abs(tr["traffic_bytes"] - tr["fetch_bytes"] - tr["write_bytes"]) <= 1.0
