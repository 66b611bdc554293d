"""Writes tests/golden/reference_packings.npz from the reference's committed
packings (assets/initial-configs/*/x_relaxed.txt: rod endpoints, header
rod_radius) and their contact counts (assets/packings_metadata.csv, column
n_contacts: pairs whose segment distance is below 1.01 d, free space; ebar0:
mean |linking number| over all rod pairs, rounded to 6 digits; ent_pairs: the
pair count).  One packing per aspect ratio of the first seed group.  Run in a container that
has the reference tree at /root/reference."""
import csv
import re
import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent / "reference_packings.npz"


def read_endpoints(path):
    radius, rows = None, []
    for line in path.read_text().splitlines():
        line = line.strip()
        if not line:
            continue
        if line.startswith("#"):
            m = re.search(r"rod_radius\s*=\s*([0-9.eE+-]+)", line)
            if m:
                radius = float(m.group(1))
            continue
        v = line.split()
        if len(v) >= 6:
            rows.append([float(t) for t in v[:6]])
    return np.asarray(rows), radius


def main():
    rows = list(csv.DictReader(open(REF / "assets" / "packings_metadata.csv")))
    group = rows[0]["seed_group"]
    pick = [r for r in rows if r["seed_group"] == group]
    seen, chosen = set(), []
    for r in pick:
        if r["alpha_nominal"] not in seen:
            seen.add(r["alpha_nominal"])
            chosen.append(r)
    out = {}
    for k, r in enumerate(chosen):
        ends, radius = read_endpoints(REF / r["path"])
        out[f"ends_{k}"] = ends
        out[f"radius_{k}"] = np.float64(radius)
        out[f"n_contacts_{k}"] = np.int64(r["n_contacts"])
        out[f"ebar0_{k}"] = np.float64(r["ebar0"])
        out[f"ent_pairs_{k}"] = np.int64(r["ent_pairs"])
        out[f"id_{k}"] = np.array(r["packing_id"])
    out["count"] = np.int64(len(chosen))
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({len(chosen)} packings)", file=sys.stderr)


if __name__ == "__main__":
    main()
