"""Writes the golden data tables held in the reference's tests as JSON fixtures.

Reads the reference test files as text (no import, nothing executed) and keeps
only numbers: the 160-digit Beta path-factor table and the 3D tile-id known
answers. Run once in the build container:
    python tests/golden/extract_reference_tables.py /path/to/reference
The GPU box never runs this script (the reference is not present there).
"""

import ast
import json
import re
import sys
from pathlib import Path


def main(reference: Path) -> None:
    out = Path(__file__).resolve().parent
    src = (reference / "tests" / "test_beta_stability.py").read_text()
    start = src.index("REFERENCES = [")
    end = src.index("]", src.index("(1e30"))
    rows = ast.literal_eval(src[start + len("REFERENCES = "):end + 1])
    json.dump({"source": "tests/test_beta_stability.py:11-22 of the reference (160-digit Gamma/digamma values)",
               "columns": ["gamma", "ratio B(1/2,gamma+1)/sqrt(2pi)", "dlogB/dgamma"], "rows": rows},
              open(out / "beta_path_factor.json", "w"), indent=1)
    src = (reference / "tests" / "test_tile_bounds.py").read_text()
    pattern = r'pytest\.param\(\(([^)]*)\), (\d+), (\[[^\]]*\]|list\(range\(24\)\)), id="([^"]+)"\)'
    cases = []
    for mean, radius, ids, name in re.findall(pattern, src):
        ids = list(range(24)) if ids.startswith("list") else ast.literal_eval(ids)
        cases.append({"id": name, "mean": [float(v) for v in mean.split(",")], "radius": int(radius),
                      "expected_ids": ids})
    json.dump({"source": "tests/test_tile_bounds.py:16-28 of the reference; grid x/y/z = 3/2/4 tiles of size 4",
               "tile_size": 4, "n_tiles_xyz": [3, 2, 4], "cases": cases},
              open(out / "tile_bounds_3d.json", "w"), indent=1)


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
