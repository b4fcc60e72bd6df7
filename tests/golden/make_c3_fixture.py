"""Writes c3_shapes.json: per shape of the batched-polyhedra workload (hipgisaxs_amd.meshes.c3_shapes,
64 shapes, seed 20261206), its semi-axes, target triangle count, grid side and the triangle count of
the hull the library's quickhull builds (2n - 4 for the n = target/2 + 2 points on the ellipsoid).

    python tests/golden/make_c3_fixture.py
"""
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from hipgisaxs_amd import meshes  # noqa: E402


def main():
    rows = []
    for i, s in enumerate(meshes.c3_shapes()):
        rows.append({"shape": i, "axes": [round(float(a), 12) for a in s["axes"]], "target": s["target"],
                     "side": s["side"], "ntri": int(s["tri"].shape[0])})
    out = Path(__file__).with_name("c3_shapes.json")
    out.write_text(json.dumps({"seed": 20261206, "shapes": rows}, indent=1) + "\n")
    print(f"wrote {out} ({sum(r['ntri'] for r in rows)} triangles)")


if __name__ == "__main__":
    main()
