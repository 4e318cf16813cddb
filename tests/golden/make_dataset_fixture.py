"""Writes the B-spline dataset fixture used by the tests: the first 300
curves of dataset/BSpline/Bundle_0.txt (data only, verbatim records) and the
interval / curve totals of the whole dataset as the loader reads it.

Run in a tree that has the reference checkout at /root/reference:
    python tests/golden/make_dataset_fixture.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from gcurval_amd import dataset as D  # noqa: E402

SRC = "/root/reference/dataset/BSpline"


def main():
    with open(os.path.join(SRC, "Bundle_0.txt")) as f:
        text = f.read()
    records = [r for r in text.strip().split("\n\n") if r.strip()]
    head = "\n\n".join(records[:300]) + "\n"
    with open(os.path.join(HERE, "dataset_bundle0_head.txt"), "w") as f:
        f.write(head)
    paths = D.bundle_paths(SRC)
    b = D.load_bspline_dataset(paths)
    meta = {"source": "dataset/BSpline/Bundle_{0..99}.txt", "n_curve": int(b.n_curve),
            "n_intervals": int(b.n_intervals), "fixture_curves": 300,
            "degrees": sorted(set(int(p) for p in b.degree))}
    with open(os.path.join(HERE, "dataset_meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(meta)


if __name__ == "__main__":
    main()
