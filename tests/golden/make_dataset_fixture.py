"""Writes the B-spline dataset fixtures used by the tests: the first 300
curves of dataset/BSpline/Bundle_0.txt (data only, verbatim records), the
files Bundle_0..9.txt (10,000 curves) verbatim and gzip-compressed, and the
interval / curve totals of the whole dataset and of those ten files as the
loader reads them.

Run in a tree that has the reference checkout at /root/reference:
    python tests/golden/make_dataset_fixture.py
"""
import gzip
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from gcurval_amd import dataset as D  # noqa: E402

SRC = "/root/reference/dataset/BSpline"
N_BUNDLES = 10  # Bundle_0..9 as compressed fixtures


def main():
    with open(os.path.join(SRC, "Bundle_0.txt")) as f:
        text = f.read()
    records = [r for r in text.strip().split("\n\n") if r.strip()]
    head = "\n\n".join(records[:300]) + "\n"
    with open(os.path.join(HERE, "dataset_bundle0_head.txt"), "w") as f:
        f.write(head)
    for j in range(N_BUNDLES):
        with open(os.path.join(SRC, f"Bundle_{j}.txt"), "rb") as f:
            raw = f.read()
        # mtime 0: the same bytes on every run
        with open(os.path.join(HERE, f"dataset_bundle{j}.txt.gz"), "wb") as f:
            f.write(gzip.compress(raw, compresslevel=9, mtime=0))
    paths = D.bundle_paths(SRC)
    b = D.load_bspline_dataset(paths)
    b10 = D.load_bspline_dataset([os.path.join(SRC, f"Bundle_{j}.txt") for j in range(N_BUNDLES)])
    meta = {"source": "dataset/BSpline/Bundle_{0..99}.txt", "n_curve": int(b.n_curve),
            "n_intervals": int(b.n_intervals), "fixture_curves": 300,
            "degrees": sorted(set(int(p) for p in b.degree)),
            "fixture_bundles": N_BUNDLES, "bundles_n_curve": int(b10.n_curve),
            "bundles_n_intervals": int(b10.n_intervals)}
    with open(os.path.join(HERE, "dataset_meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(meta)


if __name__ == "__main__":
    main()
