"""Writes tests/golden/opensps-test-data: the libraries of the reference's test data release that the batched path
reads (reference_spectra/vega, reference_spectra/sun, filters/allfilters), copied file for file from its
tests/data/opensps-test-data, and a root index opensps-data.json listing them. Data only (manifests and raw arrays);
run where the reference is present:  python tests/golden/make_datalib_subset.py [reference root]"""
import json
import os
import shutil
import sys

LIBS = [("reference_spectra", "vega"), ("reference_spectra", "sun"), ("filters", "allfilters")]


def main(ref_root="/root/reference"):
    src = os.path.join(ref_root, "tests", "data", "opensps-test-data")
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "opensps-test-data")
    with open(os.path.join(src, "opensps-data.json"), encoding="utf-8") as f:
        idx = json.load(f)
    keep = []
    for lib in idx["libraries"]:
        if (lib["kind"], lib["name"]) in LIBS:
            keep.append(lib)
            d = os.path.join(dst, lib["path"])
            os.makedirs(d, exist_ok=True)
            for fn in sorted(os.listdir(os.path.join(src, lib["path"]))):
                shutil.copyfile(os.path.join(src, lib["path"], fn), os.path.join(d, fn))
    assert len(keep) == len(LIBS), keep
    out = {k: v for k, v in idx.items() if k != "libraries"}
    out["libraries"] = keep
    out["title"] = idx.get("title", "") + " (subset: " + ", ".join(f"{k}/{n}" for k, n in LIBS) + ")"
    with open(os.path.join(dst, "opensps-data.json"), "w", encoding="utf-8") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main(*sys.argv[1:])
