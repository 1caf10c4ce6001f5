"""Regenerate tests/golden/ref_fixtures.json from a from-source build of the
reference (oracle/_ref/ref_dump, built by `make -C oracle ref`).

Fixture generation only; needs the reference sources at /root/reference
(or REF=...).  The GPU-side tests only read the committed JSON.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("REF", "/root/reference")


def main() -> int:
    subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "ref", f"REF={REF}", "-j8"])
    data = os.path.join(HERE, "data")
    os.makedirs(data, exist_ok=True)
    # The reference's bundled MatrixMarket fixtures and its published results.
    for f in sorted(os.listdir(os.path.join(REF, "data"))):
        if f.endswith(".mtx"):
            shutil.copy(os.path.join(REF, "data", f), data)
    shutil.copy(os.path.join(REF, "dataset_results.csv"), data)
    out = os.path.join(HERE, "ref_fixtures.json")
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.check_call([os.path.join(ROOT, "oracle", "_ref", "ref_dump"), out, data], env=env)
    print("wrote", out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
