"""Copies the reference's own test data files (inputs and expected outputs
of test_fftcorr_aniso.py) into tests/golden/ as fixtures.  Run once in a
container that has the reference checkout; the GPU box uses the copies.

  python tests/golden/collect_reference_fixtures.py /root/reference
"""

import os
import shutil
import sys

FILES = [
    "test_data/north/corrDD.dat",
    "test_data/north/corrRR.dat",
    "test_data/north/nonperiodic/corrDD.dat.out",
    "test_data/north/nonperiodic/corrRR.dat.out",
    "test_data/north/nonperiodic/xi.txt",
    "test_data/south/corrDD.dat",
    "test_data/south/corrRR.dat",
    "test_data/south/nonperiodic/corrDD.dat.out",
    "test_data/south/nonperiodic/corrRR.dat.out",
    "test_data/south/nonperiodic/xi.txt",
]


def main(ref_root):
    here = os.path.dirname(os.path.abspath(__file__))
    for rel in FILES:
        dst = os.path.join(here, rel.replace("test_data/", "", 1))
        os.makedirs(os.path.dirname(dst), exist_ok=True)
        shutil.copyfile(os.path.join(ref_root, rel), dst)
        print("copied", rel)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
