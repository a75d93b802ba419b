"""Copies pSi3D's sample input files (MIT licence, see the reference's LICENSE)
from the reference tree into tests/golden/psi3d/ as fixtures of the input
readers (si3d_amd/inputs.py).  Data only; run once where the reference tree
is present:

    python tests/golden/make_psi3d_inputs.py /root/reference

The met series sample_files/si3d_surfbc.txt (8917 records) and the Clear
Lake met-station series CLWaterQuality/si3d_surfbc.txt (7 stations, 2252
records) are cut to their first 48 records, with the record count of their
header line rewritten to 48; every other file is copied unchanged."""
import os
import shutil
import sys

CASES = {
    "sample_files": ["si3d_inp.txt", "si3d_bathy", "si3d_layer.txt", "si3d_init.txt", "si3d_surfbc.txt"],
    "windDrivenFlow": ["si3d_inp.txt", "h", "si3d_init.txt", "readme.txt"],
    "CLWaterQuality": ["si3d_inp.txt", "si3d_layer.txt", "si3d_init.txt", "si3d_surfbc.txt"],
}
SRC = {"sample_files": "sample_files", "windDrivenFlow": "sample_runs/windDrivenFlow",
       "CLWaterQuality": "sample_runs/CLWaterQuality"}
KEEP = 48


def main(ref):
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "psi3d")
    for case, files in CASES.items():
        dst = os.path.join(here, case)
        os.makedirs(dst, exist_ok=True)
        for f in files:
            src = os.path.join(ref, SRC[case], f)
            if f == "si3d_surfbc.txt":
                with open(src) as fi:
                    lines = fi.readlines()
                nh = 7 if case == "sample_files" else 9  # the met-station file has 2 more header records
                head, rec = lines[:nh], lines[nh:nh + KEEP]
                head[-1] = head[-1].replace(head[-1].split("=")[1].strip(), str(KEEP).rjust(4))
                with open(os.path.join(dst, f), "w") as fo:
                    fo.writelines(head + rec)
            else:
                shutil.copyfile(src, os.path.join(dst, f))
    shutil.copyfile(os.path.join(ref, "LICENSE"), os.path.join(here, "LICENSE"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
