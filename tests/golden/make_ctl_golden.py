"""Writes tests/golden/ctl/<case>.ctl with the reference's own
ctl_file_write (oracle/_ref/ctl_harness, built by make -C oracle ref) for
every case of ctl_cases.py.  Run from the repository root."""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from ctl_cases import CASES, write_case  # noqa: E402

HARNESS = os.path.join(os.path.dirname(HERE), "..", "oracle", "_ref", "ctl_harness")

if __name__ == "__main__":
    out = os.path.join(HERE, "ctl")
    os.makedirs(out, exist_ok=True)
    for c in CASES:
        with tempfile.TemporaryDirectory() as tmp:
            case = os.path.join(tmp, "case.txt")
            write_case(case, c, os.path.join(tmp, c["name"] + ".dat"))
            subprocess.run([os.path.abspath(HARNESS), case], check=True, cwd=tmp)
            shutil.copy(os.path.join(tmp, c["name"] + ".ctl"), os.path.join(out, c["name"] + ".ctl"))
        print("wrote", c["name"])
