"""Ahead-of-time build of liblinac_amd.so for gfx950 (hipcc, in-tree)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")


def build(jobs=8, verbose=False, force=True):
    """Compile every HIP source into linac_amd/liblinac_amd.so.

    ``force`` rebuilds from scratch (the driver's "does it build" check).
    """
    if force:
        subprocess.run(["make", "-s", "-C", CSRC, "clean"], check=True)
    cmd = ["make", "-C", CSRC, f"-j{jobs}"]
    if not verbose:
        cmd.insert(1, "-s")
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout[-4000:] + res.stderr[-8000:])
        raise RuntimeError("building liblinac_amd.so failed")
    return os.path.join(HERE, "liblinac_amd.so")


if __name__ == "__main__":
    print(build(verbose=True))
