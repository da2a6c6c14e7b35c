"""Build libphx.so (HIP, gfx950) in-tree: ``python parahpop_amd/build.py``."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs=8, verbose=False, force=True):
    """Compile libphx.so for gfx950; force recompiles every source even when
    the objects are newer than their sources."""
    csrc = os.path.join(HERE, "csrc")
    cmd = ["make", "-C", csrc, "-j%d" % jobs] + (["-B"] if force else [])
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if verbose or out.returncode != 0:
        sys.stdout.write(out.stdout)
    if out.returncode != 0:
        raise RuntimeError("libphx build failed")
    return os.path.join(HERE, "libphx.so")


if __name__ == "__main__":
    print(build(verbose=True))
