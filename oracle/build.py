"""ORACLE (test infrastructure): build liboracle.so with gcc: ``python oracle/build.py``."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(verbose=False):
    out = subprocess.run(["make", "-C", HERE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if verbose or out.returncode != 0:
        sys.stdout.write(out.stdout)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed")
    return os.path.join(HERE, "liboracle.so")


if __name__ == "__main__":
    print(build(verbose=True))
