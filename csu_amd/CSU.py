"""Drop-in driver with the reference's CLI (CSU.py:14-27):

    python3 csu_amd/CSU.py <case>/

removes the old executable and <case>/solution.csv, builds libcsu.so (host C +
gfx950 HIP kernels) and the executable, runs ``executable <case>/`` and then
``<case>/analisys.py <case>/`` when the case has one.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main(argv):
    if len(argv) < 2:
        print("Insert the input file")
        return 0
    path = argv[1]
    print("\nCSU - A CFD code for unstructured meshs:\n")
    from csu_amd import build
    exe = os.path.join(build.LIB, "executable")
    for f in (exe, os.path.join(path, "solution.csv")):
        if os.path.exists(f):
            os.remove(f)
    build.build()
    rc = subprocess.call([exe, path])
    post = os.path.join(path, "analisys.py")
    if os.path.exists(post):
        subprocess.call([sys.executable, post, path])
    return rc


if __name__ == "__main__":
    sys.exit(main(sys.argv))
