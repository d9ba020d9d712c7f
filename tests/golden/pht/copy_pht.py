"""Copies the reference's example trajectory files (data the reference ships, written by its own cache encoder)
into this folder as fixtures for tests/test_cache_files.py.  Run once in a tree that has the reference:
    python tests/golden/pht/copy_pht.py /root/reference"""
import os
import shutil
import sys

NAMES = ["wuhan_static", "wuhan_north", "wuhan_drive", "wuhan_segments", "wuhan_route"]

if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    for n in NAMES:
        shutil.copyfile(os.path.join(ref, "examples", "trajectories", n + ".pht"), os.path.join(here, n + ".pht"))
        print("copied", n + ".pht")
