"""Build libvespahip.so in-tree with hipcc for gfx950 (run from anywhere)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs: int = 4) -> str:
    csrc = os.path.join(HERE, "csrc")
    subprocess.run(["make", "-s", "-B", f"-j{jobs}", "-C", csrc], check=True)
    return os.path.join(HERE, "libvespahip.so")


if __name__ == "__main__":
    print(build(int(sys.argv[1]) if len(sys.argv) > 1 else 4))
