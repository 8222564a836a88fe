"""Compile every native piece in-tree: the HIP library (gfx950), the C++14
host drop-in layer and the FP64 oracle (test infrastructure)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _make(path, jobs=8):
    cmd = ["make", "-C", path, f"-j{jobs}"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stdout.write(r.stdout)
        raise RuntimeError(f"build failed in {path}")
    return r.stdout


def build_all(force=False, jobs=8):
    targets = [os.path.join(ROOT, "quetzalpic_amd", "csrc")]
    host = os.path.join(ROOT, "quetzalpic_amd", "host")
    if os.path.exists(os.path.join(host, "Makefile")):
        targets.append(host)
    targets.append(os.path.join(ROOT, "oracle"))
    for t in targets:
        if force:
            subprocess.run(["make", "-C", t, "clean"], stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        _make(t, jobs)


if __name__ == "__main__":
    build_all(force="--force" in sys.argv)
