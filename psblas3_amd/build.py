"""In-tree build of libpsb_hip.so for gfx950 (hipcc, no JIT cache)."""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(_HERE, "csrc")


def build(jobs=8, force=False):
    """Compile every HIP source for gfx950 into psblas3_amd/lib/libpsb_hip.so."""
    args = ["make", "-s", "-C", CSRC, f"-j{jobs}"]
    if force:
        subprocess.check_call(["make", "-s", "-C", CSRC, "clean"])
    subprocess.check_call(args)
    return os.path.join(_HERE, "lib", "libpsb_hip.so")


if __name__ == "__main__":
    print(build())
