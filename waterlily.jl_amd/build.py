"""Compile libwaterlily_hip.so for gfx950 (hipcc) in place."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(force=True):
    csrc = os.path.join(HERE, "csrc")
    args = ["make", "-s", "-C", csrc]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)
    return os.path.join(HERE, "libwaterlily_hip.so")
