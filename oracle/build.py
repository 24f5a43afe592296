"""Compile the oracle's C restatements into oracle/_build/ (test infrastructure only)."""
import os
import os.path as osp
import subprocess

HERE = osp.dirname(osp.abspath(__file__))
BUILD = osp.join(HERE, "_build")
LIB = osp.join(BUILD, "libmc_oracle.so")


def build():
    os.makedirs(BUILD, exist_ok=True)
    src = osp.join(HERE, "mc_lewiner.c")
    cmd = ["gcc", "-O2", "-std=c99", "-fPIC", "-shared", "-o", LIB + ".tmp", src, "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build())
