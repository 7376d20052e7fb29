"""Builds the oracle's C part (ghost.c) into oracle/_build/libghost_oracle.so.

TEST INFRASTRUCTURE (see oracle/__init__.py). Plain gcc, contraction off so
the fixed exp/log sequences are evaluated exactly as written.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "libghost_oracle.so")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-std=c11",
           os.path.join(HERE, "ghost.c"), "-o", OUT + ".tmp", "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build())
