"""Build the C restatement of the oracle (TEST INFRASTRUCTURE ONLY):
oracle/c/lbm64.c -> oracle/_lib/liboracle64.so with gcc."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "c", "lbm64.c")
OUT = os.path.join(HERE, "_lib", "liboracle64.so")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cc = shutil.which("gcc") or "gcc"
    tmp = OUT + ".tmp"
    # -ffp-contract=off: no fused multiply-adds, plain IEEE double arithmetic
    subprocess.run([cc, "-O3", "-ffp-contract=off", "-fPIC", "-shared", SRC, "-o", tmp, "-lm"], check=True)
    os.replace(tmp, OUT)
    return [OUT]


if __name__ == "__main__":
    print(build())
