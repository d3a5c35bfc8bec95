"""Build the C restatement of the oracle (TEST INFRASTRUCTURE ONLY):
oracle/c/lbm64.c -> oracle/_lib/liboracle64.so with gcc."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "c", "lbm64.c")
OUT = os.path.join(HERE, "_lib", "liboracle64.so")
OUT32 = os.path.join(HERE, "_lib", "liboracle32.so")


def build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cc = shutil.which("gcc") or "gcc"
    outs = []
    for out, real in ((OUT, "double"), (OUT32, "float")):
        tmp = out + ".tmp"
        # -ffp-contract=off: no fused multiply-adds, plain IEEE arithmetic
        subprocess.run([cc, "-O3", "-ffp-contract=off", "-fPIC", "-shared", "-DO64_REAL=" + real, SRC,
                        "-o", tmp, "-lm"], check=True)
        os.replace(tmp, out)
        outs.append(out)
    return outs


if __name__ == "__main__":
    print(build())
