"""Build the CPU oracle (liboracle.so) with gcc.

-ffp-contract=off keeps every a*b+c a separate multiply and add, matching the
gfx950 build of the shared arithmetic header bit for bit.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build():
    src = os.path.join(HERE, "kaos_oracle.c")
    out = os.path.join(HERE, "liboracle.so")
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-std=c11",
           "-Wall", "-Wno-unused-function", src, "-o", out, "-lm"]
    subprocess.run(cmd, check=True)
    return out


if __name__ == "__main__":
    print(build())
