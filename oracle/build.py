"""Builds the oracle's C restatement (test infrastructure only) into oracle/_build/."""
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent


def build():
    out = HERE / "_build"
    out.mkdir(exist_ok=True)
    lib = out / "libprescreen_oracle.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-o", str(lib),
                    str(HERE / "prescreen.c"), "-lm"], check=True)
    return lib


if __name__ == "__main__":
    print(build())
