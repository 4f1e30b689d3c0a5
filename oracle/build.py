"""Builds the oracle's C restatement (test infrastructure only) into oracle/_build/."""
import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent


def build():
    """Compiles into a private temporary file and renames it into place, so that test processes building at the
    same time (pytest -n) never load a half-written library."""
    out = HERE / "_build"
    out.mkdir(exist_ok=True)
    lib = out / "libprescreen_oracle.so"
    tmp = out / f"libprescreen_oracle.{os.getpid()}.tmp.so"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-o", str(tmp),
                    str(HERE / "prescreen.c"), "-lm"], check=True)
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build())
