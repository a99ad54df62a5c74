"""Build recipe: compiles libftdx.so / libftdx_api.so (hipcc, gfx950) in place,
the C++ API test program, and the CPU oracle used by the tests."""
from __future__ import annotations

import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)


def _make(path: str, jobs: int = 8, force: bool = True) -> None:
    if force:
        subprocess.run(["make", "-s", "-C", path, "clean"], check=True)
    subprocess.run(["make", "-s", f"-j{jobs}", "-C", path], check=True)


def build_all(force: bool = True) -> None:
    _make(PKG, force=force)
    _make(os.path.join(ROOT, "oracle"), force=force)
    cpp = os.path.join(ROOT, "tests", "cpp")
    if os.path.exists(os.path.join(cpp, "Makefile")):
        _make(cpp, force=force)


if __name__ == "__main__":
    build_all()
