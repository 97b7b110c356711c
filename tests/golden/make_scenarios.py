"""Copies the reference's scenario files into tests/golden/scenarios/.

    python tests/golden/make_scenarios.py [--reference /root/reference]

The scenario files (scenarios/*.json in the reference) are data: the inputs
that its headless --assert runner plays (tools/headless.cpp:179-250). They are
kept byte for byte, comments included, because the scenario reader must accept
them as they are. Run it in a checkout that has the reference; the tests read
the committed copies.
"""
from __future__ import annotations

import argparse
import shutil
from pathlib import Path

HERE = Path(__file__).resolve().parent


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    a = ap.parse_args()
    src = Path(a.reference) / "scenarios"
    dst = HERE / "scenarios"
    dst.mkdir(exist_ok=True)
    for f in sorted(src.glob("*.json")):
        shutil.copyfile(f, dst / f.name)
        print(dst / f.name)


if __name__ == "__main__":
    main()
