"""Writes tests/golden/mtx/*.mtx.xz: xz-compressed copies of the two
MatrixMarket files the reference holds as SpMV test data
(DLC/Cuper/data/matrices/{nasa4704,sit100}), byte for byte.  Run once in a
checkout next to the reference; the tests only read the compressed copies."""
from __future__ import annotations

import lzma
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent / "mtx"

for name in ("nasa4704", "sit100"):
    src = REF / "DLC" / "Cuper" / "data" / "matrices" / name / f"{name}.mtx"
    OUT.mkdir(exist_ok=True)
    data = src.read_bytes()
    (OUT / f"{name}.mtx.xz").write_bytes(lzma.compress(data, preset=9 | lzma.PRESET_EXTREME))
    print(f"{name}: {len(data)} bytes")
