"""Pauli-string encodings (the reference's quantum.py:1025-1068)."""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence


def ps2xyz(ps: Sequence[int]) -> Dict[str, List[int]]:
    """[1, 2, 2, 0] -> {"x": [0], "y": [1, 2], "z": []} (quantum.py:1025-1044)."""
    xyz: Dict[str, List[int]] = {"x": [], "y": [], "z": []}
    for i, j in enumerate(ps):
        if j == 1:
            xyz["x"].append(i)
        if j == 2:
            xyz["y"].append(i)
        if j == 3:
            xyz["z"].append(i)
    return xyz


def xyz2ps(xyz: Dict[str, List[int]], n: Optional[int] = None) -> List[int]:
    """Inverse of ps2xyz (quantum.py:1047-1068)."""
    if n is None:
        n = max(xyz.get("x", []) + xyz.get("y", []) + xyz.get("z", [])) + 1
    ps = [0 for _ in range(n)]
    for i in range(n):
        if i in xyz.get("x", []):
            ps[i] = 1
        elif i in xyz.get("y", []):
            ps[i] = 2
        elif i in xyz.get("z", []):
            ps[i] = 3
    return ps
