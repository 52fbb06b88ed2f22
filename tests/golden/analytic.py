"""Analytic packings used as certifier fixtures, built from the formulas of the
reference's generators (mpmath, 60+ digits):

  tilted_five   src/squares/optimizer.py:13-36  four corners + a central diamond,
                centres dilated by 1.000000001
  goebel(a)     cuda/scheduler.py:54-74  staircase corners + 45-degree strip,
                L = a + 1 + 1/sqrt(2), centres dilated by 1 + 1e-40
  strip(k)      cuda/scheduler.py:25-52  tilted five in the lower-left plus unit
                strips at pitch 1.000000001 along the top and the right
  grid(n)       src/squares/core.py:27-35
"""
from __future__ import annotations

import math


def grid(n):
    width = math.isqrt(n - 1) + 1
    return [(f"{i % width}.5", f"{i // width}.5", "0") for i in range(n)]


def tilted_five():
    import mpmath as mp
    with mp.workdps(60):
        side = 2 + 1 / mp.sqrt(2)
        f = mp.mpf("1.000000001")
        half = mp.mpf(".5")
        pts = [(half, half, 0), (side - half, half, 0), (half, side - half, 0), (side - half, side - half, 0),
               (side / 2, side / 2, mp.pi / 4)]
        return [(mp.nstr(x * f, 55), mp.nstr(y * f, 55), mp.nstr(a, 55)) for x, y, a in pts]


def goebel(a):
    import mpmath as mp
    with mp.workdps(65):
        h = 1 / mp.sqrt(2)
        L = a + 1 + h
        dil = mp.mpf("1.0000000000000000000000000000000000000001")
        rows = []

        def add(x, y, ang=0):
            rows.append((mp.nstr(x * dil, 60), mp.nstr(y * dil, 60), mp.nstr(ang, 60)))

        for i in range(a):
            for j in range(a - i):
                x, y = i + mp.mpf(".5"), j + mp.mpf(".5")
                add(x, y)
                add(L - x, L - y)
        add(mp.mpf(".5"), L - mp.mpf(".5"))
        add(L - mp.mpf(".5"), mp.mpf(".5"))
        for j in range(1 + int(mp.floor((a - 1) * mp.sqrt(2)))):
            c = 1 + h / 2 + j * h
            add(c, L - c, mp.pi / 4)
        return rows


def strip(k, base_L: str):
    """base_L: certified side of tilted_five() as a decimal string."""
    import mpmath as mp
    with mp.workdps(65):
        rows = list(tilted_five())
        L = mp.mpf(base_L)
        t = k - 2
        pitch = mp.mpf("1.000000001")

        def add(x, y):
            rows.append((mp.nstr(x, 60), mp.nstr(y, 60), "0"))

        for row in range(t):
            for col in range(2):
                add((col + mp.mpf(".5")) * pitch, L + (row + mp.mpf(".5")) * pitch)
        for col in range(t):
            for row in range(k):
                add(L + (col + mp.mpf(".5")) * pitch, (row + mp.mpf(".5")) * pitch)
        return rows
