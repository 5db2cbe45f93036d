"""Periodic-table position (row, column) per nuclear charge, the "PTP" table
of fml/kernels/kernels.py:41-66 (also fml/arad.py:43-68,
fml/math/distance.py:31-56), used for the ARAD type channels.

The table is kept as the same mapping, written as per-period column lists:
s-block columns 1-2, p-block 3-8, d-block 10-18 (9 for period 4/5 d-block
start), f-block 19-33.  Two literal quirks of the reference dictionary are
kept: key 101 appears twice and the later entry [7, 32] wins, and 102 maps
to [7, 14]."""


def _build():
    t = {1: (1, 1), 2: (1, 8)}
    # periods 2 and 3: Z0..Z0+7 -> columns 1..8
    for period, z0 in ((2, 3), (3, 11)):
        for k in range(8):
            t[z0 + k] = (period, k + 1)
    # periods 4 and 5: s (1,2), d (9..18), p (3..8)
    for period, z0 in ((4, 19), (5, 37)):
        t[z0] = (period, 1)
        t[z0 + 1] = (period, 2)
        for k in range(10):
            t[z0 + 2 + k] = (period, 9 + k)
        for k in range(6):
            t[z0 + 12 + k] = (period, 3 + k)
    # periods 6 and 7: s (1,2), f (19..33), d (10..18), p (3..8)
    for period, z0 in ((6, 55), (7, 87)):
        t[z0] = (period, 1)
        t[z0 + 1] = (period, 2)
        for k in range(15):
            t[z0 + 2 + k] = (period, 19 + k)
        for k in range(9):
            t[z0 + 17 + k] = (period, 10 + k)
        for k in range(6):
            t[z0 + 26 + k] = (period, 3 + k)
    t[101] = (7, 32)
    t[102] = (7, 14)
    return {z: list(v) for z, v in t.items()}


PTP = _build()
