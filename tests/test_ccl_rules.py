"""The two union-find rules of the GPU labelling (lsmpqs-1.0_amd/csrc/lsm_ccl.hip),
restated sequentially in numpy and checked against scipy.ndimage.label on random
binary volumes, 26- and 6-connectivity:

* local pass: every foreground voxel starts at the end of its x run, and a union
  with a diagonal neighbour is skipped when it is implied through a foreground
  x (or y) neighbour; any union (l, m) is skipped when l - 1 and m - 1 (same
  tile row) are both foreground: l - 1 takes the same-direction pair
  (l - 1, m - 1), and the x runs join l - 1 to l and m - 1 to m;
* cross-tile pass: a pair of neighbouring voxels in different tiles is united
  only by the voxel whose neighbour's tile is lexicographically larger in
  (z, y, x) -- each such voxel lies on a +x/+y/+z tile face.

Labels must equal the largest linear index of each component (the reference's
rule, src/connectivity/connectivity.c:66-83).  Union order is shuffled: the
GPU runs the unions concurrently.  Planted errors in either rule must fail.
"""
import itertools

import numpy as np
import pytest

ndi = pytest.importorskip("scipy.ndimage")


def _label(fg, tiles, conn, skip_bug=None, seed=0):
    CX, CY, CZ = tiles
    nz, ny, nx = fg.shape
    flat = fg.reshape(-1)
    n = flat.size
    P = -np.ones(n, dtype=np.int64)

    def find(x):
        q = P[x]
        while q != x:
            g = P[q]
            if g != q:
                P[x] = g
            x, q = q, g
        return x

    def unite(a, b):
        a, b = find(a), find(b)
        if a == b:
            return
        if (a > b) != (skip_bug == "link"):
            a, b = b, a
        P[a] = b

    at = lambda i, j, k: i + nx * (j + ny * k)
    # local pass, tile by tile: run initialisation
    for k in range(nz):
        for j in range(ny):
            i = 0
            while i < nx:
                if not fg[k, j, i]:
                    i += 1
                    continue
                e = i
                while e + 1 < nx and fg[k, j, e + 1] and (e + 1) // CX == i // CX:
                    e += 1
                for t in range(i, e + 1):
                    P[at(t, j, k)] = at(t if skip_bug == "runs" else e, j, k)
                i = e + 1
    rng = np.random.default_rng(seed)
    vox = np.argwhere(fg)
    rng.shuffle(vox)
    for k, j, i in vox:
        l = at(i, j, k)
        lx, ly = i % CX, j % CY
        xm = lx > 0 and fg[k, j, i - 1]
        xp = lx < CX - 1 and i + 1 < nx and fg[k, j, i + 1]
        ym = ly > 0 and fg[k, j - 1, i]
        yp = ly < CY - 1 and j + 1 < ny and fg[k, j + 1, i]
        for dk in (0, 1):
            for dj in ((-1, 0, 1) if dk else (0, 1)):
                for di in ((-1, 0, 1) if (dk or dj) else (1,)):
                    if dk == 0 and dj == 0:
                        continue
                    if conn == 6 and abs(di) + abs(dj) + abs(dk) != 1:
                        continue
                    if conn != 6:
                        if (di == 1 and xp) or (di == -1 and xm):
                            continue
                        if di == 0 and dk == 1 and ((dj == 1 and yp) or (dj == -1 and ym)):
                            continue
                    i2, j2, k2 = i + di, j + dj, k + dk
                    if not (0 <= i2 < nx and 0 <= j2 < ny and k2 < nz):
                        continue
                    if (i2 // CX, j2 // CY, k2 // CZ) != (i // CX, j // CY, k // CZ):
                        continue
                    if not fg[k2, j2, i2]:
                        continue
                    if xm and i2 % CX >= 1 and (fg[k2, j2, i2 - 1] or skip_bug == "prevx"):
                        continue
                    unite(l, at(i2, j2, k2))
    T = np.array([find(l) if flat[l] else -1 for l in range(n)])
    # cross-tile pass
    offs = [o for o in itertools.product((-1, 0, 1), repeat=3) if o != (0, 0, 0)]
    for k, j, i in vox:
        l = at(i, j, k)
        for dk, dj, di in offs:
            if conn == 6 and abs(di) + abs(dj) + abs(dk) != 1:
                continue
            i2, j2, k2 = i + di, j + dj, k + dk
            if not (0 <= i2 < nx and 0 <= j2 < ny and 0 <= k2 < nz) or not fg[k2, j2, i2]:
                continue
            o = (k2 // CZ - k // CZ, j2 // CY - j // CY, i2 // CX - i // CX)
            take = o > (0, 0, 0)       # tuple order = (z, y, x) lexicographic
            if skip_bug == "border":
                take = take and o[2] == 0
            if take:
                unite(T[l], T[at(i2, j2, k2)])
    return np.array([find(l) if flat[l] else -1 for l in range(n)]).reshape(fg.shape)


def _expected(fg, conn):
    ref, nc = ndi.label(fg, ndi.generate_binary_structure(3, 3 if conn == 26 else 1))
    idx = np.arange(fg.size).reshape(fg.shape)
    exp = -np.ones(fg.shape, dtype=np.int64)
    for c in range(1, nc + 1):
        exp[ref == c] = idx[ref == c].max()
    return exp


CASES = [((9, 13, 21), (8, 4, 4), 0.3, 26), ((10, 11, 19), (8, 4, 4), 0.55, 26),
         ((9, 9, 17), (8, 4, 4), 0.4, 6), ((12, 10, 20), (4, 4, 2), 0.7, 26),
         ((20, 18, 70), (64, 16, 16), 0.45, 26), ((18, 17, 66), (64, 16, 16), 0.35, 6),
         ((12, 14, 40), (16, 8, 4), 0.85, 26), ((12, 14, 40), (16, 8, 4), 0.8, 6)]


@pytest.mark.parametrize("shape,tiles,p,conn", CASES)
def test_labelling_rules_match_scipy(shape, tiles, p, conn):
    fg = np.random.default_rng(hash((shape, p)) % 2**32).random(shape) < p
    assert (_label(fg, tiles, conn) == _expected(fg, conn)).all()


@pytest.mark.parametrize("bug,conn", [("link", 26), ("runs", 26), ("border", 26),
                                      ("prevx", 6)])
def test_planted_rule_errors_are_caught(bug, conn):
    """Larger root linked under the smaller (labels become minima); runs not
    initialised although +x unions are still skipped; pairs crossing an x tile
    face never united; the x-previous rule applied without checking that
    m - 1 is foreground (26-connectivity mostly finds another path)."""
    fg = np.random.default_rng(11).random((10, 12, 20)) < 0.3
    assert not (_label(fg, (8, 4, 4), conn, skip_bug=bug) == _expected(fg, conn)).all()
