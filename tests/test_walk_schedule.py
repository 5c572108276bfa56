"""The walking three-step kernel's schedule (inv.cu_amd/csrc/fdm_walk.hip),
replayed on the CPU: eight rows per iteration, skews A 8, B 6, C 5, D 3,
E 2, F 0, LDS rings of 25 rows with slots 0-2 copied at 25-27, the stress
rows of the next iteration written while F still reads.  Checks, for chunks
at the top, middle and bottom of grids of several sizes:

  * every LDS read (velocity at row x: stress x-2 .. x+1 and szz at x;
    stress at x: velocity x-1 .. x+2 and its own stress) finds the right
    row at the right time level in the ring;
  * the ring writes that overlap phase F never touch a slot F reads;
  * each phase updates exactly the rows a three-step tile of the chunk's
    rows would (velocity n+1 on -7 .. +7 around the chunk, stress n+1 on
    -6 .. +5, ..., stress n+3 on the chunk);
  * row x of every phase belongs to wavefront x mod 8 (two register rows).

No counterpart in the reference (one launch per half step,
solver/fdm.h:534-551); the GPU parity tests (test_gpu_walk.py) check the
kernel's values.
"""
import pytest

R, RR, NW = 25, 28, 8
SKEW = {"A": 8, "B": 6, "C": 5, "D": 3, "E": 2, "F": 0}
LOW = {"A": -7, "B": -6, "C": -4, "D": -3, "E": -1, "F": 0}
HIGH = {"A": 8, "B": 6, "C": 5, "D": 3, "E": 2, "F": 0}


def replay(xa, ch, nx):
    xe = min(xa + ch, (nx + 7) & ~7)
    K = (xe - xa) // NW + 3

    def slot(x):
        return (x - xa + 5 * R) % R

    stress, vel = [None] * RR, [None] * RR

    def put(ring, x, lev):
        s = slot(x)
        ring[s] = (x, lev)
        if s < 3:
            ring[s + R] = (x, lev)

    def read(ring, x, lev, base):
        s = slot(base) + (x - base)
        assert 0 <= s < RR
        assert ring[s] == (x, lev), (x, lev, ring[s])

    cover = {p: [] for p in SKEW}
    for k in range(K):
        i0 = xa - 24 + NW * k
        puts = [i0 + 9 + w for w in range(NW) if i0 + 9 + w >= xa - 9]
        if k:
            f_rows = [i0 - NW + w for w in range(NW) if i0 - NW + w >= xa]
            busy = {slot(x) for x in f_rows}
            for x in puts:
                assert slot(x) not in busy, "ring write over a row F reads"
        for x in puts:
            put(stress, x, 0)
        for p in "ABCDEF":
            rows = []
            for w in range(NW):
                x = i0 + w if (p == "F" or (p != "A" and w >= SKEW[p])) else i0 + NW + w
                if x >= xa + LOW[p]:
                    rows.append(x)
            assert sorted(r % NW for r in rows) == sorted({r % NW for r in rows})
            lev = "ABCDEF".index(p) // 2
            for x in rows:
                if p in "ACE":
                    for d in (-2, -1, 0, 1):
                        read(stress, x + d, lev, x - 2)
                    put(vel, x, lev + 1)
                else:
                    read(stress, x, lev, x)
                    for d in (-1, 0, 1, 2):
                        read(vel, x + d, lev + 1, x - 1)
                    if p != "F":
                        put(stress, x, lev + 1)
            cover[p] += rows
    for p in SKEW:
        assert sorted(cover[p]) == list(range(xa + LOW[p], xe + HIGH[p])), p
    return K


@pytest.mark.parametrize("xa,ch,nx", [(0, 16, 70), (16, 16, 70), (64, 16, 70), (0, 64, 70),
                                      (0, 40, 37), (128, 128, 200), (0, 128, 4096),
                                      (1920, 128, 4096), (3968, 128, 4096), (0, 4096, 4096)])
def test_walk_schedule(xa, ch, nx):
    assert replay(xa, ch, nx) == (min(xa + ch, (nx + 7) & ~7) - xa) // NW + 3


def test_walk_schedule_catches_a_short_ring():
    """The same replay with a ring one row short fails: the check has teeth."""
    global R, RR
    saved = R, RR
    R, RR = 24, 27
    try:
        with pytest.raises(AssertionError):
            replay(0, 128, 4096)
    finally:
        R, RR = saved


def test_ring_size_matches_the_kernel():
    import re
    from pathlib import Path
    h = (Path(__file__).resolve().parents[1] / "inv.cu_amd" / "csrc" / "step.h").read_text()
    assert int(re.search(r"constexpr int WALK_RING = (\d+);", h).group(1)) == R
