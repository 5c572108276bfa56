"""Write tests/golden/checker_fixture.npz from the reference's checker example.

Inputs (data files of the reference): example/checker/model_true/*.bin,
sources.dat, stations.dat and the golden SH velocity traces
example/checker/traces/vy_{000000,000012,000024}.su (every 8th receiver,
plus the first receiver's 240-byte SU header).  Run once where the reference
checkout exists:  python tests/golden/make_checker_fixtures.py [REF_ROOT]
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "example" / "checker"
NT = 5000
SHOTS = (0, 12, 24)
REC_STEP = 8


def read_bin(p):
    raw = p.read_bytes()
    npt = int(np.frombuffer(raw[:4], np.int32)[0])
    return np.frombuffer(raw[4:4 + 4 * npt], np.float32).copy()


def read_su(p):
    raw = np.fromfile(p, np.uint8)
    rec = 240 + 4 * NT
    a = raw.reshape(-1, rec)
    return a[:, :240].copy(), a[:, 240:].copy().view(np.float32)


def main():
    out = {}
    for c in ("x", "z", "vp", "vs", "rho"):
        out[f"model_{c}"] = read_bin(ref / "model_true" / f"proc000000_{c}.bin")
    out["sources_txt"] = np.frombuffer((ref / "sources.dat").read_bytes(), np.uint8)
    out["stations_txt"] = np.frombuffer((ref / "stations.dat").read_bytes(), np.uint8)
    out["shots"] = np.array(SHOTS, np.int32)
    out["rec_step"] = np.int32(REC_STEP)
    for s in SHOTS:
        hdr, data = read_su(ref / "traces" / f"vy_{s:06d}.su")
        out[f"vy_{s}"] = data[::REC_STEP].copy()
        out[f"hdr_{s}"] = hdr[:2].copy()
    np.savez_compressed(HERE / "checker_fixture.npz", **out)
    print("wrote", HERE / "checker_fixture.npz")


if __name__ == "__main__":
    main()
