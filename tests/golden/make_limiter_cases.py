"""Writes tests/golden/limiter_cases.npz: the reference's PhaseField::NormalizeIncrements on
pair increments given directly (test infrastructure).

Random interface nodes (3-6 entries) get random pair increments spread over four orders of
magnitude, added with NodeAB::add_asym1 in random order and orientation; the reference driver
(oracle/_ref/ref_driver --normalize-only, built from the reference sources) limits them. Stored:
each node's finalized list, the injected pairs, the pairs the reference keeps (storage order,
limited values) and the plausibility warnings it prints per node. Many nodes reach the 25-pass cap
of the limiter; some end outside [0, 1] and are warned about.

    python tests/golden/make_limiter_cases.py
"""
from __future__ import annotations

import re
import struct
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

from openphase_amd import rawio  # noqa: E402
from openphase_amd.engine import StepConfig  # noqa: E402
from openphase_amd.microstructure import Lists, storage_shape  # noqa: E402
import oracle as O  # noqa: E402

HERE = Path(__file__).resolve().parent
N = 6           # 6 x 6 periodic 2D grid per run, every node an interface node
NGRAINS = 12
DT = 1e-4


def one_run(rng, td: Path):
    cfg = StepConfig(N, N, 0, slots=16)
    cnt = rng.integers(3, 7, size=N * N).astype(np.int32)
    idx, val = [], []
    for c in cnt:
        ids = rng.choice(NGRAINS, size=c, replace=False)
        v = rng.random(c) ** 4 + 1e-3
        v /= v.sum()
        idx += list(ids)
        val += list(v)
    lists = Lists(storage_shape(N, N, 0), cnt, np.array(idx, np.int32), np.array(val))
    # pairs: every pair of a node's own entries, random order and orientation
    pairs = []
    blob = [struct.pack("<I", N * N)]
    off = 0
    for c in range(N * N):
        own = idx[off:off + cnt[c]]
        off += cnt[c]
        pl = []
        for a in range(len(own)):
            for b in range(a + 1, len(own)):
                v = rng.normal() * 10.0 ** rng.uniform(-2, 2) / DT
                ia, ib = int(own[a]), int(own[b])
                if rng.random() < 0.5:
                    ia, ib, v = ib, ia, -v
                pl.append((ia, ib, v))
        rng.shuffle(pl)
        pairs.append(pl)
        blob.append(struct.pack("<II", c, len(pl)))
        for ia, ib, v in pl:
            blob.append(struct.pack("<IId", ia, ib, v))
    (td / "pairs.bin").write_bytes(b"".join(blob))
    O.write_opi(td / "input.opi", cfg, 1)
    rawio.write_phasefield(str(td / "state.dat"), lists)
    rawio.write_grains(str(td / "grains.bin"), np.zeros(NGRAINS, np.int32), None)
    cmd = [str(O.REF_DRIVER), str(td / "input.opi"), str(td / "state.dat"), str(td / "grains.bin"), "0", repr(DT),
           str(td / "out"), "0", "--threads", "1", "--normalize-only", str(td / "pairs.bin")]
    res = subprocess.run(cmd, capture_output=True, text=True, cwd=td, timeout=600)
    if res.returncode != 0:
        raise RuntimeError(res.stdout[-2000:] + res.stderr[-2000:])
    fin = rawio.read_phasefield(str(td / "out_0.dat"))
    # warnings per node: "... failed in point (i,j,k)"
    warn = np.zeros(N * N, np.int64)
    for m in re.finditer(r"failed in point \((\d+),(\d+),(\d+)\)", res.stderr):
        i, j, k = (int(x) for x in m.groups())
        warn[i * N + j] += 1
    out = (td / "out_pairs.bin").read_bytes()
    kept = []
    p = 0
    for c in range(N * N):
        (npr,) = struct.unpack_from("<I", out, p)
        p += 4
        pl = []
        for _ in range(npr):
            ia, ib, v = struct.unpack_from("<IId", out, p)
            p += 16
            pl.append((ia, ib, v))
        kept.append(pl)
    return fin, pairs, kept, warn


def main():
    rng = np.random.default_rng(0x11A1)
    rec = {"dt": DT, "runs": 6}
    for r in range(6):
        with tempfile.TemporaryDirectory() as td:
            fin, pairs, kept, warn = one_run(rng, Path(td))
        rec[f"r{r}_count"] = fin.count
        rec[f"r{r}_index"] = fin.index
        rec[f"r{r}_value"] = fin.value
        for name, pl in (("in", pairs), ("out", kept)):
            rec[f"r{r}_{name}_n"] = np.array([len(x) for x in pl], np.int32)
            flat = [q for x in pl for q in x]
            rec[f"r{r}_{name}_a"] = np.array([q[0] for q in flat], np.int32)
            rec[f"r{r}_{name}_b"] = np.array([q[1] for q in flat], np.int32)
            rec[f"r{r}_{name}_v"] = np.array([q[2] for q in flat], np.float64)
        rec[f"r{r}_warnings"] = warn
        print("run", r, "warnings", int(warn.sum()), "nodes warned", int((warn > 0).sum()))
    np.savez_compressed(HERE / "limiter_cases.npz", **rec)


if __name__ == "__main__":
    main()
