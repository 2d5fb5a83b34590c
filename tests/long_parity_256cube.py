"""Long GPU-vs-oracle parity run at the 3D benchmark size (not collected by pytest; minutes of
oracle time): 256^3 periodic, 1000 Voronoi grains (one-cell overlap start), isotropic, FP64, 16
slots. At each checkpoint the device state is compared with the oracle's: index sets must be
equal and values within 1e-12 (reported: the number of values that differ at all, the largest
difference, grain volumes and VolumeRatios).

    python tests/long_parity_256cube.py [steps] [threads]      (default 500 steps, 16 threads)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))

import oracle as O  # noqa: E402
from openphase_amd.engine import Engine, StepConfig  # noqa: E402
from openphase_amd.microstructure import voronoi_grains  # noqa: E402


def main():
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 500
    threads = int(sys.argv[2]) if len(sys.argv) > 2 else 16
    checks = sorted({c for c in (1, 20, 100, 250, 500, 1000) if c <= steps} | {steps})
    cfg = StepConfig(256, 256, 256, slots=16)
    lists = voronoi_grains(256, 256, 256, 1000, seed=0x0FE2, sides="plus")
    phase = np.zeros(1000, np.int32)
    dt = cfg.max_time_step()
    eng, orc = Engine(cfg), O.Oracle(cfg, threads=threads)
    for e in (eng, orc):
        e.set_grains(phase)
        e.upload(lists)
        e.finalize_initialization()
    print(f"256^3, 1000 grains, dt {dt:.6e}, oracle on {threads} threads", flush=True)
    done, t0, ok = 0, time.perf_counter(), True
    for c in checks:
        while done < c:
            n = min(25, c - done)
            orc.step(dt, n)
            eng.step(dt, n)
            done += n
            print(f"  step {done} ({time.perf_counter() - t0:.0f} s)", flush=True)
        a, b = eng.download(), orc.download()
        same_sets = np.array_equal(a.count, b.count) and np.array_equal(a.index, b.index)
        line = f"step {c}: index sets {'equal' if same_sets else 'DIFFER'}"
        if same_sets:
            d = np.abs(a.value - b.value)
            line += f", values: {int((d > 0).sum())} of {d.size} differ, max {d.max():.3e}"
            ok = ok and d.max() <= 1e-12
        else:
            ok = False
        (gv, gr), (ov, orr) = eng.grains(), orc.grains()
        line += f"; grain volumes max rel {np.max(np.abs(gv - ov) / np.maximum(ov, 1e-300)):.2e}, " \
                f"VolumeRatios equal: {np.array_equal(gr, orr)}"
        print(line, flush=True)
    eng.close()
    orc.close()
    print("PASS" if ok else "FAIL", flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
