"""Profiling driver for the batched soil-column pool: the C4 workload (4096 PressureIndependMultiYield
columns, N ~ U{20..400}, Newmark + Newton + PCG, 0.3 g base record) for a few steps, pool only.

    python femorax_amd/profile_columns.py STEPS [NCOL] [compact] [mixed] [w3|w4]
"""
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import bench  # noqa: E402
from femorax_amd import core  # noqa: E402
from femorax_amd import synthetic as S  # noqa: E402

if __name__ == "__main__":
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 10
    ncol = int(sys.argv[2]) if len(sys.argv) > 2 else 4096
    ctx = core.Context(0)
    print(json.dumps(bench.bench_c4(core, S, ctx, steps, ncol=ncol, cpu=False, plain=False,
                                    compact="compact" in sys.argv[3:], mixed="mixed" in sys.argv[3:],
                                    waves=3 if "w3" in sys.argv[3:] else 4 if "w4" in sys.argv[3:] else 1 if "w1" in sys.argv[3:] else 0)),
          flush=True)
    ctx.close()
