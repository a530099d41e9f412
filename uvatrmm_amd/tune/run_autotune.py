"""Runs the FP64 tile-shape autotuner on GPU 0 and writes its JSON table.

    python uvatrmm_amd/tune/run_autotune.py [OUT.json]

The table maps (n, m) buckets to the fastest FP64 tile shape (128x128,
64x64 or 128x32) measured with the batch path's persistent kernel; the batch
path reads it through uvt_batch_create(..., tune_table=path).
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import uvatrmm_amd as uvt  # noqa: E402


def main():
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "f64_tiles.json")
    h = uvt.Handle(0)
    uvt.autotune(h, out)
    h.close()
    print("wrote", out, flush=True)


if __name__ == "__main__":
    main()
