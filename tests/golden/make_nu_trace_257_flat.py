"""Writes tests/golden/nu_trace_257_flat.npz: the checks (step count, wall
time, hot-wall Nu) of one cold run of the 257 x 257 cavity
sample_cavities(8, 7, n=257)[0] (Ra 7.5e4, Pr 5.72, A 2.49) from rest to
the stopping rule (eps 1e-6, checks every 2000 steps), and its converged Nu.
Its Nu flattens for about 20 checks between two modes on its way to the
steady value, which makes it a regression case for the Nusselt stopping
rule (surrogate.settle_entry).  Needs the GPU library."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "mpi-multi-process-version_amd"))
import surrogate as S  # noqa: E402


def main():
    cav = S.sample_cavities(8, 7, n=257)[0]
    run = S.traced_run(cav, None, itc_max=2000000)
    np.savez_compressed(os.path.join(HERE, "nu_trace_257_flat.npz"), trace=run["trace"],
                        nu_converged=run["nu"], cavity=json.dumps(cav))


if __name__ == "__main__":
    main()
