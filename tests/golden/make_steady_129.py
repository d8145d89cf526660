"""Writes tests/golden/steady_2d_129_ra1e4.npz: the FP64 oracle's steady
state for the 129x129, Ra=1e4, Pr=0.71, A=3.2 cavity (run to both residuals
<= 1e-8, checked every 2000 steps), with the residual history and the mean
Nusselt numbers.  The reference reached 1e-6 at itc=392000 and 1e-8 at
itc=536000 on this case, with hot/cold Nu 2.2421/2.2422 on its steady field.
Takes ~20 minutes on 6 threads.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import lbm_oracle as lo  # noqa: E402


def main(threads=6):
    lo.set_threads(threads)
    o = lo.Oracle(2, 129, 1e4)
    hist = []
    itc1e6 = None
    t0 = time.time()
    while True:
        o.step(2000)
        eu, et = o.check()
        hist.append((o.itc, eu, et))
        if itc1e6 is None and eu <= 1e-6 and et <= 1e-6:
            itc1e6 = o.itc
        print(o.itc, eu, et, f"{time.time() - t0:.0f}s", flush=True)
        if (eu <= 1e-8 and et <= 1e-8) or o.itc >= 1_000_000:
            break
    rho, u, v, w, T = o.fields()
    nh, nc = o.nusselt()
    np.savez_compressed(os.path.join(HERE, "steady_2d_129_ra1e4.npz"),
                        u=u, v=v, T=T, rho=rho, history=np.array(hist),
                        itc=o.itc, itc_1e6=itc1e6, nu_hot=nh, nu_cold=nc)
    print("done itc", o.itc, "itc(1e-6)", itc1e6, "Nu", nh, nc)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 6)
