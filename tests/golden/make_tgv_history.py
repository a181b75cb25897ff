"""Writes tgv22p8_history.txt: the Taylor-Green kinetic-energy history of the
FP64 oracle at production size (22^3 hex8 elements, p = 8, Re = 1600,
Pr = 0.71, Mach 0.1, CFL 0.5, isothermal density, 200 steps), with the same
definitions as dsem_amd/tgv.py: Ek (turbk), -dEk/dt (first difference from
the initial Ek) and the enstrophy dissipation eps, one row per step.

The oracle is the CPU restatement of the reference algorithm (oracle/); this
fixture lets the GPU history be compared with it at full size without the
~30 CPU-minutes the oracle needs for it.

    python tests/golden/make_tgv_history.py [--threads 8] [--steps 200]
    python tests/golden/make_tgv_history.py --n 8 --p 4 --steps 8266 --every 20 \
        --out tests/golden/tgv8p4_history.txt     # to t = 10, through the dissipation peak
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O                      # noqa: E402
from dsem_amd import cases              # noqa: E402
from dsem_amd import mesh as M          # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=22)
    ap.add_argument("--p", type=int, default=8)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--every", type=int, default=1, help="keep every k-th row (and the last)")
    ap.add_argument("--out", default=os.path.join(HERE, "tgv22p8_history.txt"))
    a = ap.parse_args()
    re, pr, cfl = 1600.0, 0.71, 0.5
    ph = O.phys(re=re, pr=pr)
    mesh = M.periodic_box(a.n, a.p)
    g = O.Geometry(mesh)
    JQ = cases.tgv_state(g.xg, g.jacob, "isothermal")
    dt = cfl * O.dtcalc(g, JQ, ph)
    stag = np.zeros((mesh.K, 3, 5, g.ns))
    prev = O.kinetic_energy(g, JQ)
    rows = [(0, 0.0, prev, np.nan, np.nan)]
    t = 0.0
    t0 = time.time()
    for k in range(1, a.steps + 1):
        O.rk_step(g, JQ, dt, ph, nthreads=a.threads, stag_out=stag)
        ek = O.kinetic_energy(g, JQ)
        eps = O.enstrophy(g, JQ, stag, ph)
        t += dt
        rows.append((k, t, ek, (prev - ek) / dt, eps))
        prev = ek
        print(f"step {k} t={t:.6f} Ek={ek:.15f} dissp={rows[-1][3]:.8e} eps={eps:.8e} "
              f"({time.time() - t0:.0f} s)", flush=True)
    hdr = (f"FP64 oracle TGV history: n={a.n} p={a.p} Re={re} Pr={pr} CFL={cfl} dt={dt!r} variant=isothermal\n"
           "step time Ek -dEk/dt eps")
    rows = [r for r in rows if r[0] % a.every == 0 or r[0] == a.steps]
    np.savetxt(a.out, np.array(rows), fmt=["%d", "%.17e", "%.17e", "%.17e", "%.17e"], header=hdr)


if __name__ == "__main__":
    main()
