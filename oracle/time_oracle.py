"""Rates of the oracle on the two CPU cases whose reference rates were timed
on the same host: the reference's published FD case (150^3, radius 4,
20-point sponge) and its SEM case `50 2000` (50^3 elements of order 2,
1,030,301 nodes), both with THREADS OpenMP threads (default 8).

TEST INFRASTRUCTURE: the oracle is timed here only to state how fast the
restatement runs next to the reference on one machine (the ratio bench.py
reports beside its CPU baselines).  Usage:

    python3 oracle/time_oracle.py [THREADS]
"""
import os
import sys
import time

THREADS = int(sys.argv[1]) if len(sys.argv) > 1 else 8
os.environ["OMP_NUM_THREADS"] = str(THREADS)  # read when liboracle loads
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import oracle  # noqa: E402


def fd_rate(steps=40):
    P = oracle.FdProblem(nx=150, nsamples=steps + 2)
    P.run(2)
    t0 = time.perf_counter()
    P.run(steps)
    return 150 ** 3 * steps / (time.perf_counter() - t0) / 1e6


def sem_rate(steps=4):
    P = oracle.SemProblem(ex=50, lx=2000.0, order=2, dt=1e-3, f0=5.0, nsamples=steps + 2)
    P.set_threads(THREADS)
    P.run(1)
    t0 = time.perf_counter()
    P.run(steps)
    return P.nN * steps / (time.perf_counter() - t0) / 1e6


if __name__ == "__main__":
    print("threads %d: FD 150^3 %.1f Mpoint-updates/s; SEM 50^3 order 2 %.2f Mnode-updates/s"
          % (THREADS, fd_rate(), sem_rate()), flush=True)
