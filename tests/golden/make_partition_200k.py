"""Writes tests/golden/partition_200k.npz: the oracle's full asynchronous-Gibbs partition (seed 0) of the
200,000-vertex workload graph (directed DCSBM, 500 lognormal(0.5) blocks, power-law out-degree 2.1 on [1, 2000]
with mean 20, p_in 0.6, generator seed 0): assignment, block count, description length and the run's counters.
The oracle takes a few minutes on 8 threads, so the GPU test compares against this file.

    python tests/golden/make_partition_200k.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402
from sbp_amd.graph import dcsbm  # noqa: E402

GRAPH = dict(N=200000, B=500, gamma=2.1, kmin=1, kmax=2000, mean_degree=20, p_in=0.6, sigma=0.5, seed=0)


def graph():
    k = dict(GRAPH)
    return dcsbm(k.pop("N"), k.pop("B"), **k)


def main():
    g = graph()
    t0 = time.perf_counter()
    a, nb, st = O.sbp(O.OGraph.from_csr(g), O.params("async_gibbs", seed=0, threads=min(16, os.cpu_count() or 1)))
    print(f"oracle: B={nb} mdl={st.mdl!r} in {time.perf_counter() - t0:.1f}s")
    np.savez_compressed(os.path.join(HERE, "partition_200k.npz"), assignment=a.astype(np.int16), num_blocks=nb,
                        mdl=st.mdl, merge_phases=st.merge_phases, sweeps=st.sweeps, moves=st.moves,
                        proposals=st.proposals, E=g.E)


if __name__ == "__main__":
    main()
