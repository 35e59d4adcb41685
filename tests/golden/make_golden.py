"""Writes the golden fixtures of tests/golden/ (run from the repository root).

* tiny_graphs.json -- complete target arrays of small graphs from the
  pure-Python restatement in oracle/oracle.py (independent of the C oracle).
* n1e6_d4_p0.5_digest.json -- SHA-256 of the n=1e6, d=4, p=0.5, seed=42 target array
  from the C oracle, plus sampled row slices.
* ref_dd_n<N>_d<D>_p<P>.csv -- "Degree,Count" distributions written by the
  reference's own serial copy model (src/PA/SPA/SerialCopyModel3.cpp built by
  `make -C oracle ref`; MT19937-64 draws with its default seed), used for
  statistical parity of our degree distributions.
* ref_dd_ba_n<N>_d<D>.csv -- the same from the reference's serial
  Barabasi-Albert generator (src/PA/SPA/SerialBA.cpp, also built by
  `make -C oracle ref`).
"""
import hashlib
import json
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

TINY = [(2, 1, 0.5, 0), (5, 2, 0.5, 1), (12, 3, 0.3, 2), (20, 4, 0.0, 3), (24, 4, 1.0, 4),
        (32, 5, 0.5, 5), (32, 8, 0.2, 6), (30, 16, 0.6, 7), (32, 1, 0.5, 2**40 + 1)]
REF_DD = [(1_000_000, 4, 0.5), (1_000_000, 8, 0.2), (300_000, 16, 0.5), (300_000, 2, 0.9)]
REF_DD_BA = [(1_000_000, 1), (1_000_000, 4), (300_000, 16)]


def main():
    tiny = []
    for n, d, p, seed in TINY:
        T = oracle.py_p_threshold(p)
        tiny.append({"n": n, "d": d, "p": p, "seed": seed, "p_threshold": T,
                     "targets": oracle.py_generate(n, d, T, seed).tolist()})
    with open(os.path.join(HERE, "tiny_graphs.json"), "w") as f:
        json.dump(tiny, f)

    n, d, p, seed = 1_000_000, 4, 0.5, 42
    G, _, retries = oracle.generate(n, d, oracle.p_threshold(p), seed)
    slices = {str(s): G[s * d:(s + 16) * d].tolist() for s in (5, 1000, 123_457, 500_000, n - 16)}
    with open(os.path.join(HERE, "n1e6_d4_p0.5_digest.json"), "w") as f:
        json.dump({"n": n, "d": d, "p": p, "seed": seed, "retries": retries,
                   "sha256": hashlib.sha256(G.tobytes()).hexdigest(), "slices": slices}, f)

    if not os.path.exists(oracle.REF_BIN):
        print("reference build missing; skipping ref_dd fixtures (make -C oracle ref)")
        return
    for n, d, p in REF_DD:
        with tempfile.TemporaryDirectory() as tmp:
            oracle.run_reference_scm(n, d, p, tmp)
            shutil.copy(os.path.join(tmp, "scm_dd.csv"), os.path.join(HERE, f"ref_dd_n{n}_d{d}_p{p}.csv"))
    for n, d in REF_DD_BA:
        with tempfile.TemporaryDirectory() as tmp:
            oracle.run_reference_sba(n, d, tmp)
            shutil.copy(os.path.join(tmp, "sba_dd.csv"), os.path.join(HERE, f"ref_dd_ba_n{n}_d{d}.csv"))


if __name__ == "__main__":
    main()
