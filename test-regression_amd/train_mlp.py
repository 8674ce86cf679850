"""Trains the warm-start network of a workload on the GPU and writes its
parameters (the files under weights/ are made by this script):

    python test-regression_amd/train_mlp.py --workload c2 --candidates 4 --out weights/mlp_c2.npy
    python test-regression_amd/train_mlp.py --workload c3 --candidates 4 --out weights/mlp_c3.npy

--workload names the benchmark workload the network serves; it is trained on
that workload's training distribution (c2: the c4 channels; c3: its own
mixture under a held-out seed).  With --candidates K it trains K networks
and keeps the one that predicts the exact passive set most often on
validation channels of their own seed.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import sys
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent


def _load_pkg():
    name = "test_regression_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, PKG / "__init__.py",
                                                  submodule_search_locations=[str(PKG)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--candidates", type=int, default=4)
    ap.add_argument("--first-seed", type=int, default=0,
                    help="seed of the first candidate (the others follow)")
    ap.add_argument("--n-train", type=int, default=4_000_000)
    ap.add_argument("--n-val", type=int, default=500_000)
    ap.add_argument("--epochs", type=int, default=400)
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    _load_pkg()
    from test_regression_amd import engine, warmstart, workloads

    def log(msg):
        print(msg, flush=True)

    w = workloads.config(args.workload)
    eng = engine.Engine(list(w.chtypes()))
    # trained in memory (path None: nothing is loaded or saved there); the
    # output file is written only once training has finished
    info = warmstart.ensure_trained(eng, warmstart.training_workload(args.workload),
                                    n_train=args.n_train, epochs=args.epochs, path=None,
                                    log=log, candidates=args.candidates, n_val=args.n_val,
                                    first_seed=args.first_seed)
    out = Path(args.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    np.save(out, eng.mlp_get_params())
    info.pop("report", None)
    info["out"] = str(out)
    print(json.dumps(info), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
