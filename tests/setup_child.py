"""Builds a host setup in a fresh process and pickles it to a file (test infrastructure).

A process whose HIP runtime is initialised must not fork, so the GPU tests build their large setups
here: this child process has no GPU context and runs the local eigenproblems on a pool of forked
workers. Usage: python setup_child.py <out.pkl> <json arguments>."""
import json
import os
import pickle
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from conftest import load_package  # noqa: E402

load_package()
from amp_amd import hostsetup as hs  # noqa: E402


def main():
    out, args = sys.argv[1], json.loads(sys.argv[2])
    A, g, xyz, sym, bs = hs.generate(args["name"], **args["kw"])
    b = hs.random_rhs(A.shape[0], args.get("seed", 12345)) if args.get("rhs", "random") == "random" else g
    st = hs.build_setup(A, b, args["N"], xyz=xyz, symmetric=sym, block_size=bs, eig_workers=args.get("workers", 16),
                        **args.get("opts", {}))
    with open(out, "wb") as f:
        pickle.dump(st, f, protocol=pickle.HIGHEST_PROTOCOL)


if __name__ == "__main__":
    main()
