"""Writes tests/golden/velocity_filter_example.npz: the oracle's VelocityFilter::velocityFilter on the
reference's example input (example/mesh1.obj -> example/mesh2.obj, kept here as velocity_filter_start.obj /
velocity_filter_end.obj) with the parameters of example/testVelocityFilter.cpp:100-107 (unit inverse
masses, outerEta 2e-8, innerEta 1e-8, default stiffness, substep, CoR and rollbacks).

The oracle runs with the reference's own root finder (oracle/_ref, built from src/rpoly.h) when it is
available, so the file holds the reference's result: the return value, the filtered end positions and
the run's counters.  Takes about two minutes.  Run from the repository root:
    python tests/golden/make_velocity_filter_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def load_obj(path):
    """'v x y z' and 'f a b c' lines (1-based indices) -> (qstart (3V,), faces (3F,) int32)."""
    v, f = [], []
    with open(path) as fh:
        for line in fh:
            t = line.split()
            if t and t[0] == "v":
                v.extend(float(x) for x in t[1:4])
            elif t and t[0] == "f":
                f.extend(int(x) - 1 for x in t[1:4])
    return np.array(v, dtype=np.float64), np.array(f, dtype=np.int32)


def main():
    oracle.build()
    if not oracle.ref_rpoly_available():
        raise SystemExit("oracle/_ref is not built: the golden needs the reference's root finder")
    oracle.use_rootfinder("rpoly")
    q1, f = load_obj(os.path.join(HERE, "velocity_filter_start.obj"))
    q2, f2 = load_obj(os.path.join(HERE, "velocity_filter_end.obj"))
    assert (f == f2).all()
    im = np.ones_like(q1)
    r, qout, st, closest = oracle.velocity_filter(q1, q2, f, im, 2e-8, 1e-8)
    np.savez(os.path.join(HERE, "velocity_filter_example.npz"), result=np.int64(r), qend=qout,
             closest=np.float64(closest), **{k: np.int64(v) for k, v in st.items()})
    print("result", r, "closest", closest, st)


if __name__ == "__main__":
    main()
