"""Writes the golden fixtures of tests/golden/ from the oracle (FP64 CPU
restatement), after the oracle has met its pins (tests/test_oracle.py):

  rhs_2d_p4.npz    D=2, P=4, E=2, roots=2 periodic: seeded state y, RHS du
  rhs_ball_p1.npz  seven-patch ball P=1, E=2, roots=(2,2), plane wave
                   k=(1.1,-0.7,0.5), phi=0.3 as Dirichlet data at t=0.25:
                   seeded state y, RHS du, truncated y
  results.json     wave_errors / reflect_errors / two-level / sevenpatch_errors
                   results (l2, linf, energies, step counts)

Run from the repository root: python tests/golden/gen_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def main():
    rng = np.random.default_rng(20261203)
    pr = O.Problem.brick(2, 4, 2, roots=2)
    y = rng.standard_normal(pr.size())
    np.savez_compressed(os.path.join(HERE, "rhs_2d_p4.npz"), y=y, du=pr.rhs(y))
    pb = O.Problem.ball(1, 2, roots=(2, 2))
    yb = rng.standard_normal(pb.size())
    np.savez_compressed(os.path.join(HERE, "rhs_ball_p1.npz"), y=yb, du=pb.rhs(yb, 0.25),
                        ytrunc=pb.truncate(yb))
    res = {}
    keys = ("l2", "l2u", "l2v", "linf", "energy0", "energy", "nsteps")
    r = O.wave_errors(2, 4, 2, roots=8)
    res["wave_d2_p4_r8"] = {k: r[k] for k in keys}
    r = O.wave_errors(2, 2, 2, roots=4, reflecting=True)
    res["reflect_d2_p2_r4"] = {k: r[k] for k in keys}
    r = O.wave_errors(2, 3, 2, roots=4, regions=[((0.25, 0.75),) * 2])
    res["twolevel_d2_p3_r4"] = {k: r[k] for k in keys}
    r = O.sevenpatch_errors(3, 2, roots=(2, 2))
    res["sevenpatch_p3_s1"] = {k: r[k] for k in keys}
    with open(os.path.join(HERE, "results.json"), "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
