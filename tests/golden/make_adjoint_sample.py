"""Records the float64 fixed-contact model's answers (oracle/fixed_contacts.py) for the seeded sample of
the 1,024 adjoint throw worlds that the full-size GPU test checks: per sampled world its centres of mass
after 32 steps, the gradients of the squared distance to its seeded target with respect to the initial
velocities and the friction coefficients, and the two conditioning probes that decide whether the world
is compared at all (float32 drift of the velocities for the gradient, float32 rounding of the
separations for the forward). The model takes about 45 s a world in pure Python, far beyond what a GPU
test should spend on the host, so its answers are kept here; tests/test_adjoint.py checks one of them
against the model again on the CPU.

Run in the build container: python tests/golden/make_adjoint_sample.py
"""
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

OUT = os.path.join(HERE, "adjoint-throw1024-sample.npz")
WORLDS, STEPS, SAMPLE, SEED0 = 1024, 32, 12, 0xAD000000


def sample():
    rng = np.random.default_rng(0xAD)
    return sorted(rng.choice(WORLDS, size=SAMPLE, replace=False).tolist())


def answer(k):
    """The model's answers for world k of the workload, alone."""
    from adjoint_cases import Case
    case = Case([SEED0 + k], steps=STEPS)
    m = case.model
    mv = m.moving
    g1 = m.adjoint(case.target, steps=STEPS)
    g2 = m.adjoint(case.target, v0=m.v0 * (1 + 1e-5), steps=STEPS)
    gradient_drift = np.linalg.norm(g1[0][mv] - g2[0][mv]) / np.linalg.norm(g1[0][mv])
    com = m.forward(steps=STEPS)
    moved = np.linalg.norm(com - m.com0)
    keep = m.adj0.copy()
    m.adj0 = keep + 1e-6 * np.random.default_rng(k).standard_normal(keep.shape)
    probe = m.forward(steps=STEPS)
    m.adj0 = keep
    forward_drift = np.linalg.norm(probe - com) / moved
    print(f"world {k}: gradient drift {gradient_drift:.2e}, forward drift {forward_drift:.2e}", flush=True)
    return dict(com=com, moved=moved, gv=g1[0], gw=g1[1], gmu_sum=np.sum(g1[2]), n_contacts=len(case.contacts),
                gradient_drift=gradient_drift, forward_drift=forward_drift)


def main():
    ks = sample()
    with Pool(min(8, len(ks))) as pool:
        rows = pool.map(answer, ks)
    arrays = dict(worlds=np.array(WORLDS), steps=np.array(STEPS), seed0=np.array(SEED0), sample=np.array(ks))
    for k, r in zip(ks, rows):
        for name, v in r.items():
            arrays[f"w{k}_{name}"] = np.asarray(v)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
