"""Writes tests/golden/control128_fd.json: FP64 central finite differences of the density-control
loss at 32 entries of (u0, rho0) on the control128 scene (128^3 box, sphere R = 60, 1,177,088
voxels, 8 seeded Gaussian blobs, curl-noise control velocity; 16 steps of advect + 80 red-black
Gauss-Seidel iterations), computed with the project's own FP64 oracle (oracle/hns_oracle.cpp,
diffForward / diffLossGrad). The GPU test compares the checkpointed FP32 adjoint against these values;
a GPU run cannot afford the 128 FP64 forwards (about 30 minutes on 8 cores).

Entries: 16 velocity components and 16 densities drawn (seeded) from the 300 entries of largest
|gradient| of the FP64 tape. For each, the loss difference quotient at eps = 1e-3 and eps / 2; an
entry counts as smooth when the two agree (no clamp switch or cell crossing within +-eps).

    python tests/golden/make_control128_fd.py
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from hnanosolver_amd import scenes  # noqa: E402
from oracle import Oracle, num_threads  # noqa: E402

STEPS, K, EPS, N_ENTRIES = 16, 80, 1e-3, 32
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "control128_fd.json")


def omega(dx):
    return float(np.float32(2.0 / (1.0 + np.sin(np.float32(3.14159) * np.float32(dx)))))


def main():
    sc = scenes.control128()
    dx, dt = sc.voxel_size, sc.dt
    o = Oracle(sc.coords, np.float64)
    u64 = sc.velocity.astype(np.float64)
    r64 = sc.fields["density"].astype(np.float64)
    tg = sc.sources["target_density"].astype(np.float64)
    t0 = time.perf_counter()
    loss, gv, gr = o.diff_loss_grad(u64, r64, tg, STEPS, dt, dx, K, omega(dx))
    print("FP64 tape: loss %.10e (%.1f s, %d threads)" % (loss, time.perf_counter() - t0, num_threads()), flush=True)

    def L(uu, rr):
        _, r = o.diff_forward(uu, rr, STEPS, dt, dx, K, omega(dx))
        return float(((r - tg) ** 2).sum())

    rng = np.random.default_rng(2024)
    cand_v = np.argsort(-np.abs(gv).max(1))[:300]
    cand_r = np.argsort(-np.abs(gr))[:300]
    meta = {"scene": "control128", "seed": 3, "target_seed": 33, "voxels": int(sc.n), "steps": STEPS,
            "iterations": K, "dt": dt, "voxel_size": dx, "omega": omega(dx), "eps": EPS, "loss_fp64": loss,
            "grad_scale_fp64": float(max(np.abs(gv).max(), np.abs(gr).max())),
            "grad_norm_fp64": {"velocity": float(np.linalg.norm(gv)), "density": float(np.linalg.norm(gr))}}
    entries = []
    if os.path.exists(OUT):  # resume: entries already written for the same scene and loss
        with open(OUT) as fh:
            old = json.load(fh)
        if old.get("loss_fp64") == loss and old.get("voxels") == meta["voxels"]:
            entries = old["entries"]
            print("resuming after %d entries" % len(entries), flush=True)
    for j in range(N_ENTRIES):
        if j % 2 == 0:
            i = int(rng.choice(cand_v))
            k = int(np.argmax(np.abs(gv[i])))
            g = float(gv[i, k])

            def fd(e):
                up, um = u64.copy(), u64.copy()
                up[i, k] += e
                um[i, k] -= e
                return (L(up, r64) - L(um, r64)) / (2 * e)
            kind = "velocity"
        else:
            i = int(rng.choice(cand_r))
            k = 0
            g = float(gr[i])

            def fd(e):
                rp, rm = r64.copy(), r64.copy()
                rp[i] += e
                rm[i] -= e
                return (L(u64, rp) - L(u64, rm)) / (2 * e)
            kind = "density"
        if j < len(entries):  # the seeded draws above are replayed, so a resumed run picks the same entries
            assert entries[j]["index"] == i and entries[j]["field"] == kind
            continue
        t1 = time.perf_counter()
        f1, f2 = fd(EPS), fd(EPS / 2)
        entries.append({"field": kind, "index": i, "component": k, "fd_eps": f1, "fd_half_eps": f2, "tape_fp64": g})
        print("%2d %-8s i=%7d k=%d fd %.9e fd/2 %.9e tape %.9e (%.0f s)" % (j, kind, i, k, f1, f2, g,
                                                                          time.perf_counter() - t1), flush=True)
        with open(OUT, "w") as fh:
            json.dump(dict(meta, complete=len(entries) == N_ENTRIES, entries=entries), fh, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
