"""Write per-step state digests of the oracle on the configs' workloads.

The GPU parity tests at the configs' full sizes (tests/test_gpu_configs.py)
run the device step from the same seeded field and compare each step's
digest (tests/helpers.py state_digest: SHA-256 of the contour store and of
qs, plus t and the counts) with these, so a bit-level difference is caught
at the step where it first appears, without the oracle's minutes of CPU
time on the GPU machine.  Run in the build container:

    make -C oracle && python tests/golden/gen_digests.py 256 100
    python tests/golden/gen_digests.py 1024 500

Workload: random_pv(ng, seed=12345) (k^3 exp(-2(k/6)^2) spectrum, rms
4pi), 80 PV levels, first contours from recontouring the field, then one
pass of the advect loop per step (caps.f90:83-91, evolution.f90:55-80: run
one step at a time, surgery and recontouring included), adaptive dt.
Digest after every step up to 100, then every 10th step.
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from hydra_amd.ic import random_pv  # noqa: E402
from oracle.hyo import Oracle  # noqa: E402
from tests.helpers import state_digest  # noqa: E402

SEED = 12345


def recorded(step):
    return step <= 100 or step % 10 == 0


def main(ng, nsteps):
    o = Oracle(ng)
    o.init_from_field(random_pv(ng, seed=SEED))
    out = {"ng": ng, "seed": SEED, "ncontq": 80, "steps": nsteps, "records": []}
    t0 = time.time()
    for step in range(1, nsteps + 1):
        o.run(1)
        if recorded(step):
            rec = state_digest(o.contours(), o.d("qs"), o.getd("t"))
            rec["step"] = step
            out["records"].append(rec)
        if step % 10 == 0:
            print(f"ng={ng} step {step}: nodes {o.get('nptq')} t={o.getd('t'):.5f} "
                  f"{time.time() - t0:.0f}s", flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"digests_ng{ng}.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0)
    print(path)


if __name__ == "__main__":
    main(int(sys.argv[1]), int(sys.argv[2]))
