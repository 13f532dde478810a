"""Writes tests/golden/batched_512_sad_candidates.json: the exact number of
motion-search candidates the reference evaluates (valid coarse + valid
refine candidates, src/megpu_kernels.cuh:67-92, 304-446; each is one 16x16
SAD = 256 byte-SADs) in every frame step of the benchmark's batched job
(256 CIF + 256 QCIF streams, h261enc_cuda_amd.workloads.batched_streams),
counted by the oracle over every stream's whole sequence.

Step t covers frame t of every stream whose length exceeds t.  Step 0 is the
intra frame: the engine does not search it (mode decision makes every MB
intra whatever the costs), so it counts 0 here.  The benchmark divides the
candidates of the steps it times by the launches to get the algorithmic
byte-SADs per motion-estimation launch.

Run from the repository root:  python tests/golden/make_sad_counts.py
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from h261enc_cuda_amd.workloads import batched_streams  # noqa: E402
from oracle_pool import OracleRuns  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "batched_512_sad_candidates.json")


def main():
    specs = batched_streams()
    job = max(s["length"] for s in specs)
    t0 = time.time()
    runs = OracleRuns(specs, job)
    res = runs.result()
    per_step = [0] * job
    cif = [0] * job
    for spec, frames in zip(specs, res):
        for t, f in enumerate(frames):
            if t == 0:
                continue
            per_step[t] += f.sad
            if spec["size"][0] == 352:
                cif[t] += f.sad
    out = {"source": "tests/golden/make_sad_counts.py",
           "workload": "batched: 256 CIF + 256 QCIF streams (h261enc_cuda_amd.workloads.batched_streams, seed 2026)",
           "unit": "16x16 SAD candidates (256 byte-SADs each)",
           "steps": job,
           "candidates_per_step": per_step,
           "cif_candidates_per_step": cif}
    with open(OUT, "w") as f:
        json.dump(out, f, separators=(",", ":"))
        f.write("\n")
    print(f"wrote {OUT}: {sum(per_step)} candidates over {job} steps in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
