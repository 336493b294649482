"""Writes tests/golden/drift_16x16_beta8.txt: the compiled reference's own
stabilisation drift (stab_drift_file, dqmc_record_stab_drift,
src/dqmc.c:64-98 / src/replica_run.c:58-88) on the metric's workload,
16x16 periodic, U=4, dtau=0.1, beta=8, stab=8: 4 serial replicas
(replica_seed(1, 0, r)), 2 warm-up and 4 measured sweeps.  Run in the build
container, where oracle/_ref/dqmc_omp is built from the reference sources:

    python tests/golden/make_drift16.py
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402

INPUT = ("lattice=square\nLx=16\nLy=16\nU=4\ndtau=0.1\nbeta_list=8\nstab=8\nnwarm=2\nnmeas=4\nnbin=2\n"
         "seed=1\nparallel=serial\nnrep=4\nszz_q=none\nsperp_q=none\nstab_drift_file=drift.dat\n")


def main():
    with tempfile.TemporaryDirectory() as d:
        inp = os.path.join(d, "in.txt")
        open(inp, "w").write(INPUT)
        env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1")
        subprocess.run([pyoracle.REF_OMP, inp], cwd=d, check=True, env=env, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)
        shutil.copy(os.path.join(d, "drift.dat"), os.path.join(HERE, "drift_16x16_beta8.txt"))


if __name__ == "__main__":
    main()
