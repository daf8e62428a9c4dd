"""Copy the numbers of the reference's archived result CSVs into
tests/golden/reference_runs.json (data only, with provenance).

Run in the build container, where the reference checkout is mounted:
    python tests/golden/make_fixtures.py /root/reference
The JSON it writes is committed; tests never read the reference tree.
"""
import csv
import json
import os
import sys

RUNS = {
    # 4x4, k=2, CFL 2, T=0.02, --method both --check-update (current code version)
    "be_4x4_k2": ("speed_test/results/2d_n512_restart30_20261020T220633_rt2c3wx6/job_7687730/checks", "be",
                  ["--nx", "4", "--ny", "4", "--k", "2", "--cfl", "2", "--method", "both", "--check-update",
                   "--ntol", "1e-12", "--ltol", "1e-12", "--integrator", "be", "--T", "0.02"]),
    "dirk3_4x4_k2": ("speed_test/results/2d_n512_restart30_20261020T220633_rt2c3wx6/job_7687730/checks", "dirk3",
                     ["--nx", "4", "--ny", "4", "--k", "2", "--cfl", "2", "--method", "both", "--check-update",
                      "--ntol", "1e-12", "--ltol", "1e-12", "--T", "0.02"]),
    "entropy_4x4_k2": ("speed_test/results/2d_n512_restart30_20261020T220633_rt2c3wx6/job_7687730/checks",
                       "entropy",
                       ["--nx", "4", "--ny", "4", "--k", "2", "--cfl", "2", "--method", "both", "--check-update",
                        "--ntol", "1e-12", "--ltol", "1e-12", "--testcase", "wave", "--T", "0.02"]),
    # 256^2, k=8, BE, CFL 5, GMRES(20), 4 fixed Newton updates (speed protocol)
    "k8_256_be": ("speed_test/results/2d_submission_20261019T205542_KhFC6z/job_7635952/data", "k8_dg",
                  ["--nx", "256", "--ny", "256", "--k", "8", "--method", "dg", "--integrator", "be", "--steps", "1",
                   "--newton-steps", "4", "--warmup-updates", "1", "--cfl", "5", "--ntol", "1e-12", "--ltol",
                   "1e-12", "--restart", "20", "--max-linear", "1500", "--safety", "1.1", "--timing-only"]),
    "k8_256_be_hdg": ("speed_test/results/2d_submission_20261019T205542_KhFC6z/job_7635952/data", "k8_hdg", None),
    "k2_256_be": ("speed_test/results/2d_submission_20261019T205542_KhFC6z/job_7635952/data", "k2_dg", None),
    # 512^2, k=8, BE, CFL 5, GMRES(30)
    "k8_512_be": ("speed_test/results/2d_n512_restart30_20261020T220633_rt2c3wx6/job_7718103/data", "k8_dg", None),
}
CONVERGENCE = "accuracy_test/results/2d_submission_20261019T234902_AtKkgL/job_7640912/data/convergence.csv"
# 1D, k=2, DIRK3, CFL 2, T 0.2, GMRES(20) + block Jacobi (accuracy_test/1d/convergence.py:19-25)
CONVERGENCE_1D = "accuracy_test/results/1d_slurm_7622710/convergence_slurm_7622710.csv"


def read_csv(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def main(ref):
    out = {"provenance": "numbers copied from YueWu2002/DG_HDG_Euler_Cpp archived result CSVs", "runs": {}}
    for key, (folder, stem, args) in RUNS.items():
        it_path = os.path.join(ref, folder, stem + "_iterations.csv")
        sm_path = os.path.join(ref, folder, stem + "_summary.csv")
        run = {"source": [os.path.join(folder, stem + "_iterations.csv"), os.path.join(folder, stem + "_summary.csv")],
               "args": args, "iterations": [], "summary": []}
        for row in read_csv(it_path):
            run["iterations"].append({k: row[k] for k in ["method", "step", "stage", "iteration", "dt", "cfl",
                                                           "residual_physical_rms", "krylov_iterations",
                                                           "linear_relative_residual", "dg_linear_relative_residual"]})
        if os.path.exists(sm_path):
            for row in read_csv(sm_path):
                run["summary"].append({k: row.get(k) for k in ["method", "steps", "final_time", "newton_iterations",
                                                                "krylov_iterations", "rho_L2", "momentum_x_L2",
                                                                "momentum_y_L2", "energy_L2",
                                                                "max_conservation_drift", "lx", "ly",
                                                                "mean_newton_s"]})
        out["runs"][key] = run
    conv = []
    for row in read_csv(os.path.join(ref, CONVERGENCE)):
        conv.append({k: row[k] for k in ["method", "nx", "k", "steps", "newton_iterations", "krylov_iterations",
                                          "rho_L2", "momentum_x_L2", "momentum_y_L2", "energy_L2", "rho_order",
                                          "max_conservation_drift"]})
    out["convergence"] = {"source": CONVERGENCE, "rows": conv,
                          "args": ["--k", "2", "--cfl", "2", "--integrator", "ssp-dirk3", "--T", "0.2",
                                   "--ntol", "1e-12", "--ltol", "1e-12", "--method", "both"]}
    conv1 = []
    for row in read_csv(os.path.join(ref, CONVERGENCE_1D)):
        conv1.append({k: row[k] for k in ["method", "cells", "k", "steps", "newton_iterations", "krylov_iterations",
                                           "max_linear_relative_residual", "min_step", "max_step", "rho_L2",
                                           "momentum_L2", "energy_L2", "max_conservation_drift", "rho_order",
                                           "mean_newton_s", "difference_max_nodal"]})
    out["convergence_1d"] = {"source": CONVERGENCE_1D, "rows": conv1,
                             "args": ["--k", "2", "--integrator", "ssp-dirk3", "--cfl", "2", "--T", "0.2",
                                      "--method", "both", "--precond", "block-jacobi", "--ordering", "natural",
                                      "--ntol", "1e-12", "--ltol", "1e-12", "--restart", "20", "--max-linear",
                                      "1500", "--max-newton", "30", "--safety", "1.1"]}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_runs.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
