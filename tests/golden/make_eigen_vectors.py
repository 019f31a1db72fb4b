"""Writes tests/golden/eigen_cg.json: small Jacobi-CG runs of the Eigen
ConjugateGradient vendored by the reference (through oracle/_ref, built from
the reference tree in this container). The fixture holds only inputs and
outputs (data); run:  python tests/golden/make_eigen_vectors.py"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402
from cumat_amd.problems import poisson2d  # noqa: E402


def main():
    cases = []
    for g, shift, tol in [(8, 0.0, 1e-10), (16, 0.0, 1e-10), (32, 0.0, 1e-10), (16, 0.5, 1e-12), (32, 1.0, 1e-8)]:
        JA, IA, V = poisson2d(g, shift, np.float64)
        n = g * g
        b = np.random.default_rng(1000 + g).uniform(-1, 1, n)
        x64, it64, err64 = oracle.eigen_pcg(JA, IA, V, b, n, tol=tol, max_iter=10 * n)
        x32, it32, err32 = oracle.eigen_pcg(JA, IA, V.astype(np.float32), b.astype(np.float32), n,
                                            tol=max(tol, 1e-6), max_iter=10 * n)
        cases.append({"g": g, "shift": shift, "tol": tol, "max_iter": 10 * n, "b": b.tolist(),
                      "iterations_f64": it64, "error_f64": err64, "x_f64": x64.tolist(),
                      "tol_f32": max(tol, 1e-6), "iterations_f32": it32, "error_f32": err32,
                      "x_f32": x32.astype(np.float64).tolist()})
    out = {"_source": "Eigen 3.3.90 ConjugateGradient<SparseMatrix<S,RowMajor>, Lower|Upper, "
                      "DiagonalPreconditioner> (third-party/Eigen/src/IterativeLinearSolvers/"
                      "ConjugateGradient.h:28-91) on 2D Poisson + shift*I, b ~ U[-1,1] from "
                      "numpy default_rng(1000+g)",
           "cases": cases}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "eigen_cg.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
