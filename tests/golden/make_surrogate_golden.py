"""Writes tests/golden/surrogate_adjoint.npz: input / output vectors of the
reference's 2-D conduction surrogate adjoint, solve_thermal_adjoint in
AdjointMethod/thermal_adjoint.py (run here, in this container only, as the
source of truth for the one mode nothing else pins):

  external-T mode  T, k, P given (how run_coldplate_adjoint.py:112-117 calls
                   it with the 3-D solver's bottom temperature): the result
                   is not the derivative of a consistent objective, so only
                   the reference's own output fixes its numbers;
  default mode     no inputs: its built-in 40 x 40 case (k uniform
                   (400 + 0.6) / 2, a 100 W block source), T solved by it.

The reference module saves a plot into the working directory; it runs in a
temporary directory (tempfile) that is removed afterwards.

usage: python tests/golden/make_surrogate_golden.py
"""
import importlib.util
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference/AdjointMethod/thermal_adjoint.py"


def ramp(g, k_s=400.0, k_f=0.6, q=1.0):
    return k_f + (k_s - k_f) * (1.0 - g) / (1.0 + q * g)


def main():
    spec = importlib.util.spec_from_file_location("ref_thermal_adjoint", REF)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    rng = np.random.default_rng(2024)
    ny, nx = 12, 10
    gamma = rng.uniform(0.0, 1.0, (ny, nx))
    k = ramp(gamma)
    T = 30.0 + 50.0 * rng.uniform(size=(ny, nx))
    P = 1e3 * rng.uniform(size=(ny, nx))
    cwd = os.getcwd()
    # the script writes its outputs into the working directory
    with tempfile.TemporaryDirectory() as tmp:
        os.chdir(tmp)
        try:
            sens_ext = mod.solve_thermal_adjoint(T_in=T.copy(), k_in=k.copy(), P_in=P.copy())
            sens_def = mod.solve_thermal_adjoint()
        finally:
            os.chdir(cwd)
    k_def = np.full((40, 40), (400.0 + 0.6) / 2.0)
    P_def = np.zeros((40, 40))
    P_def[10:30, 10:30] = 100.0
    out = os.path.join(HERE, "surrogate_adjoint.npz")
    np.savez_compressed(out, ext_gamma=gamma, ext_k=k, ext_T=T, ext_P=P, ext_sens=sens_ext,
                        def_k=k_def, def_P=P_def, def_sens=sens_def)
    print("wrote", out)


if __name__ == "__main__":
    sys.exit(main())
