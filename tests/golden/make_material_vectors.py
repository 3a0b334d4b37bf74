"""Writes tests/golden/material_ref.npz: PK1 stresses and approximate-
Jacobian coefficients computed by the reference's OWN constitutive sources
(src/material/{NeoHookean,MooneyRivlin,Yeoh}.cpp, compiled unmodified into
oracle/_ref/libref_materials.so by oracle/Makefile).

Inputs: seeded deformation gradients -- near identity, large stretch,
shear, and compression -- for each law and two parameter sets.
Run from the repository root:  python tests/golden/make_material_vectors.py
"""
from pathlib import Path
import sys

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import oracle as O  # noqa: E402

PARAMS = {
    0: [[114000.0, 0.0, 0.0, 1e-6], [80000.0, 0.0, 0.0, 1e-3]],
    1: [[80000.0, 20000.0, 0.0, 1e-6], [57000.0, 57000.0, 0.0, 1e-4]],
    2: [[500000.0, -10000.0, 1000.0, 1e-7], [114000.0, 16000.0, 16000.0, 1e-6]],
}


def main():
    if O.ref_lib() is None:
        raise SystemExit("oracle/_ref/libref_materials.so missing: make -C oracle ref")
    rng = np.random.default_rng(7)
    Fs = [np.eye(3)]
    for scale in (1e-4, 1e-2, 0.3):
        for _ in range(8):
            Fs.append(np.eye(3) + scale * rng.standard_normal((3, 3)))
    Fs.append(np.diag([2.0, 0.7, 0.7]))
    Fs.append(np.array([[1.0, 0.8, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]))
    Fs.append(np.diag([0.6, 1.1, 1.2]))
    F = np.array(Fs)
    laws, params, P, coeff = [], [], [], []
    for law, plist in PARAMS.items():
        for p in plist:
            p = np.array(p)
            for Fi in F:
                laws.append(law)
                params.append(p)
                P.append(O.ref_piola(law, p, Fi))
            R = O.ref_lib()
            c = R.ref_elasticity_coeff(law, p.ctypes.data_as(O._dp))
            coeff.append([law, *p, c])
    out = Path(__file__).with_name("material_ref.npz")
    np.savez(out, F=F, law=np.array(laws), params=np.array(params), P=np.array(P),
             coeff=np.array(coeff))
    print(f"wrote {out}: {len(P)} stress vectors")


if __name__ == "__main__":
    main()
