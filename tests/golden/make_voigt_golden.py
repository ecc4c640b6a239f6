"""Writes voigt_wofz.npz: the Faddeeva function w(u + i a) from
scipy.special.wofz on a grid of damping a and reduced wavelength u, giving
H = Re w (Voigt) and F = Im(w)/2 (Faraday-Voigt) as the external truth for
the Voigt routines.  Run from the repository root:
    python tests/golden/make_voigt_golden.py
"""
import os

import numpy as np
from scipy.special import wofz

a = np.array([1e-4, 1e-3, 0.01, 0.05, 0.1, 0.2, 0.5, 1.0, 2.0, 5.0, 10.0])
u = np.linspace(-20.0, 20.0, 401)
A, U = np.meshgrid(a, u, indexing="ij")
w = wofz(U + 1j * A)
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "voigt_wofz.npz")
np.savez(out, a=a, u=u, H=w.real, F=w.imag / 2.0)
print(out)
