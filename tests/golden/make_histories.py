"""Oracle histories of the decaying-isotropic LES runs, as small npz fixtures.

The GPU tests (``tests/test_gpu_histories.py``) step the same seeded fields on the
device and compare, step by step, the time t, the CFL dt, the energy E, the
enstrophy Omega and the resolved energy E_res that the reference's driver records
after every ``advanceQ_RK4`` (``src/PySpec_3d.py:121-148``), the dissipation -dE/dt
formed from them, and the shell spectrum E(k) (``src/Classes.py:762-797``) at a few
steps.

The field is the seeded k^4 exp(-2 (k/4)^2) spectrum with u_rms = 1
(``oracle.random_isotropic``, seed 0), nu = 1e-3, CFL 0.5.  The oracle runs here, on
the CPU of the build container; only the resulting arrays are committed.

    python tests/golden/make_histories.py            # every case
    python tests/golden/make_histories.py smag64     # one case
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import spectral_oracle as so  # noqa: E402

# name: (N, turb_model, steps, steps at which E(k) is recorded)
CASES = {
    "smag64": (64, 1, 1000, (0, 250, 500, 750, 1000)),
    # the t-model's memory term Ct t PLQL grows with t: this run blows up at t = 1.50
    # (step 225), so its history stops before that
    "tmodel64": (64, 2, 220, (0, 55, 110, 165, 220)),
    "smag128": (128, 1, 100, (0, 50, 100)),
    "tmodel128": (128, 2, 16, (0, 8, 16)),
}
NU = 1e-3
CFL = 0.5
SEED = 0


def path(name: str) -> str:
    return os.path.join(HERE, f"history_{name}.npz")


def make(name: str, workers: int | None = None) -> str:
    N, model, steps, spec_at = CASES[name]
    uh = so.random_isotropic(N, seed=SEED)
    g = so.Grid(N, N, N)
    st = so.State(g, *uh, nu=NU, turb_model=model, cfl=CFL)
    fft = so.FFTBackend(workers=workers or so.default_workers())
    spectra = {}
    dts = []
    if 0 in spec_at:
        spectra[0] = so.compute_spectrum(st)[1]
    t0 = time.perf_counter()

    def cb(s, it, t):
        dts.append(s.dt)
        if it in spec_at:
            spectra[it] = so.compute_spectrum(s)[1]
        if it % 50 == 0 or it == steps:
            print(f"{name}: step {it}/{steps} t={t:.5f} E={so.compute_energy(s):.10g} "
                  f"({time.perf_counter() - t0:.0f} s)", flush=True)

    h = so.run(st, fft, et=1e30, max_steps=steps, callback=cb)
    out = path(name)
    np.savez_compressed(out, N=N, turb_model=model, nu=NU, cfl=CFL, seed=SEED, t=h["t"], dt=np.array(dts),
                        Energy=h["Energy"], Enstrophy=h["Enstrophy"], Energy_resolved=h["Energy_resolved"],
                        Dissipation=h["Dissipation"], spec_steps=np.array(sorted(spectra)),
                        spectra=np.stack([spectra[k] for k in sorted(spectra)]))
    print(f"wrote {out}", flush=True)
    return out


if __name__ == "__main__":
    for n in (sys.argv[1:] or list(CASES)):
        make(n)
