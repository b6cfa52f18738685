"""Write tests/golden/config2_golden.npz: the decaying-turbulence case at 1024^2
computed by the CPU oracle (FP64, the reference algorithm with the project FFT).

Case: [0, 2 pi]^2, N = 1024, Re = 1e4, random-phase vorticity with
E(k) ~ k^4 exp(-(k/10)^2), unit energy, seed 42; del^8 hyperviscosity
nu4 = 1e-16 applied as an integrating factor (coefficient nu4 * kmax^8 with
kmax = sqrt(2) N / 2, the corner of the stored modes, hyperviscosity.c:68-73);
adaptive dt = CFL(0.5) before every step (main_spectral.c:46-48);
1000 steps.

Stored (small): t, dt, energy and enstrophy every 100 steps; the shell
spectrum of spectral_analysis.c:103-142 after the last step; the final
vorticity modes with |kx| <= 16, |ky| <= 16.

Runs in a few minutes on 8 cores: python tests/golden/make_config2.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

N, RE, SEED, STEPS, EVERY, KLOW = 1024, 1e4, 42, 1000, 100, 16


def main():
    o = O.Oracle(N, N, 2 * np.pi, 2 * np.pi, RE)
    o.init_grids()
    kmax = np.sqrt(2.0) * N / 2.0
    o.hv_set(enabled=1, coefficient=1e-16 * kmax ** 8, order=8, cutoff_ratio=0.0)
    o.d.cfl_factor = 0.5
    o.init_random_phase(SEED, 10.0, 1.0)
    rows = [(0, o.d.t, 0.0, o.energy(), o.enstrophy())]
    t0 = time.time()
    for step in range(1, STEPS + 1):
        o.d.dt = o.cfl()
        o.advance_if()
        if step % EVERY == 0:
            rows.append((step, o.d.t, o.d.dt, o.energy(), o.enstrophy()))
            print(f"step {step} t {o.d.t:.6f} E {rows[-1][3]:.12e} Z {rows[-1][4]:.12e} "
                  f"({time.time() - t0:.0f} s)", flush=True)
    ek, ck = o.analyze_spectrum()
    w = o.spec("omega_hat")
    low = np.concatenate([w[: KLOW + 1, : KLOW + 1], w[N - KLOW:, : KLOW + 1]], axis=0)
    np.savez_compressed(os.path.join(HERE, "config2_golden.npz"),
                        history=np.array(rows), spectrum=ek, counts=ck, omega_low=low,
                        meta=np.array([N, RE, SEED, STEPS, EVERY, KLOW]))


if __name__ == "__main__":
    main()
