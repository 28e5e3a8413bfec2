"""The benchmark configurations of the lattice path as data: lattice shapes,
couplings and seeds of the plain-method run (C1), the 32^3 gauge-fixing run
(C2), the production runs (C3), the replica scan (C4) and the learned
proposal (C5).  Couplings are given in the trace convention

    S = beta_G sum_P (1 - Tr U_P / 2) - beta_A sum_{x,i} Tr[A(x) U A(x+i) U^+]
        + sum_x [beta_2 Tr A^2 + beta_4 (Tr A^2)^2],   A = A^a sigma^a / 2,

and mapped to the reference normalisation (f/agauge.f:142-169,
f/ahiggs1.f:77-78) by :func:`lceqcd_amd.couplings_from_trace_convention`.
Tr A^2 = |a|^2 / 2 per site, i.e. half of the reference's a02 observable."""
import numpy as np

from .lattice import couplings_from_trace_convention

BETA_G = 9.0
BETA_4 = 0.1

# C1: plain method, 16^3, cold start (unitconf, f/confget.f:295-307), 200
# sweeps of 1 KP + 4 OR + 1 Higgs Metropolis (step 0.4), plaquette, Tr A^2
# and the Polyakov loop every sweep
C1 = {"shape": (16, 16, 16), "beta_a": 0.35, "beta_2": -0.9, "sweeps": 200, "seed": 12345,
      "delta": 0.4}

# C2: 32^3, Landau gauge fixing by overrelaxation (f/landau_gauge.f,
# f/or_gauge_fixing1.f) with omega = 1.7, alpha = 1 to da < 1e-12, then the
# z-correlators of f/correlator.f and f/scalarcorr.f
C2 = {"shape": (32, 32, 32), "beta_a": 0.35, "beta_2": -0.9, "lam": 1.7, "alpha": 1.0,
      "limit_or": 1e-12, "configs": 10, "seed": 12345}

# C3: production, hot start, 1 KP + 4 OR + ahiggs1 per sweep, gauge fixing and
# correlators every 50 sweeps (the measurement cadence of f/main.f:58-79)
C3 = {"shapes": [(64, 64, 64), (32, 32, 128)], "beta_a": [0.30, 0.33, 0.36], "beta_2": -0.9,
      "sweeps": 5000, "meas_every": 50, "lam": 1.7, "alpha": 1.0, "limit_or": 1e-10,
      "seed": 12345}

# C4: 96 replicas cycling these shapes over the grid below, seed 12345 + r
C4_SHAPES = [(16, 16, 16), (24, 24, 24), (32, 32, 32), (40, 40, 40), (48, 48, 48),
             (24, 24, 96)]
C4_BETA_A = np.linspace(0.28, 0.40, 12)
C4_BETA_2 = np.linspace(-1.2, -0.6, 8)
# The symmetric/broken transition (24^3 scan, bench.py --workload scan:
# steepest fall of <Tr A^2> and peak of its susceptibility) lies at beta_2
# of about -0.1, 0.2 and 0.5 for beta_A = 0.28, 0.34 and 0.40, so the grid
# above is broken phase throughout; this beta_2 range straddles it at every
# beta_A of the grid.
C4_STRADDLE_BETA_2 = np.linspace(-0.6, 1.2, 8)

# C5: learned proposal at beta_A = 0.34: the configuration's beta_2 = -0.9
# (broken phase, <Tr A^2> about 9) and beta_2 = 0.2, at the transition
C5 = {"shapes": [(48, 48, 48), (96, 96, 96)], "beta_a": 0.34, "beta_2": -0.9,
      "beta_2_transition": 0.2}


def coupling(beta_a, beta_2, beta_4=BETA_4, beta_g=BETA_G):
    """(beta, beta_a, beta_2, beta_4) in the reference normalisation."""
    return couplings_from_trace_convention(beta_g, beta_a, beta_2, beta_4)


def c4_replicas(nrep=96, beta_a=None, beta_2=None):
    """(shapes, couplings, seeds, grid) of the replica scan: replica r has
    shape C4_SHAPES[r % 6], beta_A = beta_a[r % len(beta_a)], beta_2 =
    beta_2[r // len(beta_a)] (trace convention), seed 12345 + r.  The default
    grid is the configuration's literal one (all of it in the broken phase at
    beta_4 = 0.1, where <Tr A^2> is 6 to 10); C4_STRADDLE_BETA_2 as beta_2
    crosses the transition."""
    ba = C4_BETA_A if beta_a is None else np.asarray(beta_a, dtype=np.float64)
    b2 = C4_BETA_2 if beta_2 is None else np.asarray(beta_2, dtype=np.float64)
    if nrep > len(ba) * len(b2):
        raise ValueError("grid has %d points, %d replicas asked" % (len(ba) * len(b2), nrep))
    shapes = [C4_SHAPES[r % len(C4_SHAPES)] for r in range(nrep)]
    grid = [(float(ba[r % len(ba)]), float(b2[r // len(ba)])) for r in range(nrep)]
    coup = [coupling(a, b) for a, b in grid]
    seeds = [12345 + r for r in range(nrep)]
    return shapes, coup, seeds, grid


def tr_a2(a02):
    """Tr A^2 per site (trace convention) from the reference's a02 = <|a|^2>."""
    return np.asarray(a02) / 2.0
