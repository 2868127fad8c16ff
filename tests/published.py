"""Published LoKI-MC results used as known answers, and the model gases they
are quoted for (Documentation/benchmarkCalculations_LoKI-MC_v1.0.0.pdf)."""
import os

import numpy as np

import ekmc
from helpers import mixture

# Table 1, Reid ramp (1e5 electrons, stopping at 0.1 % / 0.4 % / 0.8 %
# relative standard deviations): mean energy (eV), bulk drift velocity (m/s),
# reduced bulk transverse and longitudinal diffusion (1/(m s)), each with its
# relative standard deviation in %.
REID_PUBLISHED = {
    1.0: ((0.1014, 0.06), (1.264e4, 0.43), (0.9637e24, 0.70), (0.7542e24, 0.68)),
    12.0: ((0.2691, 0.10), (6.838e4, 0.09), (1.137e24, 0.44), (0.5679e24, 0.35)),
    24.0: ((0.4081, 0.10), (8.881e4, 0.07), (1.138e24, 0.53), (0.4595e24, 0.56)),
}

# Table 2, Lucas-Saelee model gas (E/N = 10 Td, N = 1e20 m-3, M = 1000 m_e),
# keyed by (F, a, p): mean energy (eV), flux and bulk drift velocity (1e4 m/s),
# flux and bulk reduced transverse diffusion, flux and bulk reduced
# longitudinal diffusion (1e25 1/(m s)); relative standard deviations in %.
LS_PUBLISHED = {
    (0.0, 0.0, 0.0): ((5.56, 0.05), (7.33, 0.40), (7.33, 0.53), (2.73, 0.59), (2.73, 0.78), (2.66, 0.57), (2.68, 0.86)),
    (0.5, 0.0, 0.0): ((5.22, 0.05), (7.28, 0.40), (8.50, 0.52), (2.54, 0.58), (2.70, 1.35), (2.48, 0.62), (2.85, 1.18)),
    (1.0, 0.0, 0.0): ((4.97, 0.05), (7.36, 0.40), (9.47, 0.53), (2.44, 0.60), (2.75, 1.99), (2.38, 0.62), (2.89, 1.38)),
    (0.0, 5e-4, 0.5): ((5.44, 0.06), (7.35, 0.40), (7.04, 0.68), (2.67, 0.59), (2.61, 1.56), (2.62, 0.59), (2.74, 1.53)),
    (0.0, 2e-3, -0.5): ((5.55, 0.05), (7.34, 0.40), (7.31, 0.47), (2.71, 0.59), (2.77, 1.40), (2.65, 0.60), (2.67, 1.39)),
}
LS_NAMES = ("mean energy", "flux drift", "bulk drift", "flux N D_T", "bulk N D_T", "flux N D_L", "bulk N D_L")


def reid_mixture():
    """The reference's Reid-ramp LXCat file (4 amu)."""
    path = os.path.join(os.path.dirname(__file__), "golden", "reid_ramp_LXCat.txt")
    procs = []
    for ent in ekmc.read_lxcat(path):
        procs.append(dict(type=0, elastic=ent["type"] == "Elastic", threshold=ent["threshold"],
                          mass_amu=4.0, energy=ent["energy"], sigma=ent["sigma"]))
    return mixture(procs)


def lucas_saelee_processes(F, a, p):
    """Elastic 4/sqrt(eps) A^2, an inelastic/ionization pair at 15.6 eV with
    slope 0.1 A^2/eV split by F, attachment a eps^p A^2; M = 1000 m_e."""
    e = np.concatenate([[0.0], np.logspace(-4, 3, 1401)])
    with np.errstate(divide="ignore"):
        el = np.where(e > 0, 4.0 / np.sqrt(np.maximum(e, 1e-300)), 4.0 / np.sqrt(1e-4)) * 1e-20
    m_amu = 1000.0 * ekmc.EMASS / ekmc.AMU
    procs = [dict(type=0, elastic=True, mass_amu=m_amu, energy=e, sigma=el)]
    eth = np.concatenate([[15.6], e[e > 15.6]])
    if F < 1.0:
        procs.append(dict(type=0, threshold=15.6, mass_amu=m_amu, energy=eth, sigma=0.1 * (1 - F) * (eth - 15.6) * 1e-20))
    if F > 0.0:
        procs.append(dict(type=1, threshold=15.6, w=15.6, mass_amu=m_amu, energy=eth, sigma=0.1 * F * (eth - 15.6) * 1e-20))
    if a > 0.0:
        ea = e[e > 0]
        procs.append(dict(type=2, mass_amu=m_amu, energy=ea, sigma=a * ea ** p * 1e-20))
    return procs


def half_digit(v, digits):
    """Half a unit of the last printed digit of a value given to `digits`
    significant digits."""
    return 0.5 * 10 ** (np.floor(np.log10(abs(v))) - (digits - 1))


def within(v, rel, pv, prel_pct, digits, k=3.0):
    """|v - pv| within k combined standard deviations plus the rounding of
    the published value."""
    tol = k * np.hypot(rel * v, prel_pct / 100.0 * pv) + half_digit(pv, digits)
    return abs(v - pv) <= tol, tol


def ls_values(r, N=1e20):
    """The seven published Lucas-Saelee quantities of a results struct (E
    along -z), each with its relative standard error."""
    def rel(v, e):
        return e / abs(v)
    fd, fe = np.asarray(r.flux_diffusion), np.asarray(r.flux_diffusion_error)
    bd, be = np.asarray(r.bulk_diffusion), np.asarray(r.bulk_diffusion_error)
    return [(r.mean_energy, rel(r.mean_energy, r.mean_energy_error)),
            (abs(r.flux_drift_velocity[2]) / 1e4, rel(r.flux_drift_velocity[2], r.flux_drift_velocity_error[2])),
            (abs(r.bulk_drift_velocity[2]) / 1e4, rel(r.bulk_drift_velocity[2], r.bulk_drift_velocity_error[2])),
            (N * (fd[0] + fd[4]) / 2 / 1e25, max(rel(fd[0], fe[0]), rel(fd[4], fe[4]))),
            (N * (bd[0] + bd[4]) / 2 / 1e25, max(rel(bd[0], be[0]), rel(bd[4], be[4]))),
            (N * fd[8] / 1e25, rel(fd[8], fe[8])),
            (N * bd[8] / 1e25, rel(bd[8], be[8]))]


def reid_values(r, N):
    """Mean energy, bulk drift, reduced bulk transverse and longitudinal
    diffusion (E along -z) with relative standard errors."""
    bd, be = np.asarray(r.bulk_diffusion), np.asarray(r.bulk_diffusion_error)
    return [(r.mean_energy, r.mean_energy_error / r.mean_energy),
            (abs(r.bulk_drift_velocity[2]), r.bulk_drift_velocity_error[2] / abs(r.bulk_drift_velocity[2])),
            (N * (bd[0] + bd[4]) / 2, max(be[0] / abs(bd[0]), be[4] / abs(bd[4]))),
            (N * bd[8], be[8] / abs(bd[8]))]
