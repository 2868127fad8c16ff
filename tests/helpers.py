"""Small synthetic mixtures and independent physics references for tests."""
import numpy as np

import ekmc

AMU = ekmc.AMU


def mixture(procs, fractions=(1.0,)):
    """Build a MixtureData from a list of dicts with keys: type, elastic,
    threshold, mass_amu, energy, sigma, scattering, params, gas, super, statw."""
    cols = {k: [] for k in ("type", "is_elastic", "is_superelastic", "gas_index", "threshold",
                            "target_mass", "rel_density", "stat_weight_ratio", "w_parameter",
                            "scattering", "scattering_params")}
    offsets, en, sg = [0], [], []
    for p in procs:
        g = p.get("gas", 0)
        cols["type"].append(p.get("type", 0))
        cols["is_elastic"].append(1 if p.get("elastic", False) else 0)
        cols["is_superelastic"].append(1 if p.get("super", False) else 0)
        cols["gas_index"].append(g)
        cols["threshold"].append(p.get("threshold", 0.0))
        cols["target_mass"].append(p.get("mass_amu", 40.0) * AMU)
        cols["rel_density"].append(p.get("rel_density", fractions[g]))
        cols["stat_weight_ratio"].append(p.get("statw", 0.0))
        cols["w_parameter"].append(p.get("w", p.get("threshold", 0.0)))
        cols["scattering"].append(ekmc.SCATTERING[p.get("scattering", "isotropic")])
        prm = list(p.get("params", ())) + [0.0, 0.0]
        cols["scattering_params"] += prm[:2]
        e = np.asarray(p.get("energy", []), np.float64)
        s = np.asarray(p.get("sigma", []), np.float64)
        en.append(e)
        sg.append(s)
        offsets.append(offsets[-1] + len(e))
    arrays = {k: np.array(v, np.int32 if k in ("type", "is_elastic", "is_superelastic",
                                                "gas_index", "scattering") else np.float64)
              for k, v in cols.items()}
    arrays["table_offset"] = np.array(offsets, np.int32)
    arrays["table_energy"] = np.concatenate(en) if en else np.zeros(0)
    arrays["table_sigma"] = np.concatenate(sg) if sg else np.zeros(0)
    arrays["gas_fraction"] = np.array(fractions, np.float64)
    ng = max(p.get("gas", 0) for p in procs) + 1
    return ekmc.MixtureData(arrays, ng, ["p%d" % i for i in range(len(procs))])


def condition(en=100.0, angle=180.0, freq=0.0, bn=0.0, tg=300.0, n_e=1000, p=133.32):
    return ekmc.Condition(reduced_elec_field=en, elec_field_angle=angle, excitation_frequency=freq,
                          reduced_mag_field=bn, gas_temperature=tg,
                          gas_density=p / (ekmc.BOLTZMANN * tg), n_electrons=n_e)


def unit_vectors(rng, n):
    d = rng.normal(size=(n, 3))
    return d / np.linalg.norm(d, axis=1)[:, None]


def speeds(energy):
    return np.sqrt(2 * np.asarray(energy) * ekmc.CHARGE / ekmc.EMASS)


def lorentz_rhs(field):
    """m dv/dt = -e (E(t) + v x B), B along z, E(t) = E0 cos(w t)."""
    q_m = ekmc.CHARGE / ekmc.EMASS
    E0 = np.array(field.elec_field)
    w, W = field.omega, field.cyclotron

    def rhs(t, y):
        v = y[3:]
        Et = E0 * (np.cos(w * t) if w != 0 else 1.0)
        # v x (B z) = (vy B, -vx B, 0); q_m*B = W
        vxB = np.array([v[1], -v[0], 0.0]) * W / q_m
        a = -q_m * (Et + vxB)
        return np.concatenate([v, a])
    return rhs


def ks_statistic(samples, cdf):
    x = np.sort(samples)
    n = len(x)
    F = cdf(x)
    return max(np.max(np.arange(1, n + 1) / n - F), np.max(F - np.arange(n) / n))
