"""Python host side of the engine: ctypes bindings of include/ekmc.h, an LXCat
reader that builds the flattened process list the ABI takes, and the
workload definitions (synthetic gases under ../gases/).

The reader follows LoKI-MC's own (Parse.C:43-295): a ``PROCESS:`` line opens a
block, the threshold is read from the ``E = x eV`` of the next line, the line
after that holds the ``[...]`` description whose last comma field is the
collision type (an optional trailing ``integral`` / ``momentum-transfer``
subtype is dropped), ``<->`` marks a reversible collision, and the table sits
between two ``-----`` rulers.  The process list is flattened as
BoltzmannMC::allocateEvaluateVariablesFirstTime does (BoltzmannMC.C:29-271):
gases in order, their collisions in file order, "Effective" skipped, and a
superelastic entry right after each reversible collision.

Everything here is plain data plumbing; the engine itself lives in
``libekmc.so`` (HIP kernels for gfx950 plus the C++ host engine).
"""
import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
GASES = os.path.join(PKG, "gases")
# EKMC_LIBRARY selects another build of the engine (e.g. a tuning variant)
LIB_PATH = os.environ.get("EKMC_LIBRARY") or os.path.join(PKG, "lib", "libekmc.so")

AMU = 1.660539040e-27          # Constant.h:9
BOLTZMANN = 1.38064852e-23     # Constant.h:6
CHARGE = 1.6021766208e-19      # Constant.h:7
EMASS = 9.10938356e-31         # Constant.h:8

CONSERVATIVE, IONIZATION, ATTACHMENT = 0, 1, 2
SCATTERING = {"isotropic": 0, "forward": 1, "bornDipole": 2, "surendra": 3,
              "coulombScreen": 4, "momentumConservationIonization": 5}
IONIZ = {"equalSharing": 0, "oneTakesAll": 1, "usingSDCS": 2, "randomUniform": 3}
GASTEMP = {"false": 0, "true": 1, "smartActivation": 2}
FP64, MIXED = 0, 1
VARIATES_PER_EVENT = 11
NOALLOC_MAX_EVENTS = 32768  # EKMC_NOALLOC_MAX_EVENTS
NON_DEF = -123456789
NULL_COLLISION, PARTIAL_FLIGHT = -1, -2  # pass outcomes, GeneralDefinitions.h:38-40

# ----------------------------------------------------------------- structs
_p = C.POINTER


class Mixture(C.Structure):
    _fields_ = [("n_processes", C.c_int32), ("n_gases", C.c_int32),
                ("type", _p(C.c_int32)), ("is_elastic", _p(C.c_int32)),
                ("is_superelastic", _p(C.c_int32)), ("gas_index", _p(C.c_int32)),
                ("threshold", _p(C.c_double)), ("target_mass", _p(C.c_double)),
                ("rel_density", _p(C.c_double)), ("stat_weight_ratio", _p(C.c_double)),
                ("w_parameter", _p(C.c_double)), ("scattering", _p(C.c_int32)),
                ("scattering_params", _p(C.c_double)), ("table_offset", _p(C.c_int32)),
                ("table_energy", _p(C.c_double)), ("table_sigma", _p(C.c_double)),
                ("gas_fraction", _p(C.c_double))]


class Condition(C.Structure):
    _fields_ = [("reduced_elec_field", C.c_double), ("elec_field_angle", C.c_double),
                ("excitation_frequency", C.c_double), ("reduced_mag_field", C.c_double),
                ("gas_temperature", C.c_double), ("gas_density", C.c_double),
                ("n_electrons", C.c_int64)]


class Numerics(C.Structure):
    _fields_ = [("precision", C.c_int32), ("gas_temperature_effect", C.c_int32),
                ("ionization_operator", C.c_int32), ("n_interp_points", C.c_int32),
                ("initial_elec_temp_over_gas_temp", C.c_double),
                ("min_collisions_before_ss", C.c_double), ("max_collisions_before_ss", C.c_double),
                ("max_collisions_after_ss", C.c_double), ("n_energy_cells", C.c_int32),
                ("n_cos_angle_cells", C.c_int32), ("n_radial_velocity_cells", C.c_int32),
                ("n_axial_velocity_cells", C.c_int32), ("errors_to_be_checked", C.c_int32),
                ("n_integration_phases", C.c_int32), ("rel_error_mean_energy", C.c_double),
                ("rel_error_flux_drift_velocity", C.c_double),
                ("rel_error_bulk_drift_velocity", C.c_double),
                ("rel_error_flux_diffusion_coeffs", C.c_double),
                ("rel_error_bulk_diffusion_coeffs", C.c_double),
                ("rel_error_power_balance", C.c_double), ("n_integration_points", C.c_double),
                ("n_integrated_ss_times", C.c_double), ("integrated_absolute_time", C.c_double),
                ("sync_time_x_max_coll_freq", C.c_double), ("sync_over_sampling", C.c_int32),
                ("max_sync_points", C.c_int32), ("seed", C.c_uint64),
                ("work_chunk", C.c_int32), ("park_lanes", C.c_int32), ("env_coarsen", C.c_int32),
                ("evdf3d_cells", C.c_int32)]


class Results(C.Structure):
    _fields_ = ([("mean_energy", C.c_double), ("mean_energy_error", C.c_double)] +
                [(n, C.c_double * 3) for n in ("flux_drift_velocity", "flux_drift_velocity_error")] +
                [(n, C.c_double * 9) for n in ("flux_diffusion", "flux_diffusion_error")] +
                [(n, C.c_double * 3) for n in ("bulk_drift_velocity", "bulk_drift_velocity_error")] +
                [(n, C.c_double * 9) for n in ("bulk_diffusion", "bulk_diffusion_error")] +
                [(n, C.c_double * 3) for n in ("rotated_flux_drift_velocity",
                                                "rotated_flux_drift_velocity_error",
                                                "rotated_bulk_drift_velocity",
                                                "rotated_bulk_drift_velocity_error")] +
                [(n, C.c_double * 9) for n in ("rotated_flux_diffusion", "rotated_flux_diffusion_error",
                                                "rotated_bulk_diffusion", "rotated_bulk_diffusion_error")] +
                [(n, C.c_double) for n in ("power_field", "power_growth", "power_balance_rel_error",
                                           "total_ion_rate_coeff", "total_att_rate_coeff",
                                           "flux_red_mobility", "bulk_red_mobility",
                                           "flux_red_transv_diff", "flux_red_long_diff",
                                           "bulk_red_transv_diff", "bulk_red_long_diff",
                                           "steady_state_time", "final_time", "total_integrated_time",
                                           "trial_collision_frequency", "max_elec_energy",
                                           "max_eedf_energy", "elapsed_time")] +
                [(n, C.c_uint64) for n in ("real_collisions", "null_collisions",
                                           "real_collisions_at_ss", "null_collisions_at_ss",
                                           "escape_events")] +
                [(n, C.c_int32) for n in ("n_sampling_points", "n_integration_points",
                                          "good_statistical_errors", "stopped_by_max_collisions")])

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class CondStatus(C.Structure):
    _fields_ = [("time", C.c_double), ("trial_collision_frequency", C.c_double),
                ("steady_state_time", C.c_double), ("last_mean_energy", C.c_double),
                ("max_elec_energy", C.c_double), ("table_max_energy", C.c_double),
                ("real_collisions", C.c_uint64),
                ("null_collisions", C.c_uint64), ("n_sampling_points", C.c_int32),
                ("n_integration_points", C.c_int32), ("finished", C.c_int32), ("pad", C.c_int32),
                ("escape_events", C.c_uint64)]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_uint64), ("total_ms", C.c_double)]


# ------------------------------------------------------------ LXCat reader
def read_lxcat(path):
    """Entries of an LXCat file as LoKI-MC's reader sees them."""
    entries = []
    with open(path) as f:
        lines = f.read().split("\n")
    i = 0
    while i < len(lines):
        line = lines[i]
        if "PROCESS:" not in line:
            i += 1
            continue
        process_line = line
        i += 1
        thr_line = lines[i] if i < len(lines) else ""
        threshold = 0.0
        b, e = thr_line.find("E ="), thr_line.find("eV")
        if b >= 0 and e >= 0:
            threshold = float(thr_line[b + 3:e])
        elif "Elastic" not in process_line and "Effective" not in process_line:
            raise ValueError("no threshold after: " + process_line)
        i += 1
        desc = lines[i].replace(" ", "") if i < len(lines) else ""
        i += 1
        while i < len(lines) and "-----" not in lines[i]:
            i += 1
        i += 1
        energies, sigmas = [], []
        while i < len(lines) and "-----" not in lines[i]:
            tok = lines[i].split()
            if len(tok) == 2:
                energies.append(float(tok[0]))
                sigmas.append(float(tok[1]))
            elif tok:
                raise ValueError("bad table row: " + lines[i])
            i += 1
        i += 1
        body = desc[desc.find("[") + 1:desc.rfind("]")]
        fields = body.split(",")
        subtype = None
        if fields[-1].lower() in ("momentum-transfer", "integral"):
            subtype = fields[-1].lower()
            fields = fields[:-1]
        kind = fields[-1]
        if kind not in ("Elastic", "Effective", "Attachment", "Ionization", "Excitation",
                        "Vibrational", "Rotational"):
            raise ValueError("invalid collision type: " + kind)
        if subtype is None:
            subtype = "momentum-transfer" if kind in ("Elastic", "Effective") else "integral"
        collision = ",".join(fields[:-1])
        entries.append({"type": kind, "threshold": threshold, "energy": np.array(energies),
                        "sigma": np.array(sigmas), "reverse": "<->" in collision,
                        "subtype": subtype, "description": collision})
    return entries


def _new_columns():
    return {k: [] for k in ("type", "is_elastic", "is_superelastic", "gas_index", "threshold",
                            "target_mass", "rel_density", "stat_weight_ratio", "w_parameter",
                            "scattering", "scattering_params")}


def _add_process(cols, ent, g, mass_amu, fraction, scattering="isotropic", params=()):
    """One LXCat entry as a process of gas `g` (ground-state target)."""
    ptype = {"Ionization": IONIZATION, "Attachment": ATTACHMENT}.get(ent["type"], CONSERVATIVE)
    cols["type"].append(ptype)
    cols["is_elastic"].append(1 if ent["type"] == "Elastic" else 0)
    cols["is_superelastic"].append(0)
    cols["gas_index"].append(g)
    cols["threshold"].append(ent["threshold"])
    cols["target_mass"].append(mass_amu * AMU)
    cols["rel_density"].append(fraction)
    cols["stat_weight_ratio"].append(0.0)
    cols["w_parameter"].append(ent["threshold"] if ptype == IONIZATION else 0.0)
    cols["scattering"].append(SCATTERING[scattering])
    prm = list(params) + [0.0, 0.0]
    cols["scattering_params"] += prm[:2]


def _mixture(cols, energies, sigmas, fractions, n_gases, names):
    offsets = np.concatenate([[0], np.cumsum([len(e) for e in energies])]).astype(np.int32)
    arrays = {
        "type": np.array(cols["type"], np.int32),
        "is_elastic": np.array(cols["is_elastic"], np.int32),
        "is_superelastic": np.array(cols["is_superelastic"], np.int32),
        "gas_index": np.array(cols["gas_index"], np.int32),
        "threshold": np.array(cols["threshold"], np.float64),
        "target_mass": np.array(cols["target_mass"], np.float64),
        "rel_density": np.array(cols["rel_density"], np.float64),
        "stat_weight_ratio": np.array(cols["stat_weight_ratio"], np.float64),
        "w_parameter": np.array(cols["w_parameter"], np.float64),
        "scattering": np.array(cols["scattering"], np.int32),
        "scattering_params": np.array(cols["scattering_params"], np.float64),
        "table_offset": offsets,
        "table_energy": np.concatenate(energies).astype(np.float64),
        "table_sigma": np.concatenate(sigmas).astype(np.float64),
        "gas_fraction": np.array(fractions[:n_gases], np.float64),
    }
    return MixtureData(arrays, n_gases, names)


def build_mixture(gas_sets, fractions):
    """Flattened process arrays for a mixture of synthetic gas sets.

    ``gas_sets`` names files ``<set>_LXCat.txt`` under ../gases (each may hold
    several gases, listed in gases.json); ``fractions`` gives one fraction per
    gas in order.  Targets are ground states with population 1, so each
    process's relative density is its gas fraction (State.h:72-85)."""
    meta = json.load(open(os.path.join(GASES, "gases.json")))
    cols = _new_columns()
    energies, sigmas, names = [], [], []
    g = 0
    for set_name in gas_sets:
        entries = read_lxcat(os.path.join(GASES, "%s_LXCat.txt" % set_name))
        k = 0
        for gas in meta[set_name]:
            procs = gas["processes"]
            for pm in procs:
                ent = entries[k]
                k += 1
                if ent["type"] == "Effective":
                    continue
                _add_process(cols, ent, g, gas["mass_amu"], fractions[g], pm["scattering"], pm["params"])
                energies.append(ent["energy"])
                sigmas.append(ent["sigma"])
                names.append("%s %s %.4g eV" % (gas["name"], ent["type"], ent["threshold"]))
            g += 1
    return _mixture(cols, energies, sigmas, fractions, g, names)


def lxcat_mixture(path, mass_amu):
    """One gas from an LXCat file: every Elastic, Excitation, Ionization and
    Attachment entry (Effective ones skipped) as an isotropically scattering
    process of a ground-state target of `mass_amu`, population 1."""
    cols = _new_columns()
    energies, sigmas, names = [], [], []
    for ent in read_lxcat(path):
        if ent["type"] == "Effective":
            continue
        _add_process(cols, ent, 0, mass_amu, 1.0)
        energies.append(ent["energy"])
        sigmas.append(ent["sigma"])
        names.append("%s %.4g eV" % (ent["type"], ent["threshold"]))
    return _mixture(cols, energies, sigmas, [1.0], 1, names)


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class MixtureData:
    """Keeps the numpy arrays alive behind an ``ekmc_mixture`` struct."""

    def __init__(self, arrays, n_gases, names):
        self.arrays = arrays
        self.names = names
        self.n_gases = n_gases
        a = arrays
        self.struct = Mixture(
            n_processes=len(a["type"]), n_gases=n_gases,
            type=_ptr(a["type"], C.c_int32), is_elastic=_ptr(a["is_elastic"], C.c_int32),
            is_superelastic=_ptr(a["is_superelastic"], C.c_int32),
            gas_index=_ptr(a["gas_index"], C.c_int32), threshold=_ptr(a["threshold"], C.c_double),
            target_mass=_ptr(a["target_mass"], C.c_double),
            rel_density=_ptr(a["rel_density"], C.c_double),
            stat_weight_ratio=_ptr(a["stat_weight_ratio"], C.c_double),
            w_parameter=_ptr(a["w_parameter"], C.c_double),
            scattering=_ptr(a["scattering"], C.c_int32),
            scattering_params=_ptr(a["scattering_params"], C.c_double),
            table_offset=_ptr(a["table_offset"], C.c_int32),
            table_energy=_ptr(a["table_energy"], C.c_double),
            table_sigma=_ptr(a["table_sigma"], C.c_double),
            gas_fraction=_ptr(a["gas_fraction"], C.c_double))

    @property
    def n_processes(self):
        return len(self.arrays["type"])


def load_workloads():
    return json.load(open(os.path.join(GASES, "workloads.json")))


def workload(name, n_electrons=None, conditions=None):
    """(MixtureData, [Condition], Numerics) for a named workload.  The
    electron count of every condition can be overridden, and a subset of the
    conditions picked by index."""
    w = load_workloads()[name]
    mix = build_mixture(w["gases"], w["fractions"])
    conds = w["conditions"] if conditions is None else [w["conditions"][i] for i in conditions]
    cs = []
    for c in conds:
        d = dict(c)
        if n_electrons is not None:
            d["n_electrons"] = n_electrons
        cs.append(Condition(**d))
    num = numerics_default()
    num.gas_temperature_effect = GASTEMP[w["gas_temperature_effect"]]
    num.ionization_operator = IONIZ[w["ionization_operator"]]
    num.initial_elec_temp_over_gas_temp = w.get("initial_te_over_tg", 0.01)
    if "max_collisions_before_ss" in w:
        num.max_collisions_before_ss = w["max_collisions_before_ss"]
    if "max_collisions_after_ss" in w:
        num.max_collisions_after_ss = w["max_collisions_after_ss"]
    if "n_integration_phases" in w:
        num.n_integration_phases = w["n_integration_phases"]
    if "n_energy_cells" in w:
        num.n_energy_cells = w["n_energy_cells"]
    if "evdf3d_cells" in w:
        num.evdf3d_cells = w["evdf3d_cells"]
    return mix, cs, num


def numerics_default():
    """Reference defaults of numericsMC (BoltzmannMC.h:52-80)."""
    n = Numerics()
    n.precision = FP64
    n.gas_temperature_effect = 0
    n.ionization_operator = 0
    n.n_interp_points = 10000
    n.initial_elec_temp_over_gas_temp = 0.01
    n.min_collisions_before_ss = 0.0
    n.max_collisions_before_ss = 1e100
    n.max_collisions_after_ss = 1e100
    n.n_energy_cells = 1000
    n.n_cos_angle_cells = 100
    n.n_radial_velocity_cells = 200
    n.n_axial_velocity_cells = 200
    n.errors_to_be_checked = 0
    n.n_integration_phases = 100
    for f in ("rel_error_mean_energy", "rel_error_flux_drift_velocity",
              "rel_error_bulk_drift_velocity", "rel_error_flux_diffusion_coeffs",
              "rel_error_bulk_diffusion_coeffs", "rel_error_power_balance"):
        setattr(n, f, 1e100)
    n.n_integration_points = 200
    n.n_integrated_ss_times = 0.0
    n.integrated_absolute_time = 0.0
    n.sync_time_x_max_coll_freq = 1.0
    n.sync_over_sampling = 1
    n.max_sync_points = 0
    n.seed = 0
    return n


# ------------------------------------------------------------- the library
_lib = None


class EngineError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("ekmc status %d: %s" % (status, msg))
        self.status = status


def lib():
    """The built engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError("libekmc.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        D, I32, I64, U64 = C.c_double, C.c_int32, C.c_int64, C.c_uint64
        pD, pI32, pU64, pU32 = _p(D), _p(I32), _p(U64), _p(C.c_uint32)
        L.ekmc_numerics_default.argtypes = [_p(Numerics)]
        L.ekmc_create.argtypes = [_p(Mixture), _p(Condition), I32, _p(Numerics), _p(C.c_void_p)]
        L.ekmc_destroy.argtypes = [C.c_void_p]
        L.ekmc_destroy.restype = None
        L.ekmc_autotune.argtypes = [_p(Mixture), _p(Condition), I32, _p(Numerics), pI32, I32, I32, I32, pI32, pD]
        L.ekmc_autotune_launch.argtypes = [_p(Mixture), _p(Condition), I32, _p(Numerics), I32, I32, _p(Numerics), pD]
        for name in ("ekmc_solve", "ekmc_initialize", "ekmc_finalize", "ekmc_reset_kernel_stats"):
            getattr(L, name).argtypes = [C.c_void_p]
        L.ekmc_step.argtypes = [C.c_void_p, I32, pI32]
        L.ekmc_get_results.argtypes = [C.c_void_p, I32, _p(Results)]
        L.ekmc_get_process_data.argtypes = [C.c_void_p, I32, pD, pU64, pD, pD]
        L.ekmc_get_eedf.argtypes = [C.c_void_p, I32, pD, pD]
        L.ekmc_get_eedf_periodic.argtypes = [C.c_void_p, I32, pD, pD, pD]
        L.ekmc_get_swarm_periodic.argtypes = [C.c_void_p, I32, pD, pD, pD, pD, pD, pD]
        L.ekmc_get_evdf.argtypes = [C.c_void_p, I32, pD, pD, pD, pD, pD]
        L.ekmc_get_evdf3d.argtypes = [C.c_void_p, I32, pD, pD]
        L.ekmc_get_series.argtypes = [C.c_void_p, I32, pI32, pD, pD, pD, pD, pD, pD, pD]
        L.ekmc_get_electrons.argtypes = [C.c_void_p, I32, pD, pD, pD, pD]
        L.ekmc_get_kernel_stats.argtypes = [C.c_void_p, _p(KernelStat), pI32]
        L.ekmc_get_status.argtypes = [C.c_void_p, I32, _p(CondStatus)]
        L.ekmc_get_event_count.argtypes = [C.c_void_p, pU64, pU64]
        L.ekmc_last_error.restype = C.c_char_p
        L.ekmc_build_tables.argtypes = [_p(Mixture), D, I32, D, pD, pD, pD, pD, pD]
        L.ekmc_op_accelerate.argtypes = [C.c_void_p, I32, I64, pD, pD, pD, pD, pD, pD]
        L.ekmc_op_collide.argtypes = [C.c_void_p, I32, D, I64, pD, pD, pD, pD, pI32, pD, pD, pD]
        L.ekmc_op_collide_noalloc.argtypes = [C.c_void_p, I32, D, I64, pD, pD, pD, pD, pI32, pD, pD, pD]
        L.ekmc_op_philox.argtypes = [I64, pU32, pU32, pU32]
        L.ekmc_op_interval_pool.argtypes = [C.c_void_p, I32, _p(I64), pD, pD, pD, pD, pU64, pU32, pD, pD]
        L.ekmc_op_tables.argtypes = [C.c_void_p, I32, D, pD, pD, pD, pD, pD]
        _lib = L
    return _lib


def check(status):
    if status != 0:
        raise EngineError(status, lib().ekmc_last_error().decode())


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def autotune(mix, conditions, numerics, candidates=None, warmup=100, steps=100):
    """ekmc_autotune: the work unit (main electrons per wavefront work unit)
    with the most collision events/s on short runs of this batch.  Returns
    (best work_chunk, {candidate: events/s})."""
    conds = list(conditions)
    arr = (Condition * len(conds))(*conds)
    cand = np.asarray(candidates if candidates is not None else [256, 512, 768, 1024, 1536, 2048], np.int32)
    rates = np.zeros(len(cand))
    best = C.c_int32()
    check(lib().ekmc_autotune(C.byref(mix.struct), arr, len(conds), C.byref(numerics),
                              cand.ctypes.data_as(C.POINTER(C.c_int32)), len(cand), warmup, steps,
                              C.byref(best), dptr(rates)))
    return best.value, {int(c): float(r) for c, r in zip(cand, rates)}


def autotune_launch(mix, conditions, numerics, warmup=100, steps=100):
    """ekmc_autotune_launch: work unit, collision-step lane threshold and
    envelope coarseness of the fastest short run.  Returns (numerics with
    the three set, events/s)."""
    conds = list(conditions)
    arr = (Condition * len(conds))(*conds)
    best = Numerics()
    rate = C.c_double()
    check(lib().ekmc_autotune_launch(C.byref(mix.struct), arr, len(conds), C.byref(numerics), warmup, steps,
                                     C.byref(best), C.byref(rate)))
    return best, rate.value


class Engine:
    """One engine handle (ekmc_create ... ekmc_destroy)."""

    def __init__(self, mix, conditions, numerics):
        self.mix = mix
        self.conditions = list(conditions)
        arr = (Condition * len(self.conditions))(*self.conditions)
        self._conds = arr
        self.numerics = numerics
        h = C.c_void_p()
        check(lib().ekmc_create(C.byref(mix.struct), arr, len(self.conditions),
                                C.byref(numerics), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().ekmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self):
        check(lib().ekmc_solve(self.h))

    def initialize(self):
        check(lib().ekmc_initialize(self.h))

    def step(self, n_sync=1):
        n = C.c_int32()
        check(lib().ekmc_step(self.h, n_sync, C.byref(n)))
        return n.value

    def finalize(self):
        check(lib().ekmc_finalize(self.h))

    def results(self, cond=0):
        r = Results()
        check(lib().ekmc_get_results(self.h, cond, C.byref(r)))
        return r

    def process_data(self, cond=0):
        P = self.mix.n_processes
        rates, counts = np.zeros(P), np.zeros(P, np.uint64)
        gain, loss = np.zeros(P), np.zeros(P)
        check(lib().ekmc_get_process_data(self.h, cond, dptr(rates),
                                          counts.ctypes.data_as(C.POINTER(C.c_uint64)),
                                          dptr(gain), dptr(loss)))
        return rates, counts, gain, loss

    def eedf(self, cond=0):
        n = self.numerics.n_energy_cells
        cells, f = np.zeros(n), np.zeros(n)
        check(lib().ekmc_get_eedf(self.h, cond, dptr(cells), dptr(f)))
        return cells, f

    def evdf3d(self, cond=0):
        """Cell centres of one axis and the normalised 3D EVDF [n, n, n] over
        (v_x, v_y, v_z), for numerics.evdf3d_cells = n > 0."""
        n = self.numerics.evdf3d_cells
        cells, f = np.zeros(n), np.zeros(n * n * n)
        check(lib().ekmc_get_evdf3d(self.h, cond, dptr(cells), dptr(f)))
        return cells, f.reshape(n, n, n)

    def series(self, cond=0, capacity=1 << 20):
        n = C.c_int32(0)
        check(lib().ekmc_get_series(self.h, cond, C.byref(n), None, None, None, None, None, None, None))
        m = n.value
        out = {k: np.zeros(m * w) for k, w in (("times", 1), ("mean_energy", 1), ("mean_position", 3),
                                                ("mean_velocity", 3), ("flux_diffusion", 9),
                                                ("bulk_velocity", 3), ("bulk_diffusion", 9))}
        n = C.c_int32(m)
        check(lib().ekmc_get_series(self.h, cond, C.byref(n), *[dptr(out[k]) for k in (
            "times", "mean_energy", "mean_position", "mean_velocity", "flux_diffusion",
            "bulk_velocity", "bulk_diffusion")]))
        return out

    def electrons(self, cond=0):
        n = self.conditions[cond].n_electrons
        pos, vel, e, t = np.zeros(3 * n), np.zeros(3 * n), np.zeros(n), np.zeros(n)
        check(lib().ekmc_get_electrons(self.h, cond, dptr(pos), dptr(vel), dptr(e), dptr(t)))
        return pos.reshape(n, 3), vel.reshape(n, 3), e, t

    def status(self, cond=0):
        st = CondStatus()
        check(lib().ekmc_get_status(self.h, cond, C.byref(st)))
        return st

    def kernel_stats(self):
        n = C.c_int32(64)
        arr = (KernelStat * 64)()
        check(lib().ekmc_get_kernel_stats(self.h, arr, C.byref(n)))
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(n.value)}

    def reset_kernel_stats(self):
        check(lib().ekmc_reset_kernel_stats(self.h))

    def event_count(self):
        ev, real = C.c_uint64(), C.c_uint64()
        check(lib().ekmc_get_event_count(self.h, C.byref(ev), C.byref(real)))
        return ev.value, real.value

    # batch operators
    def op_accelerate(self, cond, time, dt, pos, vel):
        n = len(time)
        pos = np.ascontiguousarray(pos, np.float64).copy()
        vel = np.ascontiguousarray(vel, np.float64).copy()
        energy, gain = np.zeros(n), np.zeros(n)
        check(lib().ekmc_op_accelerate(self.h, cond, n, dptr(np.ascontiguousarray(time, np.float64)),
                                       dptr(np.ascontiguousarray(dt, np.float64)), dptr(pos),
                                       dptr(vel), dptr(energy), dptr(gain)))
        return pos, vel, energy, gain

    def op_tables(self, cond, max_energy):
        """Tables of `cond` rebuilt on the device at max_energy (engine precision)."""
        P, I = self.mix.n_processes, self.numerics.n_interp_points
        interp, cumul = np.zeros(I * P), np.zeros(I * P)
        tot, mx, step = np.zeros(I), np.zeros(I), C.c_double()
        check(lib().ekmc_op_tables(self.h, cond, max_energy, dptr(interp), dptr(cumul), dptr(tot),
                                   dptr(mx), C.byref(step)))
        return interp.reshape(I, P), cumul.reshape(I, P), tot, mx, step.value

    def op_collide(self, cond, max_energy, vel, energy, trial_freq, variates, noalloc=False):
        """ekmc_op_collide, or with noalloc=True ekmc_op_collide_noalloc."""
        fn = lib().ekmc_op_collide_noalloc if noalloc else lib().ekmc_op_collide
        n = len(energy)
        vel = np.ascontiguousarray(vel, np.float64).copy()
        energy = np.ascontiguousarray(energy, np.float64).copy()
        pid = np.zeros(n, np.int32)
        ejv, eje, de = np.zeros(3 * n), np.zeros(n), np.zeros(n)
        check(fn(self.h, cond, max_energy, n, dptr(vel), dptr(energy),
                 dptr(np.ascontiguousarray(trial_freq, np.float64)),
                 dptr(np.ascontiguousarray(variates, np.float64)),
                 pid.ctypes.data_as(C.POINTER(C.c_int32)), dptr(ejv), dptr(eje), dptr(de)))
        return pid, vel, energy, ejv, eje, de


    def op_interval_pool(self, cond, capacity=None):
        """One interval; the pool of `cond` after its flights, before
        population control (dict of arrays), with t_sync and the trial
        frequency of the interval."""
        if capacity is None:
            capacity = 4 * self.conditions[cond].n_electrons + 4096
        out = {"pos": np.zeros(3 * capacity), "vel": np.zeros(3 * capacity), "energy": np.zeros(capacity),
               "time": np.zeros(capacity), "stream": np.zeros(capacity, np.uint64),
               "event": np.zeros(capacity, np.uint32)}
        n = C.c_int64(capacity)
        ts, tr = C.c_double(), C.c_double()
        check(lib().ekmc_op_interval_pool(self.h, cond, C.byref(n), dptr(out["pos"]), dptr(out["vel"]),
                                          dptr(out["energy"]), dptr(out["time"]),
                                          out["stream"].ctypes.data_as(C.POINTER(C.c_uint64)),
                                          out["event"].ctypes.data_as(C.POINTER(C.c_uint32)),
                                          C.byref(ts), C.byref(tr)))
        m = n.value
        res = {k: (v[:3 * m].reshape(m, 3) if k in ("pos", "vel") else v[:m]) for k, v in out.items()}
        return res, ts.value, tr.value


def build_tables(mix, gas_density, n_interp, max_energy):
    P = mix.n_processes
    interp, cumul = np.zeros(n_interp * P), np.zeros(n_interp * P)
    tot, mx, step = np.zeros(n_interp), np.zeros(n_interp), C.c_double()
    check(lib().ekmc_build_tables(C.byref(mix.struct), gas_density, n_interp, max_energy,
                                  dptr(interp), dptr(cumul), dptr(tot), dptr(mx), C.byref(step)))
    return interp.reshape(n_interp, P), cumul.reshape(n_interp, P), tot, mx, step.value


def philox(counter, key):
    counter = np.ascontiguousarray(counter, np.uint32)
    key = np.ascontiguousarray(key, np.uint32)
    n = counter.shape[0]
    out = np.zeros((n, 4), np.uint32)
    P32 = C.POINTER(C.c_uint32)
    check(lib().ekmc_op_philox(n, counter.ctypes.data_as(P32), key.ctypes.data_as(P32),
                               out.ctypes.data_as(P32)))
    return out
