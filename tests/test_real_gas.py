"""Real-gas state model of the setup reader: wildcard states, state property
functions, state families, densities, Elastic derived from Effective, and
momentum-transfer cross sections from angular distributions.

The reference's own input files for N2 and CO2 (setup files, LXCat files,
property databases) are kept as data fixtures under tests/golden/input.  Each
check recomputes the expected values here from the physics the reference
states (Boltzmann populations, the rotational degeneracy of N2, the rigid
rotor, Klein-Rosseland, the closed-form angular integral of the Born-dipole
distribution) rather than from the C++ code:
StatePropertyFunctions.h:22-866, State.h:72-85, EedfState.C:38-95,
EedfGas.C:29-301, Collision.C:124-353, AngularDistributionFunctions.h:32-105.
"""
import json
import math
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "loki-mc_amd", "bin", "lokimc_gpu")
INPUT = os.path.join(ROOT, "tests", "golden", "input")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="lokimc_gpu not built")

K_EV = 1.38064852e-23 / 1.6021766208e-19   # Constant::boltzmannInEV


def fl16(x):
    """A number as the reference's parser returns it (%.16g round trip)."""
    return float("%.16g" % x)


def dry(setup, inp=INPUT, cwd=None):
    # a failing run writes errorLog.txt into its working directory
    with tempfile.TemporaryDirectory() as tmp:
        return subprocess.run([EXE, setup, "--input", inp, "--dry-run"], capture_output=True, text=True, timeout=300,
                              cwd=cwd or tmp)


def states_of(description):
    """Target and product state names of a collision description."""
    body = description.split(",")[:-1]
    body = ",".join(body)
    arrow = "<->" if "<->" in body else "->"
    left, right = body.split(arrow)

    def side(text):
        parts, depth, cur = [], 0, ""
        for ch in text:
            if ch == "+" and depth == 0:
                parts.append(cur)
                cur = ""
                continue
            depth += ch == "("
            depth -= ch == ")"
            cur += ch
        parts.append(cur)
        return [re.sub(r"^[0-9]+", "", q) for q in parts if q and q not in ("e", "E")]

    l, r = side(left), side(right)
    return l[0], r


def parse_state(name):
    gas, inner = name[:-1].split("(", 1)
    fields = inner.split(",")
    out = {"gas": gas, "ele": fields[0], "v": None, "J": None}
    for f in fields[1:]:
        if f.startswith("v="):
            out["v"] = f[2:]
        elif f.startswith("J="):
            out["J"] = f[2:]
    return out


def property_file(path):
    out = {}
    for line in open(path):
        line = line.split("%")[0].strip()
        if line:
            k, v = line.split()
            out[k] = v
    return out


# ------------------------------------------------------------------ N2

@pytest.fixture(scope="module")
def n2():
    r = dry(os.path.join(INPUT, "Nitrogen", "N2_swarm_setup.in"))
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def test_n2_vibrational_and_rotational_densities_are_boltzmann_at_the_gas_temperature(n2):
    T = 300.0
    names = set()
    for c in n2["collisions"]:
        t, prods = states_of(c["description"])
        names.add(t)
        names.update(prods)
    vib = sorted({n for n in names if parse_state(n)["v"] is not None and parse_state(n)["J"] is None and
                  parse_state(n)["ele"] == "X"}, key=lambda n: int(parse_state(n)["v"]))
    rot = sorted({n for n in names if parse_state(n)["J"] is not None}, key=lambda n: int(parse_state(n)["J"]))
    # every rotational state of the LXCat file belongs to N2(X,v=0)
    assert rot and all(parse_state(n)["v"] == "0" for n in rot)
    # vibrational levels: energies from the file, statistical weight 1
    energies = {k: fl16(float(v)) for k, v in property_file(os.path.join(INPUT, "Nitrogen", "N2energyVibLevels.txt")).items()}
    ev = np.array([energies[n] for n in vib])
    wv = np.exp(-(ev - ev.min()) / (K_EV * T))
    pv = dict(zip(vib, wv / wv.sum()))
    # rotational levels of N2(X,v=0): rigid rotor with B from the database
    # (an expression in cm-1 converted to eV), N2 nuclear-spin degeneracy
    B = fl16(6.626070040e-34 * 299792458 * 1.99824 * 100 / 1.6021766208e-19)
    J = np.array([int(parse_state(n)["J"]) for n in rot], dtype=float)
    er = B * J * (J + 1.0)
    gJ = 3.0 * (1.0 + 0.5 * (1.0 + (-1.0) ** J)) * (2.0 * J + 1.0)
    wr = gJ * np.exp(-(er - er.min()) / (K_EV * T))
    pr = dict(zip(rot, wr / wr.sum()))
    gw = dict(zip(rot, gJ))

    def density(n):
        s = parse_state(n)
        if s["J"] is not None:
            return pr[n] * pv["N2(X,v=0)"] * 1.0 * 1.0
        if s["v"] is not None:
            return pv[n] * 1.0 * 1.0
        return 1.0 if s["ele"] == "X" else 0.0

    checked_rot = 0
    for c in n2["collisions"]:
        t, prods = states_of(c["description"])
        assert c["target_density"] == pytest.approx(density(t), rel=1e-12, abs=0), c["description"]
        if "<->" in c["description"]:
            assert c["product_density"] == pytest.approx(density(prods[0]), rel=1e-12, abs=0), c["description"]
            if c["type"] == "Rotational":
                assert c["stat_weight_ratio"] == pytest.approx(gw[t] / gw[prods[0]], rel=1e-15)
                checked_rot += 1
    assert checked_rot > 10
    # every process's density is its collision's target (or, superelastic, product) density
    coll = {c["description"]: c for c in n2["collisions"]}
    for p in n2["processes"]:
        c = coll[p["description"]]
        want = c["product_density"] if p["superelastic"] else c["target_density"]
        assert p["rel_density"] == want


def test_n2_rotational_populations_sum_to_one_and_alternate_with_nuclear_spin(n2):
    dens = {}
    for c in n2["collisions"]:
        t, prods = states_of(c["description"])
        if parse_state(t)["J"] is not None:
            dens[int(parse_state(t)["J"])] = c["target_density"]
    # even J carry nuclear-spin weight 6, odd J weight 3 (rotationalDegeneracy_N2)
    assert all(j % 2 == 0 for j in dens) or any(j % 2 == 1 for j in dens)
    assert 0.0 < sum(dens.values()) <= 1.0


def _n2_variant(tmp_path, replace):
    inp = tmp_path / "in"
    shutil.copytree(INPUT, inp, copy_function=shutil.copyfile)
    p = inp / "Nitrogen" / "N2_swarm_setup.in"
    text = p.read_text()
    for a, b in replace:
        assert a in text
        text = text.replace(a, b)
    p.write_text(text)
    return dry(str(p), str(inp), cwd=str(tmp_path))


def test_unnormalized_vibrational_distribution_is_rejected(tmp_path):
    r = _n2_variant(tmp_path, [("N2(X,v=*) = boltzmannPopulation@gasTemperature", "N2(X,v=*) = 0.5")])
    assert r.returncode != 0
    assert "Vibrational distribution N2(X,v=*) is not properly normalized" in r.stderr


def test_unknown_property_function_is_rejected(tmp_path):
    r = _n2_variant(tmp_path, [("N2(X,v=*) = boltzmannPopulation@gasTemperature", "N2(X,v=*) = fooPopulation@gasTemperature")])
    assert r.returncode != 0
    assert "Trying to use the property function 'fooPopulation'" in r.stderr


def test_population_function_with_a_numeric_temperature(tmp_path):
    # the argument is a number instead of a working-condition name: 2000 K
    r = _n2_variant(tmp_path, [("N2(X,v=*) = boltzmannPopulation@gasTemperature", "N2(X,v=*) = boltzmannPopulation@2000")])
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    energies = {k: fl16(float(v)) for k, v in property_file(os.path.join(INPUT, "Nitrogen", "N2energyVibLevels.txt")).items()}
    vib = [n for n in energies if any(states_of(c["description"])[0] == n or n in states_of(c["description"])[1]
                                      for c in d["collisions"])]
    ev = np.array([energies[n] for n in vib])
    w = np.exp(-(ev - ev.min()) / (K_EV * 2000.0))
    p = dict(zip(vib, w / w.sum()))
    for c in d["collisions"]:
        t, _ = states_of(c["description"])
        s = parse_state(t)
        if s["v"] is not None and s["J"] is None:
            assert c["target_density"] == pytest.approx(p[t], rel=1e-12)


# ------------------------------------------------------------------ CO2

def _lxcat_block(path, kind, target):
    """Energies and values of the first `kind` collision of `target`."""
    lines = open(path).read().splitlines()
    for i, l in enumerate(lines):
        if l.strip() == kind and lines[i + 1].strip() == target:
            j = i
            while not lines[j].startswith("-----"):
                j += 1
            e, s = [], []
            j += 1
            while not lines[j].startswith("-----"):
                a, b = lines[j].split()
                e.append(float(a))
                s.append(float(b))
                j += 1
            return e, s
    raise AssertionError("no %s block for %s" % (kind, target))


def _lin(x, y, v):
    """Linear interpolation as gsl_interp_linear, zero outside [x0, xn]."""
    if v < x[0] or v > x[-1]:
        return 0.0
    i = int(np.searchsorted(x, v, side="right")) - 1
    i = min(max(i, 0), len(x) - 2)
    return y[i] + (v - x[i]) / (x[i + 1] - x[i]) * (y[i + 1] - y[i])


def test_co2_elastic_is_the_effective_minus_weighted_inelastic_cross_sections():
    r = dry(os.path.join(INPUT, "CO2", "CO2_swarm_setup.in"))
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    elastic = [c for c in d["collisions"] if c["type"] == "Elastic"]
    assert len(elastic) == 1 and elastic[0]["description"] == "e+CO2(X)->e+CO2(X),Elastic"
    # derived Elastic is the last process of the gas (appended after the LXCat collisions)
    assert d["processes"][-1]["elastic"] == 1
    e, s = _lxcat_block(os.path.join(INPUT, "CO2", "CO2_LXCat.txt"), "EFFECTIVE", "CO2")
    if e[0] != 0:
        e, s = [0.0] + e, [s[0]] + s
    pops = {k: fl16(float(v)) for k, v in property_file(os.path.join(INPUT, "CO2", "CO2_effectivePop.txt")).items()}
    weights = {"CO2(X,v=000)": 1, "CO2(X,v=010)": 2}
    want = np.array(s, dtype=float)
    for c in d["collisions"]:
        if c["type"] in ("Elastic", "Effective"):
            continue
        t, prods = states_of(c["description"])
        x, y, thr = c["integral_energy"], c["integral"], c["threshold"]
        for i, v in enumerate(e):
            if v > thr:
                want[i] -= pops.get(t, 0.0) * _lin(x, y, v)
        if "<->" in c["description"]:
            ratio = weights[t] / weights[prods[0]]
            for i, v in enumerate(e):
                if v == 0:
                    continue
                sig = _lin(x, y, v + thr) if v + thr > thr else 0.0
                want[i] -= pops.get(prods[0], 0.0) * (1.0 + thr / v) * (sig * ratio)
    want = np.maximum(want, 0.0)
    got = np.array(elastic[0]["integral"])
    assert elastic[0]["integral_energy"] == e
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-40)
    assert np.all(got <= np.array(s) + 1e-40) and np.all(got >= 0)


# -------------------------------------------------------- anisotropic

SYN_LXCAT = """\
PROCESS: E + X -> E + X, Elastic
PARAM.:  m/M = 0.0000136
COMMENT: [e + X(X) -> e + X(X), Elastic] synthetic
-----------------------------
 0.0   1.0e-19
 10.0  1.0e-19
 1000.0  1.0e-19
-----------------------------

PROCESS: E + X <-> E + X (J=0-J=1), Rotational
PARAM.:  E = 0.01 eV
COMMENT: [e + X(X,v=0,J=0) <-> e + X(X,v=0,J=1), Rotational] synthetic
-----------------------------
 0.01   0.0
 0.02   2.0e-20
 0.5    5.0e-20
 5.0    3.0e-20
 100.0  1.0e-20
-----------------------------
"""


def test_born_dipole_momentum_transfer_matches_the_closed_form_angular_integral(tmp_path):
    inp = tmp_path / "in"
    inp.mkdir()
    (inp / "X_LXCat.txt").write_text(SYN_LXCAT)
    (inp / "t.in").write_text("""workingConditions:
  reducedElecField: 10
  gasPressure: 133.32
  gasTemperature: 300
  excitationFrequency: 0
  elecFieldAngle: 180
  reducedMagField: 0
electronKinetics:
  isOn: true
  eedfType: boltzmannMC
  ionizationOperatorType: equalSharing
  LXCatFiles:
    - X_LXCat.txt
  gasProperties:
    mass:
      - X = 4e-26
    fraction:
      - X = 1
  stateProperties:
    statisticalWeight:
      - X(X,v=0,J=0) = 1
      - X(X,v=0,J=1) = 3
    population:
      - X(X) = 1
      - X(X,v=0) = 1
      - X(X,v=0,J=0) = 1
  anisotropicScattering:
    isOn: true
    angleNumber: 5000
    collisions:
      - group;X;Rotational;bornDipole
  numericsMC:
    nElectrons: 1E3
    gasTemperatureEffect: false
    nIntegrationPoints: 1E3
    nIntegratedSSTimes: 2
""")
    r = dry(str(inp / "t.in"), str(inp), cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    rot = [c for c in d["collisions"] if c["type"] == "Rotational"]
    assert len(rot) == 1
    c = rot[0]
    assert c["mom_transf_energy"] == c["integral_energy"]
    thr = c["threshold"]

    def pdf(E, theta):
        after = E - thr
        if after <= 0:
            return np.full_like(theta, 1.0 / (4.0 * math.pi))
        q2 = after + E - 2.0 * math.sqrt(E * after) * np.cos(theta)
        return (1.0 / (4.0 * math.pi) * math.sqrt(E) * math.sqrt(after) / q2 /
                math.log((math.sqrt(after) + math.sqrt(E)) / math.sqrt(thr)))

    def trapezoid(E, n):
        theta = np.linspace(0.0, math.pi, n)
        f = np.cos(theta) * np.sin(theta) * pdf(E, theta)
        return math.pi * (math.pi / (n - 1)) * (f[0] + 2.0 * f[1:-1].sum() + f[-1])

    for E, si, sm in zip(c["integral_energy"], c["integral"], c["mom_transf"]):
        after = E - thr
        if after <= 0:
            exact = 0.0   # isotropic at and below threshold: no momentum-transfer correction
        else:
            a, b = E + after, 2.0 * math.sqrt(E * after)
            L = math.log((math.sqrt(after) + math.sqrt(E)) / math.sqrt(thr))
            # 2 pi Int_0^pi cos sin I d(theta), I = sqrt(E E')/(4 pi q^2 L), q^2 = a - b cos
            exact = math.sqrt(E * after) / (2.0 * L) * ((a / b ** 2) * math.log((a + b) / (a - b)) - 2.0 / b)
        # the rule the reference applies (5000-point trapezoid), and its limit
        assert sm == pytest.approx(si * (1.0 - trapezoid(E, 5000)), rel=1e-9, abs=1e-40), E
        if after > 0:   # the rule converges to the closed form as the angles are refined
            assert abs(trapezoid(E, 40000) - exact) < abs(trapezoid(E, 5000) - exact) + 1e-15, E
        # where the 5000 angles resolve the forward peak (its width in angle
        # is sqrt(q2_min / sqrt(E E')), narrowing as E/threshold grows) the
        # rule is within 1e-2 of the closed form; above, only the rule holds
        if si > 0 and after > 0:
            width = math.sqrt((math.sqrt(E) - math.sqrt(after)) ** 2 / math.sqrt(E * after))
            if width > 10 * math.pi / 4999:
                assert 1.0 - sm / si == pytest.approx(exact, abs=1e-2), E
    # the integral cross section (the one the flights sample) is the LXCat one
    assert c["integral"] == [0.0, 2.0e-20, 5.0e-20, 3.0e-20, 1.0e-20]
    # target and product of the rotational collision: J=0 holds the whole population
    assert c["target_density"] == 1.0 and c["product_density"] == 0.0
    assert c["stat_weight_ratio"] == pytest.approx(1.0 / 3.0)


# ---------------------------------------------------------------- GPU runs

def _run_variant(tmp_path, setup_rel, replace, folder):
    inp = tmp_path / "in"
    shutil.copytree(INPUT, inp, copy_function=shutil.copyfile)
    p = inp / setup_rel
    text = p.read_text()
    for a, b in replace:
        assert a in text, a
        text = text.replace(a, b)
    p.write_text(text)
    r = subprocess.run([EXE, str(p), "--input", str(inp), "--output", str(tmp_path / "Output"), "--seed", "11"],
                       capture_output=True, text=True, timeout=600, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    return tmp_path / "Output" / folder


def _check_job(sub):
    for name in ("eedf.txt", "swarmParameters.txt", "rateCoefficients.txt", "rateCoefficientsMC.txt",
                 "powerBalance.txt", "MCSimDetails.txt"):
        assert (sub / name).exists(), name
    eedf = np.loadtxt(sub / "eedf.txt", skiprows=1)
    de = eedf[1, 0] - eedf[0, 0]
    assert abs(np.sum(eedf[:, 1] * np.sqrt(eedf[:, 0])) * de - 1.0) < 1e-6
    pb = (sub / "powerBalance.txt").read_text()
    assert float(re.search(r"Relative Power Balance = (\S+)%", pb).group(1)) < 1.0
    sw = (sub / "swarmParameters.txt").read_text()
    return float(re.search(r"Mean energy = (\S+) \(eV\)", sw).group(1)), pb


@pytest.mark.gpu
def test_n2_with_vibrational_and_rotational_states_runs_to_a_balanced_steady_state(tmp_path):
    out = _run_variant(tmp_path, "Nitrogen/N2_swarm_setup.in",
                       [("reducedElecField: [5,10,50,100]", "reducedElecField: [5,50]"),
                        ("nIntegrationPoints: 1E4", "nIntegrationPoints: 2E3")], "swarm_N2_short")
    m5, pb5 = _check_job(out / "reducedElecField_5")
    m50, _ = _check_job(out / "reducedElecField_50")
    assert 0.0 < m5 < m50
    # at 5 Td, superelastic rotational collisions return power to the electrons
    assert float(re.search(r"Rotational superelastic collisions = (\S+)", pb5).group(1)) > 0
    assert float(re.search(r"Rotational inelastic collisions = (\S+)", pb5).group(1)) < 0
    rc = (out / "reducedElecField_5" / "rateCoefficientsMC.txt").read_text()
    assert "e+N2(X,v=0,J=0)<->e+N2(X,v=0,J=2),Rotational" in rc


@pytest.mark.gpu
def test_co2_with_an_elastic_cross_section_derived_from_the_effective_runs(tmp_path):
    out = _run_variant(tmp_path, "CO2/CO2_swarm_setup.in",
                       [("reducedElecField: [1,10,50,100,1000]", "reducedElecField: [10,100]"),
                        ("nIntegrationPoints: 1E4", "nIntegrationPoints: 2E3")], "swarm_CO2_short")
    m10, _ = _check_job(out / "reducedElecField_10")
    m100, _ = _check_job(out / "reducedElecField_100")
    assert 0.0 < m10 < m100
    rc = (out / "reducedElecField_10" / "rateCoefficientsMC.txt").read_text()
    assert "e+CO2(X)->e+CO2(X),Elastic" in rc


def test_rate_coefficient_ids_follow_the_reference_collision_order():
    # Collision::ID counts collisions as they are created: the LXCat entries
    # in file order (the Effective one included), then the Elastic derived
    # from the Effective; the rate-coefficient files print ID + 1 and list the
    # collisions of the flights first, then the Effective ones
    r = dry(os.path.join(INPUT, "CO2", "CO2_swarm_setup.in"))
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    kinds = [(c["kind"], c["ref_id"], c["type"]) for c in d["collisions"]]
    assert kinds[0] == (0, 0, "Attachment")
    assert [k[1] for k in kinds[1:15]] == list(range(2, 16))
    assert kinds[15] == (0, 16, "Elastic")
    assert kinds[16] == (1, 1, "Effective")
    assert len(kinds) == 17
