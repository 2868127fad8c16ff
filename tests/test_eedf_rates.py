"""EEDF-convolution rate coefficients of the host post-processing
(Collision::evaluateRateCoeff, Collision.C:355-397, written to
rateCoefficients.txt) against closed forms on a Maxwellian EEDF
f(e) = 2 / sqrt(pi) kT^-3/2 exp(-e / kT), normalised as Int f sqrt(e) de = 1:

  constant sigma0:                 k = sigma0 sqrt(8 e kT / (pi m_e))
  sigma = a (e - E) above E:       k = sqrt(2 e / m_e) a 2 / sqrt(pi) kT^1/2 exp(-E / kT) (2 kT + E)
  superelastic of a reversible one: k_sup / k_ine = (g_lower / g_upper) exp(E / kT)
                                   (detailed balance, exact on the cell grid when E is a node)

CPU only: a small harness links the driver's host sources and calls the
function that writes those files."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "loki-mc_amd", "host")
LIB = os.path.join(ROOT, "loki-mc_amd", "lib")
Q, ME = 1.6021766208e-19, 9.10938356e-31

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(LIB, "libekmc.so")), reason="libekmc.so not built")

SIGMA0, SLOPE, SLOPE2, E1, E2 = 1e-19, 2e-21, 5e-22, 1.5, 4.0

LXCAT = """\
PROCESS: E + X(gnd) -> E + X(gnd), Elastic
PARAM.:  m/M = 1.0e-05
COMMENT: [e + X(gnd) -> e + X(gnd), Elastic]
-----------------------------
0.0 {s0:.6e}
1000.0 {s0:.6e}
-----------------------------
PROCESS: E + X(exc) -> E + X(exc), Elastic
PARAM.:  m/M = 1.0e-05
COMMENT: [e + X(exc) -> e + X(exc), Elastic]
-----------------------------
0.0 {s0:.6e}
1000.0 {s0:.6e}
-----------------------------
PROCESS: E + X(gnd) <-> E + X(exc), Excitation
PARAM.:  E = {e1} eV
COMMENT: [e + X(gnd) <-> e + X(exc), Excitation]
-----------------------------
{e1} 0.0
1000.0 {s1:.6e}
-----------------------------
PROCESS: E + X(gnd) -> E + X(hi), Excitation
PARAM.:  E = {e2} eV
COMMENT: [e + X(gnd) -> e + X(hi), Excitation]
-----------------------------
{e2} 0.0
1000.0 {s2:.6e}
-----------------------------
""".format(s0=SIGMA0, e1=E1, s1=SLOPE * (1000 - E1), e2=E2, s2=SLOPE2 * (1000 - E2))

SETUP = """\
workingConditions:
  gasPressure: 133.32
  gasTemperature: 300
  reducedElecField: 10
  reducedMagField: 0
  elecFieldAngle: 180
  excitationFrequency: 0
electronKinetics:
  isOn: true
  eedfType: boltzmannMC
  ionizationOperatorType: equalSharing
  LXCatFiles:
    - X_LXCat.txt
  gasProperties:
    mass:
      - X = 40*1.66053904e-27
    fraction:
      - X = 1
  stateProperties:
    population:
      - X(gnd) = 1.0
      - X(exc) = 0.0
    statisticalWeight:
      - X(gnd) = 1
      - X(exc) = 3
  numericsMC:
    nElectrons: 1000
"""


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    d = tmp_path_factory.mktemp("probe")
    exe = d / "probe"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), "-I", HOST,
                    os.path.join(ROOT, "tests", "native", "eedf_rates_probe.cpp")] +
                   [os.path.join(HOST, s) for s in ("outputs.cpp", "gas_model.cpp", "setup_file.cpp")] +
                   ["-L", LIB, "-lekmc", "-Wl,-rpath," + LIB, "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib",
                    "-o", str(exe)], check=True)
    (d / "X_LXCat.txt").write_text(LXCAT)
    (d / "t.in").write_text(SETUP)
    return exe, d


def _rates(probe, kT, step, n):
    exe, d = probe
    out = subprocess.run([str(exe), str(d / "t.in"), str(d), repr(kT), repr(step), str(n)], check=True,
                         stdout=subprocess.PIPE, text=True).stdout
    rows = [line.split() for line in out.strip().split("\n")]
    return {r[0]: (float(r[1]), float(r[2])) for r in rows}


@pytest.mark.parametrize("kT", [0.5, 1.0, 2.0])
def test_maxwellian_rate_coefficients_match_closed_forms(probe, kT):
    step = kT / 64.0                       # E1 = 1.5 eV and E2 = 4 eV lie on nodes
    n = int(40 * kT / step)                # up to 40 kT
    r = _rates(probe, kT, step, n)
    el = r["e+X(gnd)->e+X(gnd),Elastic"][0]
    ex, sup = r["e+X(gnd)<->e+X(exc),Excitation"]
    hi, hi_sup = r["e+X(gnd)->e+X(hi),Excitation"]
    v = np.sqrt(2 * Q / ME)
    lin = lambda a, E: v * a * 2 / np.sqrt(np.pi) * np.sqrt(kT) * np.exp(-E / kT) * (2 * kT + E)
    assert el == pytest.approx(SIGMA0 * np.sqrt(8 * Q * kT / (np.pi * ME)), rel=1e-4)
    assert ex == pytest.approx(lin(SLOPE, E1), rel=1e-4)
    assert hi == pytest.approx(lin(SLOPE2, E2), rel=1e-4)
    assert np.isnan(hi_sup)
    # detailed balance: g(gnd) / g(exc) = 1 / 3
    assert sup / ex == pytest.approx(np.exp(E1 / kT) / 3.0, rel=1e-12)
