"""The C-ABI library: it loads, exports every entry point include/ekmc.h
declares, its host-side table builder reproduces the oracle's tables bit for
bit, and its struct layouts match the header (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import ekmc
import oracle_py as O
from conftest import HAS_GPU, ROOT
from helpers import condition


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "ekmc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ekmc_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    lib = C.CDLL(ekmc.LIB_PATH)
    names = _declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_sizes_match_header():
    # sizes fixed by the field lists of include/ekmc.h on LP64
    assert C.sizeof(ekmc.Condition) == 56
    assert C.sizeof(ekmc.Mixture) == 8 + 15 * 8
    assert C.sizeof(ekmc.KernelStat) == 48 + 16
    n = ekmc.Numerics()
    ekmc.lib().ekmc_numerics_default(C.byref(n))
    ref = ekmc.numerics_default()
    for name, _ in ekmc.Numerics._fields_:
        assert getattr(n, name) == getattr(ref, name), name


ABI_STRUCTS = {"ekmc_mixture": "Mixture", "ekmc_condition": "Condition", "ekmc_numerics": "Numerics",
               "ekmc_results": "Results", "ekmc_kernel_stat": "KernelStat", "ekmc_cond_status": "CondStatus"}


def test_struct_layouts_match_a_c_compiler(tmp_path):
    """sizeof and offsetof of every field of every ABI struct, as gcc lays
    out include/ekmc.h, equal the ctypes mirrors the tests and the bench use."""
    import subprocess
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "ekmc.h"', "int main(void) {"]
    for cname, pyname in ABI_STRUCTS.items():
        st = getattr(ekmc, pyname)
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for field, _ in st._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, field, cname, field))
    lines += ["return 0;", "}"]
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, stdout=subprocess.PIPE, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line.strip():
            a, b, v = line.split()
            got[(a, b)] = int(v)
    assert C.sizeof(ekmc.Mixture) == got[("ekmc_mixture", "size")]
    for cname, pyname in ABI_STRUCTS.items():
        st = getattr(ekmc, pyname)
        assert C.sizeof(st) == got[(cname, "size")], cname
        for field, _ in st._fields_:
            assert getattr(st, field).offset == got[(cname, field)], (cname, field)


@pytest.mark.parametrize("wl", ["cfg1_atomic40_dc", "cfg3_o2like_batched", "cfg4_air_ac"])
def test_host_tables_bit_exact_with_oracle(wl):
    mix, conds, num = ekmc.workload(wl)
    num.n_interp_points = 3000
    for maxE in (3.7, 45.0, 999.0):
        interp, cumul, tot, mx, step = ekmc.build_tables(mix, conds[0].gas_density, 3000, maxE)
        m = O.Model(mix, conds[0], num)
        i2, c2, t2, m2, s2 = m.tables(maxE)
        assert step == s2
        assert np.array_equal(interp, i2)
        assert np.array_equal(cumul, c2)
        assert np.array_equal(tot, t2)
        assert np.array_equal(mx, m2)


def test_superelastic_tables_bit_exact_with_oracle():
    from helpers import mixture
    e = np.logspace(-3, 3, 300)
    procs = [dict(type=0, elastic=True, energy=e, sigma=1e-20 * np.ones_like(e))]
    for thr in (0.5, 3.0):
        ee = e[e >= thr]
        procs.append(dict(type=0, threshold=thr, energy=ee, sigma=2e-21 * np.log(ee / thr) / (ee / thr)))
        procs.append(dict(type=0, super=True, threshold=thr, statw=1.0 / 3.0))
    mix = mixture(procs)
    cond = condition()
    num = ekmc.numerics_default()
    num.n_interp_points = 4000
    a = ekmc.build_tables(mix, cond.gas_density, 4000, 25.0)
    b = O.Model(mix, cond, num).tables(25.0)
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x, y)


def test_invalid_mixture_is_rejected_with_a_message():
    from helpers import mixture
    mix = mixture([dict(type=0, super=True, threshold=1.0)])
    with pytest.raises(ekmc.EngineError) as ei:
        ekmc.build_tables(mix, 3e22, 100, 10.0)
    assert ei.value.status == 1
    assert "superelastic" in str(ei.value)


@pytest.mark.skipif(HAS_GPU, reason="checks the no-device error path")
def test_create_without_device_reports_no_device():
    mix, conds, num = ekmc.workload("cfg1_atomic40_dc", n_electrons=100)
    with pytest.raises(ekmc.EngineError) as ei:
        ekmc.Engine(mix, conds, num)
    assert ei.value.status == 6


def test_lxcat_reader_on_generated_gases():
    for name, counts in (("atomic40", [4]), ("n2like", [18]), ("o2like", [22]), ("mix5", [40] * 5)):
        ents = ekmc.read_lxcat(os.path.join(ekmc.GASES, "%s_LXCat.txt" % name))
        assert len(ents) == sum(counts)
        for ent in ents:
            assert np.all(np.diff(ent["energy"]) > 0)
            if ent["type"] not in ("Elastic", "Effective"):
                assert ent["energy"][0] == pytest.approx(ent["threshold"])


def test_lxcat_reader_on_reference_format_fixture():
    """The Reid ramp cross sections shipped with the reference
    (Code/Input/ReidRampGas/ReidRampGas_LXCat.txt), kept as a data fixture."""
    path = os.path.join(ROOT, "tests", "golden", "reid_ramp_LXCat.txt")
    ents = ekmc.read_lxcat(path)
    assert [e["type"] for e in ents] == ["Elastic", "Excitation"]
    assert ents[0]["subtype"] == "momentum-transfer" and ents[1]["subtype"] == "integral"
    assert ents[0]["threshold"] == 0.0 and ents[1]["threshold"] == 0.2
    assert ents[0]["energy"].tolist() == [0.0, 1000.0] and ents[0]["sigma"].tolist() == [6e-20, 6e-20]
    assert ents[1]["sigma"][-1] == 9.998e-17 and not ents[1]["reverse"]


def test_lxcat_reader_reversible_and_subtype(tmp_path):
    text = ("PROCESS: E + B -> E + B(v1), Vibrational\nPARAM.:  E = 0.3 eV\n"
            "COMMENT: [e + B(X,v=0) <-> e + B(X,v=1), Vibrational, integral]\n-----\n"
            "0.3 0\n1.0 1e-20\n-----\n")
    p = tmp_path / "x.txt"
    p.write_text(text)
    (ent,) = ekmc.read_lxcat(str(p))
    assert ent["reverse"] and ent["type"] == "Vibrational" and ent["subtype"] == "integral"
    assert ent["threshold"] == 0.3
