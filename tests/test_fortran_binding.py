"""The Fortran host boundary: fortran/yamc_ewald.f90 binds every entry point
of include/yamc_ewald.h through iso_c_binding, and the Fortran host
fortran/mc_host drives the library like the YAMC run loop (GPU test)."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
HOST = os.path.join(ROOT, "fortran", "mc_host")


def header_functions():
    text = open(os.path.join(ROOT, "include", "yamc_ewald.h")).read()
    return set(re.findall(r"^(?:int|const char \*)\s*(yamc_\w+)\(", text, re.M))


def test_binding_covers_every_c_entry_point():
    text = open(os.path.join(ROOT, "fortran", "yamc_ewald.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(yamc_\w+)"\)', text))
    assert bound == header_functions()


def test_move_record_and_descriptor_fields_match_header():
    """The bind(C) types list the C structs' fields in the same order."""
    f90 = open(os.path.join(ROOT, "fortran", "yamc_ewald.f90")).read()
    h = open(os.path.join(ROOT, "include", "yamc_ewald.h")).read()

    def c_fields(struct):
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (struct, struct), h, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        names = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            decl = re.sub(r"^(const\s+)?\w+\s*\**\s*", "", decl)
            names += [re.sub(r"[\s*]|\[\d+\]", "", n) for n in decl.split(",")]
        return names

    def f_fields(tname):
        body = re.search(r"type, bind\(C\) :: %s\n(.*?)end type" % tname, f90, re.S).group(1)
        names = []
        for line in body.strip().split("\n"):
            names += [re.sub(r"\(\d+\)", "", n).strip() for n in line.split("::")[1].split(",")]
        return names

    assert f_fields("yamc_system_desc") == c_fields("yamc_system_desc")
    assert f_fields("yamc_move_rec") == c_fields("yamc_move_rec")


def test_fortran_host_is_built():
    if not os.path.exists("/opt/rocm/lib/llvm/bin/flang"):
        pytest.skip("no Fortran compiler in this image")
    assert os.path.exists(HOST), "fortran/mc_host missing: run __graft_entry__.build()"


def parse(out):
    res = {}
    for line in out.splitlines():
        tok = line.split()
        if tok and tok[0] in ("E_INIT", "E_RUN", "FORCE_SUM"):
            res[tok[0]] = np.array([float(v) for v in tok[1:]])
        elif tok and tok[0] == "ACCEPTED":
            res["ACCEPTED"] = int(tok[1])
    return res


@pytest.mark.gpu
def test_fortran_host_matches_oracle_and_python(tmp_path):
    """Fortran host -> iso_c_binding -> C ABI -> HIP: the initial energies
    and a 3-cycle run equal the oracle's (1e-10) and the Python host's."""
    import torch
    from oracle import pyoracle
    from yamc_amd import config
    from yamc_amd.ewald import ENERGY_NAMES, EwaldMC
    from yamc_amd.system import write_host_case
    assert torch.cuda.is_available()
    s = config("small", seed=5)
    case = str(tmp_path / "small.case")
    write_host_case(s, case, precision=0, seed=123, stream=2, ncycles=3)
    print("fortran host: small system, seed 123 stream 2, 3 cycles", flush=True)
    r = subprocess.run([HOST, case], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = parse(r.stdout)
    o = pyoracle.Oracle(s)
    eo, _ = o.mc_init()
    want = np.array([eo[k] for k in ENERGY_NAMES])
    assert np.max(np.abs(got["E_INIT"] - want) / np.maximum(np.abs(want), 1.0)) < 1e-10
    o.set_rng(123, 2)
    no, _ = o.run(3)
    assert got["ACCEPTED"] == no
    eo = o.energies()
    want = np.array([eo[k] for k in ENERGY_NAMES])
    assert np.max(np.abs(got["E_RUN"] - want) / np.maximum(np.abs(want), 1.0)) < 1e-10
    g = EwaldMC(s)
    g.mc_init()
    g.set_rng(123, 2)
    ng, _ = g.run(3)
    assert ng == got["ACCEPTED"]
    eg = g.energies()
    assert np.array_equal(np.array([eg[k] for k in ENERGY_NAMES]), got["E_RUN"])
