"""The drop-in Fortran modules of fortran/shim/ (no GPU needed).

Each module there replaces the reference module of the same name: it must
define every procedure and derived type the reference's module defines, with
the same dummy-argument names, and the reference's MC.f90, runmc.f90 and the
golden-vector driver must compile and link against it and libyamc_ewald.so
(oracle/build_shim.sh).  These checks read the reference's sources and are
skipped where they are absent.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
FC = "/opt/rocm/lib/llvm/bin/flang"
SHIM = ["FourierGrid", "SumSinCosKR", "RhoSquared", "EwaldSumKSpace", "ForceKSpace", "RealSumLocal",
        "EwaldSumRealSpace", "EwaldSumTails"]

PROC = re.compile(r"^\s*(?:pure\s+|elemental\s+|recursive\s+)*(subroutine|function)\s+(\w+)\s*\(([^)]*)\)",
                  re.IGNORECASE | re.MULTILINE)
TYPE = re.compile(r"^\s*type\s*(?:,[^:]*)?::\s*(\w+)", re.IGNORECASE | re.MULTILINE)


def strip_comments(src):
    return "\n".join(line.split("!")[0] for line in src.splitlines())


def interface(path):
    src = strip_comments(open(path, errors="replace").read())
    procs = {m.group(2).lower(): [a.strip().lower() for a in m.group(3).split(",") if a.strip()]
             for m in PROC.finditer(src)}
    types = {m.group(1).lower() for m in TYPE.finditer(src)}
    return procs, types


needs_ref = pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")


@needs_ref
@pytest.mark.parametrize("mod", SHIM)
def test_shim_module_has_the_reference_interface(mod):
    rp, rt = interface(os.path.join(REF, f"{mod}.f90"))
    sp, st = interface(os.path.join(ROOT, "fortran", "shim", f"{mod}.f90"))
    assert rp, mod
    for name, args in rp.items():
        assert name in sp, f"{mod}: {name} missing"
        assert sp[name] == args, f"{mod}: {name} dummy arguments {sp[name]} != {args}"
    assert rt <= st, f"{mod}: types {rt - st} missing"


def type_fields(path, tname):
    src = strip_comments(open(path, errors="replace").read())
    m = re.search(r"^\s*type\s*(?:,[^:]*)?::\s*" + tname + r"\b(.*?)^\s*end\s*type", src,
                  re.IGNORECASE | re.MULTILINE | re.DOTALL)
    fields = []
    for line in m.group(1).splitlines():
        if "::" in line:
            decl, names = line.split("::", 1)
            for n in re.sub(r"\([^)]*\)", "", names).split(","):
                n = n.split("=")[0].strip().lower()
                if n:
                    fields.append((n, re.sub(r"\s+", "", decl.lower())))
    return fields


@needs_ref
@pytest.mark.parametrize("mod", SHIM)
def test_shim_types_keep_the_reference_fields(mod):
    """Every field of every derived type, in the reference's order and with
    its declaration (kind, rank, pointer/allocatable); a shim type may add
    fields after them."""
    _, rt = interface(os.path.join(REF, f"{mod}.f90"))
    for t in rt:
        ref = type_fields(os.path.join(REF, f"{mod}.f90"), t)
        mine = type_fields(os.path.join(ROOT, "fortran", "shim", f"{mod}.f90"), t)
        norm = lambda d: d.replace("type(", "type(").replace("integer,", "integer,")  # noqa: E731
        assert [n for n, _ in mine[:len(ref)]] == [n for n, _ in ref], (mod, t)
        for (n, d1), (_, d2) in zip(mine, ref):
            assert norm(d1) == norm(d2), (mod, t, n, d1, d2)


@needs_ref
@pytest.mark.skipif(not os.access(FC, os.X_OK), reason="flang not present")
def test_reference_mc_builds_against_the_shims(tmp_path):
    lib = os.path.join(ROOT, "yamc_amd", "libyamc_ewald.so")
    if not os.path.exists(lib):
        pytest.skip("libyamc_ewald.so not built")
    env = dict(os.environ, OUT=str(tmp_path))
    r = subprocess.run([os.path.join(ROOT, "oracle", "build_shim.sh")], env=env, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    exe = tmp_path / "shim_driver"
    assert exe.exists()
    nm = shutil.which("nm")
    if nm:
        syms = subprocess.run([nm, "-D", "--undefined-only", str(exe)], capture_output=True, text=True).stdout
        assert "yamc_op_rspace" in syms and "yamc_op_du_coulomb" in syms
