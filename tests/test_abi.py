"""C ABI checks that need no GPU: libyamc_ewald.so loads, exports every
function include/yamc_ewald.h declares, the ctypes layouts match the C
structs (compiled with gcc against the header), and creating a context
without a HIP device fails with YAMC_ERROR_DEVICE instead of computing."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "yamc_ewald.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(yamc_\w+)\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    from yamc_amd import _lib
    lib = _lib.load()
    names = declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTS)


def test_struct_layouts_match_header():
    from yamc_amd import _lib
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "yamc_ewald.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(yamc_system_desc), offsetof(yamc_system_desc, box_length),
         offsetof(yamc_system_desc, charge), offsetof(yamc_system_desc, tail_counts), sizeof(yamc_move_rec));
  printf("%zu %zu %zu\n", offsetof(yamc_move_rec, du_real), offsetof(yamc_move_rec, f_old),
         offsetof(yamc_move_rec, angle));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "l.c")
        open(c, "w").write(src)
        exe = os.path.join(td, "l")
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    v = [int(x) for x in out]
    D, M = _lib.SystemDesc, _lib.MoveRec
    assert v[0] == C.sizeof(D) and v[1] == D.box_length.offset and v[2] == D.charge.offset
    assert v[3] == D.tail_counts.offset and v[4] == C.sizeof(M)
    assert v[5] == M.du_real.offset and v[6] == M.f_old.offset and v[7] == M.angle.offset


def test_no_device_is_an_error_not_a_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    from yamc_amd import YamcError, config
    from yamc_amd.ewald import EwaldMC
    with pytest.raises(YamcError) as e:
        EwaldMC(config("small"))
    assert e.value.code == 6


def test_bad_descriptors_are_rejected():
    """Error conventions (error.f90:47-51 codes) for malformed input, checked
    before any device work."""
    import numpy as np
    from yamc_amd import YamcError, config, _lib
    from yamc_amd.ewald import _desc
    s = config("small")
    lib = _lib.load()
    desc, keep = _desc(s, 0)
    desc.kmax = 99
    ctx = C.c_void_p()
    assert lib.yamc_ctx_create(C.byref(desc), C.byref(ctx)) == 2  # ERROR_LIMITS
    desc, keep = _desc(s, 7)
    assert lib.yamc_ctx_create(C.byref(desc), C.byref(ctx)) == 3  # ERROR_PARAMETER
    assert lib.yamc_ctx_create(None, C.byref(ctx)) == 3
