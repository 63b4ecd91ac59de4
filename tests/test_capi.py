"""The C-ABI library: it loads and exports every function include/*.h
declare; host-only calls work without a GPU."""
import ctypes as ct
import os
import re

import pytest

from common import REPO

LIB = os.path.join(REPO, "funwave-tvd_amd", "libfwgpu.so")


def declared(header="fwgpu.h"):
    src = open(os.path.join(REPO, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(fw_\w+)\s*\(", src, re.M)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfwgpu.so not built")
def test_library_exports_every_declared_symbol():
    lib = ct.CDLL(LIB)
    names = declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    import fwgpu
    assert set(fwgpu.EXPORTS) == set(names)
    cnn = declared("fwcnn.h")
    assert len(cnn) >= 8
    missing = [n for n in cnn if not hasattr(lib, n)]
    assert not missing, missing
    import fwcnn
    assert set(fwcnn.EXPORTS) == set(cnn)
    assert sorted(os.listdir(os.path.join(REPO, "include"))) == ["fwcnn.h", "fwgpu.h"]


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfwgpu.so not built")
def test_field_ids_cover_reference_arrays():
    import fwgpu
    lib = fwgpu.load_library()
    ids = [lib.fw_field_id(n.encode()) for n in fwgpu.ALL_FIELDS]
    assert ids == list(range(len(fwgpu.ALL_FIELDS)))
    assert lib.fw_field_id(b"NoSuchArray") < 0
    import oracle
    assert oracle.ALL_FIELDS == fwgpu.ALL_FIELDS


@pytest.mark.skipif(not os.path.exists(LIB), reason="libfwgpu.so not built")
def test_create_rejects_bad_configuration():
    import fwgpu
    lib = fwgpu.load_library()
    c = fwgpu.Config()
    c.real_bytes = 3
    ctx = ct.c_void_p()
    assert lib.fw_create(ct.byref(c), ct.byref(ctx)) == fwgpu.FW_E_INVALID
    c.real_bytes = 4
    c.Mloc, c.Nloc, c.Ibeg, c.Jbeg, c.Iend, c.Jend = 10, 10, 4, 4, 9, 7  # wrong Iend
    assert lib.fw_create(ct.byref(c), ct.byref(ctx)) == fwgpu.FW_E_INVALID
    assert not ctx.value
