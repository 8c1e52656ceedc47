"""CPU-side checks of the C ABI: the library loads, exports every entry point
include/imexlbm.h declares, and fails cleanly where there is no GPU.  No
compute call is made here."""
import ctypes
import os
import re
import subprocess

import pytest

from tests.conftest import ROOT, gpu_available


def _declared():
    txt = open(os.path.join(ROOT, "include", "imexlbm.h")).read()
    decl = r"^(?:int|void|const char)\s*\*?\s*(imex_[a-z_0-9]+)\s*\("
    return sorted(set(re.findall(decl, txt, flags=re.M)))


def test_header_and_binding_agree():
    from imexlbm_amd import capi
    assert sorted(capi.EXPORTS) == _declared()


def test_library_exports_every_declared_symbol():
    from imexlbm_amd import capi
    lib = ctypes.CDLL(capi.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_library_is_gfx950_code_object():
    from imexlbm_amd import capi
    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", capi.LIB_PATH], capture_output=True, text=True)
    if out.returncode != 0:
        pytest.skip("roc-obj-ls unavailable")
    assert "gfx950" in out.stdout


def test_strerror_and_defaults():
    from imexlbm_amd import capi
    L = capi.load()
    assert L.imex_strerror(0) == b"ok"
    assert L.imex_strerror(-2) == b"out of memory"
    c = capi.default_config()
    assert c.precision == capi.F64 and c.mode == capi.WRAP and c.ntiles == 8
    assert c.halo_depth == 1 and c.dir_pad == -1 and c.nontemporal == -1 and c.xcd_groups == 0
    assert c.xcd_run == -1 and c.io16 == -1
    assert c.two_step == -1 and c.two_step_tile == -1 and c.two_step_zseg == 0 and c.two_step_xcds == 0
    assert c.two_step_prefetch == -1 and c.halo_direct == -1 and c.two_step_io == -1
    assert c.whole_x == 0


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_create_without_gpu_fails_cleanly():
    from imexlbm_amd import capi
    with pytest.raises(capi.ImexError) as e:
        capi.Context({"nx": 8, "ny": 8, "nz": 8, "tau0": 0.01, "rho0": 1.0, "u0": 0.05})
    assert e.value.status in (-4, -3)


def test_argument_validation_runs_before_device_use():
    from imexlbm_amd import capi
    dom = {"nx": 0, "ny": 8, "nz": 8, "tau0": 0.01, "rho0": 1.0, "u0": 0.05}
    with pytest.raises(capi.ImexError) as e:
        capi.Context(dom)
    assert e.value.status == -1
    dom = {"nx": 8, "ny": 8, "nz": 8, "tau0": 0.01, "rho0": 1.0, "u0": 0.05}
    with pytest.raises(capi.ImexError):
        capi.Context(dom, halo_depth=4, mode=capi.TILED)
    with pytest.raises(capi.ImexError):
        capi.Context([dom, dom], mode=capi.TILED)


def test_driver_aborts_without_deck(tmp_path):
    exe = os.path.join(ROOT, "imexlbm_amd", "bin", "imexlbm")
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode != 0
    assert "cannot open input.in" in r.stderr


def _struct_fields(name):
    txt = open(os.path.join(ROOT, "include", "imexlbm.h")).read()
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), txt, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        typ, names = re.match(r"(long long|int|double)\s+(.*)", decl).groups()
        for n in names.split(","):
            fields.append((n.strip(), typ))
    return fields


@pytest.mark.parametrize("cname,pyname", [("imex_config", "Config"), ("imex_info", "Info"),
                                          ("imex_domain", "Domain")])
def test_ctypes_structs_match_header(cname, pyname):
    import ctypes
    from imexlbm_amd import capi
    want = _struct_fields(cname)
    if cname == "imex_domain":
        want = [(n, t) for n, t in want]
    got = getattr(capi, pyname)._fields_
    cmap = {"int": ctypes.c_int, "long long": ctypes.c_longlong, "double": ctypes.c_double}
    assert [n for n, _ in got] == [n for n, _ in want]
    assert all(t == cmap[w] for (_, t), (_, w) in zip(got, want))
