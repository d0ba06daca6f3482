"""The C-ABI library loads and exports every symbol include/vxp.h declares,
the ctypes mirror matches the C layout, and the default configuration matches
the reference's defaults.  No compute calls: these run without a GPU."""
import ctypes as C
import os
import re
import subprocess

import pytest

import vxp
from vxp import _abi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "vxp.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vxp_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_abi.PRODUCT_SO):
        pytest.fail("libvxp.so is not built: run __graft_entry__.build()")
    return C.CDLL(_abi.PRODUCT_SO)


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_oracle_exports_the_same_api(oracle):
    # timing, launch tuning, recorded steps, the serving loop and the
    # allocation-free entry point exist only in the product
    names = [n for n in declared_functions()
             if not n.startswith(("vxp_profile", "vxp_server", "vxp_tuning", "vxp_noalloc",
                                  "vxp_graph"))
             and n not in ("vxp_config_default", "vxp_collider_mass_properties_noalloc")]
    missing = [n for n in names if not hasattr(oracle.dll, "vxo_" + n[4:])]
    assert not missing, missing


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "vxp.h"\n'
                   'int main(){printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(vxp_config),'
                   'sizeof(vxp_stats), sizeof(vxp_contact), sizeof(vxp_manifold), sizeof(vxp_handle),'
                   'sizeof(vxp_body_state), offsetof(vxp_contact, feature_id));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    got = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()))
    want = [C.sizeof(_abi.Config), C.sizeof(_abi.Stats), C.sizeof(_abi.Contact),
            C.sizeof(_abi.Manifold), C.sizeof(_abi.Handle), C.sizeof(_abi.BodyState),
            _abi.Contact.feature_id.offset]
    assert got == want


def test_config_default_matches_reference_defaults(lib):
    fn = lib.vxp_config_default
    fn.argtypes = [C.POINTER(_abi.Config)]
    c = _abi.Config()
    assert fn(C.byref(c)) == 0
    want = _abi.default_config()
    for name, _ in _abi.Config._fields_:
        a, b = getattr(c, name), getattr(want, name)
        if hasattr(a, "__len__"):
            a, b = list(a), list(b)
        assert a == b, name


def test_invalid_arguments_rejected_without_device(lib):
    fn = lib.vxp_system_create
    fn.argtypes = [C.POINTER(_abi.Config), C.c_uint32, C.POINTER(C.c_void_p)]
    c = _abi.default_config()
    p = C.c_void_p()
    assert fn(C.byref(c), 0, C.byref(p)) == -5          # zero worlds
    c.pair_order = 7
    assert fn(C.byref(c), 1, C.byref(p)) == -5          # unknown pair order
    c = _abi.default_config(brick_pair_cache_size=4096)
    assert fn(C.byref(c), 1, C.byref(p)) == -5


def test_header_functions_cite_the_reference():
    text = open(HEADER).read()
    for name in ("vxp_system_create", "vxp_collider_create", "vxp_body_create", "vxp_system_update",
                 "vxp_manifolds_get_w", "vxp_pairs_get_w", "vxp_islands_get_w"):
        i = text.index(name)
        assert "src/" in text[max(0, i - 600):i], name


def test_set_config_accepts_only_gravity_and_solver_fields(oracle):
    import vxp as _v
    s = _v.VoxelPhysicsSystem(lib=oracle, max_bodies=8)
    s.setGravity((0.0, -3.0, 0.0))
    assert s.getGravity() == (0.0, -3.0, 0.0)
    s.setSolverConfig(substeps=2, velocity_iterations=5, hertz=15.0, enable_friction=0)
    with pytest.raises(_v.VxpError):
        s.setSolverConfig(max_bodies=16)          # a PhysicsConfig field
    with pytest.raises(_v.VxpError):
        s.setSolverConfig(substeps=0)
    assert s.cfg.max_bodies == 8 and s.cfg.substeps == 2
    s.close()
