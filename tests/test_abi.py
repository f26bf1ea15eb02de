"""The C-ABI library (libjitterhip.so) loads, exports every entry point declared in
include/jitterhip.h, and agrees with the numpy struct layouts.  No compute calls: CPU only."""
import ctypes
import os
import re

import numpy as np

from jitterphysics2_amd import _abi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "jitterhip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(jh_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_abi():
    names = declared_symbols()
    for required in ("jh_world_create", "jh_world_step", "jh_world_destroy", "jh_world_set_bodies",
                     "jh_world_set_shapes", "jh_world_compute_pairs", "jh_world_get_arbiters",
                     "jh_world_get_solve_order", "jh_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_abi.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"


def test_struct_sizes_match_numpy_layouts():
    lib = _abi.load_library()
    for k, size in _abi.STRUCT_SIZES.items():
        assert lib.jh_struct_size(k) == size
    assert lib.jh_struct_size(99) == -1
    assert lib.jh_version() == 1


def test_config_defaults_match_reference_world_defaults():  # World.cs:346-412
    lib = _abi.load_library()
    c = np.zeros(1, _abi.CONFIG_DTYPE)
    assert lib.jh_config_defaults(_abi.ptr(c)) == 0
    ref = _abi.default_config()
    for f in _abi.CONFIG_DTYPE.names:
        assert np.array_equal(c[f], ref[f]), f


def test_null_arguments_are_rejected_without_a_device():
    lib = _abi.load_library()
    assert lib.jh_world_step(None, ctypes.c_float(0.01)) == -6
    assert lib.jh_world_destroy(None) == -1
    assert lib.jh_config_defaults(None) == -1
    assert b"NULL" in lib.jh_last_error()


def test_scene_builders_are_seeded_and_consistent():
    b1, s1, _, cfg, dt = scenes.heap(300, seed=9)
    b2, s2, _, _, _ = scenes.heap(300, seed=9)
    assert np.array_equal(b1, b2) and np.array_equal(s1, s2)
    assert np.all(np.diff(s1["shape_id"].astype(np.int64)) > 0)
    assert int(cfg["substep_count"][0]) == 4 and int(cfg["solver_iterations"][0]) == 10
    assert abs(dt - 1 / 60) < 1e-12
    q = b1["orientation"].astype(np.float64)
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-6)
    # unit box: mass 1, inverse inertia 6 (BoxShape.CalculateMassInertia)
    m, inertia = scenes.box_mass((1, 1, 1))
    assert m == 1.0 and abs(scenes._sym_inverse(*inertia)[0] - 6.0) < 1e-6


def test_wait_control_rejects_null_world_without_a_device():
    lib = _abi.load_library()
    assert lib.jh_world_set_spin_limit(None, 1) == -6
    assert lib.jh_world_cancel(None) == -6
    assert _abi.ERRORS[_abi.JH_ERR_TIMEOUT].startswith("device wait timed out")
