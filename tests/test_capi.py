"""CPU-only checks of the C ABI: the library loads, exports every symbol the
header declares, and validates arguments before touching a device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "motion_estimation_amd.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:[\w\*]+\s+)+\*?(me_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "me_rigid_transform" in names and "me_last_error" in names
    assert len(names) >= 5


def test_library_exports_every_declared_symbol(me):
    lib = me._lib.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_matches_header(me):
    assert set(me._lib.declared_symbols()) == set(_declared())


def test_abi_version(me):
    assert me._lib.load().me_version() == 1


def test_argument_validation_without_device(me):
    lib = me._lib.load()
    n = me._lib.int3((16, 16, 16))
    st = me._lib.ll3((256, 16, 1))
    # null input -> error code and message, no device work
    rc = lib.me_rigid_transform(None, 0, 1, None, 0, None, 0, 1, n, st, None, None, None, None, 6, None, 0, 0,
                                None, None)
    assert rc != 0
    assert b"null" in lib.me_last_error()
    rc = lib.me_rigid_deriv(C.c_void_p(16), 0, C.c_void_p(16), 0, None, 0, 1, n, st, None, None, None,
                            None, 6, None, 7, None)
    assert rc != 0 and b"p_idx" in lib.me_last_error()


def test_grid_builder_matches_oracle(me):
    from oracle import motion_oracle as O

    for shape, ds in [((16, 12, 8), None), ((64, 64, 48), (16, 16, 48))]:
        kg, rkg = me.compute_transform_grids_voxel(shape, (1, 1, 1), ds)
        okg, orkg = O.compute_transform_grids_voxel(shape, (1, 1, 1), ds)
        for k in "xyz":
            assert np.array_equal(kg[k], okg[k])
            for m in ("tan", "sin"):
                assert np.array_equal(rkg[m][k], orkg[m][k])
        # r vectors recovered from reference-style plain dicts agree
        shp, rv = me.utils.grid_vectors(dict(okg), {m: dict(orkg[m]) for m in orkg})
        assert shp == tuple(kg.shape)
        for a, b in zip(rv, kg.r):
            assert np.allclose(a, b, atol=1e-12)


def test_encoder_tuning_table_without_device(me):
    """The forward-encoding variant table is host state: set, get, refuse an
    unknown variant, clear."""
    lib = me._lib.load()
    n = me._lib.int3((64, 64, 48))
    assert lib.me_encode_tuning_get(n, 9, 256) == 0
    assert lib.me_encode_tuning_set(n, 9, 256, 2) == 0
    assert lib.me_encode_tuning_get(n, 9, 256) == 2
    assert lib.me_encode_tuning_get(n, 10, 256) == 0
    assert lib.me_encode_tuning_set(n, 9, 256, 4) != 0 and b"variant" in lib.me_last_error()
    lib.me_encode_tuning_clear()
    assert lib.me_encode_tuning_get(n, 9, 256) == 0
    # the partials workspace covers every variant: 8-coil groups need more at 9 coils
    assert lib.me_encode_partials_size(48, 9, 256, 64) >= 48 * 2
