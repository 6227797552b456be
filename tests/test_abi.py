"""The C-ABI library: it loads, exports every symbol include/mhs_amd.h
declares, refuses bad descriptors before any GPU call, and its per-material
coefficients equal the oracle's evaluation of the exact rational table."""
import ctypes as C
import re

import numpy as np
import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def lib():
    from mhs_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libmhs_amd.so not built (run __graft_entry__.build())")
    return _lib.load()


def test_exports_every_declared_symbol(lib):
    header = (ROOT / "include" / "mhs_amd.h").read_text()
    declared = set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(mhs_\w+)\(", header, re.M))
    assert declared >= {"mhs_assemble", "mhs_gather", "mhs_strain", "mhs_histogram", "mhs_fill_material"}
    for name in declared:
        assert hasattr(lib, name), name


def test_sizes(lib):
    from mhs_amd import geometry
    assert lib.mhs_src_stride() == geometry.SRC_STRIDE
    assert lib.mhs_ncoef() == 2167
    from mhs_amd import _lib
    assert lib.mhs_abi_version() == _lib.ABI_VERSION == 6
    # the launch descriptor mirror has the header's layout: 31 fields, the
    # flattened Gauss rule pointer right after the 1-D tables
    names = [f[0] for f in _lib.Launch._fields_]
    assert names.index("gq") == names.index("gl_w") + 1
    assert geometry.GQ_TOTAL == sum(n * n for n in range(1, 193))
    assert names[-4:] == ["fold_work", "fold_work_items", "rule", "near_tiles"]
    assert C.sizeof(_lib.Rule) == 8 + 8 * _lib.RULE_BANDS + 4 * _lib.RULE_BANDS + 8
    assert _lib.Rule.fold_near_ratio.offset == 8 + 12 * _lib.RULE_BANDS
    from mhs_amd import runtime
    for p in range(3):
        assert 0 < lib.mhs_fold_item_bytes(p) <= runtime.FOLD_ITEM_BYTES


def test_bad_descriptors_refused_on_host(lib):
    from mhs_amd import _lib
    assert lib.mhs_assemble(None, None) != 0
    assert b"null launch" in lib.mhs_last_error()
    L = _lib.Launch(abi_version=99)
    assert lib.mhs_assemble(C.byref(L), None) != 0
    assert b"ABI" in lib.mhs_last_error()
    L = _lib.Launch(abi_version=_lib.ABI_VERSION, order=5)
    assert lib.mhs_assemble(C.byref(L), None) != 0
    assert b"order" in lib.mhs_last_error()
    assert lib.mhs_fill_material(0.0, 1.0, None) != 0


def test_material_coefficients_match_oracle(lib):
    """The library's coefficient vector against the oracle's own evaluation
    of the records (image.py:95-104, 276-297)."""
    from oracle import image as I
    mu, lam = 30.0, 45.0
    c = np.zeros(lib.mhs_ncoef())
    assert lib.mhs_fill_material(mu, lam, c.ctypes.data) == 0
    nu = lam / (2 * (lam + mu))
    scale = 2 * (lam + mu) / (np.pi * mu * (lam + 2 * mu))
    t = I.table()
    want = []
    for key in ("r_ddg", "r_dg"):
        for rec in t[key]:
            by_pz = {}
            for pz, pnu, num, den in rec[-1]:
                by_pz[pz] = by_pz.get(pz, 0.0) + (num / den) * nu ** pnu
            want += [v * scale for _, v in sorted(by_pz.items())]
    for key in ("q_ddg", "q_dg"):
        for rec in t[key]:
            want.append(sum((num / den) * nu ** pnu for _, pnu, num, den in rec[-1]) * scale)
    assert np.allclose(c, np.array(want), rtol=1e-14, atol=0)


def test_rule_struct_fold_switch():
    """A rule's optional third entry is the fold points' own switch; two
    entries leave it 0 (the library then takes near_ratio)."""
    from mhs_amd import runtime
    two = runtime.rule_struct((3.0, ((5.0, 4), (1e300, 1))))
    three = runtime.rule_struct((3.0, ((5.0, 4), (1e300, 1)), 2.0))
    assert two.near_ratio == three.near_ratio == 3.0
    assert two.fold_near_ratio == 0.0 and three.fold_near_ratio == 2.0
    assert list(three.order)[:3] == [4, 1, 1] and three.band_hi[0] == 5.0
    assert runtime._rule_name((3.0, ((5.0, 4), (1e300, 1)), 2.0)).endswith("fold points near 2 L")
    assert "fold" not in runtime._rule_name((3.0, ((5.0, 4), (1e300, 1))))
