"""CPU-side checks of the C-ABI library and the host mirror (no GPU calls)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import dqmc_amd
from dqmc_amd import _capi
from dqmc_amd import model as PM
from oracle import dqmc_oracle as O
from tests.conftest import ROOT
from tests.conftest import load_fixture as ld

HEADER = os.path.join(ROOT, "include", "dqmc_hip.h")


def header_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(dqmc_\w+)\s*\(", txt, re.M)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_capi.LIB_PATH):
        from dqmc_amd import build
        build.build()
    return _capi.load()


def test_library_exports_every_header_symbol(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (dqmc_\w+)", out))
    declared = header_functions()
    assert declared, "no declarations parsed"
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    # the ctypes table covers exactly the declared surface
    assert sorted(_capi.SIGNATURES) == declared


def test_struct_layouts_match_header():
    # dqmc_params: 6 ints + 5 doubles; dqmc_operators: 6 pointers, an int (padded), 3 pointers
    assert ctypes.sizeof(_capi.Params) == 6 * 4 + 5 * 8
    assert ctypes.sizeof(_capi.Operators) == 10 * 8


def test_invalid_arguments_fail_loudly(lib):
    # argument validation happens before any device call
    p = _capi.Params(L=3, slices=10, safe_mult=10, checkerboard=1, edrun=0, all_checks=1,
                     delta_tau=0.1, lam=1.0, r=1.0, c=3.0, u=1.0)
    o = _capi.Operators()
    h = ctypes.c_void_p()
    rc = lib.dqmc_ctx_create(ctypes.byref(p), ctypes.byref(o), 0, ctypes.byref(h))
    assert rc == -1
    assert b"neighbors" in lib.dqmc_last_error() or b"even L" in lib.dqmc_last_error()
    p.L = 4
    p.slices = 15
    rc = lib.dqmc_ctx_create(ctypes.byref(p), ctypes.byref(o), 0, ctypes.byref(h))
    assert rc == -1 and b"multiple of safe_mult" in lib.dqmc_last_error()
    with pytest.raises(_capi.DQMCError):
        _capi.check(rc)


def test_missing_library_raises(tmp_path):
    saved = _capi._lib
    try:
        _capi._lib = None
        with pytest.raises(_capi.DQMCError):
            _capi.load(str(tmp_path / "nope.so"))
    finally:
        _capi._lib = saved


@pytest.mark.parametrize("bfield", [False, True])
def test_product_operators_match_oracle(bfield):
    """The setup operators the host passes to the library equal the oracle's (itself
    pinned by the reference fixtures)."""
    p = PM.Params(L=4, slices=10, hoppings=(1.0, 0.5, -0.5, -1.0), lam=0.5, r=2.0, bfield=bfield)
    po = O.Params(4, 10, hoppings=(1.0, 0.5, -0.5, -1.0), lam=0.5, r=2.0, bfield=bfield, chkr=True)
    ops, oo = PM.build_operators(p), O.Operators(po)
    n, N = p.n, p.N
    for v, ref in enumerate([oo.chkr_hop_half, oo.chkr_hop_half_inv, oo.chkr_hop, oo.chkr_hop_inv]):
        for g in range(2):
            D = np.zeros((n, n), complex)
            for b in range(4):
                for q in range(N // 4):
                    st = ops.plaq_sites[g, q]
                    D[np.ix_(b * N + st, b * N + st)] += ops.plaq[v, g, b, q]
            assert np.abs(D - ref[g].toarray()).max() < 1e-15
    assert np.abs(ops.expmu - oo.chkr_mu.diagonal()).max() == 0
    assert (ops.neighbors == O.neighbors(4)).all()
    pd = PM.Params(L=4, slices=10, hoppings=(1.0, 0.5, -0.5, -1.0), lam=0.5, r=2.0, bfield=bfield,
                   checkerboard=False)
    od = PM.build_operators(pd)
    assert np.abs(od.hop_exp - oo.hopping_matrix_exp).max() < 1e-14


def test_params_from_reference_xml(tmp_path):
    xml = tmp_path / "c1.in.xml"
    xml.write_text("""<?xml version="1.0" encoding="utf-8"?>
<SIMULATION><PARAMETERS>
 <PARAMETER name="L">10</PARAMETER><PARAMETER name="BETA">20</PARAMETER>
 <PARAMETER name="DELTA_TAU">0.1</PARAMETER><PARAMETER name="SAFE_MULT">10</PARAMETER>
 <PARAMETER name="WARMUP">100</PARAMETER><PARAMETER name="SWEEPS">100</PARAMETER>
 <PARAMETER name="CHECKERBOARD">TRUE</PARAMETER><PARAMETER name="HOPPINGS">1.0,0.5,0.5,1.0</PARAMETER>
 <PARAMETER name="MU">-0.5</PARAMETER><PARAMETER name="LAMBDA">1</PARAMETER><PARAMETER name="U">1.0</PARAMETER>
 <PARAMETER name="R">1.0</PARAMETER><PARAMETER name="C">3.0</PARAMETER>
 <PARAMETER name="GLOBAL_UPDATES">FALSE</PARAMETER><PARAMETER name="BFIELD">FALSE</PARAMETER>
</PARAMETERS></SIMULATION>""")
    p = dqmc_amd.params_from_xml(str(xml))
    # BETA/DELTA_TAU evaluated exactly (parameters.jl:101-104): 20/0.1 = 200 slices
    assert p.slices == 200 and p.L == 10 and p.checkerboard and not p.global_updates
    assert p.thermalization == 50 and p.measurements == 50
    assert tuple(p.t[:, 1]) == (0.5, 1.0)


def test_uniform_stream_is_order_preserving():
    from dqmc_amd.dqmc import UniformStream
    s = UniformStream(7)
    a = s.take(10).copy()
    s.advance(4)
    b = s.take(10).copy()
    assert np.array_equal(a[4:], b[:6])
    assert np.array_equal(np.concatenate([a[:4], b]), O.host_uniforms(7, 14))


def test_fixture_manifest_hashes():
    import hashlib
    import json
    man = json.load(open(os.path.join(ROOT, "tests", "golden", "data", "MANIFEST.json")))
    for e in man:
        data = open(os.path.join(ROOT, "tests", "golden", "data", e["name"] + ".npy"), "rb").read()
        assert hashlib.sha256(data).hexdigest() == e["sha256"], e["name"]
    assert len(man) >= 50
    assert ld("greens").shape == (64, 64)


@pytest.mark.parametrize("L,bfield,force", [(5, False, False), (3, True, False), (4, False, True)])
def test_generic_checkerboard_operators_match_oracle(L, bfield, force):
    """The host's generic-checkerboard operators (bond groups and 2x2 blocks handed to the
    library) assemble into the oracle's chkr_hop_half / _inv / chkr_hop / _inv matrices."""
    p = PM.Params(L=L, slices=10, hoppings=(1.0, 0.5, 0.5, 1.0), checkerboard=True, bfield=bfield,
                  generic_chkr=force)
    assert p.generic
    ops = PM.build_operators(p)
    op = O.Operators(O.Params(L, 10, hoppings=(1.0, 0.5, 0.5, 1.0), chkr=True, bfield=bfield, generic=force))
    G = len(ops.group_offsets) - 1
    assert G == len(op.groups)
    N, n = p.N, p.n
    for v, ref in enumerate((op.chkr_hop_half, op.chkr_hop_half_inv, op.chkr_hop, op.chkr_hop_inv)):
        for g in range(G):
            D = np.eye(n, dtype=np.complex128)
            for q in range(ops.group_offsets[g], ops.group_offsets[g + 1]):
                a, b = ops.bond_sites[q]
                for blk in range(4):
                    idx = [blk * N + a, blk * N + b]
                    D[np.ix_(idx, idx)] = ops.bond_blocks[v, q, blk]
            assert np.abs(D - ref[g].toarray()).max() < 1e-14, (v, g)
