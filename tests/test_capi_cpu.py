"""CPU checks of the C-ABI library: it loads, it exports every entry point
that include/lbm_capi.h declares, the Params struct matches the C layout,
and argument validation rejects bad inputs before any device call."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbm_capi.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int32_t|const char \*|lbm_handle)\s*(lbm_\w+)\s*\(",
                                 text, flags=re.M)))


def test_header_declares_the_binding_list(hip):
    assert declared_functions() == sorted(hip.EXPORTS)


def test_library_exports_every_declared_symbol(hip):
    lib = hip.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_params_struct_layout(hip):
    # dim,nx,ny,nz (16) + 6 doubles (48) + 2 int32 (8) + 2 doubles (16)
    # + 2 int32 (8) + int64 (8) = 104 bytes, no padding
    assert ctypes.sizeof(hip.Params) == 104
    assert hip.Params.itc_max.offset == 96
    assert hip.Params.epsU.offset == 72


@pytest.mark.parametrize("kw", [dict(dim=4), dict(n=4), dict(Ra=-1.0),
                                dict(Ma=1e-12), dict(Thot=0.0, Tcold=0.0),
                                dict(precision=7), dict(grid_kind=9),
                                dict(check_every=0), dict(device=-1)])
def test_create_rejects_bad_parameters(hip, kw):
    args = dict(dim=2, n=17, Ra=1e4)
    args.update(kw)
    with pytest.raises(hip.LbmError) as e:
        hip.Solver(**args)
    assert e.value.code == hip.LBM_ERR_ARG


def test_status_strings(hip):
    lib = hip.load()
    for code in range(7):
        assert lib.lbm_status_string(code)


def test_batch_rejects_bad_parameters(hip):
    ok = dict(n=17, Ra=1e4)
    with pytest.raises(hip.LbmError):
        hip.Batch([])                                    # no cavity
    with pytest.raises(hip.LbmError):
        hip.Batch([ok, dict(n=3, Ra=1e4)])               # cavity 1 invalid
    with pytest.raises(TypeError):
        hip.Batch([dict(n=17, Ra=1e4, colour="red")])
