"""The C-ABI library loads on the CPU and exports every declared symbol."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, 'include', 'gecco_hip.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(?:void|int)\s+(\w+)\s*\(', txt)))


def test_header_declares_reference_entry_points():
    names = declared_functions()
    assert 'contr_blk1blk2_blocked_mm_' in names
    assert 'contr_blk1blk2_wmaps_c_' in names
    assert 'gecco_hip_plan_execute' in names


def test_library_exports_every_declared_symbol():
    import gecco_hip
    lib = ctypes.CDLL(gecco_hip.LIBPATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_error_path_without_device_work():
    """A plan with an invalid map_info is rejected on the host with a message
    (replaces quit() in contract/contr_blk1blk2_blocked_mm.f:218-233)."""
    import gecco_hip
    import pytest
    from models import orbitals, ladder
    c, *_ = ladder(orbitals(2, 3, 1))
    blk = dict(c.blocks(False)[0])
    blk['map_info_1a'] = [2, 1, 1, 0]       # two sources into one component
    p = gecco_hip.Plan()
    with pytest.raises(gecco_hip.GeccoHipError, match='map_info'):
        p.add(blk)
    blk = dict(c.blocks(False)[0])
    blk['map_ex1cntc'] = [10 ** 6] * len(blk['map_ex1cntc'])   # out of the block
    with pytest.raises(gecco_hip.GeccoHipError, match='addresses element'):
        p.add(blk)
