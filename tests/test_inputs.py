"""Input checks of the library that need no GPU: search options the reference
rejects, binary databases that are truncated or carry impossible lengths, and
profile sets with empty profiles.  Each is refused with an error message, not
read out of bounds or computed silently with other settings."""
import ctypes
import struct

import numpy as np
import pytest

import cother_amd
from cother_amd import SearchOpts, lib

SIG = b"COTHER profile bin database v2.5"
NSECT = 179  # sections of the v2.2 address table (PMProfileModel.h:42-48)


def _opts(**kw):
    o = SearchOpts()
    o.eval, o.ssemodel, o.do_map, o.minpp = 10.0, 2, 1, 0.28
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _check(o):
    L = lib()
    L.cth_search_opts_check.argtypes = [ctypes.POINTER(SearchOpts), ctypes.c_char_p]
    rc = L.cth_search_opts_check(ctypes.byref(o), b"test")
    return rc, L.cth_last_error().decode()


@pytest.mark.parametrize("model", [0, 1, 2])
def test_ssemodel_range_accepted(model):
    """SSEMODEL takes 0, 1 and 2 (libpro/srcpro/MOptions.cpp:123)."""
    assert _check(_opts(ssemodel=model))[0] == 0


@pytest.mark.parametrize("model", [-1, 3, 7])
def test_ssemodel_out_of_range_refused(model):
    rc, msg = _check(_opts(ssemodel=model))
    assert rc != 0 and "ssemodel" in msg


def test_global_database_size_options():
    assert _check(_opts(dbsize=1000, ndb=10))[0] == 0
    rc, msg = _check(_opts(dbsize=1000, ndb=0))
    assert rc != 0 and "together" in msg
    rc, msg = _check(_opts(dbsize=5, ndb=10))
    assert rc != 0 and "ndb" in msg
    rc, _ = _check(_opts(contraction=3))
    assert rc != 0


def _db_file(path, nprof, dbsize, lens, body_bytes):
    """A binary database whose data sections all start at one address right
    after the address table (so every section holds the same bytes); the
    description sections (the last two) hold one empty description."""
    base = len(SIG) + 24 + 8 * NSECT
    desc_end = base + body_bytes          # 8-byte end address of description 0
    desc = desc_end + 8                   # its text: one NUL
    addr = [base] * (NSECT - 2) + [desc_end, desc]
    hdr = SIG + struct.pack("<QQQ", nprof, dbsize, 0) + struct.pack(f"<{NSECT}Q", *addr)
    body = bytearray(body_bytes)
    body[:4 * len(lens)] = struct.pack(f"<{len(lens)}i", *lens)
    tail = struct.pack("<Q", desc + 1) + b"\0"
    path.write_bytes(hdr + bytes(body) + tail)
    return str(path)


def test_truncated_sections_refused(tmp_path):
    # one profile of 50 positions needs 4 * 51 transition bytes per section; 64 bytes are there
    p = _db_file(tmp_path / "t.bin", 1, 50, [50], 64)
    with pytest.raises(RuntimeError, match="truncated"):
        cother_amd.read_binary_db(p)


def test_header_counts_beyond_the_file_refused(tmp_path):
    p = _db_file(tmp_path / "h.bin", 1 << 40, 1 << 41, [50], 64)
    with pytest.raises(RuntimeError, match="header counts"):
        cother_amd.read_binary_db(p)


@pytest.mark.parametrize("length", [-5, 0])
def test_nonpositive_length_refused(tmp_path, length):
    # every section fits: one profile, one position
    p = _db_file(tmp_path / "n.bin", 1, 1, [length], 256)
    with pytest.raises(RuntimeError, match="profile length"):
        cother_amd.read_binary_db(p)


def test_well_formed_minimal_file_reads(tmp_path):
    """The same construction with a valid length reads (the checks above are
    not refusing everything)."""
    p = _db_file(tmp_path / "ok.bin", 1, 1, [1], 4096)
    ps = cother_amd.read_binary_db(p)
    assert ps.nprof == 1 and ps.npos == 1


@pytest.mark.parametrize("lens", [[10, 0, 5], [-3]])
def test_pset_alloc_refuses_empty_profiles(lens):
    a = np.asarray(lens, np.int32)
    p = lib().cth_pset_alloc(len(a), a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    assert not p
    assert "length" in lib().cth_last_error().decode()


def test_lengths_over_maxcolumns_refused():
    """profiles of more than MAXCOLUMNS = 50000 positions are refused, as the
    reference's readers refuse them (liblib/mylimits.h:50,
    libmycu/cupro/IOProfileModelBin.cu:166)"""
    import ctypes
    lens = np.array([50001], np.int32)
    p = cother_amd.lib().cth_pset_alloc(1, lens.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    assert not p
    assert b"50000" in cother_amd.lib().cth_last_error()
    lens = np.array([50000], np.int32)
    p = cother_amd.lib().cth_pset_alloc(1, lens.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    assert p
    cother_amd.lib().cth_pset_free(p)
