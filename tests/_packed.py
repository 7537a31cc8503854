"""The reference's packed profile-model buffers (test helper).

TdDataReader::ReadDataChunkHelper (libmycu/cupro/TdDataReader.cpp:703-980)
packs a chunk of database profiles into one contiguous host buffer: the 181
fields of TPM2DVectorFields (cupro/PM2DVectorFields.h:53-87), each an array
over profiles, transition vectors or positions, each starting at a multiple
of 32 bytes (ALIGN_UP to CUL2CLINESIZE, cucom/cucommon.h:18); bdbCpmbeg[k] /
bdbCpmend[k] bound field k and szCpm2dvfields[k] is its byte offset from the
start (entry 181: the total).  pack() builds such a buffer from a profile
set, pack_sections() from the sections of a db2bin file as the reference's
reader copies them.
"""
import ctypes

import numpy as np

NAA, NTRN, NCV, NDD = 20, 7, 19, 100
F_BKG, F_ENO, F_LEN, F_DIST, F_TRN = 0, 20, 21, 22, 23
F_TRG = F_TRN + NTRN
F_CV = F_TRG + NAA
F_CVPRIOR = F_CV + NCV
F_CVNORM2 = F_CVPRIOR + 1
F_DD = F_CVNORM2 + 1
F_SS = F_DD + NDD
F_HDPPRB = F_SS + 3
F_HDPIND = F_HDPPRB + 1
F_ADDRPRO = F_HDPIND + 1
F_ADDRCV = F_ADDRPRO + 1
F_ADDRSS = F_ADDRCV + 1
F_AA = F_ADDRSS + 1
F_SSSTATE = F_AA + 1
NFIELDS = F_SSSTATE + 1
ALIGN = 32


class Packed:
    """One contiguous buffer with the 181 field arrays; beg/end/sz as ctypes arrays."""

    def __init__(self, arrays):
        assert len(arrays) == NFIELDS
        offs, total = [], 0
        for a in arrays:
            offs.append(total)
            total += (a.nbytes + ALIGN - 1) // ALIGN * ALIGN
        self.buf = np.zeros(total + ALIGN, np.uint8)
        base = self.buf.ctypes.data
        pad = (-base) % ALIGN  # the field offsets are relative to an aligned start
        self.start = base + pad
        for a, o in zip(arrays, offs):
            raw = np.ascontiguousarray(a).view(np.uint8).ravel()
            self.buf[pad + o:pad + o + raw.size] = raw
        P = ctypes.c_char_p
        self.beg = (ctypes.c_void_p * NFIELDS)(*[self.start + o for o in offs])
        self.end = (ctypes.c_void_p * NFIELDS)(*[self.start + o + a.nbytes for a, o in zip(arrays, offs)])
        self.sz = (ctypes.c_size_t * (NFIELDS + 1))(*(offs + [total]))


def pack(ps, lens_override=None):
    """The packed buffers of a whole profile set (a query batch or a database
    chunk); lens_override {k: n} writes length n for profile k (malformed
    buffers for the error tests)."""
    lens = ps.lens().astype(np.int32)
    nprof, npos = ps.nprof, ps.npos
    dist = np.zeros(nprof, np.uint32)
    if nprof:
        dist[1:] = np.cumsum(lens[:-1])
    bkg = ps.field("bkg").reshape(nprof, NAA)
    trn = ps.field("trn").reshape(NTRN, npos + nprof)
    trg = ps.field("trg").reshape(NAA, npos)
    cv = ps.field("cv").reshape(NCV, npos)
    ss = ps.field("ss").reshape(3, npos)
    dd = ps.field("dd").reshape(npos, -1)
    ddv = np.full((NDD, npos), 255.0, np.float32)
    ddv[0] = dd[:, 0]
    for k in range(1, NDD):
        has = dd[:, 0] >= k
        ddv[k, has] = dd[has, k]
    addrpro = np.repeat(np.arange(nprof, dtype=np.int32), lens)
    addrss = np.zeros(npos, np.uint32)
    arrays = [np.ascontiguousarray(bkg[:, r], np.float32) for r in range(NAA)]
    stored = lens.copy()
    for k, v in (lens_override or {}).items():
        stored[k] = v
    arrays += [ps.field("eno").astype(np.float32), stored, dist]
    arrays += [np.ascontiguousarray(trn[t], np.float32) for t in range(NTRN)]
    arrays += [np.ascontiguousarray(trg[r], np.float32) for r in range(NAA)]
    arrays += [np.ascontiguousarray(cv[k], np.float32) for k in range(NCV)]
    arrays += [ps.field("cvprior").astype(np.float32), ps.field("cvnorm2").astype(np.float32)]
    arrays += [ddv[k] for k in range(NDD)]
    arrays += [np.ascontiguousarray(ss[q], np.float32) for q in range(3)]
    arrays += [np.zeros(npos, np.float32), np.zeros(npos, np.int32), addrpro, np.zeros(npos, np.uint32), addrss]
    arrays += [ps.field("aa").view(np.uint8).astype(np.uint8), ps.field("sss").view(np.uint8).astype(np.uint8)]
    return Packed(arrays)


def skip(pk, lens, k):
    """Field pointers advanced past the first k profiles, as
    PMBatchProData::PMDataSkipNPros advances them (profile fields by k
    entries, transition fields by the positions plus k beginning vectors,
    position fields by the positions)."""
    npos = int(np.asarray(lens[:k]).sum())
    out = []
    for f in range(NFIELDS):
        b = pk.beg[f]
        if f < F_TRN:
            out.append(b + 4 * k)
        elif f < F_TRG:
            out.append(b + 4 * (npos + k))
        elif f in (F_AA, F_SSSTATE):
            out.append(b + npos)
        else:
            out.append(b + 4 * npos)
    return out


# section order of the db2bin file's address table (libpro/srcpro/
# PMProfileModel.cpp:1250-1400, layout v2.2): lengths, ENO, 20 background
# probabilities, 7 transitions, 20 target frequencies, 19 CV entries, CV prior,
# CV norm, 100 distance values, 3 SS probabilities, HDP1 probability and
# index, residues, SS states, description end addresses, descriptions
S_LEN, S_ENO, S_BKG = 0, 1, 2
S_TRN = S_BKG + NAA
S_TRG = S_TRN + NTRN
S_CV = S_TRG + NAA
S_CVPRIOR = S_CV + NCV
S_CVNORM2 = S_CVPRIOR + 1
S_DST = S_CVNORM2 + 1
S_SS = S_DST + NDD
S_HDPPRB = S_SS + 3
S_HDPIND = S_HDPPRB + 1
S_AA = S_HDPIND + 1
S_SSSTATE = S_AA + 1
NSECTIONS = S_SSSTATE + 3
SIGNATURE = b"COTHER profile bin database v2.5"


def pack_sections(path, p0, p1):
    """The chunk buffer of profiles [p0,p1) of a db2bin file, its sections
    copied as TdDataReader::ReadDataChunkHelper copies them (the same byte
    ranges of each section; offsets, profile addresses and SS addresses filled
    as its final loop fills them)."""
    raw = np.fromfile(path, np.uint8)
    ls = len(SIGNATURE)
    assert raw[:ls].tobytes() == SIGNATURE
    nprof, npos = (int(v) for v in raw[ls:ls + 16].view(np.uint64))
    addr = raw[ls + 24:ls + 24 + 8 * NSECTIONS].view(np.uint64).astype(np.int64)
    lens_all = raw[addr[S_LEN]:addr[S_LEN] + 4 * nprof].view(np.int32)
    pos0 = int(lens_all[:p0].sum())
    lens = lens_all[p0:p1].copy()
    n, tot = p1 - p0, int(lens.sum())

    def sec(s, itemsize, start, count):
        a = int(addr[s]) + itemsize * start
        return raw[a:a + itemsize * count].copy()

    arrays = [sec(S_BKG + r, 4, p0, n) for r in range(NAA)]
    arrays += [sec(S_ENO, 4, p0, n), lens.view(np.uint8).copy()]
    dist = np.zeros(n, np.uint32)
    if n:
        dist[1:] = np.cumsum(lens[:-1])
    arrays.append(dist.view(np.uint8))
    arrays += [sec(S_TRN + t, 4, p0 + pos0, n + tot) for t in range(NTRN)]
    arrays += [sec(S_TRG + r, 4, pos0, tot) for r in range(NAA)]
    arrays += [sec(S_CV + k, 4, pos0, tot) for k in range(NCV)]
    arrays += [sec(S_CVPRIOR, 4, pos0, tot), sec(S_CVNORM2, 4, pos0, tot)]
    arrays += [sec(S_DST + k, 4, pos0, tot) for k in range(NDD)]
    ss = [sec(S_SS + q, 4, pos0, tot) for q in range(3)]
    arrays += ss
    arrays += [sec(S_HDPPRB, 4, pos0, tot), sec(S_HDPIND, 4, pos0, tot)]
    arrays.append(np.repeat(np.arange(n, dtype=np.int32), lens).view(np.uint8))
    arrays.append(np.zeros(tot, np.uint32).view(np.uint8))
    addrss = np.zeros(tot, np.uint32)
    arrays.append(addrss.view(np.uint8))
    arrays += [sec(S_AA, 1, pos0, tot), sec(S_SSSTATE, 1, pos0, tot)]
    return Packed(arrays)
