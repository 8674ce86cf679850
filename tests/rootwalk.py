"""TEST HELPER: a structural reader of ROOT files in pure Python, independent
of the library's reader (csrc/rootio.cpp), for checking what the library's
writer (csrc/rootwrite.cpp) produces against the reference's own files.

It lists a file's records, decompresses them (zlib chunks), and decodes a
TTree record into plain dicts: the tree's fixed members, and per branch its
class, class versions, TNamed, the TBranch members, the leaf and, for a
TBranchElement, its class name, checksum and streamer fields.  Objects are
decoded as ROOT writes them: byte count (bit 30), class tag (0xFFFFFFFF and
the name, or 0x80000000 | key-relative offset + 2), the object; references
to objects already written are checked against the offsets seen.
"""
from __future__ import annotations

import struct
import zlib

KBC = 0x40000000


def _rstr(b, o):
    n = b[o]
    o += 1
    if n == 255:
        n = struct.unpack(">i", b[o:o + 4])[0]
        o += 4
    return b[o:o + n].decode("latin1"), o + n


def header(b: bytes) -> dict:
    ver, begin = struct.unpack(">ii", b[4:12])
    big = ver > 1000000
    fmt = ">qq" if big else ">ii"
    n = struct.calcsize(fmt)
    end, seekfree = struct.unpack(fmt, b[12:12 + n])
    o = 12 + n
    nbfree, nfree, nbname = struct.unpack(">iii", b[o:o + 12])
    o += 12
    units = b[o]
    comp = struct.unpack(">i", b[o + 1:o + 5])[0]
    o += 5
    if big:
        seekinfo = struct.unpack(">q", b[o:o + 8])[0]
        o += 8
    else:
        seekinfo = struct.unpack(">i", b[o:o + 4])[0]
        o += 4
    nbinfo = struct.unpack(">i", b[o:o + 4])[0]
    return dict(magic=b[:4], version=ver, begin=begin, end=end, seekfree=seekfree,
                nbytesfree=nbfree, nfree=nfree, nbytesname=nbname, units=units, compress=comp,
                seekinfo=seekinfo, nbytesinfo=nbinfo)


def records(b: bytes) -> list[dict]:
    """Every record from fBEGIN to the end of the file."""
    pos = header(b)["begin"]
    out = []
    while pos + 18 <= len(b):
        nbytes, kv, objlen, datime, keylen, cycle = struct.unpack(">ihiIhh", b[pos:pos + 18])
        if nbytes < 0:
            pos -= nbytes
            continue
        if nbytes == 0:
            break
        o = pos + 18
        if kv > 1000:
            seekkey, seekpdir = struct.unpack(">qq", b[o:o + 16])
            o += 16
        else:
            seekkey, seekpdir = struct.unpack(">ii", b[o:o + 8])
            o += 8
        cls, o = _rstr(b, o)
        name, o = _rstr(b, o)
        title, o = _rstr(b, o)
        out.append(dict(pos=pos, nbytes=nbytes, version=kv, objlen=objlen, keylen=keylen,
                        cycle=cycle, seekkey=seekkey, seekpdir=seekpdir, cls=cls, name=name,
                        title=title, title_end=o - pos))
        pos += nbytes
    return out


def payload(b: bytes, k: dict) -> bytes:
    """The record's object, decompressed (ObjLen bytes)."""
    d = b[k["pos"] + k["keylen"]:k["pos"] + k["nbytes"]]
    if k["objlen"] <= len(d):
        return d
    res, o = b"", 0
    while o < len(d):
        csz = d[o + 3] | (d[o + 4] << 8) | (d[o + 5] << 16)
        if d[o:o + 2] != b"ZL":
            raise ValueError("only zlib records are decoded")
        res += zlib.decompress(d[o + 9:o + 9 + csz])
        o += 9 + csz
    return res


def basket_header(b: bytes, k: dict) -> dict:
    """fVersion, fBufferSize, fNevBufSize, fNevBuf, fLast and the flag of a
    TBasket record (after the key header's title)."""
    o = k["pos"] + k["title_end"]
    v, bufsize, nevbufsize, nevbuf, last = struct.unpack(">hiiii", b[o:o + 18])
    return dict(version=v, buffersize=bufsize, nevbufsize=nevbufsize, nevbuf=nevbuf, last=last,
                flag=b[o + 18])


class _R:
    def __init__(self, d: bytes, keylen: int):
        self.d, self.o, self.kl = d, 0, keylen
        self.classes, self.objects = {}, set()

    def take(self, fmt):
        v = struct.unpack(fmt, self.d[self.o:self.o + struct.calcsize(fmt)])
        self.o += struct.calcsize(fmt)
        return v[0] if len(v) == 1 else v

    def string(self):
        s, self.o = _rstr(self.d, self.o)
        return s

    def versioned(self):
        bc = self.take(">I")
        start = self.o
        assert bc & KBC, "byte count expected"
        return start + (bc & ~KBC), self.take(">h")

    def tobject(self):
        v, uid, bits = self.take(">hII")
        return dict(version=v, bits=bits)

    def tnamed(self):
        end, v = self.versioned()
        t = self.tobject()
        out = dict(version=v, bits=t["bits"], name=self.string(), title=self.string())
        assert self.o == end
        return out

    def skip(self):
        end, v = self.versioned()
        raw = self.d[self.o:end]
        self.o = end
        return dict(version=v, raw=raw)

    def pointer(self):
        """(class name, decoded object) or ("null", None) / ("ref", offset)."""
        start = self.o
        bc = self.take(">I")
        if bc == 0:
            return "null", None
        if not bc & KBC:
            assert bc in self.objects, f"reference {bc:#x} to no object"
            return "ref", bc
        self.objects.add(start + self.kl + 2)
        tagpos = self.o + self.kl + 2
        tag = self.take(">I")
        if tag == 0xFFFFFFFF:
            e = self.d.index(b"\0", self.o)
            name = self.d[self.o:e].decode()
            self.o = e + 1
            self.classes[tagpos] = name
        else:
            name = self.classes[tag & 0x7FFFFFFF]
        end = start + 4 + (bc & ~KBC)
        obj = getattr(self, "c_" + name, None)
        val = obj() if obj else None
        assert self.o == end or obj is None, f"{name}: {self.o} != {end}"
        self.o = end
        return name, val

    def objarray(self):
        end, v = self.versioned()
        t = self.tobject()
        name = self.string()
        n, lb = self.take(">ii")
        items = [self.pointer() for _ in range(n)]
        assert self.o == end
        return dict(version=v, bits=t["bits"], name=name, lower=lb, items=items)

    def iofeatures(self):
        end, v = self.versioned()
        raw = self.d[self.o:end]
        self.o = end
        return dict(version=v, raw=raw)

    def c_TBranch(self):
        end, v = self.versioned()
        out = dict(version=v, named=self.tnamed(), attfill=self.skip())
        for n, f in (("fCompress", ">i"), ("fBasketSize", ">i"), ("fEntryOffsetLen", ">i"),
                     ("fWriteBasket", ">i"), ("fEntryNumber", ">q")):
            out[n] = self.take(f)
        out["fIOFeatures"] = self.iofeatures()
        for n, f in (("fOffset", ">i"), ("fMaxBaskets", ">i"), ("fSplitLevel", ">i"),
                     ("fEntries", ">q"), ("fFirstEntry", ">q"), ("fTotBytes", ">q"),
                     ("fZipBytes", ">q")):
            out[n] = self.take(f)
        out["fBranches"] = self.objarray()
        out["fLeaves"] = self.objarray()
        out["fBaskets"] = self.objarray()
        mb = out["fMaxBaskets"]
        for n, f in (("fBasketBytes", ">i"), ("fBasketEntry", ">q"), ("fBasketSeek", ">q")):
            sb = self.take(">B")
            out[n] = [self.take(f) for _ in range(mb)] if sb else []
        out["fFileName"] = self.string()
        assert self.o == end
        return out

    def c_TBranchElement(self):
        end, v = self.versioned()
        out = dict(version=v, branch=self.c_TBranch())
        out["fClassName"], out["fParentName"], out["fClonesName"] = (self.string(), self.string(),
                                                                      self.string())
        out["fCheckSum"], out["fClassVersion"] = self.take(">I"), self.take(">h")
        out["fID"], out["fType"], out["fStreamerType"], out["fMaximum"] = self.take(">iiii")
        out["fBranchCount"], out["fBranchCount2"] = self.pointer(), self.pointer()
        assert self.o == end
        return out

    def _tleaf(self):
        end, v = self.versioned()
        out = dict(version=v, named=self.tnamed())
        out["fLen"], out["fLenType"], out["fOffset"] = self.take(">iii")
        out["fIsRange"], out["fIsUnsigned"] = self.take(">BB")
        out["fLeafCount"] = self.pointer()
        assert self.o == end
        return out

    def _leaf(self, fmt):
        end, v = self.versioned()
        out = dict(version=v, leaf=self._tleaf())
        out["extra"] = self.take(fmt)
        assert self.o == end
        return out

    def c_TLeafF(self):
        return self._leaf(">ff")

    def c_TLeafD(self):
        return self._leaf(">dd")

    def c_TLeafI(self):
        return self._leaf(">ii")

    def c_TLeafElement(self):
        return self._leaf(">ii")


def tree(b: bytes, name: str) -> dict:
    """The TTree record `name`, decoded (ROOT 6 layout: TTree version 20)."""
    k = [k for k in records(b) if k["cls"] == "TTree" and k["name"] == name][-1]
    r = _R(payload(b, k), k["keylen"])
    end, v = r.versioned()
    assert v == 20, f"TTree version {v}"
    out = dict(version=v, key=k, named=r.tnamed(), attline=r.skip(), attfill=r.skip(),
               attmarker=r.skip())
    for n in ("fEntries", "fTotBytes", "fZipBytes", "fSavedBytes", "fFlushedBytes"):
        out[n] = r.take(">q")
    out["fWeight"] = r.take(">d")
    for n in ("fTimerInterval", "fScanField", "fUpdate", "fDefaultEntryOffsetLen",
              "fNClusterRange"):
        out[n] = r.take(">i")
    for n in ("fMaxEntries", "fMaxEntryLoop", "fMaxVirtualSize", "fAutoSave", "fAutoFlush",
              "fEstimate"):
        out[n] = r.take(">q")
    out["cluster_speedbumps"] = r.take(">BB")
    out["fIOFeatures"] = r.iofeatures()
    out["fBranches"] = r.objarray()
    out["fLeaves"] = r.objarray()
    out["tail"] = r.d[r.o:end]
    return out


def branches(t: dict) -> dict:
    """name -> (class, decoded branch) of a decoded tree."""
    out = {}
    for cls, obj in t["fBranches"]["items"]:
        br = obj["branch"] if cls == "TBranchElement" else obj
        out[br["named"]["name"]] = (cls, obj)
    return out
