"""The ROOT-free writer (mf_io_write_tree, csrc/rootwrite.cpp): the CPU
driver's output tree (multifit_cpu/src/multifit.cpp:107-129,266-279).

Checked three ways:
  - read back with the library's reader, every kind, bit for bit;
  - the file's own bookkeeping, decoded by an independent Python reader
    (tests/rootwalk.py): header, directory, keys list, free segments, and
    every basket the tree's branches point at;
  - against the reference's own file written by ROOT 6
    (data/mysample_100_...root): given the values the reference's Samples
    tree holds, the writer's baskets equal ROOT's (decompressed) byte for
    byte, and its tree record decodes to the same classes, class versions
    and members, branch by branch.
No ROOT installation is available to open the files; the reference's file
stands in for it.
"""
import struct
from pathlib import Path

import numpy as np
import pytest

import rootwalk

REF = Path("/root/reference/data")
F100 = REF / "mysample_100_0.000_0.000_10_25.00_10.00_0.00_1.000_1.00_CRRC43_0.00_slew_1.00.root"
needs_ref = pytest.mark.skipif(not F100.exists(), reason="reference data files not present")


def _branches(n, rng):
    from test_regression_amd import _abi
    return {
        "chi2": (_abi.MF_IO_F32, rng.normal(size=n) * 1e3),
        "ipulseintime": (_abi.MF_IO_I32, rng.integers(-5, 100, size=n)),
        "amplitudeTruth": (_abi.MF_IO_F64, rng.normal(size=n)),
        "samplesReco": (_abi.MF_IO_VEC_F64, [rng.normal(size=10) for _ in range(n)]),
        "activeBXs": (_abi.MF_IO_VEC_I32, [rng.integers(-5, 5, size=rng.integers(0, 11))
                                           for _ in range(n)]),
        "complete_chi2": (_abi.MF_IO_VEC_F64, [np.zeros(0)] * n),
        "long": (_abi.MF_IO_VEC_F64, [rng.normal(size=1500) for _ in range(n)]),
    }


def _expected(kind, vals):
    from test_regression_amd import _abi
    if kind in (_abi.MF_IO_VEC_F64, _abi.MF_IO_VEC_I32):
        flat = np.concatenate([np.asarray(v, dtype=np.float64) for v in vals])
        if kind == _abi.MF_IO_VEC_I32:
            flat = np.trunc(flat)
        return flat, np.array([len(v) for v in vals])
    v = np.asarray(vals, dtype=np.float64)
    if kind == _abi.MF_IO_F32:
        v = v.astype(np.float32).astype(np.float64)
    elif kind == _abi.MF_IO_I32:
        v = np.trunc(v)
    return v, np.ones(v.size, dtype=np.int64)


def test_round_trip_every_kind(pkg, tmp_path):
    from test_regression_amd import refdata
    rng = np.random.default_rng(5)
    br = _branches(300, rng)
    out = tmp_path / "t.root"
    refdata.write_tree(str(out), "RecoAndSim", br)
    with refdata.RootFile(str(out)) as f:
        assert f.tree_entries("RecoAndSim") == 300
        for name, (kind, vals) in br.items():
            v, c = f.read_branch("RecoAndSim", name, kind)
            ev, ec = _expected(kind, vals)
            assert np.array_equal(v, ev), name
            assert np.array_equal(c, ec), name


def test_file_bookkeeping(pkg, tmp_path):
    """Header, top directory, keys list, free segments and every basket, as an
    independent reader decodes them."""
    from test_regression_amd import refdata
    rng = np.random.default_rng(6)
    br = _branches(120, rng)
    out = tmp_path / "t.root"
    refdata.write_tree(str(out), "RecoAndSim", br)
    b = out.read_bytes()
    h = rootwalk.header(b)
    assert h["magic"] == b"root" and h["version"] < 1000000 and h["begin"] == 100
    assert h["end"] == len(b) and h["units"] == 4 and h["compress"] == 0 and h["nfree"] == 1
    recs = {r["pos"]: r for r in rootwalk.records(b)}
    top = recs[100]
    assert top["cls"] == "TFile" and top["name"] == str(out) and top["seekkey"] == 100
    assert h["nbytesname"] == top["keylen"] + 1 + len(str(out)) + 1
    # the directory: its keys list, its own seek
    d = rootwalk.payload(b, top)
    o = 1 + len(str(out)) + 1
    v, _, _, nbkeys, nbname, seekdir, seekparent, seekkeys = struct.unpack(">hIIiiiii",
                                                                          d[o:o + 30])
    assert (v, seekdir, seekparent, nbname) == (5, 100, 0, h["nbytesname"])
    keys = recs[seekkeys]
    assert keys["nbytes"] == nbkeys and keys["cls"] == "TFile"
    kd = rootwalk.payload(b, keys)
    assert struct.unpack(">i", kd[:4])[0] == 1
    tr = [r for r in recs.values() if r["cls"] == "TTree"][0]
    assert kd[4:] == b[tr["pos"]:tr["pos"] + tr["keylen"]]   # the tree's key header
    # free segments and streamer info
    fr = recs[h["seekfree"]]
    assert fr["nbytes"] == h["nbytesfree"] and fr["pos"] + fr["nbytes"] == len(b)
    assert rootwalk.payload(b, fr) == struct.pack(">hii", 1, len(b), 2000000000)
    si = recs[h["seekinfo"]]
    assert (si["cls"], si["name"], si["nbytes"]) == ("TList", "StreamerInfo", h["nbytesinfo"])
    # every basket the branches point at
    t = rootwalk.tree(b, "RecoAndSim")
    assert t["fEntries"] == 120
    bs = rootwalk.branches(t)
    assert list(bs) == list(br)
    total = 0
    for name, (cls, obj) in bs.items():
        tb = obj["branch"] if cls == "TBranchElement" else obj
        nb = tb["fWriteBasket"]
        assert nb >= 1 and tb["fEntries"] == 120
        assert tb["fBasketEntry"][0] == 0 and tb["fBasketEntry"][nb] == 120
        for i in range(nb):
            k = recs[tb["fBasketSeek"][i]]
            assert (k["cls"], k["name"], k["title"]) == ("TBasket", name, "RecoAndSim")
            assert k["nbytes"] == tb["fBasketBytes"][i]
            bh = rootwalk.basket_header(b, k)
            assert bh["nevbuf"] == tb["fBasketEntry"][i + 1] - tb["fBasketEntry"][i]
            assert bh["last"] <= k["keylen"] + k["objlen"]
            total += k["nbytes"]
        assert tb["fTotBytes"] == sum(tb["fBasketBytes"][:nb])
    assert t["fTotBytes"] == total
    # the 1500-value vectors take several baskets of at most 32000 bytes
    long_b = bs["long"][1]["branch"]
    assert long_b["fWriteBasket"] > 1
    assert all(recs[s]["objlen"] <= 32000 + 4 * 20 for s in long_b["fBasketSeek"][:long_b["fWriteBasket"]])


def test_errors(pkg, tmp_path):
    import ctypes as C
    from test_regression_amd import _abi, refdata
    L = _abi.lib()
    assert L.mf_io_write_tree(str(tmp_path / "x.root").encode(), b"T", b"", None, 0, 0) == \
        _abi.MF_ERR_INVALID
    with pytest.raises(_abi.MFError):
        refdata.write_tree(str(tmp_path / "y.root"), "T",
                           {"a": (_abi.MF_IO_F64, np.zeros(3)), "b": (_abi.MF_IO_F64, np.zeros(4))})
    vals = np.zeros(4)
    counts = np.array([1, -1], dtype=np.int64)
    col = _abi.IoColumn(b"v", _abi.MF_IO_VEC_F64, 0, vals.ctypes.data_as(C.c_void_p),
                        counts.ctypes.data_as(C.c_void_p))
    assert L.mf_io_write_tree(str(tmp_path / "z.root").encode(), b"T", b"", C.byref(col), 1,
                              2) == _abi.MF_ERR_INVALID


# the reference's Samples branches written in one basket by ROOT
REF_BRANCHES = [("pulse_shift", "MF_IO_F32"), ("nSmpl", "MF_IO_I32"),
                ("amplitudeTruth", "MF_IO_F64"), ("samples", "MF_IO_VEC_F64"),
                ("samples_noise", "MF_IO_VEC_F64"), ("BX0", "MF_IO_I32"),
                ("signalTruth", "MF_IO_F64"), ("input_pedestal", "MF_IO_F32")]


def _ref_copy(tmp_path):
    """The reference's values of REF_BRANCHES, written by the library as a
    tree of the same name."""
    from test_regression_amd import _abi, refdata
    br = {}
    with refdata.RootFile(str(F100)) as f:
        for name, kind in REF_BRANCHES:
            k = getattr(_abi, kind)
            v, c = f.read_branch("Samples", name, k)
            if k == _abi.MF_IO_VEC_F64:
                br[name] = (k, list(np.split(v, np.cumsum(c)[:-1])))
            else:
                br[name] = (k, v)
    out = tmp_path / "copy.root"
    refdata.write_tree(str(out), "Samples", br)
    return out.read_bytes(), F100.read_bytes()


@needs_ref
def test_baskets_equal_root_bytes(pkg, tmp_path):
    """Given the same values, each branch's basket is ROOT's byte for byte
    (after ROOT's zlib compression is undone): key length, basket header and
    the entries with their offset array."""
    ours, ref = _ref_copy(tmp_path)
    ro = {r["name"]: r for r in rootwalk.records(ours) if r["cls"] == "TBasket"}
    rr = {r["name"]: r for r in rootwalk.records(ref) if r["cls"] == "TBasket"}
    for name, _ in REF_BRANCHES:
        a, b = ro[name], rr[name]
        assert a["keylen"] == b["keylen"] and a["objlen"] == b["objlen"], name
        assert rootwalk.basket_header(ours, a) == rootwalk.basket_header(ref, b), name
        assert rootwalk.payload(ours, a) == rootwalk.payload(ref, b), name


@needs_ref
def test_tree_record_matches_root(pkg, tmp_path):
    """The tree record decodes to ROOT's classes, class versions and members,
    branch by branch; only what describes compression, sizes on disk and
    positions differs (the writer does not compress)."""
    ours, ref = _ref_copy(tmp_path)
    to, tr = rootwalk.tree(ours, "Samples"), rootwalk.tree(ref, "Samples")
    for key in ("version", "named", "attline", "attfill", "attmarker", "fEntries", "fWeight",
                "fTimerInterval", "fScanField", "fUpdate", "fDefaultEntryOffsetLen",
                "fNClusterRange", "fMaxEntries", "fMaxEntryLoop", "fMaxVirtualSize",
                "fAutoSave", "fAutoFlush", "fEstimate", "cluster_speedbumps", "fIOFeatures",
                "tail"):
        assert to[key] == tr[key], key
    bo, br = rootwalk.branches(to), rootwalk.branches(tr)
    skip = {"fCompress", "fZipBytes", "fBasketBytes", "fBasketSeek"}
    for name, _ in REF_BRANCHES:
        (co, oo), (cr, orr) = bo[name], br[name]
        assert co == cr, name
        if co == "TBranchElement":
            for k in ("version", "fClassName", "fParentName", "fClonesName", "fCheckSum",
                      "fClassVersion", "fID", "fType", "fStreamerType", "fMaximum",
                      "fBranchCount", "fBranchCount2"):
                assert oo[k] == orr[k], (name, k)
            oo, orr = oo["branch"], orr["branch"]
        for k, v in orr.items():
            if k in skip:
                continue
            if k in ("fBranches", "fLeaves", "fBaskets"):
                assert [c for c, _ in oo[k]["items"]] == [c for c, _ in v["items"]], (name, k)
                assert oo[k]["bits"] == v["bits"], (name, k)
                continue
            assert oo[k] == v, (name, k)
        (lc_o, lo), (lc_r, lr) = oo["fLeaves"]["items"][0], orr["fLeaves"]["items"][0]
        assert lc_o == lc_r and lo == lr, name


def test_samples_tree_round_trip(pkg, tmp_path):
    """write_samples -> read_samples gives the samples and amplitudes back bit
    for bit; the extra generator branches read back as written."""
    from test_regression_amd import _abi, refdata, workloads
    w = workloads.config("c1")
    s, _, truth = w.generate(500, with_truth=True)
    amp = truth[5].copy()   # the in-time (BX 0) amplitude
    out = tmp_path / "samples.root"
    rng = np.random.default_rng(3)
    extra = {"BX0": rng.integers(0, 10, 500), "pulse_tau": rng.normal(43, 1, 500),
             "energyPU": [rng.normal(size=rng.integers(0, 12)) for _ in range(500)]}
    refdata.write_samples(str(out), s, amp, extra)
    s2, a2 = refdata.read_samples(str(out))
    assert np.array_equal(s2, s) and np.array_equal(a2, amp)
    with refdata.RootFile(str(out)) as f:
        assert f.tree_entries("Samples") == 500
        v, _ = f.read_branch("Samples", "nSmpl", _abi.MF_IO_I32)
        assert np.array_equal(v, np.full(500, 10))
        v, _ = f.read_branch("Samples", "BX0", _abi.MF_IO_I32)
        assert np.array_equal(v, extra["BX0"])
        v, _ = f.read_branch("Samples", "pulse_tau", _abi.MF_IO_F32)
        assert np.array_equal(v, extra["pulse_tau"].astype(np.float32))
        v, c = f.read_branch("Samples", "energyPU", _abi.MF_IO_VEC_F64)
        assert np.array_equal(c, [len(e) for e in extra["energyPU"]])
        assert np.array_equal(v, np.concatenate(extra["energyPU"]))
    names = list(rootwalk.branches(rootwalk.tree(out.read_bytes(), "Samples")))
    assert names == ["nSmpl", "amplitudeTruth", "samples", "BX0", "energyPU", "pulse_tau"]
    with pytest.raises(_abi.MFError):
        refdata.write_samples(str(tmp_path / "bad.root"), s, amp, {"no_such_branch": amp})


@needs_ref
def test_samples_branches_follow_the_generator_file(pkg):
    """The writer's table of Samples branches is the reference file's branch
    list, in its order and with its leaf classes (wf_name, a string, is the
    one branch not written)."""
    from test_regression_amd import _abi, refdata
    t = rootwalk.tree(F100.read_bytes(), "Samples")
    leaf = {_abi.MF_IO_F32: "TLeafF", _abi.MF_IO_F64: "TLeafD", _abi.MF_IO_I32: "TLeafI",
            _abi.MF_IO_VEC_F64: None}
    ref = []
    for name, (cls, obj) in rootwalk.branches(t).items():
        if name == "wf_name":
            continue
        tb = obj["branch"] if cls == "TBranchElement" else obj
        ref.append((name, cls, tb["fLeaves"]["items"][0][0]))
    assert [r[0] for r in ref] == [n for n, _ in refdata.SAMPLES_BRANCHES]
    for (name, cls, lc), (_, kind) in zip(ref, refdata.SAMPLES_BRANCHES):
        if kind == _abi.MF_IO_VEC_F64:
            assert cls == "TBranchElement", name
        else:
            assert (cls, lc) == ("TBranch", leaf[kind]), name
