// rootwrite.cpp -- ROOT-free writer of one tree per file (mf_io.h), for the
// CPU driver's output tree (multifit_cpu/src/multifit.cpp:107-129,266-279:
// RecoAndSim, branches of /F, /I and std::vector<double|int>).
//
// The file is laid out as ROOT 6 writes a small file (32-bit seeks), in the
// class versions ROOT 6 holds for its own classes, so that it needs no
// streamer info: TTree 20, TBranch 13, TBranchElement 10, TLeaf 2,
// TLeafF/D/I/Element 1, TObjArray 3, TNamed 1, TAttLine/Fill/Marker 2,
// TDirectory 5, TList 5.  All integers and floats are big-endian.
//   header (100 bytes): "root", version, fBEGIN = 100, fEND, fSeekFree,
//     fNbytesFree, nfree, fNbytesName, units 4, compression 0, fSeekInfo,
//     fNbytesInfo, the UUID;
//   at 100 the top directory's record (class TFile): its name and title,
//     then the TDirectory (datimes, its keys list and its own seek), with
//     12 reserved bytes;
//   one TBasket record per basket (uncompressed): the key header, then
//     fVersion 3, fBufferSize, fNevBufSize, fNevBuf, fLast and a flag byte;
//     the entries; for entries of varying size the offset array (count
//     fNevBuf + 1, the key-relative offsets, a trailing 0);
//   the TTree record: the tree, its branches (each with its leaf and the
//     size, first entry and seek of each basket) and its leaves, written as
//     references to the branches' leaves;
//   an empty StreamerInfo list, the keys list (the tree's key header) and
//     the free-segments record (fEND .. 2000000000).
// Objects inside a record are written as ROOT does: a byte count (bit 30
// set), a class tag (0xFFFFFFFF and the class name the first time, then
// 0x80000000 | the key-relative offset of that first tag + 2), the object;
// a second pointer to an object already written is its key-relative offset
// + 2.
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/mf_io.h"

namespace mf {
mf_status set_error(mf_status st, const char *fmt, ...);
}

namespace {

constexpr int32_t BEGIN = 100;
constexpr uint32_t KBC = 0x40000000u;     // byte-count flag
constexpr uint32_t KCLASS = 0x80000000u;  // class-tag flag
constexpr uint32_t KNEWCLASS = 0xFFFFFFFFu;
constexpr int32_t BASKET_SIZE = 32000;    // ROOT's default basket size
constexpr int32_t VEC_OFFSET_LEN = 400;   // fEntryOffsetLen of a vector branch
constexpr int32_t TREE_OFFSET_LEN = 1000; // fDefaultEntryOffsetLen
constexpr uint16_t VEC_STREAM_VERSION = 9;  // std::vector entry version
constexpr uint32_t CHK_VEC_F64 = 0x100e4364u;  // ROOT's checksums of the classes
constexpr uint32_t CHK_VEC_I32 = 0x00983b88u;
constexpr uint32_t CHK_IOFEATURES = 0x1aa12f10u;

// big-endian output buffer
struct Out {
    std::vector<uint8_t> b;
    size_t size() const { return b.size(); }
    void u8(uint8_t v) { b.push_back(v); }
    void be(uint64_t v, int k)
    {
        for (int i = k - 1; i >= 0; --i) b.push_back((uint8_t)(v >> (8 * i)));
    }
    void i16(int16_t v) { be((uint16_t)v, 2); }
    void i32(int32_t v) { be((uint32_t)v, 4); }
    void u32(uint32_t v) { be(v, 4); }
    void i64(int64_t v) { be((uint64_t)v, 8); }
    void f32(float v)
    {
        uint32_t u;
        memcpy(&u, &v, 4);
        u32(u);
    }
    void f64(double v)
    {
        uint64_t u;
        memcpy(&u, &v, 8);
        be(u, 8);
    }
    void str(const std::string &s)  // ROOT string: length byte (or 255 + int32)
    {
        if (s.size() < 255) {
            u8((uint8_t)s.size());
        } else {
            u8(255);
            i32((int32_t)s.size());
        }
        b.insert(b.end(), s.begin(), s.end());
    }
    void cstr(const std::string &s)
    {
        b.insert(b.end(), s.begin(), s.end());
        u8(0);
    }
    void put_u32(size_t at, uint32_t v)
    {
        for (int i = 0; i < 4; ++i) b[at + (size_t)i] = (uint8_t)(v >> (8 * (3 - i)));
    }
    void put_i64(size_t at, int64_t v)
    {
        for (int i = 0; i < 8; ++i) b[at + (size_t)i] = (uint8_t)((uint64_t)v >> (8 * (7 - i)));
    }
    void append(const Out &o) { b.insert(b.end(), o.b.begin(), o.b.end()); }
};

size_t rstr_len(const std::string &s) { return (s.size() < 255 ? 1 : 5) + s.size(); }

uint32_t datime_now()
{
    time_t t = time(nullptr);
    struct tm g;
    gmtime_r(&t, &g);
    return (uint32_t)(g.tm_year + 1900 - 1995) << 26 | (uint32_t)(g.tm_mon + 1) << 22 |
           (uint32_t)g.tm_mday << 17 | (uint32_t)g.tm_hour << 12 | (uint32_t)g.tm_min << 6 |
           (uint32_t)g.tm_sec;
}

// Key header: Nbytes, version, ObjLen, Datime, KeyLen, Cycle, SeekKey,
// SeekPdir (int32, or int64 for key version > 1000), class, name, title.
// extra: bytes of the object's own header that KeyLen counts (a basket's).
struct KeyHdr {
    int32_t nbytes = 0, objlen = 0;
    int16_t version = 4, cycle = 1;
    uint32_t datime = 0;
    int64_t seekkey = 0, seekpdir = 0;
    size_t extra = 0;
    std::string cls, name, title;
    size_t keylen() const
    {
        return 18 + (version > 1000 ? 16 : 8) + rstr_len(cls) + rstr_len(name) + rstr_len(title) +
               extra;
    }
    void write(Out &o) const
    {
        o.i32(nbytes);
        o.i16(version);
        o.i32(objlen);
        o.u32(datime);
        o.i16((int16_t)keylen());
        o.i16(cycle);
        if (version > 1000) {
            o.i64(seekkey);
            o.i64(seekpdir);
        } else {
            o.i32((int32_t)seekkey);
            o.i32((int32_t)seekpdir);
        }
        o.str(cls);
        o.str(name);
        o.str(title);
    }
};

// An object record being written: positions are key-relative (keylen + the
// offset in the object), as ROOT's buffers count them for tags and refs.
struct ObjWriter {
    Out o;
    size_t keylen;
    std::map<std::string, uint32_t> cls_tag;  // class name -> its first tag's offset + 2
    explicit ObjWriter(size_t kl) : keylen(kl) {}
    uint32_t pos() const { return (uint32_t)(keylen + o.size()); }
    // byte count + version of a versioned streamer; returns the count's position
    size_t begin(int16_t version)
    {
        const size_t at = o.size();
        o.u32(0);
        o.i16(version);
        return at;
    }
    void end(size_t at) { o.put_u32(at, (uint32_t)(o.size() - at - 4) | KBC); }
    // pointer to a new object of class cls: byte count, class tag; returns
    // the byte count's position (the object's reference is pos + 2)
    size_t begin_ptr(const std::string &cls, uint32_t *ref)
    {
        const size_t at = o.size();
        *ref = pos() + 2;
        o.u32(0);
        auto it = cls_tag.find(cls);
        if (it == cls_tag.end()) {
            cls_tag[cls] = pos() + 2;
            o.u32(KNEWCLASS);
            o.cstr(cls);
        } else {
            o.u32(it->second | KCLASS);
        }
        return at;
    }
    void end_ptr(size_t at) { o.put_u32(at, (uint32_t)(o.size() - at - 4) | KBC); }
    void tobject(uint32_t bits)
    {
        o.i16(1);
        o.u32(0);
        o.u32(bits);
    }
    void tnamed(const std::string &name, const std::string &title, uint32_t bits)
    {
        const size_t at = begin(1);
        tobject(bits);
        o.str(name);
        o.str(title);
        end(at);
    }
    void iofeatures()
    {
        // version 0: followed by the class checksum; fIOBits 0
        const size_t at = o.size();
        o.u32(0);
        o.i16(0);
        o.u32(CHK_IOFEATURES);
        o.u8(0);
        end(at);
    }
    // TObjArray header: count, lower bound; the caller writes the elements
    size_t objarray(uint32_t bits, int32_t n)
    {
        const size_t at = begin(3);
        tobject(bits);
        o.str("");
        o.i32(n);
        o.i32(0);
        return at;
    }
};

struct Column {
    std::string name;
    int32_t kind;
    const double *values;
    const int64_t *counts;
    // baskets: first entry, entries, key position, Nbytes, ObjLen + KeyLen
    std::vector<int64_t> first, nent, seek;
    std::vector<int32_t> nbytes;
    int64_t tot = 0;
};

bool is_vec(int32_t k) { return k == MF_IO_VEC_F64 || k == MF_IO_VEC_I32; }
int elem_size(int32_t k) { return (k == MF_IO_F64 || k == MF_IO_VEC_F64) ? 8 : 4; }

void put_value(Out &o, int32_t k, double v)
{
    if (k == MF_IO_F32) o.f32((float)v);
    else if (k == MF_IO_F64 || k == MF_IO_VEC_F64) o.f64(v);
    else o.i32((int32_t)v);
}

// Baskets of one column, appended to the file at `file` (its size is the
// next record's seek).  An entry goes into the current basket while the
// basket's entries stay within BASKET_SIZE bytes (at least one entry).
void write_baskets(Out &file, Column &c, const std::string &tree, int64_t nentries,
                   uint32_t datime)
{
    std::vector<int64_t> vfirst(1, 0);  // vector kinds: first value of each entry
    if (is_vec(c.kind))
        for (int64_t e = 0; e < nentries; ++e) vfirst.push_back(vfirst.back() + c.counts[e]);
    auto entry_bytes = [&](int64_t e) -> int64_t {
        return is_vec(c.kind) ? 10 + c.counts[e] * elem_size(c.kind) : elem_size(c.kind);
    };
    int64_t e0 = 0;
    int cycle = 0;
    while (e0 < nentries) {
        int64_t e1 = e0, bytes = 0;
        while (e1 < nentries && (e1 == e0 || bytes + entry_bytes(e1) <= BASKET_SIZE))
            bytes += entry_bytes(e1++);
        KeyHdr k;
        k.version = 1004;
        k.cycle = (int16_t)cycle++;
        k.datime = datime;
        k.cls = "TBasket";
        k.name = c.name;
        k.title = tree;
        k.extra = 19;  // the basket header
        const size_t keylen = k.keylen();
        Out data;
        std::vector<int32_t> offs;
        for (int64_t e = e0; e < e1; ++e) {
            offs.push_back((int32_t)(keylen + data.size()));
            if (is_vec(c.kind)) {
                const int64_t n = c.counts[e];
                data.u32((uint32_t)(6 + n * elem_size(c.kind)) | KBC);
                data.i16((int16_t)VEC_STREAM_VERSION);
                data.i32((int32_t)n);
                for (int64_t i = 0; i < n; ++i) put_value(data, c.kind, c.values[vfirst[e] + i]);
            } else {
                put_value(data, c.kind, c.values[e]);
            }
        }
        const int32_t last = (int32_t)(keylen + data.size());
        if (is_vec(c.kind)) {
            data.i32((int32_t)(e1 - e0 + 1));
            for (int32_t v : offs) data.i32(v);
            data.i32(0);
        }
        k.objlen = (int32_t)data.size();
        k.nbytes = (int32_t)(keylen + data.size());
        k.seekkey = (int64_t)file.size();
        k.seekpdir = BEGIN;
        Out rec;
        k.write(rec);
        rec.i16(3);                                             // fVersion
        rec.i32(BASKET_SIZE);                                   // fBufferSize
        rec.i32(is_vec(c.kind) ? TREE_OFFSET_LEN : elem_size(c.kind));  // fNevBufSize
        rec.i32((int32_t)(e1 - e0));                            // fNevBuf
        rec.i32(last);                                          // fLast
        rec.u8(0);
        rec.append(data);
        c.first.push_back(e0);
        c.nent.push_back(e1 - e0);
        c.seek.push_back(k.seekkey);
        c.nbytes.push_back(k.nbytes);
        c.tot += k.nbytes;
        file.append(rec);
        e0 = e1;
    }
}

// The TTree object (TTree::Streamer, version 20).
void write_tree_object(ObjWriter &w, const std::string &tree, const std::string &title,
                       std::vector<Column> &cols, int64_t nentries)
{
    Out &o = w.o;
    const size_t at_tree = w.begin(20);
    w.tnamed(tree, title, 0x03000008u);
    {   // TAttLine, TAttFill, TAttMarker: ROOT's defaults for a tree
        size_t a = w.begin(2);
        o.i16(602);
        o.i16(1);
        o.i16(1);
        w.end(a);
        a = w.begin(2);
        o.i16(0);
        o.i16(1001);
        w.end(a);
        a = w.begin(2);
        o.i16(1);
        o.i16(1);
        o.f32(1.0f);
        w.end(a);
    }
    int64_t tot = 0;
    for (const Column &c : cols) tot += c.tot;
    o.i64(nentries);  // fEntries
    o.i64(tot);       // fTotBytes
    o.i64(tot);       // fZipBytes (uncompressed)
    o.i64(0);         // fSavedBytes
    o.i64(0);         // fFlushedBytes
    o.f64(1.0);       // fWeight
    o.i32(0);         // fTimerInterval
    o.i32(25);        // fScanField
    o.i32(0);         // fUpdate
    o.i32(TREE_OFFSET_LEN);
    o.i32(0);         // fNClusterRange
    o.i64(1000000000000LL);  // fMaxEntries
    o.i64(1000000000000LL);  // fMaxEntryLoop
    o.i64(0);                // fMaxVirtualSize
    o.i64(-300000000LL);     // fAutoSave
    o.i64(-30000000LL);      // fAutoFlush
    o.i64(1000000);          // fEstimate
    o.u8(0);                 // fClusterRangeEnd (none)
    o.u8(0);                 // fClusterSize (none)
    w.iofeatures();
    // fBranches
    std::vector<uint32_t> leaf_ref;
    const size_t at_br = w.objarray(0x03004000u, (int32_t)cols.size());
    for (const Column &c : cols) {
        const bool vec = is_vec(c.kind);
        uint32_t ref;
        const size_t at_obj = w.begin_ptr(vec ? "TBranchElement" : "TBranch", &ref);
        const size_t at_be = vec ? w.begin(10) : 0;
        const size_t at_b = w.begin(13);
        const char *suffix = c.kind == MF_IO_F32 ? "/F" : c.kind == MF_IO_F64 ? "/D" : "/I";
        w.tnamed(c.name, vec ? c.name : c.name + suffix, 0x03400000u);
        {
            const size_t a = w.begin(2);  // TAttFill
            o.i16(0);
            o.i16(1001);
            w.end(a);
        }
        const int32_t nb = (int32_t)c.seek.size();
        const int32_t maxb = nb + 1 > 10 ? nb + 1 : 10;
        o.i32(0);                         // fCompress
        o.i32(BASKET_SIZE);               // fBasketSize
        o.i32(vec ? VEC_OFFSET_LEN : 0);  // fEntryOffsetLen
        o.i32(nb);                        // fWriteBasket
        o.i64(nentries);                  // fEntryNumber
        w.iofeatures();
        o.i32(0);                   // fOffset
        o.i32(maxb);                // fMaxBaskets
        o.i32(vec ? 99 : 0);        // fSplitLevel
        o.i64(nentries);            // fEntries
        o.i64(0);                   // fFirstEntry
        o.i64(c.tot);               // fTotBytes
        o.i64(c.tot);               // fZipBytes
        w.end(w.objarray(0x03000000u, 0));  // fBranches
        {   // fLeaves: one leaf
            const size_t a = w.objarray(0x03000000u, 1);
            const char *lc = vec ? "TLeafElement"
                                 : c.kind == MF_IO_F32 ? "TLeafF"
                                 : c.kind == MF_IO_F64 ? "TLeafD" : "TLeafI";
            uint32_t lref;
            const size_t al = w.begin_ptr(lc, &lref);
            leaf_ref.push_back(lref);
            const size_t a1 = w.begin(1);
            const size_t a2 = w.begin(2);  // TLeaf
            w.tnamed(c.name, c.name, 0x03000000u);
            o.i32(1);                                   // fLen
            o.i32(vec ? 4 : elem_size(c.kind));         // fLenType
            o.i32(0);                                   // fOffset
            o.u8(0);                                    // fIsRange
            o.u8(0);                                    // fIsUnsigned
            o.u32(0);                                   // fLeafCount (none)
            w.end(a2);
            if (vec) {
                o.i32(-1);  // fID
                o.i32(-1);  // fType
            } else if (c.kind == MF_IO_F64) {
                o.f64(0.0);
                o.f64(0.0);
            } else if (c.kind == MF_IO_F32) {
                o.f32(0.0f);
                o.f32(0.0f);
            } else {
                o.i32(0);
                o.i32(0);
            }
            w.end(a1);
            w.end_ptr(al);
            w.end(a);
        }
        {   // fBaskets: none left in memory
            const size_t a = w.objarray(0x03000000u, nb + 1);
            for (int32_t i = 0; i <= nb; ++i) o.u32(0);
            w.end(a);
        }
        o.u8(1);  // fBasketBytes
        for (int32_t i = 0; i < maxb; ++i) o.i32(i < nb ? c.nbytes[(size_t)i] : 0);
        o.u8(1);  // fBasketEntry
        for (int32_t i = 0; i < maxb; ++i)
            o.i64(i < nb ? c.first[(size_t)i] : i == nb ? nentries : 0);
        o.u8(1);  // fBasketSeek
        for (int32_t i = 0; i < maxb; ++i) o.i64(i < nb ? c.seek[(size_t)i] : 0);
        o.str("");  // fFileName
        w.end(at_b);
        if (vec) {
            o.str(c.kind == MF_IO_VEC_F64 ? "vector<double>" : "vector<int>");  // fClassName
            o.str("");                                                          // fParentName
            o.str("");                                                          // fClonesName
            o.u32(c.kind == MF_IO_VEC_F64 ? CHK_VEC_F64 : CHK_VEC_I32);         // fCheckSum
            o.i16(6);    // fClassVersion
            o.i32(-1);   // fID
            o.i32(0);    // fType
            o.i32(-1);   // fStreamerType
            o.i32(0);    // fMaximum
            o.u32(0);    // fBranchCount
            o.u32(0);    // fBranchCount2
            w.end(at_be);
        }
        w.end_ptr(at_obj);
    }
    w.end(at_br);
    {   // fLeaves: references to the branches' leaves
        const size_t a = w.objarray(0x03000000u, (int32_t)leaf_ref.size());
        for (uint32_t r : leaf_ref) o.u32(r);
        w.end(a);
    }
    o.u32(0);  // fAliases
    o.i32(0);  // fIndexValues
    o.i32(0);  // fIndex
    o.u32(0);  // fTreeIndex
    o.u32(0);  // fFriends
    o.u32(0);  // fUserInfo
    o.u32(0);  // fBranchRef
    w.end(at_tree);
}

}  // namespace

extern "C" mf_status mf_io_write_tree(const char *path, const char *tree, const char *title,
                                      const mf_io_column *cols, int32_t ncols, int64_t nentries)
{
    if (!path || !tree || !cols || ncols < 1 || nentries < 0)
        return mf::set_error(MF_ERR_INVALID, "mf_io_write_tree: bad argument");
    std::vector<Column> cs;
    for (int32_t i = 0; i < ncols; ++i) {
        const mf_io_column &c = cols[i];
        if (!c.name || !c.name[0] || c.kind < MF_IO_F32 || c.kind > MF_IO_VEC_I32 ||
            (nentries > 0 && !c.values) || (is_vec(c.kind) && nentries > 0 && !c.counts))
            return mf::set_error(MF_ERR_INVALID, "mf_io_write_tree: bad column %d", i);
        if (is_vec(c.kind))
            for (int64_t e = 0; e < nentries; ++e)
                if (c.counts[e] < 0 || c.counts[e] > (int64_t)100000000)
                    return mf::set_error(MF_ERR_INVALID,
                                         "mf_io_write_tree: column %s: bad count at entry %lld",
                                         c.name, (long long)e);
        cs.push_back(Column{c.name, c.kind, c.values, c.counts, {}, {}, {}, {}, 0});
    }
    const std::string fname = path, ftitle = title ? title : "", tname = tree;
    const uint32_t dt = datime_now();

    // the top directory's record (class TFile): its size is fixed by the names
    KeyHdr top;
    top.cls = "TFile";
    top.name = fname;
    top.title = ftitle;
    top.datime = dt;
    top.seekkey = BEGIN;
    const size_t top_obj = rstr_len(fname) + rstr_len(ftitle) + 2 + 4 * 7 + 18 + 12;
    top.objlen = (int32_t)top_obj;
    top.nbytes = (int32_t)(top.keylen() + top_obj);

    Out file;
    file.b.resize((size_t)BEGIN + (size_t)top.nbytes, 0);
    for (Column &c : cs) write_baskets(file, c, tname, nentries, dt);

    // the tree
    KeyHdr tk;
    tk.cls = "TTree";
    tk.name = tname;
    tk.title = "";
    tk.datime = dt;
    tk.seekpdir = BEGIN;
    ObjWriter tw(tk.keylen());
    write_tree_object(tw, tname, "", cs, nentries);
    tk.objlen = (int32_t)tw.o.size();
    tk.nbytes = (int32_t)(tk.keylen() + tw.o.size());
    tk.seekkey = (int64_t)file.size();
    {
        Out rec;
        tk.write(rec);
        rec.append(tw.o);
        file.append(rec);
    }
    // an empty StreamerInfo list
    KeyHdr sk;
    sk.cls = "TList";
    sk.name = "StreamerInfo";
    sk.title = "Doubly linked list";
    sk.datime = dt;
    sk.seekpdir = BEGIN;
    {
        ObjWriter w(sk.keylen());
        const size_t a = w.begin(5);
        w.tobject(0x02000000u);
        w.o.str("");
        w.o.i32(0);
        w.end(a);
        sk.objlen = (int32_t)w.o.size();
        sk.nbytes = (int32_t)(sk.keylen() + w.o.size());
        sk.seekkey = (int64_t)file.size();
        Out rec;
        sk.write(rec);
        rec.append(w.o);
        file.append(rec);
    }
    // the keys list: the tree's key header
    KeyHdr kk;
    kk.cls = "TFile";
    kk.name = fname;
    kk.title = ftitle;
    kk.datime = dt;
    kk.seekpdir = BEGIN;
    {
        Out body;
        body.i32(1);
        tk.write(body);
        kk.objlen = (int32_t)body.size();
        kk.nbytes = (int32_t)(kk.keylen() + body.size());
        kk.seekkey = (int64_t)file.size();
        Out rec;
        kk.write(rec);
        rec.append(body);
        file.append(rec);
    }
    // free segments: from the end of the file on
    KeyHdr fk = kk;
    {
        fk.objlen = 10;
        fk.nbytes = (int32_t)(fk.keylen() + 10);
        fk.seekkey = (int64_t)file.size();
        const int64_t end = fk.seekkey + fk.nbytes;
        Out rec;
        fk.write(rec);
        rec.i16(1);
        rec.i32((int32_t)end);
        rec.i32(2000000000);
        file.append(rec);
    }
    const int64_t fend = (int64_t)file.size();
    if (fend > 2000000000LL)
        return mf::set_error(MF_ERR_UNSUPPORTED, "mf_io_write_tree: more than 2 GB (large files not written)");

    // UUID: version 1 and 16 bytes from the clock and the file's size
    uint8_t uuid[16];
    {
        const uint64_t a = ((uint64_t)time(nullptr) << 20) ^ (uint64_t)clock();
        const uint64_t b2 = (uint64_t)fend * 0x9E3779B97F4A7C15ull;
        for (int i = 0; i < 8; ++i) {
            uuid[i] = (uint8_t)(a >> (8 * i));
            uuid[8 + i] = (uint8_t)(b2 >> (8 * i));
        }
    }
    // header
    {
        Out h;
        h.b.insert(h.b.end(), {'r', 'o', 'o', 't'});
        h.i32(61206);  // a ROOT 6 version number (small file: below 1000000)
        h.i32(BEGIN);
        h.i32((int32_t)fend);
        h.i32((int32_t)fk.seekkey);
        h.i32(fk.nbytes);
        h.i32(1);  // nfree
        h.i32((int32_t)(top.keylen() + rstr_len(fname) + rstr_len(ftitle)));  // fNbytesName
        h.u8(4);   // fUnits
        h.i32(0);  // fCompress
        h.i32((int32_t)sk.seekkey);
        h.i32(sk.nbytes);
        h.i16(1);
        for (int i = 0; i < 16; ++i) h.u8(uuid[i]);
        memcpy(file.b.data(), h.b.data(), h.b.size());
    }
    // the top directory's record
    {
        Out r;
        top.write(r);
        r.str(fname);
        r.str(ftitle);
        r.i16(5);  // TDirectory version
        r.u32(dt);
        r.u32(dt);
        r.i32(kk.nbytes);
        r.i32((int32_t)(top.keylen() + rstr_len(fname) + rstr_len(ftitle)));
        r.i32(BEGIN);  // fSeekDir
        r.i32(0);      // fSeekParent
        r.i32((int32_t)kk.seekkey);
        r.i16(1);
        for (int i = 0; i < 16; ++i) r.u8(uuid[i]);
        for (int i = 0; i < 12; ++i) r.u8(0);
        if (r.size() != (size_t)top.nbytes)
            return mf::set_error(MF_ERR_INVALID, "mf_io_write_tree: directory record size");
        memcpy(file.b.data() + BEGIN, r.b.data(), r.size());
    }
    FILE *f = fopen(path, "wb");
    if (!f) return mf::set_error(MF_ERR_INVALID, "mf_io_write_tree: cannot open %s", path);
    const size_t w = fwrite(file.b.data(), 1, file.b.size(), f);
    const int rc = fclose(f);
    if (w != file.b.size() || rc != 0)
        return mf::set_error(MF_ERR_INVALID, "mf_io_write_tree: write to %s failed", path);
    return MF_OK;
}
