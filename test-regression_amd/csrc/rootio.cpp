// rootio.cpp -- ROOT-free reader of the reference's sample files (mf_io.h).
//
// A ROOT file is a sequence of records ("keys").  Each record starts with a
// header: Nbytes (int32, negative for a free gap), key version (int16),
// ObjLen (int32, uncompressed object size), Datime (uint32), KeyLen (int16,
// header size), Cycle (int16), SeekKey and SeekPdir (int32, or int64 when the
// key version is above 1000), then class name, object name and title as
// ROOT strings (one length byte, or 255 and an int32 length).  All integers
// and floats are big-endian.  The object follows the header: ObjLen bytes
// as is, or compressed in chunks of a 9-byte header (2-byte algorithm tag,
// method byte, 3-byte little-endian compressed and uncompressed sizes).
//
// A tree's data sits in TBasket records: title = tree name, name = branch
// name.  Their header continues after the title with fVersion (int16),
// fBufferSize, fNevBufSize, fNevBuf, fLast (int32) and a flag byte.  In the
// basket's own coordinates the object starts at KeyLen; entries occupy
// [KeyLen, fLast), and when entries vary in size an entry-offset array
// follows (int32 count = fNevBuf + 1, then the offsets, key-relative).  A
// std::vector<T> entry is a 4-byte byte count (bit 30 set), a 2-byte class
// version, an int32 element count and the elements.
//
// A TGraph object is its byte count and version, TNamed, TAttLine, TAttFill
// and TAttMarker (each skipped by its byte count), fNpoints (int32), then
// fX and fY, each a presence byte and fNpoints doubles.
#include <fcntl.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/mf_io.h"

namespace mf {
mf_status set_error(mf_status st, const char *fmt, ...);
}

namespace {

struct Key {
    int64_t pos = 0;
    int32_t nbytes = 0, objlen = 0;
    int32_t keylen = 0, cycle = 0, version = 0;
    int64_t seekpdir = 0;
    int32_t title_end = 0;  // offset (from pos) of the first byte after the title
    std::string cls, name, title;
};

// bounds-checked big-endian cursor over [p, p + n)
struct Cursor {
    const uint8_t *p;
    size_t n, off = 0;
    bool ok = true;
    Cursor(const uint8_t *p_, size_t n_, size_t off_ = 0) : p(p_), n(n_), off(off_) {}
    bool need(size_t k)
    {
        if (!ok || off + k > n || off + k < off) ok = false;
        return ok;
    }
    uint64_t be(int k)
    {
        if (!need((size_t)k)) return 0;
        uint64_t v = 0;
        for (int i = 0; i < k; ++i) v = (v << 8) | p[off + (size_t)i];
        off += (size_t)k;
        return v;
    }
    uint8_t u8() { return (uint8_t)be(1); }
    int16_t i16() { return (int16_t)be(2); }
    int32_t i32() { return (int32_t)be(4); }
    uint32_t u32() { return (uint32_t)be(4); }
    int64_t i64() { return (int64_t)be(8); }
    double f64()
    {
        const uint64_t b = be(8);
        double d;
        memcpy(&d, &b, 8);
        return d;
    }
    float f32()
    {
        const uint32_t b = (uint32_t)be(4);
        float f;
        memcpy(&f, &b, 4);
        return f;
    }
    std::string str()
    {
        size_t len = u8();
        if (len == 255) len = u32();
        if (!need(len)) return {};
        std::string s((const char *)p + off, len);
        off += len;
        return s;
    }
    // skips a versioned sub-object by its byte count (bit 30 marks it)
    void skip_counted()
    {
        const uint32_t bc = u32();
        if (!(bc & 0x40000000u)) {
            ok = false;
            return;
        }
        const size_t len = bc & 0x3fffffffu;
        if (need(len)) off += len;
    }
};

}  // namespace

struct mf_rootfile {
    int fd = -1;
    const uint8_t *data = nullptr;
    size_t size = 0;
    std::vector<Key> keys;
    std::string path;
};

namespace {

mf_status fail(mf_status st, const char *fmt, ...)
{
    char buf[400];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return mf::set_error(st, "%s", buf);
}

// reads the record list from fBEGIN on; stops at the end of the file or at
// the first header that does not describe a record inside it
void scan_keys(mf_rootfile &f, int64_t begin)
{
    int64_t pos = begin;
    while (pos >= 0 && (size_t)pos + 18 <= f.size) {
        Cursor c(f.data, f.size, (size_t)pos);
        const int32_t nbytes = c.i32();
        if (nbytes < 0) {
            pos += -(int64_t)nbytes;
            continue;
        }
        if (nbytes == 0 || (size_t)pos + (size_t)nbytes > f.size) break;
        Key k;
        k.pos = pos;
        k.nbytes = nbytes;
        k.version = c.i16();
        k.objlen = c.i32();
        (void)c.u32();  // datime
        k.keylen = c.i16();
        k.cycle = c.i16();
        if (k.version > 1000) {
            (void)c.i64();
            k.seekpdir = c.i64();
        } else {
            (void)c.i32();
            k.seekpdir = c.i32();
        }
        k.cls = c.str();
        k.name = c.str();
        k.title = c.str();
        if (!c.ok || k.keylen < 18 || k.keylen > nbytes || k.objlen < 0) break;
        k.title_end = (int32_t)(c.off - (size_t)pos);
        f.keys.push_back(std::move(k));
        pos += nbytes;
    }
}

// the record's object, decompressed
mf_status object_bytes(const mf_rootfile &f, const Key &k, std::vector<uint8_t> &out)
{
    const uint8_t *src = f.data + k.pos + k.keylen;
    const size_t clen = (size_t)(k.nbytes - k.keylen);
    out.resize((size_t)k.objlen);
    if ((size_t)k.objlen == clen) {
        memcpy(out.data(), src, clen);
        return MF_OK;
    }
    size_t in = 0, done = 0;
    while (in < clen && done < out.size()) {
        if (in + 9 > clen) return fail(MF_ERR_INVALID, "%s: truncated compressed record %s",
                                       f.path.c_str(), k.name.c_str());
        const uint8_t *h = src + in;
        const size_t csz = (size_t)h[3] | (size_t)h[4] << 8 | (size_t)h[5] << 16;
        const size_t usz = (size_t)h[6] | (size_t)h[7] << 8 | (size_t)h[8] << 16;
        if (h[0] != 'Z' || h[1] != 'L')
            return fail(MF_ERR_UNSUPPORTED,
                        "%s: record %s is compressed with '%c%c' (only zlib is read)",
                        f.path.c_str(), k.name.c_str(), h[0], h[1]);
        if (in + 9 + csz > clen || done + usz > out.size())
            return fail(MF_ERR_INVALID, "%s: bad compressed chunk in %s", f.path.c_str(),
                        k.name.c_str());
        uLongf dl = (uLongf)usz;
        if (uncompress(out.data() + done, &dl, h + 9, (uLong)csz) != Z_OK || dl != usz)
            return fail(MF_ERR_INVALID, "%s: zlib error in %s", f.path.c_str(), k.name.c_str());
        in += 9 + csz;
        done += usz;
    }
    if (done != out.size())
        return fail(MF_ERR_INVALID, "%s: record %s decompresses to %zu of %d bytes",
                    f.path.c_str(), k.name.c_str(), done, k.objlen);
    return MF_OK;
}

// positions of the directory records named `dir` (keys inside point at them
// through SeekPdir)
std::vector<int64_t> dir_positions(const mf_rootfile &f, const std::string &dir)
{
    std::vector<int64_t> v;
    for (const Key &k : f.keys)
        if ((k.cls == "TDirectory" || k.cls == "TDirectoryFile") && k.name == dir) v.push_back(k.pos);
    return v;
}

// splits "dir/name"; a key matches when its parent directory is `dir`
struct Path {
    std::string dir, name;
    explicit Path(const char *s)
    {
        const std::string p(s ? s : "");
        const size_t slash = p.rfind('/');
        if (slash == std::string::npos) {
            name = p;
        } else {
            dir = p.substr(0, slash);
            name = p.substr(slash + 1);
            const size_t inner = dir.rfind('/');
            if (inner != std::string::npos) dir = dir.substr(inner + 1);
        }
    }
    bool in_dir(const mf_rootfile &f, const Key &k) const
    {
        if (dir.empty()) return true;
        for (int64_t d : dir_positions(f, dir))
            if (k.seekpdir == d) return true;
        return false;
    }
};

const Key *find_object(const mf_rootfile &f, const Path &p, const char *cls)
{
    const Key *best = nullptr;
    for (const Key &k : f.keys)
        if (k.cls == cls && k.name == p.name && p.in_dir(f, k))
            if (!best || k.cycle > best->cycle) best = &k;
    return best;
}

}  // namespace

extern "C" mf_status mf_io_open(const char *path, mf_rootfile **file)
{
    if (!path || !file) return fail(MF_ERR_INVALID, "mf_io_open: null argument");
    *file = nullptr;
    const int fd = open(path, O_RDONLY);
    if (fd < 0) return fail(MF_ERR_INVALID, "mf_io_open: cannot open %s", path);
    struct stat st;
    if (fstat(fd, &st) != 0 || st.st_size < 64) {
        close(fd);
        return fail(MF_ERR_INVALID, "mf_io_open: %s is not a ROOT file (too short)", path);
    }
    void *m = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) {
        close(fd);
        return fail(MF_ERR_INVALID, "mf_io_open: cannot map %s", path);
    }
    auto *f = new mf_rootfile;
    f->fd = fd;
    f->data = static_cast<const uint8_t *>(m);
    f->size = (size_t)st.st_size;
    f->path = path;
    Cursor c(f->data, f->size);
    if (memcmp(f->data, "root", 4) != 0) {
        mf_io_close(f);
        return fail(MF_ERR_INVALID, "mf_io_open: %s is not a ROOT file (no 'root' magic)", path);
    }
    c.off = 4;
    (void)c.i32();  // file format version
    const int32_t begin = c.i32();
    scan_keys(*f, begin);
    if (f->keys.empty()) {
        mf_io_close(f);
        return fail(MF_ERR_INVALID, "mf_io_open: %s holds no readable record", path);
    }
    *file = f;
    return MF_OK;
}

extern "C" mf_status mf_io_close(mf_rootfile *file)
{
    if (!file) return MF_OK;
    if (file->data) munmap(const_cast<uint8_t *>(file->data), file->size);
    if (file->fd >= 0) close(file->fd);
    delete file;
    return MF_OK;
}

extern "C" mf_status mf_io_tree_entries(mf_rootfile *file, const char *tree, int64_t *entries)
{
    if (!file || !tree || !entries) return fail(MF_ERR_INVALID, "mf_io_tree_entries: null argument");
    *entries = -1;
    const Key *k = find_object(*file, Path(tree), "TTree");
    if (!k) return MF_OK;
    std::vector<uint8_t> obj;
    mf_status s = object_bytes(*file, *k, obj);
    if (s != MF_OK) return s;
    // TTree: byte count + version, TNamed, TAttLine, TAttFill, TAttMarker,
    // then fEntries (int64)
    Cursor c(obj.data(), obj.size());
    (void)c.u32();
    (void)c.i16();
    for (int i = 0; i < 4; ++i) c.skip_counted();
    const int64_t n = c.i64();
    if (!c.ok || n < 0) return fail(MF_ERR_INVALID, "%s: cannot read the header of tree %s",
                                    file->path.c_str(), tree);
    *entries = n;
    return MF_OK;
}

extern "C" mf_status mf_io_read_branch(mf_rootfile *file, const char *tree, const char *branch,
                                       int32_t kind, double *out, int64_t cap, int64_t *nvalues,
                                       int64_t *counts, int64_t counts_cap, int64_t *nentries)
{
    if (!file || !tree || !branch || !nvalues || !nentries)
        return fail(MF_ERR_INVALID, "mf_io_read_branch: null argument");
    if (kind < MF_IO_F32 || kind > MF_IO_VEC_I32)
        return fail(MF_ERR_INVALID, "mf_io_read_branch: unknown kind %d", kind);
    if (cap < 0 || counts_cap < 0) return fail(MF_ERR_INVALID, "mf_io_read_branch: negative cap");
    const Path tp(tree);
    *nvalues = 0;
    *nentries = 0;
    std::vector<const Key *> baskets;
    for (const Key &k : file->keys)
        if (k.cls == "TBasket" && k.title == tp.name && k.name == branch) baskets.push_back(&k);
    if (baskets.empty())
        return fail(MF_ERR_INVALID, "%s: no basket of branch %s in tree %s", file->path.c_str(),
                    branch, tree);
    // records are written in entry order
    std::sort(baskets.begin(), baskets.end(), [](const Key *a, const Key *b) { return a->pos < b->pos; });
    const bool vec = kind == MF_IO_VEC_F64 || kind == MF_IO_VEC_I32;
    const int esize = (kind == MF_IO_F64 || kind == MF_IO_VEC_F64) ? 8 : 4;
    int64_t nv = 0, ne = 0;
    std::vector<uint8_t> obj;
    auto put = [&](Cursor &c) {
        double v;
        switch (kind) {
        case MF_IO_F32: v = c.f32(); break;
        case MF_IO_I32:
        case MF_IO_VEC_I32: v = c.i32(); break;
        default: v = c.f64(); break;
        }
        if (out && nv < cap) out[nv] = v;
        ++nv;
    };
    for (const Key *bk : baskets) {
        Cursor h(file->data, file->size, (size_t)(bk->pos + bk->title_end));
        (void)h.i16();  // basket version
        (void)h.i32();  // buffer size
        const int32_t nevbufsize = h.i32();
        const int32_t nevbuf = h.i32();
        const int32_t last = h.i32();
        if (!h.ok || nevbuf < 0 || last < bk->keylen || last - bk->keylen > bk->objlen)
            return fail(MF_ERR_INVALID, "%s: bad basket header of %s", file->path.c_str(), branch);
        mf_status s = object_bytes(*file, *bk, obj);
        if (s != MF_OK) return s;
        const size_t dlen = (size_t)(last - bk->keylen);
        if (!vec) {
            if (nevbufsize != esize || dlen != (size_t)nevbuf * (size_t)esize)
                return fail(MF_ERR_INVALID,
                            "%s: branch %s holds %d-byte entries (%zu bytes for %d entries), "
                            "not the kind asked for",
                            file->path.c_str(), branch, nevbufsize, dlen, nevbuf);
            Cursor c(obj.data(), dlen);
            for (int32_t e = 0; e < nevbuf; ++e) {
                put(c);
                if (counts && ne < counts_cap) counts[ne] = 1;
                ++ne;
            }
            continue;
        }
        // variable-size entries: the offset array follows the data
        Cursor oc(obj.data(), obj.size(), dlen);
        const int32_t noff = oc.i32();
        if (!oc.ok || noff != nevbuf + 1)
            return fail(MF_ERR_INVALID, "%s: branch %s has no entry offsets (not a vector branch?)",
                        file->path.c_str(), branch);
        std::vector<int64_t> off((size_t)nevbuf + 1);
        for (int32_t e = 0; e < nevbuf; ++e) off[(size_t)e] = (int64_t)oc.i32() - bk->keylen;
        off[(size_t)nevbuf] = (int64_t)dlen;
        if (!oc.ok) return fail(MF_ERR_INVALID, "%s: truncated offsets in %s", file->path.c_str(), branch);
        for (int32_t e = 0; e < nevbuf; ++e) {
            const int64_t b = off[(size_t)e], en = off[(size_t)e + 1];
            if (b < 0 || en < b || en > (int64_t)dlen)
                return fail(MF_ERR_INVALID, "%s: bad entry offset in %s", file->path.c_str(), branch);
            Cursor c(obj.data(), (size_t)en, (size_t)b);
            const uint32_t bc = c.u32();
            (void)c.i16();  // class version
            const int32_t n = c.i32();
            if (!c.ok || !(bc & 0x40000000u) || n < 0 ||
                (size_t)(bc & 0x3fffffffu) != 6 + (size_t)n * (size_t)esize ||
                (int64_t)(4 + (bc & 0x3fffffffu)) != en - b)
                return fail(MF_ERR_INVALID,
                            "%s: entry %lld of %s is not a vector of %d-byte values",
                            file->path.c_str(), (long long)(ne), branch, esize);
            for (int32_t i = 0; i < n; ++i) put(c);
            if (counts && ne < counts_cap) counts[ne] = n;
            ++ne;
        }
    }
    *nvalues = nv;
    *nentries = ne;
    int64_t header = -1;
    if (mf_io_tree_entries(file, tree, &header) == MF_OK && header >= 0 && header != ne)
        return fail(MF_ERR_UNSUPPORTED,
                    "%s: tree %s has %lld entries but the basket records of %s hold %lld "
                    "(baskets stored inside the tree record are not read)",
                    file->path.c_str(), tree, (long long)header, branch, (long long)ne);
    if ((out && nv > cap) || (counts && ne > counts_cap))
        return fail(MF_ERR_INVALID, "mf_io_read_branch: %s needs %lld values / %lld entries",
                    branch, (long long)nv, (long long)ne);
    return MF_OK;
}

extern "C" mf_status mf_io_read_graph(mf_rootfile *file, const char *name, double *x, double *y,
                                      int64_t cap, int64_t *n)
{
    if (!file || !name || !n) return fail(MF_ERR_INVALID, "mf_io_read_graph: null argument");
    *n = 0;
    const Key *k = find_object(*file, Path(name), "TGraph");
    if (!k) return fail(MF_ERR_INVALID, "%s: no TGraph %s", file->path.c_str(), name);
    std::vector<uint8_t> obj;
    mf_status s = object_bytes(*file, *k, obj);
    if (s != MF_OK) return s;
    Cursor c(obj.data(), obj.size());
    (void)c.u32();
    (void)c.i16();
    for (int i = 0; i < 4; ++i) c.skip_counted();
    const int32_t np = c.i32();
    if (!c.ok || np < 0) return fail(MF_ERR_INVALID, "%s: bad TGraph %s", file->path.c_str(), name);
    const uint8_t hasx = c.u8();
    std::vector<double> xs((size_t)np), ys((size_t)np);
    if (hasx)
        for (int32_t i = 0; i < np; ++i) xs[(size_t)i] = c.f64();
    const uint8_t hasy = c.u8();
    if (hasy)
        for (int32_t i = 0; i < np; ++i) ys[(size_t)i] = c.f64();
    if (!c.ok || !hasx || !hasy)
        return fail(MF_ERR_INVALID, "%s: truncated TGraph %s", file->path.c_str(), name);
    *n = np;
    for (int32_t i = 0; i < np && i < cap; ++i) {
        if (x) x[i] = xs[(size_t)i];
        if (y) y[i] = ys[(size_t)i];
    }
    if ((x || y) && np > cap)
        return fail(MF_ERR_INVALID, "mf_io_read_graph: %s has %d points", name, np);
    return MF_OK;
}
