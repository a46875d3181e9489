/*
 * Hdf5.cpp -- the HDF5 subset of Hdf5.hpp, from the HDF5 file format
 * specification (superblock v0/v1, v1 object headers, symbol-table groups).
 */
#include "Hdf5.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <fstream>

namespace h5 {

namespace {

constexpr uint64_t kUndef = ~0ull;
const uint8_t kSignature[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};

/* symbol-table node capacity of written files: 2 * kLeafK entries per group */
constexpr int kLeafK = 16;
constexpr int kInternalK = 16;
constexpr uint64_t kBtreeSize = 24 + (2 * kInternalK + 1) * 8 + 2 * kInternalK * 8;
constexpr uint64_t kSnodSize = 8 + 2 * kLeafK * 40;
constexpr uint64_t kEntrySize = 40;

uint64_t align8(uint64_t v) { return (v + 7) & ~7ull; }

}  // namespace

/* ------------------------------------------------------------------ reader */
bool Reader::fail(const std::string& msg) {
    err_ = msg;
    return false;
}

bool Reader::range(uint64_t addr, uint64_t len) const { return addr <= buf_.size() && len <= buf_.size() - addr; }

uint64_t Reader::u(uint64_t addr, int bytes) const {
    uint64_t v = 0;
    for (int b = 0; b < bytes; ++b) v |= (uint64_t)buf_[addr + b] << (8 * b);
    return v;
}

bool Reader::open(const std::string& path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) return fail("cannot open " + path);
    const std::streamoff size = f.tellg();
    if (size <= 0) return fail(path + " is empty");
    buf_.resize((size_t)size);
    f.seekg(0);
    if (!f.read((char*)buf_.data(), size)) return fail("cannot read " + path);
    /* the superblock sits at 0, 512, 1024, 2048, ... (after a user block) */
    uint64_t sb = kUndef;
    for (uint64_t at = 0; at + 8 <= buf_.size(); at = at ? at * 2 : 512)
        if (std::memcmp(buf_.data() + at, kSignature, 8) == 0) {
            sb = at;
            break;
        }
    if (sb == kUndef) return fail(path + " is not an HDF5 file (no superblock signature)");
    if (!range(sb, 96)) return fail("truncated superblock");
    const int version = buf_[sb + 8];
    if (version > 1) return fail("HDF5 superblock version " + std::to_string(version) + " is not supported (0 or 1)");
    if (buf_[sb + 13] != 8 || buf_[sb + 14] != 8) return fail("only 8-byte offsets and lengths are supported");
    const uint64_t addrs = sb + (version == 0 ? 24 : 28);
    if (!range(addrs, 32 + kEntrySize)) return fail("truncated superblock");
    base_ = u(addrs, 8);
    const uint64_t root_entry = addrs + 32;
    root_ = base_ + u(root_entry + 8, 8);
    return true;
}

bool Reader::parseMessage(int type, const uint8_t* p, uint64_t len, Header& h,
                          std::vector<std::pair<uint64_t, uint64_t>>& more) {
    auto rd = [&](uint64_t off, int bytes) {
        uint64_t v = 0;
        for (int b = 0; b < bytes; ++b) v |= (uint64_t)p[off + b] << (8 * b);
        return v;
    };
    switch (type) {
    case 0x0001: { /* dataspace */
        if (len < 4) return fail("short dataspace message");
        const int version = p[0], rank = p[1];
        const uint64_t at = version == 1 ? 8 : 4;
        if (version != 1 && version != 2) return fail("dataspace message version " + std::to_string(version));
        if (at + 8ull * rank > len) return fail("short dataspace message");
        h.dims.clear();
        for (int d = 0; d < rank; ++d) h.dims.push_back(rd(at + 8ull * d, 8));
        if (version == 2 && p[3] == 2) h.dims.assign(1, 0); /* null dataspace */
        h.have_space = true;
        break;
    }
    case 0x0003: { /* datatype */
        if (len < 8) return fail("short datatype message");
        h.type_class = p[0] & 0x0F;
        const uint32_t bits = (uint32_t)p[1] | ((uint32_t)p[2] << 8) | ((uint32_t)p[3] << 16);
        h.type_size = (int)rd(4, 4);
        h.byte_order = (int)(bits & 1u);
        if (h.type_class == 0) {
            h.is_signed = (bits >> 3) & 1u;
        } else if (h.type_class == 1) {
            if (bits & 0x40u) return fail("VAX floating point is not supported");
            if (len < 20) return fail("short floating-point datatype message");
            const int prec = (int)rd(10, 2), eloc = p[12], esize = p[13], mloc = p[14], msize = p[15];
            const uint64_t bias = rd(16, 4);
            const bool ieee32 = h.type_size == 4 && prec == 32 && eloc == 23 && esize == 8 && mloc == 0 &&
                                msize == 23 && bias == 127;
            const bool ieee64 = h.type_size == 8 && prec == 64 && eloc == 52 && esize == 11 && mloc == 0 &&
                                msize == 52 && bias == 1023;
            if (!ieee32 && !ieee64) return fail("only IEEE binary32 / binary64 floats are supported");
        }
        h.have_type = true;
        break;
    }
    case 0x0007:
        h.external = true;
        break;
    case 0x0008: { /* data layout */
        const int version = p[0];
        if (version == 1 || version == 2) {
            const int rank = p[1], cls = p[2];
            h.layout_class = cls;
            if (cls == 1) {
                h.data_addr = rd(8, 8);
                h.data_size = kUndef; /* from the dataspace */
            } else if (cls == 0) {
                const uint64_t at = 8 + 4ull * rank;
                if (at + 4 > len) return fail("short layout message");
                const uint64_t n = rd(at, 4);
                if (at + 4 + n > len) return fail("short layout message");
                h.compact.assign(p + at + 4, p + at + 4 + n);
            }
        } else if (version == 3 || version == 4) {
            const int cls = p[1];
            h.layout_class = cls;
            if (cls == 0) {
                const uint64_t n = rd(2, 2);
                if (4 + n > len) return fail("short layout message");
                h.compact.assign(p + 4, p + 4 + n);
            } else if (cls == 1) {
                if (len < 18) return fail("short layout message");
                h.data_addr = rd(2, 8);
                h.data_size = rd(10, 8);
            }
        } else {
            return fail("layout message version " + std::to_string(version));
        }
        h.have_layout = true;
        break;
    }
    case 0x000B:
        h.filtered = true;
        break;
    case 0x0010: /* continuation */
        if (len < 16) return fail("short continuation message");
        more.emplace_back(base_ + rd(0, 8), rd(8, 8));
        break;
    case 0x0011: /* symbol table */
        if (len < 16) return fail("short symbol table message");
        h.btree = rd(0, 8);
        h.heap = rd(8, 8);
        h.group = true;
        break;
    default:
        break;
    }
    return true;
}

bool Reader::parseHeader(uint64_t addr, Header& h) {
    if (!range(addr, 16)) return fail("object header outside the file");
    if (std::memcmp(buf_.data() + addr, "OHDR", 4) == 0) return fail("version 2 object headers are not supported");
    if (buf_[addr] != 1) return fail("object header version " + std::to_string(buf_[addr]));
    const uint64_t nmsgs = u(addr + 2, 2);
    std::vector<std::pair<uint64_t, uint64_t>> chunks{{addr + 16, u(addr + 8, 4)}};
    uint64_t seen = 0;
    for (size_t c = 0; c < chunks.size() && seen < nmsgs; ++c) {
        if (c > 4096) return fail("too many object header continuation blocks");
        const uint64_t start = chunks[c].first, len = chunks[c].second;
        if (!range(start, len)) return fail("object header block outside the file");
        uint64_t p = start;
        while (p + 8 <= start + len && seen < nmsgs) {
            const int type = (int)u(p, 2);
            const uint64_t msize = u(p + 2, 2);
            const int flags = buf_[p + 4];
            if (p + 8 + msize > start + len) return fail("object header message overruns its block");
            if ((flags & 0x02) && (type == 0x0001 || type == 0x0003 || type == 0x0008 || type == 0x0011))
                return fail("shared object header messages are not supported");
            if (!parseMessage(type, buf_.data() + p + 8, msize, h, chunks)) return false;
            p += 8 + msize;
            ++seen;
        }
    }
    return true;
}

bool Reader::btreeEntries(uint64_t node, uint64_t heap_data, uint64_t heap_size,
                          std::map<std::string, uint64_t>& out, int depth) {
    if (depth > 64) return fail("group B-tree too deep");
    if (budget_ == 0) return fail("group B-tree loops or is too large");
    --budget_;
    if (!range(node, 24) || std::memcmp(buf_.data() + node, "TREE", 4) != 0) return fail("bad group B-tree node");
    if (buf_[node + 4] != 0) return fail("not a group B-tree node");
    const int level = buf_[node + 5];
    const uint64_t used = u(node + 6, 2);
    if (!range(node + 24, 16 * used + 8)) return fail("group B-tree node outside the file");
    for (uint64_t i = 0; i < used; ++i) {
        const uint64_t child = base_ + u(node + 32 + 16 * i, 8);
        if (level > 0) {
            if (!btreeEntries(child, heap_data, heap_size, out, depth + 1)) return false;
            continue;
        }
        if (!range(child, 8) || std::memcmp(buf_.data() + child, "SNOD", 4) != 0) return fail("bad symbol table node");
        const uint64_t nsym = u(child + 6, 2);
        if (!range(child + 8, nsym * kEntrySize)) return fail("symbol table node outside the file");
        if (nsym > budget_) return fail("group B-tree loops or is too large");
        budget_ -= nsym;
        for (uint64_t s = 0; s < nsym; ++s) {
            const uint64_t e = child + 8 + s * kEntrySize;
            const uint64_t name_off = u(e, 8);
            if (name_off >= heap_size) return fail("link name outside the local heap");
            const char* name = (const char*)buf_.data() + heap_data + name_off;
            const size_t max_len = (size_t)(heap_size - name_off);
            out[std::string(name, strnlen(name, max_len))] = base_ + u(e + 8, 8);
        }
    }
    return true;
}

bool Reader::groupEntries(const Header& g, std::map<std::string, uint64_t>& out) {
    if (!g.group) return fail("not a symbol-table group (new-style groups are not supported)");
    const uint64_t heap = base_ + g.heap;
    if (!range(heap, 32) || std::memcmp(buf_.data() + heap, "HEAP", 4) != 0) return fail("bad local heap");
    const uint64_t size = u(heap + 8, 8), data = base_ + u(heap + 24, 8);
    if (!range(data, size)) return fail("local heap outside the file");
    budget_ = 1u << 22;
    return btreeEntries(base_ + g.btree, data, size, out, 0);
}

bool Reader::locate(const std::string& path, Header& h) {
    h = Header();
    if (!parseHeader(root_, h)) return false;
    size_t at = 0;
    while (at < path.size()) {
        size_t end = path.find('/', at);
        if (end == std::string::npos) end = path.size();
        const std::string name = path.substr(at, end - at);
        at = end + 1;
        if (name.empty()) continue;
        std::map<std::string, uint64_t> entries;
        if (!groupEntries(h, entries)) return false;
        auto it = entries.find(name);
        if (it == entries.end()) return fail("no object named " + name + " in " + path);
        h = Header();
        if (!parseHeader(it->second, h)) return false;
    }
    return true;
}

bool Reader::exists(const std::string& path) {
    Header h;
    return locate(path, h);
}

bool Reader::raw(const std::string& path, Header& h, const uint8_t*& data, uint64_t& count) {
    if (!locate(path, h)) return false;
    if (h.group) return fail(path + " is a group");
    if (!h.have_space || !h.have_type || !h.have_layout) return fail(path + ": incomplete dataset header");
    if (h.filtered) return fail(path + ": filtered (compressed) datasets are not supported");
    if (h.external) return fail(path + ": external storage is not supported");
    if (h.type_class != 0 && h.type_class != 1) return fail(path + ": only integer and float datasets are supported");
    if (h.byte_order != 0) return fail(path + ": big-endian data is not supported");
    if (h.type_size != 1 && h.type_size != 2 && h.type_size != 4 && h.type_size != 8)
        return fail(path + ": unsupported element size");
    count = 1;
    for (uint64_t d : h.dims) {
        if (d && count > (1ull << 40) / d) return fail(path + ": dataset too large");
        count *= d;
    }
    const uint64_t nbytes = count * (uint64_t)h.type_size;
    if (h.layout_class == 1) {
        if (h.data_addr == kUndef || nbytes == 0) {
            /* never written: the fill value, zero */
            if (count > (1ull << 31)) return fail(path + ": unwritten dataset too large");
            data = nullptr;
            return true;
        }
        if (h.data_size != kUndef && h.data_size < nbytes) return fail(path + ": storage smaller than the dataspace");
        const uint64_t at = base_ + h.data_addr;
        if (!range(at, nbytes)) return fail(path + ": data outside the file (truncated?)");
        data = buf_.data() + at;
        return true;
    }
    if (h.layout_class == 0) {
        if (h.compact.size() < nbytes) return fail(path + ": compact storage smaller than the dataspace");
        data = h.compact.data();
        return true;
    }
    return fail(path + ": chunked or virtual storage is not supported");
}

bool Reader::readFloat(const std::string& path, std::vector<float>& out, std::vector<uint64_t>& dims) {
    Header h;
    const uint8_t* p = nullptr;
    uint64_t n = 0;
    if (!raw(path, h, p, n)) return false;
    dims = h.dims;
    out.assign(n, 0.f);
    if (!p) return true;
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t* e = p + i * h.type_size;
        if (h.type_class == 1) {
            if (h.type_size == 4) {
                float v;
                std::memcpy(&v, e, 4);
                out[i] = v;
            } else {
                double v;
                std::memcpy(&v, e, 8);
                out[i] = (float)v; /* round to nearest, as H5Dread's double -> float */
            }
        } else {
            uint64_t raw = 0;
            std::memcpy(&raw, e, h.type_size);
            if (h.is_signed) {
                const int shift = 64 - 8 * h.type_size;
                out[i] = (float)((int64_t)(raw << shift) >> shift);
            } else {
                out[i] = (float)raw;
            }
        }
    }
    return true;
}

bool Reader::readInt32(const std::string& path, std::vector<int32_t>& out, std::vector<uint64_t>& dims) {
    Header h;
    const uint8_t* p = nullptr;
    uint64_t n = 0;
    if (!raw(path, h, p, n)) return false;
    if (h.type_class != 0) return fail(path + ": a float dataset read as integers is not supported");
    dims = h.dims;
    out.assign(n, 0);
    if (!p) return true;
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t raw = 0;
        std::memcpy(&raw, p + i * h.type_size, h.type_size);
        int64_t v;
        if (h.is_signed) {
            const int shift = 64 - 8 * h.type_size;
            v = (int64_t)(raw << shift) >> shift;
        } else {
            v = raw > (uint64_t)INT64_MAX ? INT64_MAX : (int64_t)raw;
        }
        out[i] = (int32_t)std::min<int64_t>(std::max<int64_t>(v, INT32_MIN), INT32_MAX); /* H5T clamps */
    }
    return true;
}

/* ------------------------------------------------------------------ writer */
void Writer::addFloat(const std::string& group, const std::string& name, const std::vector<uint64_t>& dims,
                      const float* data) {
    Dataset d;
    d.name = name;
    d.dims = dims;
    d.is_float = true;
    uint64_t n = 1;
    for (uint64_t x : dims) n *= x;
    d.bytes.resize(4 * n);
    if (n) std::memcpy(d.bytes.data(), data, 4 * n);
    groups_[group].push_back(std::move(d));
}

void Writer::addInt32(const std::string& group, const std::string& name, const std::vector<uint64_t>& dims,
                      const int32_t* data) {
    Dataset d;
    d.name = name;
    d.dims = dims;
    d.is_float = false;
    uint64_t n = 1;
    for (uint64_t x : dims) n *= x;
    d.bytes.resize(4 * n);
    if (n) std::memcpy(d.bytes.data(), data, 4 * n);
    groups_[group].push_back(std::move(d));
}

namespace {

struct Out {
    std::vector<uint8_t> b;
    void put(uint64_t at, uint64_t v, int bytes) {
        for (int i = 0; i < bytes; ++i) b[at + i] = (uint8_t)(v >> (8 * i));
    }
    void str(uint64_t at, const std::string& s) { std::memcpy(b.data() + at, s.data(), s.size()); }
};

/* heap data of one group: "" at 0, then the names, each padded to 8 */
struct Heap {
    std::vector<std::pair<std::string, uint64_t>> names; /* name, offset */
    uint64_t size = 8;
    uint64_t add(const std::string& s) {
        names.emplace_back(s, size);
        const uint64_t at = size;
        size += align8(s.size() + 1);
        return at;
    }
};

uint64_t dataspace_size(size_t rank) { return 8 + 16 * rank; }
uint64_t datatype_size(bool is_float) { return is_float ? 24 : 16; }
constexpr uint64_t kFillSize = 8, kLayoutSize = 24;
uint64_t dataset_header_size(size_t rank, bool is_float) {
    return 16 + 4 * 8 + dataspace_size(rank) + datatype_size(is_float) + kFillSize + kLayoutSize;
}

/* version 1 object header prefix */
void header_prefix(Out& o, uint64_t at, int nmsgs, uint64_t body) {
    o.put(at, 1, 1);
    o.put(at + 2, (uint64_t)nmsgs, 2);
    o.put(at + 4, 1, 4);
    o.put(at + 8, body, 4);
}

/* message header; returns the data offset */
uint64_t message(Out& o, uint64_t at, int type, uint64_t size, int flags) {
    o.put(at, (uint64_t)type, 2);
    o.put(at + 2, size, 2);
    o.put(at + 4, (uint64_t)flags, 1);
    return at + 8;
}

void group_header(Out& o, uint64_t at, uint64_t btree, uint64_t heap) {
    header_prefix(o, at, 1, 24);
    const uint64_t d = message(o, at + 16, 0x0011, 16, 0);
    o.put(d, btree, 8);
    o.put(d + 8, heap, 8);
}

void symbol_entry(Out& o, uint64_t at, uint64_t name_off, uint64_t header, bool group, uint64_t btree,
                  uint64_t heap) {
    o.put(at, name_off, 8);
    o.put(at + 8, header, 8);
    o.put(at + 16, group ? 1 : 0, 4);
    if (group) {
        o.put(at + 24, btree, 8);
        o.put(at + 32, heap, 8);
    }
}

/* one group: B-tree with one symbol table node holding every entry */
void group_tables(Out& o, uint64_t btree, uint64_t heap, uint64_t snod, const Heap& hp,
                  const std::vector<uint64_t>& headers, const std::vector<bool>& is_group,
                  const std::vector<std::pair<uint64_t, uint64_t>>& sub) {
    o.str(btree, "TREE");
    o.put(btree + 4, 0, 1);
    o.put(btree + 5, 0, 1);
    o.put(btree + 6, 1, 2);
    o.put(btree + 8, kUndef, 8);
    o.put(btree + 16, kUndef, 8);
    o.put(btree + 24, 0, 8);
    o.put(btree + 32, snod, 8);
    o.put(btree + 40, hp.names.back().second, 8);
    o.str(heap, "HEAP");
    o.put(heap + 8, hp.size, 8);
    o.put(heap + 16, 1, 8); /* no free block */
    o.put(heap + 24, heap + 32, 8);
    for (const auto& nm : hp.names) o.str(heap + 32 + nm.second, nm.first);
    o.str(snod, "SNOD");
    o.put(snod + 4, 1, 1);
    o.put(snod + 6, hp.names.size(), 2);
    for (size_t i = 0; i < hp.names.size(); ++i)
        symbol_entry(o, snod + 8 + kEntrySize * i, hp.names[i].second, headers[i], is_group[i], sub[i].first,
                     sub[i].second);
}

}  // namespace

bool Writer::write(const std::string& path, std::string& err) const {
    if (groups_.empty()) {
        err = "nothing to write";
        return false;
    }
    /* names in byte order, as the B-tree keys require */
    std::vector<std::string> gnames;
    for (const auto& g : groups_) gnames.push_back(g.first);
    std::sort(gnames.begin(), gnames.end());
    if (gnames.size() > 2 * kLeafK) {
        err = "too many groups";
        return false;
    }
    struct GroupLayout {
        uint64_t header, btree, heap, snod;
        Heap names;
        std::vector<const Dataset*> sets;
        std::vector<uint64_t> headers, data;
    };
    Heap root_names;
    for (const auto& g : gnames) {
        if (g.empty() || g.find('/') != std::string::npos) {
            err = "bad group name '" + g + "'";
            return false;
        }
        root_names.add(g);
    }
    uint64_t off = 96; /* superblock with the root entry */
    const uint64_t root_header = off;
    off += 40;
    const uint64_t root_btree = off;
    off += kBtreeSize;
    const uint64_t root_heap = off;
    off += 32 + root_names.size;
    const uint64_t root_snod = off;
    off += kSnodSize;
    std::vector<GroupLayout> gl(gnames.size());
    for (size_t g = 0; g < gnames.size(); ++g) {
        GroupLayout& L = gl[g];
        for (const Dataset& d : groups_.at(gnames[g])) L.sets.push_back(&d);
        std::sort(L.sets.begin(), L.sets.end(), [](const Dataset* a, const Dataset* b) { return a->name < b->name; });
        if (L.sets.size() > 2 * kLeafK) {
            err = "too many datasets in " + gnames[g];
            return false;
        }
        for (size_t i = 0; i + 1 < L.sets.size(); ++i)
            if (L.sets[i]->name == L.sets[i + 1]->name) {
                err = "duplicate dataset " + gnames[g] + "/" + L.sets[i]->name;
                return false;
            }
        for (const Dataset* d : L.sets) {
            if (d->name.empty() || d->name.find('/') != std::string::npos || d->dims.empty() || d->dims.size() > 2) {
                err = "bad dataset " + gnames[g] + "/" + d->name;
                return false;
            }
            L.names.add(d->name);
        }
        L.header = off;
        off += 40;
        L.btree = off;
        off += kBtreeSize;
        L.heap = off;
        off += 32 + L.names.size;
        L.snod = off;
        off += kSnodSize;
        for (const Dataset* d : L.sets) {
            L.headers.push_back(off);
            off += dataset_header_size(d->dims.size(), d->is_float);
        }
    }
    for (auto& L : gl)
        for (const Dataset* d : L.sets) {
            off = align8(off);
            L.data.push_back(d->bytes.empty() ? kUndef : off);
            off += d->bytes.size();
        }
    Out o;
    o.b.assign(off, 0);
    /* superblock version 0 */
    std::memcpy(o.b.data(), kSignature, 8);
    o.put(13, 8, 1);
    o.put(14, 8, 1);
    o.put(16, kLeafK, 2);
    o.put(18, kInternalK, 2);
    o.put(24, 0, 8);
    o.put(32, kUndef, 8);
    o.put(40, off, 8);
    o.put(48, kUndef, 8);
    symbol_entry(o, 56, 0, root_header, true, root_btree, root_heap);
    /* root group */
    group_header(o, root_header, root_btree, root_heap);
    {
        std::vector<uint64_t> headers;
        std::vector<bool> is_group;
        std::vector<std::pair<uint64_t, uint64_t>> sub;
        for (auto& L : gl) {
            headers.push_back(L.header);
            is_group.push_back(true);
            sub.emplace_back(L.btree, L.heap);
        }
        group_tables(o, root_btree, root_heap, root_snod, root_names, headers, is_group, sub);
    }
    for (auto& L : gl) {
        group_header(o, L.header, L.btree, L.heap);
        group_tables(o, L.btree, L.heap, L.snod, L.names, L.headers, std::vector<bool>(L.sets.size(), false),
                     std::vector<std::pair<uint64_t, uint64_t>>(L.sets.size(), {0, 0}));
        for (size_t i = 0; i < L.sets.size(); ++i) {
            const Dataset* d = L.sets[i];
            const uint64_t h = L.headers[i];
            const size_t rank = d->dims.size();
            header_prefix(o, h, 4, dataset_header_size(rank, d->is_float) - 16);
            uint64_t at = h + 16;
            /* dataspace v1 with maximum dimensions */
            uint64_t p = message(o, at, 0x0001, dataspace_size(rank), 0);
            o.put(p, 1, 1);
            o.put(p + 1, rank, 1);
            o.put(p + 2, 1, 1);
            for (size_t k = 0; k < rank; ++k) {
                o.put(p + 8 + 8 * k, d->dims[k], 8);
                o.put(p + 8 + 8 * (rank + k), d->dims[k], 8);
            }
            at = p + dataspace_size(rank);
            /* datatype v1: IEEE binary32 little-endian, or int32 little-endian signed */
            p = message(o, at, 0x0003, datatype_size(d->is_float), 1);
            if (d->is_float) {
                o.put(p, 0x11, 1);
                o.put(p + 1, 0x20, 1);
                o.put(p + 2, 31, 1);
                o.put(p + 4, 4, 4);
                o.put(p + 8, 0, 2);
                o.put(p + 10, 32, 2);
                o.put(p + 12, 23, 1);
                o.put(p + 13, 8, 1);
                o.put(p + 14, 0, 1);
                o.put(p + 15, 23, 1);
                o.put(p + 16, 127, 4);
            } else {
                o.put(p, 0x10, 1);
                o.put(p + 1, 0x08, 1);
                o.put(p + 4, 4, 4);
                o.put(p + 8, 0, 2);
                o.put(p + 10, 32, 2);
            }
            at = p + datatype_size(d->is_float);
            /* fill value v2: allocated late, written if set, undefined */
            p = message(o, at, 0x0005, kFillSize, 1);
            o.put(p, 2, 1);
            o.put(p + 1, 2, 1);
            o.put(p + 2, 2, 1);
            o.put(p + 3, 0, 1);
            at = p + kFillSize;
            /* layout v3, contiguous */
            p = message(o, at, 0x0008, kLayoutSize, 0);
            o.put(p, 3, 1);
            o.put(p + 1, 1, 1);
            o.put(p + 2, L.data[i], 8);
            o.put(p + 10, d->bytes.size(), 8);
            if (!d->bytes.empty()) std::memcpy(o.b.data() + L.data[i], d->bytes.data(), d->bytes.size());
        }
    }
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) {
        err = "cannot create " + path;
        return false;
    }
    f.write((const char*)o.b.data(), (std::streamsize)o.b.size());
    if (!f) {
        err = "cannot write " + path;
        return false;
    }
    return true;
}

}  // namespace h5
