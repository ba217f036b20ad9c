// mfm_chombo.cpp -- reader of Chombo HDF5 checkpoint and plot files
// (the files AMR_MHDAM::setupForRestart reads, srcIdeal/AMR_MHDAM.cpp:242,
// 584-747, and AMRLevelIdealMHD::readCheckpointHeader / readCheckpointLevel,
// srcIdeal/AMRLevelIdealMHDIO.cpp:36-200), so that a level's boxes and FABs
// can be handed to the device update unchanged.
//
// The HDF5 library is not linked: this file reads the subset of the HDF5
// file format (version 0 / 2 / 3 superblocks, version 1 and 2 object
// headers, symbol-table and compact link-message groups, fixed-point,
// floating-point, string and compound datatypes, compact and contiguous
// storage) that Chombo's writer produces with the library's default
// settings.  Chunked or filtered datasets, variable-length types and dense
// (fractal-heap) groups are refused with MFM_ERR_UNSUPPORTED.
//
// Chombo's layout (AMRIO / HDF5HeaderData of Chombo 3.2.4, not vendored in
// the reference): the root group holds "num_levels", "num_components",
// "component_<c>", "time", "iteration" and the group "Chombo_global" with
// "SpaceDim"; each level is a group "level_<l>" with the attributes "dx",
// "dt", "ref_ratio" and "prob_domain" (compound lo_i.. hi_k), the dataset
// "boxes" (compound of the same six integers), "data:offsets=0" (first
// value of each box in "data:datatype=0", one more entry than boxes) and
// the group "data_attributes" with "comps" and "outputGhost" (compound
// intvecti..k).  A box's values are its box grown by outputGhost,
// component-major with x fastest: the FAB layout of the device API.

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mfm.h"

namespace {

struct H5Error : std::runtime_error {
  int code;
  H5Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void bad(const std::string& m) { throw H5Error(MFM_ERR_ARG, m); }
[[noreturn]] void unsupported(const std::string& m) { throw H5Error(MFM_ERR_UNSUPPORTED, m); }

const uint64_t kUndef = ~0ull;

struct Datatype {
  int cls = -1;          // 0 fixed, 1 float, 3 string, 6 compound
  size_t size = 0;
  bool sign = false, big_endian = false;
  struct Member {
    std::string name;
    size_t offset;
    std::shared_ptr<Datatype> type;
  };
  std::vector<Member> members;
};

struct Attribute {
  std::string name;
  Datatype type;
  std::vector<uint64_t> dims;
  std::vector<uint8_t> data;
};

struct Object {
  bool symtab = false;
  uint64_t btree = kUndef, heap = kUndef;
  std::vector<std::pair<std::string, uint64_t>> links;  // link messages
  bool has_type = false, has_space = false, has_layout = false;
  Datatype type;
  std::vector<uint64_t> dims;
  int layout = -1;  // 0 compact, 1 contiguous, 2 chunked
  uint64_t addr = kUndef, bytes = 0;
  std::vector<uint8_t> compact;
  std::vector<Attribute> attrs;
};

// A read-only view of the file with bounds-checked little-endian reads.
struct File {
  const uint8_t* p = nullptr;
  size_t n = 0;
  int fd = -1;
  uint64_t base = 0;     // absolute offset of the superblock
  int so = 8, sl = 8;    // sizes of offsets and lengths
  uint64_t root = kUndef;
  std::map<uint64_t, Object> cache;

  ~File() {
    if (p) munmap(const_cast<uint8_t*>(p), n);
    if (fd >= 0) close(fd);
  }
  const uint8_t* at(uint64_t off, uint64_t len) const {
    if (off > n || len > n - off) bad("read beyond the end of the file");
    return p + off;
  }
  uint64_t u(uint64_t off, int len) const {
    const uint8_t* q = at(off, len);
    uint64_t v = 0;
    for (int i = len - 1; i >= 0; --i) v = (v << 8) | q[i];
    return v;
  }
  // file address (relative to the base) -> absolute offset
  uint64_t abs(uint64_t a) const {
    if (a == kUndef || (so < 8 && a == ((1ull << (8 * so)) - 1))) bad("undefined address");
    return base + a;
  }
  uint64_t addr(uint64_t off) const {
    uint64_t a = u(off, so);
    if (so < 8 && a == ((1ull << (8 * so)) - 1)) return kUndef;
    return a;
  }
  std::string cstr(uint64_t off, uint64_t maxlen) const {
    const uint8_t* q = at(off, 1);
    uint64_t lim = std::min<uint64_t>(maxlen, n - off), k = 0;
    while (k < lim && q[k]) ++k;
    return std::string(reinterpret_cast<const char*>(q), k);
  }

  void open(const char* path) {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) throw H5Error(MFM_ERR_ARG, std::string("cannot open ") + path);
    struct stat st;
    if (fstat(fd, &st) != 0 || st.st_size <= 0) bad("cannot stat the file");
    n = (size_t)st.st_size;
    void* m = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) bad("cannot map the file");
    p = static_cast<const uint8_t*>(m);
    superblock();
  }

  // Superblock: at 0 or after a user block of 512 * 2^k bytes.
  void superblock() {
    static const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    uint64_t off = 0;
    for (;;) {
      if (off + 8 <= n && std::memcmp(p + off, sig, 8) == 0) break;
      off = off == 0 ? 512 : off * 2;
      if (off >= n) bad("no HDF5 superblock");
    }
    int ver = (int)u(off + 8, 1);
    if (ver == 0 || ver == 1) {
      so = (int)u(off + 13, 1);
      sl = (int)u(off + 14, 1);
      if (so < 2 || so > 8 || sl < 2 || sl > 8) bad("bad offset/length sizes");
      uint64_t q = off + 24 + (ver == 1 ? 4 : 0);
      base = u(q, so);
      q += 4 * so;  // base, free-space, end of file, driver info
      // root group symbol table entry: name offset, object header address
      root = addr(q + so);
    } else if (ver == 2 || ver == 3) {
      so = (int)u(off + 9, 1);
      sl = (int)u(off + 10, 1);
      if (so < 2 || so > 8 || sl < 2 || sl > 8) bad("bad offset/length sizes");
      uint64_t q = off + 12;
      base = u(q, so);
      root = addr(q + 3 * so);  // base, extension, end of file, root
    } else {
      unsupported("superblock version " + std::to_string(ver));
    }
    if (base == 0 && off != 0) base = off;  // a user block counts as base
  }

  // ------------------------------------------------------------ datatypes --
  size_t datatype(uint64_t q, uint64_t end, Datatype& t, int depth = 0) const {
    if (depth > 16) bad("datatype nested too deeply");
    uint8_t b0 = *at(q, 1);
    int cls = b0 & 15, ver = b0 >> 4;
    uint32_t bits = (uint32_t)u(q + 1, 3);
    t.cls = cls;
    t.size = (size_t)u(q + 4, 4);
    t.big_endian = bits & 1;
    uint64_t r = q + 8;
    switch (cls) {
      case 0:  // fixed point
        t.sign = (bits >> 3) & 1;
        return 8 + 4;
      case 1:  // floating point
        if (t.size != 4 && t.size != 8) unsupported("float of size " + std::to_string(t.size));
        return 8 + 12;
      case 3:  // string
        return 8;
      case 6: {  // compound
        int nm = (int)(bits & 0xffff);
        for (int m = 0; m < nm; ++m) {
          Datatype::Member mem;
          mem.name = cstr(r, end - r);
          if (ver == 3) {
            r += mem.name.size() + 1;
            int ob = t.size < 256 ? 1 : (t.size < 65536 ? 2 : (t.size < (1ull << 24) ? 3 : 4));
            mem.offset = (size_t)u(r, ob);
            r += ob;
          } else {
            r += ((mem.name.size() + 1 + 7) / 8) * 8;
            mem.offset = (size_t)u(r, 4);
            r += 4;
            if (ver == 1) r += 1 + 3 + 4 + 4 + 16;  // array form, unused
          }
          mem.type = std::make_shared<Datatype>();
          r += datatype(r, end, *mem.type, depth + 1);
          if (mem.offset + mem.type->size > t.size) bad("compound member beyond its type");
          t.members.push_back(mem);
        }
        return (size_t)(r - q);
      }
      default:
        unsupported("datatype class " + std::to_string(cls));
    }
  }

  std::vector<uint64_t> dataspace(uint64_t q) const {
    int ver = (int)u(q, 1), rank = (int)u(q + 1, 1), flags = (int)u(q + 2, 1);
    std::vector<uint64_t> d;
    uint64_t r;
    if (ver == 1) {
      r = q + 8;
    } else if (ver == 2) {
      int type = (int)u(q + 3, 1);
      if (type == 2) return d;  // null dataspace
      r = q + 4;
    } else {
      unsupported("dataspace version " + std::to_string(ver));
    }
    (void)flags;
    for (int i = 0; i < rank; ++i) d.push_back(u(r + (uint64_t)i * sl, sl));
    return d;
  }

  // ------------------------------------------------------------- messages --
  void message(Object& o, int type, uint64_t q, uint64_t len,
               std::vector<std::pair<uint64_t, uint64_t>>& cont) {
    uint64_t end = q + len;
    switch (type) {
      case 0x01:
        o.dims = dataspace(q);
        o.has_space = true;
        break;
      case 0x03:
        datatype(q, end, o.type);
        o.has_type = true;
        break;
      case 0x06: {  // link message (compact new-style groups)
        int flags = (int)u(q + 1, 1);
        uint64_t r = q + 2;
        int ltype = 0;
        if (flags & 8) ltype = (int)u(r++, 1);
        if (flags & 4) r += 8;
        if (flags & 16) r += 1;
        int lsize = 1 << (flags & 3);
        uint64_t nlen = u(r, lsize);
        r += lsize;
        std::string name(reinterpret_cast<const char*>(at(r, nlen)), nlen);
        r += nlen;
        if (ltype == 0) o.links.push_back({name, addr(r)});
        break;
      }
      case 0x08: {
        int ver = (int)u(q, 1);
        if (ver == 3) {
          o.layout = (int)u(q + 1, 1);
          if (o.layout == 0) {
            uint64_t sz = u(q + 2, 2);
            const uint8_t* d = at(q + 4, sz);
            o.compact.assign(d, d + sz);
          } else if (o.layout == 1) {
            o.addr = addr(q + 2);
            o.bytes = u(q + 2 + so, sl);
          }
        } else if (ver == 1 || ver == 2) {
          int rank = (int)u(q + 1, 1);
          o.layout = (int)u(q + 2, 1);
          uint64_t r = q + 8;
          if (o.layout != 0) { o.addr = addr(r); r += so; }
          uint64_t cnt = 1;
          for (int i = 0; i < rank; ++i) cnt *= u(r + 4 * (uint64_t)i, 4);
          r += 4 * (uint64_t)rank;
          if (o.layout == 0) {
            uint64_t sz = u(r, 4);
            const uint8_t* d = at(r + 4, sz);
            o.compact.assign(d, d + sz);
          } else {
            o.bytes = cnt;  // element count for version 1/2; bytes set below
          }
        } else {
          unsupported("data layout version " + std::to_string(ver));
        }
        o.has_layout = true;
        break;
      }
      case 0x0C: {  // attribute
        int ver = (int)u(q, 1);
        uint64_t nsz = u(q + 2, 2), tsz = u(q + 4, 2), ssz = u(q + 6, 2);
        uint64_t r = q + 8 + (ver == 3 ? 1 : 0);
        auto pad = [&](uint64_t x) { return ver == 1 ? ((x + 7) / 8) * 8 : x; };
        Attribute a;
        a.name = cstr(r, nsz);
        r += pad(nsz);
        datatype(r, r + tsz, a.type);
        r += pad(tsz);
        a.dims = dataspace(r);
        r += pad(ssz);
        uint64_t cnt = 1;
        for (uint64_t d : a.dims) cnt *= d;
        uint64_t bytes = cnt * a.type.size;
        const uint8_t* d = at(r, bytes);
        a.data.assign(d, d + bytes);
        o.attrs.push_back(std::move(a));
        break;
      }
      case 0x10:
        if (cont.size() >= 4096) bad("too many header continuation blocks");
        cont.push_back({addr(q), u(q + so, sl)});
        break;
      case 0x11:
        o.symtab = true;
        o.btree = addr(q);
        o.heap = addr(q + so);
        break;
      case 0x02: {  // link info: dense storage when its heap address is set
        int flags = (int)u(q + 1, 1);
        uint64_t r = q + 2 + ((flags & 1) ? 8 : 0);
        if (addr(r) != kUndef) unsupported("dense (fractal heap) group storage");
        break;
      }
      case 0x0B:
        unsupported("filtered (compressed) dataset");
      default:
        break;
    }
  }

  const Object& object(uint64_t a) {
    auto it = cache.find(a);
    if (it != cache.end()) return it->second;
    Object o;
    std::vector<std::pair<uint64_t, uint64_t>> cont;
    uint64_t q = abs(a);
    if (n >= q + 4 && std::memcmp(at(q, 4), "OHDR", 4) == 0) {
      int flags = (int)u(q + 5, 1);
      uint64_t r = q + 6;
      if (flags & 32) r += 16;
      if (flags & 16) r += 4;
      int cs = 1 << (flags & 3);
      uint64_t size0 = u(r, cs);
      r += cs;
      auto walk = [&](uint64_t s, uint64_t e) {
        while (s + 4 <= e) {
          int type = (int)u(s, 1);
          uint64_t len = u(s + 1, 2);
          int mflags = (int)u(s + 3, 1);
          s += 4 + ((flags & 4) ? 2 : 0);
          if (s + len > e) bad("object header message overruns its chunk");
          message(o, type, s, len, cont);
          s += len;
          (void)mflags;
        }
      };
      walk(r, r + size0);
      for (size_t c = 0; c < cont.size(); ++c) {
        uint64_t s = abs(cont[c].first), len = cont[c].second;
        if (std::memcmp(at(s, 4), "OCHK", 4) != 0) bad("bad continuation block");
        walk(s + 4, s + len - 4);
      }
    } else {
      int ver = (int)u(q, 1);
      if (ver != 1) unsupported("object header version " + std::to_string(ver));
      int nmsg = (int)u(q + 2, 2);
      uint64_t hsize = u(q + 8, 4);
      auto walk = [&](uint64_t s, uint64_t e, int& left) {
        while (left > 0 && s + 8 <= e) {
          int type = (int)u(s, 2);
          uint64_t len = u(s + 2, 2);
          s += 8;
          if (s + len > e) bad("object header message overruns its block");
          message(o, type, s, len, cont);
          s += len;
          --left;
        }
      };
      int left = nmsg;
      walk(q + 16, q + 16 + hsize, left);
      for (size_t c = 0; c < cont.size() && left > 0; ++c)
        walk(abs(cont[c].first), abs(cont[c].first) + cont[c].second, left);
    }
    if (o.has_layout && o.layout == 1 && o.has_type) {
      uint64_t cnt = 1;
      for (uint64_t d : o.dims) cnt *= d;
      o.bytes = cnt * o.type.size;
    }
    return cache.emplace(a, std::move(o)).first->second;
  }

  // ---------------------------------------------------------------- groups --
  void btree_members(uint64_t node, uint64_t heap_data,
                     std::vector<std::pair<std::string, uint64_t>>& out, int depth) {
    if (depth > 64) bad("group B-tree too deep");
    uint64_t q = abs(node);
    if (std::memcmp(at(q, 4), "TREE", 4) != 0) bad("bad group B-tree node");
    int type = (int)u(q + 4, 1), level = (int)u(q + 5, 1), used = (int)u(q + 6, 2);
    if (type != 0) bad("not a group B-tree");
    uint64_t r = q + 8 + 2 * (uint64_t)so + sl;  // after key 0
    for (int c = 0; c < used; ++c) {
      uint64_t child = addr(r);
      r += so + sl;
      if (level > 0) {
        btree_members(child, heap_data, out, depth + 1);
      } else {
        uint64_t s = abs(child);
        if (std::memcmp(at(s, 4), "SNOD", 4) != 0) bad("bad symbol table node");
        int nsym = (int)u(s + 6, 2);
        uint64_t e = s + 8;
        for (int k = 0; k < nsym; ++k, e += 2 * so + 24) {
          uint64_t noff = u(e, so);
          out.push_back({cstr(heap_data + noff, 1 << 20), addr(e + so)});
        }
      }
    }
  }

  std::vector<std::pair<std::string, uint64_t>> members(uint64_t obj) {
    const Object& o = object(obj);
    if (!o.symtab) return o.links;
    uint64_t h = abs(o.heap);
    if (std::memcmp(at(h, 4), "HEAP", 4) != 0) bad("bad local heap");
    uint64_t data = abs(addr(h + 8 + 2 * (uint64_t)sl));
    std::vector<std::pair<std::string, uint64_t>> out;
    btree_members(o.btree, data, out, 0);
    return out;
  }

  // "/a/b/c" -> object address
  uint64_t resolve(const std::string& path) {
    uint64_t cur = root;
    size_t i = 0;
    while (i < path.size()) {
      while (i < path.size() && path[i] == '/') ++i;
      if (i >= path.size()) break;
      size_t j = path.find('/', i);
      std::string part = path.substr(i, j == std::string::npos ? std::string::npos : j - i);
      bool found = false;
      for (auto& m : members(cur))
        if (m.first == part) { cur = m.second; found = true; break; }
      if (!found) throw H5Error(MFM_ERR_ARG, "no object " + path);
      if (j == std::string::npos) break;
      i = j;
    }
    return cur;
  }

  const Attribute* attribute(const Object& o, const std::string& name) const {
    for (auto& a : o.attrs)
      if (a.name == name) return &a;
    return nullptr;
  }

  // Raw bytes of a dataset (contiguous or compact).
  std::pair<const uint8_t*, uint64_t> raw(const Object& o) const {
    if (!o.has_layout) bad("not a dataset");
    if (o.layout == 0) return {o.compact.data(), (uint64_t)o.compact.size()};
    if (o.layout == 1) {
      if (o.addr == kUndef) return {nullptr, 0};  // never written
      return {at(abs(o.addr), o.bytes), o.bytes};
    }
    unsupported("chunked dataset");
  }
};

// ------------------------------------------------------- value conversion --
int64_t as_int(const uint8_t* q, const Datatype& t) {
  if (t.cls != 0 || t.big_endian || t.size > 8 || t.size == 0) unsupported("integer type");
  uint64_t v = 0;
  for (int i = (int)t.size - 1; i >= 0; --i) v = (v << 8) | q[i];
  if (t.sign && t.size < 8 && (v >> (8 * t.size - 1)) & 1) v |= ~0ull << (8 * t.size);
  return (int64_t)v;
}
double as_real(const uint8_t* q, const Datatype& t) {
  if (t.cls == 0) return (double)as_int(q, t);
  if (t.cls != 1 || t.big_endian) unsupported("real type");
  if (t.size == 8) { double d; std::memcpy(&d, q, 8); return d; }
  float f; std::memcpy(&f, q, 4); return f;
}

}  // namespace

// ============================================================ C entry ======
struct mfm_h5_file {
  File f;
  std::string err;
};

struct mfm_chombo_file {
  File f;
  std::string err;
  int num_levels = 0, num_components = 0, space_dim = 3;
  double time = 0.0;
  std::vector<std::string> names;
};

namespace {

template <class H, class Fn>
int guard(H* h, Fn fn) {
  if (!h) return MFM_ERR_ARG;
  try {
    fn();
    h->err.clear();
    return MFM_OK;
  } catch (const H5Error& e) {
    h->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    h->err = e.what();
    return MFM_ERR_ARG;
  }
}

const Datatype* member(const Datatype& t, const char* name) {
  for (auto& m : t.members)
    if (m.name == name) return m.type.get();
  return nullptr;
}
int member_int(const uint8_t* q, const Datatype& t, const char* name) {
  for (auto& m : t.members)
    if (m.name == name) return (int)as_int(q + m.offset, *m.type);
  throw H5Error(MFM_ERR_ARG, std::string("compound has no member ") + name);
}
bool has_member(const Datatype& t, const char* name) { return member(t, name) != nullptr; }

// Chombo box compound (lo_i, lo_j[, lo_k], hi_i, hi_j[, hi_k])
mfm_box box_of(const uint8_t* q, const Datatype& t) {
  // callers hold at least t.size bytes at q
  static const char* lo[3] = {"lo_i", "lo_j", "lo_k"};
  static const char* hi[3] = {"hi_i", "hi_j", "hi_k"};
  mfm_box b;
  for (int d = 0; d < 3; ++d) {
    if (has_member(t, lo[d])) {
      b.lo[d] = member_int(q, t, lo[d]);
      b.hi[d] = member_int(q, t, hi[d]);
    } else {
      b.lo[d] = b.hi[d] = 0;
    }
  }
  return b;
}

const Attribute& need_attr(File& f, const Object& o, const std::string& name) {
  const Attribute* a = f.attribute(o, name);
  if (!a) throw H5Error(MFM_ERR_ARG, "missing attribute " + name);
  return *a;
}
// First element of an attribute, checked to be present.
const uint8_t* first(const Attribute& a) {
  if (a.type.size == 0 || a.data.size() < a.type.size)
    throw H5Error(MFM_ERR_ARG, "attribute " + a.name + " holds no value");
  return a.data.data();
}
std::string attr_string(const Attribute& a) {
  if (a.type.cls != 3) throw H5Error(MFM_ERR_ARG, "attribute " + a.name + " is not a string");
  std::string s(reinterpret_cast<const char*>(a.data.data()), a.data.size());
  size_t z = s.find('\0');
  if (z != std::string::npos) s.resize(z);
  while (!s.empty() && s.back() == ' ') s.pop_back();
  return s;
}

}  // namespace

extern "C" {

int mfm_h5_open(const char* path, mfm_h5_file** out) {
  if (!path || !out) return MFM_ERR_ARG;
  *out = nullptr;
  auto* h = new mfm_h5_file;
  int rc = guard(h, [&] { h->f.open(path); });
  if (rc != MFM_OK) { delete h; return rc; }
  *out = h;
  return MFM_OK;
}

int mfm_h5_close(mfm_h5_file* h) {
  delete h;
  return MFM_OK;
}

const char* mfm_h5_last_error(const mfm_h5_file* h) { return h ? h->err.c_str() : "null handle"; }

int mfm_h5_dataset_info(mfm_h5_file* h, const char* path, int* rank, long* dims,
                        int* type_class, size_t* type_size) {
  return guard(h, [&] {
    const Object& o = h->f.object(h->f.resolve(path ? path : "/"));
    if (!o.has_layout) throw H5Error(MFM_ERR_ARG, "not a dataset");
    if (o.dims.size() > 8) unsupported("rank above 8");
    if (rank) *rank = (int)o.dims.size();
    if (dims) for (size_t i = 0; i < o.dims.size(); ++i) dims[i] = (long)o.dims[i];
    if (type_class) *type_class = o.type.cls;
    if (type_size) *type_size = o.type.size;
  });
}

int mfm_h5_dataset_read(mfm_h5_file* h, const char* path, void* dst, size_t bytes) {
  return guard(h, [&] {
    const Object& o = h->f.object(h->f.resolve(path ? path : "/"));
    auto r = h->f.raw(o);
    if (!dst || bytes != r.second) throw H5Error(MFM_ERR_SHAPE, "buffer size differs from the dataset");
    if (r.second) std::memcpy(dst, r.first, r.second);
  });
}

int mfm_h5_attribute_read(mfm_h5_file* h, const char* obj_path, const char* name, void* dst,
                          size_t capacity, int* type_class, size_t* bytes) {
  return guard(h, [&] {
    const Object& o = h->f.object(h->f.resolve(obj_path ? obj_path : "/"));
    const Attribute& a = need_attr(h->f, o, name ? name : "");
    if (type_class) *type_class = a.type.cls;
    if (bytes) *bytes = a.data.size();
    if (dst) {
      if (capacity < a.data.size()) throw H5Error(MFM_ERR_SHAPE, "attribute larger than the buffer");
      std::memcpy(dst, a.data.data(), a.data.size());
    }
  });
}

int mfm_h5_group_size(mfm_h5_file* h, const char* path, int* count) {
  return guard(h, [&] { *count = (int)h->f.members(h->f.resolve(path ? path : "/")).size(); });
}

int mfm_h5_group_name(mfm_h5_file* h, const char* path, int index, char* name, size_t capacity) {
  return guard(h, [&] {
    auto m = h->f.members(h->f.resolve(path ? path : "/"));
    if (index < 0 || index >= (int)m.size()) throw H5Error(MFM_ERR_ARG, "index out of range");
    if (!name || capacity < m[index].first.size() + 1)
      throw H5Error(MFM_ERR_SHAPE, "name buffer too small");
    std::memcpy(name, m[index].first.c_str(), m[index].first.size() + 1);
  });
}

// ------------------------------------------------------------- Chombo ----
int mfm_chombo_open(const char* path, mfm_chombo_file** out) {
  if (!path || !out) return MFM_ERR_ARG;
  *out = nullptr;
  auto* c = new mfm_chombo_file;
  int rc = guard(c, [&] {
    c->f.open(path);
    const Object& root = c->f.object(c->f.root);
    const Attribute& nl = need_attr(c->f, root, "num_levels");
    c->num_levels = (int)as_int(first(nl), nl.type);
    const Attribute& ncomp = need_attr(c->f, root, "num_components");
    c->num_components = (int)as_int(first(ncomp), ncomp.type);
    if (const Attribute* t = c->f.attribute(root, "time")) c->time = as_real(first(*t), t->type);
    for (int k = 0; k < c->num_components; ++k)
      c->names.push_back(attr_string(need_attr(c->f, root, "component_" + std::to_string(k))));
    try {
      const Object& g = c->f.object(c->f.resolve("/Chombo_global"));
      if (const Attribute* s = c->f.attribute(g, "SpaceDim"))
        c->space_dim = (int)as_int(first(*s), s->type);
    } catch (const H5Error&) {
    }
    if (c->num_levels < 1 || c->num_components < 1) throw H5Error(MFM_ERR_ARG, "empty file");
  });
  if (rc != MFM_OK) {
    delete c;
    return rc;
  }
  *out = c;
  return MFM_OK;
}

int mfm_chombo_close(mfm_chombo_file* c) {
  delete c;
  return MFM_OK;
}

const char* mfm_chombo_last_error(const mfm_chombo_file* c) { return c ? c->err.c_str() : "null handle"; }

int mfm_chombo_header(mfm_chombo_file* c, int* num_levels, int* num_components, int* space_dim,
                      double* time) {
  return guard(c, [&] {
    if (num_levels) *num_levels = c->num_levels;
    if (num_components) *num_components = c->num_components;
    if (space_dim) *space_dim = c->space_dim;
    if (time) *time = c->time;
  });
}

const char* mfm_chombo_component_name(const mfm_chombo_file* c, int comp) {
  if (!c || comp < 0 || comp >= (int)c->names.size()) return nullptr;
  return c->names[comp].c_str();
}

int mfm_chombo_level(mfm_chombo_file* c, int level, int* nboxes, double* dx, double* dt,
                     int* ref_ratio, mfm_box* domain, int ghost[3]) {
  return guard(c, [&] {
    if (level < 0 || level >= c->num_levels) throw H5Error(MFM_ERR_ARG, "level out of range");
    std::string g = "/level_" + std::to_string(level);
    const Object& lo = c->f.object(c->f.resolve(g));
    if (dx) { const Attribute& a = need_attr(c->f, lo, "dx"); *dx = as_real(first(a), a.type); }
    if (dt) {
      const Attribute* a = c->f.attribute(lo, "dt");
      *dt = a ? as_real(first(*a), a->type) : 0.0;
    }
    if (ref_ratio) {
      const Attribute* a = c->f.attribute(lo, "ref_ratio");
      *ref_ratio = a ? (int)as_int(first(*a), a->type) : 1;
    }
    if (domain) {
      const Attribute& a = need_attr(c->f, lo, "prob_domain");
      if (a.type.cls != 6) throw H5Error(MFM_ERR_ARG, "prob_domain is not a box");
      *domain = box_of(first(a), a.type);
    }
    const Object& bx = c->f.object(c->f.resolve(g + "/boxes"));
    if (bx.type.cls != 6 || bx.dims.size() != 1) throw H5Error(MFM_ERR_ARG, "boxes is not a box list");
    if (nboxes) *nboxes = (int)bx.dims[0];
    if (ghost) {
      ghost[0] = ghost[1] = ghost[2] = 0;
      try {
        const Object& da = c->f.object(c->f.resolve(g + "/data_attributes"));
        const Attribute* a = c->f.attribute(da, "outputGhost");
        if (!a) a = c->f.attribute(da, "ghost");
        if (a) {
          static const char* nm[3] = {"intvecti", "intvectj", "intvectk"};
          for (int d = 0; d < 3; ++d)
            if (has_member(a->type, nm[d])) ghost[d] = member_int(first(*a), a->type, nm[d]);
        }
      } catch (const H5Error& e) {
        if (e.code != MFM_ERR_ARG) throw;
      }
    }
  });
}

int mfm_chombo_boxes(mfm_chombo_file* c, int level, mfm_box* boxes, int capacity) {
  return guard(c, [&] {
    std::string g = "/level_" + std::to_string(level);
    const Object& bx = c->f.object(c->f.resolve(g + "/boxes"));
    if (bx.type.cls != 6 || bx.dims.size() != 1) throw H5Error(MFM_ERR_ARG, "boxes is not a box list");
    int nb = (int)bx.dims[0];
    if (!boxes || capacity < nb) throw H5Error(MFM_ERR_SHAPE, "box buffer too small");
    auto r = c->f.raw(bx);
    if ((uint64_t)nb * bx.type.size > r.second) throw H5Error(MFM_ERR_ARG, "short box list");
    for (int b = 0; b < nb; ++b) boxes[b] = box_of(r.first + (size_t)b * bx.type.size, bx.type);
  });
}

// Values of box b of a level: (box grown by the file's ghost) x num_components
// doubles, component-major with x fastest.
int mfm_chombo_read_box(mfm_chombo_file* c, int level, int b, double* dst, size_t count) {
  return guard(c, [&] {
    std::string g = "/level_" + std::to_string(level);
    const Object& off = c->f.object(c->f.resolve(g + "/data:offsets=0"));
    const Object& dat = c->f.object(c->f.resolve(g + "/data:datatype=0"));
    if (off.dims.size() != 1 || dat.dims.size() != 1) throw H5Error(MFM_ERR_ARG, "bad data layout");
    if (b < 0 || (uint64_t)b + 1 >= off.dims[0]) throw H5Error(MFM_ERR_ARG, "box out of range");
    auto ro = c->f.raw(off);
    if ((uint64_t)(b + 2) * off.type.size > ro.second) throw H5Error(MFM_ERR_ARG, "short offsets");
    int64_t o0 = as_int(ro.first + (size_t)b * off.type.size, off.type);
    int64_t o1 = as_int(ro.first + (size_t)(b + 1) * off.type.size, off.type);
    if (o0 < 0 || o1 < o0 || (uint64_t)o1 > dat.dims[0]) throw H5Error(MFM_ERR_ARG, "bad offsets");
    if (!dst || count != (size_t)(o1 - o0)) throw H5Error(MFM_ERR_SHAPE, "value count differs from the box");
    auto rd = c->f.raw(dat);
    if ((uint64_t)o1 * dat.type.size > rd.second) throw H5Error(MFM_ERR_ARG, "short data");
    const uint8_t* q = rd.first + (size_t)o0 * dat.type.size;
    if (dat.type.cls == 1 && dat.type.size == 8 && !dat.type.big_endian) {
      std::memcpy(dst, q, count * 8);
    } else {
      for (size_t k = 0; k < count; ++k) dst[k] = as_real(q + k * dat.type.size, dat.type);
    }
  });
}

}  // extern "C"
