// gs_checkpoint.cpp -- reads the reference's HDF5 checkpoints without an HDF5
// library: the subset of the file format that its FeVariable_SaveToFile,
// SwarmDump and timeInfo writers produce (superblock version 0 or 1, version-1
// object headers, symbol-table groups, contiguous or compact little- or
// big-endian integer and floating-point datasets, root-group attributes).
// Anything outside that subset (chunked or filtered storage, version-2 object
// headers, other datatypes) is refused with GS_H5_UNSUPPORTED, never guessed.
// Every offset read from the file is bounds-checked against its size.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gale_checkpoint.h"

namespace {

struct File {
  std::vector<uint8_t> b;
  uint64_t root = 0;  // object header address of the root group

  bool in(uint64_t off, uint64_t len) const { return off <= b.size() && len <= b.size() - off; }
  uint64_t u(uint64_t off, int n) const {  // little-endian unsigned of n bytes (caller checked bounds)
    uint64_t v = 0;
    for (int i = n - 1; i >= 0; i--) v = (v << 8) | b[off + i];
    return v;
  }
};

struct Msg {
  uint16_t type;
  uint64_t off, size;
};

struct Dataset {
  int rank = -1;
  int64_t dims[GS_H5_MAX_RANK] = {0, 0, 0, 0};
  int type = 0;          // GS_H5_*
  int elem = 0;          // bytes per element
  bool big_endian = false;
  uint64_t addr = 0;     // first byte of the raw data
  uint64_t bytes = 0;
  bool have_layout = false;
};

int load(const char* path, File& f) {
  if (!path) return GS_H5_INVALID;
  FILE* fp = std::fopen(path, "rb");
  if (!fp) return GS_H5_IO;
  std::fseek(fp, 0, SEEK_END);
  const long n = std::ftell(fp);
  std::fseek(fp, 0, SEEK_SET);
  if (n < 0) {
    std::fclose(fp);
    return GS_H5_IO;
  }
  f.b.resize((size_t)n);
  const size_t got = n ? std::fread(f.b.data(), 1, (size_t)n, fp) : 0;
  std::fclose(fp);
  if (got != (size_t)n) return GS_H5_IO;
  static const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
  if (!f.in(0, 100) || std::memcmp(f.b.data(), sig, 8) != 0) return GS_H5_FORMAT;
  const int sb = f.b[8];
  if (sb != 0 && sb != 1) return GS_H5_UNSUPPORTED;
  if (f.b[13] != 8 || f.b[14] != 8) return GS_H5_UNSUPPORTED;  // offsets and lengths of 8 bytes
  // v0: 16 bytes of versions/sizes, 4 of group K's, 4 of flags, 4 addresses; v1 adds 4 bytes
  const uint64_t entry = (sb == 0 ? 24 : 28) + 32;
  if (!f.in(entry, 40)) return GS_H5_FORMAT;
  f.root = f.u(entry + 8, 8);
  return 0;
}

// Messages of a version-1 object header, following continuation blocks.
int messages(const File& f, uint64_t addr, std::vector<Msg>& out) {
  if (!f.in(addr, 16)) return GS_H5_FORMAT;
  if (f.b[addr] != 1) return GS_H5_UNSUPPORTED;
  std::vector<std::pair<uint64_t, uint64_t>> blocks{{addr + 16, f.u(addr + 8, 4)}};
  for (size_t k = 0; k < blocks.size(); k++) {
    if (blocks.size() > 1024) return GS_H5_FORMAT;
    uint64_t o = blocks[k].first;
    const uint64_t end = o + blocks[k].second;
    if (!f.in(blocks[k].first, blocks[k].second)) return GS_H5_FORMAT;
    while (o + 8 <= end) {
      const uint16_t type = (uint16_t)f.u(o, 2);
      const uint64_t size = f.u(o + 2, 2);
      const uint64_t data = o + 8;
      if (data + size > end) return GS_H5_FORMAT;
      if (type == 0x10) {  // continuation: address and length of the next block
        if (size < 16) return GS_H5_FORMAT;
        blocks.push_back({f.u(data, 8), f.u(data + 8, 8)});
      } else if (type != 0) {
        out.push_back({type, data, size});
      }
      o = data + size;
    }
  }
  return 0;
}

// Object header address of `name` in the root group (symbol table, B-tree v1
// of group nodes, symbol-table nodes, names in the local heap).
int lookup(const File& f, const char* name, uint64_t& addr) {
  if (!name) return GS_H5_INVALID;
  while (*name == '/') name++;
  std::vector<Msg> ms;
  int st = messages(f, f.root, ms);
  if (st) return st;
  const Msg* stab = nullptr;
  for (const Msg& m : ms)
    if (m.type == 0x11 && m.size >= 16) stab = &m;
  if (!stab) return GS_H5_UNSUPPORTED;
  const uint64_t tree = f.u(stab->off, 8), heap = f.u(stab->off + 8, 8);
  if (!f.in(heap, 32) || std::memcmp(&f.b[heap], "HEAP", 4) != 0) return GS_H5_FORMAT;
  const uint64_t seg_size = f.u(heap + 8, 8), seg = f.u(heap + 24, 8);
  if (!f.in(seg, seg_size)) return GS_H5_FORMAT;
  std::vector<uint64_t> stack{tree};
  int visited = 0;
  while (!stack.empty()) {
    const uint64_t node = stack.back();
    stack.pop_back();
    if (++visited > 100000) return GS_H5_FORMAT;
    if (!f.in(node, 8)) return GS_H5_FORMAT;
    if (std::memcmp(&f.b[node], "TREE", 4) == 0) {
      if (!f.in(node, 24)) return GS_H5_FORMAT;
      const int level = f.b[node + 5];
      const uint64_t n = f.u(node + 6, 2);
      if (f.b[node + 4] != 0 || !f.in(node + 24, 8 + n * 16)) return GS_H5_FORMAT;
      for (uint64_t i = 0; i < n; i++) stack.push_back(f.u(node + 24 + 8 + i * 16, 8));
      (void)level;
    } else if (std::memcmp(&f.b[node], "SNOD", 4) == 0) {
      const uint64_t n = f.u(node + 6, 2);
      if (!f.in(node + 8, n * 40)) return GS_H5_FORMAT;
      for (uint64_t j = 0; j < n; j++) {
        const uint64_t e = node + 8 + j * 40;
        const uint64_t noff = f.u(e, 8);
        if (noff >= seg_size) return GS_H5_FORMAT;
        const char* s = (const char*)&f.b[seg + noff];
        const size_t maxlen = (size_t)(seg_size - noff);
        const size_t len = strnlen(s, maxlen);
        if (len == maxlen) return GS_H5_FORMAT;
        if (std::strcmp(s, name) == 0) {
          addr = f.u(e + 8, 8);
          return 0;
        }
      }
    } else {
      return GS_H5_FORMAT;
    }
  }
  return GS_H5_NOT_FOUND;
}

// Datatype message: fixed-point (class 0) or floating point (class 1) of 4 or 8 bytes.
int parse_type(const File& f, uint64_t o, uint64_t size, int& type, int& elem, bool& be) {
  if (size < 8) return GS_H5_FORMAT;
  const int cls = f.b[o] & 0x0f;
  const uint8_t bits = f.b[o + 1];
  elem = (int)f.u(o + 4, 4);
  be = (bits & 1) != 0;
  if (cls == 0) {
    const bool sgn = (bits & 8) != 0;
    if (!sgn) return GS_H5_UNSUPPORTED;
    if (elem == 4) type = GS_H5_I32;
    else if (elem == 8) type = GS_H5_I64;
    else return GS_H5_UNSUPPORTED;
  } else if (cls == 1) {
    if (bits & 0x40) return GS_H5_UNSUPPORTED;  // VAX order
    if (elem == 4) type = GS_H5_F32;
    else if (elem == 8) type = GS_H5_F64;
    else return GS_H5_UNSUPPORTED;
  } else {
    return GS_H5_UNSUPPORTED;
  }
  return 0;
}

// Dataspace message (versions 1 and 2): rank and current dimensions.
int parse_space(const File& f, uint64_t o, uint64_t size, int& rank, int64_t* dims) {
  if (size < 4) return GS_H5_FORMAT;
  const int ver = f.b[o];
  rank = f.b[o + 1];
  if (rank > GS_H5_MAX_RANK) return GS_H5_UNSUPPORTED;
  const uint64_t d0 = o + (ver == 1 ? 8 : 4);
  if (ver != 1 && ver != 2) return GS_H5_UNSUPPORTED;
  if (d0 + 8 * (uint64_t)rank > o + size) return GS_H5_FORMAT;
  for (int i = 0; i < rank; i++) dims[i] = (int64_t)f.u(d0 + 8 * i, 8);
  return 0;
}

int dataset(const File& f, const char* name, Dataset& d) {
  uint64_t addr = 0;
  int st = lookup(f, name, addr);
  if (st) return st;
  std::vector<Msg> ms;
  if ((st = messages(f, addr, ms))) return st;
  for (const Msg& m : ms) {
    if (m.type == 0x0001) {
      if ((st = parse_space(f, m.off, m.size, d.rank, d.dims))) return st;
    } else if (m.type == 0x0003) {
      if ((st = parse_type(f, m.off, m.size, d.type, d.elem, d.big_endian))) return st;
    } else if (m.type == 0x000B) {
      return GS_H5_UNSUPPORTED;  // filter pipeline
    } else if (m.type == 0x0008) {
      if (m.size < 1) return GS_H5_FORMAT;
      const int ver = f.b[m.off];
      if (ver == 1 || ver == 2) {
        if (m.size < 8) return GS_H5_FORMAT;
        const int nd = f.b[m.off + 1], cls = f.b[m.off + 2];
        if (cls == 1) {  // contiguous: address, then nd sizes of 4 bytes (the last the element size)
          if (m.size < 16 + 4 * (uint64_t)nd) return GS_H5_FORMAT;
          d.addr = f.u(m.off + 8, 8);
          uint64_t n = 1;
          for (int i = 0; i < nd; i++) n *= f.u(m.off + 16 + 4 * i, 4);
          d.bytes = n;
        } else if (cls == 0) {  // compact: sizes, then the data size and the data
          const uint64_t s = m.off + 8 + 4 * (uint64_t)nd;
          if (s + 4 > m.off + m.size) return GS_H5_FORMAT;
          d.bytes = f.u(s, 4);
          d.addr = s + 4;
        } else {
          return GS_H5_UNSUPPORTED;
        }
      } else if (ver == 3) {
        if (m.size < 2) return GS_H5_FORMAT;
        const int cls = f.b[m.off + 1];
        if (cls == 1) {
          if (m.size < 18) return GS_H5_FORMAT;
          d.addr = f.u(m.off + 2, 8);
          d.bytes = f.u(m.off + 10, 8);
        } else if (cls == 0) {
          if (m.size < 4) return GS_H5_FORMAT;
          d.bytes = f.u(m.off + 2, 2);
          d.addr = m.off + 4;
          if (d.bytes > m.size - 4) return GS_H5_FORMAT;
        } else {
          return GS_H5_UNSUPPORTED;
        }
      } else {
        return GS_H5_UNSUPPORTED;
      }
      d.have_layout = true;
    }
  }
  if (d.rank < 0 || !d.type || !d.have_layout) return GS_H5_FORMAT;
  uint64_t n = 1;
  for (int i = 0; i < d.rank; i++) {
    if (d.dims[i] < 0 || (uint64_t)d.dims[i] > (1ull << 40)) return GS_H5_FORMAT;
    n *= (uint64_t)d.dims[i];
    if (n > f.b.size()) return GS_H5_FORMAT;  // more elements than bytes in the file
  }
  if (d.bytes < n * (uint64_t)d.elem || !f.in(d.addr, n * (uint64_t)d.elem)) return GS_H5_FORMAT;
  d.bytes = n * (uint64_t)d.elem;
  return 0;
}

int64_t count_of(const Dataset& d) {
  int64_t n = 1;
  for (int i = 0; i < d.rank; i++) n *= d.dims[i];
  return n;
}

// Element k of a dataset converted to double / int64.
double as_double(const File& f, const Dataset& d, int64_t k) {
  uint8_t t[8];
  const uint64_t o = d.addr + (uint64_t)k * d.elem;
  for (int i = 0; i < d.elem; i++) t[i] = f.b[o + (d.big_endian ? d.elem - 1 - i : i)];
  switch (d.type) {
    case GS_H5_F64: {
      double v;
      std::memcpy(&v, t, 8);
      return v;
    }
    case GS_H5_F32: {
      float v;
      std::memcpy(&v, t, 4);
      return v;
    }
    case GS_H5_I32: {
      int32_t v;
      std::memcpy(&v, t, 4);
      return v;
    }
    default: {
      int64_t v;
      std::memcpy(&v, t, 8);
      return (double)v;
    }
  }
}

int64_t as_int(const File& f, const Dataset& d, int64_t k) {
  uint8_t t[8];
  const uint64_t o = d.addr + (uint64_t)k * d.elem;
  for (int i = 0; i < d.elem; i++) t[i] = f.b[o + (d.big_endian ? d.elem - 1 - i : i)];
  if (d.type == GS_H5_I32) {
    int32_t v;
    std::memcpy(&v, t, 4);
    return v;
  }
  if (d.type == GS_H5_I64) {
    int64_t v;
    std::memcpy(&v, t, 8);
    return v;
  }
  return (int64_t)as_double(f, d, k);
}

// Root-group attribute `name` (attribute message versions 1-3) as integers.
int attribute(const File& f, const char* name, int64_t* out, int capacity) {
  std::vector<Msg> ms;
  int st = messages(f, f.root, ms);
  if (st) return st;
  for (const Msg& m : ms) {
    if (m.type != 0x000C || m.size < 8) continue;
    const int ver = f.b[m.off];
    const uint64_t nlen = f.u(m.off + 2, 2), tlen = f.u(m.off + 4, 2), slen = f.u(m.off + 6, 2);
    auto pad = [&](uint64_t x) { return ver == 1 ? (x + 7) / 8 * 8 : x; };
    uint64_t o = m.off + 8 + (ver == 3 ? 1 : 0);
    if (ver < 1 || ver > 3 || o + pad(nlen) + pad(tlen) + pad(slen) > m.off + m.size) return GS_H5_FORMAT;
    const char* s = (const char*)&f.b[o];
    if (nlen == 0 || strnlen(s, nlen) != nlen - 1 || std::strcmp(s, name) != 0) continue;
    o += pad(nlen);
    Dataset d;
    if ((st = parse_type(f, o, tlen, d.type, d.elem, d.big_endian))) return st;
    o += pad(tlen);
    if ((st = parse_space(f, o, slen, d.rank, d.dims))) return st;
    o += pad(slen);
    d.addr = o;
    for (int i = 0; i < d.rank; i++)
      if (d.dims[i] < 0 || (uint64_t)d.dims[i] > m.size) return GS_H5_FORMAT;
    const int64_t n = count_of(d);
    if (n < 0 || (uint64_t)n > m.size || !f.in(o, (uint64_t)n * d.elem) || o + (uint64_t)n * d.elem > m.off + m.size) return GS_H5_FORMAT;
    for (int64_t k = 0; k < n && k < capacity; k++) out[k] = as_int(f, d, k);
    return (int)n;
  }
  return GS_H5_NOT_FOUND;
}

}  // namespace

extern "C" int32_t gs_h5_dataset_info(const char* path, const char* name, int32_t* rank, int64_t* dims,
                                      int32_t* type) {
  File f;
  int st = load(path, f);
  if (st) return st;
  Dataset d;
  if ((st = dataset(f, name, d))) return st;
  if (rank) *rank = d.rank;
  if (dims)
    for (int i = 0; i < GS_H5_MAX_RANK; i++) dims[i] = i < d.rank ? d.dims[i] : 0;
  if (type) *type = d.type;
  return 0;
}

extern "C" int64_t gs_h5_read_f64(const char* path, const char* name, double* out, int64_t capacity) {
  File f;
  int st = load(path, f);
  if (st) return st;
  Dataset d;
  if ((st = dataset(f, name, d))) return st;
  const int64_t n = count_of(d);
  if (!out || capacity < n) return GS_H5_INVALID;
  for (int64_t k = 0; k < n; k++) out[k] = as_double(f, d, k);
  return n;
}

extern "C" int64_t gs_h5_read_i32(const char* path, const char* name, int32_t* out, int64_t capacity) {
  File f;
  int st = load(path, f);
  if (st) return st;
  Dataset d;
  if ((st = dataset(f, name, d))) return st;
  if (d.type != GS_H5_I32 && d.type != GS_H5_I64) return GS_H5_UNSUPPORTED;
  const int64_t n = count_of(d);
  if (!out || capacity < n) return GS_H5_INVALID;
  for (int64_t k = 0; k < n; k++) {
    const int64_t v = as_int(f, d, k);
    if (v < INT32_MIN || v > INT32_MAX) return GS_H5_UNSUPPORTED;
    out[k] = (int32_t)v;
  }
  return n;
}

extern "C" int32_t gs_h5_attribute_i64(const char* path, const char* name, int64_t* out, int32_t capacity) {
  File f;
  int st = load(path, f);
  if (st) return st;
  if (!name || (capacity > 0 && !out)) return GS_H5_INVALID;
  return attribute(f, name, out, capacity);
}

extern "C" int64_t gs_h5_read_field(const char* path, int32_t dim, int32_t dofs, int64_t n_nodes, double* out) {
  if ((dim != 2 && dim != 3) || dofs < 1 || n_nodes < 0 || !out) return GS_H5_INVALID;
  File f;
  int st = load(path, f);
  if (st) return st;
  // FeVariable.cxx:2233-2262, :2326-2338: no "checkpoint file version"
  // attribute means the first layout, whose rows start with the node index.
  int64_t ver = 1;
  st = attribute(f, "checkpoint file version", &ver, 1);
  if (st == GS_H5_NOT_FOUND) ver = 1;
  else if (st < 0) return st;
  if (ver != 1 && ver != 2) return GS_H5_UNSUPPORTED;
  const int noffset = ver == 1 ? 1 : 0;
  Dataset d;
  if ((st = dataset(f, "data", d))) return st;
  if (d.rank != 2 || d.dims[0] != n_nodes) return GS_H5_SHAPE;
  // FeVariable.cxx:2348-2351: extra columns are the saved vertex coordinates
  const int64_t cols = d.dims[1];
  const bool coords = cols > dofs + noffset;
  if (cols != dofs + noffset + (coords ? dim : 0)) return GS_H5_SHAPE;
  const int64_t first = noffset + (coords ? dim : 0);
  for (int64_t i = 0; i < n_nodes; i++)
    for (int c = 0; c < dofs; c++) out[i * dofs + c] = as_double(f, d, i * cols + first + c);
  return n_nodes;
}

extern "C" int64_t gs_h5_read_swarm(const char* path, const char* swarm, int32_t dim, int64_t capacity,
                                    double* coord, int32_t* material) {
  if ((dim != 2 && dim != 3) || !swarm) return GS_H5_INVALID;
  File f;
  int st = load(path, f);
  if (st) return st;
  // FileParticleLayout.cxx:395-409: "/<variable>" first, then the older
  // "/<swarm>-<variable>" names
  std::string pos = "Position", mat = "MaterialIndex";
  Dataset dp;
  st = dataset(f, pos.c_str(), dp);
  if (st == GS_H5_NOT_FOUND) {
    pos = std::string(swarm) + "-Position";
    mat = std::string(swarm) + "-MaterialIndex";
    dp = Dataset();
    st = dataset(f, pos.c_str(), dp);
  }
  if (st) return st;
  if (dp.rank != 2 || dp.dims[1] != dim) return GS_H5_SHAPE;
  const int64_t P = dp.dims[0];
  if (!coord && !material) return P;
  if (capacity < P) return GS_H5_INVALID;
  Dataset dm;
  if (material) {
    if ((st = dataset(f, mat.c_str(), dm))) return st;
    if (count_of(dm) != P || (dm.type != GS_H5_I32 && dm.type != GS_H5_I64)) return GS_H5_SHAPE;
  }
  for (int64_t p = 0; p < P; p++) {
    if (coord) {
      for (int c = 0; c < dim; c++) coord[3 * p + c] = as_double(f, dp, p * dim + c);
      if (dim == 2) coord[3 * p + 2] = 0.0;
    }
    if (material) material[p] = (int32_t)as_int(f, dm, p);
  }
  return P;
}

// ---------------------------------------------------------------------------
// Writing: files in the same subset (superblock 0, version-1 object headers,
// one symbol-table root group holding contiguous datasets, root attributes),
// laid out as FeVariable_SaveToFile (checkpoint version 2), SwarmDump and the
// timeInfo writer name and shape their contents.
// ---------------------------------------------------------------------------
namespace {

constexpr uint64_t kUndef = ~0ull;
constexpr int kLeafK = 4, kInternalK = 16;  // the library defaults the reference's files carry

struct Buf {
  std::vector<uint8_t> b;
  uint64_t at() const { return b.size(); }
  void u(uint64_t v, int n) {
    for (int i = 0; i < n; i++) b.push_back((uint8_t)(v >> (8 * i)));
  }
  void bytes(const void* p, size_t n) {
    const uint8_t* c = (const uint8_t*)p;
    b.insert(b.end(), c, c + n);
  }
  void pad8() {
    while (b.size() % 8) b.push_back(0);
  }
  void zeros(size_t n) { b.insert(b.end(), n, 0); }
  void put(uint64_t off, uint64_t v, int n) {
    for (int i = 0; i < n; i++) b[off + i] = (uint8_t)(v >> (8 * i));
  }
};

// Datatype message bodies: little-endian IEEE double, native int32, and the
// big-endian int32 the reference uses for its attributes.
std::vector<uint8_t> type_f64() {
  Buf t;
  t.u(0x11, 1);                  // version 1, class 1 (floating point)
  t.u(0x20, 1);                  // little endian, implied mantissa msb
  t.u(63, 1);                    // sign bit position
  t.u(0, 1);
  t.u(8, 4);                     // size
  t.u(0, 2);                     // bit offset
  t.u(64, 2);                    // precision
  t.u(52, 1);                    // exponent location
  t.u(11, 1);                    // exponent size
  t.u(0, 1);                     // mantissa location
  t.u(52, 1);                    // mantissa size
  t.u(1023, 4);                  // exponent bias
  return t.b;
}
std::vector<uint8_t> type_i32(bool big_endian) {
  Buf t;
  t.u(0x10, 1);                  // version 1, class 0 (fixed point)
  t.u(big_endian ? 0x09 : 0x08, 1);  // byte order, signed
  t.u(0, 2);
  t.u(4, 4);
  t.u(0, 2);                     // bit offset
  t.u(32, 2);                    // precision
  return t.b;
}
std::vector<uint8_t> space_v1(const std::vector<uint64_t>& dims) {
  Buf s;
  s.u(1, 1);
  s.u(dims.size(), 1);
  s.u(0, 1);                     // no maximum dimensions
  s.zeros(5);
  for (uint64_t d : dims) s.u(d, 8);
  return s.b;
}

struct Item {
  std::string name;
  std::vector<uint64_t> dims;
  bool is_int = false, big_endian = false;
  const void* data = nullptr;    // row-major, native layout
  uint64_t nbytes = 0;
};

uint64_t count(const std::vector<uint64_t>& dims) {
  uint64_t n = 1;
  for (uint64_t d : dims) n *= d;
  return n;
}

// One object-header message of type `type` with body `body`, padded to 8.
void message(Buf& h, int type, const std::vector<uint8_t>& body) {
  const uint64_t n = (body.size() + 7) / 8 * 8;
  h.u(type, 2);
  h.u(n, 2);
  h.u(0, 1);
  h.zeros(3);
  h.bytes(body.data(), body.size());
  h.zeros(n - body.size());
}

std::vector<uint8_t> attribute_body(const Item& a) {
  Buf m;
  const std::vector<uint8_t> t = a.is_int ? type_i32(a.big_endian) : type_f64();
  const std::vector<uint8_t> s = space_v1(a.dims);
  m.u(1, 1);
  m.u(0, 1);
  m.u(a.name.size() + 1, 2);
  m.u(t.size(), 2);
  m.u(s.size(), 2);
  m.bytes(a.name.c_str(), a.name.size() + 1);
  m.pad8();
  m.bytes(t.data(), t.size());
  m.pad8();
  m.bytes(s.data(), s.size());
  m.pad8();
  const uint8_t* p = (const uint8_t*)a.data;
  const int elem = a.is_int ? 4 : 8;
  for (uint64_t k = 0; k < count(a.dims); k++)
    for (int i = 0; i < elem; i++) m.u(p[k * elem + (a.big_endian ? elem - 1 - i : i)], 1);
  return m.b;
}

void object_header(Buf& f, const std::vector<std::pair<int, std::vector<uint8_t>>>& msgs) {
  Buf h;
  for (auto& m : msgs) message(h, m.first, m.second);
  f.u(1, 1);            // version
  f.u(0, 1);
  f.u(msgs.size(), 2);
  f.u(1, 4);            // reference count
  f.u(h.b.size(), 4);   // header data size
  f.zeros(4);
  f.bytes(h.b.data(), h.b.size());
}

int write_file(const char* path, std::vector<Item> sets, const std::vector<Item>& attrs) {
  if (!path) return GS_H5_INVALID;
  // symbol-table entries are kept sorted by name (the B-tree key order)
  std::sort(sets.begin(), sets.end(), [](const Item& a, const Item& b) { return a.name < b.name; });
  if ((int)sets.size() > 2 * kLeafK) return GS_H5_UNSUPPORTED;
  Buf f;
  f.zeros(96);                                  // superblock, filled in last
  // local heap: the empty name at 0, then each dataset name padded to 8
  std::vector<uint64_t> name_off;
  Buf heap_data;
  heap_data.zeros(8);
  for (auto& s : sets) {
    name_off.push_back(heap_data.at());
    heap_data.bytes(s.name.c_str(), s.name.size() + 1);
    heap_data.pad8();
  }
  // root object header: symbol table message + attributes
  const uint64_t root = f.at();
  std::vector<std::pair<int, std::vector<uint8_t>>> rmsgs;
  rmsgs.push_back({0x11, std::vector<uint8_t>(16, 0)});
  for (auto& a : attrs) rmsgs.push_back({0x0C, attribute_body(a)});
  object_header(f, rmsgs);
  const uint64_t stab_body = root + 16 + 8;     // first message's body
  // B-tree (group node, level 0, one child), sized for 2K entries
  const uint64_t tree = f.at();
  f.bytes("TREE", 4);
  f.u(0, 1);
  f.u(0, 1);
  f.u(sets.empty() ? 0 : 1, 2);
  f.u(kUndef, 8);
  f.u(kUndef, 8);
  const uint64_t keys = f.at();
  f.zeros((2 * kInternalK + 1) * 8 + 2 * kInternalK * 8);
  // local heap header and data segment
  const uint64_t heap = f.at();
  f.bytes("HEAP", 4);
  f.zeros(4);
  f.u(heap_data.at(), 8);
  f.u(1, 8);                                    // no free block
  f.u(heap + 32, 8);
  f.bytes(heap_data.b.data(), heap_data.at());
  // dataset object headers
  std::vector<uint64_t> ds_hdr, layout_at;
  for (auto& s : sets) {
    ds_hdr.push_back(f.at());
    std::vector<uint8_t> lay(18, 0);
    lay[0] = 3;
    lay[1] = 1;                                 // contiguous; address and size patched below
    object_header(f, {{0x0001, space_v1(s.dims)}, {0x0003, s.is_int ? type_i32(false) : type_f64()}, {0x0008, lay}});
    layout_at.push_back(f.at() - 24 + 2);       // the layout message body is the last 24 bytes
  }
  // symbol-table node sized for 2 * leaf K entries
  const uint64_t snod = f.at();
  f.bytes("SNOD", 4);
  f.u(1, 1);
  f.u(0, 1);
  f.u(sets.size(), 2);
  for (size_t i = 0; i < sets.size(); i++) {
    f.u(name_off[i], 8);
    f.u(ds_hdr[i], 8);
    f.u(0, 4);
    f.zeros(4 + 16);
  }
  f.zeros((2 * kLeafK - sets.size()) * 40);
  // raw data
  for (size_t i = 0; i < sets.size(); i++) {
    f.pad8();
    f.put(layout_at[i], f.at(), 8);
    f.put(layout_at[i] + 8, sets[i].nbytes, 8);
    f.bytes(sets[i].data, sets[i].nbytes);
  }
  // patch B-tree keys/child, the symbol table message and the superblock
  if (!sets.empty()) {
    f.put(keys, 0, 8);
    f.put(keys + 8, snod, 8);
    f.put(keys + 16, name_off.back(), 8);
  }
  f.put(stab_body, tree, 8);
  f.put(stab_body + 8, heap, 8);
  Buf sb;
  const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
  sb.bytes(sig, 8);
  sb.u(0, 1);   // superblock version
  sb.u(0, 1);   // free-space version
  sb.u(0, 1);   // root group symbol table entry version
  sb.u(0, 1);
  sb.u(0, 1);   // shared header message format version
  sb.u(8, 1);   // size of offsets
  sb.u(8, 1);   // size of lengths
  sb.u(0, 1);
  sb.u(kLeafK, 2);
  sb.u(kInternalK, 2);
  sb.u(0, 4);   // file consistency flags
  sb.u(0, 8);   // base address
  sb.u(kUndef, 8);
  sb.u(f.at(), 8);  // end of file
  sb.u(kUndef, 8);
  sb.u(0, 8);   // root entry: link name offset
  sb.u(root, 8);
  sb.u(1, 4);   // cache type 1: scratch holds the B-tree and heap addresses
  sb.u(0, 4);
  sb.u(tree, 8);
  sb.u(heap, 8);
  std::memcpy(f.b.data(), sb.b.data(), 96);
  FILE* fp = std::fopen(path, "wb");
  if (!fp) return GS_H5_IO;
  const size_t put = std::fwrite(f.b.data(), 1, f.b.size(), fp);
  const int closed = std::fclose(fp);
  return (put == f.b.size() && closed == 0) ? 0 : GS_H5_IO;
}

}  // namespace

extern "C" int32_t gs_h5_write_field(const char* path, int32_t dim, int32_t dofs, int64_t n_nodes,
                                     const double* values, const double* coords, const int32_t* mesh_res) {
  if ((dim != 2 && dim != 3) || dofs < 1 || n_nodes < 0 || (n_nodes > 0 && !values)) return GS_H5_INVALID;
  // FeVariable.cxx:1938-2130, checkpoint version 2: rows of [coordinates,] components
  const int cols = dofs + (coords ? dim : 0);
  std::vector<double> rows((size_t)n_nodes * cols);
  for (int64_t i = 0; i < n_nodes; i++) {
    int c = 0;
    if (coords)
      for (int d = 0; d < dim; d++) rows[i * cols + c++] = coords[3 * i + d];
    for (int k = 0; k < dofs; k++) rows[i * cols + c++] = values[i * dofs + k];
  }
  const int32_t ver = 2, dims = dim;
  std::vector<Item> attrs;
  attrs.push_back({"checkpoint file version", {1}, true, true, &ver, 4});
  attrs.push_back({"dimensions", {1}, true, true, &dims, 4});
  if (mesh_res) attrs.push_back({"mesh resolution", {(uint64_t)dim}, true, true, mesh_res, 4ull * dim});
  return write_file(path, {{"data", {(uint64_t)n_nodes, (uint64_t)cols}, false, false, rows.data(),
                            rows.size() * sizeof(double)}},
                    attrs);
}

extern "C" int32_t gs_h5_write_swarm_vars(const char* path, int32_t dim, int64_t P, const double* coord,
                                          const int32_t* material, int32_t n_vars, const char* const* names,
                                          const double* const* values) {
  if ((dim != 2 && dim != 3) || P < 0 || P > INT32_MAX || (P > 0 && (!coord || !material))) return GS_H5_INVALID;
  if (n_vars < 0 || n_vars > 2 * kLeafK - 2 || (n_vars > 0 && (!names || !values))) return GS_H5_INVALID;
  // SwarmDump.cxx:213-330: "/Position" (P x dim), "/MaterialIndex" (P x 1),
  // one "/<variable>" per further swarm variable, and the particle count as
  // a root attribute
  std::vector<double> pos((size_t)P * dim);
  for (int64_t p = 0; p < P; p++)
    for (int d = 0; d < dim; d++) pos[p * dim + d] = coord[3 * p + d];
  const int32_t n = (int32_t)P;
  std::vector<Item> sets = {
      {"Position", {(uint64_t)P, (uint64_t)dim}, false, false, pos.data(), pos.size() * sizeof(double)},
      {"MaterialIndex", {(uint64_t)P, 1}, true, false, material, (uint64_t)P * 4}};
  for (int v = 0; v < n_vars; v++) {
    if (!names[v] || !names[v][0] || (P > 0 && !values[v])) return GS_H5_INVALID;
    for (const Item& it : sets)
      if (it.name == names[v]) return GS_H5_INVALID;
    sets.push_back({names[v], {(uint64_t)P, 1}, false, false, values[v], (uint64_t)P * 8});
  }
  return write_file(path, sets, {{"Swarm Particle Count", {1}, true, true, &n, 4}});
}

extern "C" int32_t gs_h5_write_swarm(const char* path, int32_t dim, int64_t P, const double* coord,
                                     const int32_t* material) {
  return gs_h5_write_swarm_vars(path, dim, P, coord, material, 0, nullptr, nullptr);
}

extern "C" int64_t gs_h5_read_swarm_var(const char* path, const char* swarm, const char* name, int64_t capacity,
                                        double* out) {
  if (!swarm || !name) return GS_H5_INVALID;
  File f;
  int st = load(path, f);
  if (st) return st;
  // FileParticleLayout.cxx:395-409 naming: "/<variable>", then "/<swarm>-<variable>"
  Dataset d;
  st = dataset(f, name, d);
  if (st == GS_H5_NOT_FOUND) {
    d = Dataset();
    st = dataset(f, (std::string(swarm) + "-" + name).c_str(), d);
  }
  if (st) return st;
  const int64_t n = count_of(d);
  if (d.rank < 1 || (d.rank == 2 && d.dims[1] != 1) || d.rank > 2) return GS_H5_SHAPE;
  if (!out) return n;
  if (capacity < n) return GS_H5_INVALID;
  for (int64_t k = 0; k < n; k++) out[k] = as_double(f, d, k);
  return n;
}

extern "C" int32_t gs_h5_write_time(const char* path, double dt, double time, int32_t nproc) {
  // the timeInfo file FileParticleLayout.cxx:565-597 reads back: Dt, currentTime, nproc
  return write_file(path,
                    {{"Dt", {1}, false, false, &dt, 8},
                     {"currentTime", {1}, false, false, &time, 8},
                     {"nproc", {1}, true, false, &nproc, 4}},
                    {});
}
