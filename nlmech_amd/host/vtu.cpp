// VTK XML UnstructuredGrid (.vtu) writer and reader (see vtu.h).
#include "vtu.h"

#include <zlib.h>

#include <algorithm>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace nlm {
namespace rw {

namespace {

constexpr size_t kBlock = 32768;  // vtkZLibDataCompressor block size

const char* kB64 =
    "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

std::string b64_encode(const uint8_t* p, size_t n) {
  std::string out;
  out.reserve((n + 2) / 3 * 4);
  size_t i = 0;
  for (; i + 2 < n; i += 3) {
    uint32_t v = (uint32_t(p[i]) << 16) | (uint32_t(p[i + 1]) << 8) | p[i + 2];
    out += kB64[(v >> 18) & 63];
    out += kB64[(v >> 12) & 63];
    out += kB64[(v >> 6) & 63];
    out += kB64[v & 63];
  }
  if (i < n) {
    uint32_t v = uint32_t(p[i]) << 16;
    if (i + 1 < n) v |= uint32_t(p[i + 1]) << 8;
    out += kB64[(v >> 18) & 63];
    out += kB64[(v >> 12) & 63];
    out += i + 1 < n ? kB64[(v >> 6) & 63] : '=';
    out += '=';
  }
  return out;
}

int b64_value(char c) {
  if (c >= 'A' && c <= 'Z') return c - 'A';
  if (c >= 'a' && c <= 'z') return c - 'a' + 26;
  if (c >= '0' && c <= '9') return c - '0' + 52;
  if (c == '+') return 62;
  if (c == '/') return 63;
  return -1;
}

// Decodes `nchars` base64 characters starting at s[pos] (whitespace already
// removed).
std::vector<uint8_t> b64_decode(const std::string& s, size_t pos,
                                size_t nchars) {
  if (pos > s.size() || nchars > s.size() - pos)
    throw std::runtime_error("vtu: base64 data truncated");
  std::vector<uint8_t> out;
  out.reserve(nchars / 4 * 3);
  uint32_t acc = 0;
  int bits = 0;
  for (size_t i = pos; i < pos + nchars; i++) {
    char c = s[i];
    if (c == '=') break;
    int v = b64_value(c);
    if (v < 0) throw std::runtime_error("vtu: bad base64 character");
    acc = (acc << 6) | uint32_t(v);
    bits += 6;
    if (bits >= 8) {
      bits -= 8;
      out.push_back(uint8_t((acc >> bits) & 0xFF));
    }
  }
  return out;
}

size_t b64_len(size_t nbytes) { return (nbytes + 2) / 3 * 4; }

size_t type_size(const std::string& t) {
  if (t == "Float64" || t == "Int64" || t == "UInt64") return 8;
  if (t == "Float32" || t == "Int32" || t == "UInt32") return 4;
  if (t == "Int16" || t == "UInt16") return 2;
  if (t == "Int8" || t == "UInt8") return 1;
  throw std::runtime_error("vtu: unsupported data type " + t);
}

template <class T>
std::vector<uint8_t> bytes_of(const T* p, size_t n) {
  std::vector<uint8_t> b(n * sizeof(T));
  if (n) std::memcpy(b.data(), p, b.size());
  return b;
}

std::string attr(const std::string& tag, const std::string& name) {
  const std::string key = " " + name + "=\"";
  size_t p = tag.find(key);
  if (p == std::string::npos) {
    // attribute at the start of the tag text after the element name
    return std::string();
  }
  p += key.size();
  size_t e = tag.find('"', p);
  if (e == std::string::npos) return std::string();
  return tag.substr(p, e - p);
}

std::string strip_ws(const std::string& s) {
  std::string out;
  out.reserve(s.size());
  for (char c : s)
    if (!(c == ' ' || c == '\n' || c == '\r' || c == '\t')) out += c;
  return out;
}

}  // namespace

int vtk_cell_vertices(int type) {
  switch (type) {
    case 1:
      return 1;
    case kVtkLine:
      return 2;
    case kVtkTriangle:
      return 3;
    case kVtkQuad:
    case kVtkTetra:
      return 4;
    default:
      return 0;
  }
}

// ---------------------------------------------------------------------------
// Writer
// ---------------------------------------------------------------------------

namespace {
// stream formatting of the legacy writer (std::ofstream defaults)
template <class T>
std::string legacy_rows(const T* data, int64_t n, int ncomp) {
  std::ostringstream o;
  for (int64_t i = 0; i < n; i++)
    for (int c = 0; c < ncomp; c++)
      o << data[i * ncomp + c] << (c + 1 < ncomp ? " " : "\n");
  return o.str();
}
void legacy_warning() {
  std::cerr << "Warning: Field data is not implementend in the legacy vtk "
               "writer"
            << std::endl;
}
}  // namespace

VtuWriter::VtuWriter(const std::string& filename, const std::string& compress,
                     bool legacy)
    : file_(filename + (legacy ? ".vtk" : ".vtu")),
      zlib_(compress == "zlib"),
      legacy_(legacy) {
  if (legacy_)
    legacy_text_ =
        "# vtk DataFile Version 3.0\nNLMech legacy vtk writer\nASCII\n";
}

void VtuWriter::appendNodes(const std::vector<double>& nodes, const double* u) {
  npoints_ = int64_t(nodes.size() / 3);
  if (legacy_) {
    std::ostringstream o;
    o << "DATASET UNSTRUCTURED_GRID\nPOINTS " << npoints_ << " double\n";
    for (int64_t i = 0; i < npoints_; i++) {
      double p[3];
      for (int d = 0; d < 3; d++)
        p[d] = u ? nodes[3 * i + d] + u[3 * i + d] : nodes[3 * i + d];
      o << p[0] << " " << p[1] << " " << p[2] << "\n";
    }
    o << "CELLS " << npoints_ << " " << 2 * npoints_ << "\n";
    for (int64_t i = 0; i < npoints_; i++) o << "1 " << i << "\n";
    o << "CELL_TYPES " << npoints_ << "\n";
    for (int64_t i = 0; i < npoints_; i++) o << "1\n";
    o << "POINT_DATA " << npoints_ << "\n";
    legacy_text_ += o.str();
    return;
  }
  std::vector<float> p(nodes.size());
  for (size_t k = 0; k < nodes.size(); k++) {
    double v = nodes[k];
    if (u) v = v + u[k];
    p[k] = float(v);
  }
  points_ = {"Points", "Float32", 3, npoints_, bytes_of(p.data(), p.size())};
}

void VtuWriter::appendCells(int type, const std::vector<int64_t>& conn) {
  if (legacy_) {
    legacy_warning();
    return;
  }
  const int nv = vtk_cell_vertices(type);
  if (nv == 0) throw std::runtime_error("vtu: unsupported cell type");
  ncells_ = int64_t(conn.size()) / nv;
  std::vector<int64_t> off(ncells_);
  for (int64_t c = 0; c < ncells_; c++) off[c] = (c + 1) * nv;
  std::vector<uint8_t> ty(ncells_, uint8_t(type));
  conn_ = {"connectivity", "Int64", 1, int64_t(conn.size()),
           bytes_of(conn.data(), conn.size())};
  offsets_ = {"offsets", "Int64", 1, ncells_, bytes_of(off.data(), off.size())};
  types_ = {"types", "UInt8", 1, ncells_, ty};
}

void VtuWriter::appendPointData(const std::string& name, const double* data,
                                int ncomp) {
  if (legacy_) {
    if (ncomp == 6) {
      // (xx, yy, zz, yz, xz, xy) as the full 3 x 3 matrix, row by row
      std::vector<double> m(size_t(npoints_) * 9);
      const int at[9] = {0, 5, 4, 5, 1, 3, 4, 3, 2};
      for (int64_t i = 0; i < npoints_; i++)
        for (int k = 0; k < 9; k++) m[size_t(9 * i + k)] = data[6 * i + at[k]];
      legacy_text_ += "TENSORS " + name + " double \n";
      legacy_text_ += legacy_rows(m.data(), 3 * npoints_, 3);
    } else {
      legacy_text_ += "SCALARS " + name + " double " + std::to_string(ncomp) +
                      "\nLOOKUP_TABLE default\n";
      legacy_text_ += legacy_rows(data, npoints_, ncomp);
    }
    return;
  }
  point_.push_back({name, "Float64", ncomp, npoints_,
                    bytes_of(data, size_t(npoints_) * ncomp)});
}

void VtuWriter::appendPointData(const std::string& name, const float* data,
                                int ncomp) {
  if (legacy_) {
    legacy_text_ += "SCALARS " + name + " float " + std::to_string(ncomp) +
                    "\nLOOKUP_TABLE default\n";
    legacy_text_ += legacy_rows(data, npoints_, ncomp);
    return;
  }
  std::vector<double> d(size_t(npoints_) * ncomp);
  for (size_t k = 0; k < d.size(); k++) d[k] = data[k];
  appendPointData(name, d.data(), ncomp);
}

void VtuWriter::appendPointData(const std::string& name, const uint8_t* data) {
  if (legacy_) {
    // (as integers; the reference streams the bytes as characters)
    std::vector<int> d(data, data + npoints_);
    legacy_text_ += "SCALARS " + name + " int 1\nLOOKUP_TABLE default\n";
    legacy_text_ += legacy_rows(d.data(), npoints_, 1);
    return;
  }
  std::vector<double> d(static_cast<size_t>(npoints_));
  for (size_t k = 0; k < d.size(); k++) d[k] = data[k];
  appendPointData(name, d.data(), 1);
}

void VtuWriter::appendFieldData(const std::string& name, double value) {
  if (legacy_) {
    legacy_warning();
    return;
  }
  field_.push_back({name, "Float64", 1, 1, bytes_of(&value, 1)});
}

void VtuWriter::appendFieldData(const std::string& name,
                                const std::vector<uint8_t>& bytes) {
  if (legacy_) {
    legacy_warning();
    return;
  }
  field_.push_back({name, "UInt8", 1, int64_t(bytes.size()), bytes});
}

void VtuWriter::close() {
  if (closed_) return;
  closed_ = true;
  if (legacy_) {
    std::ofstream f(file_);
    if (!f) throw std::runtime_error("cannot write " + file_);
    f << legacy_text_;
    if (!f) throw std::runtime_error("write error on " + file_);
    return;
  }
  std::vector<const Array*> order;
  for (const auto& a : field_) order.push_back(&a);
  for (const auto& a : point_) order.push_back(&a);
  order.push_back(&points_);
  if (ncells_ > 0) {
    order.push_back(&conn_);
    order.push_back(&offsets_);
    order.push_back(&types_);
  }
  bool big = false;
  for (const Array* a : order)
    if (a->bytes.size() >= (uint64_t(1) << 32)) big = true;
  const size_t hs = big ? 8 : 4;

  auto put_header = [&](std::vector<uint8_t>* h, uint64_t v) {
    if (big) {
      const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
      h->insert(h->end(), p, p + 8);
    } else {
      uint32_t w = uint32_t(v);
      const uint8_t* p = reinterpret_cast<const uint8_t*>(&w);
      h->insert(h->end(), p, p + 4);
    }
  };

  std::string data;
  std::vector<int64_t> offsets;
  for (const Array* a : order) {
    offsets.push_back(int64_t(data.size()));
    const size_t n = a->bytes.size();
    if (zlib_) {
      const size_t nb = (n + kBlock - 1) / kBlock;
      std::vector<uint8_t> comp;
      std::vector<uint64_t> csize(nb);
      for (size_t b = 0; b < nb; b++) {
        const size_t len = std::min(kBlock, n - b * kBlock);
        uLongf clen = compressBound(uLong(len));
        std::vector<uint8_t> buf(clen);
        if (compress2(buf.data(), &clen, a->bytes.data() + b * kBlock,
                      uLong(len), 5) != Z_OK)
          throw std::runtime_error("vtu: zlib compression failed");
        csize[b] = clen;
        comp.insert(comp.end(), buf.begin(), buf.begin() + clen);
      }
      std::vector<uint8_t> hdr;
      put_header(&hdr, nb);
      put_header(&hdr, kBlock);
      put_header(&hdr, n % kBlock);
      for (uint64_t c : csize) put_header(&hdr, c);
      data += b64_encode(hdr.data(), hdr.size());
      data += b64_encode(comp.data(), comp.size());
    } else {
      std::vector<uint8_t> buf;
      buf.reserve(hs + n);
      put_header(&buf, n);
      buf.insert(buf.end(), a->bytes.begin(), a->bytes.end());
      data += b64_encode(buf.data(), buf.size());
    }
  }

  std::ostringstream x;
  size_t k = 0;
  auto array_tag = [&](const Array& a, bool field) {
    x << "<DataArray type=\"" << a.type << "\" Name=\"" << a.name << "\"";
    if (field) x << " NumberOfTuples=\"" << a.ntuples << "\"";
    if (a.ncomp > 1) x << " NumberOfComponents=\"" << a.ncomp << "\"";
    x << " format=\"appended\" offset=\"" << offsets[k++] << "\"/>\n";
  };
  x << "<?xml version=\"1.0\"?>\n";
  x << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" "
       "byte_order=\"LittleEndian\" header_type=\""
    << (big ? "UInt64" : "UInt32") << "\"";
  if (zlib_) x << " compressor=\"vtkZLibDataCompressor\"";
  x << ">\n  <UnstructuredGrid>\n";
  if (!field_.empty()) {
    x << "    <FieldData>\n";
    for (const auto& a : field_) {
      x << "      ";
      array_tag(a, true);
    }
    x << "    </FieldData>\n";
  }
  x << "    <Piece NumberOfPoints=\"" << npoints_ << "\" NumberOfCells=\""
    << ncells_ << "\">\n      <PointData>\n";
  for (const auto& a : point_) {
    x << "        ";
    array_tag(a, false);
  }
  x << "      </PointData>\n      <CellData>\n      </CellData>\n"
       "      <Points>\n        ";
  array_tag(points_, false);
  x << "      </Points>\n      <Cells>\n";
  if (ncells_ > 0) {
    for (const Array* a : {&conn_, &offsets_, &types_}) {
      x << "        ";
      array_tag(*a, false);
    }
  }
  x << "      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n"
       "  <AppendedData encoding=\"base64\">\n   _";
  std::ofstream f(file_, std::ios::binary);
  if (!f) throw std::runtime_error("vtu: cannot write " + file_);
  const std::string head = x.str();
  f.write(head.data(), std::streamsize(head.size()));
  f.write(data.data(), std::streamsize(data.size()));
  const std::string tail = "\n  </AppendedData>\n</VTKFile>\n";
  f.write(tail.data(), std::streamsize(tail.size()));
  if (!f) throw std::runtime_error("vtu: write failed for " + file_);
}

// ---------------------------------------------------------------------------
// Reader
// ---------------------------------------------------------------------------

VtuReader::VtuReader(const std::string& filename) : file_(filename) {
  std::ifstream f(filename, std::ios::binary);
  if (!f) throw std::runtime_error("vtu: cannot open " + filename);
  std::stringstream ss;
  ss << f.rdbuf();
  std::string all = ss.str();
  size_t app = all.find("<AppendedData");
  if (app != std::string::npos) {
    size_t tag_end = all.find('>', app);
    const std::string tag = all.substr(app, tag_end - app);
    appended_b64_ = attr(tag, "encoding") != "raw";
    size_t us = all.find('_', tag_end);
    if (us == std::string::npos)
      throw std::runtime_error("vtu: appended data without '_'");
    size_t end = all.rfind("</AppendedData>");
    if (end == std::string::npos || end < us)
      throw std::runtime_error("vtu: appended data not closed");
    appended_ = all.substr(us + 1, end - us - 1);
    if (appended_b64_) appended_ = strip_ws(appended_);
    text_ = all.substr(0, app);
  } else {
    text_ = all;
  }
  size_t vf = text_.find("<VTKFile");
  if (vf == std::string::npos) throw std::runtime_error("vtu: not a VTK file");
  const std::string vtag = text_.substr(vf, text_.find('>', vf) - vf);
  if (attr(vtag, "type") != "UnstructuredGrid")
    throw std::runtime_error("vtu: not an UnstructuredGrid file");
  if (attr(vtag, "byte_order") == "BigEndian")
    throw std::runtime_error("vtu: big-endian files are not supported");
  header64_ = attr(vtag, "header_type") == "UInt64";
  const std::string comp = attr(vtag, "compressor");
  zlib_ = !comp.empty();
  if (zlib_ && comp != "vtkZLibDataCompressor")
    throw std::runtime_error("vtu: unsupported compressor " + comp);

  size_t pc = text_.find("<Piece");
  if (pc == std::string::npos) throw std::runtime_error("vtu: no Piece");
  const std::string ptag = text_.substr(pc, text_.find('>', pc) - pc);
  npoints_ = std::stoll(attr(ptag, "NumberOfPoints"));
  std::string nc = attr(ptag, "NumberOfCells");
  ncells_ = nc.empty() ? 0 : std::stoll(nc);

  // walk the DataArray elements, tracking the enclosing section
  size_t pos = 0;
  while (true) {
    size_t a = text_.find("<DataArray", pos);
    if (a == std::string::npos) break;
    size_t te = text_.find('>', a);
    std::string tag = text_.substr(a, te - a);
    bool selfclose = tag.back() == '/';
    ArrayRef r;
    r.type = attr(tag, "type");
    std::string ncs = attr(tag, "NumberOfComponents");
    r.ncomp = ncs.empty() ? 1 : std::stoi(ncs);
    r.format = attr(tag, "format");
    std::string off = attr(tag, "offset");
    r.offset = off.empty() ? 0 : std::stoll(off);
    pos = te + 1;
    if (!selfclose) {
      size_t ce = text_.find("</DataArray>", pos);
      r.inline_text = text_.substr(pos, ce - pos);
      pos = ce + 12;
    }
    const std::string name = attr(tag, "Name");
    // the innermost open section before this array
    auto last = [&](const char* open) {
      size_t p = text_.rfind(open, a);
      return p == std::string::npos ? -1 : int64_t(p);
    };
    const int64_t s_field = last("<FieldData"), s_point = last("<PointData"),
                  s_pts = last("<Points"), s_cells = last("<Cells"),
                  s_cell = last("<CellData");
    int64_t best = std::max(std::max(s_field, s_point),
                            std::max(std::max(s_pts, s_cells), s_cell));
    if (best < 0) continue;
    auto closed = [&](const char* close, int64_t open) {
      size_t p = text_.find(close, size_t(open));
      return p != std::string::npos && int64_t(p) < int64_t(a);
    };
    if (best == s_field && !closed("</FieldData>", s_field)) {
      field_[name] = r;
    } else if (best == s_point && !closed("</PointData>", s_point)) {
      point_[name] = r;
    } else if (best == s_pts && !closed("</Points>", s_pts)) {
      points_ = r;
      has_points_ = true;
    } else if (best == s_cells && !closed("</Cells>", s_cells)) {
      has_cells_ = true;
      if (name == "connectivity") conn_ = r;
      else if (name == "offsets") offsets_ = r;
      else if (name == "types") types_ = r;
    }
  }
}

std::vector<uint8_t> VtuReader::decode(const ArrayRef& a) const {
  const size_t hs = header64_ ? 8 : 4;
  // header words and sizes come from the file: every read of one is bounds
  // checked and every size compared without overflow before it is used
  auto hval = [&](const std::vector<uint8_t>& b, size_t k) -> uint64_t {
    if (k >= b.size() / hs) throw std::runtime_error("vtu: header truncated");
    if (header64_) {
      uint64_t v;
      std::memcpy(&v, b.data() + 8 * k, 8);
      return v;
    }
    uint32_t v;
    std::memcpy(&v, b.data() + 4 * k, 4);
    return v;
  };
  if (a.format == "ascii") {
    throw std::runtime_error("vtu: decode() called on ascii data");
  }
  const bool b64 = a.format == "binary" || appended_b64_;
  std::string inline_b64;
  if (a.format == "binary") inline_b64 = strip_ws(a.inline_text);
  const std::string& src = a.format == "binary" ? inline_b64 : appended_;
  const size_t start = a.format == "binary" ? 0 : size_t(a.offset);

  std::vector<uint8_t> out;
  if (!zlib_) {
    std::vector<uint8_t> h;
    size_t nbytes;
    if (b64) {
      h = b64_decode(src, start, b64_len(hs));
      nbytes = size_t(hval(h, 0));
      if (nbytes > src.size())
        throw std::runtime_error("vtu: array data truncated");
      std::vector<uint8_t> all = b64_decode(src, start, b64_len(hs + nbytes));
      if (all.size() < hs + nbytes)
        throw std::runtime_error("vtu: array data truncated");
      out.assign(all.begin() + hs, all.begin() + hs + nbytes);
    } else {
      if (start > src.size() || hs > src.size() - start)
        throw std::runtime_error("vtu: raw data truncated");
      h.assign(src.begin() + start, src.begin() + start + hs);
      nbytes = size_t(hval(h, 0));
      if (nbytes > src.size() - start - hs)
        throw std::runtime_error("vtu: raw data truncated");
      out.assign(src.begin() + start + hs, src.begin() + start + hs + nbytes);
    }
    return out;
  }
  // compressed: header [nblocks, block, last, csize...]
  std::vector<uint8_t> first;
  size_t hdr_chars = 0;
  std::vector<uint8_t> hdr;
  if (start > src.size()) throw std::runtime_error("vtu: truncated");
  if (b64) {
    first = b64_decode(src, start, b64_len(hs));
    const size_t nb = size_t(hval(first, 0));
    if (nb > src.size() / hs) throw std::runtime_error("vtu: truncated");
    hdr_chars = b64_len(hs * (3 + nb));
    hdr = b64_decode(src, start, hdr_chars);
  } else {
    if (hs > src.size() - start) throw std::runtime_error("vtu: truncated");
    first.assign(src.begin() + start, src.begin() + start + hs);
    const size_t nb = size_t(hval(first, 0));
    if (nb > src.size() / hs) throw std::runtime_error("vtu: truncated");
    hdr_chars = hs * (3 + nb);
    if (hdr_chars > src.size() - start)
      throw std::runtime_error("vtu: truncated");
    hdr.assign(src.begin() + start, src.begin() + start + hdr_chars);
  }
  const size_t nb = size_t(hval(hdr, 0));
  const size_t bs = size_t(hval(hdr, 1));
  const size_t last = size_t(hval(hdr, 2));
  // (zlib expands a block at most ~1032 times; the writer's blocks are
  // 32 KiB)
  if (bs > (size_t(1) << 31) || last > bs)
    throw std::runtime_error("vtu: bad compressed block size");
  size_t total = 0;
  for (size_t b = 0; b < nb; b++) {
    const size_t cs = size_t(hval(hdr, 3 + b));
    if (cs > src.size() - total) throw std::runtime_error("vtu: truncated");
    total += cs;
  }
  std::vector<uint8_t> comp;
  if (b64) {
    comp = b64_decode(src, start + hdr_chars, b64_len(total));
  } else {
    if (total > src.size() - start - hdr_chars)
      throw std::runtime_error("vtu: truncated");
    comp.assign(src.begin() + start + hdr_chars,
                src.begin() + start + hdr_chars + total);
  }
  size_t pos = 0;
  for (size_t b = 0; b < nb; b++) {
    const size_t cs = size_t(hval(hdr, 3 + b));
    size_t us = (b + 1 == nb && last != 0) ? last : bs;
    std::vector<uint8_t> buf(us);
    uLongf ul = uLongf(us);
    if (pos + cs > comp.size() ||
        uncompress(buf.data(), &ul, comp.data() + pos, uLong(cs)) != Z_OK ||
        ul != us)
      throw std::runtime_error("vtu: zlib block failed in " + file_);
    out.insert(out.end(), buf.begin(), buf.end());
    pos += cs;
  }
  return out;
}

bool VtuReader::to_double(const ArrayRef& a, std::vector<double>* out) const {
  out->clear();
  if (a.format == "ascii") {
    std::istringstream is(a.inline_text);
    double v;
    while (is >> v) out->push_back(v);
    return true;
  }
  const std::vector<uint8_t> b = decode(a);
  const size_t ts = type_size(a.type);
  const size_t n = b.size() / ts;
  out->resize(n);
  const uint8_t* p = b.data();
  for (size_t k = 0; k < n; k++, p += ts) {
    double v = 0;
    if (a.type == "Float64") { double t; std::memcpy(&t, p, 8); v = t; }
    else if (a.type == "Float32") { float t; std::memcpy(&t, p, 4); v = t; }
    else if (a.type == "Int64") { int64_t t; std::memcpy(&t, p, 8); v = double(t); }
    else if (a.type == "UInt64") { uint64_t t; std::memcpy(&t, p, 8); v = double(t); }
    else if (a.type == "Int32") { int32_t t; std::memcpy(&t, p, 4); v = t; }
    else if (a.type == "UInt32") { uint32_t t; std::memcpy(&t, p, 4); v = t; }
    else if (a.type == "Int16") { int16_t t; std::memcpy(&t, p, 2); v = t; }
    else if (a.type == "UInt16") { uint16_t t; std::memcpy(&t, p, 2); v = t; }
    else if (a.type == "Int8") { int8_t t; std::memcpy(&t, p, 1); v = t; }
    else if (a.type == "UInt8") { v = *p; }
    (*out)[k] = v;
  }
  return true;
}

bool VtuReader::readNodes(std::vector<double>* xyz) const {
  if (!has_points_) return false;
  return to_double(points_, xyz) && xyz->size() == size_t(npoints_) * 3;
}

bool VtuReader::hasPointData(const std::string& name) const {
  return point_.count(name) != 0;
}

bool VtuReader::readPointData(const std::string& name, std::vector<double>* out,
                              int* ncomp) const {
  auto it = point_.find(name);
  if (it == point_.end()) return false;
  if (ncomp) *ncomp = it->second.ncomp;
  return to_double(it->second, out);
}

bool VtuReader::readFieldData(const std::string& name,
                              std::vector<double>* out) const {
  auto it = field_.find(name);
  if (it == field_.end()) return false;
  return to_double(it->second, out);
}

bool VtuReader::readFieldBytes(const std::string& name,
                               std::vector<uint8_t>* out) const {
  auto it = field_.find(name);
  if (it == field_.end()) return false;
  if (it->second.type != "UInt8" && it->second.type != "Int8") return false;
  if (it->second.format == "ascii") {
    std::vector<double> d;
    to_double(it->second, &d);
    out->assign(d.begin(), d.end());
    return true;
  }
  *out = decode(it->second);
  return true;
}

bool VtuReader::readCells(std::vector<int64_t>* conn,
                          std::vector<int64_t>* offsets,
                          std::vector<uint8_t>* types) const {
  if (!has_cells_ || ncells_ == 0) return false;
  std::vector<double> c, o, t;
  to_double(conn_, &c);
  to_double(offsets_, &o);
  to_double(types_, &t);
  conn->assign(c.begin(), c.end());
  offsets->assign(o.begin(), o.end());
  types->assign(t.begin(), t.end());
  return true;
}

}  // namespace rw
}  // namespace nlm
