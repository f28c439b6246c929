// Mesh and result files of libehip, host side.
//
// Reading: the ElmerGrid mesh directory that LoadMesh reads through EIO
// (fem/src/MeshUtils.src:313-1009): mesh.header (counts), mesh.nodes
// ("tag partition x y z"), mesh.elements ("tag body type nodes..."),
// mesh.boundary ("tag bc parent1 parent2 type nodes...").  Nodes are
// numbered 1..N in file order and element node references are mapped
// through the node tags, as LoadMesh does with LocalPerm (:493-499,
// :566-569, :789-792).  Nodes per element are type mod 100.
//
// Writing: the same four files (what ElmerGrid writes for a serial mesh),
// and SaveResult's ASCII format 3 (fem/src/ModelDescription.src:2442-2865):
// header, variable table, then per saved step a "Time:" line, the
// variable name, its Perm and the values of the nodes whose Perm > 0.
#include <algorithm>
#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/ehip.h"

namespace ehip {
void set_error(const char *fmt, ...);  // system.hip
}

struct ehip_mesh {
  int nnodes = 0;
  std::vector<double> xyz;          // nnodes * 3
  int max_bulk = 0, max_bnd = 0;    // nodes of the largest element type
  std::vector<int> btype, bbody;    // bulk elements
  std::vector<int> bnodes;          // nbulk * max_bulk, 1-based, 0-padded
  std::vector<int> ftype, ftag, fpar;  // boundary elements; parents 2 per element (0: none)
  std::vector<int> fnodes;          // nbnd * max_bnd
};

namespace {

using ehip::set_error;

bool slurp(const std::string &path, std::string &out) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  out.resize(sz > 0 ? (size_t)sz : 0);
  const size_t got = sz > 0 ? std::fread(&out[0], 1, (size_t)sz, f) : 0;
  std::fclose(f);
  return got == out.size();
}

// Whitespace-separated numbers of a text buffer, line by line.
struct Scan {
  const char *p, *end;
  explicit Scan(const std::string &s) : p(s.data()), end(s.data() + s.size()) {}
  void skip_blank() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
  }
  bool next_line() {  // to the first character of the next non-empty line
    while (p < end && *p != '\n') ++p;
    while (p < end && (*p == '\n' || *p == '\r' || *p == ' ' || *p == '\t')) ++p;
    return p < end;
  }
  bool at_eol() {
    skip_blank();
    return p >= end || *p == '\n';
  }
  bool get(long &v) {
    skip_blank();
    auto r = std::from_chars(p, end, v);
    if (r.ec != std::errc()) return false;
    p = r.ptr;
    // an element type may carry a suffix (e.g. "303e1"): skip to the next blank
    while (p < end && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') ++p;
    return true;
  }
  bool get(double &v) {
    skip_blank();
    const char *q = p;
    bool fortran = false;
    while (q < end && *q != ' ' && *q != '\t' && *q != '\n' && *q != '\r') {
      fortran |= (*q == 'D' || *q == 'd');
      ++q;
    }
    if (!fortran) {
      auto r = std::from_chars(p, q, v);
      if (r.ec != std::errc() || r.ptr != q) return false;
      p = q;
      return true;
    }
    // Fortran exponent letter (1.0D0)
    char buf[64];
    size_t k = 0;
    for (const char *c = p; c < q && k + 1 < sizeof(buf); ++c) buf[k++] = (*c == 'D' || *c == 'd') ? 'e' : *c;
    buf[k] = 0;
    char *e = nullptr;
    v = std::strtod(buf, &e);
    if (e != buf + k) return false;
    p = q;
    return true;
  }
};

int read_mesh(const std::string &dir, ehip_mesh &m) {
  std::string hdr, nod, ele, bnd;
  if (!slurp(dir + "/mesh.header", hdr) || !slurp(dir + "/mesh.nodes", nod) ||
      !slurp(dir + "/mesh.elements", ele)) {
    set_error("ehip_mesh_read: %s: mesh.header, mesh.nodes and mesh.elements are required", dir.c_str());
    return EHIP_EARG;
  }
  const bool has_bnd = slurp(dir + "/mesh.boundary", bnd);
  long nn = 0, ne = 0, nb = 0;
  {
    Scan s(hdr);
    if (!s.get(nn) || !s.get(ne) || !s.get(nb) || nn <= 0 || ne <= 0 || nb < 0) {
      set_error("ehip_mesh_read: bad mesh.header");
      return EHIP_EARG;
    }
  }
  // nodes in file order; tags mapped to positions
  std::vector<long> tag((size_t)nn);
  m.nnodes = (int)nn;
  m.xyz.assign((size_t)nn * 3, 0.0);
  {
    Scan s(nod);
    for (long i = 0; i < nn; ++i) {
      long t, part;
      double x, y, z;
      if (!s.get(t) || !s.get(part) || !s.get(x) || !s.get(y) || !s.get(z)) {
        set_error("ehip_mesh_read: mesh.nodes line %ld unreadable", i + 1);
        return EHIP_EARG;
      }
      tag[i] = t;
      m.xyz[3 * i] = x;
      m.xyz[3 * i + 1] = y;
      m.xyz[3 * i + 2] = z;
      s.next_line();
    }
  }
  const long tmin = *std::min_element(tag.begin(), tag.end());
  const long tmax = *std::max_element(tag.begin(), tag.end());
  if (tmax - tmin + 1 > 4 * nn + 1024) {
    set_error("ehip_mesh_read: node tags span %ld..%ld for %ld nodes", tmin, tmax, nn);
    return EHIP_ECAP;
  }
  std::vector<int> local((size_t)(tmax - tmin + 1), 0);
  for (long i = 0; i < nn; ++i) local[tag[i] - tmin] = (int)(i + 1);
  auto map_node = [&](long t, int &out) {
    if (t < tmin || t > tmax || local[t - tmin] == 0) return false;
    out = local[t - tmin];
    return true;
  };
  // bulk elements: two passes (the widest type sets the row width)
  struct Raw {
    std::vector<int> type, aux, par, nodes;
    std::vector<size_t> start;
  };
  auto read_list = [&](const std::string &text, long count, bool boundary, Raw &r, int &maxn) -> int {
    Scan s(text);
    r.start.push_back(0);
    for (long e = 0; e < count; ++e) {
      long t, a, ty, p1 = 0, p2 = 0;
      if (!s.get(t) || !s.get(a)) return -(int)(e + 1);
      if (boundary && (!s.get(p1) || !s.get(p2))) return -(int)(e + 1);
      if (!s.get(ty)) return -(int)(e + 1);
      const int np = (int)(ty % 100);
      if (ty < 101 || np < 1 || np > 27) return -(int)(e + 1);
      for (int k = 0; k < np; ++k) {
        long v;
        int li;
        if (!s.get(v) || !map_node(v, li)) return -(int)(e + 1);
        r.nodes.push_back(li);
      }
      r.type.push_back((int)ty);
      r.aux.push_back((int)a);
      if (boundary) {
        r.par.push_back((int)std::max(0L, p1));
        r.par.push_back((int)std::max(0L, p2));
      }
      r.start.push_back(r.nodes.size());
      maxn = std::max(maxn, np);
      s.next_line();
    }
    return 0;
  };
  Raw rb, rf;
  int rc = read_list(ele, ne, false, rb, m.max_bulk);
  if (rc) {
    set_error("ehip_mesh_read: mesh.elements element %d unreadable or names an unknown node", -rc);
    return EHIP_EARG;
  }
  if (has_bnd && nb > 0) {
    rc = read_list(bnd, nb, true, rf, m.max_bnd);
    if (rc) {
      set_error("ehip_mesh_read: mesh.boundary element %d unreadable or names an unknown node", -rc);
      return EHIP_EARG;
    }
  }
  auto pack = [](const Raw &r, int width, std::vector<int> &out) {
    const size_t n = r.type.size();
    out.assign(n * (size_t)width, 0);
    for (size_t e = 0; e < n; ++e)
      std::copy(r.nodes.begin() + r.start[e], r.nodes.begin() + r.start[e + 1], out.begin() + e * width);
  };
  m.btype = rb.type;
  m.bbody = rb.aux;
  pack(rb, m.max_bulk, m.bnodes);
  m.ftype = rf.type;
  m.ftag = rf.aux;
  m.fpar = rf.par;
  pack(rf, std::max(m.max_bnd, 1), m.fnodes);
  return EHIP_OK;
}

bool write_text(const std::string &path, const std::string &text) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
  return std::fclose(f) == 0 && ok;
}

void put_int(std::string &s, long v) {
  char b[24];
  auto r = std::to_chars(b, b + sizeof(b), v);
  s.append(b, r.ptr);
}

void put_double(std::string &s, double v) {
  char b[40];
  auto r = std::to_chars(b, b + sizeof(b), v);  // shortest round-trip form
  s.append(b, r.ptr);
}

}  // namespace

using namespace ehip;

extern "C" int ehip_mesh_read(const char *dir, ehip_mesh **out) {
  if (!dir || !out) return EHIP_EARG;
  ehip_mesh *m = new ehip_mesh();
  const int rc = read_mesh(dir, *m);
  if (rc) {
    delete m;
    return rc;
  }
  *out = m;
  return EHIP_OK;
}

extern "C" int ehip_mesh_sizes(const ehip_mesh *m, int *nnodes, int *nbulk, int *nbnd, int *max_npe_bulk,
                               int *max_npe_bnd) {
  if (!m) return EHIP_EARG;
  if (nnodes) *nnodes = m->nnodes;
  if (nbulk) *nbulk = (int)m->btype.size();
  if (nbnd) *nbnd = (int)m->ftype.size();
  if (max_npe_bulk) *max_npe_bulk = m->max_bulk;
  if (max_npe_bnd) *max_npe_bnd = m->max_bnd;
  return EHIP_OK;
}

extern "C" int ehip_mesh_get(const ehip_mesh *m, double *coords, int *bulk_types, int *bulk_bodies, int *bulk_nodes,
                             int *bnd_types, int *bnd_tags, int *bnd_parents, int *bnd_nodes) {
  if (!m) return EHIP_EARG;
  if (coords) std::copy(m->xyz.begin(), m->xyz.end(), coords);
  if (bulk_types) std::copy(m->btype.begin(), m->btype.end(), bulk_types);
  if (bulk_bodies) std::copy(m->bbody.begin(), m->bbody.end(), bulk_bodies);
  if (bulk_nodes) std::copy(m->bnodes.begin(), m->bnodes.end(), bulk_nodes);
  if (bnd_types) std::copy(m->ftype.begin(), m->ftype.end(), bnd_types);
  if (bnd_tags) std::copy(m->ftag.begin(), m->ftag.end(), bnd_tags);
  if (bnd_parents) std::copy(m->fpar.begin(), m->fpar.end(), bnd_parents);
  if (bnd_nodes && m->max_bnd > 0) std::copy(m->fnodes.begin(), m->fnodes.end(), bnd_nodes);
  return EHIP_OK;
}

extern "C" int ehip_mesh_free(ehip_mesh *m) {
  delete m;
  return EHIP_OK;
}

extern "C" int ehip_mesh_write(const char *dir, int nnodes, const double *coords, int nbulk, int bulk_type,
                               const int *bulk_nodes, const int *bulk_bodies, int nbnd, int bnd_type,
                               const int *bnd_nodes, const int *bnd_tags, const int *bnd_parents) {
  if (!dir || nnodes <= 0 || !coords || nbulk <= 0 || !bulk_nodes || bulk_type < 101 || (nbnd > 0 && !bnd_nodes) ||
      (nbnd > 0 && bnd_type < 101)) {
    set_error("ehip_mesh_write: bad arguments");
    return EHIP_EARG;
  }
  const int np = bulk_type % 100, nq = nbnd > 0 ? bnd_type % 100 : 0;
  std::string h, s;
  put_int(h, nnodes); h += ' '; put_int(h, nbulk); h += ' '; put_int(h, nbnd); h += '\n';
  put_int(h, nbnd > 0 && bnd_type != bulk_type ? 2 : 1); h += '\n';
  if (nbnd > 0 && bnd_type != bulk_type) {
    put_int(h, bnd_type); h += ' '; put_int(h, nbnd); h += '\n';
  }
  put_int(h, bulk_type); h += ' '; put_int(h, nbulk + (bnd_type == bulk_type ? nbnd : 0)); h += '\n';
  if (!write_text(std::string(dir) + "/mesh.header", h)) goto fail;
  s.reserve((size_t)nnodes * 64);
  for (int i = 0; i < nnodes; ++i) {
    put_int(s, i + 1);
    s += " -1 ";
    put_double(s, coords[3 * (size_t)i]); s += ' ';
    put_double(s, coords[3 * (size_t)i + 1]); s += ' ';
    put_double(s, coords[3 * (size_t)i + 2]); s += '\n';
  }
  if (!write_text(std::string(dir) + "/mesh.nodes", s)) goto fail;
  s.clear();
  for (int e = 0; e < nbulk; ++e) {
    put_int(s, e + 1); s += ' ';
    put_int(s, bulk_bodies ? bulk_bodies[e] : 1); s += ' ';
    put_int(s, bulk_type);
    for (int k = 0; k < np; ++k) { s += ' '; put_int(s, bulk_nodes[(size_t)e * np + k]); }
    s += '\n';
  }
  if (!write_text(std::string(dir) + "/mesh.elements", s)) goto fail;
  s.clear();
  for (int e = 0; e < nbnd; ++e) {
    put_int(s, e + 1); s += ' ';
    put_int(s, bnd_tags ? bnd_tags[e] : 1); s += ' ';
    put_int(s, bnd_parents ? bnd_parents[2 * (size_t)e] : 0); s += ' ';
    put_int(s, bnd_parents ? bnd_parents[2 * (size_t)e + 1] : 0); s += ' ';
    put_int(s, bnd_type);
    for (int k = 0; k < nq; ++k) { s += ' '; put_int(s, bnd_nodes[(size_t)e * nq + k]); }
    s += '\n';
  }
  if (!write_text(std::string(dir) + "/mesh.boundary", s)) goto fail;
  return EHIP_OK;
fail:
  set_error("ehip_mesh_write: cannot write into %s", dir);
  return EHIP_EARG;
}

extern "C" int ehip_result_write(const char *path, const char *var, const char *equation, int nnodes, const int *perm,
                                 int dofs, const double *values, int append, int step, double time) {
  if (!path || !var || nnodes <= 0 || !values || dofs != 1) {
    set_error("ehip_result_write: bad arguments (one scalar variable per call)");
    return EHIP_EARG;
  }
  FILE *f = std::fopen(path, append ? "ab" : "wb");
  if (!f) {
    set_error("ehip_result_write: cannot open %s", path);
    return EHIP_EARG;
  }
  int nv = 0;
  for (int i = 0; i < nnodes; ++i) nv = std::max(nv, perm ? perm[i] : i + 1);
  if (!append) {
    char date[64];
    const std::time_t now = std::time(nullptr);
    std::strftime(date, sizeof(date), "%Y/%m/%d %H:%M:%S", std::localtime(&now));
    std::fprintf(f, " ASCII 3\n!File started at: %s\n Degrees of freedom: \n", date);
    if (std::strlen(var) <= 25)
      std::fprintf(f, "%-24s : %8d%8d%4d : %s\n", var, nv, perm ? nnodes : 0, dofs, equation ? equation : "no equation");
    else
      std::fprintf(f, "%s : %8d%8d%4d : %s\n", var, nv, perm ? nnodes : 0, dofs, equation ? equation : "no equation");
    std::fprintf(f, " Total DOFs:            1\n Number Of Nodes: %12d\n", nnodes);
  }
  {  // Fortran ES17.8E3: d.dddddddd, then E, sign and a three-digit exponent
    char m[32];
    std::snprintf(m, sizeof(m), "%.8E", time);
    char *e = std::strchr(m, 'E');
    const int ex = e ? std::atoi(e + 1) : 0;
    if (e) *e = 0;
    char es[48];
    std::snprintf(es, sizeof(es), "%sE%c%03d", m, ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
    std::fprintf(f, "Time: %7d%7d%17s\n", step, step, es);
  }
  std::fprintf(f, "%s\n", var);
  if (perm) {
    int cnt = 0;
    for (int i = 0; i < nnodes; ++i) cnt += perm[i] > 0;
    std::fprintf(f, "Perm: %12d %12d\n", nnodes, cnt);
    for (int i = 0; i < nnodes; ++i)
      if (perm[i] > 0) std::fprintf(f, "%11d%11d\n", i + 1, perm[i]);
  } else {
    std::fprintf(f, "Perm: NULL\n");
  }
  for (int i = 0; i < nnodes; ++i) {
    const int k = perm ? perm[i] : i + 1;
    if (k > 0) std::fprintf(f, " %23.15E\n", values[k - 1]);
  }
  const bool ok = std::fclose(f) == 0;
  if (!ok) {
    set_error("ehip_result_write: write to %s failed", path);
    return EHIP_EARG;
  }
  return EHIP_OK;
}
