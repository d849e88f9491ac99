// FEMSpace.h -- drop-in FE space of the assemble-and-solve path.
//
// Mirrors the interface the reference path uses (include/AFEPack/
// {Geometry,EasyMesh,TemplateElement,FEMSpace}.h): meshes, template elements
// described by their data files (triangle.1 / triangle.2 / tetrahedron.1),
// FEMSpace::buildElement/buildDof/buildDofBoundaryMark and the Element
// queries a getElementMatrix override needs.  The DOF numbering and pattern
// come from the C ABI (afe_space_create); basis functions are evaluated on the
// host here only for user-written element matrices.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../csrc/element_geom.hpp"
#include "../../csrc/quadrature.hpp"
#include "Miscellaneous.h"
#include "SparseMatrix.h"

namespace AFEPack {

// ---- device context shared by the drop-in classes -------------------------
namespace detail {
inline afe_ctx *device() {
  static afe_ctx *ctx = nullptr;
  if (!ctx) {
    const char *d = getenv("AFE_DEVICE");
    if (afe_ctx_create(d ? atoi(d) : 0, &ctx) != AFE_OK) {
      std::cerr << "afepack_amd: " << afe_last_error() << std::endl;
      abort();
    }
  }
  return ctx;
}

inline void check(int rc) {
  if (rc != AFE_OK) throw std::runtime_error(afe_last_error());
}

// where the reference prints a message and calls abort() (e.g. lazyReinit
// with a foreign pattern, src/AMGSolver.cpp:509-513), the drop-in prints the
// same message and calls this hook; it aborts unless a program installs its
// own handler (which must not return)
using fatal_handler = void (*)(const char *msg);
inline fatal_handler &fatal_hook() {
  static fatal_handler h = nullptr;
  return h;
}
[[noreturn]] inline void fatal(const char *msg) {
  std::cerr << msg << std::endl;
  if (fatal_hook()) fatal_hook()(msg);
  abort();
}

// HBM vector owned by the drop-in layer
struct DeviceVector {
  double *p = nullptr;
  size_t n = 0;
  explicit DeviceVector(size_t count) : n(count) {
    check(afe_dev_alloc(device(), (int64_t)(n * sizeof(double)), (void **)&p));
  }
  DeviceVector(const DeviceVector &) = delete;
  DeviceVector &operator=(const DeviceVector &) = delete;
  ~DeviceVector() {
    if (p) afe_dev_free(device(), p);
  }
  void upload(const double *h) { check(afe_dev_upload(device(), p, h, (int64_t)(n * sizeof(double)))); }
  void download(double *h) const { check(afe_dev_download(device(), h, p, (int64_t)(n * sizeof(double)))); }
};
} // namespace detail

template <int DIM> struct Point {
  double x[DIM] = {};
  Point() = default;
  double &operator[](int i) { return x[i]; }
  double operator[](int i) const { return x[i]; }
  operator const double *() const { return x; }
};

// ---- meshes -----------------------------------------------------------------
template <int DIM, int DOW = DIM> class Mesh {
public:
  int n_point() const { return (int)(coords.size() / DOW); }
  int n_element() const { return (int)(elem.size() / (DIM + 1)); }
  int n_geometry(int dim) const {
    return dim == 0 ? n_point() : dim == DIM ? n_element() : dim == 1 ? (int)side_mark.size() : 0;
  }
  std::vector<double> coords;      // vertex coordinates, n_point x DOW
  std::vector<int32_t> elem;       // element vertices, n_element x (DIM+1)
  std::vector<int32_t> vmark;      // boundary mark per vertex
  std::vector<int32_t> side_vert;  // 1-dimensional geometries (sides / edges), 2 vertices each
  std::vector<int32_t> side_mark;  // and their boundary marks (empty: none read)

  // Mesh::readData (Geometry.templates.h:417-423, 470-497): the text format
  // of Mesh::writeData -- the points, then for each dimension 0..DIM the
  // number of geometries and each geometry as "index, vertex count,
  // vertices, boundary count, boundaries, boundary mark" (GeometryBM,
  // src/Geometry.cpp:111-125,174-179).  Geometries may come in any order
  // (each is stored at its index).  Vertices here are the 0-dimensional
  // geometries, placed at the point each one names.  As the reference, the
  // element orientation is kept as read.
  void readData(const std::string &filename) {
    std::ifstream in(filename);
    if (!in) throw std::runtime_error("cannot open " + filename);
    int np = 0;
    in >> np;
    std::vector<double> pts((size_t)np * DOW);
    for (auto &x : pts) in >> x;
    struct G {
      std::vector<int32_t> v;
      int32_t bm = 0;
      bool seen = false;
    };
    std::vector<std::vector<G>> geo(DIM + 1);
    for (int d = 0; d <= DIM; d++) {
      int n = 0;
      in >> n;
      if (!in || n < 0) throw std::runtime_error("bad geometry count in " + filename);
      geo[d].resize(n);
      for (int k = 0; k < n; k++) {
        int idx = 0, nv = 0, nb = 0;
        in >> idx >> nv;
        if (!in || idx < 0 || idx >= n || nv < 0) throw std::runtime_error("bad geometry record in " + filename);
        G &g = geo[d][idx];
        g.v.resize(nv);
        for (auto &x : g.v) in >> x;
        in >> nb;
        for (int b = 0, t = 0; b < nb; b++) in >> t;
        in >> g.bm;
        g.seen = true;
      }
      for (const G &g : geo[d])
        if (!g.seen) throw std::runtime_error("a geometry index is missing in " + filename);
    }
    if (!in) throw std::runtime_error("truncated mesh file " + filename);
    const int nv = (int)geo[0].size();
    coords.resize((size_t)nv * DOW);
    vmark.resize(nv);
    for (int i = 0; i < nv; i++) {
      const G &g = geo[0][i];
      if (g.v.size() != 1 || g.v[0] < 0 || g.v[0] >= np) throw std::runtime_error("bad vertex geometry");
      for (int c = 0; c < DOW; c++) coords[(size_t)i * DOW + c] = pts[(size_t)g.v[0] * DOW + c];
      vmark[i] = g.bm;
    }
    side_vert.clear();
    side_mark.clear();
    for (const G &g : geo[1]) {
      if (g.v.size() != 2) throw std::runtime_error("a 1-dimensional geometry needs 2 vertices");
      side_vert.push_back(g.v[0]);
      side_vert.push_back(g.v[1]);
      side_mark.push_back(g.bm);
    }
    elem.clear();
    for (const G &g : geo[DIM]) {
      if (g.v.size() != (size_t)(DIM + 1)) throw std::runtime_error("only simplex elements are supported");
      for (int32_t x : g.v) {
        if (x < 0 || x >= nv) throw std::runtime_error("element vertex out of range");
        elem.push_back(x);
      }
    }
  }
};

// EasyMesh .n/.s/.e reader (EasyMesh.templates.h:16-94): triangles stored
// clockwise are flipped to counter-clockwise; node marks give the vertex DOF
// marks and side marks the P2 edge DOF marks
class EasyMesh : public Mesh<2, 2> {
public:
  void readData(const std::string &base) {
    std::ifstream in(base + ".n");
    if (!in) throw std::runtime_error("cannot open " + base + ".n");
    int nn, ne, ns;
    in >> nn >> ne >> ns;
    std::string rest;
    std::getline(in, rest);
    coords.resize(2 * (size_t)nn);
    vmark.resize(nn);
    for (int i = 0; i < nn; i++) {
      int idx;
      in >> idx >> coords[2 * i] >> coords[2 * i + 1] >> vmark[i];
    }
    std::ifstream is(base + ".s");
    if (!is) throw std::runtime_error("cannot open " + base + ".s");
    int ns2;
    is >> ns2;
    side_vert.assign(2 * (size_t)ns2, 0);
    side_mark.assign(ns2, 0);
    for (int i = 0; i < ns2; i++) {
      int idx, a, b, ea, eb, m;
      is >> idx >> a >> b >> ea >> eb >> m;
      if (idx < 0 || idx >= ns2) throw std::runtime_error("bad side index in " + base + ".s");
      side_vert[2 * idx] = a;
      side_vert[2 * idx + 1] = b;
      side_mark[idx] = m;
    }
    std::ifstream ie(base + ".e");
    if (!ie) throw std::runtime_error("cannot open " + base + ".e");
    int ne2, nn2, ns3;
    ie >> ne2 >> nn2 >> ns3;
    std::getline(ie, rest);
    elem.resize(3 * (size_t)ne2);
    for (int i = 0; i < ne2; i++) {
      int idx, v0, v1, v2, n0, n1, n2, s0, s1, s2;
      ie >> idx >> v0 >> v1 >> v2 >> n0 >> n1 >> n2 >> s0 >> s1 >> s2;
      const double *p0 = &coords[2 * v0], *p1 = &coords[2 * v1], *p2 = &coords[2 * v2];
      if ((p1[0] - p0[0]) * (p2[1] - p0[1]) - (p1[1] - p0[1]) * (p2[0] - p0[0]) < 0) std::swap(v1, v2);
      elem[3 * i] = v0;
      elem[3 * i + 1] = v1;
      elem[3 * i + 2] = v2;
    }
  }
};

// structured meshes of the benchmark problems (no counterpart in the
// reference, which reads mesh files): n x n squares split in two triangles,
// or n^3 cubes split in six tetrahedra, boundary vertices marked 1
inline void buildSquareMesh(Mesh<2, 2> &m, int n, double jitter = 0.0, uint64_t seed = 0) {
  const size_t nv = (size_t)(n + 1) * (n + 1), ne = 2 * (size_t)n * n;
  m.coords.resize(2 * nv);
  m.elem.resize(3 * ne);
  m.vmark.resize(nv);
  detail::check(afe_mesh_square(n, jitter, seed, m.coords.data(), m.elem.data(), m.vmark.data()));
}
inline void buildCubeMesh(Mesh<3, 3> &m, int n, double jitter = 0.0, uint64_t seed = 0) {
  const size_t nv = (size_t)(n + 1) * (n + 1) * (n + 1), ne = 6 * (size_t)n * n * n;
  m.coords.resize(3 * nv);
  m.elem.resize(4 * ne);
  m.vmark.resize(nv);
  detail::check(afe_mesh_cube(n, jitter, seed, m.coords.data(), m.elem.data(), m.vmark.data()));
}

// ---- template elements --------------------------------------------------------
// The reference reads geometry / DOF / basis descriptions from data files and
// dlopens their plugins (TemplateElement.templates.h).  Here readData looks
// the file up as the reference does (the directories of
// AFEPACK_TEMPLATE_PATH, then ".", src/Miscellaneous.cpp:145-185) and, when
// it is found, reads what selects the element: the template geometry's
// vertex count, the DOF layout per geometry dimension and the basis
// functions' count and polynomial order.  TemplateElement::reinit picks the
// supported element from that content (any file name), and fails loudly on
// content no supported element has.  A file that is not found leaves only
// its name, and the name then selects the element (triangle.1,
// triangle.2, tetrahedron.1), so that programs run where the template
// directory is absent; the numerics of the three elements live in the
// library either way.
namespace detail {
inline std::string find_template_file(const std::string &name) {
  std::vector<std::string> dirs;
  if (const char *p = std::getenv("AFEPACK_TEMPLATE_PATH")) {
    std::string all = p, cur;
    for (char c : all) {
      if (c == ':') {
        if (!cur.empty()) dirs.push_back(cur);
        cur.clear();
      } else {
        cur += c;
      }
    }
    if (!cur.empty()) dirs.push_back(cur);
  }
  dirs.push_back(".");
  for (const auto &d : dirs) {
    const std::string f = d + "/" + name;
    std::ifstream in(f);
    if (in) return f;
  }
  return "";
}

// whitespace-separated tokens with '#' comments (to the end of the line)
// removed, as the reference's filtered stream reads these files
inline std::vector<std::string> template_tokens(const std::string &path) {
  std::ifstream in(path);
  std::vector<std::string> tok;
  std::string line;
  while (std::getline(in, line)) {
    const size_t h = line.find('#');
    if (h != std::string::npos) line.resize(h);
    std::istringstream ls(line);
    std::string t;
    while (ls >> t) tok.push_back(t);
  }
  return tok;
}

inline int template_int(const std::vector<std::string> &tok, size_t i, const std::string &file) {
  if (i >= tok.size()) throw std::runtime_error("template file ends early: " + file);
  try {
    return std::stoi(tok[i]);
  } catch (const std::exception &) {
    throw std::runtime_error("template file: integer expected at token " + std::to_string(i) + " of " + file);
  }
}
} // namespace detail

// TemplateGeometry::readData (.tmp_geo, TemplateElement.templates.h): the
// volume function's library and name, then the vertex count and vertices
template <int TDIM> class TemplateGeometry {
public:
  void readData(const std::string &f) {
    file = f;
    n_vertex = -1;
    const std::string path = detail::find_template_file(f);
    if (path.empty()) return;
    const auto tok = detail::template_tokens(path);
    n_vertex = detail::template_int(tok, 2, path);
  }
  std::string file;
  int n_vertex = -1; // -1: no file read
};
template <int TDIM, int DOW = TDIM> class CoordTransform {
public:
  void readData(const std::string &f) { file = f; }
  std::string file;
};
// TemplateDOF::readData (.tmp_dof, TemplateElement.templates.h:420-430): the
// entry count, then (dimension, geometry index, DOF count) per entry
template <int TDIM> class TemplateDOF {
public:
  TemplateDOF() = default;
  explicit TemplateDOF(TemplateGeometry<TDIM> &) {}
  void reinit(TemplateGeometry<TDIM> &) {}
  void readData(const std::string &f) {
    file = f;
    n_dof = -1;
    per_dim.fill(0);
    const std::string path = detail::find_template_file(f);
    if (path.empty()) return;
    const auto tok = detail::template_tokens(path);
    const int n = detail::template_int(tok, 0, path);
    if (n < 0) throw std::runtime_error("negative DOF entry count: " + path);
    n_dof = 0;
    for (int e = 0; e < n; e++) {
      const int d = detail::template_int(tok, 1 + 3 * e, path), c = detail::template_int(tok, 3 + 3 * e, path);
      if (d < 0 || d > TDIM) throw std::runtime_error("template DOF of dimension " + std::to_string(d) + ": " + path);
      per_dim[d] += c;
      n_dof += c;
    }
  }
  std::string file;
  int n_dof = -1; // -1: no file read
  std::array<int, 4> per_dim{};
};
// BasisFunctionAdmin::readData (.bas_fun): the basis count, then per basis
// its DOF's geometry (dimension, index), the interpolation point, the
// identity (order, TDIM exponents, flag), the plugin and its two functions
template <class value_type, int DOW, int TDIM = DOW> class BasisFunctionAdmin {
public:
  BasisFunctionAdmin() = default;
  explicit BasisFunctionAdmin(TemplateDOF<TDIM> &) {}
  void reinit(TemplateDOF<TDIM> &) {}
  void readData(const std::string &f) {
    file = f;
    n_basis = -1;
    max_order = 0;
    const std::string path = detail::find_template_file(f);
    if (path.empty()) return;
    const auto tok = detail::template_tokens(path);
    const int n = detail::template_int(tok, 0, path);
    if (n < 0) throw std::runtime_error("negative basis count: " + path);
    const size_t per = 2 + TDIM + (TDIM + 2) + 3;
    for (int b = 0; b < n; b++)
      max_order = std::max(max_order, detail::template_int(tok, 1 + b * per + 2 + TDIM, path));
    if (tok.size() < 1 + (size_t)n * per) throw std::runtime_error("template file ends early: " + path);
    n_basis = n;
  }
  std::string file;
  int n_basis = -1; // -1: no file read
  int max_order = 0;
};

template <int TDIM> struct QuadratureInfo {
  const afe::QuadRule *rule = nullptr;
  int n_quadraturePoint() const { return rule->n; }
  double weight(int l) const { return afe::kQuadData[rule->off + l * (TDIM + 1) + TDIM]; }
  Point<TDIM> quadraturePoint(int l) const {
    Point<TDIM> p;
    for (int d = 0; d < TDIM; d++) p[d] = afe::kQuadData[rule->off + l * (TDIM + 1) + d];
    return p;
  }
  std::vector<Point<TDIM>> quadraturePoint() const {
    std::vector<Point<TDIM>> v(rule->n);
    for (int l = 0; l < rule->n; l++) v[l] = quadraturePoint(l);
    return v;
  }
};

template <class value_type, int DOW, int TDIM = DOW> class TemplateElement {
public:
  void reinit(TemplateGeometry<TDIM> &g, TemplateDOF<TDIM> &d, CoordTransform<TDIM, DOW> &,
              BasisFunctionAdmin<value_type, DOW, TDIM> &b) {
    if (b.n_basis >= 0 || d.n_dof >= 0 || g.n_vertex >= 0) {
      kind = from_content(g, d, b);
      return;
    }
    const std::string name = b.file.empty() ? d.file : b.file;
    if (TDIM == 3 && name.find("tetrahedron.1") != std::string::npos)
      kind = AFE_P1_TET;
    else if (TDIM == 2 && name.find("triangle.2") != std::string::npos)
      kind = AFE_P2_TRI;
    else if (TDIM == 2 && name.find("triangle.1") != std::string::npos)
      kind = AFE_P1_TRI;
    else
      throw std::runtime_error("template element not supported on this path: " + name + " / " + g.file);
  }
  double volume() const { return TDIM == 2 ? 0.5 : 1.0 / 6.0; }
  QuadratureInfo<TDIM> findQuadratureInfo(int acc) const {
    QuadratureInfo<TDIM> q;
    q.rule = afe::quad_find(TDIM, acc);
    if (!q.rule) throw std::runtime_error("no quadrature rule of that accuracy");
    return q;
  }
  int n_dof() const { return kind == AFE_P1_TRI ? 3 : kind == AFE_P2_TRI ? 6 : 4; }
  int kind = 0;

private:
  // the supported elements as their data files describe them: a simplex
  // (TDIM + 1 vertices); P1: one DOF per vertex, TDIM + 1 bases of order 1;
  // P2 (triangles): one DOF per vertex and per edge, 6 bases of order 2
  static int from_content(const TemplateGeometry<TDIM> &g, const TemplateDOF<TDIM> &d,
                          const BasisFunctionAdmin<value_type, DOW, TDIM> &b) {
    const std::string what = g.file + " / " + d.file + " / " + b.file;
    if (g.n_vertex >= 0 && g.n_vertex != TDIM + 1)
      throw std::runtime_error("template element not supported on this path (not a simplex): " + what);
    if (d.n_dof >= 0 && b.n_basis >= 0 && d.n_dof != b.n_basis)
      throw std::runtime_error("template DOF and basis files disagree on the DOF count: " + what);
    const int n = b.n_basis >= 0 ? b.n_basis : d.n_dof;
    const bool vertex_dofs = d.n_dof < 0 || d.per_dim[0] == TDIM + 1;
    if (n == TDIM + 1 && vertex_dofs && (d.n_dof < 0 || d.per_dim[1] == 0) && (b.n_basis < 0 || b.max_order == 1))
      return TDIM == 2 ? AFE_P1_TRI : AFE_P1_TET;
    if (TDIM == 2 && n == 6 && vertex_dofs && (d.n_dof < 0 || d.per_dim[1] == 3) && (b.n_basis < 0 || b.max_order == 2))
      return AFE_P2_TRI;
    throw std::runtime_error("template element not supported on this path: " + what);
  }
};

template <class value_type, int DIM, int DOW = DIM, int TDIM = DIM> class FEMSpace;

template <class value_type, int DIM, int DOW = DIM, int TDIM = DIM> class Element {
public:
  using space_t = FEMSpace<value_type, DIM, DOW, TDIM>;
  void reinit(space_t &sp, int index, int tmpl) {
    sp_ = &sp;
    idx_ = index;
    tmpl_ = tmpl;
  }
  int index() const { return idx_; }
  const std::vector<int> &dof() const { return dof_; }
  std::vector<int> &dof() { return dof_; }
  const TemplateElement<value_type, DOW, TDIM> &templateElement() const;
  QuadratureInfo<TDIM> findQuadratureInfo(int acc) const { return templateElement().findQuadratureInfo(acc); }
  // physical vertex coordinates, v[j][d]
  void vertices(double v[][DOW]) const;
  // CoordTransform::local_to_global at reference points
  std::vector<Point<DOW>> local_to_global(const std::vector<Point<TDIM>> &lp) const {
    double v[TDIM + 1][DOW];
    vertices(v);
    std::vector<Point<DOW>> out(lp.size());
    for (size_t q = 0; q < lp.size(); q++) afe::local_to_global<DOW>(lp[q].x, v, out[q].x);
    return out;
  }
  // CoordTransform::local_to_global_jacobian (constant on a simplex)
  std::vector<double> local_to_global_jacobian(const std::vector<Point<TDIM>> &lp) const {
    double v[TDIM + 1][DOW];
    vertices(v);
    return std::vector<double>(lp.size(), afe::l2g_jacobian<DOW>(v));
  }
  // basis values [basis][point] at physical points
  std::vector<std::vector<value_type>> basis_function_value(const std::vector<Point<DOW>> &p) const {
    double v[TDIM + 1][DOW], val[6];
    vertices(v);
    const int nl = templateElement().n_dof();
    std::vector<std::vector<value_type>> out(nl, std::vector<value_type>(p.size()));
    for (size_t q = 0; q < p.size(); q++) {
      eval_val(p[q].x, v, val);
      for (int j = 0; j < nl; j++) out[j][q] = val[j];
    }
    return out;
  }
  // basis gradients [basis][point][component] at physical points
  std::vector<std::vector<std::vector<value_type>>> basis_function_gradient(const std::vector<Point<DOW>> &p) const {
    double v[TDIM + 1][DOW], g[6][DOW];
    vertices(v);
    const int nl = templateElement().n_dof();
    std::vector<std::vector<std::vector<value_type>>> out(
        nl, std::vector<std::vector<value_type>>(p.size(), std::vector<value_type>(DOW)));
    for (size_t q = 0; q < p.size(); q++) {
      eval_grad(p[q].x, v, g);
      for (int j = 0; j < nl; j++)
        for (int d = 0; d < DOW; d++) out[j][q][d] = g[j][d];
    }
    return out;
  }

private:
  void eval_val(const double *p, const double v[][DOW], double *val) const {
    if constexpr (DOW == 3) {
      afe::basis_val<afe::P1_TET>(p, v, val);
    } else if (templateElement().kind == AFE_P2_TRI) {
      afe::basis_val<afe::P2_TRI>(p, v, val);
    } else {
      afe::basis_val<afe::P1_TRI>(p, v, val);
    }
  }
  void eval_grad(const double *p, const double v[][DOW], double g[][DOW]) const {
    if constexpr (DOW == 3) {
      afe::basis_grad<afe::P1_TET>(p, v, g);
    } else if (templateElement().kind == AFE_P2_TRI) {
      afe::basis_grad<afe::P2_TRI>(p, v, g);
    } else {
      afe::basis_grad<afe::P1_TRI>(p, v, g);
    }
  }
  space_t *sp_ = nullptr;
  int idx_ = 0, tmpl_ = 0;
  std::vector<int> dof_;
};

template <class value_type, int DIM, int DOW, int TDIM> class FEMSpace {
public:
  typedef int bmark_t;
  using element_t = Element<value_type, DIM, DOW, TDIM>;
  using template_t = TemplateElement<value_type, DOW, TDIM>;
  using ElementIterator = typename std::vector<element_t>::iterator;
  using ConstElementIterator = typename std::vector<element_t>::const_iterator;
  FEMSpace() = default;
  FEMSpace(Mesh<DIM, DOW> &m, std::vector<template_t> &t) { reinit(m, t); }
  ~FEMSpace() {
    if (h_) afe_space_destroy(h_);
  }
  FEMSpace(const FEMSpace &) = delete;
  FEMSpace &operator=(const FEMSpace &) = delete;
  void reinit(Mesh<DIM, DOW> &m, std::vector<template_t> &t) {
    mesh_ = &m;
    tmpl_ = &t;
  }
  std::vector<element_t> &element() { return elem_; }
  element_t &element(int i) { return elem_[i]; }
  const element_t &element(int i) const { return elem_[i]; }
  ElementIterator beginElement() { return elem_.begin(); }
  ElementIterator endElement() { return elem_.end(); }
  ConstElementIterator beginElement() const { return elem_.begin(); }
  ConstElementIterator endElement() const { return elem_.end(); }
  int n_element() const { return (int)elem_.size(); }
  Mesh<DIM, DOW> &mesh() const { return *mesh_; }
  const template_t &templateElement(int i = 0) const { return (*tmpl_)[i]; }
  void buildElement(bool = false) {}
  // FEMSpace::buildDof (FEMSpace.templates.h:705-880) through afe_space_create
  void buildDof() {
    if (h_) afe_space_destroy(h_);
    const int kind = (*tmpl_)[0].kind;
    // the side marks, when the mesh file had sides, mark the P2 edge DOFs
    detail::check(afe_space_create_sides(detail::device(), kind, mesh_->n_point(), mesh_->coords.data(),
                                         mesh_->n_element(), mesh_->elem.data(), mesh_->vmark.data(),
                                         (int64_t)mesh_->side_mark.size(),
                                         mesh_->side_vert.empty() ? nullptr : mesh_->side_vert.data(),
                                         mesh_->side_mark.empty() ? nullptr : mesh_->side_mark.data(), &h_));
    int64_t nd, nnz;
    int32_t nl;
    detail::check(afe_space_info(h_, &nd, &nnz, &nl));
    n_dof_ = (int)nd;
    std::vector<int32_t> ed((size_t)n_element() * nl);
    mark_.resize(nd);
    point_.resize((size_t)nd * DOW);
    detail::check(afe_space_dofs(h_, ed.data(), mark_.data(), point_.data()));
    for (int e = 0; e < n_element(); e++) elem_[e].dof().assign(ed.begin() + (size_t)e * nl, ed.begin() + (size_t)(e + 1) * nl);
    uploaded_ = mesh_->coords;
  }
  void buildDofBoundaryMark() {}
  int n_dof() const { return n_dof_; }
  int dofBoundaryMark(int i) const { return mark_[i]; }
  Point<DOW> dofInterpPoint(int i) const {
    Point<DOW> p;
    for (int d = 0; d < DOW; d++) p[d] = point_[(size_t)i * DOW + d];
    return p;
  }
  // the device space, with the mesh's current vertex coordinates: a moved
  // mesh (same connectivity, as in the moving-mesh solvers) is pushed to
  // the device before the next device use, so the pattern, the gather map
  // and any lazy AMG hierarchy built on this space stay valid
  afe_space *handle() const {
    if (h_ && mesh_ && mesh_->coords != uploaded_) {
      detail::check(afe_space_update_coords(h_, mesh_->coords.data()));
      uploaded_ = mesh_->coords;
    }
    return h_;
  }

private:
  mutable std::vector<double> uploaded_;
  Mesh<DIM, DOW> *mesh_ = nullptr;
  std::vector<template_t> *tmpl_ = nullptr;
  std::vector<element_t> elem_;
  afe_space *h_ = nullptr;
  int n_dof_ = 0;
  std::vector<int32_t> mark_;
  std::vector<double> point_;
};

template <class value_type, int DIM, int DOW, int TDIM>
const TemplateElement<value_type, DOW, TDIM> &Element<value_type, DIM, DOW, TDIM>::templateElement() const {
  return sp_->templateElement(tmpl_);
}

template <class value_type, int DIM, int DOW, int TDIM>
void Element<value_type, DIM, DOW, TDIM>::vertices(double v[][DOW]) const {
  const auto &m = sp_->mesh();
  for (int j = 0; j <= TDIM; j++)
    for (int d = 0; d < DOW; d++) v[j][d] = m.coords[(size_t)m.elem[(size_t)idx_ * (TDIM + 1) + j] * DOW + d];
}

template <class value_type, int DIM> class FEMFunction : public Vector<double> {
public:
  FEMFunction() = default;
  explicit FEMFunction(FEMSpace<value_type, DIM> &sp) { reinit(sp); }
  void reinit(FEMSpace<value_type, DIM> &sp) {
    sp_ = &sp;
    Vector<double>::reinit(sp.n_dof());
  }
  FEMSpace<value_type, DIM> &femSpace() const { return *sp_; }
  // OpenDX field of the values at the mesh vertices (vertex DOFs), one
  // simplex per element
  void writeOpenDXData(const std::string &file) const {
    const auto &m = sp_->mesh();
    std::vector<double> vv(m.n_point(), 0.0);
    for (int e = 0; e < sp_->n_element(); e++) {
      const auto &d = sp_->element(e).dof();
      for (int j = 0; j <= DIM; j++) vv[m.elem[(size_t)e * (DIM + 1) + j]] = (*this)(d[j]);
    }
    std::ofstream os(file);
    os.precision(12);
    os << "object 1 class array type float rank 1 shape " << DIM << " item " << m.n_point() << " data follows\n";
    for (int i = 0; i < m.n_point(); i++) {
      for (int d = 0; d < DIM; d++) os << m.coords[(size_t)i * DIM + d] << (d + 1 < DIM ? " " : "\n");
    }
    os << "object 2 class array type int rank 1 shape " << DIM + 1 << " item " << m.n_element() << " data follows\n";
    for (int e = 0; e < m.n_element(); e++)
      for (int j = 0; j <= DIM; j++) os << m.elem[(size_t)e * (DIM + 1) + j] << (j < DIM ? " " : "\n");
    os << "attribute \"element type\" string \"" << (DIM == 2 ? "triangles" : "tetrahedra") << "\"\n";
    os << "attribute \"ref\" string \"positions\"\n";
    os << "object 3 class array type float rank 0 item " << m.n_point() << " data follows\n";
    for (int i = 0; i < m.n_point(); i++) os << vv[i] << "\n";
    os << "attribute \"dep\" string \"positions\"\n";
    os << "object \"FEMFunction\" class field\ncomponent \"positions\" value 1\n"
       << "component \"connections\" value 2\ncomponent \"data\" value 3\nend\n";
  }

private:
  FEMSpace<value_type, DIM> *sp_ = nullptr;
};

// ---- Dirichlet boundary conditions (FEMSpace.h:380-520,
// FEMSpace.templates.h:1303-1373 of the reference) --------------------------
struct BoundaryConditionInfo {
  typedef int Type;
  static constexpr Type DIRICHLET = 1, NEUMANN = 2, ROBIN = 3;
};

template <class value_type, int DIM, int DOW = DIM, int TDIM = DIM, typename Number = double>
class BoundaryCondition : public BoundaryConditionInfo {
public:
  typedef int bmark_t;
  BoundaryCondition() = default;
  BoundaryCondition(Type t, bmark_t m) : type_(t), mark_(m) {}
  virtual ~BoundaryCondition() = default;
  const Type &boundaryType() const { return type_; }
  Type &boundaryType() { return type_; }
  const bmark_t &boundaryMark() const { return mark_; }
  bmark_t &boundaryMark() { return mark_; }
  virtual Number value(const Point<DOW> &) const { return 0; }

private:
  Type type_ = DIRICHLET;
  bmark_t mark_ = 0;
};

template <class value_type, int DIM, int DOW = DIM, int TDIM = DIM, typename Number = double>
class BoundaryFunction : public BoundaryCondition<value_type, DIM, DOW, TDIM, Number> {
  using base = BoundaryCondition<value_type, DIM, DOW, TDIM, Number>;

public:
  BoundaryFunction(typename base::Type t, typename base::bmark_t m, const Function<Number> &f) : base(t, m), f_(&f) {}
  BoundaryFunction(typename base::Type t, typename base::bmark_t m, value_type (*fun)(const double *),
                   std::vector<value_type> (*grad)(const double *) = nullptr)
      : base(t, m), own_(new FunctionFunction<Number>(fun, grad)), f_(own_.get()) {}
  Number value(const Point<DOW> &p) const override { return f_->value(p); }

private:
  std::shared_ptr<Function<Number>> own_;
  const Function<Number> *f_;
};

template <class value_type, int DIM, int DOW = DIM, int TDIM = DIM, typename Number = double>
class BoundaryConditionAdmin {
public:
  using bc_t = BoundaryCondition<value_type, DIM, DOW, TDIM, Number>;
  explicit BoundaryConditionAdmin(const FEMSpace<value_type, DIM, DOW, TDIM> &sp) : sp_(&sp) {}
  void reinit(const FEMSpace<value_type, DIM, DOW, TDIM> &sp) {
    sp_ = &sp;
    bcs_.clear();
  }
  void add(const bc_t &b) {
    if (b.boundaryType() != BoundaryConditionInfo::DIRICHLET)
      throw std::runtime_error("only Dirichlet boundary conditions can be applied");
    if (b.boundaryMark() < 0) throw std::runtime_error("boundary marks must be non-negative");
    for (const bc_t *c : bcs_)
      if (c->boundaryMark() == b.boundaryMark()) throw std::runtime_error("a condition for this mark exists already");
    bcs_.push_back(&b);
  }
  // the elimination on the device copy of the matrix (FEMSpace.templates.h:
  // 1338-1373); preserve_symmetry = false changes the boundary rows only
  void apply(SparseMatrix<double> &A, Vector<double> &u, Vector<double> &f, bool preserve_symmetry = true) {
    const size_t n = (size_t)sp_->n_dof();
    if (A.m() != n || A.n() != n || u.size() != n || f.size() != n) throw std::runtime_error("apply: size mismatch");
    std::vector<int32_t> marks;
    std::vector<double> g(n, 0.0);
    for (const bc_t *c : bcs_) marks.push_back(c->boundaryMark());
    for (size_t i = 0; i < n; i++)
      for (const bc_t *c : bcs_)
        if (c->boundaryMark() == sp_->dofBoundaryMark((int)i)) {
          g[i] = c->value(sp_->dofInterpPoint((int)i));
          break;
        }
    afe_matrix *a = A.device_matrix();
    if (!a) throw std::runtime_error("apply: the matrix must be assembled by a BilinearOperator of this space");
    A.push_to_device();
    detail::DeviceVector du(n), df(n);
    du.upload(u.data());
    df.upload(f.data());
    const int rc = afe_apply_dirichlet_bc(a, marks.data(), (int32_t)marks.size(), g.data(), du.p, df.p,
                                          preserve_symmetry ? 1 : 0);
    A.mark_device_newer();
    if (rc == AFE_OK) {
      du.download(u.data());
      df.download(f.data());
    }
    detail::check(rc);
  }

private:
  const FEMSpace<value_type, DIM, DOW, TDIM> *sp_;
  std::vector<const bc_t *> bcs_;
};

} // namespace AFEPack

using namespace AFEPack;
