// pybind11 module _ngsamd: the host FE setup (meshes, H1 space, integration
// rules, element classes, CSR pattern/assembly, linear forms) and C++ handles
// for device vectors and operators that call libngsamd through its C ABI.
// Python-side names follow the reference's (linalg/python_linalg.cpp,
// comp/python_comp.cpp).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <memory>
#include <stdexcept>

#include "fe.hpp"
#include "h1shapes.hpp"
#include "ngsamd.h"

namespace py = pybind11;
using namespace ngsamd;

template <class T>
using Arr = py::array_t<T, py::array::c_style | py::array::forcecast>;

template <class T>
static py::array_t<T> ToNumpy(const std::vector<T>& v) {
  py::array_t<T> a(v.size());
  if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(T));
  return a;
}

static void Check(int rc) {
  if (rc != NGSAMD_OK) throw std::runtime_error(std::string("ngsamd: ") + ngsamd_last_error());
}

// ------------------------------------------------------------ device vector
struct DeviceVector {
  void* ptr = nullptr;
  int64_t size = 0;
  int itemsize = 8;  // 8: float64, 4: float32
  bool owner = true;
  DeviceVector(int64_t n, int isz) : size(n), itemsize(isz) {
    if (isz != 4 && isz != 8) throw std::runtime_error("DeviceVector: itemsize must be 4 or 8");
    Check(ngsamd_malloc(&ptr, (size_t)n * isz));
  }
  DeviceVector(uintptr_t p, int64_t n, int isz) : ptr((void*)p), size(n), itemsize(isz), owner(false) {}
  ~DeviceVector() {
    if (owner && ptr) ngsamd_free(ptr);
  }
  DeviceVector(const DeviceVector&) = delete;
  double* d() { return (double*)ptr; }
};

struct Operator {
  ngsamd_op_t op = nullptr;
  std::string kind;
  int64_t height = 0, width = 0;
  int64_t nnz = 0;  // CSR operators: number of stored values
  int64_t ne = 0;   // matrix-free operators: number of global elements
  // matrix-free operators keep the pointer of an external coefficient, not a
  // copy: the vector is held here for as long as the operator may read it
  std::shared_ptr<DeviceVector> kappa_keep;
  // operators a composite uses (not owned by its handle): released after
  // the composite's handle is destroyed in ~Operator
  std::vector<std::shared_ptr<Operator>> parts;
  Operator(ngsamd_op_t o, std::string k) : op(o), kind(std::move(k)) {
    Check(ngsamd_op_height(op, &height));
    Check(ngsamd_op_width(op, &width));
    if (kind == "matfree") Check(ngsamd_mf_num_elements(op, &ne));
  }
  ~Operator() {
    if (op) ngsamd_op_destroy(op);
  }
  Operator(const Operator&) = delete;
};

static void RequireF64(const DeviceVector& v, int64_t n, const char* what) {
  if (v.itemsize != 8) throw std::runtime_error(std::string(what) + ": expected a float64 device vector");
  if (v.size != n) throw std::runtime_error(std::string(what) + ": vector size does not match the operator");
}

PYBIND11_MODULE(_ngsamd, m) {
  m.doc() = "ngsolve_amd native host module";

  // ------------------------------------------------------------------- mesh
  py::class_<Mesh, std::shared_ptr<Mesh>>(m, "Mesh")
      .def_readonly("dim", &Mesh::dim)
      .def_readonly("nv", &Mesh::nv)
      .def_property_readonly("ne", &Mesh::NE)
      .def_property_readonly("nedges", &Mesh::NEdges)
      .def_property_readonly("nfaces", &Mesh::NFaces)
      .def_property_readonly("nbnd", &Mesh::NBnd)
      .def_property_readonly("coords", [](const Mesh& mm) {
        auto a = ToNumpy(mm.coords);
        return a.reshape({(py::ssize_t)mm.nv, (py::ssize_t)mm.dim});
      })
      .def_property_readonly("elements", [](const Mesh& mm) {
        auto a = ToNumpy(mm.els);
        return a.reshape({(py::ssize_t)mm.NE(), (py::ssize_t)mm.NVE()});
      })
      .def_property_readonly("edges", [](const Mesh& mm) {
        return ToNumpy(mm.edges).reshape({(py::ssize_t)mm.NEdges(), (py::ssize_t)2});
      })
      .def_property_readonly("faces", [](const Mesh& mm) {
        return ToNumpy(mm.faces).reshape({(py::ssize_t)mm.NFaces(), (py::ssize_t)3});
      })
      .def_property_readonly("boundary", [](const Mesh& mm) {
        return ToNumpy(mm.bnd).reshape({(py::ssize_t)mm.NBnd(), (py::ssize_t)mm.dim});
      })
      .def_property_readonly("boundary_index", [](const Mesh& mm) { return ToNumpy(mm.bnd_index); })
      .def_readonly("boundary_names", &Mesh::bnd_names)
      .def_property_readonly("element_edges", [](const Mesh& mm) {
        return ToNumpy(mm.el_edges).reshape({(py::ssize_t)mm.NE(), (py::ssize_t)mm.NElEdges()});
      })
      .def_property_readonly("element_faces", [](const Mesh& mm) {
        return ToNumpy(mm.el_faces).reshape({(py::ssize_t)mm.NE(), (py::ssize_t)4});
      })
      .def_property_readonly("boundary_edges", [](const Mesh& mm) {
        return ToNumpy(mm.bnd_edges).reshape({(py::ssize_t)mm.NBnd(), (py::ssize_t)(mm.dim == 2 ? 1 : 3)});
      })
      .def_property_readonly("boundary_faces", [](const Mesh& mm) { return ToNumpy(mm.bnd_faces); });
  m.def("MakeStructured2DMesh", [](int nx, int ny) { return std::make_shared<Mesh>(MakeStructured2DMesh(nx, ny)); },
        py::arg("nx"), py::arg("ny"), py::call_guard<py::gil_scoped_release>());
  m.def("MakeStructured3DMesh",
        [](int nx, int ny, int nz) { return std::make_shared<Mesh>(MakeStructured3DMesh(nx, ny, nz)); },
        py::arg("nx"), py::arg("ny"), py::arg("nz"), py::call_guard<py::gil_scoped_release>());
  m.def("ReadVolMesh", [](const std::string& text) { return std::make_shared<Mesh>(ReadVolMesh(text)); });

  // ------------------------------------------------------------ integration
  py::class_<IntegrationRule>(m, "IntegrationRule")
      .def_readonly("dim", &IntegrationRule::dim)
      .def_readonly("order", &IntegrationRule::order)
      .def_property_readonly("points", [](const IntegrationRule& ir) {
        return ToNumpy(ir.pts).reshape({(py::ssize_t)ir.Size(), (py::ssize_t)ir.dim});
      })
      .def_property_readonly("weights", [](const IntegrationRule& ir) { return ToNumpy(ir.wts); })
      .def("__len__", &IntegrationRule::Size);
  m.def("TrigRule", &TrigRule);
  m.def("TetRule", &TetRule);
  m.def("GaussJacobi01", [](int n, int alpha) {
    std::vector<double> x, w;
    GaussJacobi01(n, alpha, x, w);
    return py::make_tuple(ToNumpy(x), ToNumpy(w));
  });

  // shape functions on the reference element (for tests and B tensors)
  m.def("ReferenceShapes",
        [](int dim, int order, std::vector<int> vnums, Arr<double> pts) {
          if ((int)vnums.size() != dim + 1) throw std::runtime_error("ReferenceShapes: need dim+1 vertex numbers");
          const int nd = dim == 2 ? TrigNDof(order) : TetNDof(order);
          const int np = (int)pts.shape(0);
          py::array_t<double> val({np, nd}), grad({np, nd, dim});
          auto P = pts.unchecked<2>();
          auto V = val.mutable_unchecked<2>();
          auto G = grad.mutable_unchecked<3>();
          for (int q = 0; q < np; q++) {
            if (dim == 2) {
              std::vector<AD<2>> s(nd);
              TrigShapes(order, vnums.data(), AD<2>::var(P(q, 0), 0), AD<2>::var(P(q, 1), 1), s.data());
              for (int i = 0; i < nd; i++) {
                V(q, i) = s[i].v;
                for (int d = 0; d < 2; d++) G(q, i, d) = s[i].d[d];
              }
            } else {
              std::vector<AD<3>> s(nd);
              TetShapes(order, vnums.data(), AD<3>::var(P(q, 0), 0), AD<3>::var(P(q, 1), 1), AD<3>::var(P(q, 2), 2),
                        s.data());
              for (int i = 0; i < nd; i++) {
                V(q, i) = s[i].v;
                for (int d = 0; d < 3; d++) G(q, i, d) = s[i].d[d];
              }
            }
          }
          return py::make_tuple(val, grad);
        },
        py::arg("dim"), py::arg("order"), py::arg("vnums"), py::arg("points"));

  // ------------------------------------------------------------------ space
  py::class_<H1Space, std::shared_ptr<H1Space>>(m, "H1Space")
      .def(py::init([](std::shared_ptr<Mesh> mesh, int order, uint64_t mask) {
             auto s = std::make_shared<H1Space>(*mesh, order, mask);
             return s;
           }),
           py::arg("mesh"), py::arg("order"), py::arg("dirichlet_mask"), py::keep_alive<1, 2>())
      .def_readonly("order", &H1Space::order)
      .def_readonly("ndof", &H1Space::ndof)
      .def_readonly("nd_el", &H1Space::nd_el)
      .def_property_readonly("el_dofs", [](const H1Space& f) {
        return ToNumpy(f.el_dofs).reshape({(py::ssize_t)f.mesh->NE(), (py::ssize_t)f.nd_el});
      })
      .def_property_readonly("freedofs", [](const H1Space& f) { return ToNumpy(f.freedofs); })
      .def_property_readonly("classnr", [](const H1Space& f) { return ToNumpy(f.classnr); })
      .def_property_readonly("first_edge_dof", [](const H1Space& f) { return ToNumpy(f.first_edge_dof); })
      .def_property_readonly("first_face_dof", [](const H1Space& f) { return ToNumpy(f.first_face_dof); })
      .def_property_readonly("first_el_dof", [](const H1Space& f) { return ToNumpy(f.first_el_dof); });

  m.def("ComputeGeometry", [](std::shared_ptr<Mesh> mesh) {
    ElementGeometry g = ComputeGeometry(*mesh);
    const int d = mesh->dim;
    return py::make_tuple(ToNumpy(g.detj), ToNumpy(g.ginv).reshape({(py::ssize_t)mesh->NE(), (py::ssize_t)d, (py::ssize_t)d}));
  });

  m.def("BuildClasses", [](std::shared_ptr<H1Space> fes, const IntegrationRule& ir) {
    std::vector<ElementClass> cl = BuildClasses(*fes, ir);
    py::list out;
    const int nd = fes->nd_el, dim = fes->mesh->dim, nq = ir.Size();
    for (auto& c : cl) {
      py::dict d;
      d["classnr"] = c.classnr;
      d["elements"] = ToNumpy(c.elements);
      d["bmat"] = ToNumpy(c.bmat).reshape({(py::ssize_t)nq, (py::ssize_t)nd, (py::ssize_t)dim});
      d["shape"] = ToNumpy(c.shape).reshape({(py::ssize_t)nq, (py::ssize_t)nd});
      out.append(d);
    }
    return out;
  });

  m.def("ElementColoring", [](std::shared_ptr<Mesh> mesh) { return ToNumpy(ElementColoring(*mesh)); });
  m.def("BuildPattern", [](std::shared_ptr<H1Space> fes) {
    CSRPattern p;
    {
      py::gil_scoped_release rel;
      p = BuildPattern(*fes);
    }
    return py::make_tuple(ToNumpy(p.rowptr), ToNumpy(p.cols));
  });

  m.def("AssembleCSR",
        [](std::shared_ptr<H1Space> fes, Arr<int64_t> rowptr, Arr<int> cols, const IntegrationRule& ir, bool stiffness,
           py::object kappa, bool mass, py::object rho) {
          CSRPattern p;
          p.n = fes->ndof;
          if (rowptr.size() != p.n + 1) throw std::runtime_error("AssembleCSR: row pointer size");
          p.rowptr.assign(rowptr.data(), rowptr.data() + rowptr.size());
          p.cols.assign(cols.data(), cols.data() + cols.size());
          const int ne = fes->mesh->NE();
          std::vector<double> kap, rh;
          if (!kappa.is_none()) {
            auto a = kappa.cast<Arr<double>>();
            if (a.size() != ne) throw std::runtime_error("AssembleCSR: kappa must have one value per element");
            kap.assign(a.data(), a.data() + ne);
          }
          if (!rho.is_none()) {
            auto a = rho.cast<Arr<double>>();
            if (a.size() != ne) throw std::runtime_error("AssembleCSR: rho must have one value per element");
            rh.assign(a.data(), a.data() + ne);
          }
          py::array_t<double> vals(p.cols.size());
          double* out = vals.mutable_data();
          {
            py::gil_scoped_release rel;
            AssembleCSR(*fes, p, ir, stiffness, kap.empty() ? nullptr : kap.data(), mass,
                        rh.empty() ? nullptr : rh.data(), out);
          }
          return vals;
        },
        py::arg("fes"), py::arg("rowptr"), py::arg("cols"), py::arg("ir"), py::arg("stiffness") = true,
        py::arg("kappa") = py::none(), py::arg("mass") = false, py::arg("rho") = py::none());

  // physical quadrature points [ne][nq][dim] and measure weights w_q |det F| [ne][nq]
  m.def("QuadraturePoints", [](std::shared_ptr<Mesh> mesh, const IntegrationRule& ir) {
    const int dim = mesh->dim, ne = mesh->NE(), nq = ir.Size();
    ElementGeometry g = ComputeGeometry(*mesh);
    py::array_t<double> pts({ne, nq, dim}), wts({ne, nq});
    auto P = pts.mutable_unchecked<3>();
    auto W = wts.mutable_unchecked<2>();
    for (int e = 0; e < ne; e++) {
      const int* v = &mesh->els[(size_t)e * (dim + 1)];
      const double* p0 = &mesh->coords[(size_t)dim * v[dim]];
      for (int q = 0; q < nq; q++) {
        for (int k = 0; k < dim; k++) {
          double x = p0[k];
          for (int j = 0; j < dim; j++) x += (mesh->coords[(size_t)dim * v[j] + k] - p0[k]) * ir.pts[q * dim + j];
          P(e, q, k) = x;
        }
        W(e, q) = ir.wts[q] * g.detj[e];
      }
    }
    return py::make_tuple(pts, wts);
  });

  // f_i = sum_e sum_q w_q |det F_e| f(e,q) phi_i(q)  (fem/symbolicintegrator.cpp:1036-1110)
  m.def("AssembleLinearForm", [](std::shared_ptr<H1Space> fes, const IntegrationRule& ir, Arr<double> fvals) {
    const Mesh& mesh = *fes->mesh;
    const int ne = mesh.NE(), nq = ir.Size(), nd = fes->nd_el;
    if (fvals.size() != (py::ssize_t)ne * nq) throw std::runtime_error("AssembleLinearForm: fvals must be [ne, nq]");
    std::vector<ElementClass> cl = BuildClasses(*fes, ir);
    ElementGeometry g = ComputeGeometry(mesh);
    std::vector<double> f(fes->ndof, 0.0);
    const double* fv = fvals.data();
    for (auto& c : cl)
      for (int e : c.elements)
        for (int i = 0; i < nd; i++) {
          double acc = 0.0;
          for (int q = 0; q < nq; q++) acc += ir.wts[q] * fv[(size_t)e * nq + q] * c.shape[(size_t)q * nd + i];
          f[fes->el_dofs[(size_t)e * nd + i]] += g.detj[e] * acc;
        }
    return ToNumpy(f);
  });

  // u_h at the quadrature points [ne][nq]
  m.def("EvaluateAtPoints", [](std::shared_ptr<H1Space> fes, const IntegrationRule& ir, Arr<double> u) {
    const int ne = fes->mesh->NE(), nq = ir.Size(), nd = fes->nd_el;
    if (u.size() != fes->ndof) throw std::runtime_error("EvaluateAtPoints: vector size");
    std::vector<ElementClass> cl = BuildClasses(*fes, ir);
    py::array_t<double> out({ne, nq});
    auto O = out.mutable_unchecked<2>();
    const double* uv = u.data();
    for (auto& c : cl)
      for (int e : c.elements)
        for (int q = 0; q < nq; q++) {
          double acc = 0.0;
          for (int i = 0; i < nd; i++) acc += c.shape[(size_t)q * nd + i] * uv[fes->el_dofs[(size_t)e * nd + i]];
          O(e, q) = acc;
        }
    return out;
  });

  // ------------------------------------------------------------------ device
  m.def("init", [](int dev) { Check(ngsamd_init(dev)); }, py::arg("device") = 0);
  m.def("synchronize", [] { Check(ngsamd_synchronize()); });
  m.def("set_stream", [](uintptr_t s) { Check(ngsamd_set_stream((void*)s)); });
  m.def("get_stream", [] { return (uintptr_t)ngsamd_get_stream(); });
  m.def("event_create", [] {
    void* e = nullptr;
    Check(ngsamd_event_create(&e));
    return (uintptr_t)e;
  });
  m.def("event_record", [](uintptr_t e) { Check(ngsamd_event_record((void*)e)); });
  m.def("event_elapsed_ms", [](uintptr_t a, uintptr_t b) {
    float ms = 0;
    Check(ngsamd_event_elapsed_ms((void*)a, (void*)b, &ms));
    return ms;
  });
  m.def("event_destroy", [](uintptr_t e) { Check(ngsamd_event_destroy((void*)e)); });
  m.def("stream_read", [](DeviceVector& x, DeviceVector& result) {
    RequireF64(x, x.size, "stream_read");
    RequireF64(result, 1, "stream_read (one result)");
    Check(ngsamd_stream_read(x.size, x.d(), result.d()));
  });

  py::class_<DeviceVector, std::shared_ptr<DeviceVector>>(m, "DeviceVector")
      .def(py::init<int64_t, int>(), py::arg("size"), py::arg("itemsize") = 8)
      .def_static("wrap", [](uintptr_t p, int64_t n, int isz) { return std::make_shared<DeviceVector>(p, n, isz); })
      .def_readonly("size", &DeviceVector::size)
      .def_readonly("itemsize", &DeviceVector::itemsize)
      .def_property_readonly("ptr", [](const DeviceVector& v) { return (uintptr_t)v.ptr; })
      .def("upload",
           [](DeviceVector& v, py::array a) {
             if ((int64_t)a.size() != v.size) throw std::runtime_error("DeviceVector.upload: size mismatch");
             if (v.itemsize == 8) {
               Arr<double> c = a.cast<Arr<double>>();
               Check(ngsamd_memcpy_htod(v.ptr, c.data(), (size_t)v.size * 8));
             } else {
               Arr<float> c = a.cast<Arr<float>>();
               Check(ngsamd_memcpy_htod(v.ptr, c.data(), (size_t)v.size * 4));
             }
           })
      .def("download",
           [](DeviceVector& v) -> py::array {
             if (v.itemsize == 8) {
               py::array_t<double> a(v.size);
               Check(ngsamd_memcpy_dtoh(a.mutable_data(), v.ptr, (size_t)v.size * 8));
               return a;
             }
             py::array_t<float> a(v.size);
             Check(ngsamd_memcpy_dtoh(a.mutable_data(), v.ptr, (size_t)v.size * 4));
             return a;
           })
      .def("copy_from",
           [](DeviceVector& v, const DeviceVector& o) {
             if (o.size != v.size || o.itemsize != v.itemsize) throw std::runtime_error("copy_from: shape mismatch");
             Check(ngsamd_memcpy_dtod(v.ptr, o.ptr, (size_t)v.size * v.itemsize));
           })
      .def("fill", [](DeviceVector& v, double a) { RequireF64(v, v.size, "fill"); Check(ngsamd_vec_set(v.size, a, v.d())); })
      .def("scale", [](DeviceVector& v, double a) { RequireF64(v, v.size, "scale"); Check(ngsamd_vec_scale(v.size, a, v.d())); })
      .def("axpy",
           [](DeviceVector& y, double a, DeviceVector& x) {
             RequireF64(y, y.size, "axpy");
             RequireF64(x, y.size, "axpy");
             Check(ngsamd_vec_axpy(y.size, a, x.d(), y.d()));
           })
      .def("axpby",
           [](DeviceVector& y, double a, DeviceVector& x, double b) {
             RequireF64(y, y.size, "axpby");
             RequireF64(x, y.size, "axpby");
             Check(ngsamd_vec_axpby(y.size, a, x.d(), b, y.d()));
           })
      .def("dot", [](DeviceVector& x, DeviceVector& y) {
        double r = 0.0;
        if (x.size != y.size || x.itemsize != y.itemsize) throw std::runtime_error("dot: shape mismatch");
        if (x.itemsize == 8) Check(ngsamd_vec_dot(x.size, x.d(), y.d(), &r));
        else Check(ngsamd_vec_dot_f32(x.size, (const float*)x.ptr, (const float*)y.ptr, &r));
        return r;
      });

  py::class_<Operator, std::shared_ptr<Operator>>(m, "Operator")
      .def_readonly("kind", &Operator::kind)
      // the ngsamd_op_t handle, for calls through the C ABI (ctypes)
      .def_property_readonly("handle", [](const Operator& A) { return (uintptr_t)A.op; })
      .def_readonly("height", &Operator::height)
      .def_readonly("width", &Operator::width)
      .def("mult",
           [](Operator& A, DeviceVector& x, DeviceVector& y) {
             RequireF64(x, A.width, "mult");
             RequireF64(y, A.height, "mult");
             Check(ngsamd_op_mult(A.op, x.d(), y.d()));
           })
      .def("mult_add",
           [](Operator& A, double s, DeviceVector& x, DeviceVector& y) {
             RequireF64(x, A.width, "mult_add");
             RequireF64(y, A.height, "mult_add");
             Check(ngsamd_op_mult_add(A.op, s, x.d(), y.d()));
           })
      .def("mult_f32",
           [](Operator& A, DeviceVector& x, DeviceVector& y) {
             if (x.itemsize != 4 || y.itemsize != 4 || x.size != A.width || y.size != A.height)
               throw std::runtime_error("mult_f32: expected float32 device vectors of the operator size");
             Check(ngsamd_mf_mult_f32(A.op, (const float*)x.ptr, (float*)y.ptr));
           })
      .def("diagonal",
           [](Operator& A, DeviceVector& d) {
             RequireF64(d, A.height, "diagonal");
             if (A.kind == "csr") Check(ngsamd_csr_diagonal(A.op, d.d()));
             else if (A.kind == "matfree") Check(ngsamd_mf_diagonal(A.op, d.d()));
             else throw std::runtime_error("diagonal: not available for this operator");
           })
      .def_readonly("ne", &Operator::ne)
      .def("set_coloring",
           [](Operator& A, Arr<int> colors) {
             if (A.kind != "matfree") throw std::runtime_error("set_coloring: not a matrix-free operator");
             Check(ngsamd_mf_set_coloring(A.op, colors.data()));
           })
      .def("set_kappa",
           [](Operator& A, std::shared_ptr<DeviceVector> k) {
             if (A.kind != "matfree") throw std::runtime_error("set_kappa: not a matrix-free operator");
             RequireF64(*k, A.ne, "set_kappa (one value per element)");
             Check(ngsamd_mf_set_kappa(A.op, k->d(), k->size));
             A.kappa_keep = std::move(k);
           })
      .def("element_grad",
           [](Operator& A, DeviceVector& lam, DeviceVector& u, DeviceVector& g) {
             if (A.kind != "matfree") throw std::runtime_error("element_grad: not a matrix-free operator");
             RequireF64(lam, A.height, "element_grad");
             RequireF64(u, A.height, "element_grad");
             RequireF64(g, A.ne, "element_grad (one value per element)");
             Check(ngsamd_mf_element_grad(A.op, lam.d(), u.d(), g.d()));
           })
      .def_readonly("nnz", &Operator::nnz)
      .def("csr_values", [](Operator& A) {
        if (A.kind != "csr") throw std::runtime_error("csr_values: not a CSR operator");
        py::array_t<double> v(A.nnz);
        Check(ngsamd_csr_values(A.op, v.mutable_data()));
        return v;
      });

  m.def("CreateCSR",
        [](int64_t h, int64_t w, Arr<int64_t> rowptr, Arr<int> cols, Arr<double> vals) {
          if (rowptr.size() != h + 1 || cols.size() != vals.size() || rowptr.data()[h] != (int64_t)cols.size())
            throw std::runtime_error("CreateCSR: inconsistent CSR arrays");
          ngsamd_op_t op = nullptr;
          Check(ngsamd_csr_create(&op, h, w, rowptr.data(), cols.data(), vals.data()));
          auto A = std::make_shared<Operator>(op, "csr");
          A->nnz = (int64_t)cols.size();
          return A;
        });
  m.def("LinearFormDevice", [](Operator& mf, Arr<double> weights, Arr<double> shape, Arr<double> detj,
                                DeviceVector& fvals, DeviceVector& out) {
    if (mf.kind != "matfree") throw std::runtime_error("LinearFormDevice: matrix-free operator expected");
    const int nq = (int)weights.size();
    if (shape.ndim() != 3 || shape.shape(1) != nq)
      throw std::runtime_error("LinearFormDevice: shape must be [nclass][nq][nd] with nq = len(weights)");
    if (fvals.itemsize != 8 || fvals.size != (int64_t)detj.size() * nq)
      throw std::runtime_error("LinearFormDevice: fvals must be float64 [ne][nq] with ne = len(detj)");
    RequireF64(out, mf.height, "LinearFormDevice");
    Check(ngsamd_mf_linear_form(mf.op, (int)shape.shape(0), nq, (int)shape.shape(2), (int64_t)detj.size(),
                                weights.data(), shape.data(), detj.data(), fvals.d(), out.d()));
  });
  m.def("AssembleMassCSRDevice", [](Operator& mf, Arr<double> weights, Arr<double> shape, Arr<double> scale,
                                     Arr<int64_t> rowptr, Arr<int> cols) {
    if (mf.kind != "matfree") throw std::runtime_error("AssembleMassCSRDevice: matrix-free operator expected");
    const int nq = (int)weights.size();
    if (shape.ndim() != 3 || shape.shape(1) != nq)
      throw std::runtime_error("AssembleMassCSRDevice: shape must be [nclass][nq][nd] with nq = len(weights)");
    if (rowptr.size() != mf.height + 1 || rowptr.data()[mf.height] != (int64_t)cols.size())
      throw std::runtime_error("AssembleMassCSRDevice: inconsistent CSR arrays");
    ngsamd_op_t op = nullptr;
    Check(ngsamd_mf_assemble_mass_csr(mf.op, (int)shape.shape(0), nq, (int)shape.shape(2), (int64_t)scale.size(),
                                      weights.data(), shape.data(), scale.data(), rowptr.data(), cols.data(), &op));
    auto A = std::make_shared<Operator>(op, "csr");
    A->nnz = (int64_t)cols.size();
    return A;
  });
  m.def("H1TablesDevice", [](int dim, int64_t nv, Arr<int> els, Arr<int> bnd, Arr<int> bnd_index, int order,
                             uint64_t dirichlet_mask) {
    if (dim != 2 && dim != 3) throw std::runtime_error("H1TablesDevice: dim must be 2 or 3");
    if (els.size() % (dim + 1) || bnd.size() % dim || bnd_index.size() * dim != bnd.size())
      throw std::runtime_error("H1TablesDevice: els must be [ne][dim+1], bnd [nbnd][dim], bnd_index [nbnd]");
    const int64_t ne = els.size() / (dim + 1), nb = bnd_index.size();
    ngsamd_h1_t h = nullptr;
    Check(ngsamd_h1_create(&h, dim, nv, ne, els.data(), nb, bnd.data(), bnd_index.data(), order, dirichlet_mask));
    std::unique_ptr<ngsamd_h1_s, int (*)(ngsamd_h1_t)> guard(h, ngsamd_h1_destroy);
    int64_t s[6];
    Check(ngsamd_h1_sizes(h, s));
    const int64_t ned = s[0], nfa = s[1], nbo = s[2], ndof = s[3], nd = s[4];
    const int nle = dim == 2 ? 3 : 6, nbe = dim == 2 ? 1 : 3;
    py::array_t<int> edges({ned, (int64_t)2}), faces({nfa, (int64_t)3}), el_edges({ne, (int64_t)nle}),
        el_faces({dim == 3 ? ne : 0, (int64_t)4}), obnd({nbo, (int64_t)dim}), obi(nbo),
        bnd_edges({nbo, (int64_t)nbe}), bnd_faces(dim == 3 ? nbo : 0), el_dofs({ne, nd}), classnr(ne);
    py::array_t<uint8_t> freedofs(ndof);
    Check(ngsamd_h1_topology(h, edges.mutable_data(), faces.mutable_data(), el_edges.mutable_data(),
                             el_faces.mutable_data(), obnd.mutable_data(), obi.mutable_data(),
                             bnd_edges.mutable_data(), bnd_faces.mutable_data()));
    Check(ngsamd_h1_dofs(h, el_dofs.mutable_data(), freedofs.mutable_data(), classnr.mutable_data(), nullptr));
    py::dict d;
    d["ndof"] = ndof;
    d["edges"] = edges;
    d["faces"] = faces;
    d["element_edges"] = el_edges;
    d["element_faces"] = el_faces;
    d["boundary"] = obnd;
    d["boundary_index"] = obi;
    d["boundary_edges"] = bnd_edges;
    d["boundary_faces"] = bnd_faces;
    d["el_dofs"] = el_dofs;
    d["freedofs"] = freedofs;
    d["classnr"] = classnr;
    return d;
  });
  m.def("BuildTopologyHost", [](int dim, int64_t nv, Arr<int> els, Arr<int> bnd, Arr<int> bnd_index) {
    // the host setup's topology from the same arrays (for comparison)
    auto mesh = std::make_shared<Mesh>();
    mesh->dim = dim;
    mesh->nv = (int)nv;
    mesh->coords.assign((size_t)nv * dim, 0.0);
    mesh->els.assign(els.data(), els.data() + els.size());
    mesh->bnd.assign(bnd.data(), bnd.data() + bnd.size());
    mesh->bnd_index.assign(bnd_index.data(), bnd_index.data() + bnd_index.size());
    mesh->BuildTopology();
    return mesh;
  });
  m.def("PatternDevice", [](Operator& mf) {
    if (mf.kind != "matfree") throw std::runtime_error("PatternDevice: matrix-free operator expected");
    py::array_t<int64_t> rp(mf.height + 1);
    Check(ngsamd_mf_pattern(mf.op, rp.mutable_data(), nullptr));
    py::array_t<int> cols(rp.data()[mf.height]);
    Check(ngsamd_mf_pattern(mf.op, rp.mutable_data(), cols.mutable_data()));
    return py::make_tuple(rp, cols);
  });
  m.def("AssembleCSRDevice", [](Operator& mf, Arr<int64_t> rowptr, Arr<int> cols) {
    if (mf.kind != "matfree") throw std::runtime_error("AssembleCSRDevice: matrix-free operator expected");
    if (rowptr.size() != mf.height + 1 || rowptr.data()[mf.height] != (int64_t)cols.size())
      throw std::runtime_error("AssembleCSRDevice: inconsistent CSR arrays");
    ngsamd_op_t op = nullptr;
    Check(ngsamd_mf_assemble_csr(mf.op, rowptr.data(), cols.data(), &op));
    auto A = std::make_shared<Operator>(op, "csr");
    A->nnz = (int64_t)cols.size();
    return A;
  });
  m.def("CreateJacobi", [](Arr<double> diag, py::object freedofs) {
    ngsamd_op_t op = nullptr;
    std::vector<uint8_t> fr;
    if (!freedofs.is_none()) {
      auto a = freedofs.cast<Arr<uint8_t>>();
      if (a.size() != diag.size()) throw std::runtime_error("CreateJacobi: freedofs size");
      fr.assign(a.data(), a.data() + a.size());
    }
    Check(ngsamd_jacobi_create(&op, diag.size(), diag.data(), fr.empty() ? nullptr : fr.data()));
    return std::make_shared<Operator>(op, "jacobi");
  });
  m.def("CreateBlockJacobi", [](int64_t n, Arr<int64_t> rowptr, Arr<int> cols, Arr<double> vals,
                                 Arr<int64_t> block_ptr, Arr<int> block_dofs) {
    if (rowptr.size() != n + 1 || cols.size() != vals.size() || rowptr.data()[n] != (int64_t)cols.size())
      throw std::runtime_error("CreateBlockJacobi: inconsistent CSR arrays");
    const int64_t nb = (int64_t)block_ptr.size() - 1;
    if (nb < 0 || block_ptr.data()[nb] != (int64_t)block_dofs.size())
      throw std::runtime_error("CreateBlockJacobi: inconsistent block arrays");
    ngsamd_op_t op = nullptr;
    Check(ngsamd_blockjacobi_create(&op, n, rowptr.data(), cols.data(), vals.data(), nb, block_ptr.data(),
                                    block_dofs.data()));
    return std::make_shared<Operator>(op, "blockjacobi");
  });
  m.def("CreateSparseCholesky", [](int64_t n, Arr<int64_t> rowptr, Arr<int> cols, Arr<double> vals,
                                   py::object freedofs) {
    if (rowptr.size() != n + 1 || cols.size() != vals.size() || rowptr.data()[n] != (int64_t)cols.size())
      throw std::runtime_error("CreateSparseCholesky: inconsistent CSR arrays");
    std::vector<uint8_t> fr;
    if (!freedofs.is_none()) {
      auto a = freedofs.cast<Arr<uint8_t>>();
      if ((int64_t)a.size() != n) throw std::runtime_error("CreateSparseCholesky: freedofs size");
      fr.assign(a.data(), a.data() + a.size());
    }
    ngsamd_op_t op = nullptr;
    {
      py::gil_scoped_release rel;
      Check(ngsamd_sparsecholesky_create(&op, n, rowptr.data(), cols.data(), vals.data(),
                                         fr.empty() ? nullptr : fr.data()));
    }
    return std::make_shared<Operator>(op, "sparsecholesky");
  });
  m.def("SparseCholeskyInfo", [](Operator& A) {
    int64_t info[4] = {0, 0, 0, 0};
    Check(ngsamd_sparsecholesky_info(A.op, info));
    py::dict d;
    d["nfree"] = info[0];
    d["supernodes"] = info[1];
    d["levels"] = info[2];
    d["nnz_factor"] = info[3];
    return d;
  });
  m.def("CreateMultiGrid", [](int dim, std::vector<int> dims, int64_t n, Arr<int64_t> rowptr, Arr<int> cols,
                              Arr<double> vals, py::object freedofs, int smoothing_steps, double omega,
                              int max_levels) {
    if (rowptr.size() != n + 1 || cols.size() != vals.size() || rowptr.data()[n] != (int64_t)cols.size())
      throw std::runtime_error("CreateMultiGrid: inconsistent CSR arrays");
    if ((int)dims.size() != dim) throw std::runtime_error("CreateMultiGrid: dims must have dim entries");
    std::vector<uint8_t> fr;
    if (!freedofs.is_none()) {
      auto a = freedofs.cast<Arr<uint8_t>>();
      if ((int64_t)a.size() != n) throw std::runtime_error("CreateMultiGrid: freedofs size");
      fr.assign(a.data(), a.data() + a.size());
    }
    ngsamd_op_t op = nullptr;
    {
      py::gil_scoped_release rel;
      Check(ngsamd_multigrid_create(&op, dim, dims.data(), n, rowptr.data(), cols.data(), vals.data(),
                                    fr.empty() ? nullptr : fr.data(), smoothing_steps, omega, max_levels));
    }
    return std::make_shared<Operator>(op, "multigrid");
  });
  m.def("MultiGridInfo", [](Operator& A) {
    int nl = 0;
    std::vector<int64_t> dofs(32, 0);
    Check(ngsamd_multigrid_info(A.op, &nl, dofs.data()));
    dofs.resize(nl);
    return dofs;
  });
  m.def("CreateTwoLevel", [](std::shared_ptr<Operator> smoother, std::shared_ptr<Operator> coarse) {
    ngsamd_op_t op = nullptr;
    Check(ngsamd_twolevel_create(&op, smoother->op, coarse->op));
    auto r = std::make_shared<Operator>(op, "twolevel");
    r->parts = {smoother, coarse};
    return r;
  });
  m.def("CreateJacobiDevice", [](DeviceVector& diag, py::object freedofs) {
    ngsamd_op_t op = nullptr;
    std::vector<uint8_t> fr;
    if (!freedofs.is_none()) {
      auto a = freedofs.cast<Arr<uint8_t>>();
      if (a.size() != diag.size) throw std::runtime_error("CreateJacobiDevice: freedofs size");
      fr.assign(a.data(), a.data() + a.size());
    }
    Check(ngsamd_jacobi_create_device(&op, diag.size, diag.d(), fr.empty() ? nullptr : fr.data()));
    return std::make_shared<Operator>(op, "jacobi");
  });
  m.def("CreateMatrixFree",
        [](int dim, int nd, int nq, int64_t ndof, Arr<int64_t> offsets, Arr<double> bmat, Arr<double> weights,
           Arr<int> dofs, Arr<double> metric, py::object kappa, py::object elnr, int precision, int scatter) {
          const int nclass = (int)offsets.size() - 1;
          if (nclass < 1) throw std::runtime_error("CreateMatrixFree: need at least one class");
          const int64_t nel = offsets.data()[nclass];
          if (bmat.size() != (py::ssize_t)nclass * nq * nd * dim || weights.size() != nq ||
              dofs.size() != nel * nd || metric.size() != nel * dim * (dim + 1) / 2)
            throw std::runtime_error("CreateMatrixFree: inconsistent array sizes");
          std::vector<double> kap;
          std::vector<int> el;
          if (!kappa.is_none()) {
            auto a = kappa.cast<Arr<double>>();
            if (a.size() != nel) throw std::runtime_error("CreateMatrixFree: kappa size");
            kap.assign(a.data(), a.data() + nel);
          }
          if (!elnr.is_none()) {
            auto a = elnr.cast<Arr<int>>();
            if (a.size() != nel) throw std::runtime_error("CreateMatrixFree: elnr size");
            el.assign(a.data(), a.data() + nel);
          }
          ngsamd_op_t op = nullptr;
          Check(ngsamd_mf_create(&op, dim, nd, nq, ndof, nclass, offsets.data(), bmat.data(), weights.data(),
                                 dofs.data(), metric.data(), kap.empty() ? nullptr : kap.data(),
                                 el.empty() ? nullptr : el.data(), precision, scatter));
          return std::make_shared<Operator>(op, "matfree");
        },
        py::arg("dim"), py::arg("nd"), py::arg("nq"), py::arg("ndof"), py::arg("offsets"), py::arg("bmat"),
        py::arg("weights"), py::arg("dofs"), py::arg("metric"), py::arg("kappa") = py::none(),
        py::arg("elnr") = py::none(), py::arg("precision") = 0, py::arg("scatter") = 0);

  m.def("CGSolve",
        [](Operator& A, Operator& pre, DeviceVector& f, DeviceVector& u, bool initialize, double tol, double atol,
           int maxiter, int chunk) {
          RequireF64(f, A.height, "CGSolve");
          RequireF64(u, A.height, "CGSolve");
          int its = 0;
          std::vector<double> res(std::max(maxiter, 1));
          {
            py::gil_scoped_release rel;
            Check(ngsamd_cg_solve(A.op, pre.op, f.d(), u.d(), initialize ? 1 : 0, tol, atol, maxiter, chunk, &its,
                                  res.data()));
          }
          res.resize(std::min(its, maxiter));
          return py::make_tuple(its, ToNumpy(res));
        },
        py::arg("mat"), py::arg("pre"), py::arg("rhs"), py::arg("sol"), py::arg("initialize") = true,
        py::arg("tol") = 1e-12, py::arg("atol") = -1.0, py::arg("maxiter") = 100, py::arg("chunk") = 16);

  m.def("CGSolveMulti",
        [](Operator& A, py::object pre, DeviceVector& f, DeviceVector& u, int nrhs, bool initialize, double tol,
           bool absolute, int maxiter) {
          if (f.itemsize != 8 || u.itemsize != 8 || f.size != A.height * nrhs || u.size != A.height * nrhs)
            throw std::runtime_error("CGSolveMulti: float64 vectors of height * nrhs entries expected");
          ngsamd_op_t p = pre.is_none() ? nullptr : pre.cast<Operator&>().op;
          int steps = 0;
          std::vector<double> res(std::max(maxiter, 0) + 1);
          {
            py::gil_scoped_release rel;
            Check(ngsamd_cg_solve_multi(A.op, p, nrhs, f.d(), u.d(), initialize ? 1 : 0, tol, absolute ? 1 : 0,
                                        maxiter, &steps, res.data()));
          }
          res.resize(std::min(steps, maxiter + 1));
          return py::make_tuple(steps, ToNumpy(res));
        },
        py::arg("mat"), py::arg("pre"), py::arg("rhs"), py::arg("sol"), py::arg("nrhs"), py::arg("initialize") = true,
        py::arg("tol") = 1e-8, py::arg("absolute") = false, py::arg("maxiter") = 200);
  m.def("CSRMultMulti", [](Operator& A, DeviceVector& x, DeviceVector& y, int nrhs) {
    if (x.itemsize != 8 || y.itemsize != 8 || x.size != A.width * nrhs || y.size != A.height * nrhs)
      throw std::runtime_error("CSRMultMulti: float64 vectors of size * nrhs entries expected");
    Check(ngsamd_csr_mult_multi(A.op, nrhs, x.d(), y.d()));
  });
  m.def("CreateBatchCSR",
        [](Arr<int64_t> rowptr, Arr<int> cols, Arr<double> vals, Arr<int64_t> seg, bool with_f32) {
          const int64_t n = rowptr.size() - 1;
          if (n < 0 || cols.size() != vals.size() || rowptr.data()[n] != (int64_t)cols.size())
            throw std::runtime_error("CreateBatchCSR: inconsistent CSR arrays");
          ngsamd_op_t op = nullptr;
          Check(ngsamd_batch_csr_create(&op, n, rowptr.data(), cols.data(), vals.data(), (int)seg.size() - 1,
                                        seg.data(), with_f32 ? 1 : 0));
          return std::make_shared<Operator>(op, "batch");
        });
  m.def("BatchInfo", [](Operator& A) {
    int ns = 0, bn = 0, nb = 0;
    Check(ngsamd_batch_info(A.op, &ns, &bn, &nb));
    return py::make_tuple(ns, bn, nb);
  });
  m.def("BatchTuning", [](Operator& A) {
    int cnt = 0;
    double ms[3] = {0, 0, 0};
    Check(ngsamd_batch_tuning(A.op, &cnt, ms));
    py::dict d;
    const int cand[3] = {1024, 2048, 4096};
    for (int k = 0; k < cnt && k < 3; k++) d[py::int_(cand[k])] = ms[k];
    return d;
  });
  m.def("BatchCGSolve",
        [](Operator& A, Operator& pre, DeviceVector& f, DeviceVector& u, double tol, int maxiter, int chunk) {
          RequireF64(f, A.height, "BatchCGSolve");
          RequireF64(u, A.height, "BatchCGSolve");
          int ns = 0;
          Check(ngsamd_batch_info(A.op, &ns, nullptr, nullptr));
          std::vector<int> its(ns);
          std::vector<double> res(ns), ms(ns);
          {
            py::gil_scoped_release rel;
            Check(ngsamd_batch_cg_solve(A.op, pre.op, f.d(), u.d(), tol, maxiter, chunk, its.data(), res.data(),
                                        ms.data()));
          }
          return py::make_tuple(ToNumpy(its), ToNumpy(res), ToNumpy(ms));
        },
        py::arg("mat"), py::arg("pre"), py::arg("rhs"), py::arg("sol"), py::arg("tol") = 1e-10,
        py::arg("maxiter") = 100000, py::arg("chunk") = 16);
  m.def("BatchIRCGSolve",
        [](Operator& A, Operator& pre, DeviceVector& f, DeviceVector& u, double tol, double inner_tol, int max_outer,
           int max_inner) {
          RequireF64(f, A.height, "BatchIRCGSolve");
          RequireF64(u, A.height, "BatchIRCGSolve");
          int ns = 0;
          Check(ngsamd_batch_info(A.op, &ns, nullptr, nullptr));
          std::vector<int> outer(ns), inner(ns);
          std::vector<double> rel(ns), ms(ns);
          {
            py::gil_scoped_release g;
            Check(ngsamd_batch_ir_cg_solve(A.op, pre.op, f.d(), u.d(), tol, inner_tol, max_outer, max_inner,
                                           outer.data(), inner.data(), rel.data(), ms.data()));
          }
          return py::make_tuple(ToNumpy(outer), ToNumpy(inner), ToNumpy(rel), ToNumpy(ms));
        },
        py::arg("mat"), py::arg("pre"), py::arg("rhs"), py::arg("sol"), py::arg("tol") = 1e-10,
        py::arg("inner_tol") = 1e-4, py::arg("max_outer") = 50, py::arg("max_inner") = 100000);

  m.def("IRCGSolve",
        [](Operator& A, Operator& inner, Operator& pre, DeviceVector& f, DeviceVector& u, double tol, double inner_tol,
           int max_outer, int max_inner) {
          RequireF64(f, A.height, "IRCGSolve");
          RequireF64(u, A.height, "IRCGSolve");
          double st[3] = {0, 0, 0};
          {
            py::gil_scoped_release rel;
            Check(ngsamd_ir_cg_solve(A.op, inner.op, pre.op, f.d(), u.d(), tol, inner_tol, max_outer, max_inner, st));
          }
          return py::make_tuple((int)st[0], (long long)st[1], st[2]);
        },
        py::arg("mat"), py::arg("inner"), py::arg("pre"), py::arg("rhs"), py::arg("sol"), py::arg("tol") = 1e-10,
        py::arg("inner_tol") = 1e-3, py::arg("max_outer") = 50, py::arg("max_inner") = 10000);
}
