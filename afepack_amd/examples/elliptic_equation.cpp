// elliptic_equation.cpp -- the reference's variable-coefficient example
// (examples/elliptic_equation/{EllipticEquation.h,EllipticEquation.cpp,
// main.cpp}) against the drop-in headers: -div(a grad u) = f with a
// piecewise-constant a, the element matrix written by the user (evaluated on
// the host, gathered into the CSR on the device), AMGSolver.
//
// usage: elliptic_equation [easymesh-base | n]
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>

#include <AFEPack/AMGSolver.h>
#include <AFEPack/BilinearOperator.h>
#include <AFEPack/EasyMesh.h>
#include <AFEPack/FEMSpace.h>
#include <AFEPack/Functional.h>
#include <AFEPack/Operator.h>
#include <AFEPack/TemplateElement.h>

#define DIM 2
#define PI 3.14159265358979323846

double _a_(const double *p) { return (p[0] * p[0] > p[1] * p[1]) ? 400. : 1.; }
double _u_(const double *p) { return cos(2.0 * PI * p[0]) * sin(3.0 * PI * p[1]); }
double _f_(const double *p) { return _a_(p) * 13 * PI * PI * _u_(p); }

class Matrix : public StiffMatrix<DIM, double> {
public:
  Matrix(FEMSpace<double, DIM> &sp) : StiffMatrix<DIM, double>(sp) {}
  virtual ~Matrix() {}
  virtual void getElementMatrix(const Element<double, 2> &ele0, const Element<double, 2> &ele1,
                                const ActiveElementPairIterator<DIM>::State state) {
    double volume = ele0.templateElement().volume();
    const QuadratureInfo<DIM> &quad_info = ele0.findQuadratureInfo(algebricAccuracy());
    std::vector<double> jacobian = ele0.local_to_global_jacobian(quad_info.quadraturePoint());
    int n_quadrature_point = quad_info.n_quadraturePoint();
    std::vector<AFEPack::Point<DIM>> q_point = ele0.local_to_global(quad_info.quadraturePoint());
    std::vector<std::vector<std::vector<double>>> basis_grad = ele0.basis_function_gradient(q_point);
    int n_ele_dof = ele0.dof().size();
    for (int l = 0; l < n_quadrature_point; l++) {
      double Jxw = quad_info.weight(l) * jacobian[l] * volume;
      double a_val = _a_(q_point[l]);
      for (int j = 0; j < n_ele_dof; j++)
        for (int k = 0; k < n_ele_dof; k++)
          elementMatrix(j, k) += Jxw * a_val * innerProduct(basis_grad[j][l], basis_grad[k][l]);
    }
  }
};

int main(int argc, char *argv[]) {
  EasyMesh mesh;
  const std::string arg = argc > 1 ? argv[1] : "64";
  if (std::ifstream(arg + ".n"))
    mesh.readData(arg);
  else
    buildSquareMesh(mesh, atoi(arg.c_str()));
  // the square [-1,1]^2 of the reference's mesh
  for (auto &c : mesh.coords) c = 2.0 * c - 1.0;

  TemplateGeometry<DIM> triangle_template_geometry;
  CoordTransform<DIM, DIM> triangle_coord_transform;
  TemplateDOF<DIM> triangle_template_dof;
  BasisFunctionAdmin<double, DIM, DIM> triangle_basis_function;
  triangle_template_geometry.readData("triangle.tmp_geo");
  triangle_coord_transform.readData("triangle.crd_trs");
  triangle_template_dof.reinit(triangle_template_geometry);
  triangle_template_dof.readData("triangle.1.tmp_dof");
  triangle_basis_function.reinit(triangle_template_dof);
  triangle_basis_function.readData("triangle.1.bas_fun");
  std::vector<TemplateElement<double, DIM, DIM>> template_element(1);
  template_element[0].reinit(triangle_template_geometry, triangle_template_dof, triangle_coord_transform,
                             triangle_basis_function);

  FEMSpace<double, DIM> fem_space;
  fem_space.reinit(mesh, template_element);
  int n_element = mesh.n_geometry(DIM);
  fem_space.element().resize(n_element);
  for (int i = 0; i < n_element; i++) fem_space.element(i).reinit(fem_space, i, 0);
  fem_space.buildElement();
  fem_space.buildDof();
  fem_space.buildDofBoundaryMark();

  // accuracy 4 (6-point rule) rather than the reference's 3: the 4-point
  // rule of accuracy 3 has a negative centroid weight, and on this mesh the
  // line x = -y cuts elements whose coefficient is 400 at the centroid and 1
  // at the other points, which makes their element matrices indefinite
  Matrix stiff_matrix(fem_space);
  stiff_matrix.algebricAccuracy() = 4;
  stiff_matrix.build();

  FEMFunction<double, DIM> u_h(fem_space);
  Vector<double> f_h;
  Operator::L2Discretize(&_f_, fem_space, f_h, 4);

  BoundaryFunction<double, DIM> boundary(BoundaryConditionInfo::DIRICHLET, 1, &_u_);
  BoundaryConditionAdmin<double, DIM> boundary_admin(fem_space);
  boundary_admin.add(boundary);
  boundary_admin.apply(stiff_matrix, u_h, f_h);

  AMGSolver solver(stiff_matrix);
  solver.solve(u_h, f_h);

  u_h.writeOpenDXData("u.dx");
  std::cerr << "L2 error = " << Functional::L2Error(u_h, FunctionFunction<double>(&_u_), 3) << std::endl;
  return 0;
}
