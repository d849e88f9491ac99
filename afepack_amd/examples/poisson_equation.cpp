// poisson_equation.cpp -- the reference's Poisson example
// (examples/poisson_equation/poisson_equation.cpp) written against the
// drop-in headers: -Delta u = 5 pi^2 u on a triangulation, u = sin(pi x)
// sin(2 pi y) on the boundary, P1 elements, AMGSolver.
//
// usage: poisson_equation [easymesh-base | n]
//   an EasyMesh base name reads <base>.n/.s/.e; a number n (default 64)
//   meshes the unit square with n x n squares.
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>

#include <AFEPack/AMGSolver.h>
#include <AFEPack/EasyMesh.h>
#include <AFEPack/FEMSpace.h>
#include <AFEPack/Functional.h>
#include <AFEPack/Geometry.h>
#include <AFEPack/Operator.h>
#include <AFEPack/TemplateElement.h>

#define PI (4.0 * atan(1.0))

double u(const double *);
double f(const double *);

int main(int argc, char *argv[]) {
  EasyMesh mesh;
  const std::string arg = argc > 1 ? argv[1] : "64";
  if (std::ifstream(arg + ".n"))
    mesh.readData(arg);
  else
    buildSquareMesh(mesh, atoi(arg.c_str()));

  TemplateGeometry<2> triangle_template_geometry;
  triangle_template_geometry.readData("triangle.tmp_geo");
  CoordTransform<2, 2> triangle_coord_transform;
  triangle_coord_transform.readData("triangle.crd_trs");
  TemplateDOF<2> triangle_template_dof(triangle_template_geometry);
  triangle_template_dof.readData("triangle.1.tmp_dof");
  BasisFunctionAdmin<double, 2, 2> triangle_basis_function(triangle_template_dof);
  triangle_basis_function.readData("triangle.1.bas_fun");

  std::vector<TemplateElement<double, 2, 2>> template_element(1);
  template_element[0].reinit(triangle_template_geometry, triangle_template_dof, triangle_coord_transform,
                             triangle_basis_function);

  FEMSpace<double, 2> fem_space;
  fem_space.reinit(mesh, template_element);

  int n_element = mesh.n_geometry(2);
  fem_space.element().resize(n_element);
  for (int i = 0; i < n_element; i++) fem_space.element(i).reinit(fem_space, i, 0);

  fem_space.buildElement();
  fem_space.buildDof();
  fem_space.buildDofBoundaryMark();

  StiffMatrix<2, double> stiff_matrix(fem_space);
  stiff_matrix.algebricAccuracy() = 4;
  stiff_matrix.build();

  FEMFunction<double, 2> solution(fem_space);
  Vector<double> right_hand_side;
  Operator::L2Discretize(&f, fem_space, right_hand_side, 4);

  BoundaryFunction<double, 2> boundary1(BoundaryConditionInfo::DIRICHLET, 1, &u);
  BoundaryFunction<double, 2> boundary2(BoundaryConditionInfo::DIRICHLET, 2, &u);
  BoundaryFunction<double, 2> boundary3(BoundaryConditionInfo::DIRICHLET, 3, &u);
  BoundaryFunction<double, 2> boundary4(BoundaryConditionInfo::DIRICHLET, 4, &u);
  BoundaryConditionAdmin<double, 2> boundary_admin(fem_space);
  boundary_admin.add(boundary1);
  boundary_admin.add(boundary2);
  boundary_admin.add(boundary3);
  boundary_admin.add(boundary4);
  boundary_admin.apply(stiff_matrix, solution, right_hand_side);

  AMGSolver solver(stiff_matrix);
  solver.solve(solution, right_hand_side, 1.0e-08, 200);

  solution.writeOpenDXData("u.dx");
  double error = Functional::L2Error(solution, FunctionFunction<double>(&u), 3);
  std::cerr << "\nL2 error = " << error << std::endl;

  return 0;
}

double u(const double *p) { return sin(PI * p[0]) * sin(2 * PI * p[1]); }

double f(const double *p) { return 5 * PI * PI * u(p); }
