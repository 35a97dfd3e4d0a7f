// Geom.h -- immersed bodies as closed counter-clockwise polygons.
//
// Role of the reference's object/ classes: Circle::Initialize
// (object/Circle.cpp:10-19) draws a circle as n Lagrangian points, and the
// mesh cuts the cells against the polygon's segments (meshes/
// CartesianMesh.cpp:100-169).  The cutting happens inside the library
// (dsmc_create builds the cut-cell tables); the host only supplies the
// outline, as x, y pairs.
#pragma once

#include <vector>

#include "Param.h"

namespace dsmc_host {

// Regular n-gon of radius r around (cx, cy); vertex k at angle 2 pi k / n.
std::vector<double> circle_polygon(int n, double cx, double cy, double r);

// Isosceles wedge, apex upstream at (apex_x, cy), base at apex_x + length.
std::vector<double> wedge_polygon(double apex_x, double cy, double length, double half_angle_deg);

// The outline the parameters name (empty for geometry "none").
std::vector<double> body_of(const Param &p);

}  // namespace dsmc_host
